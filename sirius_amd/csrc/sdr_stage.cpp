// sdr_stage.cpp — staging of segment bytes from a file into device memory.
//
// The reference (src/op/scan/duckdb_native_decoder.cpp:675-855,
// submit_and_await) sorts a split's segment reads by file offset, coalesces
// reads whose gap is at most coalesce_max_gap into pieces of pinned host
// memory, reads them, and then issues one host-to-device copy per segment
// into a device buffer where each segment is 16-byte aligned.
//
// Here the pieces are packed into one pinned buffer and leave in a single
// host-to-device copy: the device image keeps each piece's file layout, so
// a segment sits at its piece's position plus its offset in the piece.  The
// decoders take segments at any byte alignment, so no gather follows.
// Reads of up to 4 MiB are spread over up to 8 threads.  Pinned buffers are
// reused per device, up to eight of them, so one call's reads overlap the
// previous call's copy and calls from several threads run side by side.
#include "sdr_internal.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

namespace sdr {
namespace {

constexpr uint64_t PIECE_ALIGN  = 16;
constexpr uint64_t READ_BYTES   = 4ull << 20;  // one pread
constexpr size_t READ_THREADS   = 8;

struct Read {
  uint64_t file_off, len, pos;
};

struct Piece {
  uint64_t file_off, file_end, pos;
};

// Pieces of the extents, sorted by file offset; each extent's position in
// the packed image goes to ext[i].device_offset.  Returns the image size.
uint64_t plan(sdr_extent* ext, uint32_t n, uint64_t max_gap, std::vector<Piece>& pieces)
{
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return ext[a].file_offset < ext[b].file_offset; });
  pieces.clear();
  uint64_t end = 0;
  for (uint32_t i : order) {
    const uint64_t lo = ext[i].file_offset, hi = lo + ext[i].bytes;
    if (ext[i].bytes == 0) {
      ext[i].device_offset = 0;
      continue;
    }
    // a new piece unless the extent starts within max_gap of the current
    // piece's end (overlapping extents share bytes)
    if (pieces.empty() || lo > pieces.back().file_end + max_gap) {
      const uint64_t pos = (end + PIECE_ALIGN - 1u) & ~(PIECE_ALIGN - 1u);
      pieces.push_back({lo, hi, pos});
    } else if (hi > pieces.back().file_end) {
      pieces.back().file_end = hi;
    }
    const Piece& p       = pieces.back();
    ext[i].device_offset = p.pos + (lo - p.file_off);
    end                  = p.pos + (p.file_end - p.file_off);
  }
  return (end + PIECE_ALIGN - 1u) & ~(PIECE_ALIGN - 1u);
}

// Pinned buffers reused across calls, per device.  The pool's lock only
// guards choosing a buffer (non-blocking event queries); the wait for its
// last copy, a growth, the file reads and the enqueue run outside it on a
// buffer the caller holds, so calls from several threads overlap.
struct PinnedBuf {
  void* h         = nullptr;
  uint64_t cap    = 0;
  hipEvent_t done = nullptr;  // recorded after the copy that last read it
  int dev         = -1;
  bool busy       = false;    // held by a caller
};
std::mutex g_stage_mu;
std::vector<PinnedBuf*> g_bufs;  // never shrinks: held pointers stay valid
constexpr size_t BUFS_PER_DEV = 8;

// A buffer of this device for `total` bytes: an idle one whose last copy has
// finished (the largest fitting, else any), else a new one while the device
// has fewer than BUFS_PER_DEV, else the idle one with the oldest copy
// (waited for by the caller).  Returns nullptr when all are held.
PinnedBuf* take_buf(int dev, uint64_t total)
{
  std::lock_guard<std::mutex> lk(g_stage_mu);
  PinnedBuf *done = nullptr, *idle = nullptr;
  size_t n_dev = 0;
  for (PinnedBuf* b : g_bufs) {
    if (b->dev != dev) continue;
    ++n_dev;
    if (b->busy) continue;
    idle = idle ? idle : b;
    if (!b->done || hipEventQuery(b->done) == hipSuccess)
      if (!done || (b->cap >= total && done->cap < total)) done = b;
  }
  PinnedBuf* b = done;
  if (!b && n_dev < BUFS_PER_DEV) {
    b      = new PinnedBuf;
    b->dev = dev;
    g_bufs.push_back(b);
  }
  if (!b) b = idle;
  if (b) b->busy = true;
  return b;
}

void give_buf(PinnedBuf* b)
{
  std::lock_guard<std::mutex> lk(g_stage_mu);
  b->busy = false;
}

}  // namespace
}  // namespace sdr

using namespace sdr;

extern "C" {

uint64_t sdr_stage_plan(sdr_extent* extents, uint32_t n, uint64_t max_gap)
{
  if (!extents && n) return 0;
  std::vector<Piece> pieces;
  return plan(extents, n, max_gap, pieces);
}

int sdr_stage_file(const char* path, sdr_extent* extents, uint32_t n, uint64_t max_gap, void* d_dst,
                   uint64_t dst_capacity, sdr_stream_t stream)
{
  clear_error();
  if (!path) return fail(SDR_E_INVALID_ARG, "path is NULL");
  if (n && !extents) return fail(SDR_E_INVALID_ARG, "extents is NULL");
  std::vector<Piece> pieces;
  const uint64_t total = plan(extents, n, max_gap, pieces);
  if (total == 0) return SDR_OK;
  if (!d_dst) return fail(SDR_E_INVALID_ARG, "d_dst is NULL");
  if (total > dst_capacity)
    return fail(SDR_E_CAPACITY, "staged image needs %llu bytes, d_dst holds %llu",
                static_cast<unsigned long long>(total), static_cast<unsigned long long>(dst_capacity));
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(SDR_E_HIP, "device %d out of range", dev);
  const int fd = open(path, O_RDONLY | O_CLOEXEC);
  if (fd < 0) return fail(SDR_E_INVALID_ARG, "cannot open %s: %s", path, std::strerror(errno));

  PinnedBuf* held = take_buf(dev, total);
  if (!held) {
    close(fd);
    return fail(SDR_E_CAPACITY, "all %zu staging buffers of device %d are in use", BUFS_PER_DEV, dev);
  }
  PinnedBuf& pb = *held;
  hipError_t e  = hipSuccess;
  if (pb.done) e = hipEventSynchronize(pb.done);  // the copy that last read this buffer is done
  if (e == hipSuccess && pb.cap < total) {
    if (pb.h) e = hipHostFree(pb.h);
    pb.h   = nullptr;
    pb.cap = 0;
    if (e == hipSuccess) e = hipHostMalloc(&pb.h, total, hipHostMallocDefault);
    if (e == hipSuccess) pb.cap = total;
  }
  if (e == hipSuccess && !pb.done) e = hipEventCreateWithFlags(&pb.done, hipEventDisableTiming);
  if (e != hipSuccess) {
    close(fd);
    give_buf(held);
    return fail(SDR_E_HIP, "staging buffer: %s", hipGetErrorString(e));
  }
  // reads of at most READ_BYTES, spread over up to READ_THREADS threads
  auto* h = static_cast<uint8_t*>(pb.h);
  std::vector<Read> reads;
  for (const Piece& p : pieces)
    for (uint64_t o = p.file_off; o < p.file_end; o += READ_BYTES)
      reads.push_back({o, std::min(p.file_end, o + READ_BYTES) - o, p.pos + (o - p.file_off)});
  std::atomic<size_t> next{0};
  std::atomic<uint64_t> bad{UINT64_MAX};
  auto work = [&]() {
    for (size_t i; (i = next.fetch_add(1)) < reads.size();) {
      const Read& r = reads[i];
      uint64_t done = 0;
      while (done < r.len) {
        const ssize_t got = pread(fd, h + r.pos + done, r.len - done, static_cast<off_t>(r.file_off + done));
        if (got < 0 && errno == EINTR) continue;
        if (got <= 0) {
          uint64_t cur = bad.load();
          while (r.file_off + done < cur && !bad.compare_exchange_weak(cur, r.file_off + done)) {}
          break;
        }
        done += static_cast<uint64_t>(got);
      }
    }
  };
  const size_t n_threads = std::min<size_t>(READ_THREADS, (total + READ_BYTES - 1) / READ_BYTES);
  std::vector<std::thread> pool;
  for (size_t t = 1; t < n_threads; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  close(fd);
  if (bad.load() != UINT64_MAX) {
    give_buf(held);
    return fail(SDR_E_BOUNDS, "short read of %s at offset %llu", path, static_cast<unsigned long long>(bad.load()));
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  e             = hipMemcpyAsync(d_dst, pb.h, total, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipEventRecord(pb.done, s);
  if (e != hipSuccess) (void)hipStreamSynchronize(s);  // the buffer is not tracked: nothing may still read it
  give_buf(held);
  if (e != hipSuccess) return fail(SDR_E_HIP, "staging copy: %s", hipGetErrorString(e));
  return SDR_OK;
}

}  // extern "C"
