// sdr_api.cpp — C ABI of the decoder (include/sdr/decode.h): validation,
// descriptor upload and kernel dispatch for the fixed-width path.
//
// Validation mirrors the reference's viability throws
// (src/cuda/scan/gpu_native_decode.cu:86-101, :382-385, :414-417, :496-500,
// :555-570) but runs over every column before any device work is queued, so
// a failing call leaves the outputs untouched.
#include "sdr_internal.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace sdr {

namespace {
// A fixed buffer: reporting an error never allocates.
thread_local char g_err[512];
}

int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = '\0'; }
const char* last_error() { return g_err; }

// The stream-ordered allocator keeps freed blocks cached instead of
// returning them to the driver at every synchronisation.
void ensure_pool(int dev)
{
  static std::once_flag flags[64];
  if (dev < 0 || dev >= 64) return;
  std::call_once(flags[dev], [dev]() {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t thr = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
  });
}

int device_alloc(void** p, size_t bytes, hipStream_t s)
{
  *p = nullptr;
  if (bytes == 0) return SDR_OK;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  ensure_pool(dev);
  HIP_TRY(hipMallocAsync(p, bytes, s));
  return SDR_OK;
}

void device_free(void* p, hipStream_t s)
{
  if (p) (void)hipFreeAsync(p, s);
}

namespace {
// Pinned staging slots for descriptor uploads, kept for the process.  The
// list's lock only guards choosing a slot (a non-blocking event query per
// slot); allocation, the host copy and the enqueue run outside it on a slot
// the caller holds.
struct StageSlot {
  void* h          = nullptr;
  size_t cap       = 0;
  hipEvent_t done  = nullptr;  // recorded after the slot's last copy
  int dev          = -1;
  bool busy        = false;    // held by a caller between choice and enqueue
};
std::mutex g_stage_mu;
std::vector<StageSlot*> g_stage;  // never shrinks: held pointers stay valid
constexpr size_t STAGE_SLOTS_MAX = 64;
constexpr size_t STAGE_MIN_BYTES = 64u << 10;
}  // namespace

int upload_async(void* d_dst, const void* h_src, size_t bytes, hipStream_t s)
{
  if (bytes == 0) return SDR_OK;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  StageSlot* slot = nullptr;
  bool fresh      = false;
  {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    for (StageSlot* sl : g_stage)
      if (sl->dev == dev && !sl->busy && sl->cap >= bytes && hipEventQuery(sl->done) == hipSuccess) {
        slot = sl;
        break;
      }
    if (!slot && g_stage.size() < STAGE_SLOTS_MAX) {
      slot      = new StageSlot;
      slot->dev = dev;
      g_stage.push_back(slot);
      fresh = true;
    }
    if (slot) slot->busy = true;
  }
  if (!slot) {  // every slot busy: a copy from the caller's memory (the runtime stages it)
    HIP_TRY(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s));
    return SDR_OK;
  }
  hipError_t e = hipSuccess;
  if (fresh) {
    size_t cap = STAGE_MIN_BYTES;
    while (cap < bytes) cap *= 2;
    e = hipHostMalloc(&slot->h, cap, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&slot->done, hipEventDisableTiming);
    if (e == hipSuccess) slot->cap = cap;
  }
  if (e == hipSuccess) {
    std::memcpy(slot->h, h_src, bytes);
    e = hipMemcpyAsync(d_dst, slot->h, bytes, hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) e = hipEventRecord(slot->done, s);
  {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    slot->busy = false;  // a slot whose set-up failed keeps cap 0 and is never chosen
  }
  if (e != hipSuccess) return fail(SDR_E_HIP, "descriptor upload: %s", hipGetErrorString(e));
  return SDR_OK;
}

namespace {
// Descriptor slots for the per-call decode: a pinned host buffer and a
// device buffer, reused once the event recorded after the last launch that
// read them has completed.  At most DESC_SLOTS_PER_DEV are in flight per
// device; a call beyond that waits for the oldest one (the GPU still has
// the others queued), so once the slots exist a decode allocates nothing.
// The depth lets the GPU ride out host stalls of a few milliseconds.
struct DescSlot {
  void* h         = nullptr;
  void* d         = nullptr;
  size_t cap      = 0;
  hipEvent_t done = nullptr;
  hipEvent_t up   = nullptr;  // recorded after the slot's upload on the copy stream
  int dev         = -1;
  bool held       = false;  // acquired and not yet released
  uint64_t seq    = 0;      // release order
};
std::mutex g_desc_mu;
std::vector<DescSlot*> g_desc;
uint64_t g_desc_seq = 0;
constexpr size_t DESC_SLOTS_PER_DEV = 32;  // about 8 ms of headline steps queued ahead of the host
constexpr size_t DESC_MIN_BYTES     = 64u << 10;

// One copy stream per device (default priority, not synchronising with the
// null stream) carries descriptor uploads while the caller's stream is busy,
// so a call's upload runs while that stream still works on earlier calls;
// the caller's stream then waits only for that copy's event.  Created once
// per device.
hipStream_t g_copy_stream[64] = {};
std::once_flag g_copy_once[64];

int copy_stream(int dev, hipStream_t* out)
{
  if (dev < 0 || dev >= 64) return fail(SDR_E_HIP, "device %d out of range", dev);
  hipError_t e = hipSuccess;
  std::call_once(g_copy_once[dev], [&]() { e = hipStreamCreateWithFlags(&g_copy_stream[dev], hipStreamNonBlocking); });
  if (!g_copy_stream[dev]) return fail(SDR_E_HIP, "copy stream: %s", hipGetErrorString(e));
  *out = g_copy_stream[dev];
  return SDR_OK;
}
}  // namespace

int desc_acquire(const void* h_src, size_t bytes, hipStream_t s, void** d_out, void** token)
{
  *d_out = nullptr;
  *token = nullptr;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  DescSlot* pick = nullptr;
  bool fresh = false, wait = false;
  {
    // choose under the lock (non-blocking event queries only); the waits,
    // allocations and copies below run outside it on the held slot
    std::lock_guard<std::mutex> lk(g_desc_mu);
    DescSlot* oldest = nullptr;
    size_t n_dev     = 0;
    for (DescSlot* sl : g_desc) {
      if (sl->dev != dev) continue;
      ++n_dev;
      if (sl->held) continue;
      if (hipEventQuery(sl->done) == hipSuccess) {
        if (!pick || (sl->cap >= bytes && pick->cap < bytes)) pick = sl;
      } else if (!oldest || sl->seq < oldest->seq) {
        oldest = sl;
      }
    }
    if (!pick && (n_dev < DESC_SLOTS_PER_DEV || !oldest)) {
      pick      = new DescSlot;
      pick->dev = dev;
      g_desc.push_back(pick);
      fresh = true;
    } else if (!pick) {  // every slot in flight: wait for the oldest
      pick = oldest;
      wait = true;
    }
    pick->held = true;
  }
  auto unhold = [&](int rc) {
    std::lock_guard<std::mutex> lk(g_desc_mu);
    pick->held = false;
    return rc;
  };
  hipError_t e = hipSuccess;
  if (fresh) {
    e = hipEventCreateWithFlags(&pick->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pick->up, hipEventDisableTiming);
    if (e != hipSuccess) {  // leave a slot whose events exist, or none
      std::lock_guard<std::mutex> lk(g_desc_mu);
      g_desc.erase(std::find(g_desc.begin(), g_desc.end(), pick));
      if (pick->done) (void)hipEventDestroy(pick->done);
      delete pick;
      return fail(SDR_E_HIP, "descriptor slot event creation failed");
    }
  }
  if (wait) e = hipEventSynchronize(pick->done);
  if (e == hipSuccess && pick->cap < bytes) {  // its event has completed: nothing still reads the old buffers
    if (pick->h) e = hipHostFree(pick->h);
    if (e == hipSuccess && pick->d) e = hipFree(pick->d);
    pick->h = pick->d = nullptr;
    pick->cap         = 0;
    size_t cap        = DESC_MIN_BYTES;
    while (cap < bytes) cap *= 2;
    if (e == hipSuccess) e = hipHostMalloc(&pick->h, cap, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&pick->d, cap);
    if (e == hipSuccess) pick->cap = cap;
  }
  if (e != hipSuccess) return unhold(fail(SDR_E_HIP, "descriptor slot: %s", hipGetErrorString(e)));
  std::memcpy(pick->h, h_src, bytes);
  // The copy goes on the copy stream only while the caller's stream still
  // has work queued: on an idle stream the cross-stream wait would only add
  // latency.  Under stream capture it stays on the caller's stream, as a
  // node of the caller's graph.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  e = hipStreamIsCapturing(s, &cap);
  if (e == hipSuccess && cap == hipStreamCaptureStatusNone && hipStreamQuery(s) == hipErrorNotReady) {
    hipStream_t cs = nullptr;
    const int rc   = copy_stream(dev, &cs);
    if (rc) return unhold(rc);
    e = hipMemcpyAsync(pick->d, pick->h, bytes, hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = hipEventRecord(pick->up, cs);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, pick->up, 0);
  } else if (e == hipSuccess) {
    e = hipMemcpyAsync(pick->d, pick->h, bytes, hipMemcpyHostToDevice, s);
  }
  if (e != hipSuccess) return unhold(fail(SDR_E_HIP, "descriptor upload: %s", hipGetErrorString(e)));
  *d_out = pick->d;
  *token = pick;
  return SDR_OK;
}

int desc_release(void* token, hipStream_t s)
{
  if (!token) return SDR_OK;
  auto* sl = static_cast<DescSlot*>(token);
  const hipError_t e = hipEventRecord(sl->done, s);  // the slot is still held: no other caller touches it
  if (e != hipSuccess) (void)hipStreamSynchronize(s);  // the slot cannot be tracked: wait for the stream instead
  {
    std::lock_guard<std::mutex> lk(g_desc_mu);
    sl->seq  = ++g_desc_seq;
    sl->held = false;
  }
  if (e != hipSuccess) return fail(SDR_E_HIP, "descriptor slot event: %s", hipGetErrorString(e));
  return SDR_OK;
}

bool type_ok_for(uint32_t codec, uint32_t ts)
{
  switch (codec) {
    case SDR_CODEC_BITPACKING:
    case SDR_CODEC_RLE: return ts == 1 || ts == 2 || ts == 4 || ts == 8;
    case SDR_CODEC_ALP:
    case SDR_CODEC_ALPRD: return ts == 4 || ts == 8;
    case SDR_CODEC_UNCOMPRESSED:
    case SDR_CODEC_CONSTANT: return ts == 1 || ts == 2 || ts == 4 || ts == 8 || ts == 16;
    default: return false;
  }
}

bool codec_fixed(uint32_t codec)
{
  return codec == SDR_CODEC_BITPACKING || codec == SDR_CODEC_RLE || codec == SDR_CODEC_ALP ||
         codec == SDR_CODEC_ALPRD || codec == SDR_CODEC_UNCOMPRESSED || codec == SDR_CODEC_CONSTANT;
}

int check_runs(const sdr_run* runs, uint32_t n, uint32_t total_rows, const char* what)
{
  if (n && !runs) return fail(SDR_E_INVALID_ARG, "%s runs pointer is NULL", what);
  for (uint32_t r = 0; r < n; ++r) {
    if (runs[r].n_segments && !runs[r].segments)
      return fail(SDR_E_INVALID_ARG, "%s run %u has a NULL segment array", what, r);
    for (uint32_t s = 0; s < runs[r].n_segments; ++s) {
      const sdr_segment& g = runs[r].segments[s];
      if (uint64_t{g.row_offset} + g.row_count > total_rows)
        return fail(SDR_E_BOUNDS, "%s segment row_offset (%u) + row_count (%u) > total_rows (%u)", what,
                    g.row_offset, g.row_count, total_rows);
      if (g.row_count && !g.d_bytes) return fail(SDR_E_INVALID_ARG, "%s segment %u has NULL bytes", what, s);
    }
  }
  return SDR_OK;
}

int check_validity_runs(const sdr_run* runs, uint32_t n, uint32_t total_rows)
{
  int rc = check_runs(runs, n, total_rows, "validity");
  if (rc) return rc;
  for (uint32_t r = 0; r < n; ++r) {
    if (runs[r].codec != SDR_CODEC_UNCOMPRESSED)
      return fail(SDR_E_UNSUPPORTED_CODEC, "validity codec %u not implemented", runs[r].codec);
    for (uint32_t s = 0; s < runs[r].n_segments; ++s) {
      const sdr_segment& g = runs[r].segments[s];
      if (!g.row_count) continue;
      if (g.row_offset % 8u) return fail(SDR_E_ALIGNMENT, "validity row_offset (%u) not byte-aligned", g.row_offset);
      if (g.bytes_size < (g.row_count + 7u) / 8u)
        return fail(SDR_E_BOUNDS, "validity segment bytes_size (%u) < required %u", g.bytes_size,
                    (g.row_count + 7u) / 8u);
    }
  }
  return SDR_OK;
}

// Overlay the UNCOMPRESSED validity bytes on an all-valid mask and count the
// nulls into *d_count (src/cuda/scan/gpu_native_decode.cu:396-476, :519-538).
int build_null_mask(const sdr_run* runs, uint32_t n, uint32_t total_rows, uint8_t* d_mask,
                    unsigned long long* d_count, hipStream_t s)
{
  HIP_TRY(launch_fill_u8(d_mask, 0xFF, (total_rows + 7u) / 8u, s));
  for (uint32_t r = 0; r < n; ++r)
    for (uint32_t k = 0; k < runs[r].n_segments; ++k) {
      const sdr_segment& g = runs[r].segments[k];
      if (!g.row_count) continue;
      HIP_TRY(hipMemcpyAsync(d_mask + g.row_offset / 8u, g.d_bytes, (g.row_count + 7u) / 8u,
                             hipMemcpyDeviceToDevice, s));
    }
  HIP_TRY(launch_count_nulls(d_mask, total_rows, d_count, s));
  return SDR_OK;
}

int validate_fixed_columns(const sdr_column* cols, uint32_t n_cols)
{
  if (n_cols && !cols) return fail(SDR_E_INVALID_ARG, "cols is NULL");
  for (uint32_t c = 0; c < n_cols; ++c) {
    const sdr_column& col = cols[c];
    if (col.total_rows != cols[0].total_rows)
      return fail(SDR_E_BOUNDS, "all columns must share the same total_rows");
    if (col.total_rows > 0x7FFFFFFFu) return fail(SDR_E_BOUNDS, "total_rows (%u) > INT32_MAX", col.total_rows);
    if (!(col.type_size == 1 || col.type_size == 2 || col.type_size == 4 || col.type_size == 8 ||
          col.type_size == 16))
      return fail(SDR_E_BAD_WIDTH, "column %u: type_size %u is not fixed-width", c, col.type_size);
    if (col.total_rows && !col.d_out) return fail(SDR_E_INVALID_ARG, "column %u: d_out is NULL", c);
    int rc = check_runs(col.data, col.n_data, col.total_rows, "data");
    if (rc) return rc;
    for (uint32_t r = 0; r < col.n_data; ++r) {
      const sdr_run& run = col.data[r];
      if (!codec_fixed(run.codec))
        return fail(SDR_E_UNSUPPORTED_CODEC, "viability invariant violated: data codec %u not implemented", run.codec);
      if (!type_ok_for(run.codec, col.type_size))
        return fail(SDR_E_BAD_WIDTH, "viability invariant violated: %s type_size %u", sdr_codec_name(run.codec),
                    col.type_size);
      for (uint32_t s = 0; s < run.n_segments; ++s) {
        const sdr_segment& g = run.segments[s];
        if (!g.row_count) continue;
        if (run.codec == SDR_CODEC_UNCOMPRESSED && uint64_t{g.bytes_size} < uint64_t{g.row_count} * col.type_size)
          return fail(SDR_E_BOUNDS, "UNCOMPRESSED segment bytes_size (%u) < required %llu", g.bytes_size,
                      (unsigned long long)(uint64_t{g.row_count} * col.type_size));
        if (run.codec == SDR_CODEC_CONSTANT && g.bytes_size < col.type_size)
          return fail(SDR_E_BOUNDS, "CONSTANT segment bytes_size (%u) < type_size (%u)", g.bytes_size, col.type_size);
      }
    }
    if (col.has_nulls && col.n_validity) {
      rc = check_validity_runs(col.validity, col.n_validity, col.total_rows);
      if (rc) return rc;
      if (col.total_rows && !col.d_null_mask) return fail(SDR_E_INVALID_ARG, "column %u: d_null_mask is NULL", c);
    }
  }
  return SDR_OK;
}

namespace {

hipError_t launch_batch(const Batch& b, const FixedSeg* d, uint32_t base, uint32_t n, uint32_t max_rows,
                        hipStream_t s, bool host)
{
  switch (b.codec) {
    case SDR_CODEC_BITPACKING: return launch_bitpacking(d, base, n, max_rows, b.ts, s, host);
    case SDR_CODEC_RLE: return launch_rle(d, base, n, max_rows, b.ts, s, host);
    case SDR_CODEC_ALP: return launch_alp(d, base, n, max_rows, b.ts, false, s, host);
    case SDR_CODEC_ALPRD: return launch_alp(d, base, n, max_rows, b.ts, true, s, host);
    case SDR_CODEC_UNCOMPRESSED: return launch_uncompressed(d, base, n, max_rows, b.ts, s, host);
    case SDR_CODEC_CONSTANT: return launch_constant(d, base, n, max_rows, b.ts, s, host);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

// Group every live segment by (codec, type size): one device descriptor
// array, one launch per group (or one per segment in the plain mode).
int build_batches(const std::vector<RunTarget>& targets, std::vector<Batch>& batches, std::vector<FixedSeg>& all)
{
  std::vector<std::vector<FixedSeg>> per;
  for (const RunTarget& t : targets) {
    size_t bi = 0;
    while (bi < batches.size() && !(batches[bi].codec == t.run->codec && batches[bi].ts == t.type_size)) ++bi;
    if (bi == batches.size()) {
      batches.push_back({t.run->codec, t.type_size, 0, 0, 0});
      per.emplace_back();
    }
    for (uint32_t k = 0; k < t.run->n_segments; ++k) {
      const sdr_segment& g = t.run->segments[k];
      if (!g.row_count) continue;
      FixedSeg f;
      f.bytes      = g.d_bytes;
      f.out        = static_cast<uint8_t*>(t.d_out) + uint64_t{g.row_offset} * t.type_size;
      f.bytes_size = g.bytes_size;
      f.row_count  = g.row_count;
      per[bi].push_back(f);
      batches[bi].max_rows = std::max(batches[bi].max_rows, g.row_count);
    }
  }
  for (size_t i = 0; i < batches.size(); ++i) {
    batches[i].first = static_cast<uint32_t>(all.size());
    batches[i].count = static_cast<uint32_t>(per[i].size());
    all.insert(all.end(), per[i].begin(), per[i].end());
  }
  return SDR_OK;
}

// `d`: the descriptors on the device, or null to carry them in the kernel
// arguments from `all`.
hipError_t launch_batches(const std::vector<Batch>& batches, const std::vector<FixedSeg>& all, const FixedSeg* d,
                          uint32_t mode, uint32_t codec_filter, hipStream_t s, int* n_launches)
{
  hipError_t e = hipSuccess;
  int n = 0;
  const bool host      = d == nullptr;
  const FixedSeg* segs = host ? all.data() : d;
  for (size_t i = 0; e == hipSuccess && i < batches.size(); ++i) {
    const Batch& b = batches[i];
    if (!b.count || (codec_filter && b.codec != codec_filter)) continue;
    if (mode == SDR_LAUNCH_PER_SEGMENT) {
      for (uint32_t k = 0; k < b.count && e == hipSuccess; ++k, ++n)
        e = launch_batch(b, segs, b.first + k, 1, all[b.first + k].row_count, s, host);
    } else {
      e = launch_batch(b, segs, b.first, b.count, b.max_rows, s, host);
      ++n;
    }
  }
  if (n_launches) *n_launches = n;
  return e;
}

// A call's descriptors travel in the kernel arguments when each launch holds
// one segment (the plain mode) or every (codec, width) group fits two
// launches of SEGS_INLINE: no upload copy, event or cross-stream wait.
bool inline_descriptors(const std::vector<Batch>& batches, uint32_t mode)
{
  if (mode == SDR_LAUNCH_PER_SEGMENT) return true;
  for (const Batch& b : batches)
    if (b.count > 2u * SEGS_INLINE) return false;
  return true;
}

int decode_runs(const std::vector<RunTarget>& targets, hipStream_t s, uint32_t mode)
{
  std::vector<Batch> batches;
  std::vector<FixedSeg> all;
  build_batches(targets, batches, all);
  if (all.empty()) return SDR_OK;
  if (inline_descriptors(batches, mode)) {
    const hipError_t e = launch_batches(batches, all, nullptr, mode, 0, s, nullptr);
    if (e != hipSuccess) return fail(SDR_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return SDR_OK;
  }
  void* d     = nullptr;
  void* token = nullptr;
  int rc = desc_acquire(all.data(), all.size() * sizeof(FixedSeg), s, &d, &token);
  if (rc) return rc;
  const hipError_t e = launch_batches(batches, all, static_cast<FixedSeg*>(d), mode, 0, s, nullptr);
  rc                 = desc_release(token, s);
  if (e != hipSuccess) return fail(SDR_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  return rc;
}

int decode_table(sdr_column* cols, uint32_t n_cols, hipStream_t s, uint32_t mode)
{
  clear_error();
  if (mode != SDR_LAUNCH_BATCHED && mode != SDR_LAUNCH_PER_SEGMENT && mode != SDR_LAUNCH_PERSISTENT)
    return fail(SDR_E_INVALID_ARG, "unknown launch mode %u", mode);
  int rc = validate_fixed_columns(cols, n_cols);
  if (rc) return rc;
  if (n_cols == 0) return SDR_OK;
  if (mode == SDR_LAUNCH_PERSISTENT) return decode_table_persistent(cols, n_cols, s);
  std::vector<RunTarget> targets;
  for (uint32_t c = 0; c < n_cols; ++c) {
    cols[c].null_count = 0;
    for (uint32_t r = 0; r < cols[c].n_data; ++r) targets.push_back({&cols[c].data[r], cols[c].d_out, cols[c].type_size});
  }
  rc = decode_runs(targets, s, mode);
  if (rc) return rc;
  // validity: one counter per column, one synchronisation for all of them
  std::vector<uint32_t> masked;
  for (uint32_t c = 0; c < n_cols; ++c)
    if (cols[c].has_nulls && cols[c].n_validity && cols[c].total_rows) masked.push_back(c);
  if (!masked.empty()) {
    unsigned long long* d_cnt = nullptr;
    rc = device_alloc(reinterpret_cast<void**>(&d_cnt), masked.size() * sizeof(unsigned long long), s);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(d_cnt, 0, masked.size() * sizeof(unsigned long long), s));
    for (size_t i = 0; i < masked.size(); ++i) {
      sdr_column& col = cols[masked[i]];
      rc = build_null_mask(col.validity, col.n_validity, col.total_rows, col.d_null_mask, d_cnt + i, s);
      if (rc) {
        device_free(d_cnt, s);
        return rc;
      }
    }
    std::vector<unsigned long long> h(masked.size());
    hipError_t e = hipMemcpyAsync(h.data(), d_cnt, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
    device_free(d_cnt, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(SDR_E_HIP, "null count: %s", hipGetErrorString(e));
    for (size_t i = 0; i < masked.size(); ++i) cols[masked[i]].null_count = static_cast<int64_t>(h[i]);
  }
  hipError_t e = hipPeekAtLastError();
  if (e != hipSuccess) return fail(SDR_E_HIP, "sticky launch error: %s", hipGetErrorString(e));
  return SDR_OK;
}

int decode_one_run(uint32_t codec, const sdr_run* run, void* d_out, uint32_t ts, hipStream_t s)
{
  clear_error();
  if (!run) return fail(SDR_E_INVALID_ARG, "run is NULL");
  if (run->codec != codec) return fail(SDR_E_INVALID_ARG, "run codec %u does not match entry point %s", run->codec, sdr_codec_name(codec));
  if (!type_ok_for(codec, ts)) return fail(SDR_E_BAD_WIDTH, "viability invariant violated: %s type_size %u", sdr_codec_name(codec), ts);
  if (run->n_segments && !run->segments) return fail(SDR_E_INVALID_ARG, "run has a NULL segment array");
  bool any = false;
  for (uint32_t k = 0; k < run->n_segments; ++k) {
    if (!run->segments[k].row_count) continue;
    if (!run->segments[k].d_bytes) return fail(SDR_E_INVALID_ARG, "segment %u has NULL bytes", k);
    any = true;
  }
  if (any && !d_out) return fail(SDR_E_INVALID_ARG, "d_out is NULL");
  std::vector<RunTarget> t{{run, d_out, ts}};
  return decode_runs(t, s, SDR_LAUNCH_BATCHED);
}

// ---- plans ------------------------------------------------------------------

int plan_create(const sdr_column* cols, uint32_t n_cols, uint32_t mode, hipStream_t s, sdr_plan** out)
{
  clear_error();
  if (!out) return fail(SDR_E_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (mode != SDR_LAUNCH_BATCHED && mode != SDR_LAUNCH_PER_SEGMENT)
    return fail(SDR_E_INVALID_ARG, "unknown launch mode %u", mode);
  int rc = validate_fixed_columns(cols, n_cols);
  if (rc) return rc;
  auto* p = new sdr_plan();
  p->mode = mode;
  std::vector<RunTarget> targets;
  for (uint32_t c = 0; c < n_cols; ++c) {
    for (uint32_t r = 0; r < cols[c].n_data; ++r) targets.push_back({&cols[c].data[r], cols[c].d_out, cols[c].type_size});
    const bool m = cols[c].has_nulls && cols[c].n_validity && cols[c].total_rows;
    p->col_masked.push_back(m);
    if (m) p->masked.push_back(cols[c]);
  }
  p->n_cols = n_cols;
  build_batches(targets, p->batches, p->all);
  hipError_t e = hipSuccess;
  if (!p->all.empty()) e = hipMalloc(reinterpret_cast<void**>(&p->d_segs), p->all.size() * sizeof(FixedSeg));
  if (e == hipSuccess && !p->masked.empty())
    e = hipMalloc(reinterpret_cast<void**>(&p->d_counts), p->masked.size() * sizeof(unsigned long long));
  if (e == hipSuccess && p->d_segs)
    e = hipMemcpyAsync(p->d_segs, p->all.data(), p->all.size() * sizeof(FixedSeg), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete p;
    return fail(SDR_E_HIP, "plan setup: %s", hipGetErrorString(e));
  }
  *out = p;
  return SDR_OK;
}

int plan_launch(sdr_plan* p, uint32_t codec, hipStream_t s)
{
  clear_error();
  if (!p) return fail(SDR_E_INVALID_ARG, "plan is NULL");
  hipError_t e = launch_batches(p->batches, p->all, p->d_segs, p->mode, codec, s, nullptr);
  if (e != hipSuccess) return fail(SDR_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  if (codec == 0 && !p->masked.empty()) {
    HIP_TRY(launch_fill_u8(reinterpret_cast<uint8_t*>(p->d_counts), 0, p->masked.size() * sizeof(unsigned long long), s));
    for (size_t i = 0; i < p->masked.size(); ++i) {
      const sdr_column& col = p->masked[i];
      int rc = build_null_mask(col.validity, col.n_validity, col.total_rows, col.d_null_mask, p->d_counts + i, s);
      if (rc) return rc;
    }
  }
  return SDR_OK;
}

}  // namespace sdr

sdr_plan::~sdr_plan()
{
  if (d_segs) (void)hipFree(d_segs);
  if (d_counts) (void)hipFree(d_counts);
}

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
using namespace sdr;

extern "C" {

const char* sdr_last_error(void) { return last_error(); }
int sdr_abi_version(void) { return SDR_ABI_VERSION; }
const char* sdr_build_target(void) { return "gfx950"; }

static const struct {
  const char* name;
  uint32_t id;
} k_codec_names[] = {{"Uncompressed", SDR_CODEC_UNCOMPRESSED}, {"Constant", SDR_CODEC_CONSTANT},
                     {"RLE", SDR_CODEC_RLE},                   {"Dictionary", SDR_CODEC_DICTIONARY},
                     {"BitPacking", SDR_CODEC_BITPACKING},     {"FSST", SDR_CODEC_FSST},
                     {"ALP", SDR_CODEC_ALP},                   {"ALPRD", SDR_CODEC_ALPRD},
                     {"DICT_FSST", SDR_CODEC_DICT_FSST}};

int sdr_codec_from_name(const char* name)
{
  if (!name) return SDR_E_INVALID_ARG;
  for (auto& c : k_codec_names)
    if (std::strcmp(c.name, name) == 0) return static_cast<int>(c.id);
  return SDR_E_UNSUPPORTED_CODEC;
}

const char* sdr_codec_name(uint32_t codec)
{
  for (auto& c : k_codec_names)
    if (c.id == codec) return c.name;
  return "";
}

int sdr_decode_table(sdr_column* cols, uint32_t n_cols, sdr_stream_t stream)
{
  return decode_table(cols, n_cols, static_cast<hipStream_t>(stream), SDR_LAUNCH_BATCHED);
}

int sdr_decode_table_mode(sdr_column* cols, uint32_t n_cols, sdr_stream_t stream, uint32_t launch_mode)
{
  return decode_table(cols, n_cols, static_cast<hipStream_t>(stream), launch_mode);
}

int sdr_plan_create(const sdr_column* cols, uint32_t n_cols, uint32_t launch_mode, sdr_stream_t stream,
                    sdr_plan** out)
{
  return plan_create(cols, n_cols, launch_mode, static_cast<hipStream_t>(stream), out);
}

int sdr_plan_launch(sdr_plan* plan, sdr_stream_t stream) { return plan_launch(plan, 0, static_cast<hipStream_t>(stream)); }

int sdr_plan_launch_codec(sdr_plan* plan, uint32_t codec, sdr_stream_t stream)
{
  return plan_launch(plan, codec, static_cast<hipStream_t>(stream));
}

int sdr_plan_num_launches(const sdr_plan* plan)
{
  if (!plan) return SDR_E_INVALID_ARG;
  int n = 0;
  for (const auto& b : plan->batches) n += b.count ? (plan->mode == SDR_LAUNCH_PER_SEGMENT ? static_cast<int>(b.count) : 1) : 0;
  return n;
}

int sdr_plan_null_counts(sdr_plan* plan, int64_t* counts, uint32_t n_cols, sdr_stream_t stream)
{
  clear_error();
  if (!plan || (n_cols && !counts)) return fail(SDR_E_INVALID_ARG, "NULL argument");
  if (n_cols < plan->n_cols) return fail(SDR_E_INVALID_ARG, "counts holds %u entries, plan has %u columns", n_cols, plan->n_cols);
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<unsigned long long> h(plan->masked.size());
  if (!h.empty()) {
    HIP_TRY(hipMemcpyAsync(h.data(), plan->d_counts, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  for (uint32_t c = 0; c < plan->n_cols; ++c) counts[c] = 0;
  // masked columns are recorded in column order
  size_t mi = 0;
  for (uint32_t c = 0; c < plan->n_cols && mi < plan->masked.size(); ++c)
    if (plan->col_masked.size() > c && plan->col_masked[c]) counts[c] = static_cast<int64_t>(h[mi++]);
  return SDR_OK;
}

int sdr_plan_destroy(sdr_plan* plan)
{
  delete plan;
  return SDR_OK;
}

int sdr_rowids(const sdr_row_range* ranges, uint32_t n, int64_t* d_out, sdr_stream_t stream)
{
  clear_error();
  if (n == 0) return SDR_OK;
  if (!ranges) return fail(SDR_E_INVALID_ARG, "ranges is NULL");
  static_assert(sizeof(sdr_row_range) == sizeof(RowRange), "sdr_row_range layout");
  uint32_t max_count = 0;
  for (uint32_t i = 0; i < n; ++i) max_count = ranges[i].count > max_count ? ranges[i].count : max_count;
  if (max_count && !d_out) return fail(SDR_E_INVALID_ARG, "d_out is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* d       = nullptr;
  void* token   = nullptr;
  int rc        = desc_acquire(ranges, sizeof(sdr_row_range) * n, s, &d, &token);
  if (rc) return rc;
  const hipError_t e = launch_rowids(static_cast<const RowRange*>(d), n, max_count, d_out, s);
  rc                 = desc_release(token, s);
  if (e != hipSuccess) return fail(SDR_E_HIP, "row id launch: %s", hipGetErrorString(e));
  return rc;
}

int sdr_decode_bitpacking_data(const sdr_run* run, void* d_out, uint32_t ts, sdr_stream_t stream)
{
  return decode_one_run(SDR_CODEC_BITPACKING, run, d_out, ts, static_cast<hipStream_t>(stream));
}
int sdr_decode_rle_data(const sdr_run* run, void* d_out, uint32_t ts, sdr_stream_t stream)
{
  return decode_one_run(SDR_CODEC_RLE, run, d_out, ts, static_cast<hipStream_t>(stream));
}
int sdr_decode_alp_data(const sdr_run* run, void* d_out, uint32_t ts, sdr_stream_t stream)
{
  return decode_one_run(SDR_CODEC_ALP, run, d_out, ts, static_cast<hipStream_t>(stream));
}
int sdr_decode_alprd_data(const sdr_run* run, void* d_out, uint32_t ts, sdr_stream_t stream)
{
  return decode_one_run(SDR_CODEC_ALPRD, run, d_out, ts, static_cast<hipStream_t>(stream));
}

}  // extern "C"
