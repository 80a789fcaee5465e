// sdr_strings.cpp — C ABI of the VARCHAR decode path (FSST, DICTIONARY and
// UNCOMPRESSED segments, in any mix).
//
// Mirrors gpu_decode_strings_column (src/cuda/scan/gpu_decode_strings.cu:
// 100-402): per-run validation, one column-wide int32 offset space, chars
// sized by the caller, validity overlay.  The reference's length pass, CUB
// scan and gather pass are one kernel here (k_fsst.hip), after a segment
// pre-pass that imports symbol tables and bounds every chunk.
#include "sdr_fsst.hpp"
#include "sdr_internal.hpp"

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

// Device memory of the allocation-free string decode
// (sdr_decode_strings_column_noalloc): a fixed array in the code object, one
// copy per GPU, held until the process ends.  Host code never uses its name
// as an address: the per-device address comes from hipGetSymbolAddress.
static __device__ __align__(256) uint8_t scratch[SDR_NOALLOC_SCRATCH_BYTES];

namespace sdr {

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[SCRATCH_DEVICES] = {};

// This device's copy of `scratch`, looked up once per device (with that
// device current).
int scratch_address(uint8_t** out)
{
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_DEVICES) return fail(SDR_E_HIP, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SDR_OK;
}
}  // namespace

int validate_string_column(const sdr_string_column* col)
{
  if (!col) return fail(SDR_E_INVALID_ARG, "col is NULL");
  if (col->total_rows > 0x7FFFFFFFu) return fail(SDR_E_BOUNDS, "total_rows (%u) > INT32_MAX", col->total_rows);
  int rc = check_runs(col->data, col->n_data, col->total_rows, "data");
  if (rc) return rc;
  for (uint32_t r = 0; r < col->n_data; ++r) {
    const uint32_t k = col->data[r].codec;
    if (k != SDR_CODEC_FSST && k != SDR_CODEC_DICTIONARY && k != SDR_CODEC_UNCOMPRESSED)
      return fail(SDR_E_UNSUPPORTED_CODEC, "viability invariant violated: data codec %u (%s) not implemented", k,
                  sdr_codec_name(k));
  }
  if (col->has_nulls && col->n_validity) {
    rc = check_validity_runs(col->validity, col->n_validity, col->total_rows);
    if (rc) return rc;
    if (col->total_rows && !col->d_null_mask) return fail(SDR_E_INVALID_ARG, "d_null_mask is NULL");
  }
  if (col->total_rows && !col->d_offsets) return fail(SDR_E_INVALID_ARG, "d_offsets is NULL");
  // segments may come in any order but must not share rows
  std::vector<std::pair<uint32_t, uint32_t>> spans;  // row_offset, row_count
  for (uint32_t r = 0; r < col->n_data; ++r)
    for (uint32_t k = 0; k < col->data[r].n_segments; ++k)
      if (col->data[r].segments[k].row_count)
        spans.push_back({col->data[r].segments[k].row_offset, col->data[r].segments[k].row_count});
  std::sort(spans.begin(), spans.end());
  for (size_t i = 1; i < spans.size(); ++i)
    if (spans[i].first < spans[i - 1].first + spans[i - 1].second)
      return fail(SDR_E_BOUNDS, "string segments overlap at row %u", spans[i].first);
  return SDR_OK;
}

// `scratch_mem`: null to take the call's device memory from the
// stream-ordered pool, else the per-device scratch array (capacity
// SDR_NOALLOC_SCRATCH_BYTES), in which case nothing is allocated: the range
// descriptors go up with a plain copy and every clear is a fill kernel.
int decode_strings(sdr_string_column* col, hipStream_t s, uint32_t mode, uint8_t* scratch_mem = nullptr)
{
  clear_error();
  if (mode != SDR_LAUNCH_BATCHED && mode != SDR_LAUNCH_PER_SEGMENT && mode != SDR_LAUNCH_PERSISTENT)
    return fail(SDR_E_INVALID_ARG, "unknown launch mode %u", mode);
  int rc = validate_string_column(col);
  if (rc) return rc;
  if (mode == SDR_LAUNCH_PERSISTENT && col->total_rows) return decode_strings_persistent(col, s);
  col->chars_used = 0;
  col->null_count = 0;
  const uint32_t n = col->total_rows;
  if (n == 0) {
    if (col->d_offsets) HIP_TRY(launch_fill_u8(reinterpret_cast<uint8_t*>(col->d_offsets), 0, sizeof(int32_t), s));
    return SDR_OK;
  }

  // segments in row order; rows they do not cover become empty strings
  std::vector<std::pair<const sdr_segment*, uint32_t>> segs;  // segment, codec
  for (uint32_t r = 0; r < col->n_data; ++r)
    for (uint32_t k = 0; k < col->data[r].n_segments; ++k)
      if (col->data[r].segments[k].row_count) segs.push_back({&col->data[r].segments[k], col->data[r].codec});
  std::stable_sort(segs.begin(), segs.end(),
                   [](const auto& a, const auto& b) { return a.first->row_offset < b.first->row_offset; });

  // row ranges: the segments and the gaps between them; the pre-pass
  // writes their chunk lists on the device
  std::vector<FsstSeg> hsegs;
  uint32_t n_chunks = 0;
  auto add_range = [&](const uint8_t* bytes, uint32_t size, uint32_t col_row, uint32_t rows, uint32_t codec) {
    hsegs.push_back({bytes, size, rows, n_chunks, col_row, codec, 0});
    n_chunks += fsst_chunk_count(rows);
  };
  uint32_t cursor = 0;
  for (const auto& sc : segs) {
    const sdr_segment* g = sc.first;
    if (g->row_offset > cursor) add_range(nullptr, 0, cursor, g->row_offset - cursor, 0);
    add_range(g->d_bytes, g->bytes_size, g->row_offset, g->row_count, sc.second);
    cursor = g->row_offset + g->row_count;
  }
  if (cursor < n) add_range(nullptr, 0, cursor, n - cursor, 0);
  const uint32_t n_ranges = static_cast<uint32_t>(hsegs.size());

  auto al = [](size_t v) { return (v + 255u) & ~size_t{255}; };
  const size_t seg_bytes   = al(std::max<size_t>(1, hsegs.size()) * sizeof(FsstSeg));
  const size_t chunk_bytes = al(size_t{n_chunks} * sizeof(FsstChunk));
  const size_t info_bytes  = al(size_t{n_chunks} * sizeof(FsstChunkInfo));
  const size_t tab_bytes   = al(std::max<size_t>(1, hsegs.size()) * sizeof(FsstSymTab));
  const size_t state_bytes = al(sizeof(FsstState) + sizeof(uint64_t) * n_chunks);
  const size_t total_bytes = seg_bytes + chunk_bytes + info_bytes + tab_bytes + state_bytes + sizeof(unsigned long long);
  uint8_t* d = nullptr;
  if (scratch_mem) {
    if (total_bytes > SDR_NOALLOC_SCRATCH_BYTES)
      return fail(SDR_E_CAPACITY, "string decode needs %llu bytes of descriptors; the no-allocation scratch holds %u",
                  static_cast<unsigned long long>(total_bytes), SDR_NOALLOC_SCRATCH_BYTES);
    d = scratch_mem;
  } else {
    rc = device_alloc(reinterpret_cast<void**>(&d), total_bytes, s);
    if (rc) return rc;
  }
  auto release = [&]() {
    if (!scratch_mem) device_free(d, s);
  };
  auto* d_segs   = reinterpret_cast<FsstSeg*>(d);
  auto* d_chunks = reinterpret_cast<FsstChunk*>(d + seg_bytes);
  auto* d_info   = reinterpret_cast<FsstChunkInfo*>(d + seg_bytes + chunk_bytes);
  auto* d_tabs   = reinterpret_cast<FsstSymTab*>(d + seg_bytes + chunk_bytes + info_bytes);
  uint8_t* d_rest = d + seg_bytes + chunk_bytes + info_bytes + tab_bytes;
  auto* d_state  = reinterpret_cast<FsstState*>(d_rest);
  auto* d_nulls  = reinterpret_cast<unsigned long long*>(d_rest + state_bytes);

  hipError_t e = hipSuccess;
  if (scratch_mem) {
    e = hipMemcpyAsync(d_segs, hsegs.data(), hsegs.size() * sizeof(FsstSeg), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(SDR_E_HIP, "string decode: %s", hipGetErrorString(e));
  } else {
    rc = upload_async(d_segs, hsegs.data(), hsegs.size() * sizeof(FsstSeg), s);
    if (rc) {
      release();
      return rc;
    }
  }
  e = launch_fsst_reset(d_state, n_chunks, s);
  if (e == hipSuccess)
    e = launch_fsst_prep(d_segs, n_ranges, nullptr, std::min<uint32_t>(n_ranges, 4096u), d_chunks, d_info, d_tabs, s);
  if (e == hipSuccess) {
    if (mode == SDR_LAUNCH_PER_SEGMENT) {
      for (uint32_t i = 0; e == hipSuccess && i < n_ranges; ++i)
        e = launch_fsst(d_chunks, hsegs[i].first_chunk, fsst_chunk_count(hsegs[i].row_count), n_chunks, d_info,
                        d_tabs, d_state, col->d_offsets, col->d_chars, col->d_chars ? col->chars_capacity : 0, n, s);
    } else {
      e = launch_fsst(d_chunks, 0, n_chunks, n_chunks, d_info, d_tabs, d_state, col->d_offsets, col->d_chars,
                      col->d_chars ? col->chars_capacity : 0, n, s);
    }
  }
  const bool masked = col->has_nulls && col->d_null_mask;
  if (e == hipSuccess && masked) {
    e = launch_fill_u8(reinterpret_cast<uint8_t*>(d_nulls), 0, sizeof(unsigned long long), s);
    if (e == hipSuccess) {
      rc = build_null_mask(col->validity, col->n_validity, n, col->d_null_mask, d_nulls, s);
      if (rc) {
        release();
        return rc;
      }
    }
  }
  struct {
    uint32_t ticket, error;
    uint64_t total;
  } h{};
  unsigned long long h_nulls = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d_state, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && masked) e = hipMemcpyAsync(&h_nulls, d_nulls, sizeof(h_nulls), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  release();
  if (e != hipSuccess) return fail(SDR_E_HIP, "string decode: %s", hipGetErrorString(e));
  if (h.error) return fail(SDR_E_HIP, "string decode: offset look-back did not complete");
  col->chars_used = h.total;
  col->null_count = masked ? static_cast<int64_t>(h_nulls) : 0;
  if (h.total > 0x7FFFFFFFull)
    return fail(SDR_E_OVERFLOW, "total_chars (%llu) exceeds int32 max", static_cast<unsigned long long>(h.total));
  if (h.total > col->chars_capacity || (h.total && !col->d_chars))
    return fail(SDR_E_OVERFLOW, "total_chars (%llu) exceeds chars_capacity (%llu)",
                static_cast<unsigned long long>(h.total), static_cast<unsigned long long>(col->chars_capacity));
  return SDR_OK;
}

}  // namespace sdr

using namespace sdr;

extern "C" {

uint64_t sdr_strings_upper_bound(const sdr_run* runs, uint32_t n_runs)
{
  uint64_t total = 0;
  for (uint32_t r = 0; runs && r < n_runs; ++r)
    for (uint32_t k = 0; runs[r].segments && k < runs[r].n_segments; ++k) {
      const sdr_segment& g = runs[r].segments[k];
      if (!g.row_count) continue;
      if (g.max_string_length == 0) return UINT64_MAX;
      total += uint64_t{g.row_count} * g.max_string_length;
    }
  return total;
}

int sdr_decode_strings_column_mode(sdr_string_column* col, sdr_stream_t stream, uint32_t launch_mode)
{
  return decode_strings(col, static_cast<hipStream_t>(stream), launch_mode);
}

int sdr_decode_strings_column(sdr_string_column* col, sdr_stream_t stream)
{
  return decode_strings(col, static_cast<hipStream_t>(stream), SDR_LAUNCH_BATCHED);
}

int sdr_decode_strings_column_noalloc(sdr_string_column* col, sdr_stream_t stream)
{
  clear_error();
  uint8_t* mem = nullptr;
  const int rc = scratch_address(&mem);
  if (rc) return rc;
  return decode_strings(col, static_cast<hipStream_t>(stream), SDR_LAUNCH_BATCHED, mem);
}

}  // extern "C"
