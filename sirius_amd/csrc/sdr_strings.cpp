// sdr_strings.cpp — C ABI of the VARCHAR decode path (FSST, DICTIONARY,
// UNCOMPRESSED and DICT_FSST segments, in any mix).
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
    if (k != SDR_CODEC_FSST && k != SDR_CODEC_DICTIONARY && k != SDR_CODEC_UNCOMPRESSED && k != SDR_CODEC_DICT_FSST)
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

bool has_dict_fsst(const sdr_string_column* col)
{
  for (uint32_t r = 0; r < col->n_data; ++r)
    if (col->data[r].codec == SDR_CODEC_DICT_FSST && col->data[r].n_segments) return true;
  return false;
}

namespace {

// Device memory of a column's DICT_FSST segments, released with the call's.
struct DictFsstCall {
  uint8_t* mem = nullptr;  // descriptors, range map, entry offsets, results
  uint8_t* pre = nullptr;  // decoded FSST entries
  DictFsstSeg* d_desc = nullptr;
  uint32_t* d_range = nullptr;
  bool any_inline = false;
  void release(hipStream_t s)
  {
    device_free(mem, s);
    device_free(pre, s);
    mem = pre = nullptr;
  }
};

uint32_t al8(uint32_t v) { return (v + 7u) & ~7u; }

// The DICT_FSST part of a column decode (strings/dict_fsst.cu:471-690, two
// host round trips as there): read the headers of the DICT_FSST ranges,
// fill their descriptors (region offsets, validity, :516-525), run the
// dictionary pre-pass, read back each segment's decoded total and inline-
// NULL flag, predecode the FSST entries, then point the ranges at their
// descriptors (bytes) with their validity (pad) and upload them again.
int dict_fsst_phase(const sdr_string_column* col, std::vector<FsstSeg>& hsegs, FsstSeg* d_segs, FsstSymTab* d_tabs,
                    hipStream_t s, DictFsstCall& dc)
{
  const uint32_t n_ranges = static_cast<uint32_t>(hsegs.size());
  // 1. headers
  uint32_t* d_hdr = nullptr;
  int rc          = device_alloc(reinterpret_cast<void**>(&d_hdr), size_t{n_ranges} * 16u, s);
  if (rc) return rc;
  std::vector<uint32_t> hdr(size_t{n_ranges} * 4u);
  hipError_t e = launch_dfsst_headers(d_segs, n_ranges, d_hdr, s);
  if (e == hipSuccess) e = hipMemcpyAsync(hdr.data(), d_hdr, hdr.size() * 4u, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  device_free(d_hdr, s);
  if (e != hipSuccess) return fail(SDR_E_HIP, "DICT_FSST headers: %s", hipGetErrorString(e));
  // 2. descriptors
  std::vector<DictFsstSeg> ds;
  std::vector<uint32_t> range_d(n_ranges, ~0u);
  uint64_t n_e = 0;
  for (uint32_t i = 0; i < n_ranges; ++i) {
    const FsstSeg& sg = hsegs[i];
    if (sg.codec != SDR_CODEC_DICT_FSST || !sg.bytes) continue;
    const uint32_t* h = &hdr[4u * i];
    DictFsstSeg d{};
    d.bytes      = sg.bytes;
    d.bytes_size = sg.bytes_size;
    d.range      = i;
    const uint32_t dict_size = h[0], dict_count = h[1], symtab_size = h[3];
    const uint32_t mode = h[2] & 0xFFu, slw = (h[2] >> 8) & 0xFFu, diw = (h[2] >> 16) & 0xFFu;
    d.off_dict   = 16u;
    d.off_symtab = al8(d.off_dict + dict_size);
    d.off_slens  = mode == 0u ? d.off_dict + al8(dict_size) : al8(d.off_symtab + symtab_size);
    d.off_didx   = al8(d.off_slens + (dict_count * slw + 7u) / 8u);
    d.valid      = sg.bytes_size >= 16u && mode <= 2u && slw <= 32u && diw <= 32u &&
              (d.off_didx <= sg.bytes_size || mode == 2u);
    d.mode       = static_cast<uint8_t>(mode);
    d.slw        = static_cast<uint8_t>(slw);
    d.diw        = static_cast<uint8_t>(diw);
    d.dict_count = d.valid ? dict_count : 0u;
    d.e_byte     = static_cast<uint32_t>(n_e);
    if (d.valid) n_e += uint64_t{dict_count} + 1u;
    if (n_e > 0xFFFFFFFFull) return fail(SDR_E_CAPACITY, "DICT_FSST dictionaries hold more than 2^32 entries");
    range_d[i] = static_cast<uint32_t>(ds.size());
    ds.push_back(d);
  }
  const uint32_t n_d = static_cast<uint32_t>(ds.size());
  if (n_d == 0) return SDR_OK;
  auto al = [](size_t v) { return (v + 255u) & ~size_t{255}; };
  const size_t b_desc = al(n_d * sizeof(DictFsstSeg)), b_range = al(size_t{n_ranges} * 4u);
  const size_t b_e = al(std::max<uint64_t>(1, n_e) * 4u), b_out = al(size_t{n_d} * 8u);
  rc = device_alloc(reinterpret_cast<void**>(&dc.mem), b_desc + b_range + 2u * b_e + b_out, s);
  if (rc) return rc;
  dc.d_desc       = reinterpret_cast<DictFsstSeg*>(dc.mem);
  dc.d_range      = reinterpret_cast<uint32_t*>(dc.mem + b_desc);
  auto* d_e_byte  = reinterpret_cast<uint32_t*>(dc.mem + b_desc + b_range);
  auto* d_e_dec   = reinterpret_cast<uint32_t*>(dc.mem + b_desc + b_range + b_e);
  auto* d_out     = reinterpret_cast<uint32_t*>(dc.mem + b_desc + b_range + 2u * b_e);
  for (auto& d : ds) {
    d.e = d_e_dec + d.e_byte;
    if (d.mode == 0u || d.dict_count <= 1u) {  // rows copy raw dict bytes
      d.src      = d.bytes + d.off_dict;
      d.src_size = d.bytes_size > d.off_dict ? d.bytes_size - d.off_dict : 0u;
    }
  }
  rc = upload_async(dc.d_desc, ds.data(), n_d * sizeof(DictFsstSeg), s);
  if (!rc) rc = upload_async(dc.d_range, range_d.data(), size_t{n_ranges} * 4u, s);
  if (rc) return rc;
  // 3. dictionary pre-pass, per-segment totals and inline-NULL flags
  std::vector<uint32_t> out(size_t{n_d} * 2u);
  e = launch_dfsst_prep(dc.d_desc, n_d, d_e_byte, d_e_dec, d_tabs, d_out, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out.data(), d_out, out.size() * 4u, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(SDR_E_HIP, "DICT_FSST dictionary pre-pass: %s", hipGetErrorString(e));
  uint64_t n_pre = 0;
  for (uint32_t j = 0; j < n_d; ++j) {
    dc.any_inline = dc.any_inline || out[2u * j + 1u];
    if (ds[j].valid && ds[j].mode != 0u && ds[j].dict_count > 1u) n_pre += out[2u * j];
  }
  if (dc.any_inline && col->total_rows && !col->d_null_mask)
    return fail(SDR_E_INVALID_ARG, "DICT_FSST segments hold inline NULLs (entry 0) and d_null_mask is NULL");
  // 4. predecode
  if (n_pre) {
    rc = device_alloc(reinterpret_cast<void**>(&dc.pre), n_pre, s);
    if (rc) return rc;
    uint64_t at = 0;
    for (uint32_t j = 0; j < n_d; ++j) {
      auto& d = ds[j];
      if (!d.valid || d.mode == 0u || d.dict_count <= 1u) continue;
      d.src      = dc.pre + at;
      d.src_size = out[2u * j];
      at += out[2u * j];
    }
    rc = upload_async(dc.d_desc, ds.data(), n_d * sizeof(DictFsstSeg), s);
    if (rc) return rc;
    e = launch_dfsst_predecode(dc.d_desc, n_d, d_e_byte, d_tabs, s);
    if (e != hipSuccess) return fail(SDR_E_HIP, "DICT_FSST predecode: %s", hipGetErrorString(e));
  }
  // 5. the ranges now name their descriptors
  for (uint32_t i = 0; i < n_ranges; ++i)
    if (range_d[i] != ~0u) {
      hsegs[i].bytes = reinterpret_cast<const uint8_t*>(dc.d_desc + range_d[i]);
      hsegs[i].pad   = ds[range_d[i]].valid;
    }
  return upload_async(d_segs, hsegs.data(), hsegs.size() * sizeof(FsstSeg), s);
}

}  // namespace

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
  const bool dfsst = has_dict_fsst(col);
  // DICT_FSST needs host round trips: it decodes through the per-column passes
  if (mode == SDR_LAUNCH_PERSISTENT && col->total_rows && !dfsst) return decode_strings_persistent(col, s);
  if (dfsst && scratch_mem)
    return fail(SDR_E_UNSUPPORTED_CODEC,
                "DICT_FSST needs device memory sized from its headers; decode it with sdr_decode_strings_column");
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
  DictFsstCall dc;
  if (dfsst) {
    rc = dict_fsst_phase(col, hsegs, d_segs, d_tabs, s, dc);
    if (rc) {
      dc.release(s);  // stream-ordered: after every pass of the phase already enqueued
      release();
      return rc;
    }
  }
  auto release_all = [&]() {
    dc.release(s);
    release();
  };
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
  const bool masked = (col->has_nulls && col->d_null_mask) || dc.any_inline;
  if (e == hipSuccess && masked) {
    e = launch_fill_u8(reinterpret_cast<uint8_t*>(d_nulls), 0, sizeof(unsigned long long), s);
    if (e == hipSuccess) {
      rc = build_null_mask(col->validity, col->has_nulls ? col->n_validity : 0u, n, col->d_null_mask, d_nulls, s);
      if (rc) {
        release_all();
        return rc;
      }
      // DICT_FSST inline NULLs (dict_fsst.cu:426-446) on top of the validity
      if (dc.any_inline)
        e = launch_dfsst_nulls(d_segs, n_ranges, dc.d_desc, dc.d_range, col->d_null_mask, n, d_nulls, s);
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
  release_all();
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
