// k_fixed.hip — gfx950 kernels for the fixed-width codecs: BITPACKING, RLE,
// ALP, ALPRD, UNCOMPRESSED, CONSTANT and the validity mask helpers.  One
// launch covers a batch of segments: blockIdx.y selects the segment,
// blockIdx.x * 4 + wave the work unit (decoders in dec_fixed.hpp).
#include "dec_fixed.hpp"
#include "dec_rle.hpp"
#include "sdr_tune.hpp"

#include <algorithm>

namespace sdr {

// Wave w of block x decodes groups [(4x + w) * gpw, (4x + w + 1) * gpw).
template <typename T>
__global__ __launch_bounds__(BLOCK_THREADS) void k_bitpacking(const FixedSeg* __restrict__ segs,
                                                              uint32_t seg_base, uint32_t gpw,
                                                              uint32_t chunk_rows)
{
  // 4 KiB of packed bits per wave keeps 8 waves per SIMD resident
  constexpr uint32_t SB = 4096;
  __shared__ uint32_t lds[WAVES_PER_BLOCK][(SB + 48) / 4];
  const FixedSeg sd  = segs[seg_base + blockIdx.y];
  const uint32_t w   = wave_in_block();
  const uint32_t ngr = (sd.row_count + BP_GROUP - 1) / BP_GROUP;
  const uint32_t k0  = (blockIdx.x * WAVES_PER_BLOCK + w) * gpw;
  for (uint32_t k = k0; k < k0 + gpw && k < ngr; ++k) {
    const uint32_t grc = sd.row_count - k * BP_GROUP < BP_GROUP ? sd.row_count - k * BP_GROUP : BP_GROUP;
    bp_group<T, SB>(sd.bytes, sd.bytes_size, k, grc, reinterpret_cast<T*>(sd.out) + size_t{k} * BP_GROUP,
                    lds[w], chunk_rows);
  }
}

template <typename T, bool RD>
__global__ __launch_bounds__(BLOCK_THREADS) void k_alp(const FixedSeg* __restrict__ segs, uint32_t seg_base,
                                                       uint32_t vpw)
{
  __shared__ uint32_t lds[WAVES_PER_BLOCK][RD ? ALPRD_LEFT_WORDS + ALP_CHUNK_WORDS : ALP_CHUNK_WORDS];
  const FixedSeg sd  = segs[seg_base + blockIdx.y];
  const uint32_t w   = wave_in_block();
  const uint32_t nv  = (sd.row_count + ALP_VEC - 1) / ALP_VEC;
  const uint32_t v0  = (blockIdx.x * WAVES_PER_BLOCK + w) * vpw;
  for (uint32_t v = v0; v < v0 + vpw && v < nv; ++v) {
    const uint32_t fill = sd.row_count - v * ALP_VEC < ALP_VEC ? sd.row_count - v * ALP_VEC : ALP_VEC;
    T* out = reinterpret_cast<T*>(sd.out) + size_t{v} * ALP_VEC;
    if constexpr (RD) alprd_vector<T>(sd.bytes, sd.bytes_size, v, fill, out, lds[w]);
    else alp_vector<T>(sd.bytes, sd.bytes_size, v, fill, out, lds[w]);
  }
}

// One block per (segment, slice of `slice_rows` rows) over a 1-D grid of
// n_blocks (a multiple of 8).  Workgroups are dispatched to the 8 XCDs
// round-robin, so block b runs on XCD b % 8; the remap below hands each XCD
// a contiguous range of (segment, slice) pairs, so the slices of a segment
// share one L2 and its counts are fetched from HBM about once.
constexpr uint32_t N_XCD = 8;

template <uint32_t TS>
__global__ __launch_bounds__(RLE_THREADS) void k_rle(const FixedSeg* __restrict__ segs, uint32_t seg_base,
                                                    uint32_t n_segs, uint32_t slices, uint32_t slice_rows)
{
  __shared__ RleSmem sm;
  const uint32_t per_xcd = gridDim.x / N_XCD;
  const uint32_t logical = (blockIdx.x % N_XCD) * per_xcd + blockIdx.x / N_XCD;
  const uint32_t sgi     = logical / slices;
  if (sgi >= n_segs) return;
  const FixedSeg sd = segs[seg_base + sgi];
  const uint32_t ra = (logical % slices) * slice_rows;
  if (ra >= sd.row_count) return;
  const uint32_t rb = sd.row_count - ra < slice_rows ? sd.row_count : ra + slice_rows;
  rle_slice_t<TS>(sd.bytes, sd.bytes_size, sd.row_count, ra, rb, sd.out, sm);
}

// ===========================================================================
// UNCOMPRESSED copy and CONSTANT broadcast (thin paths; the reference's
// gpu_native_decode.cu:108-115 and :186-216).  A unit is 64 KiB of output:
// on a 256 MiB copy, 16 KiB units took 225 us, 64 KiB 186 us, 128 KiB 183 us
// and 256 KiB 311 us (too few blocks to fill the chip).
// ===========================================================================
constexpr uint32_t COPY_UNIT = 65536;

__global__ __launch_bounds__(BLOCK_THREADS) void k_copy(const FixedSeg* __restrict__ segs, uint32_t seg_base,
                                                        uint32_t type_size)
{
  const FixedSeg sd    = segs[seg_base + blockIdx.y];
  const uint64_t total = uint64_t{sd.row_count} * type_size;
  const uint64_t b0    = uint64_t{blockIdx.x} * COPY_UNIT;
  if (b0 >= total) return;
  const uint64_t b1 = total - b0 < COPY_UNIT ? total : b0 + COPY_UNIT;
  copy_span<BLOCK_THREADS>(sd.bytes, sd.out, b0, b1, threadIdx.x);
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_constant(const FixedSeg* __restrict__ segs, uint32_t seg_base,
                                                            uint32_t type_size)
{
  const FixedSeg sd    = segs[seg_base + blockIdx.y];
  const uint64_t total = uint64_t{sd.row_count} * type_size;
  const uint64_t b0    = uint64_t{blockIdx.x} * COPY_UNIT;
  if (b0 >= total) return;
  const uint64_t b1 = total - b0 < COPY_UNIT ? total : b0 + COPY_UNIT;
  const_span<BLOCK_THREADS>(sd.bytes, type_size, sd.out, b0, b1, threadIdx.x);
}

__global__ void k_fill_u8(uint8_t* p, uint8_t v, uint64_t n)
{
  for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < n; i += uint64_t{gridDim.x} * blockDim.x)
    p[i] = v;
}

__global__ void k_count_nulls(const uint8_t* mask, uint32_t rows, unsigned long long* count)
{
  uint32_t local = 0;
  const uint32_t nbytes = (rows + 7u) / 8u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nbytes; i += gridDim.x * blockDim.x) {
    uint32_t b = mask[i];
    if (i == nbytes - 1 && (rows & 7u)) b |= 0xFFu << (rows & 7u);  // bits past rows are not rows
    local += 8u - static_cast<uint32_t>(__builtin_popcount(b & 0xFFu));
  }
  // wave reduce, one atomic per wave
  local = wave_incl_sum_u32(local);
  if (lane_id() == 63 && local) atomicAdd(count, static_cast<unsigned long long>(local));
}

// ===========================================================================
// launch wrappers
// ===========================================================================
namespace {
constexpr uint32_t MAX_GRID_Y = 65535;

template <typename Launch>
hipError_t chunked(uint32_t seg_base, uint32_t n, Launch f)
{
  for (uint32_t s = 0; s < n; s += MAX_GRID_Y) {
    const uint32_t m = n - s < MAX_GRID_Y ? n - s : MAX_GRID_Y;
    f(seg_base + s, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
}  // namespace

hipError_t launch_bitpacking(const FixedSeg* d_segs, uint32_t seg_base, uint32_t n, uint32_t max_rows,
                             uint32_t ts, hipStream_t s)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint32_t gpw   = tune(SDR_TUNE_BP_GROUPS_PER_WAVE, ts);
  const uint32_t chunk = tune(SDR_TUNE_BP_CHUNK_ROWS, ts);
  const uint32_t units = (max_rows + BP_GROUP - 1) / BP_GROUP;
  const uint32_t gx    = (units + WAVES_PER_BLOCK * gpw - 1) / (WAVES_PER_BLOCK * gpw);
  return chunked(seg_base, n, [&](uint32_t b, uint32_t m) {
    dim3 g(gx, m);
    switch (ts) {
      case 1: hipLaunchKernelGGL(k_bitpacking<uint8_t>, g, dim3(BLOCK_THREADS), 0, s, d_segs, b, gpw, chunk); break;
      case 2: hipLaunchKernelGGL(k_bitpacking<uint16_t>, g, dim3(BLOCK_THREADS), 0, s, d_segs, b, gpw, chunk); break;
      case 4: hipLaunchKernelGGL(k_bitpacking<uint32_t>, g, dim3(BLOCK_THREADS), 0, s, d_segs, b, gpw, chunk); break;
      default: hipLaunchKernelGGL(k_bitpacking<uint64_t>, g, dim3(BLOCK_THREADS), 0, s, d_segs, b, gpw, chunk); break;
    }
  });
}

hipError_t launch_rle(const FixedSeg* d_segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t ts,
                      hipStream_t s)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint32_t sr     = tune(SDR_TUNE_RLE_SLICE_ROWS, ts);
  const uint32_t slices = (max_rows + sr - 1) / sr;
  // segments per launch so that the 1-D grid stays below 2^31 blocks
  const uint32_t per = static_cast<uint32_t>(std::min<uint64_t>(MAX_GRID_Y, (0x7FFFFFF0ull / slices)));
  for (uint32_t b = 0; b < n; b += per) {
    const uint32_t m      = n - b < per ? n - b : per;
    const uint64_t blocks = (uint64_t{m} * slices + N_XCD - 1) / N_XCD * N_XCD;
    const dim3 grid(static_cast<uint32_t>(blocks)), block(RLE_THREADS);
    switch (ts) {
      case 1: hipLaunchKernelGGL(k_rle<1>, grid, block, 0, s, d_segs, seg_base + b, m, slices, sr); break;
      case 2: hipLaunchKernelGGL(k_rle<2>, grid, block, 0, s, d_segs, seg_base + b, m, slices, sr); break;
      case 4: hipLaunchKernelGGL(k_rle<4>, grid, block, 0, s, d_segs, seg_base + b, m, slices, sr); break;
      default: hipLaunchKernelGGL(k_rle<8>, grid, block, 0, s, d_segs, seg_base + b, m, slices, sr); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_alp(const FixedSeg* d_segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t ts,
                      bool rd, hipStream_t s)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint32_t vpw   = tune(SDR_TUNE_ALP_VECTORS_PER_WAVE, ts);
  const uint32_t units = (max_rows + ALP_VEC - 1) / ALP_VEC;
  const uint32_t gx    = (units + WAVES_PER_BLOCK * vpw - 1) / (WAVES_PER_BLOCK * vpw);
  return chunked(seg_base, n, [&](uint32_t b, uint32_t m) {
    dim3 g(gx, m);
    if (ts == 8) {
      if (rd) hipLaunchKernelGGL((k_alp<double, true>), g, dim3(BLOCK_THREADS), 0, s, d_segs, b, vpw);
      else hipLaunchKernelGGL((k_alp<double, false>), g, dim3(BLOCK_THREADS), 0, s, d_segs, b, vpw);
    } else {
      if (rd) hipLaunchKernelGGL((k_alp<float, true>), g, dim3(BLOCK_THREADS), 0, s, d_segs, b, vpw);
      else hipLaunchKernelGGL((k_alp<float, false>), g, dim3(BLOCK_THREADS), 0, s, d_segs, b, vpw);
    }
  });
}

hipError_t launch_uncompressed(const FixedSeg* d_segs, uint32_t seg_base, uint32_t n, uint32_t max_rows,
                               uint32_t ts, hipStream_t s)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint64_t gx = (uint64_t{max_rows} * ts + COPY_UNIT - 1) / COPY_UNIT;
  return chunked(seg_base, n, [&](uint32_t b, uint32_t m) {
    hipLaunchKernelGGL(k_copy, dim3(static_cast<uint32_t>(gx), m), dim3(BLOCK_THREADS), 0, s, d_segs, b, ts);
  });
}

hipError_t launch_constant(const FixedSeg* d_segs, uint32_t seg_base, uint32_t n, uint32_t max_rows,
                           uint32_t ts, hipStream_t s)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint64_t gx = (uint64_t{max_rows} * ts + COPY_UNIT - 1) / COPY_UNIT;
  return chunked(seg_base, n, [&](uint32_t b, uint32_t m) {
    hipLaunchKernelGGL(k_constant, dim3(static_cast<uint32_t>(gx), m), dim3(BLOCK_THREADS), 0, s, d_segs, b, ts);
  });
}

// One y-block row per range (the reference's cudf::sequence per row group,
// duckdb_native_decoder.cpp:924-949, without the concatenation).
__global__ __launch_bounds__(BLOCK_THREADS) void k_rowids(const RowRange* __restrict__ ranges, int64_t* __restrict__ out)
{
  const RowRange r = gload(ranges + blockIdx.y);
  const gptr<int64_t> o = gbl(out + r.out_offset);
  for (uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x; i < r.count; i += gridDim.x * BLOCK_THREADS)
    o[i] = r.first_row + static_cast<int64_t>(i);
}

hipError_t launch_rowids(const RowRange* d_ranges, uint32_t n, uint32_t max_count, int64_t* out, hipStream_t s)
{
  if (n == 0 || max_count == 0) return hipSuccess;
  uint32_t gx = (max_count + BLOCK_THREADS - 1) / BLOCK_THREADS;
  if (gx > 1024) gx = 1024;
  for (uint32_t b = 0; b < n; b += 65535) {  // grid y limit
    const uint32_t m = n - b < 65535 ? n - b : 65535;
    hipLaunchKernelGGL(k_rowids, dim3(gx, m), dim3(BLOCK_THREADS), 0, s, d_ranges + b, out);
  }
  return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, uint8_t v, uint64_t n, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u8, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_count_nulls(const uint8_t* mask, uint32_t rows, unsigned long long* d_count, hipStream_t s)
{
  if (rows == 0) return hipSuccess;
  uint32_t nbytes = (rows + 7u) / 8u;
  uint32_t blocks = (nbytes + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_count_nulls, dim3(blocks), dim3(256), 0, s, mask, rows, d_count);
  return hipGetLastError();
}

}  // namespace sdr
