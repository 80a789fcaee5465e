// k_fixed.hip — gfx950 kernels for the fixed-width codecs: BITPACKING, RLE,
// ALP, ALPRD, UNCOMPRESSED, CONSTANT and the validity mask helpers.  One
// launch covers a batch of segments: blockIdx.y selects the segment,
// blockIdx.x * 4 + wave the work unit (decoders in dec_fixed.hpp).
#include "dec_fixed.hpp"
#include "dec_rle.hpp"
#include "sdr_tune.hpp"

#include <algorithm>
#include <cstring>

namespace sdr {

// Where a launch finds its segment descriptors: a device array, or the
// kernel arguments themselves (at most SEGS_INLINE descriptors, 3 KiB of the
// 4 KiB kernel-argument segment), which spares a small call the descriptor
// upload.  blockIdx-indexed, so the load is uniform per workgroup.
// A descriptor read at a workgroup-uniform index, its fields made uniform
// (scalar registers), so the decoders' address arithmetic on them stays
// scalar as with a restrict kernel-argument pointer.
__device__ __forceinline__ FixedSeg uniform_seg(const FixedSeg& v)
{
  const uint64_t b = reinterpret_cast<uint64_t>(v.bytes), o = reinterpret_cast<uint64_t>(v.out);
  FixedSeg r;
  r.bytes = reinterpret_cast<const uint8_t*>(
    (uint64_t{static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32)))} << 32) |
    static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(b))));
  r.out = reinterpret_cast<uint8_t*>(
    (uint64_t{static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(o >> 32)))} << 32) |
    static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(o))));
  r.bytes_size = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v.bytes_size)));
  r.row_count  = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v.row_count)));
  return r;
}

struct DevSegs {
  const FixedSeg* p;
  uint32_t base;
  __device__ __forceinline__ FixedSeg at(uint32_t i) const { return uniform_seg(gload(p + base + i)); }
};
struct SegsInline {
  FixedSeg seg[SEGS_INLINE];
  __device__ __forceinline__ FixedSeg at(uint32_t i) const { return uniform_seg(seg[i]); }
};

// Wave w of block x decodes groups [(4x + w) * gpw, (4x + w + 1) * gpw).
template <typename T, class Src>
__global__ __launch_bounds__(BLOCK_THREADS) void k_bitpacking(Src src, uint32_t gpw, uint32_t chunk_rows)
{
  // 4 KiB of packed bits per wave keeps 8 waves per SIMD resident
  constexpr uint32_t SB = 4096;
  __shared__ uint32_t lds[WAVES_PER_BLOCK][(SB + 48) / 4];
  const FixedSeg sd  = src.at(blockIdx.y);
  const uint32_t w   = wave_in_block();
  const uint32_t ngr = (sd.row_count + BP_GROUP - 1) / BP_GROUP;
  const uint32_t k0  = (blockIdx.x * WAVES_PER_BLOCK + w) * gpw;
  for (uint32_t k = k0; k < k0 + gpw && k < ngr; ++k) {
    const uint32_t grc = sd.row_count - k * BP_GROUP < BP_GROUP ? sd.row_count - k * BP_GROUP : BP_GROUP;
    bp_group<T, SB>(sd.bytes, sd.bytes_size, k, grc, reinterpret_cast<T*>(sd.out) + size_t{k} * BP_GROUP,
                    lds[w], chunk_rows);
  }
}

template <typename T, bool RD, class Src>
__global__ __launch_bounds__(BLOCK_THREADS) void k_alp(Src src, uint32_t vpw)
{
  __shared__ uint32_t lds[WAVES_PER_BLOCK][RD ? ALPRD_LEFT_WORDS + ALP_CHUNK_WORDS : ALP_CHUNK_WORDS];
  const FixedSeg sd  = src.at(blockIdx.y);
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

template <uint32_t TS, class Src>
__global__ __launch_bounds__(RLE_THREADS) void k_rle(Src src, uint32_t n_segs, uint32_t slices, uint32_t slice_rows)
{
  __shared__ RleSmem sm;
  const uint32_t per_xcd = gridDim.x / N_XCD;
  const uint32_t logical = (blockIdx.x % N_XCD) * per_xcd + blockIdx.x / N_XCD;
  const uint32_t sgi     = logical / slices;
  if (sgi >= n_segs) return;
  const FixedSeg sd = src.at(sgi);
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

template <class Src>
__global__ __launch_bounds__(BLOCK_THREADS) void k_copy(Src src, uint32_t type_size)
{
  const FixedSeg sd    = src.at(blockIdx.y);
  const uint64_t total = uint64_t{sd.row_count} * type_size;
  const uint64_t b0    = uint64_t{blockIdx.x} * COPY_UNIT;
  if (b0 >= total) return;
  const uint64_t b1 = total - b0 < COPY_UNIT ? total : b0 + COPY_UNIT;
  copy_span<BLOCK_THREADS>(sd.bytes, sd.out, b0, b1, threadIdx.x);
}

template <class Src>
__global__ __launch_bounds__(BLOCK_THREADS) void k_constant(Src src, uint32_t type_size)
{
  const FixedSeg sd    = src.at(blockIdx.y);
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

// Call f(src, m) for every launch over the n descriptors at
// segs[seg_base..]: device descriptors in pieces of at most max_m, or host
// descriptors copied into the kernel arguments SEGS_INLINE at a time.
template <typename F>
hipError_t for_segs(const FixedSeg* segs, uint32_t seg_base, uint32_t n, bool host, uint32_t max_m, F f)
{
  const uint32_t lim = host ? std::min(max_m, SEGS_INLINE) : max_m;
  for (uint32_t b = 0; b < n; b += lim) {
    const uint32_t m = n - b < lim ? n - b : lim;
    if (host) {
      SegsInline si{};
      std::memcpy(si.seg, segs + seg_base + b, size_t{m} * sizeof(FixedSeg));
      f(si, m);
    } else {
      f(DevSegs{segs, seg_base + b}, m);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
}  // namespace

hipError_t launch_bitpacking(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t ts,
                             hipStream_t s, bool host)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint32_t gpw   = tune(SDR_TUNE_BP_GROUPS_PER_WAVE, ts);
  const uint32_t chunk = tune(SDR_TUNE_BP_CHUNK_ROWS, ts);
  const uint32_t units = (max_rows + BP_GROUP - 1) / BP_GROUP;
  const uint32_t gx    = (units + WAVES_PER_BLOCK * gpw - 1) / (WAVES_PER_BLOCK * gpw);
  return for_segs(segs, seg_base, n, host, MAX_GRID_Y, [&](auto src, uint32_t m) {
    using S = decltype(src);
    const dim3 g(gx, m);
    switch (ts) {
      case 1: hipLaunchKernelGGL((k_bitpacking<uint8_t, S>), g, dim3(BLOCK_THREADS), 0, s, src, gpw, chunk); break;
      case 2: hipLaunchKernelGGL((k_bitpacking<uint16_t, S>), g, dim3(BLOCK_THREADS), 0, s, src, gpw, chunk); break;
      case 4: hipLaunchKernelGGL((k_bitpacking<uint32_t, S>), g, dim3(BLOCK_THREADS), 0, s, src, gpw, chunk); break;
      default: hipLaunchKernelGGL((k_bitpacking<uint64_t, S>), g, dim3(BLOCK_THREADS), 0, s, src, gpw, chunk); break;
    }
  });
}

hipError_t launch_rle(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t ts,
                      hipStream_t s, bool host)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint32_t sr     = tune(SDR_TUNE_RLE_SLICE_ROWS, ts);
  const uint32_t slices = (max_rows + sr - 1) / sr;
  // segments per launch so that the 1-D grid stays below 2^31 blocks
  const uint32_t per = static_cast<uint32_t>(std::min<uint64_t>(MAX_GRID_Y, (0x7FFFFFF0ull / slices)));
  return for_segs(segs, seg_base, n, host, per, [&](auto src, uint32_t m) {
    using S = decltype(src);
    const uint64_t blocks = (uint64_t{m} * slices + N_XCD - 1) / N_XCD * N_XCD;
    const dim3 grid(static_cast<uint32_t>(blocks)), block(RLE_THREADS);
    switch (ts) {
      case 1: hipLaunchKernelGGL((k_rle<1, S>), grid, block, 0, s, src, m, slices, sr); break;
      case 2: hipLaunchKernelGGL((k_rle<2, S>), grid, block, 0, s, src, m, slices, sr); break;
      case 4: hipLaunchKernelGGL((k_rle<4, S>), grid, block, 0, s, src, m, slices, sr); break;
      default: hipLaunchKernelGGL((k_rle<8, S>), grid, block, 0, s, src, m, slices, sr); break;
    }
  });
}

hipError_t launch_alp(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t ts, bool rd,
                      hipStream_t s, bool host)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint32_t vpw   = tune(SDR_TUNE_ALP_VECTORS_PER_WAVE, ts);
  const uint32_t units = (max_rows + ALP_VEC - 1) / ALP_VEC;
  const uint32_t gx    = (units + WAVES_PER_BLOCK * vpw - 1) / (WAVES_PER_BLOCK * vpw);
  return for_segs(segs, seg_base, n, host, MAX_GRID_Y, [&](auto src, uint32_t m) {
    using S = decltype(src);
    const dim3 g(gx, m);
    if (ts == 8) {
      if (rd) hipLaunchKernelGGL((k_alp<double, true, S>), g, dim3(BLOCK_THREADS), 0, s, src, vpw);
      else hipLaunchKernelGGL((k_alp<double, false, S>), g, dim3(BLOCK_THREADS), 0, s, src, vpw);
    } else {
      if (rd) hipLaunchKernelGGL((k_alp<float, true, S>), g, dim3(BLOCK_THREADS), 0, s, src, vpw);
      else hipLaunchKernelGGL((k_alp<float, false, S>), g, dim3(BLOCK_THREADS), 0, s, src, vpw);
    }
  });
}

hipError_t launch_uncompressed(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t ts,
                               hipStream_t s, bool host)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint64_t gx = (uint64_t{max_rows} * ts + COPY_UNIT - 1) / COPY_UNIT;
  return for_segs(segs, seg_base, n, host, MAX_GRID_Y, [&](auto src, uint32_t m) {
    hipLaunchKernelGGL((k_copy<decltype(src)>), dim3(static_cast<uint32_t>(gx), m), dim3(BLOCK_THREADS), 0, s, src,
                       ts);
  });
}

hipError_t launch_constant(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t ts,
                           hipStream_t s, bool host)
{
  if (n == 0 || max_rows == 0) return hipSuccess;
  const uint64_t gx = (uint64_t{max_rows} * ts + COPY_UNIT - 1) / COPY_UNIT;
  return for_segs(segs, seg_base, n, host, MAX_GRID_Y, [&](auto src, uint32_t m) {
    hipLaunchKernelGGL((k_constant<decltype(src)>), dim3(static_cast<uint32_t>(gx), m), dim3(BLOCK_THREADS), 0, s,
                       src, ts);
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
