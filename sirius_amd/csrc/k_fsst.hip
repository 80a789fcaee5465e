// k_fsst.hip — per-column FSST decode kernel (one launch decodes every chunk
// of a VARCHAR column; see dec_fsst.hpp for the chunk decoder).
#include "dec_fsst.hpp"
#include "sdr_tune.hpp"

namespace sdr {

__global__ __launch_bounds__(FSST_THREADS) __attribute__((amdgpu_waves_per_eu(6))) void k_fsst(const FsstChunk* __restrict__ chunks, uint32_t n_chunks,
                                                       const FsstChunkInfo* __restrict__ info,
                                                       const FsstSymTab* __restrict__ tabs, FsstState* __restrict__ st,
                                                       int32_t* __restrict__ offsets, uint8_t* __restrict__ chars,
                                                       uint64_t chars_cap, uint32_t total_rows, uint32_t spin_limit)
{
  __shared__ FsstSmemT<FSST_OUT_COLUMN> sm;
  __shared__ uint32_t s_ticket;
  if (threadIdx.x == 0) s_ticket = atomicAdd(&st->ticket, 1u);
  __syncthreads();
  const uint32_t c = s_ticket;
  if (c >= n_chunks) return;  // grid size equals chunk count; never taken
  const FsstChunk ch = chunks[c];
  fsst_chunk<FSST_OUT_COLUMN>(ch, info[c], tabs + (ch.seg == FSST_GAP ? 0u : ch.seg), c, n_chunks, st, offsets,
                              chars, chars_cap, total_rows, spin_limit, sm);
}

// Segment pre-pass: one workgroup per segment (grid-stride).
__global__ __launch_bounds__(FSST_THREADS) void k_fsst_prep(const FsstSeg* __restrict__ segs, uint32_t n_segs,
                                                            const uint32_t* __restrict__ d_n_segs,
                                                            FsstChunk* __restrict__ chunks,
                                                            FsstChunkInfo* __restrict__ info,
                                                            FsstSymTab* __restrict__ tabs)
{
  __shared__ FsstPrepSmem sm;
  const uint32_t n = d_n_segs ? *d_n_segs : n_segs;
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) fsst_prep_segment(segs[i], i, chunks, info, tabs + i, sm);
}

__global__ void k_fsst_reset(FsstState* st, uint32_t n_chunks)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_chunks; i += gridDim.x * blockDim.x)
    st->status[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->ticket = 0;
    st->error  = 0;
    st->total  = 0;
  }
}

// ---- DICT_FSST passes ----

__global__ void k_dfsst_headers(const FsstSeg* __restrict__ segs, uint32_t n, uint32_t* __restrict__ hdr)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const FsstSeg sg = gload(segs + i);
    const bool has   = sg.codec == SDR_CODEC_DICT_FSST && sg.bytes && sg.bytes_size >= 16u;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) hdr[4u * i + w] = has ? ld_u32_any(sg.bytes + 4u * w) : 0u;
  }
}

__global__ __launch_bounds__(FSST_THREADS) void k_dfsst_prep(const DictFsstSeg* __restrict__ dseg, uint32_t n,
                                                             uint32_t* __restrict__ e_byte,
                                                             uint32_t* __restrict__ e_dec,
                                                             FsstSymTab* __restrict__ tabs, uint32_t* __restrict__ out)
{
  __shared__ FsstPrepSmem sm;
  for (uint32_t j = blockIdx.x; j < n; j += gridDim.x) {
    const DictFsstSeg d = gload(dseg + j);
    dfsst_prep_segment(d, e_byte, e_dec, tabs + d.range, out + 2u * j, sm);
    __syncthreads();  // sm is reused by the next segment
  }
}

// grid.y: segments; grid.x blocks share a segment's entries
__global__ __launch_bounds__(FSST_THREADS) void k_dfsst_predecode(const DictFsstSeg* __restrict__ dseg,
                                                                  const uint32_t* __restrict__ e_byte,
                                                                  const FsstSymTab* __restrict__ tabs)
{
  __shared__ uint64_t sym[256];
  __shared__ uint8_t len[256];
  const DictFsstSeg d = gload(dseg + blockIdx.y);
  if (!d.valid || d.mode == 0u || d.dict_count <= 1u) return;  // uniform per block
  sym[threadIdx.x] = gbl(tabs[d.range].sym)[threadIdx.x];
  len[threadIdx.x] = gbl(tabs[d.range].len)[threadIdx.x];
  __syncthreads();
  dfsst_predecode(d, e_byte, sym, len, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// One thread per byte of the null mask: each of its valid rows that lies in
// a DICT_FSST range is cleared when the range is invalid or, outside mode 2,
// selects entry 0 (dict_fsst.cu:426-446); the rows cleared are counted.
__global__ void k_dfsst_nulls(const FsstSeg* __restrict__ segs, uint32_t n_ranges,
                              const DictFsstSeg* __restrict__ dseg, const uint32_t* __restrict__ range_dseg,
                              uint8_t* __restrict__ mask, uint32_t total_rows, unsigned long long* __restrict__ count)
{
  uint32_t cleared = 0;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; uint64_t{b} * 8u < total_rows;
       b += gridDim.x * blockDim.x) {
    const uint32_t m0 = mask[b];
    uint32_t m        = m0;
    for (uint32_t j = 0; j < 8u && b * 8u + j < total_rows; ++j) {
      if (!((m >> j) & 1u)) continue;
      const uint32_t r = b * 8u + j;
      uint32_t lo = 0, hi = n_ranges;  // the range holding r: the last with col_row <= r
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (gbl(segs)[mid].col_row <= r) lo = mid;
        else hi = mid;
      }
      const uint32_t di = gbl(range_dseg)[lo];
      if (di == ~0u) continue;
      const DictFsstSeg d = gload(dseg + di);
      if (!d.valid || (d.mode != 2u && dfsst_index(d, r - gbl(segs)[lo].col_row) == 0u)) m &= ~(1u << j);
    }
    if (m != m0) {
      mask[b] = static_cast<uint8_t>(m);
      cleared += static_cast<uint32_t>(__builtin_popcount(m0 ^ m));
    }
  }
  const uint32_t tot = wave_incl_sum_u32(cleared);
  if (lane_id() == WAVE - 1 && tot) atomicAdd(count, static_cast<unsigned long long>(tot));
}

hipError_t launch_dfsst_headers(const FsstSeg* d_segs, uint32_t n_ranges, uint32_t* d_hdr, hipStream_t s)
{
  if (n_ranges == 0) return hipSuccess;
  const uint32_t blocks = (n_ranges + 255u) / 256u;
  hipLaunchKernelGGL(k_dfsst_headers, dim3(blocks), dim3(256), 0, s, d_segs, n_ranges, d_hdr);
  return hipGetLastError();
}

hipError_t launch_dfsst_prep(const DictFsstSeg* d_dseg, uint32_t n_dseg, uint32_t* d_e_byte, uint32_t* d_e_dec,
                             FsstSymTab* d_tabs, uint32_t* d_out, hipStream_t s)
{
  if (n_dseg == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dfsst_prep, dim3(n_dseg < 4096u ? n_dseg : 4096u), dim3(FSST_THREADS), 0, s, d_dseg, n_dseg,
                     d_e_byte, d_e_dec, d_tabs, d_out);
  return hipGetLastError();
}

hipError_t launch_dfsst_predecode(const DictFsstSeg* d_dseg, uint32_t n_dseg, const uint32_t* d_e_byte,
                                  const FsstSymTab* d_tabs, hipStream_t s)
{
  if (n_dseg == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dfsst_predecode, dim3(8, n_dseg), dim3(FSST_THREADS), 0, s, d_dseg, d_e_byte, d_tabs);
  return hipGetLastError();
}

hipError_t launch_dfsst_nulls(const FsstSeg* d_segs, uint32_t n_ranges, const DictFsstSeg* d_dseg,
                              const uint32_t* d_range_dseg, uint8_t* d_mask, uint32_t total_rows,
                              unsigned long long* d_count, hipStream_t s)
{
  if (total_rows == 0 || n_ranges == 0) return hipSuccess;
  uint32_t blocks = ((total_rows + 7u) / 8u + 255u) / 256u;
  if (blocks > 2048u) blocks = 2048u;
  hipLaunchKernelGGL(k_dfsst_nulls, dim3(blocks), dim3(256), 0, s, d_segs, n_ranges, d_dseg, d_range_dseg, d_mask,
                     total_rows, d_count);
  return hipGetLastError();
}

hipError_t launch_fsst_reset(FsstState* d_state, uint32_t n_chunks, hipStream_t s)
{
  uint32_t blocks = (n_chunks + 255) / 256;
  if (blocks == 0) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_fsst_reset, dim3(blocks), dim3(256), 0, s, d_state, n_chunks);
  return hipGetLastError();
}

hipError_t launch_fsst_prep(const FsstSeg* d_segs, uint32_t n_segs, const uint32_t* d_n_segs, uint32_t grid,
                            FsstChunk* d_chunks, FsstChunkInfo* d_info, FsstSymTab* d_tabs, hipStream_t s)
{
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fsst_prep, dim3(grid), dim3(FSST_THREADS), 0, s, d_segs, n_segs, d_n_segs, d_chunks, d_info,
                     d_tabs);
  return hipGetLastError();
}

hipError_t launch_fsst(const FsstChunk* d_chunks, uint32_t first, uint32_t n, uint32_t n_chunks_total,
                       const FsstChunkInfo* d_info, const FsstSymTab* d_tabs, FsstState* d_state,
                       int32_t* d_offsets, uint8_t* d_chars, uint64_t chars_cap, uint32_t total_rows,
                       hipStream_t s)
{
  if (n == 0) return hipSuccess;
  // Tickets are global chunk numbers: a launch covering [first, first + n)
  // runs after every launch of lower chunks on the same stream.
  (void)first;
  hipLaunchKernelGGL(k_fsst, dim3(n), dim3(FSST_THREADS), 0, s, d_chunks, n_chunks_total, d_info, d_tabs, d_state,
                     d_offsets, d_chars, chars_cap, total_rows, tune(SDR_TUNE_FSST_SPIN_LIMIT));
  return hipGetLastError();
}

}  // namespace sdr
