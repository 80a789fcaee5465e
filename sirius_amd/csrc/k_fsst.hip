// k_fsst.hip — per-column FSST decode kernel (one launch decodes every chunk
// of a VARCHAR column; see dec_fsst.hpp for the chunk decoder).
#include "dec_fsst.hpp"
#include "sdr_tune.hpp"

namespace sdr {

__global__ __launch_bounds__(FSST_THREADS) void k_fsst(const FsstChunk* __restrict__ chunks, uint32_t n_chunks,
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
