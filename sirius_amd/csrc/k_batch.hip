// k_batch.hip — persistent mixed-codec decode kernel.
//
// One launch decodes a whole device-resident work queue (sdr_batch.hpp):
// BITPACKING groups, RLE row slices, ALP/ALPRD vectors, UNCOMPRESSED and
// CONSTANT byte ranges and FSST chunks, from any number of columns and
// segments.  A grid sized to the resident capacity of the GPU loops: a block
// claims the next queue entry with one atomic, its four wavefronts decode the
// entry's units (RLE slices and FSST chunks are decoded by the whole block),
// and it claims again until the queue is drained.  Claiming in queue order is what
// lets FSST chunks of one column wait on their predecessors' offsets: a
// predecessor was claimed earlier by a running block.
#include "dec_fixed.hpp"
#include "dec_rle.hpp"
#include "dec_fsst.hpp"
#include "sdr_batch.hpp"

namespace sdr {

// Per-wave staging of the fixed-width codecs: 4 KiB of bitpacking or ALP
// codes, or the ALP-RD left stream plus 4 KiB of its right stream.
constexpr uint32_t BATCH_BP_STAGE    = 4096;
constexpr uint32_t BATCH_FIXED_WORDS = ALPRD_LEFT_WORDS + ALP_CHUNK_WORDS;
static_assert(BATCH_FIXED_WORDS >= (BATCH_BP_STAGE + 48) / 4, "bitpacking stage fits");

// Queues with VARCHAR entries need the FSST chunk state; fixed-width-only
// queues run a kernel without it, at half the LDS.
union __attribute__((aligned(16))) BatchSmemFixed {
  uint32_t fixed[WAVES_PER_BLOCK][BATCH_FIXED_WORDS];
  RleSmem rle;
};
union __attribute__((aligned(16))) BatchSmemAll {
  uint32_t fixed[WAVES_PER_BLOCK][BATCH_FIXED_WORDS];
  RleSmem rle;
  FsstSmemT<FSST_OUT_BATCH> fsst;
};

template <typename T, typename SM>
__device__ __forceinline__ void fixed_units(const BatchItem& it, uint32_t w, SM& sm)
{
  const uint32_t end = it.unit0 + it.n_units;
  for (uint32_t u = it.unit0 + w; u < end; u += WAVES_PER_BLOCK) {
    if (it.kind == BK_BITPACKING) {
      const uint32_t grc = it.row_count - u * BP_GROUP < BP_GROUP ? it.row_count - u * BP_GROUP : BP_GROUP;
      bp_group<T, BATCH_BP_STAGE>(it.seg, it.bytes_size, u, grc, reinterpret_cast<T*>(it.out) + size_t{u} * BP_GROUP,
                                  sm.fixed[w]);
    } else if constexpr (sizeof(T) >= 4) {
      const uint32_t fill = it.row_count - u * ALP_VEC < ALP_VEC ? it.row_count - u * ALP_VEC : ALP_VEC;
      using F = std::conditional_t<sizeof(T) == 8, double, float>;
      F* out = reinterpret_cast<F*>(it.out) + size_t{u} * ALP_VEC;
      if (it.kind == BK_ALP) alp_vector<F>(it.seg, it.bytes_size, u, fill, out, sm.fixed[w]);
      else alprd_vector<F>(it.seg, it.bytes_size, u, fill, out, sm.fixed[w]);
    }
  }
}

template <bool STRINGS, typename SM>
__device__ __forceinline__ void batch_loop(const BatchItem* __restrict__ items, BatchQueue* __restrict__ q, SM& sm,
                                           uint32_t& s_idx)
{
  const uint32_t n    = q->n_items;
  const uint32_t spin = q->spin_limit;
  const uint32_t w    = wave_in_block();
  for (;;) {
    if (threadIdx.x == 0) s_idx = atomicAdd(&q->claimed, 1u);
    __syncthreads();
    const uint32_t idx = s_idx;
    __syncthreads();
    if (idx >= n) break;
    const BatchItem it = items[idx];
    switch (it.kind) {
      case BK_FSST:
        if constexpr (STRINGS) {
          const FsstColumn col = gload(it.column);
          const FsstChunk ch   = gload(col.chunks + it.chunk);
          fsst_chunk<FSST_OUT_BATCH>(ch, gload(col.info + it.chunk), col.tabs + (ch.seg == FSST_GAP ? 0u : ch.seg),
                                     it.chunk, col.n_chunks, col.state, col.offsets, col.chars, col.chars_cap,
                                     col.total_rows, spin, sm.fsst);
        }  // else unreachable: queues with VARCHAR entries run k_batch
        break;
      case BK_RLE:
        for (uint32_t u = it.unit0; u < it.unit0 + it.n_units; ++u) {
          const uint32_t sr = uint32_t{it.reserved} * 64u;
          const uint32_t ra = u * sr;
          const uint32_t rb = it.row_count - ra < sr ? it.row_count : ra + sr;
          rle_slice(it.seg, it.bytes_size, it.row_count, ra, rb, it.out, it.ts, sm.rle);
        }
        break;
      case BK_COPY:
      case BK_CONSTANT: {
        const uint64_t total = uint64_t{it.row_count} * it.ts;
        for (uint32_t u = it.unit0 + w; u < it.unit0 + it.n_units; u += WAVES_PER_BLOCK) {
          const uint64_t b0 = uint64_t{u} * COPY_UNIT_BYTES;
          const uint64_t b1 = total - b0 < COPY_UNIT_BYTES ? total : b0 + COPY_UNIT_BYTES;
          if (it.kind == BK_COPY) copy_span<WAVE>(it.seg, it.out, b0, b1, lane_id());
          else const_span<WAVE>(it.seg, it.ts, it.out, b0, b1, lane_id());
        }
        break;
      }
      default:
        switch (it.ts) {
          case 1: fixed_units<uint8_t>(it, w, sm); break;
          case 2: fixed_units<uint16_t>(it, w, sm); break;
          case 4: fixed_units<uint32_t>(it, w, sm); break;
          default: fixed_units<uint64_t>(it, w, sm); break;
        }
        break;
    }
    __syncthreads();  // LDS reuse by the next entry
  }
}

__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(const BatchItem* __restrict__ items,
                                                                                          BatchQueue* __restrict__ q)
{
  __shared__ BatchSmemAll sm;
  __shared__ uint32_t s_idx;
  batch_loop<true>(items, q, sm, s_idx);
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_batch_fixed(const BatchItem* __restrict__ items,
                                                               BatchQueue* __restrict__ q)
{
  __shared__ BatchSmemFixed sm;
  __shared__ uint32_t s_idx;
  batch_loop<false>(items, q, sm, s_idx);
}

__global__ void k_batch_reset(BatchQueue* q)
{
  const uint32_t n = q->n_fsst_status;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) gbl(q->fsst_status)[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    q->claimed = 0;
    q->error   = 0;
  }
}

__global__ __launch_bounds__(FSST_THREADS) void k_batch_fsst_prep(const BatchQueue* __restrict__ q)
{
  __shared__ FsstPrepSmem sm;
  const uint32_t n = q->n_fsst_segs;
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x)
    fsst_prep_segment(gload(q->fsst_segs + i), i, q->fsst_chunks, q->fsst_info, q->fsst_tabs + i, sm);
}

hipError_t launch_batch_fsst_prep(BatchQueue* q, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_fsst_prep, dim3(512), dim3(FSST_THREADS), 0, s, q);
  return hipGetLastError();
}

hipError_t launch_batch_reset(BatchQueue* q, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_reset, dim3(64), dim3(256), 0, s, q);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchItem* items, BatchQueue* q, uint32_t grid, bool strings, hipStream_t s)
{
  if (strings) hipLaunchKernelGGL(k_batch, dim3(grid), dim3(BLOCK_THREADS), 0, s, items, q);
  else hipLaunchKernelGGL(k_batch_fixed, dim3(grid), dim3(BLOCK_THREADS), 0, s, items, q);
  return hipGetLastError();
}

int batch_occupancy_blocks_per_cu(bool strings)
{
  int blocks = 0;
  const void* k = strings ? reinterpret_cast<const void*>(k_batch) : reinterpret_cast<const void*>(k_batch_fixed);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, BLOCK_THREADS, 0) != hipSuccess) return 1;
  return blocks > 0 ? blocks : 1;
}

}  // namespace sdr
