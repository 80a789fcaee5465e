// k_batch.hip — persistent decode kernels of a mixed-codec batch.
//
// One launch decodes a whole device-resident work queue (sdr_batch.hpp):
// BITPACKING groups, RLE row slices, ALP/ALPRD vectors, UNCOMPRESSED and
// CONSTANT byte ranges and FSST chunks, from any number of columns and
// segments.  The queue is partitioned by codec and value width into groups
// (BatchGroup); each group is drained by its own persistent kernel,
// specialised for that one decoder, so it keeps the register and LDS
// footprint of the per-codec kernels instead of the union of all of them.
// A fixed-width group kernel's grid is its resident capacity.  Fixed-width
// entries are independent, so block b takes entries b, b + grid, ... with no
// claim at all; its four wavefronts decode an entry's units (an RLE slice is
// decoded by the whole block).  The FSST kernel runs one block per chunk,
// claimed in queue order with one atomic on the group's counter, which is
// what lets a chunk wait on its predecessors' offsets: a predecessor was
// claimed earlier by a block that started earlier.  The group kernels run as
// parallel branches (graph nodes) or in stream order, sharing the GPU as
// their queues drain.
#include "dec_fixed.hpp"
#include "dec_fsst.hpp"
#include "dec_rle.hpp"
#include "sdr_batch.hpp"

namespace sdr {

// Per-wave staging of the fixed-width codecs, as in the per-codec kernels:
// 4 KiB of bitpacking or ALP codes, or the ALP-RD left stream plus 4 KiB of
// its right stream.
constexpr uint32_t BATCH_BP_STAGE = 4096;
static_assert(BATCH_PREP_THREADS == FSST_THREADS, "pre-pass workgroup size");
static_assert(BATCH_WAVES == WAVES_PER_BLOCK, "fixed-width group block");

template <typename T>
__global__ __launch_bounds__(BLOCK_THREADS) void k_bgroup_bp(const BatchItem* __restrict__ items,
                                                             BatchQueue* __restrict__ q, uint32_t g)
{
  __shared__ uint32_t lds[WAVES_PER_BLOCK][(BATCH_BP_STAGE + 48) / 4];
  const uint32_t w     = wave_in_block();
  const uint32_t chunk = q->bp_chunk_rows[sizeof(T) == 1 ? 0 : sizeof(T) == 2 ? 1 : sizeof(T) == 4 ? 2 : 3];
  const uint32_t first = q->first[g], n = q->count[g];
  for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
    const BatchItem it = items[first + e];
    const uint32_t end = it.unit0 + it.n_units;
    for (uint32_t u = it.unit0 + w; u < end; u += WAVES_PER_BLOCK) {
      const uint32_t grc = it.row_count - u * BP_GROUP < BP_GROUP ? it.row_count - u * BP_GROUP : BP_GROUP;
      bp_group<T, BATCH_BP_STAGE>(it.seg, it.bytes_size, u, grc, reinterpret_cast<T*>(it.out) + size_t{u} * BP_GROUP,
                                  lds[w], chunk);
    }
  }
}

template <typename F, bool RD>
__global__ __launch_bounds__(BLOCK_THREADS) void k_bgroup_alp(const BatchItem* __restrict__ items,
                                                              BatchQueue* __restrict__ q, uint32_t g)
{
  __shared__ uint32_t lds[WAVES_PER_BLOCK][RD ? ALPRD_LEFT_WORDS + ALP_CHUNK_WORDS : ALP_CHUNK_WORDS];
  const uint32_t w     = wave_in_block();
  const uint32_t first = q->first[g], n = q->count[g];
  for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
    const BatchItem it = items[first + e];
    const uint32_t end = it.unit0 + it.n_units;
    for (uint32_t u = it.unit0 + w; u < end; u += WAVES_PER_BLOCK) {
      const uint32_t fill = it.row_count - u * ALP_VEC < ALP_VEC ? it.row_count - u * ALP_VEC : ALP_VEC;
      F* out = reinterpret_cast<F*>(it.out) + size_t{u} * ALP_VEC;
      if constexpr (RD) alprd_vector<F>(it.seg, it.bytes_size, u, fill, out, lds[w]);
      else alp_vector<F>(it.seg, it.bytes_size, u, fill, out, lds[w]);
    }
  }
}

template <uint32_t TS>
__global__ __launch_bounds__(RLE_THREADS) void k_bgroup_rle(const BatchItem* __restrict__ items,
                                                            BatchQueue* __restrict__ q, uint32_t g)
{
  __shared__ RleSmem sm;
  const uint32_t first = q->first[g], n = q->count[g];
  for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
    const BatchItem it = items[first + e];
    const uint32_t sr  = uint32_t{it.reserved} * 64u;
    for (uint32_t u = it.unit0; u < it.unit0 + it.n_units; ++u) {
      const uint32_t ra = u * sr;
      const uint32_t rb = it.row_count - ra < sr ? it.row_count : ra + sr;
      rle_slice_t<TS>(it.seg, it.bytes_size, it.row_count, ra, rb, it.out, sm);
    }
    __syncthreads();  // LDS reuse by the next entry
  }
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_bgroup_copy(const BatchItem* __restrict__ items,
                                                               BatchQueue* __restrict__ q, uint32_t g)
{
  const uint32_t w     = wave_in_block();
  const uint32_t first = q->first[g], n = q->count[g];
  for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
    const BatchItem it   = items[first + e];
    const uint64_t total = uint64_t{it.row_count} * it.ts;
    for (uint32_t u = it.unit0 + w; u < it.unit0 + it.n_units; u += WAVES_PER_BLOCK) {
      const uint64_t b0 = uint64_t{u} * COPY_UNIT_BYTES;
      const uint64_t b1 = total - b0 < COPY_UNIT_BYTES ? total : b0 + COPY_UNIT_BYTES;
      if (it.kind == BK_COPY) copy_span<WAVE>(it.seg, it.out, b0, b1, lane_id());
      else const_span<WAVE>(it.seg, it.ts, it.out, b0, b1, lane_id());
    }
  }
}

// One block per FSST entry (the grid is the group's entry count).  The block
// claims the next entry in queue order with one atomic, so a chunk's
// predecessors in its column are held by blocks that started earlier, which
// the look-back's progress needs (the per-column kernel's ticket does the
// same).  One chunk per block, with no loop around it, keeps the kernel at
// 80 VGPRs: six workgroups per CU with the 11 KiB window.
__global__ __launch_bounds__(FSST_THREADS) __attribute__((amdgpu_waves_per_eu(6))) void k_bgroup_fsst(
  const BatchItem* __restrict__ items, BatchQueue* __restrict__ q, uint32_t g)
{
  __shared__ FsstSmemT<FSST_OUT_BATCH> sm;
  __shared__ uint32_t s_idx;
  (void)items;  // the FSST group's entries are q->fsst_entries
  // read before the claim, so they arrive while it does
  const uint32_t first = q->first[g], count = q->count[g];
  const uint32_t spin  = q->spin_limit;
  if (threadIdx.x == 0) s_idx = atomicAdd(&q->claimed[g].v, 1u);
  __syncthreads();
  const uint32_t i = s_idx;
  if (i >= count) return;  // the grid is the entry count: never taken
  const uint32_t idx = first + i;
  // the entries list every column's chunks column by column, as the batch's
  // chunk and chunk-info arrays do: entry idx describes chunk idx of those
  // arrays, so the chunk loads need not wait for the column's; every lane
  // reads the same descriptors, kept in SGPRs
  const FsstEntry e      = uniform(gload(q->fsst_entries + idx));
  const FsstChunk ch     = uniform(gload(q->fsst_chunks + idx));
  const FsstChunkInfo ci = uniform(gload(q->fsst_info + idx));
  // the column comes through the scalar cache (read-only for the whole
  // launch), so its wait falls behind the loads of the symbol table and
  // the compressed bytes
  const FsstColumn col = sload(q->fsst_cols + e.col);
  fsst_chunk<FSST_OUT_BATCH>(ch, ci, q->fsst_tabs + (ch.seg == FSST_GAP ? 0u : ch.seg), e.chunk, col.n_chunks,
                             col.state, col.offsets, col.chars, col.chars_cap, col.total_rows, spin, sm);
}

__global__ void k_batch_reset(BatchQueue* q)
{
  const uint32_t n = q->n_fsst_status;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) gbl(q->fsst_status)[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < BG_COUNT) q->claimed[threadIdx.x].v = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) q->error = 0;
}

__global__ __launch_bounds__(FSST_THREADS) void k_batch_fsst_prep(const BatchQueue* __restrict__ q)
{
  __shared__ FsstPrepSmem sm;
  const uint32_t n = q->n_fsst_segs;
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x)
    fsst_prep_segment(gload(q->fsst_segs + i), i, q->fsst_chunks, q->fsst_info, q->fsst_tabs + i, sm);
}

namespace {

// The kernel of group g (a BatchGroup).
const void* group_kernel(uint32_t g)
{
  switch (g) {
    case BG_BP1: return reinterpret_cast<const void*>(k_bgroup_bp<uint8_t>);
    case BG_BP2: return reinterpret_cast<const void*>(k_bgroup_bp<uint16_t>);
    case BG_BP4: return reinterpret_cast<const void*>(k_bgroup_bp<uint32_t>);
    case BG_BP8: return reinterpret_cast<const void*>(k_bgroup_bp<uint64_t>);
    case BG_RLE1: return reinterpret_cast<const void*>(k_bgroup_rle<1>);
    case BG_RLE2: return reinterpret_cast<const void*>(k_bgroup_rle<2>);
    case BG_RLE4: return reinterpret_cast<const void*>(k_bgroup_rle<4>);
    case BG_RLE8: return reinterpret_cast<const void*>(k_bgroup_rle<8>);
    case BG_ALP4: return reinterpret_cast<const void*>(k_bgroup_alp<float, false>);
    case BG_ALP8: return reinterpret_cast<const void*>(k_bgroup_alp<double, false>);
    case BG_ALPRD4: return reinterpret_cast<const void*>(k_bgroup_alp<float, true>);
    case BG_ALPRD8: return reinterpret_cast<const void*>(k_bgroup_alp<double, true>);
    case BG_COPY: return reinterpret_cast<const void*>(k_bgroup_copy);
    default: return reinterpret_cast<const void*>(k_bgroup_fsst);
  }
}

}  // namespace

hipError_t batch_group_launch_params(uint32_t g, const void** func, uint32_t* block, uint32_t* blocks_per_cu)
{
  *func  = group_kernel(g);
  *block = g == BG_FSST ? FSST_THREADS : (g >= BG_RLE1 && g <= BG_RLE8) ? RLE_THREADS : BLOCK_THREADS;
  int per = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, *func, static_cast<int>(*block), 0);
  *blocks_per_cu = per > 0 ? static_cast<uint32_t>(per) : 1u;
  return e;
}

hipError_t launch_batch_fsst_prep(BatchQueue* q, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_fsst_prep, dim3(BATCH_PREP_GRID), dim3(BATCH_PREP_THREADS), 0, s, q);
  return hipGetLastError();
}

const void* batch_fsst_prep_kernel() { return reinterpret_cast<const void*>(k_batch_fsst_prep); }
const void* batch_reset_kernel() { return reinterpret_cast<const void*>(k_batch_reset); }

hipError_t launch_batch_reset(BatchQueue* q, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_reset, dim3(BATCH_RESET_GRID), dim3(256), 0, s, q);
  return hipGetLastError();
}

hipError_t launch_batch_group(const BatchItem* items, BatchQueue* q, uint32_t g, uint32_t grid, hipStream_t s)
{
  const void* f = nullptr;
  uint32_t block = 0, per = 0;
  hipError_t e = batch_group_launch_params(g, &f, &block, &per);
  if (e != hipSuccess) return e;
  void* args[] = {&items, &q, &g};
  return hipLaunchKernel(f, dim3(grid), dim3(block), args, 0, s);
}

}  // namespace sdr
