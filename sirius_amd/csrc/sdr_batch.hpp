// sdr_batch.hpp — device work queue of the persistent mixed-codec kernel.
#pragma once

#include "sdr_fsst.hpp"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdr {

enum BatchKind : uint8_t {
  BK_BITPACKING = 1,
  BK_RLE        = 2,
  BK_ALP        = 3,
  BK_ALPRD      = 4,
  BK_COPY       = 5,
  BK_CONSTANT   = 6,
  BK_FSST       = 7,
};

// Rows (RLE), bytes (COPY/CONSTANT) or values (others) one unit covers.
constexpr uint32_t COPY_UNIT_BYTES = 16384;

// One FSST column of a batch: its chunk chain and outputs.
struct FsstColumn {
  const FsstChunk* chunks;
  const FsstSeg* segs;
  const FsstChunkInfo* info;  // per chunk, from the segment pre-pass
  const FsstSymTab* tabs;     // per row range of the batch (a chunk's seg indexes it), from the pre-pass
  FsstState* state;
  int32_t* offsets;
  uint8_t* chars;
  uint64_t chars_cap;
  uint32_t total_rows;
  uint32_t n_chunks;
};

// One fixed-width queue entry (40 bytes): units [unit0, unit0 + n_units) of
// one segment, spread over the block's four wavefronts (the host also keeps
// one per segment to plan the per-codec grids).
struct BatchItem {
  const uint8_t* seg;
  uint8_t* out;  // output of the segment's row 0
  uint32_t bytes_size;
  uint32_t row_count;
  uint32_t unit0;
  uint32_t n_units;
  uint8_t kind;
  uint8_t ts;
  uint16_t reserved;  // RLE: slice rows / 64
};
static_assert(sizeof(BatchItem) == 40, "BatchItem layout");

// One FSST queue entry: chunk `chunk` of the batch's string column `col`,
// decoded by a whole block.
struct FsstEntry {
  uint32_t col;
  uint32_t chunk;
};

// Queue groups: entries of one decoder and value width, drained by one
// specialised persistent kernel (k_batch.hip).  Entries are laid out group
// by group; within a group in the order they were added.
enum BatchGroup : uint32_t {
  BG_FSST = 0,
  BG_BP1, BG_BP2, BG_BP4, BG_BP8,
  BG_RLE1, BG_RLE2, BG_RLE4, BG_RLE8,
  BG_ALP4, BG_ALP8, BG_ALPRD4, BG_ALPRD8,
  BG_COPY,  // UNCOMPRESSED and CONSTANT byte ranges
  BG_COUNT
};

// A claim counter on a cache line of its own: the group kernels run at once
// and each claims from its own counter.
struct alignas(128) BatchCounter {
  uint32_t v;
};

// Queue header read by the kernels.
struct BatchQueue {
  BatchCounter claimed[BG_COUNT];  // next unclaimed entry of each group (reset per launch)
  uint32_t first[BG_COUNT];        // each group's entries: [first, first + count)
  uint32_t count[BG_COUNT];
  uint32_t n_items;
  uint32_t error;
  uint32_t n_fsst_status;  // status words to clear in `fsst_status`
  uint32_t spin_limit;     // look-back wait bound of the FSST chunks
  uint32_t n_fsst_segs;    // row ranges of all FSST columns, for the pre-pass
  uint32_t bp_chunk_rows[4];  // most rows a bitpacking wavefront stages at once, per width 1/2/4/8
  uint64_t* fsst_status;   // contiguous FsstState storage of all FSST columns
  const FsstSeg* fsst_segs;      // their first_chunk index fsst_chunks and fsst_info
  FsstChunk* fsst_chunks;
  FsstChunkInfo* fsst_info;
  FsstSymTab* fsst_tabs;
  const FsstEntry* fsst_entries;  // the FSST group's entries (first[BG_FSST] is 0)
  const FsstColumn* fsst_cols;    // the string columns they index
};

// Fixed grids of the reset and FSST pre-pass kernels (both loop), so one
// recorded launch serves every batch.
constexpr uint32_t BATCH_RESET_GRID    = 64;
constexpr uint32_t BATCH_PREP_GRID     = 512;
constexpr uint32_t BATCH_PREP_THREADS  = 256;  // the FSST pre-pass workgroup
constexpr uint32_t BATCH_WAVES         = 4;    // wavefronts of a fixed-width group block
constexpr uint32_t BATCH_XCDS          = 8;    // XCDs workgroups are dispatched to round-robin;
                                               // fixed-width group grids are a multiple

// One per-codec grid of a batch's fixed-width segments: the segments of
// codec kind `kind` and value width `ts`, descriptors [base, base + n) of the
// batch's FixedSeg array, the largest of them max_rows rows.
struct FixedLaunch {
  uint32_t kind, ts, base, n, max_rows;
  bool operator==(const FixedLaunch& o) const
  {
    return kind == o.kind && ts == o.ts && base == o.base && n == o.n && max_rows == o.max_rows;
  }
};

hipError_t launch_batch_reset(BatchQueue* q, hipStream_t s);
// FSST segment pre-pass over the queue's segments.
hipError_t launch_batch_fsst_prep(BatchQueue* q, hipStream_t s);
// The persistent kernel of group g over `grid` blocks.
hipError_t launch_batch_group(const BatchItem* items, BatchQueue* q, uint32_t g, uint32_t grid, hipStream_t s);
// Kernel, block size and resident blocks per CU of group g (for graph nodes
// and grid sizing).
hipError_t batch_group_launch_params(uint32_t g, const void** func, uint32_t* block, uint32_t* blocks_per_cu);
const void* batch_fsst_prep_kernel();
const void* batch_reset_kernel();

}  // namespace sdr
