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

// One queue entry (64 bytes).  Fixed-width entries cover units
// [unit0, unit0 + n_units) of one segment, spread over the block's four
// wavefronts; an FSST entry is one chunk of one column, decoded by the
// whole block.
struct BatchItem {
  const uint8_t* seg;
  uint8_t* out;  // output of the segment's row 0 (fixed-width)
  uint32_t bytes_size;
  uint32_t row_count;
  uint32_t unit0;
  uint32_t n_units;
  uint8_t kind;
  uint8_t ts;
  uint16_t reserved;          // RLE: slice rows / 64
  uint32_t chunk;             // FSST: chunk index in the column chain
  const FsstColumn* column;   // FSST
  uint64_t pad[2];
};
static_assert(sizeof(BatchItem) == 64, "BatchItem layout");

// Queue header read by the kernel: the item count and the claim counter.
struct BatchQueue {
  uint32_t n_items;
  uint32_t claimed;
  uint32_t n_fsst_status;  // status words to clear in `fsst_status`
  uint32_t error;
  uint64_t* fsst_status;   // contiguous FsstState storage of all FSST columns
  uint32_t n_fsst_segs;    // row ranges of all FSST columns, for the pre-pass
  uint32_t spin_limit;     // look-back wait bound of the FSST chunks
  const FsstSeg* fsst_segs;      // their first_chunk index fsst_chunks and fsst_info
  FsstChunk* fsst_chunks;
  FsstChunkInfo* fsst_info;
  FsstSymTab* fsst_tabs;
};

hipError_t launch_batch_reset(BatchQueue* q, hipStream_t s);
// FSST segment pre-pass over the queue's segments (a fixed grid that loops,
// so one recorded launch serves every batch).
hipError_t launch_batch_fsst_prep(BatchQueue* q, hipStream_t s);
// `strings`: the queue holds VARCHAR (BK_FSST) entries; without them a
// kernel that carries no FSST state runs, at higher occupancy.
hipError_t launch_batch(const BatchItem* items, BatchQueue* q, uint32_t grid, bool strings, hipStream_t s);
// Blocks per CU the persistent kernel can keep resident.
int batch_occupancy_blocks_per_cu(bool strings);

}  // namespace sdr
