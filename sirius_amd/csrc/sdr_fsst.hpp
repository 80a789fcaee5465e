// sdr_fsst.hpp — descriptors shared by the FSST kernel and its host driver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdr {

constexpr uint32_t FSST_CHUNK_ROWS = 1024;        // rows per workgroup
constexpr uint32_t FSST_GAP        = 0xFFFFFFFFu;  // chunk of rows no segment covers
constexpr uint32_t FSST_SPIN_LIMIT = 1u << 24;     // look-back wait bound (never reached)

struct FsstSeg {
  const uint8_t* bytes;
  uint32_t bytes_size;
  uint32_t row_count;
};

// A chunk: `rows` rows starting at segment row `row_begin` (column row
// `col_row`).  Chunks are listed in column row order.
struct FsstChunk {
  uint32_t seg;
  uint32_t row_begin;
  uint32_t rows;
  uint32_t col_row;
};

// Device state of one column decode: ticket counter, look-back status per
// chunk, error flag and the decoded chars total.
struct FsstState {
  uint32_t ticket;
  uint32_t error;
  uint64_t total;
  uint64_t status[1];  // n_chunks entries
};

hipError_t launch_fsst_reset(FsstState* d_state, uint32_t n_chunks, hipStream_t s);
hipError_t launch_fsst(const FsstChunk* d_chunks, uint32_t first, uint32_t n, uint32_t n_chunks_total,
                       const FsstSeg* d_segs, FsstState* d_state, int32_t* d_offsets, uint8_t* d_chars,
                       uint64_t chars_cap, uint32_t total_rows, hipStream_t s);

}  // namespace sdr
