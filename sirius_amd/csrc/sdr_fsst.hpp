// sdr_fsst.hpp — descriptors shared by the FSST kernels and their host drivers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sdr/decode.h"

namespace sdr {

constexpr uint32_t FSST_CHUNK_ROWS = 1024;        // rows per workgroup
constexpr uint32_t FSST_GAP        = 0xFFFFFFFFu;  // chunk of rows no segment covers
constexpr uint32_t FSST_SPIN_LIMIT = 1u << 24;     // look-back wait bound (never reached)

// A row range of a column: a segment, or (bytes == nullptr) rows no
// segment covers, which decode to empty strings.  Its rows are chunks
// first_chunk, first_chunk + 1, ... of the column's chunk list, which the
// segment pre-pass writes; the range starts at column row col_row.  codec
// is the segment's sdr_codec (FSST, DICTIONARY or UNCOMPRESSED; 0 for a gap).
struct FsstSeg {
  const uint8_t* bytes;
  uint32_t bytes_size;
  uint32_t row_count;
  uint32_t first_chunk;
  uint32_t col_row;
  uint32_t codec;
  uint32_t pad;
};
static_assert(sizeof(FsstSeg) == 32, "FsstSeg layout");

// Chunks covering `rows` rows.
inline uint32_t fsst_chunk_count(uint32_t rows) { return (rows + FSST_CHUNK_ROWS - 1u) / FSST_CHUNK_ROWS; }

// A chunk: `rows` rows starting at segment row `row_begin` (column row
// `col_row`).  Chunks are listed in column row order.
struct FsstChunk {
  uint32_t seg;
  uint32_t row_begin;
  uint32_t rows;
  uint32_t col_row;
};

// What the segment pre-pass (k_fsst_prep) leaves for each chunk: the
// segment's header fields and the chunk's compressed-length prefix bounds,
// so a chunk decode issues all of its loads at once.
constexpr uint32_t FI_WIDTH = 0xFFu;      // bits 0-7: length-stream bit width
constexpr uint32_t FI_OK    = 1u << 8;    // header valid, symbol table imported
constexpr uint32_t FI_LFAST = 1u << 9;    // the chunk's length fields lie inside the segment
constexpr uint32_t FI_HUGE  = 1u << 10;   // a row of the chunk has >= 1 MiB of compressed bytes
constexpr uint32_t FI_DICT  = 1u << 11;   // a DICTIONARY segment (bits 0-7: selection width)
constexpr uint32_t FI_UNC   = 1u << 12;   // an UNCOMPRESSED string segment
constexpr uint32_t FI_DFSST = 1u << 13;   // a DICT_FSST segment (bytes: its DictFsstSeg)

struct FsstChunkInfo {
  const uint8_t* bytes;  // the segment
  uint32_t c_lo;         // compressed bytes of the segment's rows before the chunk
                         // (DICTIONARY: index_buffer_offset)
  uint32_t c_hi;         // ... up to and including the chunk's last row
                         // (DICTIONARY: index_buffer_count)
  uint32_t dict_end;
  uint32_t flags;        // width and FI_* bits
  uint32_t bytes_size;
  uint32_t pad;
};
static_assert(sizeof(FsstChunkInfo) == 32, "FsstChunkInfo layout");

// A segment's symbol table as the decoder reads it: symbol bytes (zero
// above the length) and lengths, 256 codes (255, the escape, has length 0).
struct FsstSymTab {
  uint64_t sym[256];
  uint8_t len[256];
};
static_assert(sizeof(FsstSymTab) == 2304, "FsstSymTab layout");

// A DICT_FSST segment (strings/dict_fsst.cu:17-38): the host fills the
// region offsets from the 16-byte header (:64-72, :516-525), the dictionary
// pre-pass the entry offsets, and `src` is the raw dict region (mode 0, or
// a dictionary of at most one entry) or the entries the predecode pass
// wrote.  Entry k of the segment is [e[k], e[k + 1]) of src.
struct DictFsstSeg {
  const uint8_t* bytes;  // the segment
  const uint32_t* e;     // dict_count + 1 entry offsets into src
  const uint8_t* src;
  uint64_t src_size;     // readable bytes at src
  uint32_t bytes_size;
  uint32_t dict_count;
  uint32_t off_dict, off_symtab, off_slens, off_didx;
  uint32_t range;        // its row range: index of its symbol table
  uint8_t mode, slw, diw, valid;
  uint32_t e_byte;       // offset of its byte_off entries in the pre-pass array
  uint32_t pad;
};
static_assert(sizeof(DictFsstSeg) == 72, "DictFsstSeg layout");

// DICT_FSST passes (k_fsst.hip): the 16-byte headers of the DICT_FSST
// ranges into hdr[4 * i] (ranges of other codecs are skipped); the
// per-segment dictionary pre-pass (byte and decoded entry offsets, symbol
// table, per-segment decoded total and inline-NULL flag into out[2 * j]);
// the predecode of FSST-mode entries; and the inline-NULL marking of a
// column's mask, counting the rows it clears into *d_count.
hipError_t launch_dfsst_headers(const FsstSeg* d_segs, uint32_t n_ranges, uint32_t* d_hdr, hipStream_t s);
hipError_t launch_dfsst_prep(const DictFsstSeg* d_dseg, uint32_t n_dseg, uint32_t* d_e_byte, uint32_t* d_e_dec,
                             FsstSymTab* d_tabs, uint32_t* d_out, hipStream_t s);
hipError_t launch_dfsst_predecode(const DictFsstSeg* d_dseg, uint32_t n_dseg, const uint32_t* d_e_byte,
                                  const FsstSymTab* d_tabs, hipStream_t s);
hipError_t launch_dfsst_nulls(const FsstSeg* d_segs, uint32_t n_ranges, const DictFsstSeg* d_dseg,
                              const uint32_t* d_range_dseg, uint8_t* d_mask, uint32_t total_rows,
                              unsigned long long* d_count, hipStream_t s);

// Device state of one column decode: ticket counter, look-back status per
// chunk, error flag and the decoded chars total.
struct FsstState {
  uint32_t ticket;
  uint32_t error;
  uint64_t total;
  uint64_t status[1];  // n_chunks entries
};

hipError_t launch_fsst_reset(FsstState* d_state, uint32_t n_chunks, hipStream_t s);
// Segment pre-pass over n_segs ranges (or *d_n_segs when given, for a
// launch recorded before the count is known): writes the chunk list, the
// chunks' FsstChunkInfo and the segments' symbol tables.
hipError_t launch_fsst_prep(const FsstSeg* d_segs, uint32_t n_segs, const uint32_t* d_n_segs, uint32_t grid,
                            FsstChunk* d_chunks, FsstChunkInfo* d_info, FsstSymTab* d_tabs, hipStream_t s);
hipError_t launch_fsst(const FsstChunk* d_chunks, uint32_t first, uint32_t n, uint32_t n_chunks_total,
                       const FsstChunkInfo* d_info, const FsstSymTab* d_tabs, FsstState* d_state,
                       int32_t* d_offsets, uint8_t* d_chars, uint64_t chars_cap, uint32_t total_rows,
                       hipStream_t s);

}  // namespace sdr
