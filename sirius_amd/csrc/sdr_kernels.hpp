// sdr_kernels.hpp — host-side launch wrappers of the decode kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdr {

// One fixed-width segment as the kernels see it.  `out` already points at
// the segment's first output row (d_out + row_offset * type_size).
struct FixedSeg {
  const uint8_t* bytes;
  uint8_t* out;
  uint32_t bytes_size;
  uint32_t row_count;
};
static_assert(sizeof(FixedSeg) == 24, "FixedSeg layout");

// Rows of one RLE slice (one 256-thread block decodes a slice).
constexpr uint32_t RLE_SLICE_ROWS = 16384;

// Descriptors a launch can carry in its kernel arguments.
constexpr uint32_t SEGS_INLINE = 128;

// Launch the decode of segments [seg_base, seg_base + n) of `segs`: a device
// array, or with `host` a host array whose descriptors travel in the kernel
// arguments (SEGS_INLINE per launch, no upload).  `max_rows` is the largest
// row_count among them (sizes the grid).
hipError_t launch_bitpacking(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows,
                             uint32_t type_size, hipStream_t s, bool host = false);
hipError_t launch_rle(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t type_size,
                      hipStream_t s, bool host = false);
hipError_t launch_alp(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows, uint32_t type_size,
                      bool rd, hipStream_t s, bool host = false);
hipError_t launch_uncompressed(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows,
                               uint32_t type_size, hipStream_t s, bool host = false);
hipError_t launch_constant(const FixedSeg* segs, uint32_t seg_base, uint32_t n, uint32_t max_rows,
                           uint32_t type_size, hipStream_t s, bool host = false);

// Row ids: out[r.out_offset + i] = r.first_row + i for each range r (i < r.count).
struct RowRange {
  int64_t first_row;
  uint64_t out_offset;
  uint32_t count;
  uint32_t reserved;
};
hipError_t launch_rowids(const RowRange* d_ranges, uint32_t n, uint32_t max_count, int64_t* out, hipStream_t s);

// Validity: fill mask bytes, then count zero bits among the first `rows`.
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, uint64_t n, hipStream_t s);
hipError_t launch_count_nulls(const uint8_t* mask, uint32_t rows, unsigned long long* d_count,
                              hipStream_t s);

}  // namespace sdr
