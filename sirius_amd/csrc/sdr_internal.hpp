// sdr_internal.hpp — host-side helpers shared by the C ABI translation units.
#pragma once

#include "../../include/sdr/decode.h"
#include "sdr_kernels.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace sdr {

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
const char* last_error();

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess) return ::sdr::fail(SDR_E_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

void ensure_pool(int dev);
int device_alloc(void** p, size_t bytes, hipStream_t s);
void device_free(void* p, hipStream_t s);
// Host-to-device copy of a small host array (descriptors) through a pinned
// staging slot: asynchronous for the host, ordered on s.
int upload_async(void* d_dst, const void* h_src, size_t bytes, hipStream_t s);
// Per-call descriptors in a reused slot (pinned + device buffer): copies
// h_src to the slot's device buffer on s and returns it; desc_release marks
// the slot free once the work enqueued on s so far has completed.
int desc_acquire(const void* h_src, size_t bytes, hipStream_t s, void** d_out, void** token);
int desc_release(void* token, hipStream_t s);

bool type_ok_for(uint32_t codec, uint32_t ts);
bool codec_fixed(uint32_t codec);
int check_runs(const sdr_run* runs, uint32_t n, uint32_t total_rows, const char* what);
int check_validity_runs(const sdr_run* runs, uint32_t n, uint32_t total_rows);
int build_null_mask(const sdr_run* runs, uint32_t n, uint32_t total_rows, uint8_t* d_mask,
                    unsigned long long* d_count, hipStream_t s);

// One fixed-width run and where its rows go.
struct RunTarget {
  const sdr_run* run;
  void* d_out;
  uint32_t type_size;
};
int decode_runs(const std::vector<RunTarget>& targets, hipStream_t s, uint32_t mode);
int validate_fixed_columns(const sdr_column* cols, uint32_t n_cols);
int validate_string_column(const sdr_string_column* col);
// The column has DICT_FSST segments (decoded only by the per-column passes).
bool has_dict_fsst(const sdr_string_column* col);
// SDR_LAUNCH_PERSISTENT: one persistent launch through a thread-local batch.
int decode_table_persistent(sdr_column* cols, uint32_t n_cols, hipStream_t s);
int decode_strings_persistent(sdr_string_column* col, hipStream_t s);
// Batch internals used by the server (sdr_serve.cpp); the caller validates.
int batch_add_fixed(sdr_batch* b, sdr_column* cols, uint32_t n_cols);
int batch_add_strings(sdr_batch* b, sdr_string_column* col);
int batch_launch(sdr_batch* b, hipStream_t s);
int batch_ready(sdr_batch* b);    // 1 finished, 0 running, < 0 error
int batch_collect(sdr_batch* b);  // after batch_ready: write back totals and null counts
int batch_column_status(const sdr_batch* b, const sdr_string_column* col);

// Segments grouped by (codec, type size); [first, first + count) in the
// descriptor array.
struct Batch {
  uint32_t codec, ts;
  uint32_t first, count;
  uint32_t max_rows;
};

}  // namespace sdr

struct sdr_plan {
  uint32_t mode   = 0;
  uint32_t n_cols = 0;
  std::vector<sdr::Batch> batches;
  std::vector<sdr::FixedSeg> all;
  sdr::FixedSeg* d_segs = nullptr;
  std::vector<sdr_column> masked;  // columns with a null mask, in column order
  std::vector<bool> col_masked;
  unsigned long long* d_counts = nullptr;
  ~sdr_plan();
};
