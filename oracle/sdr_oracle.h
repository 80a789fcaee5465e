/*
 * sdr_oracle.h — CPU restatement of the DuckDB segment decoders.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP decode
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped.  The product
 * library (sirius_amd/csrc) does not link or call it.
 *
 * The decode entry points take the same structs as include/sdr/decode.h,
 * with host pointers in place of device pointers.
 */
#ifndef SDR_ORACLE_H
#define SDR_ORACLE_H

#include <stdint.h>

#include "../include/sdr/decode.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Decode `row_count` rows of one fixed-width segment into `out`. */
int oracle_decode_segment(uint32_t codec, const uint8_t* seg, uint32_t bytes_size,
                          uint32_t row_count, uint32_t type_size, void* out);

/* Column-level decode with host pointers (same validation as the library). */
int oracle_decode_table(sdr_column* cols, uint32_t n_cols);
int oracle_decode_strings_column(sdr_string_column* col);

/* FSST: decoded length of every row of one segment (lengths[row_count]). */
int oracle_fsst_lengths(const uint8_t* seg, uint32_t bytes_size, uint32_t row_count,
                        uint32_t* lengths);
/* FSST: decode one segment; `chars` receives sum(lengths) bytes. */
int oracle_fsst_decode(const uint8_t* seg, uint32_t bytes_size, uint32_t row_count,
                       uint32_t* lengths, uint8_t* chars, uint64_t chars_cap);

/* DICT_FSST: decode one segment (lengths[row_count]; `chars` may be NULL
 * for lengths only); *inline_null receives the segment's inline-NULL flag. */
int oracle_dict_fsst_decode(const uint8_t* seg, uint32_t bytes_size, uint32_t row_count,
                            uint32_t* lengths, uint8_t* chars, uint64_t chars_cap, uint8_t* inline_null);

/* Multithreaded column decode used as the timed CPU baseline: segments of
 * every run are spread over `n_threads` threads (1 = scalar). */
int oracle_decode_table_mt(sdr_column* cols, uint32_t n_cols, int n_threads);
int oracle_decode_strings_column_mt(sdr_string_column* col, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
