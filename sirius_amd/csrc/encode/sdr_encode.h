/*
 * sdr_encode.h — DuckDB-format segment encoders and seeded synthetic column
 * generators (host C++, C ABI).
 *
 * These produce the bytes the decoder consumes when real DuckDB is not
 * available: every segment follows DuckDB's on-disk layout for its codec
 * (BITPACKING, RLE, ALP, ALPRD, FSST), cut at DuckDB's row-group size and
 * block size.  Encoder *choices* (bitpacking mode, ALP exponent/factor,
 * ALPRD cut, FSST symbol table) restate DuckDB's published algorithms; any
 * valid choice decodes identically, so parity does not depend on them.
 *
 * Output: one blob holding every segment (each starting on a 16-byte
 * boundary, as the reference's staging buffer places them) plus a segment
 * table whose `offset` fields index the blob.
 */
#ifndef SDR_ENCODE_H
#define SDR_ENCODE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdr_enc_segment {
  uint64_t offset; /* byte offset of the segment in the blob */
  uint32_t bytes_size;
  uint32_t row_offset;
  uint32_t row_count;
  uint32_t max_string_length;
} sdr_enc_segment;

typedef struct sdr_encoded sdr_encoded;

/* DuckDB defaults: 122880-row row groups, 262136-byte blocks. */
#define SDR_ROW_GROUP_ROWS 122880u
#define SDR_BLOCK_SIZE 262136u

/* force_mode: 0 = DuckDB's AUTO choice, else a BitpackingMode value
 * (2 CONSTANT, 3 CONSTANT_DELTA, 4 DELTA_FOR, 5 FOR) used whenever valid. */
sdr_encoded* sdr_encode_bitpacking(const void* values, uint64_t n, uint32_t type_size,
                                   int is_signed, uint32_t row_group_rows, uint32_t block_size,
                                   int force_mode);
sdr_encoded* sdr_encode_rle(const void* values, uint64_t n, uint32_t type_size,
                            uint32_t row_group_rows, uint32_t block_size);
/* type_size 4 (float) or 8 (double) */
sdr_encoded* sdr_encode_alp(const void* values, uint64_t n, uint32_t type_size,
                            uint32_t row_group_rows, uint32_t block_size);
sdr_encoded* sdr_encode_alprd(const void* values, uint64_t n, uint32_t type_size,
                              uint32_t row_group_rows, uint32_t block_size);
sdr_encoded* sdr_encode_uncompressed(const void* values, uint64_t n, uint32_t type_size,
                                     uint32_t row_group_rows, uint32_t block_size);
/* strings: chars + offsets[n+1] (uint64) */
sdr_encoded* sdr_encode_fsst(const uint8_t* chars, const uint64_t* offsets, uint64_t n,
                             uint32_t row_group_rows, uint32_t block_size, uint64_t seed);

/* DICT_FSST segments of up to seg_rows rows (halved until a segment fits
 * block_size): mode 0 raw dictionary, 1 FSST-compressed dictionary, 2 one
 * FSST entry per row.  valid (n bytes, or NULL for all valid): rows with 0
 * select entry 0, the NULL slot (modes 0 and 1). */
sdr_encoded* sdr_encode_dict_fsst(const uint8_t* chars, const uint64_t* offsets, const uint8_t* valid,
                                  uint64_t n, uint32_t mode, uint32_t seg_rows, uint32_t block_size,
                                  uint64_t seed);

const uint8_t* sdr_encoded_bytes(const sdr_encoded* e, uint64_t* size);
const sdr_enc_segment* sdr_encoded_segments(const sdr_encoded* e, uint32_t* n);
void sdr_encoded_free(sdr_encoded* e);

/* ---- seeded generators (splitmix64-seeded xoshiro256**) ---------------- */

void sdr_gen_u64(uint64_t seed, uint64_t n, uint64_t* out);
/* config 1: u32 % 2^17, except every 2048-row group with index % 10 == 9
 * which draws u32 % 2^5 */
void sdr_gen_bp_int32(uint64_t seed, uint64_t n, int32_t* out);
/* runs of Geometric(p = 1/mean_run, >= 1) length, run values u64 % value_mod */
void sdr_gen_runs_int64(uint64_t seed, uint64_t n, double mean_run, uint64_t value_mod,
                        int64_t* out);
/* lo + u32 % range */
void sdr_gen_uniform_int32(uint64_t seed, uint64_t n, int32_t lo, uint32_t range, int32_t* out);
/* days 8035 + u % 2557, sorted within each 2048-row group */
void sdr_gen_dates_sorted(uint64_t seed, uint64_t n, int32_t* out);
/* 99%: k/100 with k = 90000 + u % 10410001; 1%: u53 * 1e6 */
void sdr_gen_alp_doubles(uint64_t seed, uint64_t n, double* out);
/* u53 / 2^53 in [0, 1) */
void sdr_gen_unit_doubles(uint64_t seed, uint64_t n, double* out);
/* Strings of L ~ U[min_len, max_len] bytes: words of a seeded 2000-word
 * vocabulary (lowercase, 3-10 chars) drawn by Zipf(s) over rank, joined by
 * single spaces until >= L, truncated to L.  Two calls: with chars == NULL
 * the total byte count is returned and offsets[n+1] filled. */
uint64_t sdr_gen_zipf_strings(uint64_t seed, uint64_t n, uint32_t min_len, uint32_t max_len,
                              double zipf_s, uint64_t* offsets, uint8_t* chars);

#ifdef __cplusplus
}
#endif

#endif
