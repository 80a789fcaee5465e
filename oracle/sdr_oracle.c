/*
 * sdr_oracle.c — CPU restatement of the DuckDB segment decoders.
 *
 * TEST INFRASTRUCTURE ONLY (see sdr_oracle.h): the parity checker for the
 * HIP path, never linked into the product library.
 *
 * Every decoder follows the reference's kernel semantics, including its
 * malformed-input rules (rows of a malformed group, vector or segment decode
 * to zero).  Citations are to ducndh/sirius.  Where the reference reads
 * memory it never bounded (undefined behaviour on malformed bytes only), this
 * restatement reads the missing bytes as zero and says so at the site; no
 * well-formed segment reaches those cases.
 *
 * Build: plain C11, -O2 -ffp-contract=off (ALP products must not be fused).
 */
#include "sdr_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* byte helpers                                                              */
/* ------------------------------------------------------------------------ */

static uint16_t ld16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }
static uint32_t ld32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t ld64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }

static uint64_t ld_t(const uint8_t* p, uint32_t ts)
{
  switch (ts) {
    case 1: return p[0];
    case 2: return ld16(p);
    case 4: return ld32(p);
    default: return ld64(p);
  }
}

static void st_t(uint8_t* p, uint32_t ts, uint64_t v)
{
  switch (ts) {
    case 1: p[0] = (uint8_t)v; break;
    case 2: { uint16_t x = (uint16_t)v; memcpy(p, &x, 2); break; }
    case 4: { uint32_t x = (uint32_t)v; memcpy(p, &x, 4); break; }
    default: memcpy(p, &v, 8); break;
  }
}

static uint64_t mask_t(uint32_t ts) { return ts >= 8 ? ~0ull : ((1ull << (8u * ts)) - 1u); }

/* Read `w` bits (0..64) starting at bit `bitpos` of an LSB-first stream.
 * Bytes at or past `avail` read as zero.  Equivalent to the reference's
 * word-wise unpack_value<T> (src/include/cuda/scan/detail/bit_unpack.cuh:48-82)
 * on little-endian data, written byte-wise so the two formulations check each
 * other. */
static uint64_t read_bits(const uint8_t* p, uint64_t avail, uint64_t bitpos, uint32_t w)
{
  if (w == 0) return 0;
  uint64_t first = bitpos >> 3;
  uint32_t off   = (uint32_t)(bitpos & 7u);
  uint64_t last  = (bitpos + w - 1) >> 3;
  uint32_t n     = (uint32_t)(last - first + 1); /* <= 9 */
  uint64_t lo    = 0;
  for (uint32_t b = 0; b < n && b < 8; ++b) {
    uint64_t idx = first + b;
    uint64_t byte = idx < avail ? p[idx] : 0;
    lo |= byte << (8u * b);
  }
  uint64_t v = lo >> off;
  if (n == 9) {
    uint64_t idx  = first + 8;
    uint64_t byte = idx < avail ? p[idx] : 0;
    v |= byte << (64u - off);
  }
  if (w < 64) v &= (1ull << w) - 1u;
  return v;
}

/* DuckDB's BitpackingPrimitives::GetRequiredSize: values padded to a
 * multiple of 32, then width bits each (src/cuda/scan/gpu_decode_alp.cu:235-245). */
static uint64_t bp_required_bytes(uint64_t count, uint32_t width)
{
  if (width == 0) return 0;
  uint64_t padded = (count + 31u) / 32u * 32u;
  return padded * width / 8u;
}

/* ------------------------------------------------------------------------ */
/* BITPACKING                                                                */
/* follows src/cuda/scan/gpu_decode_bitpacking.cu:17-42 (layout) and        */
/* :119-213 (per-group parse) and :222-348 (modes)                           */
/* ------------------------------------------------------------------------ */

enum { BP_INVALID = 0, BP_AUTO = 1, BP_CONSTANT = 2, BP_CONSTANT_DELTA = 3, BP_DELTA_FOR = 4, BP_FOR = 5 };
#define BP_GROUP 2048u

static void bp_decode_group(const uint8_t* seg, uint32_t bytes, uint32_t k, uint32_t grc,
                            uint32_t ts, uint8_t* out)
{
  uint64_t tmask = mask_t(ts);
  int mode = BP_INVALID;
  uint32_t width = 0, rc = 0;
  uint64_t frame = 0, aux = 0, data_offset = 0;

  uint64_t metadata_end = bytes >= 8 ? ld64(seg) : 0;
  int metadata_ok = metadata_end >= 8u + ((uint64_t)k + 1u) * 4u && metadata_end <= bytes;
  if (metadata_ok) {
    uint32_t encoded  = ld32(seg + metadata_end - ((uint64_t)k + 1u) * 4u);
    uint32_t data_off = encoded & 0x00FFFFFFu;
    uint32_t pm       = (encoded >> 24) & 0xFFu;
    rc                = grc;
    if ((uint64_t)data_off + 2u * ts <= metadata_end) {
      uint64_t v0 = ld_t(seg + data_off, ts);
      uint64_t v1 = ld_t(seg + data_off + ts, ts);
      switch (pm) {
        case BP_CONSTANT: mode = BP_CONSTANT; aux = v0; break;
        case BP_CONSTANT_DELTA: mode = BP_CONSTANT_DELTA; frame = v0; aux = v1; break;
        case BP_FOR:
          mode = BP_FOR; frame = v0; width = (uint32_t)v1;
          data_offset = (uint64_t)data_off + 2u * ts;
          break;
        case BP_DELTA_FOR:
          if ((uint64_t)data_off + 3u * ts > metadata_end) break;
          mode = BP_DELTA_FOR; frame = v0; width = (uint32_t)v1;
          aux = ld_t(seg + data_off + 2u * ts, ts);
          data_offset = (uint64_t)data_off + 3u * ts;
          break;
        default: break;
      }
      if (width > ts * 8u) { mode = BP_INVALID; width = 0; }
      if (mode == BP_FOR || mode == BP_DELTA_FOR) {
        uint64_t words = ((uint64_t)rc * width + 31u) / 32u;
        if (data_offset + words * 4u > metadata_end) { mode = BP_INVALID; width = 0; }
      }
    }
  }

  switch (mode) {
    case BP_CONSTANT:
      for (uint32_t i = 0; i < rc; ++i) st_t(out + (uint64_t)i * ts, ts, aux);
      return;
    case BP_CONSTANT_DELTA:
      for (uint32_t i = 0; i < rc; ++i)
        st_t(out + (uint64_t)i * ts, ts, (frame + (((uint64_t)i * aux) & tmask)) & tmask);
      return;
    case BP_FOR: {
      const uint8_t* packed = seg + data_offset;
      uint64_t avail = metadata_end - data_offset;
      for (uint32_t i = 0; i < rc; ++i) {
        uint64_t u = read_bits(packed, avail, (uint64_t)i * width, width);
        st_t(out + (uint64_t)i * ts, ts, (frame + u) & tmask);
      }
      return;
    }
    case BP_DELTA_FOR: {
      /* out[i] = delta_offset + sum_{j<=i} (frame + packed_j), wrapping in T
       * (gpu_decode_bitpacking.cu:283-348; the test restatement is
       * test/cpp/scan/test_gpu_decode_bitpacking.cpp:532-537). */
      const uint8_t* packed = seg + data_offset;
      uint64_t avail = metadata_end - data_offset;
      uint64_t running = aux;
      for (uint32_t i = 0; i < rc; ++i) {
        uint64_t u = read_bits(packed, avail, (uint64_t)i * width, width);
        running = (running + frame + u) & tmask;
        st_t(out + (uint64_t)i * ts, ts, running);
      }
      return;
    }
    default:
      /* INVALID / unknown mode / failed bound: zero-fill the trusted rows. */
      memset(out, 0, (size_t)grc * ts);
      return;
  }
}

static void bp_decode_segment(const uint8_t* seg, uint32_t bytes, uint32_t rows, uint32_t ts,
                              uint8_t* out)
{
  uint32_t groups = (rows + BP_GROUP - 1u) / BP_GROUP;
  for (uint32_t k = 0; k < groups; ++k) {
    uint32_t grc = rows - k * BP_GROUP;
    if (grc > BP_GROUP) grc = BP_GROUP;
    bp_decode_group(seg, bytes, k, grc, ts, out + (uint64_t)k * BP_GROUP * ts);
  }
}

/* ------------------------------------------------------------------------ */
/* RLE                                                                       */
/* follows src/cuda/scan/gpu_decode_rle.cu:17-36 (layout), :140-171 (header  */
/* parse), :255-316 (count walk: first exact match wins, a zero count or an  */
/* overshoot in a 2048-entry tile without a match is malformed, the walk is  */
/* capped at 90112 entries) and :327-413 (run expansion)                     */
/* ------------------------------------------------------------------------ */

#define RLE_HEADER 8u
#define RLE_TILE 2048u
#define RLE_MAX_ENTRIES 90112u

static void rle_decode_segment(const uint8_t* seg, uint32_t bytes, uint32_t rows, uint32_t ts,
                               uint8_t* out)
{
  if (rows == 0) return;
  int malformed = 0;
  uint64_t counts_off = 0;
  uint32_t ec = 0;
  if (bytes < RLE_HEADER) {
    malformed = 1;
  } else {
    counts_off = ld64(seg);
    if (counts_off < RLE_HEADER || counts_off >= bytes) malformed = 1;
  }
  if (!malformed) {
    uint64_t values_bytes = counts_off - RLE_HEADER;
    uint64_t counts_bytes = bytes - counts_off;
    uint64_t n_real = (values_bytes % ts == 0) ? values_bytes / ts : counts_bytes / 2u;
    uint64_t n_max = n_real < RLE_MAX_ENTRIES ? n_real : RLE_MAX_ENTRIES;
    /* The reference reads counts[] up to n_max without bounding by the
     * segment; counts past bytes_size are not visited here (the walk then
     * ends without a match, which the reference also reports as malformed
     * unless out-of-bounds bytes happened to complete the sum). */
    if (n_max > counts_bytes / 2u) n_max = counts_bytes / 2u;
    const uint8_t* counts = seg + counts_off;
    uint32_t run = 0; /* 32-bit running sum, as the reference's uint32 scan */
    int64_t match = -1;
    malformed = 1;
    for (uint64_t t0 = 0; t0 < n_max; t0 += RLE_TILE) {
      uint64_t t1 = t0 + RLE_TILE < n_max ? t0 + RLE_TILE : n_max;
      int zero = 0, over = 0;
      for (uint64_t i = t0; i < t1; ++i) {
        uint32_t c = ld16(counts + 2u * i);
        if (c == 0) zero = 1;
        run += c;
        if (match < 0) {
          if (run == rows) match = (int64_t)i;
          else if (run > rows) over = 1;
        }
      }
      if (match >= 0) { ec = (uint32_t)(match + 1); malformed = 0; break; }
      if (zero || over) break;
    }
    /* Values of the matched entries must lie inside the segment (only a
     * counts_offset that is not a multiple of type_size can break this). */
    if (!malformed && RLE_HEADER + (uint64_t)ec * ts > bytes) malformed = 1;
  }
  if (malformed) {
    memset(out, 0, (size_t)rows * ts);
    return;
  }
  const uint8_t* values = seg + RLE_HEADER;
  const uint8_t* counts = seg + counts_off;
  uint32_t j = 0;
  uint32_t end = ld16(counts); /* prefix[0] */
  for (uint32_t r = 0; r < rows; ++r) {
    while (end <= r) { /* first run whose inclusive prefix exceeds r */
      ++j;
      end += ld16(counts + 2u * j);
    }
    memcpy(out + (uint64_t)r * ts, values + (uint64_t)j * ts, ts);
  }
}

/* ------------------------------------------------------------------------ */
/* ALP / ALPRD                                                               */
/* follows src/cuda/scan/gpu_decode_alp.cu:17-72 (layouts), :151-217 (FACT  */
/* and FRAC tables, DuckDB's AlpConstants), :251-314 (ALP parse), :368-393   */
/* (ALP decode + exceptions), :426-504 (ALPRD parse), :578-625 (ALPRD decode)*/
/* ------------------------------------------------------------------------ */

#define ALP_VECTOR 1024u
#define ALP_HEADER 4u
#define ALP_VEC_HEADER 13u
#define ALP_UNCOMPRESSED 0xFFu
#define ALP_MAX_FACTOR 18u
#define ALPRD_HEADER 7u
#define ALPRD_MAX_LEFT_BW 3u
#define ALPRD_MAX_DICT 8u
#define ALPRD_UNCOMPRESSED 0xFFFFu

static const int64_t ALP_FACT[19] = {
  1LL, 10LL, 100LL, 1000LL, 10000LL, 100000LL, 1000000LL, 10000000LL, 100000000LL,
  1000000000LL, 10000000000LL, 100000000000LL, 1000000000000LL, 10000000000000LL,
  100000000000000LL, 1000000000000000LL, 10000000000000000LL, 100000000000000000LL,
  1000000000000000000LL};
static const float ALP_FRAC_F32[11] = {1.0F, 1e-1F, 1e-2F, 1e-3F, 1e-4F, 1e-5F,
                                       1e-6F, 1e-7F, 1e-8F, 1e-9F, 1e-10F};
static const double ALP_FRAC_F64[21] = {1.0,   1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                        1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                        1e-14, 1e-15, 1e-16, 1e-17, 1e-18, 1e-19, 1e-20};

static void alp_decode_vector(const uint8_t* seg, uint32_t bytes, uint32_t v, uint32_t fill,
                              uint32_t ts, uint8_t* out)
{
  uint32_t tbits = ts * 8u;
  uint32_t max_exp = ts == 4 ? 10u : 18u;
  if (bytes < ALP_HEADER) goto invalid;
  {
    uint64_t metadata_end = ld32(seg);
    uint64_t trailer_need = ((uint64_t)v + 2u) * 4u;
    if (metadata_end > bytes || metadata_end < trailer_need) goto invalid;
    uint64_t data_off = ld32(seg + metadata_end - ((uint64_t)v + 1u) * 4u);
    if (data_off < ALP_HEADER || data_off >= metadata_end) goto invalid;
    const uint8_t* vp = seg + data_off;
    uint32_t exp = vp[0];
    if (exp == ALP_UNCOMPRESSED) {
      if (data_off + 1u + (uint64_t)fill * ts > metadata_end) goto invalid;
      memcpy(out, vp + 1, (size_t)fill * ts);
      return;
    }
    if (data_off + ALP_VEC_HEADER > metadata_end) goto invalid;
    uint32_t fac = vp[1];
    uint32_t ec = ld16(vp + 2);
    uint64_t for_v = ld64(vp + 4);
    uint32_t bw = vp[12];
    if (bw > tbits) goto invalid;
    if (exp > max_exp || fac > ALP_MAX_FACTOR) goto invalid;
    uint64_t packed_bytes = bp_required_bytes(fill, bw);
    uint64_t payload_end = data_off + ALP_VEC_HEADER + packed_bytes + (uint64_t)ec * (ts + 2u);
    if (payload_end > metadata_end) goto invalid;
    const uint8_t* packed = vp + ALP_VEC_HEADER;
    for (uint32_t i = 0; i < fill; ++i) {
      uint64_t u = read_bits(packed, packed_bytes, (uint64_t)i * bw, bw);
      int64_t enc = (int64_t)(u + for_v); /* FOR add wraps unsigned, then int64 */
      if (ts == 8) {
        double d = (double)enc * (double)ALP_FACT[fac] * ALP_FRAC_F64[exp];
        memcpy(out + (uint64_t)i * 8u, &d, 8);
      } else {
        float f = (float)enc * (float)ALP_FACT[fac] * ALP_FRAC_F32[exp];
        memcpy(out + (uint64_t)i * 4u, &f, 4);
      }
    }
    const uint8_t* exc = packed + packed_bytes;
    const uint8_t* pos = exc + (uint64_t)ec * ts;
    for (uint32_t e = 0; e < ec; ++e) { /* patched in order: the last one wins */
      uint32_t p = ld16(pos + 2u * e);
      if (p < fill) memcpy(out + (uint64_t)p * ts, exc + (uint64_t)e * ts, ts);
    }
    return;
  }
invalid:
  memset(out, 0, (size_t)fill * ts);
}

static void alprd_decode_vector(const uint8_t* seg, uint32_t bytes, uint32_t v, uint32_t fill,
                                uint32_t ts, uint8_t* out)
{
  uint32_t tbits = ts * 8u;
  if (bytes < ALPRD_HEADER) goto invalid;
  {
    uint64_t metadata_end = ld32(seg);
    uint64_t trailer_need = ALPRD_HEADER + ((uint64_t)v + 1u) * 4u;
    if (metadata_end > bytes || metadata_end < trailer_need) goto invalid;
    uint32_t rbw = seg[4], lbw = seg[5], dsz = seg[6];
    if (rbw > tbits || lbw > ALPRD_MAX_LEFT_BW || dsz > ALPRD_MAX_DICT) goto invalid;
    uint64_t dict_end = ALPRD_HEADER + (uint64_t)dsz * 2u;
    if (dict_end > metadata_end) goto invalid;
    uint64_t data_off = ld32(seg + metadata_end - ((uint64_t)v + 1u) * 4u);
    if (data_off < dict_end || data_off + 2u > metadata_end) goto invalid;
    const uint8_t* vp = seg + data_off;
    uint32_t ec = ld16(vp);
    if (ec == ALPRD_UNCOMPRESSED) {
      if (data_off + 2u + (uint64_t)fill * ts > metadata_end) goto invalid;
      memcpy(out, vp + 2, (size_t)fill * ts);
      return;
    }
    uint64_t lpb = bp_required_bytes(fill, lbw);
    uint64_t rpb = bp_required_bytes(fill, rbw);
    uint64_t payload_end = data_off + 2u + lpb + rpb + (uint64_t)ec * 4u;
    if (payload_end > metadata_end) goto invalid;
    uint16_t dict[ALPRD_MAX_DICT] = {0}; /* unused slots read as 0 */
    for (uint32_t d = 0; d < dsz; ++d) dict[d] = ld16(seg + ALPRD_HEADER + 2u * d);
    const uint8_t* left = vp + 2;
    const uint8_t* right = left + lpb;
    uint64_t tmask = mask_t(ts);
    int full_right = rbw >= tbits;
    uint64_t r_mask = full_right ? tmask : ((1ull << rbw) - 1u);
    for (uint32_t i = 0; i < fill; ++i) {
      uint64_t r = read_bits(right, rpb, (uint64_t)i * rbw, rbw);
      uint64_t bits;
      if (full_right) {
        bits = r;
      } else {
        uint32_t li = (uint32_t)read_bits(left, lpb, (uint64_t)i * lbw, lbw);
        bits = (((uint64_t)dict[li] << rbw) | r) & tmask;
      }
      st_t(out + (uint64_t)i * ts, ts, bits);
    }
    const uint8_t* exc = right + rpb;
    const uint8_t* pos = exc + (uint64_t)ec * 2u;
    for (uint32_t e = 0; e < ec; ++e) {
      uint32_t p = ld16(pos + 2u * e);
      if (p >= fill) continue;
      uint64_t existing = ld_t(out + (uint64_t)p * ts, ts);
      uint64_t right_part = existing & r_mask;
      uint64_t bits = full_right ? right_part
                                 : ((((uint64_t)ld16(exc + 2u * e)) << rbw) | right_part) & tmask;
      st_t(out + (uint64_t)p * ts, ts, bits);
    }
    return;
  }
invalid:
  memset(out, 0, (size_t)fill * ts);
}

static void alp_decode_segment(int rd, const uint8_t* seg, uint32_t bytes, uint32_t rows,
                               uint32_t ts, uint8_t* out)
{
  uint32_t nv = (rows + ALP_VECTOR - 1u) / ALP_VECTOR;
  for (uint32_t v = 0; v < nv; ++v) {
    uint32_t fill = rows - v * ALP_VECTOR;
    if (fill > ALP_VECTOR) fill = ALP_VECTOR;
    uint8_t* o = out + (uint64_t)v * ALP_VECTOR * ts;
    if (rd) alprd_decode_vector(seg, bytes, v, fill, ts, o);
    else alp_decode_vector(seg, bytes, v, fill, ts, o);
  }
}

/* ------------------------------------------------------------------------ */
/* FSST                                                                      */
/* follows src/cuda/scan/strings/fsst.cu:17-44 (layout), :69-87 (header      */
/* check), :132-188 (compressed-length prefix), :194-286 (decoded lengths),  */
/* src/include/cuda/scan/detail/fsst.cuh:50-96 (symbol-table import, a port  */
/* of DuckDB's duckdb_fsst_import) and DuckDB's duckdb_fsst_decompress.      */
/* ------------------------------------------------------------------------ */

#define FSST_HEADER 16u
#define FSST_ESC 255u
#define FSST_NSYM 255u
#define FSST_SYMTAB_MAX 8192u
#define FSST_VERSION 20190218ull
#define FSST_CORRUPT 32774747032022883ull

typedef struct fsst_table {
  uint8_t len[256];
  uint64_t sym[256];
} fsst_table;

typedef struct fsst_hdr {
  uint32_t dict_size, dict_end, width, symtab_off;
} fsst_hdr;

static int fsst_parse_header(const uint8_t* seg, uint32_t bytes, fsst_hdr* h)
{
  if (bytes < FSST_HEADER) return 0;
  h->dict_size = ld32(seg);
  h->dict_end = ld32(seg + 4);
  h->width = ld32(seg + 8);
  h->symtab_off = ld32(seg + 12);
  return h->dict_end <= bytes && h->symtab_off < h->dict_end && h->width <= 32u;
}

/* Symbol-table import.  Bytes past the segment (or past the 8 KiB blob the
 * reference stages) read as zero; the reference reads uninitialised shared
 * memory there.  A length histogram describing more than 255 symbols stops
 * at code 254; the reference writes past its table there. */
static void fsst_import(const uint8_t* seg, uint32_t bytes, uint32_t off, fsst_table* t)
{
  uint32_t avail = bytes - off;
  if (avail > FSST_SYMTAB_MAX) avail = FSST_SYMTAB_MAX;
  const uint8_t* b = seg + off;
#define BYTE_AT(i) ((uint32_t)(i) < avail ? b[(i)] : 0u)
  memset(t, 0, sizeof(*t));
  uint64_t version = 0;
  for (uint32_t i = 0; i < 8; ++i) version |= (uint64_t)BYTE_AT(i) << (8u * i);
  if ((version >> 32) != FSST_VERSION) return; /* every code decodes to nothing */
  uint32_t zero_term = BYTE_AT(8) & 1u;
  uint8_t histo[8];
  for (uint32_t i = 0; i < 8; ++i) histo[i] = (uint8_t)BYTE_AT(9 + i);
  t->len[0] = 1;
  t->sym[0] = 0;
  uint32_t code = zero_term;
  if (zero_term) histo[0]--;
  uint32_t pos = 17;
  for (uint32_t l = 1; l <= 8; ++l) {
    uint32_t hi = l & 7u;
    uint32_t sl = hi + 1u; /* 2,3,4,5,6,7,8,1 */
    for (uint32_t i = 0; i < histo[hi]; ++i, ++code) {
      uint64_t s = 0;
      for (uint32_t j = 0; j < sl; ++j) s |= (uint64_t)BYTE_AT(pos + j) << (8u * j);
      pos += sl;
      if (code < FSST_NSYM) {
        t->len[code] = (uint8_t)sl;
        t->sym[code] = s;
      }
    }
  }
  while (code < FSST_NSYM) {
    t->sym[code] = FSST_CORRUPT;
    t->len[code++] = 8;
  }
#undef BYTE_AT
}

/* Per-row inclusive prefix of the bit-packed compressed lengths (32-bit,
 * wrapping like the reference's scan).  Length bits past the segment read
 * as zero. */
static void fsst_comp_prefix(const uint8_t* seg, uint32_t bytes, uint32_t rows, uint32_t w,
                             uint32_t* cum)
{
  uint32_t run = 0;
  for (uint32_t i = 0; i < rows; ++i) {
    run += (uint32_t)read_bits(seg + FSST_HEADER, bytes - FSST_HEADER, (uint64_t)i * w, w);
    cum[i] = run;
  }
}

/* Decode one row (DuckDB's duckdb_fsst_decompress order: a code emits its
 * symbol, an escape emits the next byte literally).  A trailing escape is
 * dropped, as the reference's length pass does (fsst.cu:273-281).  Returns
 * the decoded length; writes to `dst` when it is non-null. */
static uint32_t fsst_row(const fsst_table* t, const uint8_t* src, uint32_t n, uint8_t* dst)
{
  uint32_t pos = 0, o = 0;
  while (pos < n) {
    uint32_t code = src[pos++];
    if (code < FSST_ESC) {
      uint32_t l = t->len[code];
      if (dst) {
        uint64_t s = t->sym[code];
        for (uint32_t j = 0; j < l; ++j) dst[o + j] = (uint8_t)(s >> (8u * j));
      }
      o += l;
    } else if (pos < n) {
      if (dst) dst[o] = src[pos];
      ++pos;
      ++o;
    }
  }
  return o;
}

typedef struct fsst_seg_state {
  int ok;
  fsst_hdr h;
  fsst_table t;
} fsst_seg_state;

static void fsst_prepare(const uint8_t* seg, uint32_t bytes, fsst_seg_state* s)
{
  s->ok = fsst_parse_header(seg, bytes, &s->h);
  if (s->ok) fsst_import(seg, bytes, s->h.symtab_off, &s->t);
}

/* Row i is empty when its compressed range is not inside [0, dict_end)
 * (fsst.cu:240-244).  The reference's gather kernel does not repeat this
 * check and writes such a row's bytes over the following rows; here the row
 * stays empty in both passes. */
static int fsst_row_range(const fsst_seg_state* s, const uint32_t* cum, uint32_t i,
                          uint32_t* start, uint32_t* len)
{
  uint32_t c = cum[i];
  uint32_t p = i > 0 ? cum[i - 1] : 0;
  if (c > s->h.dict_end || p > c) return 0;
  *len = c - p;
  *start = s->h.dict_end - c;
  return *len > 0;
}

int oracle_fsst_lengths(const uint8_t* seg, uint32_t bytes_size, uint32_t rows, uint32_t* lengths)
{
  fsst_seg_state* s = (fsst_seg_state*)malloc(sizeof(fsst_seg_state));
  uint32_t* cum = (uint32_t*)malloc(sizeof(uint32_t) * (rows ? rows : 1));
  if (!s || !cum) { free(s); free(cum); return SDR_E_CAPACITY; }
  fsst_prepare(seg, bytes_size, s);
  if (!s->ok) {
    memset(lengths, 0, sizeof(uint32_t) * rows);
  } else {
    fsst_comp_prefix(seg, bytes_size, rows, s->h.width, cum);
    for (uint32_t i = 0; i < rows; ++i) {
      uint32_t st, n;
      lengths[i] = fsst_row_range(s, cum, i, &st, &n) ? fsst_row(&s->t, seg + st, n, NULL) : 0;
    }
  }
  free(cum);
  free(s);
  return SDR_OK;
}

int oracle_fsst_decode(const uint8_t* seg, uint32_t bytes_size, uint32_t rows, uint32_t* lengths,
                       uint8_t* chars, uint64_t chars_cap)
{
  fsst_seg_state* s = (fsst_seg_state*)malloc(sizeof(fsst_seg_state));
  uint32_t* cum = (uint32_t*)malloc(sizeof(uint32_t) * (rows ? rows : 1));
  if (!s || !cum) { free(s); free(cum); return SDR_E_CAPACITY; }
  fsst_prepare(seg, bytes_size, s);
  uint64_t o = 0;
  int rc = SDR_OK;
  if (!s->ok) {
    memset(lengths, 0, sizeof(uint32_t) * rows);
  } else {
    fsst_comp_prefix(seg, bytes_size, rows, s->h.width, cum);
    for (uint32_t i = 0; i < rows; ++i) {
      uint32_t st, n, l = 0;
      if (fsst_row_range(s, cum, i, &st, &n)) {
        l = fsst_row(&s->t, seg + st, n, NULL);
        if (o + l > chars_cap) { rc = SDR_E_OVERFLOW; break; }
        fsst_row(&s->t, seg + st, n, chars + o);
      }
      lengths[i] = l;
      o += l;
    }
  }
  free(cum);
  free(s);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* DICTIONARY and UNCOMPRESSED strings                                       */
/* ------------------------------------------------------------------------ */

/* One row of a plain (non-FSST) string segment: its length and the segment
 * offset of its first byte.  Bytes of the row at or past bytes_size read as
 * zero when copied (plain_copy).
 *
 * DICTIONARY follows src/cuda/scan/strings/dictionary.cu:17-33 (layout),
 * :57-64 (header: dict_size, dict_end, index_buffer_offset,
 * index_buffer_count, bitpacking_width; valid if the index buffer and
 * dict_end fit and the width is <= 32), :84-108 (lengths: selection value
 * sel from the bit-packed buffer at byte 20; 0 means NULL, sel >=
 * index_buffer_count is empty, else idx[sel] - idx[sel-1]) and :117-140
 * (bytes at dict_end - idx[sel]).
 *
 * UNCOMPRESSED follows src/cuda/scan/strings/uncompressed.cu:17-33 (layout),
 * :56-81 (segment valid if bytes_size >= 8 + 4 rows; length |off[i]| -
 * |off[i-1]| of the int32 offsets at byte 8) and :88-120 (bytes at dict_end -
 * |off[i]|, dict_end a u32 at byte 4).
 *
 * Deliberate divergence, malformed input only: a row whose end passes
 * dict_end or whose end falls below its start is empty.  The reference's
 * gather would copy from outside the segment there (and, for DICTIONARY
 * with sel >= index_buffer_count, copy bytes its length pass counted as
 * zero). */
typedef struct plain_seg {
  uint32_t codec, ok, dict_end, ibo, ibc, width;
} plain_seg;

static plain_seg plain_prepare(uint32_t codec, const uint8_t* seg, uint32_t size, uint32_t rows)
{
  plain_seg h;
  memset(&h, 0, sizeof(h));
  h.codec = codec;
  if (codec == SDR_CODEC_DICTIONARY) {
    if (size < 20u) return h;
    h.dict_end = ld32(seg + 4);
    h.ibo      = ld32(seg + 8);
    h.ibc      = ld32(seg + 12);
    h.width    = ld32(seg + 16);
    h.ok = (uint64_t)h.ibo + 4ull * h.ibc <= size && h.dict_end <= size && h.width <= 32u;
  } else {
    h.ok = (uint64_t)size >= 8ull + 4ull * rows;
    h.dict_end = h.ok ? ld32(seg + 4) : 0u;
  }
  return h;
}

static uint32_t abs32(int32_t v) { return v < 0 ? (uint32_t)(-(int64_t)v) : (uint32_t)v; }

static uint32_t plain_row(const plain_seg* h, const uint8_t* seg, uint32_t size, uint32_t i, uint32_t* start)
{
  uint32_t e0, e1;
  *start = 0;
  if (!h->ok) return 0;
  if (h->codec == SDR_CODEC_DICTIONARY) {
    uint32_t sel = (uint32_t)read_bits(seg + 20, size - 20u, (uint64_t)i * h->width, h->width);
    if (sel == 0 || sel >= h->ibc) return 0;
    e1 = ld32(seg + h->ibo + 4ull * sel);
    e0 = ld32(seg + h->ibo + 4ull * (sel - 1u));
  } else {
    e1 = abs32((int32_t)ld32(seg + 8 + 4ull * i));
    e0 = i ? abs32((int32_t)ld32(seg + 8 + 4ull * (i - 1u))) : 0u;
  }
  if (e1 > h->dict_end || e1 < e0) return 0;
  *start = h->dict_end - e1;
  return e1 - e0;
}

static void plain_lengths(uint32_t codec, const uint8_t* seg, uint32_t size, uint32_t rows, uint32_t* lengths)
{
  plain_seg h = plain_prepare(codec, seg, size, rows);
  uint32_t st;
  for (uint32_t i = 0; i < rows; ++i) lengths[i] = plain_row(&h, seg, size, i, &st);
}

static void plain_decode(uint32_t codec, const uint8_t* seg, uint32_t size, uint32_t rows, uint8_t* chars,
                         uint64_t cap)
{
  plain_seg h = plain_prepare(codec, seg, size, rows);
  uint64_t o = 0;
  for (uint32_t i = 0; i < rows; ++i) {
    uint32_t st, l = plain_row(&h, seg, size, i, &st);
    for (uint32_t k = 0; k < l && o < cap; ++k, ++o) {
      uint64_t at = (uint64_t)st + k;
      chars[o] = at < size ? seg[at] : 0u; /* dict_end of UNCOMPRESSED is not bounded by the reference */
    }
  }
}

/* ------------------------------------------------------------------------ */
/* DICT_FSST strings                                                         */
/* ------------------------------------------------------------------------ */

/* DICT_FSST follows src/cuda/scan/strings/dict_fsst.cu:17-38 (layout and
 * modes), :64-72 (16-byte header: dict_size, dict_count, mode,
 * string_lengths_width, dictionary_indices_width, pad, symbol_table_size),
 * :516-525 (region offsets, all 8-byte aligned: dict at 16, symbol table at
 * align8(16 + dict_size), string lengths at 16 + align8(dict_size) in mode 0
 * or align8(symtab + symtab_size), indices at align8(lengths + their bytes);
 * a segment is valid if it has rows, 16 header bytes, mode <= 2 and, unless
 * mode 2, its index region starts inside it), :160-185 (entry k of the
 * dictionary is [byte_off[k], byte_off[k+1]) of the dict region, byte_off
 * the scanned bit-packed string lengths), :190-235 (decoded offsets: mode 0,
 * or a dictionary of at most one entry, copies byte_off; the FSST modes give
 * entry 0 (NULL) length 0 and every other entry the length its compressed
 * bytes walk to, an escape counting one byte and skipping the next even when
 * it is the entry's last byte), :238-243 (a segment has inline NULLs when it
 * is not mode 2, has more than one entry and entry 0 is empty), :246-265 (row
 * lengths: index = the bit-packed dictionary index, or row + 1 in mode 2;
 * length dec_off[idx+1] - dec_off[idx] when 0 < idx < dict_count, else 0),
 * :268-320 (mode 1 predecodes each entry once; a trailing escape writes no
 * byte), :324-400 (gather: mode 0 copies raw entry bytes, modes 1 and 2 the
 * decoded entry) and :426-446 (with inline NULLs anywhere in the column,
 * every row of a segment that is not mode 2, invalid segments included, whose
 * index is 0 is NULL).  Arithmetic on the header is 32-bit as in the
 * reference.
 *
 * Where this build differs, malformed input only:
 *  - a lengths or indices width above 32 makes the segment invalid (the
 *    reference reads a 64-bit window there);
 *  - the byte a trailing escape counts is written as 0 (the reference leaves
 *    it unwritten); parity unpinned: DuckDB's encoder never ends an entry
 *    with an escape;
 *  - a row whose index is >= dict_count decodes empty in both passes (the
 *    reference's gather copies the bytes of whatever offsets follow), and an
 *    entry whose end is below its start is empty;
 *  - bytes past the segment read as zero. */
typedef struct dfsst_seg {
  int ok, has_fsst, inline_null;
  uint32_t mode, dict_count, diw, off_dict, off_didx;
  uint32_t* E;         /* dict_count + 1 entry offsets into src */
  const uint8_t* src;  /* dict region (mode 0 / no table) or decoded entries */
  uint64_t src_size;
  uint8_t* pre;        /* decoded entries, owned */
} dfsst_seg;

static uint32_t al8(uint32_t v) { return (v + 7u) & ~7u; }

static void dfsst_free(dfsst_seg* d)
{
  free(d->E);
  free(d->pre);
  memset(d, 0, sizeof(*d));
}

static int dfsst_prepare(const uint8_t* seg, uint32_t size, uint32_t rows, dfsst_seg* d)
{
  memset(d, 0, sizeof(*d));
  if (rows == 0 || size < 16u) return SDR_OK;
  uint32_t dict_size = ld32(seg), dict_count = ld32(seg + 4), symtab_size = ld32(seg + 12);
  uint32_t mode = seg[8], slw = seg[9], diw = seg[10];
  if (mode > 2u || slw > 32u || diw > 32u) return SDR_OK;
  uint32_t off_dict   = 16u;
  uint32_t off_symtab = al8(off_dict + dict_size);
  uint32_t off_slens  = mode == 0u ? off_dict + al8(dict_size) : al8(off_symtab + symtab_size);
  uint32_t slens_bits = dict_count * slw;
  uint32_t off_didx   = al8(off_slens + (slens_bits + 7u) / 8u);
  if (off_didx > size && mode != 2u) return SDR_OK;
  d->E = (uint32_t*)calloc((size_t)dict_count + 1u, sizeof(uint32_t));
  uint32_t* bo = (uint32_t*)calloc((size_t)dict_count + 1u, sizeof(uint32_t));
  if (!d->E || !bo) { free(bo); dfsst_free(d); return SDR_E_CAPACITY; }
  d->ok = 1;
  d->mode = mode;
  d->dict_count = dict_count;
  d->diw = diw;
  d->off_dict = off_dict;
  d->off_didx = off_didx;
  uint64_t sl_avail = off_slens < size ? size - off_slens : 0u;
  for (uint32_t k = 0; k < dict_count; ++k)
    bo[k + 1] = bo[k] + (uint32_t)read_bits(seg + (sl_avail ? off_slens : 0u), sl_avail, (uint64_t)k * slw, slw);
  d->has_fsst = mode != 0u && dict_count > 1u;
  d->inline_null = mode != 2u && dict_count > 1u && bo[1] == 0u;
  uint64_t dict_avail = off_dict < size ? size - off_dict : 0u;
  const uint8_t* dict = seg + off_dict;
#define DICT_AT(i) ((uint64_t)(i) < dict_avail ? dict[(i)] : 0u)
  if (!d->has_fsst) {
    memcpy(d->E, bo, sizeof(uint32_t) * ((size_t)dict_count + 1u));
    d->src = dict;
    d->src_size = dict_avail;
    free(bo);
    return SDR_OK;
  }
  fsst_table* t = (fsst_table*)malloc(sizeof(fsst_table));
  if (!t) { free(bo); dfsst_free(d); return SDR_E_CAPACITY; }
  if (off_symtab < size) fsst_import(seg, size, off_symtab, t);
  else memset(t, 0, sizeof(*t)); /* no table bytes: every code decodes to nothing */
  d->E[0] = 0;
  d->E[1] = 0; /* entry 0 is NULL */
  for (uint32_t k = 1; k < dict_count; ++k) {
    uint32_t c0 = bo[k], n = bo[k + 1] >= bo[k] ? bo[k + 1] - bo[k] : 0u;
    uint32_t pos = 0, dl = 0;
    while (pos < n) {
      uint32_t code = DICT_AT((uint64_t)c0 + pos);
      ++pos;
      if (code < FSST_ESC) dl += t->len[code];
      else { ++pos; ++dl; }
    }
    d->E[k + 1] = d->E[k] + dl;
  }
  uint64_t total = d->E[dict_count];
  d->pre = (uint8_t*)calloc(total ? total : 1u, 1);
  if (!d->pre) { free(t); free(bo); dfsst_free(d); return SDR_E_CAPACITY; }
  for (uint32_t k = 1; k < dict_count; ++k) {
    uint32_t c0 = bo[k], n = bo[k + 1] >= bo[k] ? bo[k + 1] - bo[k] : 0u;
    uint32_t pos = 0;
    uint64_t o = d->E[k], lim = d->E[k + 1] >= d->E[k] ? d->E[k + 1] : d->E[k];
    if (lim > total) lim = total; /* offsets that wrapped (malformed lengths) */
    while (pos < n) {
      uint32_t code = DICT_AT((uint64_t)c0 + pos);
      ++pos;
      if (code < FSST_ESC) {
        for (uint32_t j = 0; j < t->len[code] && o < lim; ++j) d->pre[o++] = (uint8_t)(t->sym[code] >> (8u * j));
      } else if (pos < n) {
        if (o < lim) d->pre[o++] = (uint8_t)DICT_AT((uint64_t)c0 + pos);
        ++pos;
      }
    }
  }
#undef DICT_AT
  d->src = d->pre;
  d->src_size = total;
  free(t);
  free(bo);
  return SDR_OK;
}

/* Dictionary index of segment row i (0, which is NULL, for an invalid segment). */
static uint32_t dfsst_index(const dfsst_seg* d, const uint8_t* seg, uint32_t size, uint32_t i)
{
  if (!d->ok) return 0;
  if (d->mode == 2u) return i + 1u;
  uint64_t avail = d->off_didx < size ? size - d->off_didx : 0u;
  return (uint32_t)read_bits(seg + (avail ? d->off_didx : 0u), avail, (uint64_t)i * d->diw, d->diw);
}

static uint32_t dfsst_row(const dfsst_seg* d, const uint8_t* seg, uint32_t size, uint32_t i, uint32_t* start)
{
  uint32_t idx = dfsst_index(d, seg, size, i);
  *start = 0;
  if (!d->ok || idx == 0u || idx >= d->dict_count || d->E[idx + 1] < d->E[idx]) return 0;
  *start = d->E[idx];
  return d->E[idx + 1] - d->E[idx];
}

int oracle_dict_fsst_decode(const uint8_t* seg, uint32_t bytes_size, uint32_t rows, uint32_t* lengths,
                            uint8_t* chars, uint64_t chars_cap, uint8_t* inline_null)
{
  dfsst_seg d;
  int rc = dfsst_prepare(seg, bytes_size, rows, &d);
  if (rc) return rc;
  uint64_t o = 0;
  for (uint32_t i = 0; i < rows; ++i) {
    uint32_t st, l = dfsst_row(&d, seg, bytes_size, i, &st);
    lengths[i] = l;
    if (chars)
      for (uint32_t k = 0; k < l && o < chars_cap; ++k, ++o)
        chars[o] = (uint64_t)st + k < d.src_size ? d.src[st + k] : 0u;
  }
  if (inline_null) *inline_null = (uint8_t)d.inline_null;
  dfsst_free(&d);
  return SDR_OK;
}

/* Rows of a DICT_FSST segment that are NULL when the column has inline
 * NULLs: index 0 outside mode 2 (every row of an invalid segment). */
static void dfsst_mark_nulls(const uint8_t* seg, uint32_t size, uint32_t rows, uint32_t row_offset, uint8_t* mask)
{
  dfsst_seg d;
  if (dfsst_prepare(seg, size, rows, &d)) return;
  if (!d.ok || d.mode != 2u)
    for (uint32_t i = 0; i < rows; ++i)
      if (dfsst_index(&d, seg, size, i) == 0u) {
        uint32_t r = row_offset + i;
        mask[r >> 3] &= (uint8_t)~(1u << (r & 7u));
      }
  dfsst_free(&d);
}

static int dfsst_has_inline_nulls(const uint8_t* seg, uint32_t size, uint32_t rows)
{
  dfsst_seg d;
  if (dfsst_prepare(seg, size, rows, &d)) return 0;
  int r = d.ok && d.inline_null;
  dfsst_free(&d);
  return r;
}

/* ------------------------------------------------------------------------ */
/* UNCOMPRESSED / CONSTANT / validity                                        */
/* follows src/cuda/scan/gpu_native_decode.cu:154-170 (CONSTANT), :221-286   */
/* (UNCOMPRESSED), :396-476 (validity overlay), :519-538 (all-valid mask)    */
/* ------------------------------------------------------------------------ */

static int type_ok_for(uint32_t codec, uint32_t ts)
{
  switch (codec) {
    case SDR_CODEC_BITPACKING:
    case SDR_CODEC_RLE: return ts == 1 || ts == 2 || ts == 4 || ts == 8;
    case SDR_CODEC_ALP:
    case SDR_CODEC_ALPRD: return ts == 4 || ts == 8;
    case SDR_CODEC_UNCOMPRESSED:
    case SDR_CODEC_CONSTANT: return ts == 1 || ts == 2 || ts == 4 || ts == 8 || ts == 16;
    default: return 0;
  }
}

static int codec_known_fixed(uint32_t codec)
{
  return codec == SDR_CODEC_BITPACKING || codec == SDR_CODEC_RLE || codec == SDR_CODEC_ALP ||
         codec == SDR_CODEC_ALPRD || codec == SDR_CODEC_UNCOMPRESSED ||
         codec == SDR_CODEC_CONSTANT;
}

int oracle_decode_segment(uint32_t codec, const uint8_t* seg, uint32_t bytes, uint32_t rows,
                          uint32_t ts, void* outv)
{
  uint8_t* out = (uint8_t*)outv;
  if (!codec_known_fixed(codec)) return SDR_E_UNSUPPORTED_CODEC;
  if (!type_ok_for(codec, ts)) return SDR_E_BAD_WIDTH;
  if (rows == 0) return SDR_OK;
  switch (codec) {
    case SDR_CODEC_BITPACKING: bp_decode_segment(seg, bytes, rows, ts, out); break;
    case SDR_CODEC_RLE: rle_decode_segment(seg, bytes, rows, ts, out); break;
    case SDR_CODEC_ALP: alp_decode_segment(0, seg, bytes, rows, ts, out); break;
    case SDR_CODEC_ALPRD: alp_decode_segment(1, seg, bytes, rows, ts, out); break;
    case SDR_CODEC_UNCOMPRESSED:
      if ((uint64_t)bytes < (uint64_t)rows * ts) return SDR_E_BOUNDS;
      memcpy(out, seg, (size_t)rows * ts);
      break;
    case SDR_CODEC_CONSTANT:
      if (bytes < ts) return SDR_E_BOUNDS;
      for (uint32_t i = 0; i < rows; ++i) memcpy(out + (uint64_t)i * ts, seg, ts);
      break;
  }
  return SDR_OK;
}

static int check_runs_bounds(const sdr_run* runs, uint32_t n, uint32_t total_rows)
{
  for (uint32_t r = 0; r < n; ++r) {
    if (runs[r].n_segments && !runs[r].segments) return SDR_E_INVALID_ARG;
    for (uint32_t s = 0; s < runs[r].n_segments; ++s) {
      const sdr_segment* g = &runs[r].segments[s];
      if ((uint64_t)g->row_offset + g->row_count > total_rows) return SDR_E_BOUNDS;
      if (g->row_count && !g->d_bytes) return SDR_E_INVALID_ARG;
    }
  }
  return SDR_OK;
}

static int check_validity(const sdr_run* runs, uint32_t n)
{
  for (uint32_t r = 0; r < n; ++r) {
    if (runs[r].codec != SDR_CODEC_UNCOMPRESSED) return SDR_E_UNSUPPORTED_CODEC;
    for (uint32_t s = 0; s < runs[r].n_segments; ++s) {
      const sdr_segment* g = &runs[r].segments[s];
      if (g->row_count == 0) continue;
      if (g->row_offset % 8u) return SDR_E_ALIGNMENT;
      if (g->bytes_size < (g->row_count + 7u) / 8u) return SDR_E_BOUNDS;
    }
  }
  return SDR_OK;
}

static int64_t build_mask(const sdr_run* runs, uint32_t n, uint32_t total_rows, uint8_t* mask)
{
  uint32_t nbytes = (total_rows + 7u) / 8u;
  memset(mask, 0xFF, nbytes);
  for (uint32_t r = 0; r < n; ++r)
    for (uint32_t s = 0; s < runs[r].n_segments; ++s) {
      const sdr_segment* g = &runs[r].segments[s];
      if (g->row_count == 0) continue;
      memcpy(mask + g->row_offset / 8u, g->d_bytes, (g->row_count + 7u) / 8u);
    }
  int64_t nulls = 0;
  for (uint32_t i = 0; i < total_rows; ++i) nulls += !((mask[i >> 3] >> (i & 7u)) & 1u);
  return nulls;
}

/* ------------------------------------------------------------------------ */
/* column-level entries (validation mirrors gpu_native_decode.cu:546-608)   */
/* ------------------------------------------------------------------------ */

static int validate_fixed(const sdr_column* cols, uint32_t n_cols)
{
  if (n_cols && !cols) return SDR_E_INVALID_ARG;
  for (uint32_t c = 0; c < n_cols; ++c) {
    const sdr_column* col = &cols[c];
    if (col->total_rows != cols[0].total_rows) return SDR_E_BOUNDS;
    if (col->total_rows > 0x7FFFFFFFu) return SDR_E_BOUNDS;
    if (!(col->type_size == 1 || col->type_size == 2 || col->type_size == 4 ||
          col->type_size == 8 || col->type_size == 16))
      return SDR_E_BAD_WIDTH;
    if (col->n_data && !col->data) return SDR_E_INVALID_ARG;
    if (col->n_validity && !col->validity) return SDR_E_INVALID_ARG;
    if (col->total_rows && !col->d_out) return SDR_E_INVALID_ARG;
    int rc = check_runs_bounds(col->data, col->n_data, col->total_rows);
    if (rc) return rc;
    for (uint32_t r = 0; r < col->n_data; ++r) {
      if (!codec_known_fixed(col->data[r].codec)) return SDR_E_UNSUPPORTED_CODEC;
      if (!type_ok_for(col->data[r].codec, col->type_size)) return SDR_E_BAD_WIDTH;
      for (uint32_t s = 0; s < col->data[r].n_segments; ++s) {
        const sdr_segment* g = &col->data[r].segments[s];
        if (!g->row_count) continue;
        if (col->data[r].codec == SDR_CODEC_UNCOMPRESSED &&
            (uint64_t)g->bytes_size < (uint64_t)g->row_count * col->type_size)
          return SDR_E_BOUNDS;
        if (col->data[r].codec == SDR_CODEC_CONSTANT && g->bytes_size < col->type_size)
          return SDR_E_BOUNDS;
      }
    }
    if (col->has_nulls && col->n_validity) {
      rc = check_runs_bounds(col->validity, col->n_validity, col->total_rows);
      if (rc) return rc;
      rc = check_validity(col->validity, col->n_validity);
      if (rc) return rc;
      if (col->total_rows && !col->d_null_mask) return SDR_E_INVALID_ARG;
    }
  }
  return SDR_OK;
}

int oracle_decode_table(sdr_column* cols, uint32_t n_cols) { return oracle_decode_table_mt(cols, n_cols, 1); }

/* ---- thread pool over segments ----------------------------------------- */

typedef struct seg_job {
  uint32_t codec, ts;
  const sdr_segment* g;
  uint8_t* out;
} seg_job;

typedef struct job_pool {
  seg_job* jobs;
  size_t n;
  atomic_size_t next;
} job_pool;

static void* job_worker(void* arg)
{
  job_pool* p = (job_pool*)arg;
  for (;;) {
    size_t i = atomic_fetch_add(&p->next, 1);
    if (i >= p->n) break;
    seg_job* j = &p->jobs[i];
    oracle_decode_segment(j->codec, j->g->d_bytes, j->g->bytes_size, j->g->row_count, j->ts,
                          j->out + (uint64_t)j->g->row_offset * j->ts);
  }
  return NULL;
}

static void run_pool(job_pool* p, int n_threads)
{
  atomic_store(&p->next, 0);
  if (n_threads <= 1 || p->n <= 1) {
    job_worker(p);
    return;
  }
  pthread_t th[256];
  if (n_threads > 256) n_threads = 256;
  for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], NULL, job_worker, p);
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
}

int oracle_decode_table_mt(sdr_column* cols, uint32_t n_cols, int n_threads)
{
  int rc = validate_fixed(cols, n_cols);
  if (rc) return rc;
  size_t n = 0;
  for (uint32_t c = 0; c < n_cols; ++c)
    for (uint32_t r = 0; r < cols[c].n_data; ++r) n += cols[c].data[r].n_segments;
  job_pool p;
  p.jobs = (seg_job*)malloc(sizeof(seg_job) * (n ? n : 1));
  if (!p.jobs) return SDR_E_CAPACITY;
  p.n = 0;
  for (uint32_t c = 0; c < n_cols; ++c)
    for (uint32_t r = 0; r < cols[c].n_data; ++r)
      for (uint32_t s = 0; s < cols[c].data[r].n_segments; ++s) {
        const sdr_segment* g = &cols[c].data[r].segments[s];
        if (!g->row_count) continue;
        seg_job j = {cols[c].data[r].codec, cols[c].type_size, g, (uint8_t*)cols[c].d_out};
        p.jobs[p.n++] = j;
      }
  run_pool(&p, n_threads);
  free(p.jobs);
  for (uint32_t c = 0; c < n_cols; ++c) {
    cols[c].null_count = 0;
    if (cols[c].has_nulls && cols[c].n_validity && cols[c].total_rows)
      cols[c].null_count =
        build_mask(cols[c].validity, cols[c].n_validity, cols[c].total_rows, cols[c].d_null_mask);
  }
  return SDR_OK;
}

/* ---- strings ------------------------------------------------------------ */

typedef struct str_job {
  const sdr_segment* g;
  uint32_t codec;
  uint32_t* lengths; /* column-wide lengths */
  int32_t* offsets;  /* column-wide offsets (phase 2) */
  uint8_t* chars;
} str_job;

typedef struct str_pool {
  str_job* jobs;
  size_t n;
  atomic_size_t next;
  int phase;
} str_pool;

static void* str_worker(void* arg)
{
  str_pool* p = (str_pool*)arg;
  for (;;) {
    size_t i = atomic_fetch_add(&p->next, 1);
    if (i >= p->n) break;
    str_job* j = &p->jobs[i];
    const sdr_segment* g = j->g;
    if (j->codec == SDR_CODEC_DICT_FSST) {
      if (p->phase == 0) {
        oracle_dict_fsst_decode(g->d_bytes, g->bytes_size, g->row_count, j->lengths + g->row_offset, NULL, 0, NULL);
      } else {
        uint64_t base = (uint32_t)j->offsets[g->row_offset];
        uint64_t end = (uint32_t)j->offsets[g->row_offset + g->row_count];
        uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (g->row_count ? g->row_count : 1));
        if (tmp) {
          oracle_dict_fsst_decode(g->d_bytes, g->bytes_size, g->row_count, tmp, j->chars + base, end - base, NULL);
          free(tmp);
        }
      }
    } else if (j->codec != SDR_CODEC_FSST) {
      if (p->phase == 0) {
        plain_lengths(j->codec, g->d_bytes, g->bytes_size, g->row_count, j->lengths + g->row_offset);
      } else {
        uint64_t base = (uint32_t)j->offsets[g->row_offset];
        uint64_t end = (uint32_t)j->offsets[g->row_offset + g->row_count];
        plain_decode(j->codec, g->d_bytes, g->bytes_size, g->row_count, j->chars + base, end - base);
      }
    } else if (p->phase == 0) {
      oracle_fsst_lengths(g->d_bytes, g->bytes_size, g->row_count, j->lengths + g->row_offset);
    } else {
      uint64_t base = (uint32_t)j->offsets[g->row_offset];
      uint64_t end = (uint32_t)j->offsets[g->row_offset + g->row_count];
      uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (g->row_count ? g->row_count : 1));
      if (tmp) {
        oracle_fsst_decode(g->d_bytes, g->bytes_size, g->row_count, tmp, j->chars + base, end - base);
        free(tmp);
      }
    }
  }
  return NULL;
}

static void run_str_pool(str_pool* p, int phase, int n_threads)
{
  p->phase = phase;
  atomic_store(&p->next, 0);
  if (n_threads <= 1 || p->n <= 1) {
    str_worker(p);
    return;
  }
  pthread_t th[256];
  if (n_threads > 256) n_threads = 256;
  for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], NULL, str_worker, p);
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
}

int oracle_decode_strings_column(sdr_string_column* col) { return oracle_decode_strings_column_mt(col, 1); }

int oracle_decode_strings_column_mt(sdr_string_column* col, int n_threads)
{
  if (!col) return SDR_E_INVALID_ARG;
  uint32_t n = col->total_rows;
  col->chars_used = 0;
  col->null_count = 0;
  if (col->n_data && !col->data) return SDR_E_INVALID_ARG;
  if (n > 0x7FFFFFFFu) return SDR_E_BOUNDS;
  for (uint32_t r = 0; r < col->n_data; ++r)
    if (col->data[r].codec != SDR_CODEC_FSST && col->data[r].codec != SDR_CODEC_DICTIONARY &&
        col->data[r].codec != SDR_CODEC_UNCOMPRESSED && col->data[r].codec != SDR_CODEC_DICT_FSST)
      return SDR_E_UNSUPPORTED_CODEC;
  int rc = check_runs_bounds(col->data, col->n_data, n);
  if (rc) return rc;
  if (col->has_nulls && col->n_validity) {
    rc = check_runs_bounds(col->validity, col->n_validity, n);
    if (rc) return rc;
    rc = check_validity(col->validity, col->n_validity);
    if (rc) return rc;
  }
  if (n == 0) return SDR_OK;
  if (!col->d_offsets) return SDR_E_INVALID_ARG;
  /* DICT_FSST inline NULLs need a mask to land in (checked before any output) */
  int any_inline = 0;
  for (uint32_t r = 0; r < col->n_data; ++r)
    if (col->data[r].codec == SDR_CODEC_DICT_FSST)
      for (uint32_t s = 0; s < col->data[r].n_segments; ++s) {
        const sdr_segment* g = &col->data[r].segments[s];
        if (g->row_count && dfsst_has_inline_nulls(g->d_bytes, g->bytes_size, g->row_count)) any_inline = 1;
      }
  if (any_inline && !col->d_null_mask) return SDR_E_INVALID_ARG;
  uint32_t* lengths = (uint32_t*)calloc((size_t)n + 1u, sizeof(uint32_t));
  size_t nj = 0;
  for (uint32_t r = 0; r < col->n_data; ++r) nj += col->data[r].n_segments;
  str_pool p;
  p.jobs = (str_job*)malloc(sizeof(str_job) * (nj ? nj : 1));
  if (!lengths || !p.jobs) { free(lengths); free(p.jobs); return SDR_E_CAPACITY; }
  p.n = 0;
  for (uint32_t r = 0; r < col->n_data; ++r)
    for (uint32_t s = 0; s < col->data[r].n_segments; ++s) {
      const sdr_segment* g = &col->data[r].segments[s];
      if (!g->row_count) continue;
      str_job j = {g, col->data[r].codec, lengths, col->d_offsets, col->d_chars};
      p.jobs[p.n++] = j;
    }
  run_str_pool(&p, 0, n_threads);
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    col->d_offsets[i] = (int32_t)(uint32_t)total;
    total += lengths[i];
    if (total > 0x7FFFFFFFull) { free(lengths); free(p.jobs); return SDR_E_OVERFLOW; }
  }
  col->d_offsets[n] = (int32_t)total;
  col->chars_used = total;
  if (total > col->chars_capacity || (total && !col->d_chars)) {
    free(lengths);
    free(p.jobs);
    return SDR_E_OVERFLOW;
  }
  run_str_pool(&p, 1, n_threads);
  free(lengths);
  free(p.jobs);
  if ((col->has_nulls && col->d_null_mask) || any_inline) {
    col->null_count = build_mask(col->validity, col->has_nulls ? col->n_validity : 0u, n, col->d_null_mask);
    if (any_inline) {
      for (uint32_t r = 0; r < col->n_data; ++r)
        if (col->data[r].codec == SDR_CODEC_DICT_FSST)
          for (uint32_t s = 0; s < col->data[r].n_segments; ++s) {
            const sdr_segment* g = &col->data[r].segments[s];
            if (g->row_count) dfsst_mark_nulls(g->d_bytes, g->bytes_size, g->row_count, g->row_offset, col->d_null_mask);
          }
      int64_t nulls = 0;
      for (uint32_t i = 0; i < n; ++i) nulls += !((col->d_null_mask[i >> 3] >> (i & 7u)) & 1u);
      col->null_count = nulls;
    }
  }
  return SDR_OK;
}
