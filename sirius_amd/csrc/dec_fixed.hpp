// dec_fixed.hpp — per-unit device decoders of the fixed-width codecs
// (BITPACKING, ALP, ALPRD, UNCOMPRESSED, CONSTANT), shared by the per-codec
// kernels (k_fixed.hip) and the persistent mixed-codec kernel (k_batch.hip).
// RLE decodes block-cooperatively (dec_rle.hpp).
//
// Work partitioning: a block is 4 wavefronts; each wavefront independently
// decodes one work unit of one segment (blockIdx.y selects the segment,
// blockIdx.x * 4 + wave selects the unit):
//   BITPACKING  unit = one 2048-row metadata group
//   ALP/ALPRD   unit = one 1024-row vector
// Metadata is parsed by every lane from the same addresses, so it lives in
// scalar registers; no block barrier is ever needed.  Malformed metadata
// zero-fills the unit's trusted row range, as the reference kernels do.
#pragma once

#include "sdr_device.hpp"
#include "sdr_kernels.hpp"

#include <type_traits>

namespace sdr {

constexpr uint32_t WAVES_PER_BLOCK = 4;
constexpr uint32_t BLOCK_THREADS   = WAVES_PER_BLOCK * WAVE;
// Per-wave LDS: 8 KiB of packed bits plus the 16-byte head skew, tail block
// and two guard words read by the straddle extraction.
constexpr uint32_t STAGE_BYTES = 8192;
constexpr uint32_t STAGE_WORDS = (STAGE_BYTES + 48) / 4;

__device__ __forceinline__ uint32_t wave_in_block()
{
  return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

template <typename T>
using acc_t = std::conditional_t<sizeof(T) == 8, uint64_t, uint32_t>;

template <typename T>
__device__ __forceinline__ acc_t<T> wave_incl_sum(acc_t<T> v)
{
  if constexpr (sizeof(T) == 8) return wave_incl_sum_u64(v);
  else return wave_incl_sum_u32(v);
}

template <typename T>
__device__ __forceinline__ acc_t<T> readlane_acc(acc_t<T> v, int l)
{
  if constexpr (sizeof(T) == 8) return readlane_u64(v, l);
  else return __builtin_amdgcn_readlane(v, l);
}

// Fill rows [0, n) of `out` with gen(row).
template <typename T, typename F>
__device__ __forceinline__ void wave_fill(T* out, uint32_t n, F gen)
{
  constexpr uint32_t VPL = 16 / sizeof(T);
  const bool al = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
  for (uint32_t r = lane_id() * VPL; r < n; r += WAVE * VPL) {
    T v[VPL];
#pragma unroll
    for (uint32_t i = 0; i < VPL; ++i) v[i] = gen(r + i);
    store_vec<T>(out + r, v, n - r, al);
  }
}

// ===========================================================================
// BITPACKING
// Layout and parse rules: the reference's gpu_decode_bitpacking.cu:17-42 and
// :119-213; modes :222-348.
// ===========================================================================
enum : uint32_t { BP_INVALID = 0, BP_CONSTANT = 2, BP_CONSTANT_DELTA = 3, BP_DELTA_FOR = 4, BP_FOR = 5 };
constexpr uint32_t BP_GROUP = 2048;

// `lds` holds STAGE_B bytes of packed bits (plus the head and guard words);
// groups wider than that are staged in power-of-two row chunks.
template <typename T, uint32_t STAGE_B = STAGE_BYTES>
__device__ __forceinline__ void bp_group(const uint8_t* seg, uint32_t bytes, uint32_t k, uint32_t grc, T* out,
                         uint32_t* lds, uint32_t max_chunk = BP_GROUP)
{
  constexpr uint32_t TS  = sizeof(T);
  constexpr uint32_t VPL = 16 / TS;
  using A                = acc_t<T>;

  uint32_t mode = BP_INVALID, width = 0, rc = 0;
  uint64_t frame = 0, aux = 0, doff = 0;
  const uint64_t metadata_end = bytes >= 8 ? su64(seg) : 0;
  if (metadata_end >= 8u + (uint64_t{k} + 1u) * 4u && metadata_end <= bytes) {
    const uint32_t enc      = su32(seg + metadata_end - (uint64_t{k} + 1u) * 4u);
    const uint32_t data_off = enc & 0x00FFFFFFu;
    const uint32_t pm       = enc >> 24;
    rc                      = grc;
    if (uint64_t{data_off} + 2u * TS <= metadata_end) {
      const uint64_t v0 = su_t<TS>(seg + data_off);
      const uint64_t v1 = su_t<TS>(seg + data_off + TS);
      if (pm == BP_CONSTANT) {
        mode = pm;
        aux  = v0;
      } else if (pm == BP_CONSTANT_DELTA) {
        mode  = pm;
        frame = v0;
        aux   = v1;
      } else if (pm == BP_FOR) {
        mode  = pm;
        frame = v0;
        width = static_cast<uint32_t>(v1);
        doff  = uint64_t{data_off} + 2u * TS;
      } else if (pm == BP_DELTA_FOR && uint64_t{data_off} + 3u * TS <= metadata_end) {
        mode  = pm;
        frame = v0;
        width = static_cast<uint32_t>(v1);
        aux   = su_t<TS>(seg + data_off + 2u * TS);
        doff  = uint64_t{data_off} + 3u * TS;
      }
      if (width > TS * 8u) {
        mode  = BP_INVALID;
        width = 0;
      }
      if (mode == BP_FOR || mode == BP_DELTA_FOR) {
        const uint64_t words = (uint64_t{rc} * width + 31u) / 32u;
        if (doff + words * 4u > metadata_end) {
          mode  = BP_INVALID;
          width = 0;
        }
      }
    }
  }

  if (mode == BP_CONSTANT) {
    const T c = static_cast<T>(aux);
    wave_fill<T>(out, rc, [c](uint32_t) { return c; });
    return;
  }
  if (mode == BP_CONSTANT_DELTA) {
    const A f = static_cast<A>(frame), d = static_cast<A>(aux);
    wave_fill<T>(out, rc, [f, d](uint32_t i) { return static_cast<T>(f + static_cast<A>(i) * d); });
    return;
  }
  if (mode != BP_FOR && mode != BP_DELTA_FOR) {
    wave_fill<T>(out, grc, [](uint32_t) { return T(0); });
    return;
  }

  // FOR / DELTA_FOR: stage packed words through LDS in chunks of at most
  // STAGE_B bytes of bits and at most max_chunk rows.
  static_assert(STAGE_B * 8u >= 64u * (BP_GROUP / 4u) && STAGE_B <= STAGE_BYTES, "stage size");
  const bool al          = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
  uint32_t chunk         = max_chunk;  // 512, 1024 or 2048
  while (chunk * width > STAGE_B * 8u) chunk >>= 1;
  const T tframe         = static_cast<T>(frame);
  const uint8_t* stream  = seg + doff;
  A carry                = static_cast<A>(aux);  // DELTA_FOR running value
  for (uint32_t c0 = 0; c0 < rc; c0 += chunk) {
    const uint32_t rows = rc - c0 < chunk ? rc - c0 : chunk;
    wave_lds_sync();
    const uint32_t head = stage_to_lds(lds, stream + (c0 * width) / 8u, (rows * width + 7u) / 8u);
    wave_lds_sync();
    for (uint32_t s0 = 0; s0 < rows; s0 += WAVE * VPL) {
      const uint32_t r = s0 + lane_id() * VPL;
      T v[VPL];
#pragma unroll
      for (uint32_t i = 0; i < VPL; ++i) {
        const uint32_t row = r + i;
        T u                = 0;
        if (row < rows) {
          const uint32_t bp = head + row * width;
          if constexpr (TS == 8) u = static_cast<T>(lds_bits64(lds, bp, width));
          else u = static_cast<T>(lds_bits32(lds, bp, width));
        }
        v[i] = row < rows ? static_cast<T>(static_cast<A>(tframe) + static_cast<A>(u)) : T(0);
      }
      if (mode == BP_DELTA_FOR) {
        A loc[VPL];
        loc[0] = v[0];
#pragma unroll
        for (uint32_t i = 1; i < VPL; ++i) loc[i] = loc[i - 1] + static_cast<A>(v[i]);
        const A tot  = loc[VPL - 1];
        const A incl = wave_incl_sum<T>(tot);
        const A base = carry + (incl - tot);
#pragma unroll
        for (uint32_t i = 0; i < VPL; ++i) v[i] = static_cast<T>(base + loc[i]);
        carry += readlane_acc<T>(incl, 63);
      }
      if (r < rows) store_vec<T>(out + c0 + r, v, rows - r, al);
    }
  }
}

// ===========================================================================
// ALP / ALPRD
// Layouts and parse rules: the reference's gpu_decode_alp.cu:17-72,
// :251-314 (ALP), :426-504 (ALPRD); decode :368-393 and :578-625.
// ===========================================================================
constexpr uint32_t ALP_VEC = 1024;
// ALPRD stages the left stream (<= 1024 * 3 bits + head + guard) ahead of
// the right stream.
constexpr uint32_t ALPRD_LEFT_WORDS = 112;
constexpr uint32_t ALP_STAGE_WORDS  = ALPRD_LEFT_WORDS + STAGE_WORDS;
// ALP stages a vector's packed codes, and ALPRD its right parts, 4 KiB at a
// time (one stage unless the width exceeds 31 bits), so their kernels keep 8
// waves per SIMD.
constexpr uint32_t ALP_CHUNK_BYTES = 4096;
constexpr uint32_t ALP_CHUNK_WORDS = (ALP_CHUNK_BYTES + 48) / 4;

__constant__ int64_t c_alp_fact[19] = {1LL,
                                       10LL,
                                       100LL,
                                       1000LL,
                                       10000LL,
                                       100000LL,
                                       1000000LL,
                                       10000000LL,
                                       100000000LL,
                                       1000000000LL,
                                       10000000000LL,
                                       100000000000LL,
                                       1000000000000LL,
                                       10000000000000LL,
                                       100000000000000LL,
                                       1000000000000000LL,
                                       10000000000000000LL,
                                       100000000000000000LL,
                                       1000000000000000000LL};
__constant__ float c_alp_frac_f32[11] = {1.0F, 1e-1F, 1e-2F, 1e-3F, 1e-4F, 1e-5F,
                                         1e-6F, 1e-7F, 1e-8F, 1e-9F, 1e-10F};
__constant__ double c_alp_frac_f64[21] = {1.0,   1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                          1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                          1e-14, 1e-15, 1e-16, 1e-17, 1e-18, 1e-19, 1e-20};

__device__ __forceinline__ uint32_t bp_req_bytes(uint32_t n, uint32_t w)
{
  return w == 0 ? 0u : ((n + 31u) / 32u) * 32u * w / 8u;
}

// Patch exceptions after the main store, in rounds of 64 in list order.  A
// later exception at the same position wins (the order DuckDB's serial patch
// loop gives): inside a round a lane stores only if no higher lane of the
// round names the same position; rounds are separated by a store fence.
template <typename F>
__device__ __forceinline__ void patch_exceptions(uint32_t ec, uint32_t fill, const uint8_t* pos_base,
                                                 F store_one)
{
  for (uint32_t e0 = 0; e0 < ec; e0 += WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // earlier stores first
    const uint32_t m = ec - e0 < WAVE ? ec - e0 : WAVE;
    const uint32_t e = e0 + lane_id();
    const uint32_t pos = e < ec ? ld_u16_any(pos_base + 2u * e) : 0xFFFFFFFFu;
    bool live = e < ec && pos < fill;
    for (uint32_t l = 0; l < m; ++l) {
      const uint32_t pl = __builtin_amdgcn_readlane(pos, l);
      live = live && !(l > lane_id() && pl == pos);
    }
    if (live) store_one(e, pos);
  }
}

template <typename T>
__device__ __forceinline__ void alp_vector(const uint8_t* seg, uint32_t bytes, uint32_t v, uint32_t fill, T* out,
                           uint32_t* lds)
{
  constexpr uint32_t TS  = sizeof(T);
  constexpr uint32_t VPL = 16 / TS;
  constexpr uint32_t MAX_EXP = TS == 4 ? 10u : 18u;
  const bool al = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;

  // ---- parse (uniform) ----
  uint32_t status = 0;  // 0 invalid, 1 raw, 2 compressed
  uint32_t data_off = 0, exp = 0, fac = 0, ec = 0, bw = 0, packed = 0;
  uint64_t for_v = 0;
  if (bytes >= 4u) {
    const uint32_t metadata_end = su32(seg);
    const uint64_t need         = (uint64_t{v} + 2u) * 4u;
    if (metadata_end <= bytes && metadata_end >= need) {
      data_off = su32(seg + metadata_end - (v + 1u) * 4u);
      if (data_off >= 4u && data_off < metadata_end) {
        exp = su8(seg + data_off);
        if (exp == 0xFFu) {
          if (uint64_t{data_off} + 1u + uint64_t{fill} * TS <= metadata_end) status = 1;
        } else if (uint64_t{data_off} + 13u <= metadata_end) {
          fac   = su8(seg + data_off + 1);
          ec    = su16(seg + data_off + 2);
          for_v = su64(seg + data_off + 4);
          bw    = su8(seg + data_off + 12);
          if (bw <= TS * 8u && exp <= MAX_EXP && fac <= 18u) {
            packed = bp_req_bytes(fill, bw);
            const uint64_t end = uint64_t{data_off} + 13u + packed + uint64_t{ec} * (TS + 2u);
            if (end <= metadata_end) status = 2;
          }
        }
      }
    }
  }
  if (status == 0) {
    wave_fill<T>(out, fill, [](uint32_t) { return T(0); });
    return;
  }
  if (status == 1) {
    const uint8_t* raw = seg + data_off + 1;
    for (uint32_t i = lane_id(); i < fill; i += WAVE) {
      const uint64_t b = ld_t_any<TS>(raw + i * TS);
      T x;
      if constexpr (TS == 8) x = __builtin_bit_cast(T, b);
      else x = __builtin_bit_cast(T, static_cast<uint32_t>(b));
      st_nt(out + i, x);
    }
    return;
  }

  // ---- compressed ----
  const uint8_t* pk = seg + data_off + 13u;
  const T fact = static_cast<T>(c_alp_fact[fac]);
  T frac;
  if constexpr (TS == 8) frac = c_alp_frac_f64[exp];
  else frac = c_alp_frac_f32[exp];
  // rows per stage: the whole vector when its codes fit ALP_CHUNK_BYTES,
  // else a multiple of the store step (wide codes, bw > 31)
  constexpr uint32_t STEP = WAVE * VPL;
  uint32_t ch = fill;
  if ((fill * bw + 7u) / 8u + 16u > ALP_CHUNK_BYTES) ch = ((ALP_CHUNK_BYTES - 16u) * 8u / bw) / STEP * STEP;
  for (uint32_t r0 = 0; r0 < fill; r0 += ch) {
    const uint32_t r1   = fill - r0 < ch ? fill : r0 + ch;
    const uint32_t bit0 = r0 * bw;
    const uint32_t b0   = bit0 >> 3;
    wave_lds_sync();
    const uint32_t head = stage_to_lds(lds, pk + b0, (r1 * bw + 7u) / 8u - b0) + (bit0 & 7u);
    wave_lds_sync();
    for (uint32_t s0 = r0; s0 < r1; s0 += STEP) {
      const uint32_t r = s0 + lane_id() * VPL;
      T x[VPL];
#pragma unroll
      for (uint32_t i = 0; i < VPL; ++i) {
        const uint32_t row = r + i;
        const uint64_t u   = row < r1 ? lds_bits64(lds, head + (row - r0) * bw, bw) : 0;
        const int64_t enc  = static_cast<int64_t>(u + for_v);
        // ((T)enc * (T)FACT[fac]) * FRAC[exp], left to right, no contraction
        x[i] = static_cast<T>(enc) * fact * frac;
      }
      if (r < r1) store_vec<T>(out + r, x, r1 - r, al);
    }
  }
  if (ec) {
    const uint8_t* exc = pk + packed;
    const uint8_t* pos = exc + ec * TS;
    patch_exceptions(ec, fill, pos, [&](uint32_t e, uint32_t p) {
      const uint64_t b = ld_t_any<TS>(exc + e * TS);
      T x;
      if constexpr (TS == 8) x = __builtin_bit_cast(T, b);
      else x = __builtin_bit_cast(T, static_cast<uint32_t>(b));
      gbl(out)[p] = x;
    });
  }
}

template <typename T>
__device__ __forceinline__ void alprd_vector(const uint8_t* seg, uint32_t bytes, uint32_t v, uint32_t fill, T* out,
                             uint32_t* lds)
{
  constexpr uint32_t TS    = sizeof(T);
  constexpr uint32_t VPL   = 16 / TS;
  constexpr uint32_t TBITS = TS * 8u;
  using R = std::conditional_t<TS == 8, uint64_t, uint32_t>;
  const bool al = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;

  uint32_t status = 0, data_off = 0, rbw = 0, lbw = 0, dsz = 0, ec = 0, lpb = 0, rpb = 0;
  if (bytes >= 7u) {
    const uint32_t metadata_end = su32(seg);
    const uint64_t need         = 7u + (uint64_t{v} + 1u) * 4u;
    if (metadata_end <= bytes && metadata_end >= need) {
      rbw = su8(seg + 4);
      lbw = su8(seg + 5);
      dsz = su8(seg + 6);
      const uint32_t dict_end = 7u + dsz * 2u;
      if (rbw <= TBITS && lbw <= 3u && dsz <= 8u && dict_end <= metadata_end) {
        data_off = su32(seg + metadata_end - (v + 1u) * 4u);
        if (data_off >= dict_end && uint64_t{data_off} + 2u <= metadata_end) {
          ec = su16(seg + data_off);
          if (ec == 0xFFFFu) {
            if (uint64_t{data_off} + 2u + uint64_t{fill} * TS <= metadata_end) status = 1;
          } else {
            lpb = bp_req_bytes(fill, lbw);
            rpb = bp_req_bytes(fill, rbw);
            const uint64_t end = uint64_t{data_off} + 2u + lpb + rpb + uint64_t{ec} * 4u;
            if (end <= metadata_end) status = 2;
          }
        }
      }
    }
  }
  if (status == 0) {
    wave_fill<T>(out, fill, [](uint32_t) { return T(0); });
    return;
  }
  if (status == 1) {
    const uint8_t* raw = seg + data_off + 2;
    for (uint32_t i = lane_id(); i < fill; i += WAVE) {
      const uint64_t b = ld_t_any<TS>(raw + i * TS);
      T x;
      if constexpr (TS == 8) x = __builtin_bit_cast(T, b);
      else x = __builtin_bit_cast(T, static_cast<uint32_t>(b));
      st_nt(out + i, x);
    }
    return;
  }
  // dictionary: eight u16 entries packed two to a word (unused slots read
  // as zero), so it takes four SGPRs and the kernel keeps 8 waves per SIMD
  uint32_t dict[4];
#pragma unroll
  for (uint32_t d = 0; d < 4; ++d) {
    const uint32_t lo = 2u * d < dsz ? su16(seg + 7u + 4u * d) : 0u;
    const uint32_t hi = 2u * d + 1u < dsz ? su16(seg + 9u + 4u * d) : 0u;
    dict[d]           = lo | hi << 16;
  }

  const uint8_t* lp = seg + data_off + 2u;
  const uint8_t* rp = lp + lpb;
  // left stream (<= 384 B) staged once at word 0; the right stream after it,
  // 4 KiB at a time (one stage unless the vector's right parts pass that),
  // so the kernel keeps 8 waves per SIMD
  uint32_t* lds_l = lds;
  uint32_t* lds_r = lds + ALPRD_LEFT_WORDS;
  constexpr uint32_t STEP = WAVE * VPL;
  uint32_t ch = fill;
  if ((fill * rbw + 7u) / 8u + 16u > ALP_CHUNK_BYTES) ch = ((ALP_CHUNK_BYTES - 16u) * 8u / rbw) / STEP * STEP;
  const bool full_right = rbw >= TBITS;
  const uint8_t* exc    = rp + rpb;
  const uint8_t* pos    = exc + ec * 2u;
  wave_lds_sync();
  const uint32_t lhead = stage_to_lds(lds_l, lp, (fill * lbw + 7u) / 8u);
  for (uint32_t r0 = 0; r0 < fill; r0 += ch) {
    const uint32_t r1   = fill - r0 < ch ? fill : r0 + ch;
    const uint32_t bit0 = r0 * rbw;
    const uint32_t b0   = bit0 >> 3;
    if (r0) wave_lds_sync();  // the previous chunk's readers are done
    const uint32_t rhead = stage_to_lds(lds_r, rp + b0, (r1 * rbw + 7u) / 8u - b0) + (bit0 & 7u);
    wave_lds_sync();
    for (uint32_t s0 = r0; s0 < r1; s0 += STEP) {
      const uint32_t r = s0 + lane_id() * VPL;
      T x[VPL];
#pragma unroll
      for (uint32_t i = 0; i < VPL; ++i) {
        const uint32_t row = r + i;
        R bits             = 0;
        if (row < r1) {
          const R right = static_cast<R>(lds_bits64(lds_r, rhead + (row - r0) * rbw, rbw));
          if (full_right) {
            bits = right;
          } else {
            const uint32_t li = lds_bits32(lds_l, lhead + row * lbw, lbw);  // < 8: lbw <= 3
            const uint32_t pw = (li & 4u) ? ((li & 2u) ? dict[3] : dict[2]) : ((li & 2u) ? dict[1] : dict[0]);
            const R dv        = static_cast<R>((li & 1u) ? pw >> 16 : pw & 0xFFFFu);
            bits = static_cast<R>(dv << rbw) | right;
          }
        }
        x[i] = __builtin_bit_cast(T, bits);
      }
      if (r < r1) store_vec<T>(out + r, x, r1 - r, al);
    }
    // exceptions of this chunk, while its right parts are staged
    if (ec)
      patch_exceptions(ec, fill, pos, [&](uint32_t e, uint32_t p) {
        if (p < r0 || p >= r1) return;
        const R right = static_cast<R>(lds_bits64(lds_r, rhead + (p - r0) * rbw, rbw));
        const R bits  = full_right ? right
                                   : static_cast<R>(static_cast<R>(ld_u16_any(exc + 2u * e)) << rbw) | right;
        gbl(out)[p]   = __builtin_bit_cast(T, bits);
      });
  }
}

// ===========================================================================
// UNCOMPRESSED / CONSTANT spans: bytes [b0, b1) of a segment's output, written
// by NT cooperating threads (t = 0 .. NT-1; NT >= 16).  Bytes up to the next
// word boundary of the destination, and the last partial word, are single
// byte stores; the rest are word stores, four per thread in flight.
// ===========================================================================
template <uint32_t NT, typename W>
__device__ __forceinline__ void copy_words(gptr<const uint8_t> src, gptr<uint8_t> dst, uintptr_t da, uint64_t b0,
                                           uint64_t b1, uint32_t t)
{
  constexpr uint32_t G = sizeof(W);
  const uint64_t head  = (G - (da & (G - 1u))) & (G - 1u);
  const uint64_t h     = head < b1 - b0 ? head : b1 - b0;
  if (t < h) dst[b0 + t] = src[b0 + t];
  const uint64_t v0 = b0 + h;
  const uint64_t nv = (b1 - v0) / G;
  const gptr<const W> vs = reinterpret_cast<gptr<const W>>(src + v0);
  const gptr<W> vd       = reinterpret_cast<gptr<W>>(dst + v0);
  uint64_t k = t;
  for (; k + 3u * NT < nv; k += 4u * NT) {
    const W x0 = __builtin_nontemporal_load(vs + k);
    const W x1 = __builtin_nontemporal_load(vs + k + NT);
    const W x2 = __builtin_nontemporal_load(vs + k + 2u * NT);
    const W x3 = __builtin_nontemporal_load(vs + k + 3u * NT);
    __builtin_nontemporal_store(x0, vd + k);
    __builtin_nontemporal_store(x1, vd + k + NT);
    __builtin_nontemporal_store(x2, vd + k + 2u * NT);
    __builtin_nontemporal_store(x3, vd + k + 3u * NT);
  }
  for (; k < nv; k += NT) __builtin_nontemporal_store(__builtin_nontemporal_load(vs + k), vd + k);
  const uint64_t tb = v0 + nv * G;
  if (tb + t < b1) dst[tb + t] = src[tb + t];
}

// The word is the widest (16, 8 or 4 bytes) at which source and destination
// are at the same phase; below 4, bytes.
template <uint32_t NT>
__device__ __forceinline__ void copy_span(const uint8_t* src_p, uint8_t* dst_p, uint64_t b0, uint64_t b1, uint32_t t)
{
  const gptr<const uint8_t> src = gbl(src_p);
  const gptr<uint8_t> dst       = gbl(dst_p);
  const uintptr_t sa = reinterpret_cast<uintptr_t>(src_p) + b0, da = reinterpret_cast<uintptr_t>(dst_p) + b0;
  const uintptr_t x  = sa ^ da;
  if ((x & 15u) == 0) copy_words<NT, u32x4>(src, dst, da, b0, b1, t);
  else if ((x & 7u) == 0) copy_words<NT, uint64_t>(src, dst, da, b0, b1, t);
  else if ((x & 3u) == 0) copy_words<NT, uint32_t>(src, dst, da, b0, b1, t);
  else
    for (uint64_t i = b0 + t; i < b1; i += NT) dst[i] = src[i];
}

// The output repeats a ts-byte value (ts = 1, 2, 4, 8 or 16, so the period
// divides 16): byte i of the output is byte (i & (ts - 1)) of the value.
template <uint32_t NT>
__device__ __forceinline__ void const_span(const uint8_t* val_p, uint32_t ts, uint8_t* dst_p, uint64_t b0, uint64_t b1,
                                           uint32_t t)
{
  const gptr<const uint8_t> val = gbl(val_p);
  uint64_t lo = 0, hi = 0;  // the value's bytes 0..7 and 8..15
  for (uint32_t i = 0; i < ts && i < 8u; ++i) lo |= uint64_t{val[i]} << (8u * i);
  for (uint32_t i = 8; i < ts; ++i) hi |= uint64_t{val[i]} << (8u * (i - 8u));
  if (ts < 16u) {  // replicate the period over 16 bytes
    for (uint32_t w = ts; w < 8u; w *= 2u) lo |= lo << (8u * w);
    hi = lo;
  }
  const uint32_t m = ts - 1u;
  auto byte_at = [&](uint64_t i) -> uint8_t {
    const uint32_t j = static_cast<uint32_t>(i) & m;  // ts <= 16: j < 16
    return static_cast<uint8_t>((j < 8u ? lo >> (8u * j) : hi >> (8u * (j - 8u))) & 0xFFu);
  };
  const gptr<uint8_t> dst = gbl(dst_p);
  const uintptr_t da      = reinterpret_cast<uintptr_t>(dst_p) + b0;
  const uint64_t head     = (16u - (da & 15u)) & 15u;
  const uint64_t h        = head < b1 - b0 ? head : b1 - b0;
  if (t < h) dst[b0 + t] = byte_at(b0 + t);
  const uint64_t v0 = b0 + h;
  const uint64_t nv = (b1 - v0) / 16u;
  // the 16 bytes from v0 on: the pattern rotated right by (v0 mod 16) bytes
  const uint32_t ph = static_cast<uint32_t>(v0) & 15u;
  uint64_t a = lo, c = hi;
  if (ph >= 8u) {
    a = hi;
    c = lo;
  }
  const uint32_t r = (ph & 7u) * 8u;
  const uint64_t plo = r ? (a >> r) | (c << (64u - r)) : a;
  const uint64_t phi = r ? (c >> r) | (a << (64u - r)) : c;
  u32x4 x;
  x.x = static_cast<uint32_t>(plo);
  x.y = static_cast<uint32_t>(plo >> 32);
  x.z = static_cast<uint32_t>(phi);
  x.w = static_cast<uint32_t>(phi >> 32);
  const gptr<u32x4> vd = reinterpret_cast<gptr<u32x4>>(dst + v0);
  for (uint64_t k = t; k < nv; k += NT) __builtin_nontemporal_store(x, vd + k);
  const uint64_t tb = v0 + nv * 16u;
  if (tb + t < b1) dst[tb + t] = byte_at(tb + t);
}

}  // namespace sdr
