// dec_rle.hpp — block-cooperative RLE decode.
//
// Layout, header parse and count walk follow the reference's
// gpu_decode_rle.cu:17-36, :140-171 and :255-316: u64 counts offset at byte
// 0, values T[ec] from byte 8, u16 counts at the counts offset; the first
// entry whose running (uint32) count equals the segment's row count ends the
// segment; a zero count, or a running count above the row count, in a
// 2048-entry tile that holds no match is malformed, and the walk stops at
// 90112 entries.  Malformed segments decode to zeros.
//
// One 256-thread block decodes one slice of rows of one segment:
//   1. walk: the counts arrive 8192 entries at a time (16 KiB, 16-byte loads
//      into LDS); each thread sums 32 consecutive counts, a block scan turns
//      the sums into per-chunk row bases, and block minima locate the first
//      exact match and the first zero/overshoot.  Segments of up to 8192
//      entries cost one pass, i.e. one memory round trip.
//   2. expansion: the slice is split over the 4 wavefronts; each finds its
//      first run from the LDS chunk bases, then walks windows of 64 runs:
//      a DPP scan gives each run's end row.  A window of at most 4096 rows
//      marks its run ends in a per-wave LDS bitmap, and each 64-row step
//      reads its 64 bits, so a row's run index is the ballot count of ends
//      at or before the step plus a popcount of those bits.  Longer windows
//      build the step's mask with scalar readlanes; windows holding a
//      zero-length run take an exact per-lane count.
#pragma once

#include "sdr_device.hpp"

namespace sdr {

constexpr int ICMP_ULE = 37;  // LLVM's unsigned <= predicate, for __builtin_amdgcn_uicmp

constexpr uint32_t RLE_TILE        = 2048;
constexpr uint32_t RLE_MAX_ENTRIES = 90112;
constexpr uint32_t RLE_THREADS     = 256;
constexpr uint32_t RLE_WAVES       = RLE_THREADS / WAVE;
constexpr uint32_t RLE_PER_THREAD  = 32;
constexpr uint32_t RLE_SUPER       = RLE_THREADS * RLE_PER_THREAD;  // 8192 entries per pass
constexpr uint32_t RLE_MAX_SUPERS  = (RLE_MAX_ENTRIES + RLE_SUPER - 1) / RLE_SUPER;
constexpr uint32_t RLE_STAGE_WORDS = (RLE_SUPER * 2 + 48) / 4;
static_assert(RLE_SUPER % RLE_TILE == 0, "passes hold whole tiles");

// Per-wave bitmap of run ends over a window's 64-row steps, one 64-bit word
// per step: RLE_BM_ROWS window rows, the up to 63 rows the first step starts
// before the window, and the step a pair may overrun.
constexpr uint32_t RLE_BM_ROWS   = 4096;
constexpr uint32_t RLE_BM_QWORDS = (RLE_BM_ROWS + 63u) / 64u + 3u;

struct RleSmem {
  uint32_t stage[RLE_STAGE_WORDS];         // counts of the resident pass
  uint64_t bm[RLE_WAVES][RLE_BM_QWORDS];   // run-end bitmap of each wave's window
  uint32_t chunk_base[RLE_THREADS];        // row base of each 32-entry chunk
  uint32_t wave_sum[RLE_WAVES];
  uint32_t pass_base[RLE_MAX_SUPERS + 1];  // row base of each pass
  uint32_t match, bad;                     // block minima (entry index)
};

__device__ __forceinline__ uint32_t lds_u16(const uint32_t* lds, uint32_t byte)
{
  const uint32_t w = byte >> 2;
  return __builtin_amdgcn_alignbit(lds[w + 1], lds[w], (byte & 3u) * 8u) & 0xFFFFu;
}

// Inclusive prefix sum within each 16-lane row of the wave.
__device__ __forceinline__ uint32_t row_incl_sum_u32(uint32_t v)
{
  v += dpp_u32<0x111>(v);  // row_shr:1
  v += dpp_u32<0x112>(v);  // row_shr:2
  v += dpp_u32<0x114>(v);  // row_shr:4
  v += dpp_u32<0x118>(v);  // row_shr:8
  return v;
}

// Stage pass `t0` (n entries) and compute every chunk's row base from
// `base`.  With `detect`, also record the first exact match and the first
// zero/overshoot entry of the pass in sm.match / sm.bad (UINT32_MAX: none).
// Returns the pass's count total.  Called by all threads of the block.
//
// Wave w reads words [1024 w, 1024 w + 1024) of the staged counts (two
// counts a word), 64 consecutive words per step, so the LDS reads are
// conflict-free; a 16-lane row of a step is one 32-entry chunk, summed by a
// DPP row scan.
__device__ __forceinline__ uint32_t rle_pass(const uint8_t* counts, uint32_t t0, uint32_t n, uint32_t base,
                                             uint32_t rows, bool detect, uint32_t& head, RleSmem& sm)
{
  __syncthreads();  // earlier readers of sm.stage and sm.chunk_base are done
  if (detect && threadIdx.x == 0) {
    sm.match = UINT32_MAX;
    sm.bad   = UINT32_MAX;
  }
  head = block_stage<RLE_THREADS>(sm.stage, counts + 2u * t0, 2u * n);
  __syncthreads();

  const uint32_t t    = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wv   = t / WAVE;
  const uint32_t row  = lane >> 4;
  const uint32_t hq   = head >> 2;
  const uint32_t sh   = (head & 3u) * 8u;
  // word i of the pass (counts 2i and 2i + 1), counts past the pass reading as zero
  auto word = [&](uint32_t i) {
    const uint32_t w = __builtin_amdgcn_alignbit(sm.stage[hq + i + 1u], sm.stage[hq + i], sh);
    return w & ((2u * i + 1u < n ? 0xFFFF0000u : 0u) | (2u * i < n ? 0xFFFFu : 0u));
  };
#pragma unroll 4
  for (uint32_t k = 0; k < RLE_PER_THREAD / 2; ++k) {
    const uint32_t w  = word(1024u * wv + 64u * k + lane);
    const uint32_t rs = row_incl_sum_u32((w & 0xFFFFu) + (w >> 16));
    if ((lane & 15u) == 15u) sm.chunk_base[64u * wv + 4u * k + row] = rs;  // the chunk's sum
  }
  __syncthreads();
  const uint32_t s    = sm.chunk_base[t];  // chunk t's sum
  const uint32_t incl = wave_incl_sum_u32(s);
  if (lane == WAVE - 1) sm.wave_sum[wv] = incl;
  __syncthreads();
  uint32_t woff = 0, total = 0;
#pragma unroll
  for (uint32_t k = 0; k < RLE_WAVES; ++k) {
    const uint32_t ws = sm.wave_sum[k];
    woff += k < wv ? ws : 0u;
    total += ws;
  }
  sm.chunk_base[t] = base + woff + incl - s;  // only thread t reads or writes entry t here
  if (detect) {
    __syncthreads();  // chunk bases visible
    uint32_t m = UINT32_MAX, bad = UINT32_MAX;  // this lane's first entries (they rise with k)
#pragma unroll 4
    for (uint32_t k = 0; k < RLE_PER_THREAD / 2; ++k) {
      const uint32_t i  = 1024u * wv + 64u * k + lane;
      const uint32_t w  = word(i);
      const uint32_t c0 = w & 0xFFFFu, c1 = w >> 16;
      const uint32_t rs = row_incl_sum_u32(c0 + c1);
      const uint32_t r0 = sm.chunk_base[64u * wv + 4u * k + row] + rs - (c0 + c1) + c0;
      const uint32_t r1 = r0 + c1;
      const bool l0 = 2u * i < n, l1 = 2u * i + 1u < n;
      const uint32_t e  = t0 + 2u * i;
      const uint32_t hm = (l0 && r0 == rows) ? e : ((l1 && r1 == rows) ? e + 1u : UINT32_MAX);
      const uint32_t hb = (l0 && (c0 == 0u || r0 > rows)) ? e : ((l1 && (c1 == 0u || r1 > rows)) ? e + 1u : UINT32_MAX);
      m   = m < hm ? m : hm;
      bad = bad < hb ? bad : hb;
    }
    if (ballot(m != UINT32_MAX) != 0ull && m != UINT32_MAX) atomicMin(&sm.match, m);
    if (ballot(bad != UINT32_MAX) != 0ull && bad != UINT32_MAX) atomicMin(&sm.bad, bad);
  }
  __syncthreads();
  return total;
}

// The value width TS (1, 2, 4 or 8 bytes) is a template parameter: each
// width is its own code path with no per-row dispatch.
template <uint32_t TS>
__device__ __forceinline__ uint64_t ld_width(const uint8_t* p)
{
  if constexpr (TS == 1) return ld_u8(p);
  else if constexpr (TS == 2) return ld_u16_any(p);
  else if constexpr (TS == 4) return ld_u32_any(p);
  else return ld_u64_any(p);
}

template <uint32_t TS>
__device__ __forceinline__ void st_width(uint8_t* out, uint32_t row, uint64_t v)
{
  if constexpr (TS == 1) st_nt(out + row, static_cast<uint8_t>(v));
  else if constexpr (TS == 2) st_nt(reinterpret_cast<uint16_t*>(out) + row, static_cast<uint16_t>(v));
  else if constexpr (TS == 4) st_nt(reinterpret_cast<uint32_t*>(out) + row, static_cast<uint32_t>(v));
  else st_nt(reinterpret_cast<uint64_t*>(out) + row, v);
}

// Expand rows [lo, hi) of one wave: entries [t0, j_end) are the pass
// resident in LDS (byte skew head), chunk bases in sm.
template <uint32_t TS>
__device__ __forceinline__ void rle_expand(const uint8_t* values, uint32_t t0, uint32_t j_end,
                                           uint32_t head, uint32_t lo, uint32_t hi, uint8_t* out, RleSmem& sm)
{
  const uint32_t lane = lane_id();
  uint64_t* bm        = sm.bm[threadIdx.x / WAVE];
  // chunk holding the first run that ends after row lo: the last chunk with
  // entries whose base is <= lo (bases never decrease up to the match)
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t u = 0; u < RLE_THREADS / WAVE; ++u) {
    const uint32_t c = lane + WAVE * u;
    cnt += static_cast<uint32_t>(
      __builtin_popcountll(ballot(t0 + RLE_PER_THREAD * c < j_end && sm.chunk_base[c] <= lo)));
  }
  if (cnt == 0) return;  // unreachable for a segment whose walk matched
  const uint32_t ch = cnt - 1u;
  uint32_t j, wstart;
  {
    const uint32_t e    = t0 + RLE_PER_THREAD * ch + lane;
    const bool live     = lane < RLE_PER_THREAD && e < j_end;
    const uint32_t c    = live ? lds_u16(sm.stage, head + 2u * (e - t0)) : 0u;
    const uint32_t incl = wave_incl_sum_u32(c);
    const uint32_t end  = sm.chunk_base[ch] + incl;
    const uint64_t hit  = ballot(live && end > lo);
    if (!hit) return;  // unreachable, as above
    const int l = __builtin_ctzll(hit);
    j           = t0 + RLE_PER_THREAD * ch + static_cast<uint32_t>(l);
    wstart      = __builtin_amdgcn_readlane(end - c, l);
  }
  const unsigned long long upto = lane == WAVE - 1 ? ~0ull : ((2ull << lane) - 1ull);
  // run counts and values of a window; the next window's are fetched before
  // this one's rows are written, so the value load overlaps the stores
  auto fetch = [&](uint32_t jw, uint32_t& c, uint64_t& v) {
    const uint32_t jl = jw + lane;
    const bool lv     = jl < j_end;
    c                 = lv ? lds_u16(sm.stage, head + 2u * (jl - t0)) : 0u;
    v                 = lv ? ld_width<TS>(values + uint64_t{jl} * TS) : 0u;
  };
  uint32_t cnt_next = 0;
  uint64_t val_next = 0;
  fetch(j, cnt_next, val_next);
  while (wstart < hi && j < j_end) {
    const uint32_t jl   = j + lane;
    const bool live     = jl < j_end;
    const uint32_t cnt1 = cnt_next;
    const uint64_t val  = val_next;
    if (j + WAVE < j_end) fetch(j + WAVE, cnt_next, val_next);
    const uint32_t incl = wave_incl_sum_u32(cnt1);
    const uint32_t end  = wstart + incl;  // exclusive end row of run jl
    // wave-uniform values are pinned to scalar registers (readfirstlane), so
    // the step loop's bounds and branches are scalar work
    const uint32_t wend = __builtin_amdgcn_readfirstlane(wstart + __builtin_amdgcn_readlane(incl, WAVE - 1));
    const bool dup      = ballot(live && cnt1 == 0u) != 0;
    const uint32_t a    = __builtin_amdgcn_readfirstlane(wstart > lo ? wstart : lo);
    const uint32_t b    = __builtin_amdgcn_readfirstlane(wend < hi ? wend : hi);
    // A window of at most RLE_BM_ROWS rows marks its run ends in the wave's
    // bitmap once; a step then reads its 64 bits instead of collecting the
    // ends one readlane at a time.
    const bool use_bm = !dup && wend - wstart <= RLE_BM_ROWS && a < b;
    const uint32_t first = __builtin_amdgcn_readfirstlane(lo + ((a - lo) & ~(WAVE - 1u)));
    if (use_bm) {
      // word k holds the ends in rows [first + 64 k, first + 64 k + 64):
      // bit l of a step's word is an end at row r0 + l, so a step reads one
      // aligned word.  Ends before `first` are in every step's ballot.
      const uint32_t nq = (b - first + 63u) / 64u + 2u;
      for (uint32_t i = lane; i < nq; i += WAVE) bm[i] = 0ull;
      wave_lds_sync();
      if (live && end >= first && end - first < 64u * nq) {
        const uint32_t bit = end - first;
        __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(bm) + (bit >> 5), 1u << (bit & 31u), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
      wave_lds_sync();
      // two independent 64-row steps per iteration, so their LDS round trips
      // overlap; dead lanes' ends read as UINT32_MAX, so one unsigned compare
      // is the ballot of ends at or before r0
      const uint32_t end_cmp = live ? end : UINT32_MAX;
      auto rank_at = [&](uint32_t r0) {
        const uint64_t m = bm[(r0 - first) >> 6];
        return static_cast<uint32_t>(__builtin_popcountll(__builtin_amdgcn_uicmp(end_cmp, r0, ICMP_ULE))) +
               static_cast<uint32_t>(__builtin_popcountll(m & upto & ~1ull));
      };
      for (uint32_t r0 = first; r0 < b; r0 += 2u * WAVE) {
        // A stored row (< b <= wend) has rank <= 63; larger ranks belong to
        // rows past the window, which are not stored.  ds_bpermute reads
        // lane (index / 4) mod 64, so those need no clamp.
        const int s0 = static_cast<int>(rank_at(r0) << 2);
        const int s1 = static_cast<int>(rank_at(r0 + WAVE) << 2);
        const uint32_t vl = static_cast<uint32_t>(val), vh = static_cast<uint32_t>(val >> 32);
        const uint32_t l0 = __builtin_amdgcn_ds_bpermute(s0, static_cast<int>(vl));
        const uint32_t l1 = __builtin_amdgcn_ds_bpermute(s1, static_cast<int>(vl));
        const uint32_t h0 = TS == 8u ? __builtin_amdgcn_ds_bpermute(s0, static_cast<int>(vh)) : 0u;
        const uint32_t h1 = TS == 8u ? __builtin_amdgcn_ds_bpermute(s1, static_cast<int>(vh)) : 0u;
        const uint32_t row0 = r0 + lane, row1 = r0 + WAVE + lane;
        const uint64_t v0 = static_cast<uint64_t>(l0) | (static_cast<uint64_t>(h0) << 32);
        const uint64_t v1 = static_cast<uint64_t>(l1) | (static_cast<uint64_t>(h1) << 32);
        if (r0 >= a && r0 + 2u * WAVE <= b) {  // uniform: both steps inside [a, b)
          st_width<TS>(out, row0, v0);
          st_width<TS>(out, row1, v1);
        } else {
          if (row0 >= a && row0 < b) st_width<TS>(out, row0, v0);
          if (row1 >= a && row1 < b) st_width<TS>(out, row1, v1);
        }
      }
      j += WAVE;
      wstart = wend;
      continue;
    }
    for (uint32_t r0 = first; r0 < b; r0 += WAVE) {
      const uint32_t row = r0 + lane;
      uint32_t rank;
      if (!dup) {
        rank              = static_cast<uint32_t>(__builtin_popcountll(ballot(live && end <= r0)));
        uint64_t inside   = ballot(live && end > r0 && end - r0 < WAVE);
        unsigned long long m = 0;
        while (inside) {
          const int l = __builtin_ctzll(inside);
          inside &= inside - 1u;
          m |= 1ull << (__builtin_amdgcn_readlane(end, l) - r0);
        }
        rank += static_cast<uint32_t>(__builtin_popcountll(m & upto));
      } else {
        rank = 0;
#pragma unroll 1
        for (uint32_t l = 0; l < WAVE; ++l) {
          const uint32_t el = __builtin_amdgcn_readlane(end, l);
          rank += (j + l < j_end && el <= row) ? 1u : 0u;
        }
      }
      const int src = static_cast<int>((rank < WAVE ? rank : WAVE - 1u) << 2);
      const uint32_t xl = __builtin_amdgcn_ds_bpermute(src, static_cast<int>(static_cast<uint32_t>(val)));
      const uint32_t xh =
        TS == 8u ? __builtin_amdgcn_ds_bpermute(src, static_cast<int>(static_cast<uint32_t>(val >> 32))) : 0u;
      if (row >= a && row < b) st_width<TS>(out, row, static_cast<uint64_t>(xl) | (static_cast<uint64_t>(xh) << 32));
    }
    j += WAVE;
    wstart = wend;
  }
}

// Decode rows [ra, rb) of an RLE segment of `rows` rows of `ts`-byte values
// (1, 2, 4 or 8) into out[0, rows).  Called by all 256 threads of a block
// (the control flow is block-uniform).
template <uint32_t TS>
__device__ __forceinline__ void rle_slice_t(const uint8_t* seg, uint32_t bytes, uint32_t rows, uint32_t ra, uint32_t rb, uint8_t* out,
                            RleSmem& sm)
{
  bool malformed        = true;
  uint64_t counts_off   = 0;
  uint32_t n_max        = 0;
  if (bytes >= 8u) {
    counts_off = su64(seg);
    if (counts_off >= 8u && counts_off < bytes) {
      const uint64_t vb = counts_off - 8u, cb = bytes - counts_off;
      uint64_t n_real   = (vb % TS == 0) ? vb / TS : cb / 2u;
      if (n_real > RLE_MAX_ENTRIES) n_real = RLE_MAX_ENTRIES;
      if (n_real > cb / 2u) n_real = cb / 2u;
      n_max     = static_cast<uint32_t>(n_real);
      malformed = false;
    }
  }
  const uint8_t* counts = seg + counts_off;

  // ---- 1. walk ----
  uint32_t ec = 0, n_pass = 0, head = 0, base = 0;
  if (!malformed) {
    malformed = true;
    for (uint32_t t0 = 0; t0 < n_max; t0 += RLE_SUPER) {
      const uint32_t n = n_max - t0 < RLE_SUPER ? n_max - t0 : RLE_SUPER;
      if (threadIdx.x == 0) sm.pass_base[n_pass] = base;
      const uint32_t total = rle_pass(counts, t0, n, base, rows, true, head, sm);
      ++n_pass;
      const uint32_t m = sm.match, bad = sm.bad;
      if (m != UINT32_MAX) {
        // a zero or overshoot before the match's tile is malformed
        malformed = bad < (m & ~(RLE_TILE - 1u));
        ec        = m + 1u;
        break;
      }
      if (bad != UINT32_MAX) break;
      base += total;
    }
    if (!malformed && 8u + uint64_t{ec} * TS > bytes) malformed = true;
  }

  const uint32_t w  = threadIdx.x / WAVE;
  const uint32_t ns = rb - ra;
  const uint32_t qw = ((ns + RLE_WAVES * WAVE - 1u) / (RLE_WAVES * WAVE)) * WAVE;
  const uint32_t wa = ra + w * qw < rb ? ra + w * qw : rb;
  const uint32_t wb = wa + qw < rb ? wa + qw : rb;
  if (malformed) {
    for (uint32_t r = wa + lane_id(); r < wb; r += WAVE) st_width<TS>(out, r, 0u);
    return;
  }

  // ---- 2. expansion, pass by pass (one pass unless ec > 8192) ----
  __syncthreads();  // pass_base visible
  const uint8_t* values = seg + 8u;
  uint32_t resident     = n_pass - 1u;
  for (uint32_t k = 0; k < n_pass; ++k) {
    const uint32_t r_lo = sm.pass_base[k];
    const uint32_t r_hi = k + 1u < n_pass ? sm.pass_base[k + 1u] : rows;
    if (r_hi <= ra || r_lo >= rb) continue;  // block-uniform
    const uint32_t t0 = k * RLE_SUPER;
    if (k != resident) {
      const uint32_t n = n_max - t0 < RLE_SUPER ? n_max - t0 : RLE_SUPER;
      rle_pass(counts, t0, n, r_lo, rows, false, head, sm);
      resident = k;
    }
    const uint32_t j_end = ec < t0 + RLE_SUPER ? ec : t0 + RLE_SUPER;
    const uint32_t lo    = wa > r_lo ? wa : r_lo;
    const uint32_t hi    = wb < r_hi ? wb : r_hi;
    if (lo < hi) rle_expand<TS>(values, t0, j_end, head, lo, hi, out, sm);
  }
}

// Runtime-width entry (the persistent batch kernel): one dispatch per slice.
__device__ __forceinline__ void rle_slice(const uint8_t* seg, uint32_t bytes, uint32_t rows, uint32_t ra, uint32_t rb,
                                          uint8_t* out, uint32_t ts, RleSmem& sm)
{
  switch (ts) {
    case 1: rle_slice_t<1>(seg, bytes, rows, ra, rb, out, sm); break;
    case 2: rle_slice_t<2>(seg, bytes, rows, ra, rb, out, sm); break;
    case 4: rle_slice_t<4>(seg, bytes, rows, ra, rb, out, sm); break;
    default: rle_slice_t<8>(seg, bytes, rows, ra, rb, out, sm); break;
  }
}

}  // namespace sdr
