// sdr_encode.cpp — DuckDB-format segment encoders and seeded column
// generators.  See sdr_encode.h for the contract.
//
// Layouts written here are DuckDB's on-disk segment layouts as the reference
// decoder parses them (ducndh/sirius):
//   BITPACKING  src/cuda/scan/gpu_decode_bitpacking.cu:17-42
//   RLE         src/cuda/scan/gpu_decode_rle.cu:17-36
//   ALP, ALPRD  src/cuda/scan/gpu_decode_alp.cu:17-72
//   FSST        src/cuda/scan/strings/fsst.cu:17-44, detail/fsst.cuh:50-96
// Encoder choices follow DuckDB's compression code (bitpacking AUTO mode
// selection, ALP top-k exponent/factor sampling, ALPRD cut-point search) and,
// for FSST, the symbol-table construction of Boncz, Neumann and Leis (VLDB
// 2020) in a simplified greedy form.

#include "sdr_encode.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

struct sdr_encoded {
  std::vector<uint8_t> blob;
  std::vector<sdr_enc_segment> segs;
};

namespace {

// ---------------------------------------------------------------------------
// PRNG: splitmix64-seeded xoshiro256** (Blackman & Vigna).
// ---------------------------------------------------------------------------
inline uint64_t splitmix64(uint64_t& x)
{
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed)
  {
    uint64_t x = seed;
    for (auto& w : s) w = splitmix64(x);
  }
  uint64_t next()
  {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  uint32_t u32() { return static_cast<uint32_t>(next() >> 32); }
  double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
inline uint64_t align8(uint64_t v) { return (v + 7u) & ~uint64_t{7}; }

template <typename V>
inline void put(std::vector<uint8_t>& b, size_t off, V v)
{
  std::memcpy(b.data() + off, &v, sizeof(V));
}

inline uint32_t bit_width_u64(uint64_t v) { return v == 0 ? 0u : 64u - static_cast<uint32_t>(__builtin_clzll(v)); }

// DuckDB BitpackingPrimitives::GetRequiredSize
inline uint64_t packed_bytes(uint64_t n, uint32_t w) { return w == 0 ? 0 : (n + 31u) / 32u * 32u * w / 8u; }

// Append `n` values of `w` bits LSB-first (continuous stream, padded to a
// multiple of 32 values) at the end of `out`.
void pack_append(std::vector<uint8_t>& out, const uint64_t* v, uint64_t n, uint32_t w)
{
  uint64_t nb   = packed_bytes(n, w);
  size_t base   = out.size();
  out.resize(base + nb, 0);
  if (w == 0) return;
  uint8_t* p = out.data() + base;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t x = w >= 64 ? v[i] : (v[i] & ((uint64_t{1} << w) - 1u));
    uint64_t bit = i * w;
    uint32_t rem = w;
    while (rem) {
      uint64_t byte = bit >> 3;
      uint32_t off  = static_cast<uint32_t>(bit & 7u);
      uint32_t take = std::min<uint32_t>(8u - off, rem);
      p[byte] |= static_cast<uint8_t>((x & ((1u << take) - 1u)) << off);
      x >>= take;
      bit += take;
      rem -= take;
    }
  }
}

void add_segment(sdr_encoded* e, const std::vector<uint8_t>& seg, uint64_t row_offset, uint32_t rows,
                 uint32_t max_len = 0)
{
  size_t off = (e->blob.size() + 15u) & ~size_t{15};
  e->blob.resize(off, 0);
  e->blob.insert(e->blob.end(), seg.begin(), seg.end());
  sdr_enc_segment s{};
  s.offset            = off;
  s.bytes_size        = static_cast<uint32_t>(seg.size());
  s.row_offset        = static_cast<uint32_t>(row_offset);
  s.row_count         = rows;
  s.max_string_length = max_len;
  e->segs.push_back(s);
}

// ---------------------------------------------------------------------------
// BITPACKING — DuckDB BitpackingState::Flush mode selection.
// ---------------------------------------------------------------------------
enum BpMode : uint32_t { BP_CONSTANT = 2, BP_CONSTANT_DELTA = 3, BP_DELTA_FOR = 4, BP_FOR = 5 };
constexpr uint32_t BP_GROUP = 2048;

template <typename T>
struct BpGroup {
  uint32_t mode;
  std::vector<uint8_t> bytes;  // header Ts + packed stream
};

template <typename S>
bool sub_ok(S a, S b, S& r) { return !__builtin_sub_overflow(a, b, &r); }

template <typename T>
BpGroup<T> bp_encode_group(const T* v, uint32_t n, int force)
{
  using U = std::make_unsigned_t<T>;
  using S = std::make_signed_t<T>;
  BpGroup<T> g;
  auto put_t = [&](T x) {
    size_t o = g.bytes.size();
    g.bytes.resize(o + sizeof(T));
    std::memcpy(g.bytes.data() + o, &x, sizeof(T));
  };
  T mn = v[0], mx = v[0];
  for (uint32_t i = 1; i < n; ++i) {
    mn = std::min(mn, v[i]);
    mx = std::max(mx, v[i]);
  }
  if (mn == mx && (force == 0 || force == BP_CONSTANT)) {
    g.mode = BP_CONSTANT;
    put_t(mx);
    return g;
  }
  S for_diff{};
  bool can_for = sub_ok<S>(static_cast<S>(mx), static_cast<S>(mn), for_diff);
  // delta statistics (DuckDB CalculateDeltaStats)
  bool can_delta = n >= 2;
  std::vector<S> d(n);
  S min_d = std::numeric_limits<S>::max(), max_d = std::numeric_limits<S>::min();
  for (uint32_t i = 1; i < n && can_delta; ++i) {
    if (!sub_ok<S>(static_cast<S>(v[i]), static_cast<S>(v[i - 1]), d[i])) can_delta = false;
    min_d = std::min(min_d, d[i]);
    max_d = std::max(max_d, d[i]);
  }
  S dd{}, delta_offset{};
  if (can_delta) {
    d[0]      = min_d;
    can_delta = sub_ok<S>(max_d, min_d, dd) && sub_ok<S>(static_cast<S>(v[0]), min_d, delta_offset);
  }
  if (can_delta) {
    if (max_d == min_d && (force == 0 || force == BP_CONSTANT_DELTA)) {
      g.mode = BP_CONSTANT_DELTA;
      put_t(v[0]);
      put_t(static_cast<T>(max_d));
      return g;
    }
    uint32_t wd = bit_width_u64(static_cast<U>(dd));
    uint32_t wr = can_for ? bit_width_u64(static_cast<U>(for_diff)) : 8u * sizeof(T);
    if ((force == 0 && wd < wr) || force == BP_DELTA_FOR) {
      g.mode = BP_DELTA_FOR;
      put_t(static_cast<T>(min_d));
      put_t(static_cast<T>(wd));
      put_t(static_cast<T>(delta_offset));
      std::vector<uint64_t> p(n);
      for (uint32_t i = 0; i < n; ++i) p[i] = static_cast<U>(static_cast<U>(d[i]) - static_cast<U>(min_d));
      pack_append(g.bytes, p.data(), n, wd);
      return g;
    }
  }
  g.mode   = BP_FOR;
  T frame  = can_for ? mn : T(0);
  uint32_t w = can_for ? bit_width_u64(static_cast<U>(for_diff)) : 8u * sizeof(T);
  put_t(frame);
  put_t(static_cast<T>(w));
  std::vector<uint64_t> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = static_cast<U>(static_cast<U>(v[i]) - static_cast<U>(frame));
  pack_append(g.bytes, p.data(), n, w);
  return g;
}

template <typename T>
void bp_encode(sdr_encoded* e, const T* v, uint64_t n, uint32_t rg_rows, uint32_t block, int force)
{
  struct Pending {
    std::vector<uint8_t> data;  // bytes after the 8-byte header
    std::vector<uint32_t> entries;
    uint64_t row_start = 0;
    uint32_t rows      = 0;
  } cur;
  auto flush = [&]() {
    if (cur.rows == 0) return;
    uint64_t meta_off = align8(8 + cur.data.size());
    uint64_t total    = meta_off + 4u * cur.entries.size();
    std::vector<uint8_t> seg(total, 0);
    put<uint64_t>(seg, 0, total);
    std::memcpy(seg.data() + 8, cur.data.data(), cur.data.size());
    for (size_t k = 0; k < cur.entries.size(); ++k) put<uint32_t>(seg, total - 4u * (k + 1), cur.entries[k]);
    add_segment(e, seg, cur.row_start, cur.rows);
    cur.data.clear();
    cur.entries.clear();
    cur.rows = 0;
  };
  for (uint64_t rg = 0; rg < n; rg += rg_rows) {
    uint64_t rg_end = std::min<uint64_t>(n, rg + rg_rows);
    flush();
    cur.row_start = rg;
    for (uint64_t g0 = rg; g0 < rg_end; g0 += BP_GROUP) {
      uint32_t gn = static_cast<uint32_t>(std::min<uint64_t>(BP_GROUP, rg_end - g0));
      BpGroup<T> grp = bp_encode_group<T>(v + g0, gn, force);
      uint64_t need = align8(8 + cur.data.size() + grp.bytes.size()) + 4u * (cur.entries.size() + 1);
      if (need > block && cur.rows > 0) {
        flush();
        cur.row_start = g0;
      }
      uint32_t data_off = static_cast<uint32_t>(8 + cur.data.size());
      cur.entries.push_back((grp.mode << 24) | data_off);
      cur.data.insert(cur.data.end(), grp.bytes.begin(), grp.bytes.end());
      cur.rows += gn;
    }
  }
  flush();
}

// ---------------------------------------------------------------------------
// RLE — DuckDB RLECompressState (runs capped at 65535, counts aligned to 8).
// ---------------------------------------------------------------------------
template <typename T>
void rle_encode(sdr_encoded* e, const T* v, uint64_t n, uint32_t rg_rows, uint32_t block)
{
  const uint64_t max_entries = (block - 8u) / (sizeof(T) + 2u);
  std::vector<T> vals;
  std::vector<uint16_t> counts;
  uint64_t seg_start = 0;
  uint32_t seg_rows  = 0;
  auto flush = [&]() {
    if (seg_rows == 0) return;
    uint64_t cnt_off = align8(8 + vals.size() * sizeof(T));
    std::vector<uint8_t> seg(cnt_off + 2u * counts.size(), 0);
    put<uint64_t>(seg, 0, cnt_off);
    std::memcpy(seg.data() + 8, vals.data(), vals.size() * sizeof(T));
    std::memcpy(seg.data() + cnt_off, counts.data(), counts.size() * 2u);
    add_segment(e, seg, seg_start, seg_rows);
    vals.clear();
    counts.clear();
    seg_rows = 0;
  };
  for (uint64_t rg = 0; rg < n; rg += rg_rows) {
    uint64_t rg_end = std::min<uint64_t>(n, rg + rg_rows);
    flush();
    seg_start = rg;
    uint64_t i = rg;
    while (i < rg_end) {
      T x        = v[i];
      uint64_t j = i + 1;
      while (j < rg_end && v[j] == x && j - i < 65535u) ++j;
      if (vals.size() == max_entries) {
        flush();
        seg_start = i;
      }
      vals.push_back(x);
      counts.push_back(static_cast<uint16_t>(j - i));
      seg_rows += static_cast<uint32_t>(j - i);
      i = j;
    }
  }
  flush();
}

// ---------------------------------------------------------------------------
// ALP — DuckDB AlpCompression (alp.hpp / alp_compress.hpp) restated.
// ---------------------------------------------------------------------------
template <typename T>
struct AlpK;
template <>
struct AlpK<double> {
  static constexpr int MAX_EXP = 18;
  static constexpr double MAGIC = 6755399441055744.0;  // 2^52 + 2^51
  static constexpr double EXP[24] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                     1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                     1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22, 1e23};
  static constexpr double FRAC[21] = {1e0,   1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                      1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                      1e-14, 1e-15, 1e-16, 1e-17, 1e-18, 1e-19, 1e-20};
  using Bits = uint64_t;
};
template <>
struct AlpK<float> {
  static constexpr int MAX_EXP = 10;
  static constexpr float MAGIC = 12582912.0f;  // 2^23 + 2^22
  static constexpr float EXP[11] = {1e0f, 1e1f, 1e2f, 1e3f, 1e4f, 1e5f, 1e6f, 1e7f, 1e8f, 1e9f, 1e10f};
  static constexpr float FRAC[11] = {1e0f, 1e-1f, 1e-2f, 1e-3f, 1e-4f, 1e-5f,
                                     1e-6f, 1e-7f, 1e-8f, 1e-9f, 1e-10f};
  using Bits = uint32_t;
};
constexpr int64_t ALP_FACT[19] = {1LL,
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
constexpr double ALP_UPPER = 9223372036854774784.0;
constexpr double ALP_LOWER = -9223372036854774784.0;

template <typename T>
bool alp_impossible(T x)
{
  return !std::isfinite(x) || x > static_cast<T>(ALP_UPPER) || x < static_cast<T>(ALP_LOWER) ||
         (x == T(0) && std::signbit(x));
}

template <typename T>
int64_t alp_encode_value(T v, int e, int f)
{
  volatile T tmp = v * AlpK<T>::EXP[e] * AlpK<T>::FRAC[f];
  T t = tmp;
  if (alp_impossible(t)) return static_cast<int64_t>(ALP_UPPER);
  volatile T r = t + AlpK<T>::MAGIC;
  T rr = r - AlpK<T>::MAGIC;
  return static_cast<int64_t>(rr);
}

// decode order (T)enc * (T)FACT[f] * FRAC[e] — the decoder's formula exactly
template <typename T>
T alp_decode_value(int64_t enc, int f, int e)
{
  volatile T a = static_cast<T>(enc) * static_cast<T>(ALP_FACT[f]);
  return a * AlpK<T>::FRAC[e];
}

template <typename T>
bool alp_same(T a, T b)
{
  return std::memcmp(&a, &b, sizeof(T)) == 0;
}

// estimated bits for a sample under (e, f)
template <typename T>
uint64_t alp_cost(const T* s, uint32_t n, int e, int f)
{
  int64_t mn = std::numeric_limits<int64_t>::max(), mx = std::numeric_limits<int64_t>::min();
  uint32_t exc = 0;
  for (uint32_t i = 0; i < n; ++i) {
    int64_t enc = alp_encode_value<T>(s[i], e, f);
    if (!alp_same<T>(alp_decode_value<T>(enc, f, e), s[i])) {
      ++exc;
      continue;
    }
    mn = std::min(mn, enc);
    mx = std::max(mx, enc);
  }
  uint32_t bw = (mx >= mn) ? bit_width_u64(static_cast<uint64_t>(mx) - static_cast<uint64_t>(mn)) : 0;
  return uint64_t{bw} * n + uint64_t{exc} * (8u * sizeof(T) + 16u);
}

template <typename T>
void alp_encode(sdr_encoded* e, const T* v, uint64_t n, uint32_t rg_rows, uint32_t block)
{
  constexpr uint32_t VEC = 1024;
  constexpr uint32_t TS  = sizeof(T);
  struct Seg {
    std::vector<uint8_t> data;  // after the 4-byte header
    std::vector<uint32_t> offs;
    uint64_t row_start = 0;
    uint32_t rows      = 0;
  } cur;
  auto flush = [&]() {
    if (cur.rows == 0) return;
    uint64_t meta_off = align8(4 + cur.data.size());
    uint64_t total    = meta_off + 4u * cur.offs.size();
    std::vector<uint8_t> seg(total, 0);
    put<uint32_t>(seg, 0, static_cast<uint32_t>(total));
    std::memcpy(seg.data() + 4, cur.data.data(), cur.data.size());
    for (size_t k = 0; k < cur.offs.size(); ++k) put<uint32_t>(seg, total - 4u * (k + 1), cur.offs[k]);
    add_segment(e, seg, cur.row_start, cur.rows);
    cur.data.clear();
    cur.offs.clear();
    cur.rows = 0;
  };
  for (uint64_t rg = 0; rg < n; rg += rg_rows) {
    uint64_t rg_end = std::min<uint64_t>(n, rg + rg_rows);
    flush();
    cur.row_start = rg;
    // row-group level: best (e,f) of up to 8 sampled vectors, keep top 5
    uint64_t nvec = (rg_end - rg + VEC - 1) / VEC;
    std::vector<std::pair<int, int>> best_of;
    uint64_t step = std::max<uint64_t>(1, nvec / 8);
    for (uint64_t vi = 0; vi < nvec && best_of.size() < 8; vi += step) {
      uint64_t a = rg + vi * VEC, b = std::min<uint64_t>(rg_end, a + VEC);
      std::vector<T> smp;
      uint64_t st = std::max<uint64_t>(1, (b - a) / 32);
      for (uint64_t i = a; i < b && smp.size() < 32; i += st) smp.push_back(v[i]);
      uint64_t bc = ~uint64_t{0};
      std::pair<int, int> bp{0, 0};
      for (int ex = AlpK<T>::MAX_EXP; ex >= 0; --ex)
        for (int fa = ex; fa >= 0; --fa) {
          uint64_t c = alp_cost<T>(smp.data(), static_cast<uint32_t>(smp.size()), ex, fa);
          if (c < bc) {
            bc = c;
            bp = {ex, fa};
          }
        }
      best_of.push_back(bp);
    }
    std::vector<std::pair<std::pair<int, int>, int>> tally;
    for (auto& c : best_of) {
      auto it = std::find_if(tally.begin(), tally.end(), [&](auto& t) { return t.first == c; });
      if (it == tally.end()) tally.push_back({c, 1});
      else ++it->second;
    }
    std::sort(tally.begin(), tally.end(), [](auto& x, auto& y) {
      if (x.second != y.second) return x.second > y.second;
      return x.first > y.first;
    });
    if (tally.size() > 5) tally.resize(5);
    for (uint64_t a = rg; a < rg_end; a += VEC) {
      uint32_t cnt = static_cast<uint32_t>(std::min<uint64_t>(VEC, rg_end - a));
      const T* x   = v + a;
      std::pair<int, int> ef = tally[0].first;
      if (tally.size() > 1) {
        std::vector<T> smp;
        uint32_t st = std::max<uint32_t>(1, cnt / 32);
        for (uint32_t i = 0; i < cnt && smp.size() < 32; i += st) smp.push_back(x[i]);
        uint64_t bc = ~uint64_t{0};
        for (auto& t : tally) {
          uint64_t c = alp_cost<T>(smp.data(), static_cast<uint32_t>(smp.size()), t.first.first, t.first.second);
          if (c < bc) {
            bc = c;
            ef = t.first;
          }
        }
      }
      int ex = ef.first, fa = ef.second;
      std::vector<int64_t> enc(cnt);
      std::vector<uint16_t> epos;
      std::vector<T> evals;
      bool have_first = false;
      int64_t first_ok = 0;
      for (uint32_t i = 0; i < cnt; ++i) {
        enc[i] = alp_encode_value<T>(x[i], ex, fa);
        if (!alp_same<T>(alp_decode_value<T>(enc[i], fa, ex), x[i])) {
          epos.push_back(static_cast<uint16_t>(i));
          evals.push_back(x[i]);
        } else if (!have_first) {
          have_first = true;
          first_ok   = enc[i];
        }
      }
      for (auto p : epos) enc[p] = first_ok;
      int64_t mn = *std::min_element(enc.begin(), enc.end());
      int64_t mx = *std::max_element(enc.begin(), enc.end());
      uint32_t bw = bit_width_u64(static_cast<uint64_t>(mx) - static_cast<uint64_t>(mn));
      std::vector<uint8_t> vb;
      uint64_t comp_size = 13 + packed_bytes(cnt, bw) + epos.size() * (TS + 2u);
      if (comp_size > 1u + uint64_t{cnt} * TS || epos.size() > 0xFFFEu) {
        vb.push_back(0xFF);  // uncompressed vector
        size_t o = vb.size();
        vb.resize(o + uint64_t{cnt} * TS);
        std::memcpy(vb.data() + o, x, uint64_t{cnt} * TS);
      } else {
        vb.resize(13);
        vb[0] = static_cast<uint8_t>(ex);
        vb[1] = static_cast<uint8_t>(fa);
        put<uint16_t>(vb, 2, static_cast<uint16_t>(epos.size()));
        put<uint64_t>(vb, 4, static_cast<uint64_t>(mn));
        vb[12] = static_cast<uint8_t>(bw);
        std::vector<uint64_t> u(cnt);
        for (uint32_t i = 0; i < cnt; ++i) u[i] = static_cast<uint64_t>(enc[i]) - static_cast<uint64_t>(mn);
        pack_append(vb, u.data(), cnt, bw);
        size_t o = vb.size();
        vb.resize(o + evals.size() * TS + epos.size() * 2u);
        std::memcpy(vb.data() + o, evals.data(), evals.size() * TS);
        std::memcpy(vb.data() + o + evals.size() * TS, epos.data(), epos.size() * 2u);
      }
      uint64_t need = align8(4 + cur.data.size() + vb.size()) + 4u * (cur.offs.size() + 1);
      if (need > block && cur.rows > 0) {
        flush();
        cur.row_start = a;
      }
      cur.offs.push_back(static_cast<uint32_t>(4 + cur.data.size()));
      cur.data.insert(cur.data.end(), vb.begin(), vb.end());
      cur.rows += cnt;
    }
  }
  flush();
}

// ---------------------------------------------------------------------------
// ALPRD — DuckDB AlpRDCompression restated: per row group, the cut point
// (left width 1..16) and an up-to-8-entry left-part dictionary minimising
// the estimated size on a sample; unmatched left parts become exceptions.
// ---------------------------------------------------------------------------
template <typename T>
void alprd_encode(sdr_encoded* e, const T* v, uint64_t n, uint32_t rg_rows, uint32_t block)
{
  using Bits              = typename AlpK<T>::Bits;
  constexpr uint32_t VEC  = 1024;
  constexpr uint32_t TS   = sizeof(T);
  constexpr uint32_t TBIT = 8u * TS;
  struct Seg {
    std::vector<uint8_t> data;  // vectors
    std::vector<uint32_t> offs;
    uint64_t row_start = 0;
    uint32_t rows      = 0;
  } cur;
  uint32_t rbw = 0, lbw = 0;
  std::vector<uint16_t> dict;
  auto header_size = [&]() -> uint64_t { return 7u + 2u * dict.size(); };
  auto flush = [&]() {
    if (cur.rows == 0) return;
    uint64_t hs       = header_size();
    uint64_t meta_off = align8(hs + cur.data.size());
    uint64_t total    = meta_off + 4u * cur.offs.size();
    std::vector<uint8_t> seg(total, 0);
    put<uint32_t>(seg, 0, static_cast<uint32_t>(total));
    seg[4] = static_cast<uint8_t>(rbw);
    seg[5] = static_cast<uint8_t>(lbw);
    seg[6] = static_cast<uint8_t>(dict.size());
    for (size_t d = 0; d < dict.size(); ++d) put<uint16_t>(seg, 7 + 2 * d, dict[d]);
    std::memcpy(seg.data() + hs, cur.data.data(), cur.data.size());
    for (size_t k = 0; k < cur.offs.size(); ++k) put<uint32_t>(seg, total - 4u * (k + 1), cur.offs[k] + static_cast<uint32_t>(hs));
    add_segment(e, seg, cur.row_start, cur.rows);
    cur.data.clear();
    cur.offs.clear();
    cur.rows = 0;
  };
  auto bits_of = [](T x) {
    Bits b;
    std::memcpy(&b, &x, TS);
    return b;
  };
  for (uint64_t rg = 0; rg < n; rg += rg_rows) {
    uint64_t rg_end = std::min<uint64_t>(n, rg + rg_rows);
    flush();
    cur.row_start = rg;
    // sample
    std::vector<Bits> smp;
    uint64_t st = std::max<uint64_t>(1, (rg_end - rg) / 256);
    for (uint64_t i = rg; i < rg_end && smp.size() < 256; i += st) smp.push_back(bits_of(v[i]));
    uint64_t best = ~uint64_t{0};
    for (uint32_t lw = 1; lw <= 16; ++lw) {
      uint32_t r = TBIT - lw;
      std::unordered_map<uint32_t, uint32_t> freq;
      for (auto b : smp) ++freq[static_cast<uint32_t>(b >> r)];
      std::vector<std::pair<uint32_t, uint32_t>> fv(freq.begin(), freq.end());
      std::sort(fv.begin(), fv.end(), [](auto& a, auto& b) {
        if (a.second != b.second) return a.second > b.second;
        return a.first < b.first;
      });
      size_t ds = std::min<size_t>(8, fv.size());
      uint32_t covered = 0;
      for (size_t k = 0; k < ds; ++k) covered += fv[k].second;
      uint32_t exc = static_cast<uint32_t>(smp.size()) - covered;
      uint32_t l   = ds <= 1 ? 0u : bit_width_u64(ds - 1);
      uint64_t cost = uint64_t{r + l} * smp.size() + uint64_t{exc} * 32u;
      if (cost < best) {
        best = cost;
        rbw  = r;
        lbw  = l;
        dict.clear();
        for (size_t k = 0; k < ds; ++k) dict.push_back(static_cast<uint16_t>(fv[k].first));
      }
    }
    Bits rmask = rbw >= TBIT ? ~Bits{0} : static_cast<Bits>((Bits{1} << rbw) - 1u);
    for (uint64_t a = rg; a < rg_end; a += VEC) {
      uint32_t cnt = static_cast<uint32_t>(std::min<uint64_t>(VEC, rg_end - a));
      std::vector<uint64_t> left(cnt), right(cnt);
      std::vector<uint16_t> eleft, epos;
      for (uint32_t i = 0; i < cnt; ++i) {
        Bits b    = bits_of(v[a + i]);
        right[i]  = b & rmask;
        uint16_t lp = static_cast<uint16_t>(b >> rbw);
        size_t k  = 0;
        while (k < dict.size() && dict[k] != lp) ++k;
        if (k == dict.size()) {
          left[i] = 0;
          eleft.push_back(lp);
          epos.push_back(static_cast<uint16_t>(i));
        } else {
          left[i] = k;
        }
      }
      std::vector<uint8_t> vb;
      uint64_t comp = 2 + packed_bytes(cnt, lbw) + packed_bytes(cnt, rbw) + epos.size() * 4u;
      if (comp > 2u + uint64_t{cnt} * TS || epos.size() >= 0xFFFFu) {
        vb.resize(2 + uint64_t{cnt} * TS);
        put<uint16_t>(vb, 0, 0xFFFF);
        std::memcpy(vb.data() + 2, v + a, uint64_t{cnt} * TS);
      } else {
        vb.resize(2);
        put<uint16_t>(vb, 0, static_cast<uint16_t>(epos.size()));
        pack_append(vb, left.data(), cnt, lbw);
        pack_append(vb, right.data(), cnt, rbw);
        size_t o = vb.size();
        vb.resize(o + 4u * epos.size());
        std::memcpy(vb.data() + o, eleft.data(), 2u * eleft.size());
        std::memcpy(vb.data() + o + 2u * eleft.size(), epos.data(), 2u * epos.size());
      }
      uint64_t need = align8(header_size() + cur.data.size() + vb.size()) + 4u * (cur.offs.size() + 1);
      if (need > block && cur.rows > 0) {
        flush();
        cur.row_start = a;
      }
      cur.offs.push_back(static_cast<uint32_t>(cur.data.size()));
      cur.data.insert(cur.data.end(), vb.begin(), vb.end());
      cur.rows += cnt;
    }
  }
  flush();
}

// ---------------------------------------------------------------------------
// FSST — symbol table construction (greedy, 5 generations over a sample, as
// in the FSST paper) and DuckDB's FSST segment layout.
// ---------------------------------------------------------------------------
struct Sym {
  uint8_t len;
  uint8_t b[8];
  bool operator==(const Sym& o) const { return len == o.len && std::memcmp(b, o.b, len) == 0; }
};

struct SymTable {
  std::vector<Sym> syms;                   // in code order (after finalize)
  int16_t one[256];                         // byte -> code of 1-byte symbol or -1
  std::vector<std::vector<uint16_t>> by2;  // (b0<<8|b1) -> codes of len>=2, longest first
  void index()
  {
    std::fill(one, one + 256, int16_t{-1});
    by2.assign(65536, {});
    for (size_t c = 0; c < syms.size(); ++c) {
      if (syms[c].len == 1) one[syms[c].b[0]] = static_cast<int16_t>(c);
      else by2[(syms[c].b[0] << 8) | syms[c].b[1]].push_back(static_cast<uint16_t>(c));
    }
    for (auto& l : by2)
      std::sort(l.begin(), l.end(), [&](uint16_t a, uint16_t b) { return syms[a].len > syms[b].len; });
  }
  // longest match at p (n bytes left): code or -1; sets len
  int match(const uint8_t* p, size_t n, uint32_t& len) const
  {
    if (n >= 2) {
      for (uint16_t c : by2[(p[0] << 8) | p[1]]) {
        const Sym& s = syms[c];
        if (s.len <= n && std::memcmp(s.b, p, s.len) == 0) {
          len = s.len;
          return c;
        }
      }
    }
    if (one[p[0]] >= 0) {
      len = 1;
      return one[p[0]];
    }
    len = 1;
    return -1;
  }
};

struct SymHash {
  size_t operator()(const Sym& s) const
  {
    uint64_t x = 0;
    std::memcpy(&x, s.b, s.len);
    return std::hash<uint64_t>()(x * 31 + s.len);
  }
};

SymTable fsst_build(const std::vector<const uint8_t*>& sp, const std::vector<uint32_t>& sl)
{
  SymTable st;
  st.index();
  for (int gen = 0; gen < 5; ++gen) {
    // codes: 0..n-1 symbols, 256+byte pseudo-codes for unmatched bytes
    size_t ns = st.syms.size();
    auto sym_of = [&](int code) -> Sym {
      if (code >= 256) {
        Sym s{};
        s.len  = 1;
        s.b[0] = static_cast<uint8_t>(code - 256);
        return s;
      }
      return st.syms[code];
    };
    std::vector<uint32_t> c1(512, 0);
    std::unordered_map<uint32_t, uint32_t> c2;
    for (size_t r = 0; r < sp.size(); ++r) {
      const uint8_t* p = sp[r];
      size_t n         = sl[r];
      int prev         = -1;
      size_t i         = 0;
      while (i < n) {
        uint32_t len;
        int c = st.match(p + i, n - i, len);
        if (c < 0) c = 256 + p[i];
        ++c1[c];
        if (prev >= 0) ++c2[(static_cast<uint32_t>(prev) << 16) | static_cast<uint32_t>(c)];
        prev = c;
        i += len;
      }
    }
    (void)ns;
    std::unordered_map<Sym, uint64_t, SymHash> gain;
    for (int c = 0; c < 512; ++c)
      if (c1[c]) {
        Sym s = sym_of(c);
        gain[s] += uint64_t{c1[c]} * s.len;
      }
    for (auto& kv : c2) {
      Sym a = sym_of(static_cast<int>(kv.first >> 16)), b = sym_of(static_cast<int>(kv.first & 0xFFFF));
      if (a.len + b.len > 8) continue;
      Sym s{};
      s.len = static_cast<uint8_t>(a.len + b.len);
      std::memcpy(s.b, a.b, a.len);
      std::memcpy(s.b + a.len, b.b, b.len);
      gain[s] += uint64_t{kv.second} * s.len;
    }
    std::vector<std::pair<Sym, uint64_t>> cand(gain.begin(), gain.end());
    std::sort(cand.begin(), cand.end(), [](auto& x, auto& y) {
      if (x.second != y.second) return x.second > y.second;
      if (x.first.len != y.first.len) return x.first.len > y.first.len;
      return std::memcmp(x.first.b, y.first.b, x.first.len) < 0;
    });
    if (cand.size() > 255) cand.resize(255);
    st.syms.clear();
    for (auto& c : cand) st.syms.push_back(c.first);
    st.index();
  }
  // finalize: code order = lengths 2,3,...,8 then 1
  std::stable_sort(st.syms.begin(), st.syms.end(), [](const Sym& a, const Sym& b) {
    auto key = [](const Sym& s) { return s.len == 1 ? 9 : s.len; };
    return key(a) < key(b);
  });
  st.index();
  return st;
}

// duckdb_fsst_export serialisation (version, zero-terminated flag, length
// histogram, symbol bytes in code order).
std::vector<uint8_t> fsst_export(const SymTable& st)
{
  uint8_t histo[8] = {0};
  for (auto& s : st.syms) ++histo[s.len - 1];
  uint64_t version = (uint64_t{20190218} << 32) | (uint64_t{st.syms.size()} << 8) | 1u;
  std::vector<uint8_t> out(17, 0);
  std::memcpy(out.data(), &version, 8);
  out[8] = 0;
  for (int i = 0; i < 8; ++i) out[9 + i] = histo[i];
  for (auto& s : st.syms) out.insert(out.end(), s.b, s.b + s.len);
  return out;
}

void fsst_compress_row(const SymTable& st, const uint8_t* p, size_t n, std::vector<uint8_t>& out)
{
  size_t i = 0;
  while (i < n) {
    uint32_t len;
    int c = st.match(p + i, n - i, len);
    if (c < 0) {
      out.push_back(255);
      out.push_back(p[i]);
    } else {
      out.push_back(static_cast<uint8_t>(c));
    }
    i += len;
  }
}

void fsst_encode(sdr_encoded* e, const uint8_t* chars, const uint64_t* offs, uint64_t n, uint32_t rg_rows,
                 uint32_t block, uint64_t seed)
{
  Rng rng(seed);
  for (uint64_t rg = 0; rg < n; rg += rg_rows) {
    uint64_t rg_end = std::min<uint64_t>(n, rg + rg_rows);
    // sample ~16 KiB of rows
    std::vector<const uint8_t*> sp;
    std::vector<uint32_t> sl;
    uint64_t budget = 16384;
    uint64_t rows   = rg_end - rg;
    for (uint32_t t = 0; t < 4096 && budget > 0; ++t) {
      uint64_t r = rg + rng.next() % rows;
      uint32_t l = static_cast<uint32_t>(offs[r + 1] - offs[r]);
      sp.push_back(chars + offs[r]);
      sl.push_back(l);
      budget = budget > l ? budget - l : 0;
    }
    SymTable st = fsst_build(sp, sl);
    std::vector<uint8_t> symtab = fsst_export(st);
    // pack rows into segments
    uint64_t r = rg;
    std::vector<uint8_t> comp;
    while (r < rg_end) {
      std::vector<uint32_t> clen;
      std::vector<uint8_t> dict;  // compressed rows in row order
      uint32_t maxc = 0, maxlen = 0;
      uint64_t r0 = r;
      while (r < rg_end) {
        comp.clear();
        uint64_t len = offs[r + 1] - offs[r];
        fsst_compress_row(st, chars + offs[r], len, comp);
        uint32_t mc = std::max<uint32_t>(maxc, static_cast<uint32_t>(comp.size()));
        uint32_t w  = bit_width_u64(mc);
        uint64_t total = 16 + packed_bytes(clen.size() + 1, w) + symtab.size() + dict.size() + comp.size();
        if (total > block && !clen.empty()) break;
        maxc = mc;
        maxlen = std::max<uint32_t>(maxlen, static_cast<uint32_t>(len));
        clen.push_back(static_cast<uint32_t>(comp.size()));
        dict.insert(dict.end(), comp.begin(), comp.end());
        ++r;
      }
      uint32_t w        = bit_width_u64(maxc);
      uint64_t bp       = packed_bytes(clen.size(), w);
      uint64_t sym_off  = 16 + bp;
      uint64_t dict_beg = sym_off + symtab.size();
      uint64_t total    = dict_beg + dict.size();
      std::vector<uint8_t> seg;
      seg.reserve(total);
      seg.resize(16, 0);
      std::vector<uint64_t> cl(clen.begin(), clen.end());
      pack_append(seg, cl.data(), cl.size(), w);
      seg.insert(seg.end(), symtab.begin(), symtab.end());
      seg.resize(total, 0);
      // row i occupies [dict_end - cum[i], dict_end - cum[i-1])
      uint64_t cum = 0, src = 0;
      for (size_t i = 0; i < clen.size(); ++i) {
        cum += clen[i];
        std::memcpy(seg.data() + total - cum, dict.data() + src, clen[i]);
        src += clen[i];
      }
      put<uint32_t>(seg, 0, static_cast<uint32_t>(dict.size()));
      put<uint32_t>(seg, 4, static_cast<uint32_t>(total));
      put<uint32_t>(seg, 8, w);
      put<uint32_t>(seg, 12, static_cast<uint32_t>(sym_off));
      add_segment(e, seg, r0, static_cast<uint32_t>(clen.size()), maxlen);
    }
  }
}

// ---------------------------------------------------------------------------
// DICT_FSST — DuckDB's dictionary codec over FSST-compressed entries
// (layout of src/cuda/scan/strings/dict_fsst.cu:17-38 and the reference's
// test synthesizer, test/cpp/scan/strings_synth.hpp:437-549): a 16-byte
// header {dict_size, dict_count, mode, string_lengths_width,
// dictionary_indices_width, pad, symbol_table_size}, the dictionary entries
// front to back from byte 16 (entry 0 is the empty NULL slot), the FSST
// symbol table (modes 1 and 2), the bit-packed entry lengths and, except in
// mode 2, the bit-packed row indices, every region 8-byte aligned.
// Mode 0 keeps entries raw, mode 1 FSST-compresses them, mode 2 has one
// entry per row (row i is entry i + 1) and no index region.
// ---------------------------------------------------------------------------

// Values of `w` bits LSB-first into exactly ceil(n * w / 8) bytes at `dst`.
void pack_tight(uint8_t* dst, const std::vector<uint32_t>& v, uint32_t w)
{
  if (w == 0) return;
  uint64_t bit = 0;
  for (uint32_t x : v) {
    for (uint32_t b = 0; b < w; ++b, ++bit)
      if ((x >> b) & 1u) dst[bit >> 3] |= static_cast<uint8_t>(1u << (bit & 7u));
  }
}

inline uint32_t al8(uint64_t v) { return static_cast<uint32_t>((v + 7u) & ~uint64_t{7}); }

// One segment over rows [r0, r1); NULL rows (valid[r] == 0) select entry 0.
std::vector<uint8_t> dict_fsst_segment(const uint8_t* chars, const uint64_t* offs, const uint8_t* valid,
                                       uint64_t r0, uint64_t r1, uint32_t mode, uint64_t seed, uint32_t* max_len)
{
  std::vector<std::string> dict(1);  // entry 0: the NULL slot
  std::vector<uint32_t> sel;
  std::unordered_map<std::string, uint32_t> at;
  *max_len = 0;
  for (uint64_t r = r0; r < r1; ++r) {
    std::string v(reinterpret_cast<const char*>(chars + offs[r]), offs[r + 1] - offs[r]);
    *max_len = std::max<uint32_t>(*max_len, static_cast<uint32_t>(v.size()));
    if (mode == 2) {
      dict.push_back(std::move(v));
      continue;
    }
    if (valid && !valid[r]) {
      sel.push_back(0);
      continue;
    }
    auto it = at.find(v);
    if (it == at.end()) {
      at.emplace(v, static_cast<uint32_t>(dict.size()));
      sel.push_back(static_cast<uint32_t>(dict.size()));
      dict.push_back(std::move(v));
    } else {
      sel.push_back(it->second);
    }
  }
  const uint32_t dict_count = static_cast<uint32_t>(dict.size());
  std::vector<std::vector<uint8_t>> entry(dict_count);
  std::vector<uint8_t> symtab;
  if (mode != 0 && dict_count > 1) {
    Rng rng(seed);
    std::vector<const uint8_t*> sp;
    std::vector<uint32_t> sl;
    uint64_t budget = 16384;
    for (uint32_t t = 0; t < 4096 && budget > 0; ++t) {
      const std::string& e = dict[1 + rng.next() % (dict_count - 1)];
      sp.push_back(reinterpret_cast<const uint8_t*>(e.data()));
      sl.push_back(static_cast<uint32_t>(e.size()));
      budget = budget > e.size() ? budget - e.size() : 0;
    }
    SymTable st = fsst_build(sp, sl);
    symtab      = fsst_export(st);
    for (uint32_t k = 1; k < dict_count; ++k)
      fsst_compress_row(st, reinterpret_cast<const uint8_t*>(dict[k].data()), dict[k].size(), entry[k]);
  } else {
    for (uint32_t k = 0; k < dict_count; ++k) entry[k].assign(dict[k].begin(), dict[k].end());
  }
  uint32_t max_entry = 0;
  uint64_t dict_size = 0;
  std::vector<uint32_t> lens(dict_count);
  for (uint32_t k = 0; k < dict_count; ++k) {
    lens[k] = static_cast<uint32_t>(entry[k].size());
    max_entry = std::max(max_entry, lens[k]);
    dict_size += lens[k];
  }
  const uint32_t rows = static_cast<uint32_t>(r1 - r0);
  const uint32_t slw  = std::max(1u, bit_width_u64(max_entry));
  const uint32_t diw  = std::max(1u, bit_width_u64(dict_count - 1u));
  const uint32_t off_dict   = 16;
  const uint32_t off_symtab = mode == 0 ? 0u : al8(off_dict + dict_size);
  const uint32_t off_slens  = mode == 0 ? off_dict + al8(dict_size) : al8(uint64_t{off_symtab} + symtab.size());
  const uint32_t off_didx   = al8(off_slens + (uint64_t{dict_count} * slw + 7u) / 8u);
  const uint32_t total      = mode == 2 ? off_didx : al8(off_didx + (uint64_t{rows} * diw + 7u) / 8u);
  std::vector<uint8_t> seg(total, 0);
  put<uint32_t>(seg, 0, static_cast<uint32_t>(dict_size));
  put<uint32_t>(seg, 4, dict_count);
  seg[8]  = static_cast<uint8_t>(mode);
  seg[9]  = static_cast<uint8_t>(slw);
  seg[10] = static_cast<uint8_t>(diw);
  put<uint32_t>(seg, 12, static_cast<uint32_t>(symtab.size()));
  uint64_t cur = off_dict;
  for (uint32_t k = 0; k < dict_count; ++k) {
    if (!entry[k].empty()) std::memcpy(seg.data() + cur, entry[k].data(), entry[k].size());
    cur += entry[k].size();
  }
  if (!symtab.empty()) std::memcpy(seg.data() + off_symtab, symtab.data(), symtab.size());
  pack_tight(seg.data() + off_slens, lens, slw);
  if (mode != 2) pack_tight(seg.data() + off_didx, sel, diw);
  return seg;
}

void dict_fsst_encode(sdr_encoded* e, const uint8_t* chars, const uint64_t* offs, const uint8_t* valid, uint64_t n,
                      uint32_t mode, uint32_t seg_rows, uint32_t block, uint64_t seed)
{
  uint64_t r = 0;
  while (r < n) {
    uint64_t want = std::min<uint64_t>(seg_rows, n - r);
    for (;;) {
      uint32_t ml = 0;
      std::vector<uint8_t> seg = dict_fsst_segment(chars, offs, valid, r, r + want, mode, seed ^ r, &ml);
      if (seg.size() <= block || want == 1) {
        add_segment(e, seg, r, static_cast<uint32_t>(want), ml);
        r += want;
        break;
      }
      want = std::max<uint64_t>(1, want / 2);
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

sdr_encoded* sdr_encode_dict_fsst(const uint8_t* chars, const uint64_t* offsets, const uint8_t* valid, uint64_t n,
                                  uint32_t mode, uint32_t seg_rows, uint32_t block_size, uint64_t seed)
{
  if (mode > 2 || seg_rows == 0) return nullptr;
  auto* e = new sdr_encoded();
  dict_fsst_encode(e, chars, offsets, valid, n, mode, seg_rows, block_size, seed);
  return e;
}

sdr_encoded* sdr_encode_bitpacking(const void* values, uint64_t n, uint32_t ts, int is_signed,
                                   uint32_t rg_rows, uint32_t block, int force_mode)
{
  auto* e = new sdr_encoded();
  switch (ts * 2 + (is_signed ? 1 : 0)) {
    case 2: bp_encode<uint8_t>(e, static_cast<const uint8_t*>(values), n, rg_rows, block, force_mode); break;
    case 3: bp_encode<int8_t>(e, static_cast<const int8_t*>(values), n, rg_rows, block, force_mode); break;
    case 4: bp_encode<uint16_t>(e, static_cast<const uint16_t*>(values), n, rg_rows, block, force_mode); break;
    case 5: bp_encode<int16_t>(e, static_cast<const int16_t*>(values), n, rg_rows, block, force_mode); break;
    case 8: bp_encode<uint32_t>(e, static_cast<const uint32_t*>(values), n, rg_rows, block, force_mode); break;
    case 9: bp_encode<int32_t>(e, static_cast<const int32_t*>(values), n, rg_rows, block, force_mode); break;
    case 16: bp_encode<uint64_t>(e, static_cast<const uint64_t*>(values), n, rg_rows, block, force_mode); break;
    case 17: bp_encode<int64_t>(e, static_cast<const int64_t*>(values), n, rg_rows, block, force_mode); break;
    default: delete e; return nullptr;
  }
  return e;
}

sdr_encoded* sdr_encode_rle(const void* values, uint64_t n, uint32_t ts, uint32_t rg_rows, uint32_t block)
{
  auto* e = new sdr_encoded();
  switch (ts) {
    case 1: rle_encode<uint8_t>(e, static_cast<const uint8_t*>(values), n, rg_rows, block); break;
    case 2: rle_encode<uint16_t>(e, static_cast<const uint16_t*>(values), n, rg_rows, block); break;
    case 4: rle_encode<uint32_t>(e, static_cast<const uint32_t*>(values), n, rg_rows, block); break;
    case 8: rle_encode<uint64_t>(e, static_cast<const uint64_t*>(values), n, rg_rows, block); break;
    default: delete e; return nullptr;
  }
  return e;
}

sdr_encoded* sdr_encode_alp(const void* values, uint64_t n, uint32_t ts, uint32_t rg_rows, uint32_t block)
{
  auto* e = new sdr_encoded();
  if (ts == 8) alp_encode<double>(e, static_cast<const double*>(values), n, rg_rows, block);
  else if (ts == 4) alp_encode<float>(e, static_cast<const float*>(values), n, rg_rows, block);
  else { delete e; return nullptr; }
  return e;
}

sdr_encoded* sdr_encode_alprd(const void* values, uint64_t n, uint32_t ts, uint32_t rg_rows, uint32_t block)
{
  auto* e = new sdr_encoded();
  if (ts == 8) alprd_encode<double>(e, static_cast<const double*>(values), n, rg_rows, block);
  else if (ts == 4) alprd_encode<float>(e, static_cast<const float*>(values), n, rg_rows, block);
  else { delete e; return nullptr; }
  return e;
}

sdr_encoded* sdr_encode_uncompressed(const void* values, uint64_t n, uint32_t ts, uint32_t rg_rows, uint32_t block)
{
  auto* e = new sdr_encoded();
  uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(rg_rows, block / ts));
  const uint8_t* p = static_cast<const uint8_t*>(values);
  for (uint64_t r = 0; r < n; r += per) {
    uint64_t c = std::min<uint64_t>(per, n - r);
    std::vector<uint8_t> seg(p + r * ts, p + (r + c) * ts);
    add_segment(e, seg, r, static_cast<uint32_t>(c));
  }
  return e;
}

sdr_encoded* sdr_encode_fsst(const uint8_t* chars, const uint64_t* offsets, uint64_t n, uint32_t rg_rows,
                             uint32_t block, uint64_t seed)
{
  auto* e = new sdr_encoded();
  fsst_encode(e, chars, offsets, n, rg_rows, block, seed);
  return e;
}

const uint8_t* sdr_encoded_bytes(const sdr_encoded* e, uint64_t* size)
{
  *size = e->blob.size();
  return e->blob.data();
}
const sdr_enc_segment* sdr_encoded_segments(const sdr_encoded* e, uint32_t* n)
{
  *n = static_cast<uint32_t>(e->segs.size());
  return e->segs.data();
}
void sdr_encoded_free(sdr_encoded* e) { delete e; }

void sdr_gen_u64(uint64_t seed, uint64_t n, uint64_t* out)
{
  Rng r(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = r.next();
}

void sdr_gen_bp_int32(uint64_t seed, uint64_t n, int32_t* out)
{
  Rng r(seed);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t u   = r.u32();
    bool narrow  = ((i / 2048u) % 10u) == 9u;
    out[i]       = static_cast<int32_t>(narrow ? (u & 31u) : (u & 0x1FFFFu));
  }
}

void sdr_gen_runs_int64(uint64_t seed, uint64_t n, double mean_run, uint64_t value_mod, int64_t* out)
{
  Rng r(seed);
  double p  = 1.0 / mean_run;
  double lq = std::log1p(-p);
  uint64_t i = 0;
  while (i < n) {
    double u   = 1.0 - r.unit();  // (0, 1]
    uint64_t L = 1 + static_cast<uint64_t>(std::floor(std::log(u) / lq));
    int64_t v  = static_cast<int64_t>(r.next() % value_mod);
    for (uint64_t k = 0; k < L && i < n; ++k) out[i++] = v;
  }
}

void sdr_gen_uniform_int32(uint64_t seed, uint64_t n, int32_t lo, uint32_t range, int32_t* out)
{
  Rng r(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = lo + static_cast<int32_t>(r.u32() % range);
}

void sdr_gen_dates_sorted(uint64_t seed, uint64_t n, int32_t* out)
{
  sdr_gen_uniform_int32(seed, n, 8035, 2557, out);
  for (uint64_t g = 0; g < n; g += 2048) std::sort(out + g, out + std::min<uint64_t>(n, g + 2048));
}

void sdr_gen_alp_doubles(uint64_t seed, uint64_t n, double* out)
{
  Rng r(seed);
  for (uint64_t i = 0; i < n; ++i) {
    if (r.next() % 100u == 0) out[i] = r.unit() * 1e6;
    else out[i] = static_cast<double>(90000 + r.next() % 10410001u) / 100.0;
  }
}

void sdr_gen_unit_doubles(uint64_t seed, uint64_t n, double* out)
{
  Rng r(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = r.unit();
}

uint64_t sdr_gen_zipf_strings(uint64_t seed, uint64_t n, uint32_t min_len, uint32_t max_len, double zipf_s,
                              uint64_t* offsets, uint8_t* chars)
{
  Rng r(seed);
  constexpr int NW = 2000;
  std::vector<std::string> words(NW);
  for (auto& w : words) {
    uint32_t l = 3 + r.u32() % 8u;
    w.resize(l);
    for (auto& c : w) c = static_cast<char>('a' + r.u32() % 26u);
  }
  std::vector<double> cdf(NW);
  double acc = 0;
  for (int k = 0; k < NW; ++k) {
    acc += 1.0 / std::pow(k + 1.0, zipf_s);
    cdf[k] = acc;
  }
  uint64_t total = 0;
  std::string s;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t L = min_len + r.u32() % (max_len - min_len + 1u);
    s.clear();
    while (s.size() < L) {
      double u = r.unit() * acc;
      int k    = static_cast<int>(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
      if (k >= NW) k = NW - 1;
      if (!s.empty()) s.push_back(' ');
      s += words[k];
    }
    s.resize(L);
    if (offsets) offsets[i] = total;
    if (chars) std::memcpy(chars + total, s.data(), L);
    total += L;
  }
  if (offsets) offsets[n] = total;
  return total;
}

}  // extern "C"
