// dec_fsst.hpp — gfx950 FSST chunk decoder: VARCHAR segments to int32
// offsets + chars.  Used by the per-column kernel (k_fsst.hip) and the
// persistent mixed-codec kernel (k_batch.hip).
//
// Segment layout, header checks and symbol-table format: the reference's
// src/cuda/scan/strings/fsst.cu:17-44, :69-87 and
// src/include/cuda/scan/detail/fsst.cuh:50-96 (a device port of DuckDB's
// duckdb_fsst_import).  Decoded lengths and bytes follow DuckDB's serial
// decompressor: a code emits its symbol, an escape (255) emits the next byte
// literally, a trailing escape is dropped (fsst.cu:273-281).
//
// One 256-thread workgroup decodes one chunk of up to FSST_CHUNK_ROWS rows.
// Chunks are numbered in column row order and claimed in that order (an
// atomic ticket or the position in a claimed work queue), so a chunk only
// ever waits on chunks already running: each chunk publishes its
// decoded byte count, looks back over its predecessors for the column-wide
// exclusive prefix (a decoupled look-back), writes its rows' int32 offsets
// and then its chars.
//
// The chunk's compressed bytes are staged in LDS once and walked twice, in
// storage order, by spans: thread t takes the t-th 16-byte-aligned slice of
// the staged bytes whatever rows it cuts, so every lane of a wave does the
// same number of byte steps however the row lengths vary.  The first walk
// sums decoded lengths and records, at every 16-byte grid point, the walk
// state (decoded prefix, escape flag, row); the second walk restarts from
// those states and ORs each symbol into a zeroed LDS copy of the chunk's
// output, which leaves in 16-byte stores.  Chunks whose compressed bytes do
// not fit the stage, malformed ones and rows longer than the output window
// take a row-per-thread path that reads global memory.
#pragma once

#include "sdr_device.hpp"
#include "sdr_fsst.hpp"

namespace sdr {

constexpr uint32_t FSST_THREADS   = 256;
constexpr uint32_t FSST_WAVES     = FSST_THREADS / WAVE;
constexpr uint32_t ROWS_PER_THR   = FSST_CHUNK_ROWS / FSST_THREADS;
constexpr uint32_t FSST_PIECE     = 8192;  // LDS output piece of the row-wise path
constexpr uint32_t FSST_ESC       = 255;
constexpr uint32_t FSST_NSYM      = 255;
constexpr uint64_t FSST_VERSION   = 20190218ull;
constexpr uint64_t FSST_CORRUPT   = 32774747032022883ull;  // "corrupt\0"
constexpr uint32_t SYMTAB_MAX     = 8192;
constexpr uint64_t ST_AGG         = 1ull << 62;  // aggregate published
constexpr uint64_t ST_INC         = 2ull << 62;  // inclusive prefix published
constexpr uint64_t ST_VAL         = (1ull << 62) - 1;

// Bytes of a segment read in order through aligned words; bytes at or past
// `end` read as zero.
__device__ __forceinline__ uint32_t byte_at(const uint8_t* p, const uint8_t* end)
{
  if (p >= end) return 0;
  return *gbl(p);
}

// w-bit value at bit `bitpos` of the stream starting at `base`; bytes at or
// past `end` read as zero (only malformed segments reach them).
__device__ __forceinline__ uint32_t stream_bits32(const uint8_t* base, const uint8_t* end, uint64_t bitpos,
                                                  uint32_t w)
{
  if (w == 0) return 0;
  const uint8_t* p   = base + (bitpos >> 3);
  const uint32_t off = static_cast<uint32_t>(bitpos & 7u);
  const uint32_t nb  = (off + w + 7u) >> 3;  // bytes spanned, <= 5
  uint64_t v = 0;
  if (p + nb <= end) {
    uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint32_t sh = static_cast<uint32_t>(a & 3u);
    const gptr<const uint32_t> wp = gbl(reinterpret_cast<const uint32_t*>(a - sh));
    const uint32_t need = (sh + nb + 3u) >> 2;  // words spanned, <= 3
    uint32_t w0 = wp[0], w1 = need > 1 ? wp[1] : 0u, w2 = need > 2 ? wp[2] : 0u;
    const uint64_t lo = static_cast<uint64_t>(w0) | (static_cast<uint64_t>(w1) << 32);
    v = (lo >> (8u * sh)) | (sh ? (static_cast<uint64_t>(w2) << (64u - 8u * sh)) : 0ull);
  } else {
    for (uint32_t b = 0; b < nb; ++b) v |= static_cast<uint64_t>(byte_at(p + b, end)) << (8u * b);
  }
  v >>= off;
  return static_cast<uint32_t>(w >= 32u ? v : (v & ((1ull << w) - 1ull)));
}

// Sequential reader of w-bit fields (w <= 32) from an LSB-first stream,
// loading each aligned word only when its bits are needed.
struct BitCursor {
  gptr<const uint32_t> w;
  uint64_t acc;
  uint32_t nbits;
  __device__ __forceinline__ BitCursor(const uint8_t* base, uint64_t bitpos)
  {
    uintptr_t a = reinterpret_cast<uintptr_t>(base + (bitpos >> 3));
    const uint32_t skip = static_cast<uint32_t>((a & 3u) * 8u + (bitpos & 7u));
    w     = gbl(reinterpret_cast<const uint32_t*>(a & ~uintptr_t{3}));
    acc   = static_cast<uint64_t>(*w++) >> skip;
    nbits = 32u - skip;
  }
  __device__ __forceinline__ uint32_t next(uint32_t wd)
  {
    if (nbits < wd) {
      acc |= static_cast<uint64_t>(*w++) << nbits;
      nbits += 32u;
    }
    const uint32_t v = static_cast<uint32_t>(wd >= 32u ? acc : (acc & ((1ull << wd) - 1ull)));
    acc >>= wd;
    nbits -= wd;
    return v;
  }
};

// Sum of the w-bit fields [i0, i1) of the stream at `base`; the fields must
// lie inside [base, end) or the slow bounded reader is used.
__device__ __forceinline__ uint32_t sum_fields(const uint8_t* base, const uint8_t* end, uint32_t i0, uint32_t i1,
                                               uint32_t w)
{
  if (i0 >= i1 || w == 0) return 0;
  uint32_t s = 0;
  if (base + ((uint64_t{i1} * w + 7u) >> 3) <= end) {
    BitCursor bc(base, uint64_t{i0} * w);
    for (uint32_t i = i0; i < i1; ++i) s += bc.next(w);
  } else {
    for (uint32_t i = i0; i < i1; ++i) s += stream_bits32(base, end, uint64_t{i} * w, w);
  }
  return s;
}

// Sequential byte reader over global bytes [p, p + n) using aligned 4-byte loads.
struct ByteReader {
  gptr<const uint32_t> w;
  uint32_t cur;
  uint32_t boff;
  __device__ __forceinline__ explicit ByteReader(const uint8_t* p)
  {
    boff = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(p) & 3u);
    w    = gbl(reinterpret_cast<const uint32_t*>(p - boff));
    cur  = *w;
  }
  __device__ __forceinline__ uint32_t next()
  {
    if (boff == 4u) {
      cur  = *++w;
      boff = 0;
    }
    return (cur >> (8u * boff++)) & 0xFFu;
  }
};

// A workgroup barrier for LDS alone: each wave's LDS operations complete
// first, global loads stay in flight.  (__syncthreads() also waits for every
// pending load while the compiler tracks LDS-direct loads, which it cannot
// tell apart from register loads.)
__device__ __forceinline__ void lds_barrier() { __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.  LdsOnly: its barriers leave
// global loads in flight.
template <bool LdsOnly = false>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t* total)
{
  const uint32_t incl = wave_incl_sum_u32(v);
  const uint32_t w    = threadIdx.x >> 6;
  if (lane_id() == 63) s_wave[w] = incl;
  if constexpr (LdsOnly) lds_barrier();
  else __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (uint32_t k = 0; k < FSST_WAVES; ++k) {
    const uint32_t x = s_wave[k];
    before += k < w ? x : 0u;
    tot += x;
  }
  if constexpr (LdsOnly) lds_barrier();
  else __syncthreads();
  *total = tot;
  return before + incl - v;
}

// Shared memory of one chunk decode.
constexpr uint32_t FSST_SYMSTAGE = 17u + 8u * FSST_NSYM;  // the largest symbol table (2057 bytes)
constexpr uint32_t FSST_CSTAGE   = 16384;                  // chunk compressed bytes staged in LDS
constexpr uint32_t FSST_NGRID    = (FSST_CSTAGE + 16u) / 16u + 2u;
// Output windows.  An 11 KiB window keeps a chunk kernel's LDS at 27.0 KB,
// and with at most 80 VGPRs it runs six workgroups per CU (1.6-3% faster
// than five with 16 KiB).  Both kernels decode one chunk per block: a
// persistent loop around the chunk took the batch kernel to 96 VGPRs.
constexpr uint32_t FSST_OUT_COLUMN = 11264;  // output window of the per-column kernel (6 blocks per CU)
constexpr uint32_t FSST_OUT_BATCH  = 11264;  // ... of the batch kernel (6 blocks per CU)
// 16-byte groups of a thread's first-walk span of a chunk's compressed bytes
// (at most FSST_CSTAGE of them after a head of up to 15 bytes), loaded into
// registers at the chunk's start
constexpr uint32_t SPAN_GROUPS = (FSST_CSTAGE + 15u + 16u * FSST_THREADS - 1u) / (16u * FSST_THREADS);
static_assert(SPAN_GROUPS == 5, "first-walk span registers");
// s_waitcnt vmcnt(SPAN_GROUPS), expcnt and lgkmcnt left at their maxima
// (gfx9 encoding: vmcnt bits 3:0, expcnt 6:4, lgkmcnt 11:8)
constexpr int WAITCNT_VMEM_SPANS = 0x0F70 | static_cast<int>(SPAN_GROUPS);
// What a chunk without compressed bytes loads as its span (never used).
static __device__ __attribute__((aligned(16))) const uint32_t fsst_zero_group[4] = {0u, 0u, 0u, 0u};

// Walk state at a grid point: decoded prefix (bits 0-19), escape flag (bit
// 20), row (bits 21-31).  A staged chunk decodes to at most 8 * FSST_CSTAGE
// bytes and holds at most 1024 rows, so the fields never overflow.
constexpr uint32_t GS_ESC   = 1u << 20;
constexpr uint32_t GS_ROW   = 21;
constexpr uint32_t GS_S     = GS_ESC - 1u;
constexpr uint32_t ROW_HUGE = 1u << 20;  // a staged row is shorter than this

template <uint32_t OUT_CAP>
struct FsstSmemT {
  // the segment's FsstSymTab, staged whole
  __attribute__((aligned(16))) uint64_t sym[256];  // symbol bytes, zero above its length
  uint8_t len[256];
  uint32_t comp[FSST_CHUNK_ROWS + 1];  // segment-relative inclusive compressed prefix
  uint32_t dec[FSST_CHUNK_ROWS + 1];   // chunk-relative exclusive decoded offset
  uint32_t gst[FSST_NGRID];            // walk state at each 16-byte grid point of the stage
  uint32_t part[FSST_THREADS];         // decoded bytes before each thread's span
  // output window; before that, pieces of the compressed-length stream
  __attribute__((aligned(16))) uint8_t out[(OUT_CAP > FSST_PIECE ? OUT_CAP : FSST_PIECE) + 48];
  uint32_t wave[FSST_WAVES];
  uint64_t base;

  // The second walk's output window.  The chunk's compressed bytes are not
  // staged in LDS: the first walk holds each thread's span in registers and
  // the second reads the same 16-byte blocks again from L2.
  static constexpr uint32_t WINDOW = OUT_CAP > FSST_PIECE ? OUT_CAP : FSST_PIECE;
  __device__ __forceinline__ uint8_t* window()
  {
    // a window of nout <= WINDOW + 15 bytes (16-byte head) is cleared in 4-byte words below nout + 32
    static_assert((WINDOW + 15u + 31u) / 4u * 4u + 4u <= sizeof(out), "a window's clear stays inside out[]");
    return out;
  }
};

// Walk the n compressed bytes of one row with DuckDB's serial semantics:
// emit(byte) for a literal after an escape, sym(code) for a symbol; a
// trailing escape emits nothing.  `fetch(i)` returns byte i of the row.
template <typename Fetch, typename Lit, typename Sym>
__device__ __forceinline__ void fsst_walk(uint32_t n, Fetch fetch, Lit lit, Sym sym)
{
  bool esc = false;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t b = fetch(i);
    if (esc) {
      if (!lit(b)) return;
      esc = false;
    } else if (b == FSST_ESC) {
      esc = true;
    } else if (!sym(b)) {
      return;
    }
  }
}

// Decoded length of the n compressed bytes of one row at p (global memory),
// read a word at a time (only the aligned words holding the row).
template <class SM>
__device__ __forceinline__ uint32_t fsst_row_len(const uint8_t* p, uint32_t n, const SM& sm)
{
  const uint32_t boff = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(p) & 3u);
  const gptr<const uint32_t> w = gbl(reinterpret_cast<const uint32_t*>(p - boff));
  const uint32_t end  = boff + n;
  bool esc            = false;
  uint32_t d          = 0;
  for (uint32_t q = 0; q * 4u < end; ++q) {
    const uint32_t word = w[q];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const uint32_t idx = q * 4u + j;
      const uint32_t b   = (word >> (8u * j)) & 0xFFu;
      if (idx >= boff && idx < end) {
        d += esc ? 1u : sm.len[b];  // len[255] == 0: an escape emits nothing
        esc = !esc && b == FSST_ESC;
      }
    }
  }
  return d;
}

// ---- row-per-thread path (unstaged or malformed chunks) ----

// dec[] and the decoded total of a chunk, one row per thread at a time,
// reading the compressed bytes from global memory.  Rows whose compressed
// range leaves [0, dict_end) decode to nothing.
template <class SM>
__device__ __attribute__((noinline)) uint32_t fsst_lengths_rows(SM& sm, bool ok, const uint8_t* seg,
                                                                uint32_t dict_end, uint32_t rows)
{
  uint32_t T = 0;
#pragma unroll 1
  for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
    const uint32_t r = k * FSST_THREADS + threadIdx.x;
    uint32_t d       = 0;
    if (ok && r < rows) {
      const uint32_t cum = sm.comp[r + 1], prev = sm.comp[r];
      if (cum <= dict_end && prev <= cum && cum != prev) d = fsst_row_len(seg + dict_end - cum, cum - prev, sm);
    }
    uint32_t tot      = 0;
    const uint32_t ex = block_excl_scan(d, sm.wave, &tot);
    if (r < rows) sm.dec[r] = T + ex;
    T += tot;
  }
  if (threadIdx.x == 0) sm.dec[rows] = T;
  __syncthreads();
  return T;
}

// Chars of a chunk from dec[], one row per thread, in 8 KiB LDS pieces
// aligned to the absolute chars address: whole 16-byte blocks leave with
// vector stores, the partial first/last blocks byte-wise (they share 16
// bytes with the neighbours).
template <class SM>
__device__ __attribute__((noinline)) void fsst_emit_rows(SM& sm, const uint8_t* seg, uint32_t dict_end,
                                                        uint32_t rows, uint32_t T, uint64_t base, uint8_t* chars,
                                                        uint64_t chars_cap)
{
  const uint32_t tid = threadIdx.x;
  const uintptr_t cb = reinterpret_cast<uintptr_t>(chars);
  const uintptr_t g0 = cb + base;
  const uintptr_t g1 = cb + ((base + T < chars_cap) ? base + T : chars_cap);
  constexpr int32_t PIECE = static_cast<int32_t>(FSST_PIECE);
  for (uintptr_t pa = g0 & ~uintptr_t{15}; pa < g1; pa += FSST_PIECE) {
    const uintptr_t pb = pa + FSST_PIECE;
#pragma unroll 1
    for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
      const uint32_t r = k * FSST_THREADS + tid;
      if (r >= rows) continue;
      const uint32_t d0 = sm.dec[r], dl = sm.dec[r + 1] - d0;
      const uintptr_t o0 = g0 + d0;
      if (dl == 0 || o0 + dl <= pa || o0 >= pb) continue;
      const uint32_t cum = sm.comp[r + 1], prev = sm.comp[r];
      int32_t o = static_cast<int32_t>(static_cast<intptr_t>(o0 - pa));  // may start before the piece
      auto put = [&](uint64_t v, uint32_t l) {
        for (uint32_t j = 0; j < l; ++j, ++o)
          if (o >= 0 && o < PIECE) sm.out[o] = static_cast<uint8_t>(v >> (8u * j));
        return o < PIECE;
      };
      ByteReader br(seg + dict_end - cum);
      fsst_walk(cum - prev, [&](uint32_t) -> uint32_t { return br.next(); },
                [&](uint32_t b) { return put(b, 1u); }, [&](uint32_t b) { return put(sm.sym[b], sm.len[b]); });
    }
    __syncthreads();
    const uintptr_t lo = pa > g0 ? pa : g0;
    const uintptr_t hi = pb < g1 ? pb : g1;
    for (uintptr_t blk = pa + uintptr_t{tid} * 16u; blk < hi; blk += uintptr_t{FSST_THREADS} * 16u) {
      if (blk >= lo && blk + 16u <= hi) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(&sm.out[blk - pa]);
        __builtin_nontemporal_store(v, gbl(reinterpret_cast<u32x4*>(blk)));
      } else {
        for (uintptr_t x = blk; x < blk + 16u; ++x)
          if (x >= lo && x < hi) *gbl(reinterpret_cast<uint8_t*>(x)) = sm.out[x - pa];
      }
    }
    __syncthreads();
  }
}

// ---- span path (staged chunks) ----
//
// Stage coordinates: the chunk's compressed bytes sit at stage positions [c_h, c_h + C);
// row r occupies [top - comp[r + 1], top - comp[r]) with top = c_h + c_hi
// (rows are stored back to front, so row indices fall as the position
// rises).  A walk's row variable r is the row holding the byte it is about
// to read; `rend` is that row's end, and at rend the walk steps to the next
// lower row (several, over empty rows).

// Span of thread t over [lo, hi), cut on a 16-byte grid starting at g0.
__device__ __forceinline__ uint32_t span_bytes(uint32_t g0, uint32_t hi)
{
  return ((hi - g0 + 16u * FSST_THREADS - 1u) / (16u * FSST_THREADS)) * 16u;
}

// Byte j (0..15) of a 16-byte group.
__device__ __forceinline__ uint32_t group_byte(const u32x4& q, uint32_t j)
{
  return (q[j >> 2] >> (8u * (j & 3u))) & 0xFFu;
}

// First walk: decoded lengths.  Records the walk state at each grid point
// and, per row, the decoded prefix at its start; returns the chunk's decoded
// total and leaves dec[] as the chunk-relative exclusive decoded offsets.
// Thread t's first-walk span: the 16-byte groups [t * span, t * span + span)
// of stage positions below y_hi, stage position 0 being the 16-byte-aligned
// address `stage` (the blocks an LDS stage of the chunk would hold).
//
// Every thread issues exactly SPAN_GROUPS loads (y_hi > 0), with no branch
// around them: a group outside its span rereads the stage's last group and
// is never used.  The caller can then wait for the loads issued before these
// with vmcnt(SPAN_GROUPS).
__device__ __forceinline__ void fsst_span_load(u32x4 (&qv)[SPAN_GROUPS], const uint8_t* stage, uint32_t y_hi)
{
  const uint32_t span = span_bytes(0, y_hi);
  const uint32_t s0   = threadIdx.x * span;
  const uint32_t last = (y_hi - 1u) >> 4;
  const gptr<const u32x4> cg = gbl(reinterpret_cast<const u32x4*>(stage));
#pragma unroll
  for (uint32_t k = 0; k < SPAN_GROUPS; ++k)
    qv[k] = cg[16u * k < span && s0 + 16u * k < y_hi ? (s0 >> 4) + k : last];
}

// The last word of thread t's span (the four bytes before thread t + 1's
// first byte), for the escape parity of the thread after it.
__device__ __forceinline__ uint32_t fsst_span_tail(const u32x4 (&qv)[SPAN_GROUPS], uint32_t y_hi)
{
  const uint32_t last = span_bytes(0, y_hi) / 16u - 1u;
  uint32_t w          = 0;
#pragma unroll
  for (uint32_t k = 0; k < SPAN_GROUPS; ++k) w = k == last ? qv[k][3] : w;
  return w;
}

template <class SM>
__device__ __forceinline__ uint32_t fsst_lengths_staged(SM& sm, uint32_t rows, uint32_t c_h, uint32_t c_hi,
                                                        uint32_t C, const u32x4 (&qv)[SPAN_GROUPS],
                                                        const uint8_t* stage)
{
  const uint32_t tid  = threadIdx.x;
  const uint32_t y_hi = c_h + C;
  const uint32_t top  = c_h + c_hi;
  const uint32_t span = span_bytes(0, y_hi);
  const uint32_t ya   = tid * span > c_h ? tid * span : c_h;
  const uint32_t yb   = tid * span + span < y_hi ? tid * span + span : y_hi;
  uint32_t acc        = 0;
  if (ya < yb) {
    // the row holding byte ya: the largest r <= rows with comp[r] <= top - ya
    // (r == rows, a row above the first, when ya is the first byte)
    const uint32_t u0 = top - ya;
    uint32_t r        = 0;
#pragma unroll
    for (uint32_t s = 1024; s; s >>= 1)
      if (r + s <= rows && sm.comp[r + s] <= u0) r += s;
    uint32_t rend  = top - sm.comp[r];
    uint32_t rnext = r ? top - sm.comp[r - 1] : rend;  // end of row r - 1
    bool esc       = false;
    if (r < rows) {  // ya inside row r: an odd run of 255s before it escapes it
      // the four bytes before ya come from the previous thread's span (sm.part,
      // written before the barrier that ends the compressed-prefix scan),
      // earlier ones from L2
      const uint32_t rs = top - sm.comp[r + 1];
      const uint32_t pw = tid ? sm.part[tid - 1] : 0u;
      for (uint32_t y = ya; y > rs; --y) {
        const uint32_t d = ya - y;
        const uint32_t b = d < 4u ? (pw >> (24u - 8u * d)) & 0xFFu : ld_u8(stage + y - 1u);
        if (b != FSST_ESC) break;
        esc = !esc;
      }
    }
    u32x4 qr[SPAN_GROUPS];
#pragma unroll
    for (uint32_t k = 0; k < SPAN_GROUPS; ++k) qr[k] = qv[k];
#pragma unroll 1
    for (uint32_t g = ya & ~15u; g < yb; g += 16) {
      const u32x4 q = qr[0];
#pragma unroll
      for (uint32_t k = 0; k + 1 < SPAN_GROUPS; ++k) qr[k] = qr[k + 1];
      uint32_t L[16];
#pragma unroll
      for (uint32_t j = 0; j < 16; ++j) L[j] = sm.len[group_byte(q, j)];
      const uint32_t lo = ya > g ? ya - g : 0u;  // the group's bytes of this span: [lo, hi)
      const uint32_t hi = yb - g < 16u ? yb - g : 16u;
#pragma unroll
      for (uint32_t j = 0; j < 16; ++j) {
        const uint32_t y = g + j;
        if (y == rend && r) {  // rows starting at y
          do {
            --r;
            sm.dec[r + 1] = acc;  // the prefix at row r's start
            rend          = rnext;
            rnext         = r ? top - sm.comp[r - 1] : rend;
          } while (y == rend && r);
          esc = false;
        }
        if (j == 0 && lo == 0) sm.gst[g >> 4] = acc | (esc ? GS_ESC : 0u) | (r << GS_ROW);
        const uint32_t b  = group_byte(q, j);
        const bool valid  = j >= lo && j < hi;
        acc += valid ? (esc ? 1u : L[j]) : 0u;
        esc = valid ? (!esc && b == FSST_ESC) : esc;
      }
    }
    // rows starting at the last byte's end (empty rows below the others)
    if (yb == y_hi)
      while (rend == y_hi && r) {
        --r;
        sm.dec[r + 1] = acc;
        rend          = rnext;
        rnext         = r ? top - sm.comp[r - 1] : rend;
      }
  }
  uint32_t T         = 0;
  const uint32_t pre = block_excl_scan(acc, sm.wave, &T);
  sm.part[tid]       = pre;
  if (ya < yb)
    for (uint32_t g = (ya + 15u) & ~15u; g < yb; g += 16) sm.gst[g >> 4] += pre;
  __syncthreads();
  // dec[r + 1] = T - (decoded prefix at row r's start); the prefix was
  // recorded at dec[r + 1], span-relative, by the thread whose span holds
  // that start, so each slot is read and rewritten by one thread here
#pragma unroll
  for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
    const uint32_t r = k * FSST_THREADS + tid;
    if (r < rows) {
      const uint32_t p = top - sm.comp[r + 1];
      sm.dec[r + 1]    = T - (sm.dec[r + 1] + sm.part[(p < y_hi ? p : y_hi - 1u) / span]);
    }
  }
  if (tid == 0) sm.dec[0] = 0;
  __syncthreads();
  return T;
}

// One byte of the second walk for a lane whose next row start is in
// registers or LDS (the per-byte form, used when a 16-byte group holds more
// row starts than the prefetch covers).  `y` is the byte's stage position.
template <class SM>
__device__ __forceinline__ void emit_byte_slow(SM& sm, uint32_t y, uint32_t b, bool valid, uint32_t top,
                                               uint32_t wofs, uint32_t& r, uint32_t& rend, uint32_t& o, bool& esc)
{
  if (valid && y == rend) {  // rows starting at y
    do {
      --r;
      rend = top - sm.comp[r];
    } while (y == rend);
    o   = sm.dec[r] - wofs;
    esc = false;
  }
  const bool lit   = esc;
  const uint32_t l = lit ? 1u : sm.len[b];
  const uint64_t v = lit ? uint64_t{b} : sm.sym[b];
  esc              = !lit && b == FSST_ESC;
  if (valid) {
    const uint32_t sh = 8u * (o & 3u);
    const uint64_t w  = v << sh;
    uint32_t* d       = reinterpret_cast<uint32_t*>(&sm.window()[o & ~3u]);
    __hip_atomic_fetch_or(d, static_cast<uint32_t>(w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_or(d + 1, static_cast<uint32_t>(w >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_or(d + 2, static_cast<uint32_t>((uint64_t{static_cast<uint32_t>(v >> 32)} << sh) >> 32),
                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  o += l;
}

constexpr uint32_t EMIT_AHEAD = 4;  // row starts a lane prefetches per 16-byte group

// Second walk: chars.  Rows are emitted in windows of whole rows whose
// output fits OUT_CAP bytes; each window's output is built in LDS (zeroed,
// then every symbol ORed in at its byte position, so neighbouring spans
// may share words) and leaves in 16-byte stores.  Returns false, having
// written nothing more, when a single row's output exceeds OUT_CAP.
//
// A lane walks its span 16 bytes at a time.  At each group it loads, with
// the group's symbols, the ends and output starts of the next EMIT_AHEAD
// rows, so the bytes themselves only compare and select: a row starting at
// byte j moves the output position to its start and clears the escape
// flag (empty rows start where the next row starts; the last match wins).
// A lane whose group holds more row starts than that takes the group byte
// by byte.
//
// The window is SM::WINDOW bytes of LDS; byte i of a window is chunk output
// position wlo + i, whatever the chunk's place in the column, so a window can
// be built before the chunk's base is known.  `get_base()` returns the base;
// every thread calls it after each window's ORs, and it synchronises the
// workgroup (after the first window's ORs the look-back completes there).
// The stores shift the window to the destination's 16-byte alignment.  The
// compressed bytes come from global memory at `stage`, the address of stage
// position 0 (16-byte blocks read again from L2).  A lane loads its first
// group while the window is cleared, and each next group while it emits the
// current one.
template <class SM, class GetBase>
__device__ __forceinline__ bool fsst_emit_staged(SM& sm, uint32_t rows, uint32_t c_h, uint32_t c_hi, uint32_t T,
                                                 GetBase&& get_base, uint8_t* chars, uint64_t chars_cap,
                                                 const uint8_t* stage)
{
  constexpr uint32_t OUT_CAP = SM::WINDOW;
  uint8_t* const wb         = sm.window();
  const gptr<const u32x4> cg = gbl(reinterpret_cast<const u32x4*>(stage));
  const uint32_t tid  = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t top  = c_h + c_hi;
  const uintptr_t cb  = reinterpret_cast<uintptr_t>(chars);
  uint32_t r_hi       = rows - 1;  // highest row of the next window
  // The lowest row of the window whose highest row is hi: the lowest r_lo
  // with dec[hi + 1] - dec[r_lo] <= OUT_CAP (~0u: row hi alone is larger).
  // Every wave searches for itself (the same answer from the same dec[]),
  // so no barrier hands it out.
  auto find_window = [&](uint32_t hi) -> uint32_t {
    const uint32_t d_top = sm.dec[hi + 1];
    if (d_top <= OUT_CAP) return 0u;
    if (d_top - sm.dec[hi] > OUT_CAP) return ~0u;
    // 64-way search rounds over (a, b] (fails at a, holds at b)
    uint32_t a = 0, b = hi;
    while (b - a > 1u) {
      const uint32_t step = (b - a + WAVE - 1u) / WAVE;
      const uint32_t m    = a + (lane + 1u) * step < b ? a + (lane + 1u) * step : b;
      const uint64_t hold = ballot(d_top - sm.dec[m] <= OUT_CAP);
      const uint32_t f    = static_cast<uint32_t>(__builtin_ctzll(hold));  // lane 63 tests b, which holds
      const uint32_t nb   = a + (f + 1u) * step < b ? a + (f + 1u) * step : b;
      a                   = a + f * step;
      b                   = nb;
    }
    return b;
  };
  uint32_t r_lo = find_window(r_hi);
  for (;;) {
    if (r_lo == ~0u) return false;
    const uint32_t wlo  = sm.dec[r_lo];
    const uint32_t whi  = sm.dec[r_hi + 1];
    {
      const uint32_t nout = whi - wlo;
      const uint32_t wofs = wlo;  // chunk output position of window byte 0
      const uint32_t ya   = top - sm.comp[r_hi + 1];
      const uint32_t yb   = top - sm.comp[r_lo];
      const uint32_t g0   = ya & ~15u;
      const uint32_t span = span_bytes(g0, yb);
      const uint32_t s0   = g0 + tid * span;
      const uint32_t ys   = tid ? s0 : ya;
      const uint32_t ye   = s0 + span < yb ? s0 + span : yb;
      u32x4 qn            = u32x4{0u, 0u, 0u, 0u};
      if (ys < ye) qn = cg[ys >> 4];
      // Cleared with 4-byte stores.  Measured on gfx950: clearing with
      // 16-byte stores (ds_write_b128) and then ORing from other waves
      // after a barrier lost about 0.2% of the ORed bytes once five or more
      // workgroups shared a CU; 4-byte stores never did.
      for (uint32_t i = tid * 4u; i < nout + 32u; i += FSST_THREADS * 4u)
        *reinterpret_cast<uint32_t*>(&wb[i]) = 0u;
      __syncthreads();
      // bytes outside [ys, ye) OR into this word, which is never flushed
      const uint32_t trash = (nout + 19u) & ~3u;
      if (ys < ye) {
        // thread 0 starts above the window's highest row and steps into it
        // at ya; the others resume a grid-point state of the first walk
        uint32_t r = r_hi + 1, o = 0;
        bool esc   = false;
        if (tid) {
          const uint32_t gs = sm.gst[ys >> 4];
          r                 = gs >> GS_ROW;
          esc               = (gs & GS_ESC) != 0;
          o = sm.dec[r] + (gs & GS_S) - (T - sm.dec[r + 1]) - wofs;
        }
        uint32_t rend = top - sm.comp[r];
#pragma unroll 1
        for (uint32_t g = ys & ~15u; g < ye; g += 16) {
          const u32x4 q = qn;
          if (g + 16u < ye) qn = cg[(g >> 4) + 1u];
          // the next EMIT_AHEAD rows below r: ends e[k] and output starts
          // d[k]; row r - k - 1 starts at e[k - 1] (e[-1] = rend)
          uint32_t e[EMIT_AHEAD], d[EMIT_AHEAD];
#pragma unroll
          for (uint32_t k = 0; k < EMIT_AHEAD; ++k) {
            const bool in    = r > k;
            const uint32_t i = in ? r - k - 1u : 0u;
            e[k]             = in ? top - sm.comp[i] : ~0u;
            d[k]             = sm.dec[i] - wofs;
          }
          const uint32_t lim = ye - g < 16u ? ye - g : 16u;  // bytes of the group in the span: [lo, lim)
          const uint32_t lo  = ys > g ? ys - g : 0u;
          const bool more    = e[EMIT_AHEAD - 1] - g < lim;  // another row starts inside the group
          if (!more) {
            uint32_t st[EMIT_AHEAD];  // start offsets in the group (>= 16 when none)
#pragma unroll
            for (uint32_t k = 0; k < EMIT_AHEAD; ++k) st[k] = (k ? e[k - 1] : rend) - g;
            // symbols and lengths in two halves of eight bytes (fewer live registers)
#pragma unroll
            for (uint32_t h = 0; h < 16; h += 8) {
            uint32_t L[8];
            uint64_t V[8];
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) {
              const uint32_t bj = group_byte(q, h + i);
              L[i]              = sm.len[bj];
              V[i]              = sm.sym[bj];
            }
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) {
              const uint32_t j = h + i;
              bool start = false;
#pragma unroll
              for (uint32_t k = 0; k < EMIT_AHEAD; ++k) {
                const bool s = st[k] == j;
                o            = s ? d[k] : o;
                start |= s;
              }
              const uint32_t bj = group_byte(q, j);
              const bool lit    = esc && !start;
              const uint32_t l  = lit ? 1u : L[i];
              const uint64_t v  = lit ? uint64_t{bj} : V[i];
              esc               = !lit && bj == FSST_ESC;
              const bool valid  = j >= lo && j < lim;
              const uint32_t sh = 8u * (o & 3u);
              const uint64_t w  = v << sh;
              const uint32_t w2 = static_cast<uint32_t>((uint64_t{static_cast<uint32_t>(v >> 32)} << sh) >> 32);
              uint32_t* dw      = reinterpret_cast<uint32_t*>(&wb[valid ? (o & ~3u) : trash]);
              __hip_atomic_fetch_or(dw, static_cast<uint32_t>(w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
              __hip_atomic_fetch_or(dw + 1, static_cast<uint32_t>(w >> 32), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
              __hip_atomic_fetch_or(dw + 2, w2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
              o += l;
            }
            }
            // rows started in the group
#pragma unroll
            for (uint32_t k = 0; k < EMIT_AHEAD; ++k)
              if (st[k] < lim) {
                r    = r - 1u;
                rend = e[k];
              }
          } else {
#pragma unroll 1
            for (uint32_t j = 0; j < 16; ++j)
              emit_byte_slow(sm, g + j, group_byte(q, j), j >= lo && j < lim, top, wofs, r, rend, o, esc);
          }
        }
      }
      const uint64_t base  = get_base();  // synchronises the workgroup
      const uintptr_t g_lo = cb + base + wlo;
      const uintptr_t g_hi = cb + (base + whi < chars_cap ? base + whi : chars_cap);
      const uint32_t* ww   = reinterpret_cast<const uint32_t*>(wb);
      for (uintptr_t blk = (g_lo & ~uintptr_t{15}) + uintptr_t{tid} * 16u; blk < g_hi;
           blk += uintptr_t{FSST_THREADS} * 16u) {
        if (blk >= g_lo && blk + 16u <= g_hi) {
          // window bytes [i, i + 16), i = blk - g_lo, from five aligned words
          const uint32_t i  = static_cast<uint32_t>(blk - g_lo);
          const uint32_t w  = i >> 2, sh = i & 3u;
          const uint32_t x0 = ww[w], x1 = ww[w + 1], x2 = ww[w + 2], x3 = ww[w + 3], x4 = ww[w + 4];
          const u32x4 v     = u32x4{__builtin_amdgcn_alignbyte(x1, x0, sh), __builtin_amdgcn_alignbyte(x2, x1, sh),
                                __builtin_amdgcn_alignbyte(x3, x2, sh), __builtin_amdgcn_alignbyte(x4, x3, sh)};
          __builtin_nontemporal_store(v, gbl(reinterpret_cast<u32x4*>(blk)));
        } else {
          for (uintptr_t x = blk; x < blk + 16u; ++x)
            if (x >= g_lo && x < g_hi) *gbl(reinterpret_cast<uint8_t*>(x)) = wb[x - g_lo];
        }
      }
    }
    __syncthreads();
    if (r_lo == 0) return true;
    r_hi = r_lo - 1;
    r_lo = find_window(r_hi);
  }
}

// Decoupled look-back for the column-wide chars offset of chunk c, by wave
// 0: publish this chunk's aggregate T (fsst_lookback_publish), then sum
// predecessors 64 at a time back to the nearest published inclusive prefix
// (fsst_lookback_wait).  Between the two the workgroup may do other work.  A
// chunk waits at most `spin_limit` sleeps for an unpublished predecessor;
// past that (or at once when spin_limit is 0, which only tests set) it sets
// the column's error flag and ends, so the host reports the failure and no
// wave hangs.  A status word carries its flag and value together and is the
// only datum exchanged between chunks, so device-scope relaxed atomics
// (coherent across the XCDs' L2s) suffice; release/acquire orderings would
// add an L2 write-back / invalidate per access.
__device__ __forceinline__ void fsst_lookback_publish(uint32_t c, uint64_t T, FsstState* st, uint32_t spin_limit)
{
  if (threadIdx.x == 0) {
    const gptr<uint64_t> status = gbl(st->status);
    if (c == 0) __hip_atomic_store(&status[0], ST_INC | T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else if (spin_limit != 0) __hip_atomic_store(&status[c], ST_AGG | T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The wait half: synchronises the workgroup and returns the exclusive prefix
// to every thread.
template <class SM>
__device__ __forceinline__ uint64_t fsst_lookback_wait(uint32_t c, uint32_t n_chunks, uint64_t T, FsstState* st,
                                                       uint32_t spin_limit, SM& sm)
{
  if (threadIdx.x < WAVE) {
    const uint32_t lane = threadIdx.x;
    const gptr<uint64_t> status = gbl(st->status);
    uint64_t excl64     = 0;
    if (c == 0) {
      // published inclusive already
    } else if (spin_limit == 0) {  // no waiting allowed: report the look-back as failed
      if (lane == 0) __hip_atomic_fetch_or(gbl(&st->error), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      uint32_t p = c - 1, spins = 0;
      for (;;) {
        const bool valid = lane <= p;
        const uint64_t s =
          valid ? __hip_atomic_load(&status[p - lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
        const uint64_t flag = s & ~ST_VAL;
        const uint64_t incm = ballot(valid && flag == ST_INC);
        const uint32_t stop = incm ? static_cast<uint32_t>(__builtin_ctzll(incm)) : (p < WAVE - 1 ? p : WAVE - 1);
        const uint64_t need = stop == WAVE - 1 ? ~0ull : ((2ull << stop) - 1ull);
        if (ballot(valid && flag == 0) & need) {  // a predecessor has not published yet
          if (++spins > spin_limit) {              // never expected: give up rather than hang
            if (lane == 0) __hip_atomic_fetch_or(gbl(&st->error), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl64 += readlane_u64(wave_incl_sum_u64(lane <= stop ? (s & ST_VAL) : 0ull), WAVE - 1);
        if (incm || p < WAVE) break;  // chunk 0 always publishes an inclusive prefix
        p -= WAVE;
      }
      if (lane == 0)
        __hip_atomic_store(&status[c], ST_INC | ((excl64 + T) & ST_VAL), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) {
      sm.base = excl64;
      if (c == n_chunks - 1) *gbl(&st->total) = excl64 + T;
    }
  }
  __syncthreads();
  return sm.base;
}

// Both halves with nothing between them.
template <class SM>
__device__ __forceinline__ uint64_t fsst_lookback(uint32_t c, uint32_t n_chunks, uint64_t T, FsstState* st,
                                                  uint32_t spin_limit, SM& sm)
{
  fsst_lookback_publish(c, T, st, spin_limit);
  return fsst_lookback_wait(c, n_chunks, T, st, spin_limit, sm);
}

// Column offsets of the chunk's rows (and the column's closing offset).
template <class SM>
__device__ __forceinline__ void fsst_offsets(const FsstChunk& ch, uint32_t c, uint32_t n_chunks, uint32_t rows,
                                             uint32_t T, uint64_t base, int32_t* offsets, uint32_t total_rows,
                                             const SM& sm)
{
  const gptr<int32_t> off = gbl(offsets);
  for (uint32_t i = threadIdx.x; i < rows; i += FSST_THREADS)
    off[ch.col_row + i] = static_cast<int32_t>(static_cast<uint32_t>(base + sm.dec[i]));
  if (c == n_chunks - 1 && threadIdx.x == 0) off[total_rows] = static_cast<int32_t>(static_cast<uint32_t>(base + T));
}

// ---- segment pre-pass ----

struct FsstPrepSmem {
  __attribute__((aligned(16))) uint8_t st[FSST_SYMSTAGE + 48];
  uint8_t len[256];  // code lengths of a DICT_FSST dictionary's table
  uint32_t sum[FSST_THREADS];
  uint32_t huge[FSST_THREADS];
  uint32_t wave[FSST_WAVES];
};

// Sum of the w-bit fields [f0, f1) of the length stream at `base`, and
// whether one of them reaches ROW_HUGE.  Fields past `end` read as zero.
__device__ __forceinline__ uint32_t fields_sum(const uint8_t* base, const uint8_t* end, uint32_t f0, uint32_t f1,
                                               uint32_t w, bool* huge)
{
  uint32_t s = 0, mx = 0;
  if (f0 < f1 && w) {
    if (base + ((uint64_t{f1} * w + 7u) >> 3) <= end) {
      BitCursor bc(base, uint64_t{f0} * w);
      for (uint32_t f = f0; f < f1; ++f) {
        const uint32_t v = bc.next(w);
        s += v;
        mx = v > mx ? v : mx;
      }
    } else {
      for (uint32_t f = f0; f < f1; ++f) {
        const uint32_t v = stream_bits32(base, end, uint64_t{f} * w, w);
        s += v;
        mx = v > mx ? v : mx;
      }
    }
  }
  *huge = mx >= ROW_HUGE;
  return s;
}

// ---- DICTIONARY and UNCOMPRESSED strings ----
//
// DICTIONARY (strings/dictionary.cu:17-33): a 20-byte header {dict_size,
// dict_end, index_buffer_offset, index_buffer_count, bitpacking_width}, the
// bit-packed selection values from byte 20 (0 is NULL), u32 cumulative
// entry ends at index_buffer_offset and entry bytes stored back to front
// ending at dict_end: entry s is [dict_end - idx[s], dict_end - idx[s-1]).
// The header is valid if the index buffer and dict_end lie inside the
// segment and the width is <= 32 (:57-64).
//
// UNCOMPRESSED (strings/uncompressed.cu:17-33): {dict_size, dict_end} then
// one int32 cumulative end per row from byte 8, read as |v| (the sign is
// DuckDB's in-memory marker); row i is [dict_end - |off[i]|, dict_end -
// |off[i-1]|).  The segment is valid if its offsets fit (:66); dict_end
// must also lie inside the segment here (the reference does not check it).
//
// Malformed input only: a row whose end passes dict_end or falls below its
// start decodes empty, where the reference would copy from outside the
// segment; a DICTIONARY selection >= index_buffer_count is empty in both of
// the reference's passes' intent (its length pass gives 0).

// Per-chunk info of a plain segment, one thread per chunk.
__device__ __forceinline__ void plain_prep_segment(const FsstSeg& sg, FsstChunkInfo* __restrict__ info)
{
  const uint8_t* seg = sg.bytes;
  const bool dict    = sg.codec == SDR_CODEC_DICTIONARY;
  uint32_t dict_end = 0, ibo = 0, ibc = 0, width = 0;
  bool ok = false;
  if (dict) {
    if (sg.bytes_size >= 20u) {
      dict_end = su32(seg + 4);
      ibo      = su32(seg + 8);
      ibc      = su32(seg + 12);
      width    = su32(seg + 16);
      ok = uint64_t{ibo} + 4ull * ibc <= sg.bytes_size && dict_end <= sg.bytes_size && width <= 32u;
    }
  } else if (uint64_t{sg.bytes_size} >= 8ull + 4ull * sg.row_count) {
    dict_end = su32(seg + 4);
    ok       = dict_end <= sg.bytes_size;
  }
  const uint32_t flags = (dict ? FI_DICT | (width & FI_WIDTH) : FI_UNC) | (ok ? FI_OK : 0u);
  for (uint32_t k = threadIdx.x; k * FSST_CHUNK_ROWS < sg.row_count; k += FSST_THREADS) {
    FsstChunkInfo ci;
    ci.bytes      = seg;
    ci.c_lo       = ibo;
    ci.c_hi       = ibc;
    ci.dict_end   = dict_end;
    ci.flags      = flags;
    ci.bytes_size = sg.bytes_size;
    ci.pad        = 0;
    const gptr<u32x4> dst = gbl(reinterpret_cast<u32x4*>(info + sg.first_chunk + k));
    const u32x8 cv        = __builtin_bit_cast(u32x8, ci);
    dst[0]                = cv.lo;
    dst[1]                = cv.hi;
  }
}

// ---- DICT_FSST strings ----
//
// strings/dict_fsst.cu:17-38: a 16-byte header, the dictionary entries front
// to back from byte 16 (entry 0 is the NULL slot), in the FSST modes (1, 2)
// a symbol table, the bit-packed entry lengths and, except in mode 2, the
// bit-packed row indices; mode 0 keeps entries raw, mode 1 FSST-compresses
// them, mode 2 has one entry per row (row i is entry i + 1).  The host reads
// the headers and fills DictFsstSeg (region offsets, :516-525); a pre-pass
// per segment builds the entry offsets (:160-235), a predecode pass writes
// the FSST entries once (:268-320), and the rows copy their entry like
// DICTIONARY rows (:246-265, :324-360).  Bytes past the segment read as
// zero; the byte a trailing escape counts is written as 0.

// Dictionary index of segment row sr (row + 1 in mode 2).
__device__ __forceinline__ uint32_t dfsst_index(const DictFsstSeg& d, uint32_t sr)
{
  if (d.mode == 2u) return sr + 1u;
  return stream_bits32(d.bytes + d.off_didx, d.bytes + d.bytes_size, uint64_t{sr} * d.diw, d.diw);
}

// Decoded length of one compressed entry of n bytes at p (bytes at or past
// `end` read as zero, i.e. as code 0): a code adds its symbol's length, an
// escape adds one and skips the next byte, even past the entry's end.
__device__ __forceinline__ uint32_t dfsst_entry_len(const uint8_t* p, uint32_t n, const uint8_t* end,
                                                    const uint8_t* len)
{
  const uint64_t in = p < end ? static_cast<uint64_t>(end - p) : 0u;
  const uint32_t n_in = n < in ? n : static_cast<uint32_t>(in);
  uint32_t pos = 0, dl = 0;
  while (pos < n_in) {
    const uint32_t code = *gbl(p + pos);
    ++pos;
    if (code < FSST_ESC) {
      dl += len[code];
    } else {
      ++pos;
      ++dl;
    }
  }
  if (pos < n) dl += (n - pos) * len[0];  // the zero bytes past the segment
  return dl;
}

// Block sum of one u64 per thread (every thread gets it).
template <class SM>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, SM& sm)
{
  const uint64_t w = readlane_u64(wave_incl_sum_u64(v), WAVE - 1);
  if (lane_id() == 0) {
    sm.part[2u * (threadIdx.x / WAVE)]      = static_cast<uint32_t>(w);
    sm.part[2u * (threadIdx.x / WAVE) + 1u] = static_cast<uint32_t>(w >> 32);
  }
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (uint32_t k = 0; k < FSST_WAVES; ++k)
    t += static_cast<uint64_t>(sm.part[2u * k]) | (static_cast<uint64_t>(sm.part[2u * k + 1u]) << 32);
  __syncthreads();
  return t;
}

// Chars of a plain chunk from dec[] (row starts) and comp[] (each row's
// first byte in the segment), one row per thread, in 8 KiB LDS pieces
// flushed as in fsst_emit_rows.
template <class SM>
__device__ __forceinline__ void plain_emit_rows(SM& sm, const uint8_t* seg, const uint8_t* end,
                                                         uint32_t rows, uint32_t T, uint64_t base, uint8_t* chars,
                                                         uint64_t chars_cap)
{
  const uint32_t tid = threadIdx.x;
  const uintptr_t cb = reinterpret_cast<uintptr_t>(chars);
  const uintptr_t g0 = cb + base;
  const uintptr_t g1 = cb + ((base + T < chars_cap) ? base + T : chars_cap);
  for (uintptr_t pa = g0 & ~uintptr_t{15}; pa < g1; pa += FSST_PIECE) {
    const uintptr_t pb = pa + FSST_PIECE;
#pragma unroll 1
    for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
      const uint32_t r = k * FSST_THREADS + tid;
      if (r >= rows) continue;
      const uint32_t d0 = sm.dec[r], dl = sm.dec[r + 1] - d0;
      const uintptr_t o0 = g0 + d0;
      if (dl == 0 || o0 + dl <= pa || o0 >= pb) continue;
      const uint32_t j0 = o0 < pa ? static_cast<uint32_t>(pa - o0) : 0u;
      const uint32_t j1 = o0 + dl > pb ? static_cast<uint32_t>(pb - o0) : dl;
      const uint8_t* src = seg + sm.comp[r];
      uint8_t* dst       = sm.out + (o0 - pa);
      if (src + dl <= end) {
        ByteReader br(src + j0);
        for (uint32_t j = j0; j < j1; ++j) dst[j] = static_cast<uint8_t>(br.next());
      } else {
        for (uint32_t j = j0; j < j1; ++j) dst[j] = static_cast<uint8_t>(byte_at(src + j, end));
      }
    }
    __syncthreads();
    const uintptr_t lo = pa > g0 ? pa : g0;
    const uintptr_t hi = pb < g1 ? pb : g1;
    for (uintptr_t blk = pa + uintptr_t{tid} * 16u; blk < hi; blk += uintptr_t{FSST_THREADS} * 16u) {
      if (blk >= lo && blk + 16u <= hi) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(&sm.out[blk - pa]);
        __builtin_nontemporal_store(v, gbl(reinterpret_cast<u32x4*>(blk)));
      } else {
        for (uintptr_t x = blk; x < blk + 16u; ++x)
          if (x >= lo && x < hi) *gbl(reinterpret_cast<uint8_t*>(x)) = sm.out[x - pa];
      }
    }
    __syncthreads();
  }
}

// Decode chunk c of a valid DICTIONARY or UNCOMPRESSED segment: per-row
// lengths (dictionary.cu:84-108, uncompressed.cu:56-81), the chunk's
// look-back for its chars offset, offsets, then the bytes (dictionary.cu:
// 117-140, uncompressed.cu:88-120).  Every thread must call it.
// A leaf call with scalar arguments: descriptors passed by value, or a
// call from here, would need a stack in scratch memory.
template <class SM>
__device__ __attribute__((noinline)) void plain_chunk(const uint8_t* seg, uint32_t bytes_size, uint32_t dict_end,
                                                     uint32_t flags, uint32_t ibo, uint32_t ibc, uint32_t row_begin,
                                                     uint32_t rows, uint32_t col_row, uint32_t c, uint32_t n_chunks,
                                                     FsstState* __restrict__ st, int32_t* __restrict__ offsets,
                                                     uint8_t* __restrict__ chars, uint64_t chars_cap,
                                                     uint32_t total_rows, uint32_t spin_limit, SM& sm)
{
  const uint32_t tid  = threadIdx.x;
  const uint8_t* end  = seg + bytes_size;
  const bool dict     = flags & FI_DICT;
  const bool dfsst    = flags & FI_DFSST;
  const uint32_t w    = flags & FI_WIDTH;
  const FsstChunk ch{0u, row_begin, rows, col_row};
  DictFsstSeg dd{};
  if (dfsst) {  // seg is the segment's descriptor; rows copy from its entries
    dd  = gload(reinterpret_cast<const DictFsstSeg*>(seg));
    seg = dd.src;
    end = dd.src + dd.src_size;
  }
  __syncthreads();  // the previous user of sm is done
  uint32_t T   = 0;
  uint64_t own = 0;
#pragma unroll 1
  for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
    const uint32_t r = k * FSST_THREADS + tid;
    uint32_t len = 0, src = 0;
    if (r < rows) {
      const uint32_t sr = ch.row_begin + r;
      uint32_t e0 = 0, e1 = 0;
      if (dfsst) {
        const uint32_t idx = dfsst_index(dd, sr);
        if (idx != 0u && idx < dd.dict_count) {
          e0 = gbl(dd.e)[idx];
          e1 = gbl(dd.e)[idx + 1u];
        }
        dict_end = e1 >= e0 ? e1 : 0u;  // the entries are front to back: a row starts at e0
      } else if (dict) {
        const uint32_t sel = stream_bits32(seg + 20, end, uint64_t{sr} * w, w);
        if (sel != 0u && sel < ibc) {
          e1 = ld_u32_any(seg + ibo + 4ull * sel);
          e0 = ld_u32_any(seg + ibo + 4ull * (sel - 1u));
        }
      } else {
        const int32_t v1 = static_cast<int32_t>(ld_u32_any(seg + 8 + 4ull * sr));
        const int32_t v0 = sr ? static_cast<int32_t>(ld_u32_any(seg + 8 + 4ull * (sr - 1u))) : 0;
        e1 = v1 < 0 ? 0u - static_cast<uint32_t>(v1) : static_cast<uint32_t>(v1);
        e0 = v0 < 0 ? 0u - static_cast<uint32_t>(v0) : static_cast<uint32_t>(v0);
      }
      if (e1 <= dict_end && e1 >= e0) {
        len = e1 - e0;
        src = dfsst ? e0 : dict_end - e1;
      }
      sm.comp[r] = src;
    }
    own += len;
    uint32_t tot      = 0;
    const uint32_t ex = block_excl_scan(len, sm.wave, &tot);
    if (r < rows) sm.dec[r] = T + ex;
    T += tot;
  }
  if (tid == 0) sm.dec[rows] = T;
  // Rows are bounded by the segment, but a chunk of long rows can pass
  // 2^31 chars: publish a saturated total so the column reports overflow.
  const uint64_t T64 = block_sum_u64(own, sm);
  const uint64_t Tp  = T64 > 0x7FFFFFFFull ? 0x80000000ull : T64;
  const uint64_t b   = fsst_lookback(c, n_chunks, Tp, st, spin_limit, sm);
  fsst_offsets(ch, c, n_chunks, rows, static_cast<uint32_t>(Tp), b, offsets, total_rows, sm);
  if (T64 == 0 || Tp != T64 || b >= chars_cap) return;
  plain_emit_rows(sm, seg, end, rows, T, b, chars, chars_cap);
}

// Import the symbol table serialised at seg + symoff (at most
// FSST_SYMSTAGE bytes inside the segment are read; the rest read as zero)
// into `tab`, one code per thread (detail/fsst.cuh:50-96); returns the
// length of the calling thread's code.  A table whose version word does not
// match decodes every code to nothing.  Every thread of the block must call
// it.
__device__ __forceinline__ uint32_t import_symtab(const uint8_t* seg, uint32_t bytes_size, uint32_t symoff, bool ok,
                                                  FsstSymTab* __restrict__ tab, FsstPrepSmem& sm)
{
  const uint32_t tid = threadIdx.x;
  __syncthreads();  // the previous segment's use of sm is done
  uint32_t sym_n = 0, sym_h = 0;
  if (ok) {
    sym_n = bytes_size - symoff;
    if (sym_n > FSST_SYMSTAGE) sym_n = FSST_SYMSTAGE;
    sym_h = block_stage<FSST_THREADS>(reinterpret_cast<uint32_t*>(sm.st), seg + symoff, sym_n);
  }
  __syncthreads();
  uint32_t len = 0;
  uint64_t sym = 0;
  if (ok) {
    auto sb = [&](uint32_t i) -> uint32_t { return i < sym_n ? sm.st[sym_h + i] : 0u; };
    uint64_t version = 0;
    for (uint32_t i = 0; i < 8; ++i) version |= static_cast<uint64_t>(sb(i)) << (8u * i);
    const uint32_t code = tid;
    if ((version >> 32) == FSST_VERSION && code < FSST_NSYM) {
      const uint32_t zt = sb(8) & 1u;
      uint32_t next_code = zt, pos = 17, found = 0;
      len = 8;
      sym = FSST_CORRUPT;
      if (zt && code == 0) {
        len   = 1;
        sym   = 0;
        found = 1;
      }
      // groups in code order: lengths 2,3,...,8 then 1
      for (uint32_t l = 1; l <= 8 && !found; ++l) {
        const uint32_t h  = l & 7u;
        const uint32_t sl = h + 1u;
        uint32_t cnt      = sb(9 + h);
        if (h == 0 && zt) cnt = (cnt - 1u) & 0xFFu;
        if (code >= next_code && code < next_code + cnt) {
          const uint32_t at = pos + (code - next_code) * sl;
          sym               = 0;
          for (uint32_t j = 0; j < sl; ++j) sym |= static_cast<uint64_t>(sb(at + j)) << (8u * j);
          len   = sl;
          found = 1;
        }
        next_code += cnt;
        pos += cnt * sl;
      }
    }
  }
  gbl(tab->sym)[tid] = sym;
  gbl(tab->len)[tid] = static_cast<uint8_t>(len);
  return len;
}

// Dictionary pre-pass of one DICT_FSST segment with the calling block: the
// entry byte offsets e_byte[0..dict_count] (scanned bit-packed lengths), for
// the FSST modes with more than one entry the symbol table and the decoded
// entry offsets (entry 0 empty), otherwise the byte offsets again, into
// e_dec; out = {decoded total, inline-NULL flag}.  32-bit sums, as the
// reference's scans.
__device__ __forceinline__ void dfsst_prep_segment(const DictFsstSeg& d, uint32_t* __restrict__ e_byte,
                                                   uint32_t* __restrict__ e_dec, FsstSymTab* __restrict__ tab,
                                                   uint32_t* __restrict__ out, FsstPrepSmem& sm)
{
  const uint32_t tid = threadIdx.x;
  if (!d.valid) {  // uniform per block
    if (tid == 0) out[0] = out[1] = 0;
    return;
  }
  const uint8_t* seg  = d.bytes;
  const uint8_t* end  = seg + d.bytes_size;
  const bool fsst     = d.mode != 0u && d.dict_count > 1u;
  const uint32_t mylen = import_symtab(seg, d.bytes_size, d.off_symtab, fsst && d.off_symtab < d.bytes_size, tab, sm);
  sm.len[tid] = static_cast<uint8_t>(mylen);
  __syncthreads();
  const uint8_t* slens = seg + d.off_slens;
  const uint8_t* dict  = seg + d.off_dict;
  uint32_t carry_b = 0, carry_d = 0, first_len = 0;
  for (uint32_t k0 = 0; k0 < d.dict_count; k0 += 4u * FSST_THREADS) {
    uint32_t v[4], dl[4], sb = 0, sd = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t k = k0 + 4u * tid + i;
      v[i] = k < d.dict_count ? stream_bits32(slens, end, uint64_t{k} * d.slw, d.slw) : 0u;
      sb += v[i];
    }
    uint32_t tb = 0;
    uint32_t run = carry_b + block_excl_scan(sb, sm.wave, &tb);
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t k = k0 + 4u * tid + i;
      if (k < d.dict_count) e_byte[d.e_byte + k] = run;
      dl[i] = !fsst ? v[i] : (k == 0u || k >= d.dict_count) ? 0u : dfsst_entry_len(dict + run, v[i], end, sm.len);
      if (k == 0u) first_len = v[i];
      run += v[i];
      sd += dl[i];
    }
    uint32_t td = 0;
    uint32_t rd = carry_d + block_excl_scan(sd, sm.wave, &td);
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t k = k0 + 4u * tid + i;
      if (k < d.dict_count) e_dec[d.e_byte + k] = rd;
      rd += dl[i];
    }
    carry_b += tb;
    carry_d += td;
  }
  if (tid == 0) {
    e_byte[d.e_byte + d.dict_count] = carry_b;
    e_dec[d.e_byte + d.dict_count]  = carry_d;
    out[0] = carry_d;
    out[1] = d.mode != 2u && d.dict_count > 1u && first_len == 0u;
  }
}

// Predecode the FSST entries k >= 1 of one segment into d.src at d.e[k]
// (exactly d.e[k + 1] - d.e[k] bytes; the byte of a trailing escape and
// bytes past the decoded ones are 0), entries strided over `stride`
// threads starting at `t`.
__device__ __forceinline__ void dfsst_predecode(const DictFsstSeg& d, const uint32_t* __restrict__ e_byte,
                                                const uint64_t* sym, const uint8_t* len, uint32_t t, uint32_t stride)
{
  const uint8_t* end  = d.bytes + d.bytes_size;
  const uint8_t* dict = d.bytes + d.off_dict;
  uint8_t* out        = const_cast<uint8_t*>(d.src);
  for (uint32_t k = 1u + t; k < d.dict_count; k += stride) {
    const uint32_t c0 = gbl(e_byte)[d.e_byte + k], c1 = gbl(e_byte)[d.e_byte + k + 1u];
    const uint32_t n  = c1 >= c0 ? c1 - c0 : 0u;
    uint64_t o        = gbl(d.e)[k];
    uint64_t lim      = gbl(d.e)[k + 1u];
    if (lim < o) lim = o;
    if (lim > d.src_size) lim = d.src_size;
    const uint8_t* p  = dict + c0;
    uint32_t pos = 0;
    while (pos < n && o < lim) {
      const uint32_t code = p + pos < end ? *gbl(p + pos) : 0u;
      ++pos;
      if (code < FSST_ESC) {
        const uint64_t v = sym[code];
        for (uint32_t j = 0; j < len[code] && o < lim; ++j) *gbl(out + o++) = static_cast<uint8_t>(v >> (8u * j));
      } else if (pos < n) {
        *gbl(out + o++) = p + pos < end ? *gbl(p + pos) : 0u;
        ++pos;
      }
    }
    while (o < lim) *gbl(out + o++) = 0u;
  }
}

// Range `si` with the calling 256-thread workgroup: its chunk descriptors
// into chunks[first_chunk + k]; for a segment, header checks
// (fsst.cu:79-87), the symbol table import, one code per thread
// (detail/fsst.cuh:50-96), into `tab`, and for each chunk k its
// compressed-length prefix bounds (fsst.cu:132-188) and flags into
// info[first_chunk + k].  Bytes past the segment read as zero; no symbol
// table parse reads past byte 2057.
__device__ __forceinline__ void fsst_prep_segment(const FsstSeg sg, uint32_t si, FsstChunk* __restrict__ chunks,
                                                  FsstChunkInfo* __restrict__ info, FsstSymTab* __restrict__ tab,
                                                  FsstPrepSmem& sm)
{
  const uint32_t tid = threadIdx.x;
  const uint8_t* seg = sg.bytes;
  const uint8_t* end = seg + sg.bytes_size;
  const bool gap     = seg == nullptr;
  for (uint32_t k = tid; k * FSST_CHUNK_ROWS < sg.row_count; k += FSST_THREADS) {
    const uint32_t r0 = k * FSST_CHUNK_ROWS;
    const uint32_t n  = sg.row_count - r0 < FSST_CHUNK_ROWS ? sg.row_count - r0 : FSST_CHUNK_ROWS;
    const FsstChunk fc{gap ? FSST_GAP : si, gap ? 0u : r0, n, sg.col_row + r0};
    *gbl(reinterpret_cast<u32x4*>(chunks + sg.first_chunk + k)) = __builtin_bit_cast(u32x4, fc);
  }
  if (gap) return;  // uniform: nothing else of a gap is read
  if (sg.codec == SDR_CODEC_DICTIONARY || sg.codec == SDR_CODEC_UNCOMPRESSED) {
    plain_prep_segment(sg, info);
    return;  // uniform per segment
  }
  if (sg.codec == SDR_CODEC_DICT_FSST) {  // bytes: the range's DictFsstSeg; pad: 1 if valid
    for (uint32_t k = tid; k * FSST_CHUNK_ROWS < sg.row_count; k += FSST_THREADS) {
      FsstChunkInfo ci{sg.bytes, 0u, 0u, 0u, FI_DFSST | (sg.pad ? FI_OK : 0u), 0u, 0u};
      const gptr<u32x4> dst = gbl(reinterpret_cast<u32x4*>(info + sg.first_chunk + k));
      const u32x8 cv        = __builtin_bit_cast(u32x8, ci);
      dst[0]                = cv.lo;
      dst[1]                = cv.hi;
    }
    return;  // uniform per segment
  }
  bool ok = false;
  uint32_t dict_end = 0, width = 0, symoff = 0;
  if (sg.bytes_size >= 16u) {
    dict_end = su32(seg + 4);
    width    = su32(seg + 8);
    symoff   = su32(seg + 12);
    ok       = dict_end <= sg.bytes_size && symoff < dict_end && width <= 32u;
  }
  import_symtab(seg, sg.bytes_size, symoff, ok, tab, sm);

  // chunk bounds: wave w sums chunks w, w + 4, ..., 16 fields per lane
  const uint32_t rows = sg.row_count;
  const uint32_t nch  = (rows + FSST_CHUNK_ROWS - 1u) / FSST_CHUNK_ROWS;
  const uint32_t wv   = tid / WAVE;
  const uint32_t lane = lane_id();
  uint32_t carry      = 0;
  for (uint32_t k0 = 0; k0 < nch; k0 += FSST_THREADS) {
    const uint32_t cnt = nch - k0 < FSST_THREADS ? nch - k0 : FSST_THREADS;
    if (ok)
      for (uint32_t k = k0 + wv; k < k0 + cnt; k += FSST_WAVES) {
        const uint32_t r0 = k * FSST_CHUNK_ROWS;
        const uint32_t r1 = rows - r0 < FSST_CHUNK_ROWS ? rows : r0 + FSST_CHUNK_ROWS;
        const uint32_t f0 = r0 + 16u * lane < r1 ? r0 + 16u * lane : r1;
        const uint32_t f1 = f0 + 16u < r1 ? f0 + 16u : r1;
        bool hg           = false;
        const uint32_t s  = fields_sum(seg + 16, end, f0, f1, width, &hg);
        const uint32_t tt = __builtin_amdgcn_readlane(wave_incl_sum_u32(s), WAVE - 1);
        const bool any_hg = ballot(hg) != 0;
        if (lane == 0) {
          sm.sum[k - k0]  = tt;
          sm.huge[k - k0] = any_hg;
        }
      }
    __syncthreads();
    const uint32_t v  = (ok && tid < cnt) ? sm.sum[tid] : 0u;
    uint32_t T        = 0;
    const uint32_t ex = block_excl_scan(v, sm.wave, &T);
    if (tid < cnt) {
      const uint32_t k  = k0 + tid;
      const uint32_t r1 = rows - k * FSST_CHUNK_ROWS < FSST_CHUNK_ROWS ? rows : (k + 1u) * FSST_CHUNK_ROWS;
      uint32_t flags    = 0;
      if (ok) {
        flags = width | FI_OK;
        if (seg + 16 + ((uint64_t{r1} * width + 7u) >> 3) <= end) flags |= FI_LFAST;
        if (sm.huge[tid]) flags |= FI_HUGE;
      }
      FsstChunkInfo ci;
      ci.bytes      = seg;
      ci.c_lo       = carry + ex;
      ci.c_hi       = carry + ex + v;
      ci.dict_end   = dict_end;
      ci.flags      = flags;
      ci.bytes_size = sg.bytes_size;
      ci.pad        = 0;
      const gptr<u32x4> dst = gbl(reinterpret_cast<u32x4*>(info + sg.first_chunk + k));
      const u32x8 cv        = __builtin_bit_cast(u32x8, ci);
      dst[0]                = cv.lo;
      dst[1]                = cv.hi;
    }
    carry += T;
    __syncthreads();  // sum[] and huge[] are read before the next batch refills them
  }
}

// ---- chunk decode ----

// Decode chunk `c` (its position in the column's chunk order) with the
// calling 256-thread workgroup.  Every thread must call it, after a
// workgroup barrier that follows any earlier use of `sm` (the callers' claim
// of the chunk ends with one), so the first loads issue without waiting.
// `ci` and `tab` are the segment pre-pass's output for the chunk and its
// segment (unused for a chunk of uncovered rows).
//
// One memory round trip stages, with LDS-direct loads, the symbol table,
// the chunk's compressed-length fields and its compressed bytes; a second
// is the look-back for the column-wide chars offset.
template <uint32_t OUT_CAP>
__device__ __forceinline__ void fsst_chunk(const FsstChunk ch, const FsstChunkInfo ci, const FsstSymTab* tab,
                                           uint32_t c, uint32_t n_chunks, FsstState* __restrict__ st,
                                           int32_t* __restrict__ offsets, uint8_t* __restrict__ chars,
                                           uint64_t chars_cap, uint32_t total_rows, uint32_t spin_limit,
                                           FsstSmemT<OUT_CAP>& sm)
{
  const uint32_t tid      = threadIdx.x;
  const bool ok           = ch.seg != FSST_GAP && (ci.flags & FI_OK);
  if (ok && (ci.flags & (FI_DICT | FI_UNC | FI_DFSST))) {  // uniform per chunk
    plain_chunk(ci.bytes, ci.bytes_size, ci.dict_end, ci.flags, ci.c_lo, ci.c_hi, ch.row_begin, ch.rows, ch.col_row,
                c, n_chunks, st, offsets, chars, chars_cap, total_rows, spin_limit, sm);
    return;
  }
  const uint8_t* seg      = ok ? ci.bytes : nullptr;
  const uint8_t* end      = ok ? ci.bytes + ci.bytes_size : nullptr;
  const uint32_t dict_end = ok ? ci.dict_end : 0u;
  const uint32_t width    = ok ? (ci.flags & FI_WIDTH) : 0u;
  const uint32_t rows     = ch.rows;
  const uint32_t c_lo     = ok ? ci.c_lo : 0u;
  const uint32_t c_hi     = ok ? ci.c_hi : 0u;
  const bool lfast        = ok && width && (ci.flags & FI_LFAST);
  // Rows shorter than 1 MiB keep the prefix sums from wrapping, so comp[]
  // rises from c_lo to c_hi.
  const bool staged = ok && rows && !(ci.flags & FI_HUGE) && c_hi >= c_lo && c_hi <= dict_end &&
                      c_hi - c_lo <= FSST_CSTAGE;

  if (ok) block_issue<FSST_THREADS>(reinterpret_cast<uint32_t*>(sm.sym), reinterpret_cast<const uint8_t*>(tab),
                                    sizeof(FsstSymTab));
  uint32_t bit0 = 0;
  if (lfast) {
    const uint64_t b0 = (uint64_t{ch.row_begin} * width) >> 3;
    const uint64_t b1 = (uint64_t{ch.row_begin + rows} * width + 7u) >> 3;
    const uint32_t h  = block_issue<FSST_THREADS>(reinterpret_cast<uint32_t*>(sm.out), seg + 16 + b0,
                                                 static_cast<uint32_t>(b1 - b0));
    bit0 = h * 8u + static_cast<uint32_t>((uint64_t{ch.row_begin} * width) & 7u);
  }
  // Rows are stored back to front ending at dict_end: row r is
  // [dict_end - comp[r+1], dict_end - comp[r]).  Stage position 0 is the
  // 16-byte-aligned address at or below the chunk's first compressed byte;
  // each thread loads its first-walk span into registers.
  uint32_t c_h         = 0;
  const uint8_t* stage = nullptr;
  if (staged) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(seg + dict_end - c_hi);
    c_h               = static_cast<uint32_t>(a & 15u);
    stage             = reinterpret_cast<const uint8_t*>(a - c_h);
  }
  // The span loads go out last and unconditionally (a chunk with no
  // compressed bytes reads a zero group), so the wait below covers the
  // symbol table and length fields only: the compressed prefix runs while
  // the spans still arrive.
  const bool spans = staged && c_h + (c_hi - c_lo) != 0;
  u32x4 qv[SPAN_GROUPS];
  fsst_span_load(qv, spans ? stage : reinterpret_cast<const uint8_t*>(fsst_zero_group),
                 spans ? c_h + (c_hi - c_lo) : 16u);
  static_assert(SPAN_GROUPS == 5, "the wait below leaves the five span loads in flight");
  __asm__ volatile("s_waitcnt vmcnt(5) lgkmcnt(0)\n\ts_barrier" ::: "memory");

  // ---- compressed prefix comp[] (fsst.cu:132-188); thread t owns rows
  // 4t .. 4t + 3 here ----
  {
    const uint32_t* lw = reinterpret_cast<const uint32_t*>(sm.out);
    uint32_t cl[ROWS_PER_THR];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < ROWS_PER_THR; ++i) {
      const uint32_t r = ROWS_PER_THR * tid + i;
      uint32_t v       = 0;
      if (width && r < rows)
        v = lfast ? lds_bits32(lw, bit0 + r * width, width)
                  : stream_bits32(seg + 16, end, uint64_t{ch.row_begin + r} * width, width);
      cl[i] = v;
      sum += v;
    }
    uint32_t tot     = 0;
    uint32_t run     = c_lo + block_excl_scan<true>(sum, sm.wave, &tot);  // the spans still arrive
    if (tid == 0) sm.comp[0] = c_lo;
#pragma unroll
    for (uint32_t i = 0; i < ROWS_PER_THR; ++i) {
      const uint32_t r = ROWS_PER_THR * tid + i;
      run += cl[i];
      if (r < rows) sm.comp[r + 1] = run;
    }
  }
  if (staged) sm.part[tid] = fsst_span_tail(qv, c_h + (c_hi - c_lo));
  __syncthreads();

  if (staged) {
    const uint32_t C = c_hi - c_lo;
    uint32_t T       = 0;
    if (C) {
      T = fsst_lengths_staged(sm, rows, c_h, c_hi, C, qv, stage);
    } else {
      for (uint32_t i = tid; i <= rows; i += FSST_THREADS) sm.dec[i] = 0;
      __syncthreads();
    }
    // The aggregate is published now; the look-back's wait, and the offsets,
    // come once the first window's chars are in LDS.
    fsst_lookback_publish(c, T, st, spin_limit);
    uint64_t base = 0;
    bool based    = false;
    auto finish   = [&]() -> uint64_t {
      if (!based) {
        base = fsst_lookback_wait(c, n_chunks, T, st, spin_limit, sm);
        fsst_offsets(ch, c, n_chunks, rows, T, base, offsets, total_rows, sm);
        based = true;
      } else {
        __syncthreads();
      }
      return base;
    };
    const bool emitted = T != 0 && fsst_emit_staged(sm, rows, c_h, c_hi, T, finish, chars, chars_cap, stage);
    if (!based) finish();
    if (T == 0 || emitted || base >= chars_cap) return;
    fsst_emit_rows(sm, seg, dict_end, rows, T, base, chars, chars_cap);
  } else {
    const uint32_t T    = fsst_lengths_rows(sm, ok, seg, dict_end, rows);
    const uint64_t base = fsst_lookback(c, n_chunks, T, st, spin_limit, sm);
    fsst_offsets(ch, c, n_chunks, rows, T, base, offsets, total_rows, sm);
    if (T == 0 || base >= chars_cap) return;
    fsst_emit_rows(sm, seg, dict_end, rows, T, base, chars, chars_cap);
  }
}

}  // namespace sdr
