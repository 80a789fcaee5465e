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
// and then its chars.  The compressed bytes are read from HBM once: the
// second walk (emit) hits the L1/L2 lines the first walk (lengths) loaded.
#pragma once

#include "sdr_device.hpp"
#include "sdr_fsst.hpp"

namespace sdr {

constexpr uint32_t FSST_THREADS   = 256;
constexpr uint32_t FSST_WAVES     = FSST_THREADS / WAVE;
constexpr uint32_t ROWS_PER_THR   = FSST_CHUNK_ROWS / FSST_THREADS;
constexpr uint32_t FSST_PIECE     = 8192;  // LDS output staging bytes
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
  return *p;
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
    const uint32_t* wp = reinterpret_cast<const uint32_t*>(a - sh);
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
  const uint32_t* w;
  uint64_t acc;
  uint32_t nbits;
  __device__ __forceinline__ BitCursor(const uint8_t* base, uint64_t bitpos)
  {
    uintptr_t a = reinterpret_cast<uintptr_t>(base + (bitpos >> 3));
    const uint32_t skip = static_cast<uint32_t>((a & 3u) * 8u + (bitpos & 7u));
    w     = reinterpret_cast<const uint32_t*>(a & ~uintptr_t{3});
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

// Sequential byte reader over [p, p + n) using aligned 4-byte loads.
struct ByteReader {
  const uint32_t* w;
  uint32_t cur;
  uint32_t boff;
  // pointer arithmetic (not an integer round trip) keeps an LDS pointer's
  // address space visible to the compiler, so LDS rows read with ds_read
  __device__ __forceinline__ explicit ByteReader(const uint8_t* p)
  {
    boff = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(p) & 3u);
    w    = reinterpret_cast<const uint32_t*>(p - boff);
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

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t* total)
{
  const uint32_t incl = wave_incl_sum_u32(v);
  const uint32_t w    = threadIdx.x >> 6;
  if (lane_id() == 63) s_wave[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (uint32_t k = 0; k < FSST_WAVES; ++k) {
    const uint32_t x = s_wave[k];
    before += k < w ? x : 0u;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}


// Shared memory of one chunk decode.
constexpr uint32_t FSST_SYMSTAGE = 17u + 8u * FSST_NSYM;  // the largest symbol table (2057 bytes)
constexpr uint32_t FSST_CSTAGE   = 16384;                  // chunk compressed bytes staged in LDS

struct FsstSmem {
  uint64_t sym[256];  // symbol bytes, zero above its length
  uint8_t len[256];
  uint32_t comp[FSST_CHUNK_ROWS + 1];  // segment-relative inclusive compressed prefix
  uint32_t dec[FSST_CHUNK_ROWS + 1];   // chunk-relative exclusive decoded offset
  // emit pieces; before that, pieces of the compressed-length stream
  __attribute__((aligned(16))) uint8_t out[FSST_PIECE + 48];
  // the symbol table, then the chunk's compressed bytes
  __attribute__((aligned(16))) uint8_t cst[FSST_CSTAGE + 48];
  uint32_t wave[FSST_WAVES];
  uint64_t base;
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

// Walk the n compressed bytes of one row at p (LDS or global memory) a
// word at a time: emit(value, length) for every byte -- a symbol, the
// literal after an escape (length 1), or nothing (length 0) for an escape.
// The body has no data-dependent branch, so lanes of a wave stay together
// except for the row length.  Reads only the aligned words holding the row.
template <typename Emit>
__device__ __forceinline__ void fsst_row(const uint8_t* p, uint32_t n, const FsstSmem& sm, Emit emit)
{
  const uint32_t boff = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(p) & 3u);
  const uint32_t* w   = reinterpret_cast<const uint32_t*>(p - boff);
  const uint32_t end  = boff + n;
  bool esc            = false;
  for (uint32_t q = 0; q * 4u < end; ++q) {
    const uint32_t word = w[q];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const uint32_t idx = q * 4u + j;
      const uint32_t b   = (word >> (8u * j)) & 0xFFu;
      const bool lit     = esc;
      const uint32_t l   = lit ? 1u : sm.len[b];  // len[255] == 0: an escape emits nothing
      const uint64_t v   = lit ? uint64_t{b} : sm.sym[b];
      if (idx >= boff && idx < end) {
        emit(v, l);
        esc = !lit && b == FSST_ESC;
      }
    }
  }
}

// Decode chunk `c` (its position in the column's chunk order) with the
// calling 256-thread workgroup.  Every thread must call it.
//
// Global traffic is staged through LDS with LDS-direct loads, so each stage
// costs one memory round trip: the symbol table, the compressed-length
// stream up to the chunk's last row (in 8 KiB pieces), and the chunk's
// compressed bytes (when they fit in 16 KiB; rows outside the staged range
// read global memory).  The column-wide chars offset comes from a
// decoupled look-back done by one wavefront over 64 predecessors at a time.
__device__ __forceinline__ void fsst_chunk(const FsstChunk ch, uint32_t c, uint32_t n_chunks,
                                           const FsstSeg* __restrict__ segs, FsstState* __restrict__ st,
                                           int32_t* __restrict__ offsets, uint8_t* __restrict__ chars,
                                           uint64_t chars_cap, uint32_t total_rows, uint32_t spin_limit,
                                           FsstSmem& sm)
{
  const uint32_t tid = threadIdx.x;
  const bool gap     = ch.seg == FSST_GAP;
  const FsstSeg sg   = gap ? FsstSeg{nullptr, 0, 0} : segs[ch.seg];
  const uint8_t* seg = sg.bytes;
  const uint8_t* end = seg + sg.bytes_size;

  // ---- header (fsst.cu:79-87) ----
  bool ok = false;
  uint32_t dict_end = 0, width = 0, symoff = 0;
  if (!gap && sg.bytes_size >= 16u) {
    dict_end = su32(seg + 4);
    width    = su32(seg + 8);
    symoff   = su32(seg + 12);
    ok       = dict_end <= sg.bytes_size && symoff < dict_end && width <= 32u;
  }

  // ---- symbol table import, one code per thread (detail/fsst.cuh:50-96) ----
  // Bytes past the segment read as zero; no parse reads past byte 2057.
  __syncthreads();  // the previous user of sm is done
  uint32_t sym_n = 0, sym_h = 0;
  if (ok) {
    sym_n = sg.bytes_size - symoff;
    if (sym_n > FSST_SYMSTAGE) sym_n = FSST_SYMSTAGE;
    sym_h = block_stage<FSST_THREADS>(reinterpret_cast<uint32_t*>(sm.cst), seg + symoff, sym_n);
  }
  __syncthreads();
  if (ok) {
    auto sb = [&](uint32_t i) -> uint32_t { return i < sym_n ? sm.cst[sym_h + i] : 0u; };
    uint64_t version = 0;
    for (uint32_t i = 0; i < 8; ++i) version |= static_cast<uint64_t>(sb(i)) << (8u * i);
    const uint32_t code = tid;
    uint32_t len = 0;
    uint64_t sym = 0;
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
        uint32_t cnt = sb(9 + h);
        if (h == 0 && zt) cnt = (cnt - 1u) & 0xFFu;
        if (code >= next_code && code < next_code + cnt) {
          const uint32_t at = pos + (code - next_code) * sl;
          sym = 0;
          for (uint32_t j = 0; j < sl; ++j) sym |= static_cast<uint64_t>(sb(at + j)) << (8u * j);
          len   = sl;
          found = 1;
        }
        next_code += cnt;
        pos += cnt * sl;
      }
    }
    sm.len[tid] = static_cast<uint8_t>(len);
    sm.sym[tid] = sym;
  } else {
    sm.len[tid] = 0;
    sm.sym[tid] = 0;
  }

  // ---- compressed lengths (fsst.cu:132-188): the sum of the segment's
  // fields before the chunk, and this thread's own fields.  Thread t owns
  // rows t, t + 256, t + 512 and t + 768 of the chunk, so every wave has
  // rows in every stretch of the output the emit below walks. ----
  const uint32_t rows = ch.rows;
  auto row_of         = [&](uint32_t k) { return k * FSST_THREADS + tid; };
  uint32_t clen[ROWS_PER_THR];
#pragma unroll
  for (uint32_t k = 0; k < ROWS_PER_THR; ++k) clen[k] = 0;
  uint32_t part = 0;
  const uint32_t nf = ch.row_begin + rows;
  const bool lfast  = ok && width && seg + 16 + ((uint64_t{nf} * width + 7u) >> 3) <= end;
  if (lfast) {
    const uint32_t per = ((FSST_PIECE - 16u) * 8u) / width;  // fields per staged piece
    for (uint32_t f0 = 0; f0 < nf; f0 += per) {
      const uint32_t f1 = nf - f0 < per ? nf : f0 + per;
      const uint64_t b0 = (uint64_t{f0} * width) >> 3;
      const uint64_t b1 = (uint64_t{f1} * width + 7u) >> 3;
      __syncthreads();  // the previous piece is consumed
      const uint32_t h = block_stage<FSST_THREADS>(reinterpret_cast<uint32_t*>(sm.out), seg + 16 + b0,
                                                   static_cast<uint32_t>(b1 - b0));
      __syncthreads();
      const uint32_t bit0 = h * 8u + static_cast<uint32_t>((uint64_t{f0} * width) & 7u);
      const uint32_t* lw  = reinterpret_cast<const uint32_t*>(sm.out);
      const uint32_t pe   = f1 < ch.row_begin ? f1 : ch.row_begin;
      for (uint32_t f = f0 + tid; f < pe; f += FSST_THREADS) part += lds_bits32(lw, bit0 + (f - f0) * width, width);
#pragma unroll
      for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
        const uint32_t f = ch.row_begin + row_of(k);
        if (row_of(k) < rows && f >= f0 && f < f1) clen[k] = lds_bits32(lw, bit0 + (f - f0) * width, width);
      }
    }
  } else if (ok) {
    // a length stream running past the segment (malformed): bounded reads
    if (ch.row_begin) {
      const uint32_t q  = (ch.row_begin + FSST_THREADS - 1) / FSST_THREADS;
      const uint32_t i0 = tid * q < ch.row_begin ? tid * q : ch.row_begin;
      const uint32_t i1 = i0 + q < ch.row_begin ? i0 + q : ch.row_begin;
      part = sum_fields(seg + 16, end, i0, i1, width);
    }
#pragma unroll
    for (uint32_t k = 0; k < ROWS_PER_THR; ++k)
      if (row_of(k) < rows) clen[k] = stream_bits32(seg + 16, end, uint64_t{ch.row_begin + row_of(k)} * width, width);
  }
  uint32_t before = 0;
  block_excl_scan(part, sm.wave, &before);
  if (tid == 0) sm.comp[0] = before;
  {
    uint32_t run = before;  // rows before the current 256-row stride
#pragma unroll
    for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
      uint32_t tot    = 0;
      const uint32_t ex = block_excl_scan(clen[k], sm.wave, &tot);
      if (row_of(k) < rows) sm.comp[row_of(k) + 1] = run + ex + clen[k];
      run += tot;
    }
  }
  __syncthreads();

  // ---- stage the chunk's compressed bytes.  Rows are stored back to front
  // ending at dict_end: row r is [dict_end - comp[r+1], dict_end - comp[r]).
  const uint32_t c_lo = sm.comp[0], c_hi = sm.comp[rows];
  const bool staged   = ok && rows && c_hi >= c_lo && c_hi <= dict_end && c_hi - c_lo <= FSST_CSTAGE;
  uint32_t c_h = 0;
  if (staged) c_h = block_stage<FSST_THREADS>(reinterpret_cast<uint32_t*>(sm.cst), seg + dict_end - c_hi, c_hi - c_lo);
  __syncthreads();

  // ---- decoded lengths (fsst.cu:194-286, serial escape semantics) ----
  uint32_t dlen[ROWS_PER_THR];
#pragma unroll
  for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
    const uint32_t r = row_of(k);
    uint32_t d = 0;
    if (ok && r < rows) {
      const uint32_t cum = sm.comp[r + 1], prev = sm.comp[r];
      if (cum <= dict_end && prev <= cum && cum != prev) {
        const uint32_t n = cum - prev;
        auto add = [&](uint64_t, uint32_t l) { d += l; };
        if (staged && prev >= c_lo && cum <= c_hi) fsst_row(sm.cst + c_h + (c_hi - cum), n, sm, add);
        else fsst_row(seg + dict_end - cum, n, sm, add);
      }
    }
    dlen[k] = d;
  }
  uint32_t T = 0;
#pragma unroll
  for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
    uint32_t tot    = 0;
    const uint32_t ex = block_excl_scan(dlen[k], sm.wave, &tot);
    if (row_of(k) < rows) sm.dec[row_of(k)] = T + ex;
    T += tot;
  }
  if (tid == 0) sm.dec[rows] = T;

  // ---- decoupled look-back for the column-wide chars offset, by wave 0:
  // publish this chunk's aggregate, then sum predecessors 64 at a time back
  // to the nearest published inclusive prefix.  A status word carries its
  // flag and value together and is the only datum exchanged between chunks,
  // so device-scope relaxed atomics (coherent across the XCDs' L2s) suffice;
  // release/acquire orderings would add an L2 write-back / invalidate per
  // access. ----
  if (tid < WAVE) {
    const uint32_t lane = tid;
    uint64_t* status    = st->status;
    uint64_t excl64     = 0;
    if (c == 0) {
      if (lane == 0) __hip_atomic_store(&status[0], ST_INC | T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0) __hip_atomic_store(&status[c], ST_AGG | T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
            if (lane == 0) atomicOr(&st->error, 1u);
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
      if (c == n_chunks - 1) st->total = excl64 + T;
    }
  }
  __syncthreads();
  const uint64_t base = sm.base;

  // ---- offsets ----
  for (uint32_t i = tid; i < rows; i += FSST_THREADS)
    offsets[ch.col_row + i] = static_cast<int32_t>(static_cast<uint32_t>(base + sm.dec[i]));
  if (c == n_chunks - 1 && tid == 0) offsets[total_rows] = static_cast<int32_t>(static_cast<uint32_t>(base + T));
  if (T == 0 || base >= chars_cap) return;

  // ---- chars: decode into LDS pieces aligned to the absolute chars
  // address, flush whole 16-byte blocks with vector stores and the partial
  // first/last blocks byte-wise (they share 16 bytes with the neighbours) ----
  const uintptr_t cb = reinterpret_cast<uintptr_t>(chars);
  const uintptr_t g0 = cb + base;
  const uintptr_t g1 = cb + ((base + T < chars_cap) ? base + T : chars_cap);
  for (uintptr_t pa = g0 & ~uintptr_t{15}; pa < g1; pa += FSST_PIECE) {
    const uintptr_t pb = pa + FSST_PIECE;
#pragma unroll
    for (uint32_t k = 0; k < ROWS_PER_THR; ++k) {
      const uint32_t r = row_of(k);
      if (r >= rows || dlen[k] == 0) continue;
      const uintptr_t o0 = g0 + sm.dec[r];
      if (o0 + dlen[k] <= pa || o0 >= pb) continue;
      const uint32_t cum = sm.comp[r + 1], prev = sm.comp[r];
      // o: output position relative to the piece (may start before it)
      int32_t o = static_cast<int32_t>(static_cast<intptr_t>(o0 - pa));
      constexpr int32_t PIECE = static_cast<int32_t>(FSST_PIECE);
      const bool inside = o >= 0 && o + static_cast<int32_t>(dlen[k]) <= PIECE;
      auto walk = [&](auto&& put) {
        auto lit = [&](uint32_t b) { return put(b, 1u); };
        auto sym = [&](uint32_t b) { return put(sm.sym[b], sm.len[b]); };
        if (staged && prev >= c_lo && cum <= c_hi) {
          ByteReader lr(sm.cst + c_h + (c_hi - cum));
          fsst_walk(cum - prev, [&](uint32_t) -> uint32_t { return lr.next(); }, lit, sym);
        } else {
          ByteReader br(seg + dict_end - cum);
          fsst_walk(cum - prev, [&](uint32_t) -> uint32_t { return br.next(); }, lit, sym);
        }
      };
      if (inside) {
        // The row lies inside the piece: its first bytes up to a dword
        // boundary and its last < 4 bytes are written singly (neighbouring
        // rows share those dwords); everything between leaves a 16-byte
        // register window as aligned dword stores.  Fixed-trip loops only.
        uint64_t wa = 0, wb = 0;  // pending bytes, starting at dword-aligned oa
        uint32_t na = 0;
        int32_t oa  = o;
        auto put = [&](uint64_t v, uint32_t l) {
#pragma unroll
          for (uint32_t h = 0; h < 3; ++h) {
            if (na == 0 && (oa & 3) && l) {
              sm.out[oa++] = static_cast<uint8_t>(v);
              v >>= 8;
              --l;
            }
          }
          wa |= l ? v << (8u * na) : 0ull;
          wb |= (l && na) ? v >> (64u - 8u * na) : 0ull;
          na += l;
#pragma unroll
          for (uint32_t f = 0; f < 2; ++f) {
            if (na >= 4u) {
              *reinterpret_cast<uint32_t*>(&sm.out[oa]) = static_cast<uint32_t>(wa);
              wa = (wa >> 32) | (wb << 32);
              wb >>= 32;
              oa += 4;
              na -= 4u;
            }
          }
        };
        if (staged && prev >= c_lo && cum <= c_hi) fsst_row(sm.cst + c_h + (c_hi - cum), cum - prev, sm, put);
        else fsst_row(seg + dict_end - cum, cum - prev, sm, put);
#pragma unroll
        for (uint32_t j = 0; j < 3; ++j)
          if (j < na) sm.out[oa + static_cast<int32_t>(j)] = static_cast<uint8_t>(wa >> (8u * j));
      } else {
        // a row crossing a piece edge: bytes outside the piece are skipped
        auto put = [&](uint64_t v, uint32_t l) {
          for (uint32_t j = 0; j < l; ++j, ++o)
            if (o >= 0 && o < PIECE) sm.out[o] = static_cast<uint8_t>(v >> (8u * j));
          return o < PIECE;
        };
        walk(put);
      }
    }
    __syncthreads();
    const uintptr_t lo = pa > g0 ? pa : g0;
    const uintptr_t hi = pb < g1 ? pb : g1;
    for (uintptr_t blk = pa + uintptr_t{tid} * 16u; blk < hi; blk += uintptr_t{FSST_THREADS} * 16u) {
      if (blk >= lo && blk + 16u <= hi) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(&sm.out[blk - pa]);
        __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(blk));
      } else {
        for (uintptr_t x = blk; x < blk + 16u; ++x)
          if (x >= lo && x < hi) *reinterpret_cast<uint8_t*>(x) = sm.out[x - pa];
      }
    }
    __syncthreads();
  }
}

}  // namespace sdr
