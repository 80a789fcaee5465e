// sdr_device.hpp — gfx950 device primitives shared by the decode kernels.
//
// Wave64 throughout: a wavefront is 64 lanes; rows of 16 lanes are the DPP
// unit.  Segment bytes may start at any byte address, so every scalar header
// read goes through the aligned-word loaders below (they read only the
// 4-byte words that hold the requested bytes, which can never cross a page
// the requested bytes do not also touch).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sdr {

constexpr uint32_t WAVE = 64;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Orders LDS traffic between lanes of one wavefront.  LDS instructions of a
// wave execute in issue order; this keeps the compiler from moving accesses
// across the point and waits for outstanding LDS operations.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- global-memory pointers ------------------------------------------------
//
// Pointers read from descriptors (segment bytes, output columns, the FSST
// chars base) are generic to the compiler, which then emits flat loads and
// stores.  A flat access also counts on the LDS counter, so every later LDS
// wait would also wait for outstanding stores.  Accesses through gbl() are
// global instructions.

template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<T> gbl(T* p)
{
  return (gptr<T>)(p);
}

// A whole descriptor read through a global pointer (struct copies through an
// address-space-qualified lvalue are not allowed).
template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
  using W = std::conditional_t<alignof(T) >= 8, u32x2, uint32_t>;
  static_assert(sizeof(T) % sizeof(W) == 0, "descriptor size");
  struct alignas(alignof(T)) Words {
    W w[sizeof(T) / sizeof(W)];
  } v;
  const gptr<const W> src = gbl(reinterpret_cast<const W*>(p));
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / sizeof(W); ++i) v.w[i] = src[i];
  return __builtin_bit_cast(T, v);
}

// ---- unaligned loads -------------------------------------------------------

__device__ __forceinline__ uint32_t ld_u8(const uint8_t* p) { return *gbl(p); }

__device__ __forceinline__ uint32_t ld_u16_any(const uint8_t* p)
{
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  uint32_t o  = static_cast<uint32_t>(a & 3u);
  const gptr<const uint32_t> w = gbl(reinterpret_cast<const uint32_t*>(a - o));
  uint32_t lo = w[0];
  uint32_t v  = (o <= 2u) ? (lo >> (8u * o)) : __builtin_amdgcn_alignbit(w[1], lo, 24u);
  return v & 0xFFFFu;
}

__device__ __forceinline__ uint32_t ld_u32_any(const uint8_t* p)
{
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  uint32_t o  = static_cast<uint32_t>(a & 3u);
  const gptr<const uint32_t> w = gbl(reinterpret_cast<const uint32_t*>(a - o));
  uint32_t lo = w[0];
  if (o == 0) return lo;
  return __builtin_amdgcn_alignbit(w[1], lo, 8u * o);
}

__device__ __forceinline__ uint64_t ld_u64_any(const uint8_t* p)
{
  return static_cast<uint64_t>(ld_u32_any(p)) | (static_cast<uint64_t>(ld_u32_any(p + 4)) << 32);
}

// Load a T-sized little-endian value (1, 2, 4 or 8 bytes) zero-extended.
template <uint32_t TS>
__device__ __forceinline__ uint64_t ld_t_any(const uint8_t* p)
{
  if constexpr (TS == 1) return ld_u8(p);
  else if constexpr (TS == 2) return ld_u16_any(p);
  else if constexpr (TS == 4) return ld_u32_any(p);
  else return ld_u64_any(p);
}

// ---- wave-uniform loads of segment headers ---------------------------------
//
// Header fields are the same for every lane, and segment bytes are never
// written by a decode kernel, so they are read as constant-address-space
// words (scalar loads into SGPRs) at an address made wave-uniform with
// readfirstlane.  Same aligned-word discipline as above.

typedef __attribute__((address_space(4))) const uint32_t const_word;

__device__ __forceinline__ uint64_t uniform_addr(const uint8_t* p)
{
  const uint64_t a  = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t su_word(uint64_t aligned4) { return *(const_word*)(aligned4); }

// A whole wave-uniform descriptor read the same way, for descriptors that no
// kernel of the launch writes.  A scalar load's wait is placed at the first
// use of the value (or the next wait on the LDS/scalar counter), so loads
// issued after it overlap it.
template <typename T>
__device__ __forceinline__ T sload(const T* p)
{
  static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "whole aligned words");
  struct alignas(alignof(T)) Words {
    uint32_t w[sizeof(T) / 4];
  } x;
  const uint64_t a = uniform_addr(reinterpret_cast<const uint8_t*>(p));
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; ++i) x.w[i] = su_word(a + 4u * i);
  return __builtin_bit_cast(T, x);
}

__device__ __forceinline__ uint32_t su8(const uint8_t* p)
{
  const uint64_t a = uniform_addr(p);
  const uint32_t o = static_cast<uint32_t>(a & 3u);
  return (su_word(a - o) >> (8u * o)) & 0xFFu;
}

__device__ __forceinline__ uint32_t su16(const uint8_t* p)
{
  const uint64_t a = uniform_addr(p);
  const uint32_t o = static_cast<uint32_t>(a & 3u);
  const uint32_t lo = su_word(a - o);
  const uint32_t v  = (o <= 2u) ? (lo >> (8u * o)) : __builtin_amdgcn_alignbit(su_word(a - o + 4u), lo, 24u);
  return v & 0xFFFFu;
}

__device__ __forceinline__ uint32_t su32(const uint8_t* p)
{
  const uint64_t a = uniform_addr(p);
  const uint32_t o = static_cast<uint32_t>(a & 3u);
  const uint32_t lo = su_word(a - o);
  if (o == 0) return lo;
  return __builtin_amdgcn_alignbit(su_word(a - o + 4u), lo, 8u * o);
}

__device__ __forceinline__ uint64_t su64(const uint8_t* p)
{
  return static_cast<uint64_t>(su32(p)) | (static_cast<uint64_t>(su32(p + 4)) << 32);
}

template <uint32_t TS>
__device__ __forceinline__ uint64_t su_t(const uint8_t* p)
{
  if constexpr (TS == 1) return su8(p);
  else if constexpr (TS == 2) return su16(p);
  else if constexpr (TS == 4) return su32(p);
  else return su64(p);
}

// ---- wave scans (DPP row shifts + cross-row readlane) ----------------------

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
  return static_cast<uint32_t>(
    __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}

// Inclusive prefix sum over the 64 lanes (wrapping).
__device__ __forceinline__ uint32_t wave_incl_sum_u32(uint32_t v)
{
  v += dpp_u32<0x111>(v);  // row_shr:1
  v += dpp_u32<0x112>(v);  // row_shr:2
  v += dpp_u32<0x114>(v);  // row_shr:4
  v += dpp_u32<0x118>(v);  // row_shr:8
  uint32_t r15 = __builtin_amdgcn_readlane(v, 15);
  uint32_t r31 = __builtin_amdgcn_readlane(v, 31);
  uint32_t r47 = __builtin_amdgcn_readlane(v, 47);
  uint32_t row = lane_id() >> 4;
  uint32_t add = (row >= 1 ? r15 : 0u) + (row >= 2 ? r31 : 0u) + (row >= 3 ? r47 : 0u);
  return v + add;
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v)
{
  uint32_t lo = dpp_u32<CTRL>(static_cast<uint32_t>(v));
  uint32_t hi = dpp_u32<CTRL>(static_cast<uint32_t>(v >> 32));
  return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l)
{
  uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), l);
  uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), l);
  return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

__device__ __forceinline__ uint64_t wave_incl_sum_u64(uint64_t v)
{
  v += dpp_u64<0x111>(v);
  v += dpp_u64<0x112>(v);
  v += dpp_u64<0x114>(v);
  v += dpp_u64<0x118>(v);
  uint64_t r15 = readlane_u64(v, 15);
  uint64_t r31 = readlane_u64(v, 31);
  uint64_t r47 = readlane_u64(v, 47);
  uint32_t row = lane_id() >> 4;
  uint64_t add = (row >= 1 ? r15 : 0ull) + (row >= 2 ? r31 : 0ull) + (row >= 3 ? r47 : 0ull);
  return v + add;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// ---- LDS bit extraction ----------------------------------------------------

// `w` bits (w <= 32) at bit `bitpos` of an LSB-first stream held in LDS words.
__device__ __forceinline__ uint32_t lds_bits32(const uint32_t* lds, uint32_t bitpos, uint32_t w)
{
  uint32_t word = bitpos >> 5;
  uint32_t v    = __builtin_amdgcn_alignbit(lds[word + 1], lds[word], bitpos & 31u);
  return w >= 32u ? v : (v & ((1u << w) - 1u));
}

// `w` bits (w <= 64) at bit `bitpos`.
__device__ __forceinline__ uint64_t lds_bits64(const uint32_t* lds, uint32_t bitpos, uint32_t w)
{
  uint32_t word = bitpos >> 5;
  uint32_t sh   = bitpos & 31u;
  uint32_t w0 = lds[word], w1 = lds[word + 1], w2 = lds[word + 2];
  uint64_t v = static_cast<uint64_t>(__builtin_amdgcn_alignbit(w1, w0, sh)) |
               (static_cast<uint64_t>(__builtin_amdgcn_alignbit(w2, w1, sh)) << 32);
  return w >= 64u ? v : (v & ((1ull << w) - 1ull));
}

// ---- global -> LDS staging -------------------------------------------------
//
// Staging uses gfx950's LDS-direct loads (global_load_lds_dwordx4): lane i of
// an instruction writes its 16 bytes at LDS address M0 + 16 i without
// passing through VGPRs, so every load of a stage is in flight at once and a
// stage costs one memory round trip.  Inactive lanes write nothing.

typedef __attribute__((address_space(3))) uint32_t lds_word;

// s_waitcnt vmcnt(0), expcnt and lgkmcnt left at their maxima (gfx9 encoding)
constexpr int WAITCNT_VMEM_DONE = 0x0F70;

__device__ __forceinline__ void lds_direct_16(const uint8_t* g, uint32_t* lds_dst)
{
  __builtin_amdgcn_global_load_lds(g, (lds_word*)(lds_dst), 16, 0, 0);
}

// Copy the bytes [p, p + nbytes) of global memory into LDS as whole 16-byte
// blocks, by one wavefront: LDS byte 0 holds the byte at align_down(p, 16).
// Returns the bit offset of p in the staged words.  Every 16-byte block read
// overlaps the requested range, so the copy never touches a page the range
// does not.  The data is in LDS when this returns (the wave has waited).
__device__ __forceinline__ uint32_t stage_to_lds(uint32_t* lds, const uint8_t* p, uint32_t nbytes)
{
  const uintptr_t a   = reinterpret_cast<uintptr_t>(p);
  const uintptr_t a16 = a & ~uintptr_t{15};
  const uint32_t head = static_cast<uint32_t>(a - a16);
  const uint32_t nblk = nbytes ? (head + nbytes + 15u) >> 4 : 0u;
  const uint8_t* src  = reinterpret_cast<const uint8_t*>(a16) + 16u * lane_id();
  for (uint32_t k = 0; k * WAVE < nblk; ++k)
    if (k * WAVE + lane_id() < nblk) lds_direct_16(src + 16u * WAVE * k, lds + 4u * WAVE * k);
  __builtin_amdgcn_s_waitcnt(WAITCNT_VMEM_DONE);
  return head * 8u;
}

// Block-wide copy of [p, p + nbytes) into LDS as 16-byte blocks by NT
// threads (LDS byte 0 holds align_down(p, 16)); returns the byte skew of p.
// Issues the loads only: several stages can be in flight together, and the
// caller waits (s_waitcnt vmcnt(0)) and then synchronises the block.
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_issue(uint32_t* lds, const uint8_t* p, uint32_t nbytes)
{
  const uintptr_t a   = reinterpret_cast<uintptr_t>(p);
  const uintptr_t a16 = a & ~uintptr_t{15};
  const uint32_t head = static_cast<uint32_t>(a - a16);
  const uint32_t nblk = nbytes ? (head + nbytes + 15u) >> 4 : 0u;
  // first block of this wave per round (wave-uniform: its LDS address is an SGPR for M0)
  const uint32_t wb   = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / WAVE))) * WAVE;
  const uint8_t* src  = reinterpret_cast<const uint8_t*>(a16) + 16u * threadIdx.x;
  for (uint32_t k = 0; k * NT < nblk; ++k)
    if (k * NT + threadIdx.x < nblk) lds_direct_16(src + 16u * NT * k, lds + 4u * (NT * k + wb));
  return head;
}

// v with every 32-bit word taken from the first active lane: a descriptor
// every lane loaded from the same address, made scalar (SGPRs instead of
// VGPRs).
template <typename T>
__device__ __forceinline__ T uniform(const T& v)
{
  static_assert(sizeof(T) % 4 == 0, "whole words");
  struct alignas(alignof(T)) Words {
    uint32_t w[sizeof(T) / 4];
  } x = __builtin_bit_cast(Words, v);
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; ++i)
    x.w[i] = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x.w[i])));
  return __builtin_bit_cast(T, x);
}

// block_issue, then this wave's wait; the caller's barrier makes the whole
// stage visible to the block.
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_stage(uint32_t* lds, const uint8_t* p, uint32_t nbytes)
{
  const uint32_t head = block_issue<NT>(lds, p, nbytes);
  __builtin_amdgcn_s_waitcnt(WAITCNT_VMEM_DONE);
  return head;
}

// ---- stores ----------------------------------------------------------------

template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v)
{
  __builtin_nontemporal_store(v, gbl(p));
}

// Store VPL = 16 / sizeof(T) consecutive values; one 16-byte store when the
// destination is 16-byte aligned and all values are valid.
template <typename T>
__device__ __forceinline__ void store_vec(T* p, const T* v, uint32_t n_valid, bool aligned16)
{
  constexpr uint32_t VPL = 16 / sizeof(T);
  if (aligned16 && n_valid >= VPL) {
    u32x4 x;
    __builtin_memcpy(&x, v, 16);
    __builtin_nontemporal_store(x, gbl(reinterpret_cast<u32x4*>(p)));
  } else {
#pragma unroll
    for (uint32_t i = 0; i < VPL; ++i)
      if (i < n_valid) st_nt(p + i, v[i]);
  }
}

}  // namespace sdr
