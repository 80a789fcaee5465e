// sdr_arena.hpp — the decode server's device memory pool.
//
// Device memory is reserved in large slabs (hipMalloc, only when the pool
// grows) and carved by a host-side best-fit allocator that coalesces
// neighbouring free ranges.  Nothing here touches the device after a slab is
// reserved: allocation and release are host bookkeeping, so a warm server
// allocates and frees request outputs and resident segments without HIP
// calls.  Ranges that device work may still read or write are released
// behind a fence (an event) and rejoin the free lists once it has passed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <unordered_map>
#include <vector>

namespace sdr {

class DeviceArena {
 public:
  static constexpr uint64_t ALIGN = 256;

  DeviceArena() = default;
  DeviceArena(const DeviceArena&) = delete;
  DeviceArena& operator=(const DeviceArena&) = delete;
  ~DeviceArena()
  {
    for (auto& s : slabs_) (void)hipFree(s.first);
  }

  // Carve `bytes` (rounded up to ALIGN) from the free lists; nullptr when
  // no free range is large enough.
  uint8_t* take(uint64_t bytes)
  {
    bytes   = round_up(bytes ? bytes : 1);
    auto it = by_size_.lower_bound(bytes);
    if (it == by_size_.end()) return nullptr;
    const uint64_t size = it->first;
    uint8_t* p          = it->second;
    by_size_.erase(it);
    by_addr_.erase(p);
    if (size > bytes) insert_free(p + bytes, size - bytes);
    used_[p] = bytes;
    in_use_ += bytes;
    return p;
  }

  // Reserve a slab of at least `bytes` (and at least `min_slab`) and add it
  // to the free lists.
  hipError_t grow(uint64_t bytes, uint64_t min_slab)
  {
    uint64_t n = round_up(bytes > min_slab ? bytes : min_slab);
    void* p    = nullptr;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) return e;
    slabs_[static_cast<uint8_t*>(p)] = n;
    reserved_ += n;
    insert_free(static_cast<uint8_t*>(p), n);
    return hipSuccess;
  }

  // Size of an allocation made by take(), 0 for a pointer it did not return.
  uint64_t size_of(const uint8_t* p) const
  {
    auto it = used_.find(const_cast<uint8_t*>(p));
    return it == used_.end() ? 0 : it->second;
  }

  // Return an allocation to the free lists at once: no device work may
  // still use it.  False for a pointer take() did not return.
  bool put(uint8_t* p)
  {
    auto it = used_.find(p);
    if (it == used_.end()) return false;
    const uint64_t n = it->second;
    used_.erase(it);
    in_use_ -= n;
    // coalesce with the free neighbours inside the same slab
    uint8_t* lo = p;
    uint64_t sz = n;
    auto nx     = by_addr_.lower_bound(p);
    if (nx != by_addr_.end() && nx->first == p + n && same_slab(p, nx->first)) {
      sz += nx->second;
      erase_size(nx->first, nx->second);
      by_addr_.erase(nx);
    }
    auto pv = by_addr_.lower_bound(p);
    if (pv != by_addr_.begin()) {
      --pv;
      if (pv->first + pv->second == p && same_slab(pv->first, p)) {
        lo = pv->first;
        sz += pv->second;
        erase_size(pv->first, pv->second);
        by_addr_.erase(pv);
      }
    }
    insert_free(lo, sz);
    return true;
  }

  uint64_t reserved() const { return reserved_; }
  uint64_t in_use() const { return in_use_; }
  size_t n_slabs() const { return slabs_.size(); }

  static uint64_t round_up(uint64_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

 private:
  void insert_free(uint8_t* p, uint64_t n)
  {
    by_addr_[p] = n;
    by_size_.emplace(n, p);
  }
  void erase_size(uint8_t* p, uint64_t n)
  {
    auto r = by_size_.equal_range(n);
    for (auto it = r.first; it != r.second; ++it)
      if (it->second == p) {
        by_size_.erase(it);
        return;
      }
  }
  bool same_slab(uint8_t* a, uint8_t* b) const
  {
    auto it = slabs_.upper_bound(a);  // first slab above a
    if (it == slabs_.begin()) return false;
    --it;
    return a < it->first + it->second && b >= it->first && b < it->first + it->second;
  }

  std::map<uint8_t*, uint64_t> slabs_;             // reserved slabs by base
  std::map<uint8_t*, uint64_t> by_addr_;           // free ranges by start
  std::multimap<uint64_t, uint8_t*> by_size_;       // free ranges by size (best fit)
  std::unordered_map<uint8_t*, uint64_t> used_;    // allocations by start
  uint64_t reserved_ = 0;
  uint64_t in_use_   = 0;
};

}  // namespace sdr
