// frep_pool.hip -- caching sub-allocator for device and pinned host buffers.
//
// Tree buffers (node records, tape pools, counters) come and go with every
// new scene, regrowth and eviction.  hipMalloc costs tens of microseconds and
// hipFree synchronizes the device, which would stall every batch in flight;
// so blocks are carved from large slabs and recycled through per-size-class
// free lists instead.  A block is returned only after the work using it has
// completed (the callers synchronize first), so recycling needs no fences.
#include "frep_host.hpp"

#include <algorithm>

namespace frep {

namespace {

constexpr size_t kAlign = 256;
constexpr size_t kDeviceSlab = size_t(1) << 30;    // 1 GiB device slabs
constexpr size_t kPinnedSlab = size_t(16) << 20;   // 16 MiB pinned slabs
constexpr size_t kDirect = size_t(256) << 20;      // larger blocks bypass the pool

// Size classes: multiples of 256 bytes up to 1 KiB, then four classes per
// power of two (m * 2^e, m in 4..7), so a block wastes at most a quarter.
size_t class_size(size_t bytes, uint32_t* cls) {
    bytes = std::max(bytes, kAlign);
    if (bytes <= 1024) {
        const size_t s = (bytes + kAlign - 1) & ~(kAlign - 1);
        *cls = (uint32_t)(s / kAlign - 1);                   // 0..3
        return s;
    }
    uint32_t e = 63 - (uint32_t)__builtin_clzll(bytes) - 2;   // bytes in [4, 8) * 2^e
    size_t m = (bytes + (size_t(1) << e) - 1) >> e;
    if (m == 8) {
        m = 4;
        ++e;
    }
    *cls = 4 + (e - 8) * 4 + (uint32_t)(m - 4);              // e >= 8 since bytes > 1024
    return m << e;
}

}  // namespace

BlockPool::BlockPool(bool pinned) : pinned_(pinned) {}

BlockPool::~BlockPool() { release_all(); }

void BlockPool::release_all() {
    std::lock_guard<std::mutex> lk(mu_);
    for (Slab& s : slabs_) {
        if (pinned_) (void)hipHostFree(s.base);
        else (void)hipFree(s.base);
    }
    for (auto& kv : live_)
        if (kv.second.cls == kDirectClass) {
            if (pinned_) (void)hipHostFree(kv.first);
            else (void)hipFree(kv.first);
        }
    slabs_.clear();
    live_.clear();
    free_.clear();
    bytes_ = 0;
}

hipError_t BlockPool::alloc(size_t bytes, void** out) {
    *out = nullptr;
    std::lock_guard<std::mutex> lk(mu_);
    if (bytes > kDirect) {
        void* p = nullptr;
        const hipError_t e = pinned_ ? hipHostMalloc(&p, bytes, hipHostMallocDefault)
                                     : hipMalloc(&p, bytes);
        if (e != hipSuccess) return e;
        live_[p] = Live{kDirectClass, bytes};
        bytes_ += bytes;
        *out = p;
        return hipSuccess;
    }
    uint32_t cls = 0;
    const size_t sz = class_size(bytes, &cls);
    if (cls < free_.size() && !free_[cls].empty()) {
        *out = free_[cls].back();
        free_[cls].pop_back();
        live_[*out] = Live{cls, sz};
        return hipSuccess;
    }
    if (slabs_.empty() || slabs_.back().size - slabs_.back().used < sz) {
        const size_t want = std::max(sz, pinned_ ? kPinnedSlab : kDeviceSlab);
        void* p = nullptr;
        const hipError_t e = pinned_ ? hipHostMalloc(&p, want, hipHostMallocDefault)
                                     : hipMalloc(&p, want);
        if (e != hipSuccess) return e;
        slabs_.push_back(Slab{static_cast<char*>(p), want, 0});
        bytes_ += want;
    }
    Slab& s = slabs_.back();
    *out = s.base + s.used;
    s.used += sz;
    live_[*out] = Live{cls, sz};
    return hipSuccess;
}

void BlockPool::free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    const Live b = it->second;
    live_.erase(it);
    const uint32_t cls = b.cls;
    if (cls == kDirectClass) {
        bytes_ -= b.size;
        if (pinned_) (void)hipHostFree(p);
        else (void)hipFree(p);
        return;
    }
    if (free_.size() <= cls) free_.resize(cls + 1);
    free_[cls].push_back(p);
}

size_t BlockPool::reserved_bytes() {
    std::lock_guard<std::mutex> lk(mu_);
    return bytes_;
}

}  // namespace frep
