// Device-wide scans (three-phase: per-block local scan, scan of block sums,
// add back) and the rank directory of bitsets.  wave64 shuffles do the
// intra-wave part; one block covers 2048 items.
#include <atomic>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace its {

namespace {
constexpr int kMaxDevices = 64;
std::atomic<hipMemPool_t> g_pools[kMaxDevices];
std::mutex g_pool_mu;
}  // namespace

hipMemPool_t device_pool() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    hipMemPool_t pool = g_pools[dev].load(std::memory_order_acquire);
    if (pool) return pool;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    pool = g_pools[dev].load(std::memory_order_relaxed);
    if (pool) return pool;
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    if (hipMemPoolCreate(&pool, &props) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    g_pools[dev].store(pool, std::memory_order_release);
    return pool;
}

void device_pool_trim() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return;
    hipMemPool_t pool = g_pools[dev].load(std::memory_order_acquire);
    if (pool) (void)hipMemPoolTrimTo(pool, 0);
}

namespace {
struct BlockCache {
    std::mutex mu;
    std::unordered_map<void*, size_t> size_of;                    // blocks handed out by block_alloc: bytes
    std::map<hipStream_t, std::multimap<size_t, void*>> free_of;  // per stream: bytes -> cached block
    size_t cached = 0;                                            // bytes in the free lists
};
BlockCache g_blocks[kMaxDevices];

size_t size_class(size_t b) {
    if (b <= 256) return 256;
    size_t p = 256;
    while (p < b) p <<= 1;
    if (p <= (size_t(1) << 20)) return p;
    const size_t step = p >> 3;
    return (b + step - 1) / step * step;
}
BlockCache* cache_here() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    return &g_blocks[dev];
}
}  // namespace

void* block_alloc(size_t bytes, hipStream_t s, hipError_t* err) {
    *err = hipSuccess;
    const size_t want = size_class(bytes);
    BlockCache* c = cache_here();
    if (c) {
        std::lock_guard<std::mutex> lk(c->mu);
        auto fl = c->free_of.find(s);
        if (fl != c->free_of.end()) {
            auto it = fl->second.lower_bound(want);
            if (it != fl->second.end() && it->first <= 2 * want) {
                void* p = it->second;
                c->cached -= it->first;
                fl->second.erase(it);
                return p;
            }
        }
    }
    void* p = nullptr;
    hipMemPool_t pool = device_pool();
    hipError_t e = pool ? hipMallocFromPoolAsync(&p, want, pool, s) : hipMallocAsync(&p, want, s);
    if (e != hipSuccess) {
        *err = e;
        return nullptr;
    }
    if (c) {
        std::lock_guard<std::mutex> lk(c->mu);
        c->size_of[p] = want;
    }
    return p;
}

void block_free(void* p, hipStream_t s) {
    BlockCache* c = cache_here();
    if (c) {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->size_of.find(p);
        if (it != c->size_of.end() && c->cached + it->second <= kBlockCacheMax) {
            c->free_of[s].emplace(it->second, p);
            c->cached += it->second;
            return;
        }
        if (it != c->size_of.end()) c->size_of.erase(it);
    }
    (void)hipFreeAsync(p, s);
}

void block_free_now(void* p) {
    BlockCache* c = cache_here();
    if (c) {
        std::lock_guard<std::mutex> lk(c->mu);
        c->size_of.erase(p);
    }
    (void)hipFree(p);
}

void block_cache_trim(hipStream_t s) {
    BlockCache* c = cache_here();
    if (!c) return;
    std::multimap<size_t, void*> blocks;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto fl = c->free_of.find(s);
        if (fl == c->free_of.end()) return;
        blocks.swap(fl->second);
        c->free_of.erase(fl);
        for (auto& b : blocks) {
            c->cached -= b.first;
            c->size_of.erase(b.second);
        }
    }
    for (auto& b : blocks) (void)hipFree(b.second);
}

namespace {
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// in -> exclusive prefix within the block tile; block total to sums[block]
template <class T, class Load, class Store>
__global__ __launch_bounds__(kScanThreads) void k_scan_local(Load load, Store store, T* sums, int64_t n) {
    __shared__ T wave_tot[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < n) ? load(base + i) : T(0);
        sum += v[i];
    }
    T incl = wave_incl_scan(sum);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wave_tot[wid] = incl;
    __syncthreads();
    T woff = 0;
    for (int w = 0; w < wid; ++w) woff += wave_tot[w];
    T run = woff + incl - sum;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) store(base + i, run);
        run += v[i];
    }
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = run;
}

template <class T, class Load, class Store>
__global__ __launch_bounds__(kScanThreads) void k_scan_add(Load load, Store store, const T* offs, int64_t n) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
    const T o = offs[blockIdx.x];
    if (o == 0) return;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t j = base + (int64_t)i * kScanThreads;
        if (j < n) store(j, load(j) + o);
    }
}

template <class T>
struct ArrLoad {
    const T* p;
    __device__ T operator()(int64_t i) const { return p[i]; }
};
template <class T>
struct ArrStore {
    T* p;
    __device__ void operator()(int64_t i, T v) const { p[i] = v; }
};
struct PopcLoad {
    const RankWord* g;
    __device__ uint32_t operator()(int64_t i) const { return (uint32_t)__popc(g[i].bits); }
};
struct RankStore {
    RankWord* g;
    __device__ void operator()(int64_t i, uint32_t v) const { g[i].rank = v; }
};
struct RankLoad {
    const RankWord* g;
    __device__ uint32_t operator()(int64_t i) const { return g[i].rank; }
};

// Launches the scan levels (block scans, the scan of the block sums, add-back)
// without waiting; returns the device address of the grand total.  The
// per-level block-sum buffers stay alive in `keep` until the caller syncs.
template <class T, class Load, class Store, class AddLoad>
const T* scan_launch(Load load, Store store, AddLoad addload, int64_t n, hipStream_t s, std::vector<DBuf<T>>& keep) {
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    keep.emplace_back(nb);
    T* sums = keep.back().p;
    hipLaunchKernelGGL((k_scan_local<T, Load, Store>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, load, store, sums,
                       n);
    ITS_LAUNCH_CHECK();
    if (nb == 1) return sums;   // the single block's total
    const T* total = scan_launch<T>(ArrLoad<T>{sums}, ArrStore<T>{sums}, ArrLoad<T>{sums}, nb, s, keep);
    hipLaunchKernelGGL((k_scan_add<T, AddLoad, Store>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, addload, store,
                       (const T*)sums, n);
    ITS_LAUNCH_CHECK();
    return total;
}

template <class T, class Load, class Store, class AddLoad>
uint64_t scan_impl(Load load, Store store, AddLoad addload, int64_t n, hipStream_t s) {
    if (n <= 0) return 0;
    std::vector<DBuf<T>> keep;
    keep.reserve(8);
    const T* d_total = scan_launch<T>(load, store, addload, n, s, keep);
    T total;
    ITS_HIP(hipMemcpyAsync(&total, d_total, sizeof(T), hipMemcpyDeviceToHost, s));
    ITS_HIP(hipStreamSynchronize(s));
    return (uint64_t)total;
}
}  // namespace

uint64_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s) {
    return scan_impl<uint32_t>(ArrLoad<uint32_t>{in}, ArrStore<uint32_t>{out}, ArrLoad<uint32_t>{out}, n, s);
}
uint64_t scan_exclusive_u64(const uint64_t* in, uint64_t* out, int64_t n, hipStream_t s) {
    return scan_impl<uint64_t>(ArrLoad<uint64_t>{in}, ArrStore<uint64_t>{out}, ArrLoad<uint64_t>{out}, n, s);
}
uint64_t rank_build(RankWord* g, int64_t nwords, hipStream_t s) {
    return scan_impl<uint32_t>(PopcLoad{g}, RankStore{g}, RankLoad{g}, nwords, s);
}
void rank_build_async(RankWord* g, int64_t nwords, hipStream_t s, uint32_t* d_total) {
    if (nwords <= 0) {
        ITS_HIP(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
        return;
    }
    // the block-sum buffers are released when this returns: stream-ordered
    // after the scan when s is the allocation stream, else after a full wait
    std::vector<DBuf<uint32_t>> keep;
    keep.reserve(8);
    const uint32_t* t = scan_launch<uint32_t>(PopcLoad{g}, RankStore{g}, RankLoad{g}, nwords, s, keep);
    ITS_HIP(hipMemcpyAsync(d_total, t, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
}
void scan_exclusive_u32_async(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, uint32_t* d_total) {
    if (n <= 0) {
        ITS_HIP(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
        return;
    }
    // block-sum buffers: released as in rank_build_async
    std::vector<DBuf<uint32_t>> keep;
    keep.reserve(8);
    const uint32_t* t =
        scan_launch<uint32_t>(ArrLoad<uint32_t>{in}, ArrStore<uint32_t>{out}, ArrLoad<uint32_t>{out}, n, s, keep);
    ITS_HIP(hipMemcpyAsync(d_total, t, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
}
void scan_exclusive_u64_async(const uint64_t* in, uint64_t* out, int64_t n, hipStream_t s, uint64_t* d_total) {
    if (n <= 0) {
        ITS_HIP(hipMemsetAsync(d_total, 0, sizeof(uint64_t), s));
        return;
    }
    // block-sum buffers: released as in rank_build_async
    std::vector<DBuf<uint64_t>> keep;
    keep.reserve(8);
    const uint64_t* t =
        scan_launch<uint64_t>(ArrLoad<uint64_t>{in}, ArrStore<uint64_t>{out}, ArrLoad<uint64_t>{out}, n, s, keep);
    ITS_HIP(hipMemcpyAsync(d_total, t, sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
}

}  // namespace its
