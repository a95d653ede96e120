// Shared device infrastructure of the ITS gfx950 library: error plumbing,
// device buffers, Morton bit tricks, wave64 / block / device-wide scans.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "its_c.h"

namespace its {

// ---------------------------------------------------------------- errors ----
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& msg);

#define ITS_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::its::Error(ITS_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                              __FILE__ + ":" + std::to_string(__LINE__));               \
    } while (0)
#define ITS_REQUIRE(cond, code, msg)                      \
    do {                                                  \
        if (!(cond)) throw ::its::Error((code), (msg));   \
    } while (0)
#define ITS_LAUNCH_CHECK() ITS_HIP(hipGetLastError())

// ---------------------------------------------------------------- buffers ---
// Device allocation owned by a host object.
// Stream for stream-ordered allocation on this thread (the entry points set
// their context's stream).  When set, DBuf allocates and frees with
// hipMallocAsync / hipFreeAsync on it, so temporaries cost no device-wide
// synchronisation (hipFree waits for the whole device); when not set, plain
// hipMalloc / hipFree.  Such memory is used on other streams only after a
// synchronisation, and may be freed with hipFree once its stream is gone.
// The library's own stream-ordered pool on the current device: created on
// first use, kept until the process ends.  It keeps freed blocks (release
// threshold at the maximum), so repeated temporaries are not re-mapped; the
// device's default pool, which other code in the process may use, keeps its
// own settings.  nullptr if the pool cannot be created (then the default
// pool is used).
hipMemPool_t device_pool();
// give the unused blocks of the current device's pool back to the driver
void device_pool_trim();

// Block cache in front of the pool.  Stream-ordered frees and re-allocations
// of the pool cost tens to hundreds of microseconds each, so a block freed
// on a stream is kept in that stream's free list and handed to a later
// allocation on the same stream (stream order makes the earlier uses finish
// first, as with hipFreeAsync followed by hipMallocAsync).  Sizes are
// rounded up to classes (powers of two up to 1 MiB, then eighths of the
// power of two); a request takes a cached block of its class or of up to
// twice its size.  At most kBlockCacheMax bytes per device are kept.
constexpr size_t kBlockCacheMax = size_t(16) << 30;
void* block_alloc(size_t bytes, hipStream_t s, hipError_t* err);   // s != nullptr
void block_free(void* p, hipStream_t s);                            // s != nullptr: cached
void block_free_now(void* p);                                       // hipFree, after a device wait
// free the blocks cached for stream s (nothing may still use them: s idle)
void block_cache_trim(hipStream_t s);

inline hipStream_t& alloc_stream() {
    thread_local hipStream_t s = nullptr;
    return s;
}
struct AllocStream {
    hipStream_t prev;
    explicit AllocStream(hipStream_t s) : prev(alloc_stream()) { alloc_stream() = s; }
    ~AllocStream() { alloc_stream() = prev; }
    AllocStream(const AllocStream&) = delete;
    AllocStream& operator=(const AllocStream&) = delete;
};

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) {
            hipStream_t s = alloc_stream();
            hipError_t e;
            if (s) {
                p = (T*)block_alloc(count * sizeof(T), s, &e);
            } else {
                e = hipMalloc(&p, count * sizeof(T));
            }
            if (e != hipSuccess) {
                p = nullptr;
                n = 0;
                throw Error(ITS_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                             hipGetErrorString(e));
            }
        }
    }
    // grow-only reuse
    void ensure(size_t count) { if (count > n) alloc(count); }
    void release() {
        if (p) {
            hipStream_t s = alloc_stream();
            if (s) block_free(p, s);
            else block_free_now(p);
        }
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

inline unsigned blocks_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// ---------------------------------------------------------------- Morton ----
// x -> bit 3i, y -> bit 3i+1, z -> bit 3i+2 (coordinates < 1024)
__host__ __device__ inline uint32_t spread3(uint32_t x) {
    x &= 0x3FF;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__host__ __device__ inline uint32_t compact3(uint32_t x) {
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030C30C3u;
    x = (x ^ (x >> 4)) & 0x0300F00Fu;
    x = (x ^ (x >> 8)) & 0x030000FFu;
    x = (x ^ (x >> 16)) & 0x000003FFu;
    return x;
}
__host__ __device__ inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
__host__ __device__ inline void demorton3(uint32_t m, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = compact3(m);
    y = compact3(m >> 1);
    z = compact3(m >> 2);
}

// ---------------------------------------------------------------- bitsets ---
// A bitset with a rank directory: word w holds bits [32w, 32w+32) and the
// number of set bits in all words before it.
struct RankWord {
    uint32_t bits;
    uint32_t rank;
};

__device__ inline bool bit_test(const RankWord* g, uint32_t i) { return (g[i >> 5].bits >> (i & 31)) & 1u; }
// index of set bit i among all set bits; caller ensures the bit is set
__device__ inline uint32_t bit_rank(const RankWord* g, uint32_t i) {
    RankWord w = g[i >> 5];
    return w.rank + __popc(w.bits & ((1u << (i & 31)) - 1u));
}
// returns ITS_NONE if not set
__device__ inline uint32_t bit_find(const RankWord* g, uint32_t i) {
    RankWord w = g[i >> 5];
    uint32_t b = 1u << (i & 31);
    if (!(w.bits & b)) return ITS_NONE;
    return w.rank + __popc(w.bits & (b - 1u));
}

// ---------------------------------------------------------------- scans -----
// wave64 inclusive scan of unsigned values
template <class T>
__device__ inline T wave_incl_scan(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T n = __shfl_up(v, o, 64);
        if (lane >= o) v += n;
    }
    return v;
}
template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Device-wide exclusive scan of n values of T (uint32_t or uint64_t).
// out may alias in.  Returns the total (synchronises the stream).
uint64_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s);
uint64_t scan_exclusive_u64(const uint64_t* in, uint64_t* out, int64_t n, hipStream_t s);
// Fill the rank fields of a RankWord array from its bits; returns the total.
uint64_t rank_build(RankWord* g, int64_t nwords, hipStream_t s);
// the same, leaving the total in device memory (no wait)
void rank_build_async(RankWord* g, int64_t nwords, hipStream_t s, uint32_t* d_total);
// The same scans with the total left in device memory at d_total; no wait.
void scan_exclusive_u32_async(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, uint32_t* d_total);
void scan_exclusive_u64_async(const uint64_t* in, uint64_t* out, int64_t n, hipStream_t s, uint64_t* d_total);

template <class T>
void to_host(std::vector<T>& h, const T* d, size_t n, hipStream_t s) {
    h.resize(n);
    if (n) ITS_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    ITS_HIP(hipStreamSynchronize(s));
}
template <class T>
T read1(const T* d, hipStream_t s) {
    T v;
    ITS_HIP(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, s));
    ITS_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace its
