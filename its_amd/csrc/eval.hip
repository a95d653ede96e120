// Evaluation of f(x) = sum_i lambda_i B_i(x) for gfx950.
//
// Replaces the reference's point evaluation paths:
//   - CPU support sum mt_cpuTest_2 / mq_setInDomainPoints
//     (ThinShells.cpp:1249-1294, SVO.cu:1195-1261);
//   - GPU brute-force row sums reduceRowSumKernel / mq_reduceRowSumKernel
//     (CUDACompute.cu:435-590), which loop over all n basis functions per point.
//
// Field layout, per level d (cells of width w_d = w_0 2^d):
//   cells[d]  : bitset over the R_d^3 cells in Morton order, with a rank
//               directory.  A bit is set iff some corner of the cell is a
//               level-d vertex.
//   blocks[d] : one 32-byte block (8 binary32 coefficients, corner index
//               x | y<<1 | z<<2, absent corners 0) per set bit, in Morton order.
// The only basis functions of level d that can be non-zero at x are the
// corners of the level-d cell containing x.  So f(x) is a sum over levels of
// one trilinear interpolation of a block.  If a cell has no vertex corner,
// none of its descendants has one.  The walk therefore goes root -> leaf and
// stops at the first unmarked cell.  Far-field queries touch only the small
// coarse levels, which stay in L2.
//
// Level collapse: every level's hat sum is trilinear on each cell of any
// finer level (two-scale relation).  So the whole contribution of levels
// >= lc is the trilinear interpolation of one dense lattice grid G of its
// values (the top-down prolongation cascade).  G is stored per level-lc cell
// as one 32-byte block of its 8 corner values, so a query reads one aligned
// cache line for all coarse levels; lc is the finest level whose cell count
// stays under a tunable bound (32^3 cells, 1 MB, by default: L2 resident).
// A query inside the box then walks only the lc sparse fine levels.
//
// Child links ride in the blocks: the first marked child (32 bits) and the
// child mask (8 bits) of a cell sit in the 5 low mantissa bits of its 8
// corner values, so one 32-byte block per level gives both the values and
// the way down, and a query reads one cache line per level and nothing else.
// Scattered reads are what the walk pays for, one L1 fill each.  The
// per-level walk over all levels remains for points outside the box (DENSE
// and RAW modes).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace its {
namespace {
constexpr int kT = 256;

__global__ void k_mark_cells(int64_t nv, int64_t voff, const uint16_t* vx, const uint16_t* vy, const uint16_t* vz,
                             int R, uint32_t* bits) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    int64_t v = voff + i;
    int X = vx[v], Y = vy[v], Z = vz[v];
    for (int k = 0; k < 8; ++k) {
        int cx = X - (k & 1), cy = Y - ((k >> 1) & 1), cz = Z - ((k >> 2) & 1);
        if (cx < 0 || cy < 0 || cz < 0 || cx >= R || cy >= R || cz >= R) continue;
        uint32_t m = morton3(cx, cy, cz);
        atomicOr(&bits[m >> 5], 1u << (m & 31));
    }
}

__global__ void k_pack(const uint32_t* bits, int64_t nwords, RankWord* g) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nwords) g[w] = RankWord{bits[w], 0u};
}

__global__ void k_fill_blocks(int64_t nv, int64_t voff, const uint16_t* vx, const uint16_t* vy, const uint16_t* vz,
                              int R, const RankWord* cells, const double* lambda, float* blocks) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    int64_t v = voff + i;
    int X = vx[v], Y = vy[v], Z = vz[v];
    float lam = (float)lambda[v];
    for (int k = 0; k < 8; ++k) {
        int a = k & 1, b = (k >> 1) & 1, c = (k >> 2) & 1;
        int cx = X - a, cy = Y - b, cz = Z - c;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= R || cy >= R || cz >= R) continue;
        uint32_t idx = bit_rank(cells, morton3(cx, cy, cz));
        blocks[8 * (int64_t)idx + k] = lam;   // vertex sits at corner (a,b,c) of the cell
    }
}

struct EvalParams {
    const RankWord* cells[kMaxLevels];
    const float4* blocks[kMaxLevels];
    const float4* G;                   // per level-lc cell: 8 corner values of levels >= lc
    int L;
    int R;
    int lc;           // collapsed level
    int Rc;           // cells per axis of G (R_lc, a power of two)
    int rc_shift;     // log2(Rc)
    float ox, oy, oz, inv_w;
    double blo[3], bhi[3];   // accept box: open (SUPPORT) or closed (DENSE)
    float flo[3], fhi[3];    // the same bounds as binary32 (see SoaF32::in_box)
    int mode;
    float out_val;
};

// accept-box test in binary64 (the reference compares V3d coordinates)
__device__ inline bool box64(const EvalParams& P, double dx, double dy, double dz) {
    if (P.mode == ITS_EVAL_SUPPORT)
        return !(dx <= P.blo[0] || dy <= P.blo[1] || dz <= P.blo[2] || dx >= P.bhi[0] || dy >= P.bhi[1] ||
                 dz >= P.bhi[2]);
    if (P.mode == ITS_EVAL_DENSE)
        return !(dx < P.blo[0] || dy < P.blo[1] || dz < P.blo[2] || dx > P.bhi[0] || dy > P.bhi[1] || dz > P.bhi[2]);
    return true;
}

struct F32Box {
    // For binary32 inputs the binary64 comparisons are exact in binary32
    // against bounds rounded outward on the host (flo/fhi).
    __device__ bool in_box(const EvalParams& P, float x, float y, float z, double, double, double) const {
        // the reference's "outside" comparisons (a NaN coordinate is not
        // outside), combined with | rather than ||: lane masks, no branches
        if (P.mode == ITS_EVAL_SUPPORT)
            return !((x <= P.flo[0]) | (y <= P.flo[1]) | (z <= P.flo[2]) | (x >= P.fhi[0]) | (y >= P.fhi[1]) |
                     (z >= P.fhi[2]));
        if (P.mode == ITS_EVAL_DENSE)
            return !((x < P.flo[0]) | (y < P.flo[1]) | (z < P.flo[2]) | (x > P.fhi[0]) | (y > P.fhi[1]) |
                     (z > P.fhi[2]));
        return true;
    }
};
struct SoaF32 : F32Box {
    const float *x, *y, *z;
    __host__ __device__ SoaF32(const float* x_, const float* y_, const float* z_) : x(x_), y(y_), z(z_) {}
    __device__ void operator()(int64_t i, float& a, float& b, float& c, double& da, double& db, double& dc) const {
        a = x[i];
        b = y[i];
        c = z[i];
        da = db = dc = 0.0;
    }
};
// interleaved x, y, z per query: the layout of a caller's point array
struct AosF32 : F32Box {
    const float* p;
    __host__ __device__ explicit AosF32(const float* p_) : p(p_) {}
    __device__ void operator()(int64_t i, float& a, float& b, float& c, double& da, double& db, double& dc) const {
        a = p[3 * i];
        b = p[3 * i + 1];
        c = p[3 * i + 2];
        da = db = dc = 0.0;
    }
};
struct AosF64 {
    const double* p;
    __device__ void operator()(int64_t i, float& a, float& b, float& c, double& da, double& db, double& dc) const {
        da = p[3 * i]; db = p[3 * i + 1]; dc = p[3 * i + 2];
        a = (float)da; b = (float)db; c = (float)dc;
    }
    __device__ bool in_box(const EvalParams& P, float, float, float, double dx, double dy, double dz) const {
        return box64(P, dx, dy, dz);
    }
};

// lattice point (ix, iy, iz) = origin + i * voxelSize (x fastest)
struct LatticeLoad {
    double o[3], vs[3];
    int64_t dx, dy;
    __device__ void operator()(int64_t i, float& a, float& b, float& c, double& da, double& db, double& dc) const {
        int64_t ix = i % dx, iy = (i / dx) % dy, iz = i / (dx * dy);
        da = o[0] + (double)ix * vs[0];
        db = o[1] + (double)iy * vs[1];
        dc = o[2] + (double)iz * vs[2];
        a = (float)da; b = (float)db; c = (float)dc;
    }
    __device__ bool in_box(const EvalParams& P, float, float, float, double dx, double dy, double dz) const {
        return box64(P, dx, dy, dz);
    }
};

__device__ inline float hat0(float t) { return fmaxf(0.f, 1.f - fabsf(t)); }
__device__ inline float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

// trilinear interpolation inside a cell (t in [0,1]^3); b0 = corners z=0
// (x | y<<1), b1 = corners z=1
__device__ inline float tri_in(float4 b0, float4 b1, float tx, float ty, float tz) {
    const float lo = lerpf(lerpf(b0.x, b0.y, tx), lerpf(b0.z, b0.w, tx), ty);
    const float hi = lerpf(lerpf(b1.x, b1.y, tx), lerpf(b1.z, b1.w, tx), ty);
    return lerpf(lo, hi, tz);
}

// same with clamped hats, for points outside the clamped cell
__device__ inline float tri_clamped(float4 b0, float4 b1, float tx, float ty, float tz) {
    const float wx0 = hat0(tx), wx1 = hat0(tx - 1.f);
    const float wy0 = hat0(ty), wy1 = hat0(ty - 1.f);
    const float wz0 = hat0(tz), wz1 = hat0(tz - 1.f);
    const float lo = fmaf(wy1, fmaf(wx1, b0.w, wx0 * b0.z), wy0 * fmaf(wx1, b0.y, wx0 * b0.x));
    const float hi = fmaf(wy1, fmaf(wx1, b1.w, wx0 * b1.z), wy0 * fmaf(wx1, b1.y, wx0 * b1.x));
    return fmaf(wz1, hi, wz0 * lo);
}

__device__ inline float pow2neg(int d) { return __uint_as_float((uint32_t)(127 - d) << 23); }

// Link payload of a block: corner value k holds mask bit k in bit 0 and bits
// 4k..4k+3 of the first marked child in bits 1-4.  Values are rounded to a
// multiple of 32 ulps before the payload goes in (relative error at most
// 2^-19), and readers clear the 5 bits.
constexpr uint32_t kLinkBits = 31u;
// The low bytes of the 8 corner words, corner k in byte k (k < 4 in .x,
// k >= 4 in .y): byte gathers, a few byte permutes.
__device__ inline uint2 link_pack(float4 b0, float4 b1) {
    return make_uint2((__float_as_uint(b0.x) & 0xFFu) | ((__float_as_uint(b0.y) & 0xFFu) << 8) |
                          ((__float_as_uint(b0.z) & 0xFFu) << 16) | (__float_as_uint(b0.w) << 24),
                      (__float_as_uint(b1.x) & 0xFFu) | ((__float_as_uint(b1.y) & 0xFFu) << 8) |
                          ((__float_as_uint(b1.z) & 0xFFu) << 16) | (__float_as_uint(b1.w) << 24));
}
// mask bit of child `key` (0..7) from the packed bytes
__device__ inline uint32_t link_has(uint2 pk, uint32_t key) {
    return ((key < 4u ? pk.x : pk.y) >> (8u * (key & 3u))) & 1u;
}
// mask (bit k: bit 0 of byte k) and first child (nibble k: bits 1-4 of byte
// k) from the packed bytes, four bytes per word at once
__device__ inline uint32_t link_unpack(uint2 pk, uint32_t& base) {
    auto m4 = [](uint32_t w) {
        w &= 0x01010101u;
        return (w | (w >> 7) | (w >> 14) | (w >> 21)) & 0xFu;
    };
    auto n4 = [](uint32_t w) {
        w = (w >> 1) & 0x0F0F0F0Fu;
        w |= w >> 4;
        return (w & 0xFFu) | ((w >> 8) & 0xFF00u);
    };
    base = n4(pk.x) | (n4(pk.y) << 16);
    return m4(pk.x) | (m4(pk.y) << 4);
}
__device__ inline float4 link_clear(float4 b) {
    return make_float4(__uint_as_float(__float_as_uint(b.x) & ~kLinkBits), __uint_as_float(__float_as_uint(b.y) & ~kLinkBits),
                       __uint_as_float(__float_as_uint(b.z) & ~kLinkBits), __uint_as_float(__float_as_uint(b.w) & ~kLinkBits));
}
// mask and first child of a linked block; its values with the payload cleared
__device__ inline uint32_t link_decode(float4& b0, float4& b1, uint32_t& base) {
    const uint32_t mask = link_unpack(link_pack(b0, b1), base);
    b0 = link_clear(b0);
    b1 = link_clear(b1);
    return mask;
}

// every level, root -> leaf, clamped cells and hats: any point
__device__ float eval_walk(const EvalParams& P, float ux, float uy, float uz, int cx, int cy, int cz, uint32_t m0) {
    float f = 0.f;
    for (int d = P.L - 1; d >= 0; --d) {
        const uint32_t m = m0 >> (3 * d);
        const RankWord w = P.cells[d][m >> 5];
        const uint32_t bit = 1u << (m & 31);
        if (!(w.bits & bit)) break;
        const uint32_t idx = w.rank + __popc(w.bits & (bit - 1u));
        const float s = pow2neg(d);
        float4 b0 = P.blocks[d][2 * idx], b1 = P.blocks[d][2 * idx + 1];
        if (d > 0 && d < P.lc) {   // linked level: drop the link payload
            b0 = link_clear(b0);
            b1 = link_clear(b1);
        }
        f += tri_clamped(b0, b1, ux * s - (float)(cx >> d), uy * s - (float)(cy >> d), uz * s - (float)(cz >> d));
    }
    return f;
}

// child of the level-(d+1) cell holding voxel (cx, cy, cz): bit d of each
// coordinate, x | y<<1 | z<<2
__device__ inline uint32_t child_key(int cx, int cy, int cz, int d) {
    return __builtin_amdgcn_ubfe((uint32_t)cx, (uint32_t)d, 1u) | (__builtin_amdgcn_ubfe((uint32_t)cy, (uint32_t)d, 1u) << 1) |
           (__builtin_amdgcn_ubfe((uint32_t)cz, (uint32_t)d, 1u) << 2);
}
__device__ inline uint32_t child_index(uint32_t base, uint32_t mask, uint32_t key) {
    return base + __popc(mask & ((1u << key) - 1u));
}
// G cell of voxel (cx, cy, cz): x-fastest over the Rc^3 level-lc cells
__device__ inline uint32_t g_cell(const EvalParams& P, int cx, int cy, int cz) {
    return ((uint32_t)(cz >> P.lc) << (2 * P.rc_shift)) | ((uint32_t)(cy >> P.lc) << P.rc_shift) |
           (uint32_t)(cx >> P.lc);
}
// local coordinates in the level-d cell: u in [0, R), so floor(u 2^-d) is
// floor(u) >> d and the fraction is exact
__device__ inline float tri_level(float4 b0, float4 b1, float ux, float uy, float uz, float s) {
    return tri_in(b0, b1, __builtin_amdgcn_fractf(ux * s), __builtin_amdgcn_fractf(uy * s),
                  __builtin_amdgcn_fractf(uz * s));
}

// The query front end shared by both launch schemes: load, accept-box test,
// and either a finished value (outside the box, or outside [0, R) through the
// clamped walk) or the grid coordinates for the octree walk.
template <class Loader>
__device__ inline bool query_front(const Loader& load, const EvalParams& P, int64_t i, float& f, float& ux, float& uy,
                                   float& uz) {
    float x, y, z;
    double dx, dy, dz;
    load(i, x, y, z, dx, dy, dz);
    if (!load.in_box(P, x, y, z, dx, dy, dz)) {
        f = P.out_val;
        return false;
    }
    ux = (x - P.ox) * P.inv_w;
    uy = (y - P.oy) * P.inv_w;
    uz = (z - P.oz) * P.inv_w;
    const float Rf = (float)P.R;
    if (ux >= 0.f && uy >= 0.f && uz >= 0.f && ux < Rf && uy < Rf && uz < Rf) return true;
    // clamp in float first: far or non-finite points stay defined
    const float fr = (float)(P.R - 1);
    const int cx = (int)fminf(fmaxf(floorf(ux), 0.f), fr);
    const int cy = (int)fminf(fmaxf(floorf(uy), 0.f), fr);
    const int cz = (int)fminf(fmaxf(floorf(uz), 0.f), fr);
    f = eval_walk(P, ux, uy, uz, cx, cy, cz, morton3(cx, cy, cz));
    return false;
}

// Query loop.  Each thread carries K queries in lockstep (strided by the
// grid size, so every load stays coalesced across the wavefront), which puts
// K independent block loads in flight per level.
template <class Loader, int K>
__global__ __launch_bounds__(kT) void k_eval(Loader load, int64_t N, EvalParams P, float* out) {
    const int lc = P.lc;
    const float sc = pow2neg(lc);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < N; i0 += K * stride) {
        float ux[K], uy[K], uz[K], f[K];
        bool act[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int64_t i = i0 + j * stride;
            f[j] = ux[j] = uy[j] = uz[j] = 0.f;
            act[j] = i < N && query_front(load, P, i, f[j], ux[j], uy[j], uz[j]);
        }
        // levels >= lc: one 32-byte G block per query, with the child link
        float4 b0[K], b1[K];
        int cx[K], cy[K], cz[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (!act[j]) continue;
            cx[j] = (int)ux[j];   // u >= 0: truncation is floor
            cy[j] = (int)uy[j];
            cz[j] = (int)uz[j];
            const uint32_t g = g_cell(P, cx[j], cy[j], cz[j]);
            b0[j] = P.G[2 * (size_t)g];
            b1[j] = P.G[2 * (size_t)g + 1];
        }
        uint32_t idx[K];
        bool more = false;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (!act[j]) continue;
            if (lc == 0) {
                f[j] = tri_level(b0[j], b1[j], ux[j], uy[j], uz[j], sc);
                act[j] = false;
                continue;
            }
            uint32_t base;
            const uint32_t mask = link_decode(b0[j], b1[j], base);
            f[j] = tri_level(b0[j], b1[j], ux[j], uy[j], uz[j], sc);
            const uint32_t key = child_key(cx[j], cy[j], cz[j], lc - 1);
            act[j] = (mask >> key) & 1u;
            if (act[j]) {
                idx[j] = child_index(base, mask, key);
                more = true;
            }
        }
        // marked fine levels lc-1 .. 0 through the child links
        for (int d = lc - 1; more && d >= 0; --d) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (!act[j]) continue;
                b0[j] = P.blocks[d][2 * (size_t)idx[j]];
                b1[j] = P.blocks[d][2 * (size_t)idx[j] + 1];
            }
            const float s = pow2neg(d);
            more = false;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (!act[j]) continue;
                if (d == 0) {
                    f[j] += tri_level(b0[j], b1[j], ux[j], uy[j], uz[j], s);
                    continue;
                }
                uint32_t base;
                const uint32_t mask = link_decode(b0[j], b1[j], base);
                f[j] += tri_level(b0[j], b1[j], ux[j], uy[j], uz[j], s);
                const uint32_t key = child_key(cx[j], cy[j], cz[j], d - 1);
                act[j] = (mask >> key) & 1u;
                if (act[j]) {
                    idx[j] = child_index(base, mask, key);
                    more = true;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int64_t i = i0 + j * stride;
            if (i < N) out[i] = f[j];
        }
    }
}

// One query through the per-thread walk (the arithmetic of k_eval with one
// query per thread).
template <class Loader>
__device__ inline float eval_query(const Loader& load, const EvalParams& P, int64_t i) {
    float f = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
    if (!query_front(load, P, i, f, ux, uy, uz)) return f;
    const int lc = P.lc;
    const float sc = pow2neg(lc);
    const int cx = (int)ux, cy = (int)uy, cz = (int)uz;
    const uint32_t g = g_cell(P, cx, cy, cz);
    float4 b0 = P.G[2 * (size_t)g], b1 = P.G[2 * (size_t)g + 1];
    if (lc == 0) return tri_level(b0, b1, ux, uy, uz, sc);
    uint32_t base;
    uint32_t mask = link_decode(b0, b1, base);
    f = tri_level(b0, b1, ux, uy, uz, sc);
    uint32_t key = child_key(cx, cy, cz, lc - 1);
    if (!((mask >> key) & 1u)) return f;
    uint32_t idx = child_index(base, mask, key);
    for (int d = lc - 1; d >= 0; --d) {
        b0 = P.blocks[d][2 * (size_t)idx];
        b1 = P.blocks[d][2 * (size_t)idx + 1];
        if (d == 0) return f + tri_level(b0, b1, ux, uy, uz, pow2neg(d));
        mask = link_decode(b0, b1, base);
        f += tri_level(b0, b1, ux, uy, uz, pow2neg(d));
        key = child_key(cx, cy, cz, d - 1);
        if (!((mask >> key) & 1u)) break;
        idx = child_index(base, mask, key);
    }
    return f;
}

// Many octrees in one launch: every workgroup takes query ranges of one mesh
// at a time from the item list (mesh, begin, end).
struct BatchItem {
    int64_t begin, end;
    int32_t mesh, pad;
};
__global__ __launch_bounds__(kT) void k_eval_batch(const EvalParams* P, const BatchItem* items, int64_t n_items,
                                                   const float* x, const float* y, const float* z, float* out) {
    const SoaF32 ld{x, y, z};
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const BatchItem I = items[it];
        const EvalParams& p = P[I.mesh];
        for (int64_t i = I.begin + threadIdx.x; i < I.end; i += blockDim.x) out[i] = eval_query(ld, p, i);
    }
}

// Query loop with a per-wavefront queue.
//
// Random queries diverge: most stop at the collapsed level, a few walk every
// fine level, and a wavefront runs the longest walk of its 64 lanes.  Here
// every wavefront evaluates the collapsed level for 64 queries at a time (a
// chunk) and pushes only the queries with a marked child into a FIFO of its
// own in LDS.  Whenever the FIFO holds 64 entries the wavefront walks the fine
// levels for the oldest 64, so the fine-level instructions run on full
// wavefronts.
//
// Each output line is written once.  A chunk whose queries all stopped at the
// collapsed level is stored at once; otherwise its 64 results wait in one of
// kSlots LDS slots until its last queued query has been walked, and the whole
// chunk is stored then.  Slots are reused round-robin; the FIFO hands out the
// oldest chunk's entries first, so a slot that is still waiting when its turn
// comes is freed by walking the FIFO's head (with fewer than 64 entries if
// need be).
//
// Entries are the grid coordinates u, the partial sum, the collapsed cell's
// packed link bytes and the result's place (slot, lane); only the queued
// queries unpack the link into mask and first child.  Points on the upper faces
// (u = R) take the clamped walk with everything outside the grid.
constexpr int kWaves = kT / 64;
constexpr int kQueue = 128;   // at most 63 pending + 64 new entries
constexpr int kSlots = 4;     // chunks waiting for fine walks, per wavefront

struct WaveQueue {
    float ux[kWaves][kQueue], uy[kWaves][kQueue], uz[kWaves][kQueue], f[kWaves][kQueue];
    uint32_t tag[kWaves][kQueue];
    uint2 link[kWaves][kQueue];   // the collapsed cell's packed link bytes
    float res[kWaves][kSlots][64];
    int64_t start[kWaves][kSlots];   // first query of the chunk waiting in each slot
};

__device__ inline void lds_fence() {
    // the writes of some lanes are read by other lanes of the same wavefront
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// fine levels lc-1 .. 0 for one queued query (lk: the collapsed cell's
// packed link bytes)
__device__ inline float fine_walk(const EvalParams& P, float ux, float uy, float uz, float f, uint2 lk) {
    const int cx = (int)ux, cy = (int)uy, cz = (int)uz;   // u >= 0: truncation is floor
    uint32_t base;
    const uint32_t mask = link_unpack(lk, base);
    uint32_t idx = child_index(base, mask, child_key(cx, cy, cz, P.lc - 1));
    for (int d = P.lc - 1; d >= 0; --d) {
        float4 b0 = P.blocks[d][2 * (size_t)idx];
        float4 b1 = P.blocks[d][2 * (size_t)idx + 1];
        if (d == 0) return f + tri_level(b0, b1, ux, uy, uz, pow2neg(d));
        uint32_t nbase;
        const uint32_t m = link_decode(b0, b1, nbase);
        f += tri_level(b0, b1, ux, uy, uz, pow2neg(d));
        const uint32_t key = child_key(cx, cy, cz, d - 1);
        if (!((m >> key) & 1u)) break;
        idx = child_index(nbase, m, key);
    }
    return f;
}

// Per-wavefront state of the queue kernel; every member is uniform across the
// wavefront.  pend[s] counts the queued queries of the chunk in slot s (-1: the
// slot is free).
struct WaveState {
    uint32_t head, tail;   // FIFO entries [head, tail), modulo kQueue
    int pend[kSlots];
};

// Walk the oldest min(64, queued) entries, put their results in their slots,
// and store every chunk that has no queued query left.
__device__ inline void walk_head(const EvalParams& P, WaveQueue& Q, WaveState& S, int w, int lane, int64_t off,
                                 int64_t N, float* out) {
    const uint32_t k = min(64u, S.tail - S.head);
    lds_fence();
    uint32_t slot = kSlots;   // no slot for idle lanes
    if ((uint32_t)lane < k) {
        const uint32_t pos = (S.head + lane) & (kQueue - 1);
        const float r = fine_walk(P, Q.ux[w][pos], Q.uy[w][pos], Q.uz[w][pos], Q.f[w][pos], Q.link[w][pos]);
        const uint32_t t = Q.tag[w][pos];
        slot = t >> 6;
        Q.res[w][slot][t & 63] = r;
    }
    S.head += k;
    lds_fence();
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        S.pend[s] -= __popcll(__ballot(slot == (uint32_t)s));
        if (S.pend[s] == 0) {
            const int64_t i = Q.start[w][s] + lane;
            if (i < N) out[off + i] = Q.res[w][s][lane];
            S.pend[s] = -1;
        }
    }
}

template <class Loader>
__global__ __launch_bounds__(kT) void k_eval_q(Loader load, int64_t off, int64_t N, EvalParams P, float* out) {
    __shared__ WaveQueue Q;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lc = P.lc;
    const float sc = pow2neg(lc);
    const uint64_t below = (1ull << lane) - 1ull;
    WaveState S;
    S.head = S.tail = 0u;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) S.pend[s] = -1;
    int cur = 0;   // slot of the next chunk (round-robin)
    const int64_t step = (int64_t)gridDim.x * kT;
    int64_t c = (int64_t)blockIdx.x * kT + w * 64;
    // One loop with a single walk site: walk the FIFO's head when it holds 64
    // entries, when the next chunk's slot still waits for queued queries
    // (they are the FIFO's oldest entries), or to drain it at the end.
    for (;;) {
        const bool more = c < N;
        bool busy = false;
#pragma unroll
        for (int s = 0; s < kSlots; ++s) busy |= (s == cur) && S.pend[s] > 0;
        if (S.tail - S.head >= 64u || (more && busy) || (!more && S.tail != S.head)) {
            walk_head(P, Q, S, w, lane, off, N, out);
            continue;
        }
        if (!more) break;
        const int64_t i = c + lane;
        bool push = false;
        float f = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
        uint2 lk = make_uint2(0u, 0u);
        if (i < N && query_front(load, P, off + i, f, ux, uy, uz)) {
            const int cx = (int)ux, cy = (int)uy, cz = (int)uz;
            const uint32_t g = g_cell(P, cx, cy, cz);
            float4 g0 = P.G[2 * (size_t)g], g1 = P.G[2 * (size_t)g + 1];
            if (lc > 0) {
                // only the pushed queries' fine walks unpack the link
                lk = link_pack(g0, g1);
                push = link_has(lk, child_key(cx, cy, cz, lc - 1));
                g0 = link_clear(g0);
                g1 = link_clear(g1);
            }
            f = tri_level(g0, g1, ux, uy, uz, sc);
        }
        const uint64_t bal = __ballot(push);
        if (bal == 0ull) {
            if (i < N) out[off + i] = f;   // nothing queued: the chunk is complete
        } else {
            if (push) {
                const uint32_t pos = (S.tail + __popcll(bal & below)) & (kQueue - 1);
                Q.ux[w][pos] = ux;
                Q.uy[w][pos] = uy;
                Q.uz[w][pos] = uz;
                Q.f[w][pos] = f;
                Q.tag[w][pos] = ((uint32_t)cur << 6) | (uint32_t)lane;
                Q.link[w][pos] = lk;
            } else {
                Q.res[w][cur][lane] = f;
            }
            S.tail += (uint32_t)__popcll(bal);
#pragma unroll
            for (int s = 0; s < kSlots; ++s)
                if (s == cur) S.pend[s] = __popcll(bal);
            if (lane == 0) Q.start[w][cur] = c;
            cur = (cur + 1) & (kSlots - 1);
        }
        c += step;
    }
}

// One level of the dense prolongation cascade on the full lattice:
// g_d(X) = lambda_d(X) (0 where X is not a vertex) + trilinear interpolation
// of g_{d+1} at X/2, summed in binary64; x-fastest grids.  A thread walks a
// run of zrun (<= 32) points along z at one (X, Y), so the writes of a wave
// (consecutive X) stay contiguous.  The vertex bitset is keyed z-fastest: the
// run's keys lie in at most two of its words, both loaded before the walk.
// Consecutive points share coarse planes, so each plane's bilinear sum is
// formed once.  OutT = float rounds the values of the last level as they are
// written.
constexpr int kZRunMax = 32;
template <typename OutT>
__global__ void k_dense_level(int D, int zrun, const double* __restrict__ gc, int Dcoarse,
                              const RankWord* __restrict__ lat, int64_t voff, const double* __restrict__ lambda,
                              OutT* __restrict__ g) {
    const int64_t nzr = (D + zrun - 1) / zrun;
    const int64_t nt = (int64_t)D * D * nzr;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        const int X = (int)(t % D);
        const int64_t rest = t / D;
        const int Y = (int)(rest % D), Z0 = (int)(rest / D) * zrun;
        const int Z1 = min(Z0 + zrun, D);
        const int x0 = X >> 1, x1 = (X + 1) >> 1, y0 = Y >> 1, y1 = (Y + 1) >> 1;
        const bool two_x = X & 1, two_y = Y & 1;
        // 1 / (number of coarse points averaged): 1, 1/2 or 1/4 in the plane,
        // half that between planes; powers of two, so the products are the
        // quotients exactly
        const double rxy = (two_x ? 0.5 : 1.0) * (two_y ? 0.5 : 1.0);
        const uint32_t key0 = (uint32_t)(((int64_t)X * D + Y) * D);
        const uint32_t wlo = (key0 + (uint32_t)Z0) >> 5;
        const RankWord w0 = lat[wlo], w1 = lat[(key0 + (uint32_t)Z1 - 1u) >> 5];
        // bilinear sum (not yet divided) of the coarse values on plane p
        auto plane = [&](int p) {
            const double* r0 = gc + ((int64_t)p * Dcoarse + y0) * Dcoarse;
            double s = r0[x0];
            if (two_x) s += r0[x1];
            if (two_y) {
                const double* r1 = gc + ((int64_t)p * Dcoarse + y1) * Dcoarse;
                s += r1[x0];
                if (two_x) s += r1[x1];
            }
            return s;
        };
        double cur = gc ? plane(Z0 >> 1) : 0.0;
        for (int Z = Z0; Z < Z1; ++Z) {
            double v = 0.0;
            if (gc) {
                if (Z & 1) {
                    const double nxt = plane((Z + 1) >> 1);
                    v = (cur + nxt) * (0.5 * rxy);
                    cur = nxt;
                } else {
                    v = cur * rxy;
                }
            }
            const uint32_t key = key0 + (uint32_t)Z;
            const RankWord w = (key >> 5) == wlo ? w0 : w1;
            const uint32_t bit = 1u << (key & 31);
            if (w.bits & bit) v += lambda[voff + w.rank + __popc(w.bits & (bit - 1u))];
            g[((int64_t)Z * D + Y) * D + X] = (OutT)v;
        }
    }
}
// The small top levels of the cascade (at most kTopD points per axis) in one
// workgroup: one launch in place of one per level.  Every point's own
// coefficient (0 where the point is not a vertex) is fetched first, all
// levels at once, into LDS; the levels then follow each other in LDS with a
// barrier between them, and each level is also written out.  Same
// arithmetic per point as k_dense_level.
constexpr int kTopD = 17;
constexpr int kTopThreads = 1024;
constexpr int kTopPoints = 8 + 27 + 125 + 729 + 4913;   // D = 2, 3, 5, 9, 17: the most the top levels hold
struct TopLevels {
    int n;                          // levels, coarsest first
    int D[kMaxLevels];              // points per axis
    int base[kMaxLevels + 1];       // first LDS slot of each level
    const RankWord* lat[kMaxLevels];
    int64_t voff[kMaxLevels];
    double* out[kMaxLevels];
};
__global__ __launch_bounds__(kTopThreads) void k_dense_top(TopLevels T, const double* __restrict__ lambda) {
    __shared__ double sv[kTopPoints];
    // own coefficients of every point of every top level
    for (int l = 0; l < T.n; ++l) {
        const int D = T.D[l], n = D * D * D;
        for (int t = threadIdx.x; t < n; t += kTopThreads) {
            const int X = t % D, Y = (t / D) % D, Z = t / (D * D);
            const uint32_t r = bit_find(T.lat[l], (uint32_t)(((int64_t)X * D + Y) * D + Z));
            // below the coarsest level a point that is not a vertex adds -0.0,
            // which leaves every sum as it is (a -0.0 sum included); the
            // coarsest level's values are 0.0 + lambda, as in k_dense_level
            const double lam = r != ITS_NONE ? lambda[T.voff[l] + r] : (l ? -0.0 : 0.0);
            sv[T.base[l] + t] = l ? lam : 0.0 + lam;
        }
    }
    __syncthreads();
    for (int l = 1; l < T.n; ++l) {
        const int D = T.D[l], Dc = T.D[l - 1];
        const double* gc = sv + T.base[l - 1];
        const int n = D * D * D;
        for (int t = threadIdx.x; t < n; t += kTopThreads) {
            const int X = t % D, Y = (t / D) % D, Z = t / (D * D);
            const int xs[2] = {X >> 1, (X + 1) >> 1}, ys[2] = {Y >> 1, (Y + 1) >> 1}, zs[2] = {Z >> 1, (Z + 1) >> 1};
            const bool tx = X & 1, ty = Y & 1, tz = Z & 1;
            const double rxy = (tx ? 0.5 : 1.0) * (ty ? 0.5 : 1.0);
            auto plane = [&](int p) {
                const double* r0 = gc + (p * Dc + ys[0]) * Dc;
                double s = r0[xs[0]];
                if (tx) s += r0[xs[1]];
                if (ty) {
                    const double* r1 = gc + (p * Dc + ys[1]) * Dc;
                    s += r1[xs[0]];
                    if (tx) s += r1[xs[1]];
                }
                return s;
            };
            const double v = tz ? (plane(zs[0]) + plane(zs[1])) * (0.5 * rxy) : plane(zs[0]) * rxy;
            sv[T.base[l] + t] = v + sv[T.base[l] + t];
        }
        __syncthreads();
    }
    for (int l = 0; l < T.n; ++l) {
        const int n = T.D[l] * T.D[l] * T.D[l];
        for (int t = threadIdx.x; t < n; t += kTopThreads) T.out[l][t] = sv[T.base[l] + t];
    }
}

// lattice values (x fastest, D = Rc + 1 per axis) -> one 64-byte block per
// cell: 8 corner values (index x | y<<1 | z<<2), child base, child mask
__global__ void k_cell_blocks(int Rc, const float* lat, float* blocks) {
    const int64_t n = (int64_t)Rc * Rc * Rc;
    const int64_t D = Rc + 1;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        int64_t x = c % Rc, y = (c / Rc) % Rc, z = c / ((int64_t)Rc * Rc);
        float* o = blocks + 8 * c;
        for (int k = 0; k < 8; ++k)
            o[k] = lat[((z + ((k >> 2) & 1)) * D + (y + ((k >> 1) & 1))) * D + (x + (k & 1))];
    }
}

// every marked cell of the child level links into its parent's link word:
// the smallest child index and the bit of its key (marked cells only have
// marked parents, so the parent entry exists)
__global__ void k_link(int64_t n, const uint32_t* codes, const RankWord* parent_cells, int parent_dense_R,
                       uint2* parent_link, int* err) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = codes[i], p = c >> 3;
    uint32_t slot;
    if (parent_dense_R > 0) {
        uint32_t x, y, z;
        demorton3(p, x, y, z);
        slot = (z * (uint32_t)parent_dense_R + y) * (uint32_t)parent_dense_R + x;
    } else {
        slot = bit_find(parent_cells, p);
        if (slot == ITS_NONE) { atomicOr(err, 1); return; }
    }
    atomicMin(&parent_link[slot].x, (uint32_t)i);
    atomicOr(&parent_link[slot].y, 1u << (c & 7));
}

// link words before linking: no child
__global__ void k_link_clear(int64_t n, uint2* link) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        link[i] = make_uint2(ITS_NONE, 0u);
}

// link word of each cell into the 5 low bits of its 8 corner values (value
// rounded to the nearest multiple of 32 ulps first, halves away from zero)
__global__ void k_embed_link(int64_t n, const uint2* link, uint32_t* blocks) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint2 L = link[i];
        for (int k = 0; k < 8; ++k) {
            const uint32_t u = (blocks[8 * i + k] + 16u) & ~31u;
            blocks[8 * i + k] = u | ((L.y >> k) & 1u) | (((L.x >> (4 * k)) & 15u) << 1);
        }
    }
}

__global__ void k_emit_codes(const RankWord* g, int64_t nwords, uint32_t* out) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    RankWord rw = g[w];
    uint32_t b = rw.bits, r = rw.rank;
    while (b) {
        int i = __ffs(b) - 1;
        out[r++] = (uint32_t)(w * 32 + i);
        b &= b - 1;
    }
}

EvalParams make_params(its_svo s, const double box_end[3], int mode, float out_val) {
    ITS_REQUIRE(s->field_ready, ITS_E_STATE, "no coefficient field: run a fit or set lambda first");
    EvalParams P{};
    for (int d = 0; d < s->depth; ++d) {
        P.cells[d] = s->lv[d].cells.p;
        P.blocks[d] = s->lv[d].blocks.p;
    }
    P.G = s->dense_blocks.p;
    P.L = s->depth;
    P.R = s->R;
    P.lc = s->dense_level;
    P.Rc = s->lv[s->dense_level].R;
    P.rc_shift = 0;
    while ((1 << P.rc_shift) < P.Rc) ++P.rc_shift;
    P.ox = (float)s->origin[0];
    P.oy = (float)s->origin[1];
    P.oz = (float)s->origin[2];
    P.inv_w = (float)(1.0 / s->voxel_width);
    P.mode = mode;
    P.out_val = out_val;
    for (int a = 0; a < 3; ++a) {
        if (mode == ITS_EVAL_DENSE) {
            // minRange = boxOrigin - boxWidth, maxRange = boxEnd + boxWidth
            P.blo[a] = s->origin[a] - s->width[a];
            P.bhi[a] = box_end[a] + s->width[a];
        } else {
            P.blo[a] = s->origin[a];
            P.bhi[a] = box_end[a];
        }
        // For a binary32 x: (x <= lo) iff (x <= largest float <= lo),
        // (x >= hi) iff (x >= smallest float >= hi), (x < lo) iff
        // (x < smallest float >= lo), (x > hi) iff (x > largest float <= hi).
        float lo_dn = (float)P.blo[a], lo_up = (float)P.blo[a], hi_dn = (float)P.bhi[a], hi_up = (float)P.bhi[a];
        if ((double)lo_dn > P.blo[a]) lo_dn = std::nextafter(lo_dn, -INFINITY);
        if ((double)lo_up < P.blo[a]) lo_up = std::nextafter(lo_up, INFINITY);
        if ((double)hi_dn > P.bhi[a]) hi_dn = std::nextafter(hi_dn, -INFINITY);
        if ((double)hi_up < P.bhi[a]) hi_up = std::nextafter(hi_up, INFINITY);
        if (mode == ITS_EVAL_DENSE) {
            P.flo[a] = lo_up;
            P.fhi[a] = hi_dn;
        } else {
            P.flo[a] = lo_dn;
            P.fhi[a] = hi_up;
        }
    }
    return P;
}

// persistent grid: enough workgroups to fill every CU several times over
unsigned eval_grid(its_svo s, int64_t N) {
    int64_t want = std::max<int64_t>(1, (int64_t)s->ctx->prop.multiProcessorCount * tune_get(ITS_TUNE_EVAL_BLOCKS_PER_CU));
    return (unsigned)std::min<int64_t>(want, blocks_for(N, kT));
}
template <class Loader>
void launch_eval(Loader ld, its_svo s, int64_t N, const EvalParams& P, float* out, hipStream_t st) {
    if (tune_get(ITS_TUNE_EVAL_QUEUE)) {
        constexpr int64_t kSeg = int64_t(1) << 31;   // queue entries hold 31-bit indices
        for (int64_t off = 0; off < N; off += kSeg) {
            const int64_t n = std::min(kSeg, N - off);
            hipLaunchKernelGGL(k_eval_q<Loader>, dim3(eval_grid(s, n)), dim3(kT), 0, st, ld, off, n, P, out);
            ITS_LAUNCH_CHECK();
        }
        return;
    }
    const int64_t ilp = tune_get(ITS_TUNE_EVAL_ILP);
    const unsigned g = eval_grid(s, (N + ilp - 1) / ilp);
    if (ilp >= 4) hipLaunchKernelGGL((k_eval<Loader, 4>), dim3(g), dim3(kT), 0, st, ld, N, P, out);
    else if (ilp >= 2) hipLaunchKernelGGL((k_eval<Loader, 2>), dim3(g), dim3(kT), 0, st, ld, N, P, out);
    else hipLaunchKernelGGL((k_eval<Loader, 1>), dim3(g), dim3(kT), 0, st, ld, N, P, out);
    ITS_LAUNCH_CHECK();
}
}  // namespace

void field_build(its_svo s) {
    lambda_changing(s);
    ITS_REQUIRE(s->has_lambda, ITS_E_STATE, "no coefficients");
    ITS_REQUIRE(s->depth > 0, ITS_E_STATE, "empty octree");
    hipStream_t st = s->ctx->stream;
    for (int d = 0; d < s->depth; ++d) {
        Level& L = s->lv[d];
        const int R = L.R;
        const int64_t ncell = (int64_t)R * R * R;
        const int64_t nwords = std::max<int64_t>(1, (ncell + 31) / 32);
        DBuf<uint32_t> bits(nwords);
        ITS_HIP(hipMemsetAsync(bits.p, 0, bits.bytes(), st));
        hipLaunchKernelGGL(k_mark_cells, dim3(blocks_for(L.n_verts, kT)), dim3(kT), 0, st, L.n_verts, s->esum_verts[d],
                           s->vx.p, s->vy.p, s->vz.p, R, bits.p);
        ITS_LAUNCH_CHECK();
        L.cells.ensure(nwords);
        hipLaunchKernelGGL(k_pack, dim3(blocks_for(nwords, kT)), dim3(kT), 0, st, bits.p, nwords, L.cells.p);
        ITS_LAUNCH_CHECK();
        L.n_cells = (int64_t)rank_build(L.cells.p, nwords, st);
        L.blocks.ensure(std::max<int64_t>(2 * L.n_cells, 2));
        ITS_HIP(hipMemsetAsync(L.blocks.p, 0, 2 * L.n_cells * sizeof(float4), st));
        hipLaunchKernelGGL(k_fill_blocks, dim3(blocks_for(L.n_verts, kT)), dim3(kT), 0, st, L.n_verts,
                           s->esum_verts[d], s->vx.p, s->vy.p, s->vz.p, R, L.cells.p, s->lambda64.p,
                           (float*)L.blocks.p);
        ITS_LAUNCH_CHECK();
    }
    // dense cascade from the root level down to the collapsed level
    int lc = 0;
    const int64_t max_cells = tune_get(ITS_TUNE_DENSE_CELLS);
    while (lc < s->depth - 1 && (int64_t)s->lv[lc].R * s->lv[lc].R * s->lv[lc].R > max_cells) ++lc;
    s->dense_level = lc;
    {
        DBuf<float> lat;
        dense_cascade(s, lc, lat);
        const int Rc = s->lv[lc].R;
        const int64_t nc = (int64_t)Rc * Rc * Rc;
        s->dense_blocks.ensure(2 * nc);
        hipLaunchKernelGGL(k_cell_blocks, dim3((unsigned)std::min<int64_t>(blocks_for(nc, kT), 65536)), dim3(kT), 0, st,
                           Rc, lat.p, (float*)s->dense_blocks.p);
        ITS_LAUNCH_CHECK();
    }
    if (lc == 0) {   // no linked levels: the values stay exact
        ITS_HIP(hipStreamSynchronize(st));
        s->field_ready = true;
        return;
    }
    // child links of every level above the finest, down to the collapsed one:
    // a link word (first marked child, child mask) per cell, then folded into
    // the cell's block
    const int64_t nc_dense = (int64_t)s->lv[lc].R * s->lv[lc].R * s->lv[lc].R;
    std::vector<DBuf<uint2>> links(lc + 1);
    for (int d = 1; d <= lc; ++d) {
        const int64_t m = d == lc ? nc_dense : std::max<int64_t>(s->lv[d].n_cells, 1);
        links[d].alloc(m);
        hipLaunchKernelGGL(k_link_clear, dim3((unsigned)std::min<int64_t>(blocks_for(m, kT), 65536)), dim3(kT), 0, st,
                           m, links[d].p);
        ITS_LAUNCH_CHECK();
    }
    DBuf<int> err(1);
    ITS_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
    for (int d = 0; d < lc; ++d) {   // link the marked cells of level d into level d+1
        Level& L = s->lv[d];
        const int64_t nwords = std::max<int64_t>(1, ((int64_t)L.R * L.R * L.R + 31) / 32);
        DBuf<uint32_t> codes(std::max<int64_t>(L.n_cells, 1));
        hipLaunchKernelGGL(k_emit_codes, dim3(blocks_for(nwords, kT)), dim3(kT), 0, st, L.cells.p, nwords, codes.p);
        ITS_LAUNCH_CHECK();
        const bool dense_parent = (d + 1 == lc);
        hipLaunchKernelGGL(k_link, dim3(blocks_for(L.n_cells, kT)), dim3(kT), 0, st, L.n_cells, codes.p,
                           s->lv[d + 1].cells.p, dense_parent ? s->lv[lc].R : 0, links[d + 1].p, err.p);
        ITS_LAUNCH_CHECK();
        ITS_HIP(hipStreamSynchronize(st));
    }
    ITS_REQUIRE(read1(err.p, st) == 0, ITS_E_STATE, "marked cell without a marked parent");
    for (int d = 1; d <= lc; ++d) {
        const int64_t m = d == lc ? nc_dense : s->lv[d].n_cells;
        if (m == 0) continue;
        uint32_t* blocks = (uint32_t*)(d == lc ? s->dense_blocks.p : s->lv[d].blocks.p);
        hipLaunchKernelGGL(k_embed_link, dim3((unsigned)std::min<int64_t>(blocks_for(m, kT), 65536)), dim3(kT), 0, st,
                           m, (const uint2*)links[d].p, blocks);
        ITS_LAUNCH_CHECK();
    }
    ITS_HIP(hipStreamSynchronize(st));
    s->field_ready = true;
}

namespace {
// runs of 32 points on large levels (fewest bitset words per point); shorter
// runs on small levels, which would otherwise have too few threads
template <typename OutT>
void launch_dense_level(its_svo s, int d, const double* coarse, int Dcoarse, OutT* g, hipStream_t st) {
    const int D = s->lv[d].R + 1;
    int zrun = kZRunMax;
    while (zrun > 4 && (int64_t)D * D * ((D + zrun - 1) / zrun) < (int64_t)1 << 19) zrun >>= 1;
    const int64_t nt = (int64_t)D * D * ((D + zrun - 1) / zrun);
    const unsigned nb = (unsigned)std::min<int64_t>(blocks_for(nt, kT), 65536);
    hipLaunchKernelGGL((k_dense_level<OutT>), dim3(nb), dim3(kT), 0, st, D, zrun, coarse, Dcoarse,
                       (const RankWord*)s->lv[d].lattice.p, s->esum_verts[d], (const double*)s->lambda64.p, g);
    ITS_LAUNCH_CHECK();
}

// binary64 cascade from the root level down to level stop + 1, in the octree's
// two grow-only cascade buffers; returns the buffer holding level stop + 1
// (nullptr when stop is the root level)
const double* cascade_above(its_svo s, int stop, hipStream_t st) {
    const double* prev = nullptr;
    int Dprev = 0, k = 0, d = s->depth - 1;
    // both buffers sized for the largest level first: the pointers taken
    // below must not move
    size_t most = 1;
    for (int e = s->depth - 1; e > stop; --e) most = std::max(most, (size_t)(s->lv[e].R + 1) * (s->lv[e].R + 1) * (s->lv[e].R + 1));
    s->casc[0].ensure(most);
    s->casc[1].ensure(most);
    // the small top levels in one launch
    TopLevels T{};
    for (; d > stop && s->lv[d].R + 1 <= kTopD; --d, k ^= 1) {
        const int D = s->lv[d].R + 1;
        T.base[T.n + 1] = T.base[T.n] + D * D * D;
        T.D[T.n] = D;
        T.lat[T.n] = s->lv[d].lattice.p;
        T.voff[T.n] = s->esum_verts[d];
        T.out[T.n] = s->casc[k].p;
        ++T.n;
        prev = s->casc[k].p;
        Dprev = D;
    }
    if (T.n) {
        ITS_REQUIRE(T.base[T.n] <= kTopPoints, ITS_E_STATE, "top cascade levels exceed their LDS space");
        hipLaunchKernelGGL(k_dense_top, dim3(1), dim3(kTopThreads), 0, st, T, (const double*)s->lambda64.p);
        ITS_LAUNCH_CHECK();
    }
    for (; d > stop; --d, k ^= 1) {
        const int D = s->lv[d].R + 1;
        launch_dense_level<double>(s, d, prev, Dprev, s->casc[k].p, st);
        prev = s->casc[k].p;
        Dprev = D;
    }
    return prev;
}
}  // namespace

// f at every lattice point of level `stop` (values of levels >= stop only
// when stop > 0), x fastest; binary64 throughout, or rounded to binary32 as
// the last level is written
void dense_cascade64(its_svo s, int stop, DBuf<double>& out) {
    hipStream_t st = s->ctx->stream;
    const double* coarse = cascade_above(s, stop, st);
    const int D = s->lv[stop].R + 1;
    out.ensure((size_t)D * D * D);
    launch_dense_level<double>(s, stop, coarse, stop + 1 < s->depth ? s->lv[stop + 1].R + 1 : 0, out.p, st);
    ITS_HIP(hipStreamSynchronize(st));
}

void dense_cascade(its_svo s, int stop, DBuf<float>& out) {
    hipStream_t st = s->ctx->stream;
    const double* coarse = cascade_above(s, stop, st);
    const int D = s->lv[stop].R + 1;
    out.ensure((size_t)D * D * D);
    launch_dense_level<float>(s, stop, coarse, stop + 1 < s->depth ? s->lv[stop + 1].R + 1 : 0, out.p, st);
    ITS_HIP(hipStreamSynchronize(st));
}

void eval_points(its_svo s, const float* x, const float* y, const float* z, int64_t N, const double box_end[3],
                 int mode, float out_val, float* f, hipStream_t st) {
    EvalParams P = make_params(s, box_end, mode, out_val);
    if (N <= 0) return;
    launch_eval(SoaF32{x, y, z}, s, N, P, f, st);
}

void eval_points_aos32(its_svo s, const float* pts, int64_t N, const double box_end[3], int mode, float out_val,
                       float* f, hipStream_t st) {
    EvalParams P = make_params(s, box_end, mode, out_val);
    if (N <= 0) return;
    launch_eval(AosF32{pts}, s, N, P, f, st);
}

void eval_points_aos64(its_svo s, const double* pts, int64_t N, const double box_end[3], int mode, float out_val,
                       float* f, hipStream_t st) {
    EvalParams P = make_params(s, box_end, mode, out_val);
    if (N <= 0) return;
    launch_eval(AosF64{pts}, s, N, P, f, st);
}

namespace {
constexpr int64_t kChunkQ = ITS_NOALLOC_CHUNK;   // queries per work item of the batched launch

int64_t batch_item_count(int64_t count, const int64_t* off) {
    int64_t n = 0;
    for (int64_t m = 0; m < count; ++m) n += (off[m + 1] - off[m] + kChunkQ - 1) / kChunkQ;
    return n;
}
// per-mesh parameters and the work items (mesh, begin, end) of a batch
void batch_tables(its_svo const* svos, int64_t count, const int64_t* off, float out_val, EvalParams* P,
                  BatchItem* items) {
    int64_t k = 0;
    for (int64_t m = 0; m < count; ++m) {
        its_svo s = svos[m];
        const double end[3] = {s->origin[0] + s->width[0], s->origin[1] + s->width[1], s->origin[2] + s->width[2]};
        P[m] = make_params(s, end, ITS_EVAL_SUPPORT, out_val);
        for (int64_t b = off[m]; b < off[m + 1]; b += kChunkQ)
            items[k++] = BatchItem{b, std::min(off[m + 1], b + kChunkQ), (int32_t)m, 0};
    }
}
void launch_batch(its_ctx ctx, const EvalParams* dP, const BatchItem* dI, int64_t n_items, const float* x,
                  const float* y, const float* z, float* f, hipStream_t st) {
    const unsigned grid = (unsigned)std::min<int64_t>(
        n_items, (int64_t)ctx->prop.multiProcessorCount * tune_get(ITS_TUNE_EVAL_BLOCKS_PER_CU));
    hipLaunchKernelGGL(k_eval_batch, dim3(grid), dim3(kT), 0, st, dP, dI, n_items, x, y, z, f);
    ITS_LAUNCH_CHECK();
}
}  // namespace

// queries [off[m], off[m+1]) belong to octree m; support semantics with the
// octree's own box end (origin + width) and out_val for points outside
void eval_batch(its_svo const* svos, int64_t count, const int64_t* off, const float* x, const float* y,
                const float* z, float out_val, float* f, hipStream_t st) {
    if (count <= 0 || off[count] <= 0) return;
    const int64_t n_items = batch_item_count(count, off);
    std::vector<EvalParams> P(count);
    std::vector<BatchItem> items(n_items);
    batch_tables(svos, count, off, out_val, P.data(), items.data());
    DBuf<EvalParams> dP(count);
    DBuf<BatchItem> dI(n_items);
    ITS_HIP(hipMemcpyAsync(dP.p, P.data(), count * sizeof(EvalParams), hipMemcpyHostToDevice, st));
    ITS_HIP(hipMemcpyAsync(dI.p, items.data(), n_items * sizeof(BatchItem), hipMemcpyHostToDevice, st));
    launch_batch(svos[0]->ctx, dP.p, dI.p, n_items, x, y, z, f, st);
    ITS_HIP(hipStreamSynchronize(st));
}

// ---- batched evaluation that allocates nothing
// The per-mesh parameters and the work items go to a file-scope device array
// instead of stream-ordered buffers.  Every device holds its own copy of the
// array, at its own address, until the process ends.  Host code never uses
// the array's name as an address: eval_scratch_init takes the address on each
// device with hipGetSymbolAddress, and the launches use that.  The host side
// builds the tables in a static staging area of the same layout, so a call
// allocates neither device nor host memory.  Calls share the one array per
// device (and the staging area): the mutex orders calls of this process, and
// each call waits for its launch before it returns.
struct EvalScratch {
    EvalParams P[kNoAllocMeshes];
    BatchItem I[kNoAllocItems];
};
static __device__ __align__(256) EvalScratch scratch;

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
EvalScratch* g_scratch_dev[kMaxDevices] = {};   // address of each device's array
EvalScratch g_scratch_host;                     // staging, guarded by g_scratch_mu
}  // namespace

void eval_scratch_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return;
    void* p = nullptr;
    ITS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ITS_REQUIRE(p != nullptr, ITS_E_HIP, "no device address for the evaluation scratch");
    g_scratch_dev[device] = (EvalScratch*)p;
}

void eval_batch_noalloc(its_svo const* svos, int64_t count, const int64_t* off, const float* x, const float* y,
                        const float* z, float out_val, float* f, hipStream_t st) {
    ITS_REQUIRE(count <= kNoAllocMeshes, ITS_E_RANGE,
                "batch holds more meshes than the allocation-free evaluation takes (ITS_NOALLOC_MAX_MESHES)");
    const int64_t n_items = batch_item_count(count, off);
    ITS_REQUIRE(n_items <= kNoAllocItems, ITS_E_RANGE,
                "too many queries for the allocation-free evaluation (ITS_NOALLOC_MAX_ITEMS chunks of "
                "ITS_NOALLOC_CHUNK queries)");
    if (count <= 0 || off[count] <= 0) return;
    const its_ctx ctx = svos[0]->ctx;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    EvalScratch* d = g_scratch_dev[ctx->device];
    ITS_REQUIRE(d != nullptr, ITS_E_STATE, "evaluation scratch not set up on this device");
    batch_tables(svos, count, off, out_val, g_scratch_host.P, g_scratch_host.I);
    ITS_HIP(hipMemcpyAsync(d->P, g_scratch_host.P, count * sizeof(EvalParams), hipMemcpyHostToDevice, st));
    ITS_HIP(hipMemcpyAsync(d->I, g_scratch_host.I, n_items * sizeof(BatchItem), hipMemcpyHostToDevice, st));
    launch_batch(ctx, d->P, d->I, n_items, x, y, z, f, st);
    ITS_HIP(hipStreamSynchronize(st));
}

void eval_lattice(its_svo s, const double o[3], const double vs[3], const int64_t dims[3], float* out, hipStream_t st) {
    const double dummy[3] = {0, 0, 0};
    EvalParams P = make_params(s, dummy, ITS_EVAL_RAW, 0.f);
    int64_t N = dims[0] * dims[1] * dims[2];
    if (N <= 0) return;
    LatticeLoad ld{{o[0], o[1], o[2]}, {vs[0], vs[1], vs[2]}, dims[0], dims[1]};
    launch_eval(ld, s, N, P, out, st);
}

}  // namespace its
