// Solver state of the matrix-free collocation operator (internal).
#pragma once
#include "internal.hpp"

namespace its {

struct Coarse {
    int L, lc;
    uint32_t off[kMaxLevels + 1];   // global vertex offsets
};

// Operator gathers shared by the single-mesh and the batched solver.  Index
// lists are read a chunk at a time and then the chunk's values, so a thread
// waits for two loads per chunk instead of two per entry; the sums keep the
// list order.  Parent lists hold at most 8 entries (one chunk of 8); child
// lists up to 27 (one chunk of 32).
// sum of g over the prolongation list [k0, k1) (coarse parents of a vertex)
template <class T>
__device__ inline T gather_parents(const uint32_t* idx, uint32_t k0, uint32_t k1, const T* g) {
    T s = 0;
    for (uint32_t kb = k0; kb < k1; kb += 8) {
        uint32_t id[8];
        T gv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) id[q] = (kb + q < k1) ? idx[kb + q] : 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) gv[q] = (kb + q < k1) ? g[id[q]] : (T)0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (kb + q < k1) s += gv[q];
    }
    return s;
}
// sum of 2^-k h over the restriction list [k0, k1) (fine children, with
// log2 of the child's parent count k in bits 28 and up), weights exact
template <class T>
__device__ inline T gather_children(const uint32_t* ridx, uint32_t k0, uint32_t k1, const T* h) {
    constexpr int W = 32;
    T s = 0;
    for (uint32_t kb = k0; kb < k1; kb += W) {
        uint32_t v[W];
        T hv[W];
#pragma unroll
        for (int q = 0; q < W; ++q) v[q] = (kb + q < k1) ? ridx[kb + q] : 0u;
#pragma unroll
        for (int q = 0; q < W; ++q) hv[q] = (kb + q < k1) ? h[v[q] & 0x0FFFFFFFu] : (T)0;
#pragma unroll
        for (int q = 0; q < W; ++q)
            if (kb + q < k1) s += ldexp(hv[q], -(int)(v[q] >> 28));
    }
    return s;
}

// ---------------------------------------------------------------- solver
// Per-vertex and per-block bodies of the CGLS passes, shared by the
// single-mesh solver (lscg.hip) and the many-mesh solver (lscg_batch.hip),
// so both compute every value with the same operations in the same order.
constexpr int kFitT = 256;     // threads per block of the passes
constexpr int kFitRed = 512;   // most partials of a reduction (fixed per mesh size)
__host__ __device__ inline unsigned fit_red_blocks(int64_t n) {
    const int64_t b = (n + kFitT - 1) / kFitT;
    return (unsigned)(b < 1 ? 1 : (b < kFitRed ? b : kFitRed));
}
// device scalar slots of a solve
enum FitScal { S_ABSNEW = 0, S_ALPHA, S_BETA, S_TT, S_RN2, S_THRESH, S_RN2_0, S_ZZ, S_ZZ2, S_COUNT };

// g_i = p_i + 2^-k sum_{c in pro(i)} g_c; with UPD the new search direction
// p_i = z_i + beta p_i first
template <class T, bool UPD>
__device__ inline void prolong_at(int64_t i, const uint32_t* ptr, const uint32_t* idx, T* p, const T* z, T* g,
                                  const double* scal) {
    T pi = p[i];
    if (UPD) {
        pi = z[i] + (T)scal[S_BETA] * pi;
        p[i] = pi;
    }
    const uint32_t k0 = ptr[i], k1 = ptr[i + 1];
    const T s = gather_parents(idx, k0, k1, g);
    T w = (k1 > k0) ? (T)1 / (T)(k1 - k0) : (T)0;
    g[i] = pi + w * s;
}
// h_c += sum_{y in res(c)} 2^-k(y) h_y
template <class T>
__device__ inline void restrict_at(int64_t c, const uint32_t* ptr, const uint32_t* ridx, T* h) {
    h[c] += gather_children(ridx, ptr[c], ptr[c + 1], (const T*)h);
}
// x += alpha p; r -= alpha t along the coincidence chain of i; h_i = (D^T r)_i
__device__ inline void update_dt_at(int64_t i, const uint8_t* finest, const uint32_t* up, float* x, const float* p,
                                    float* r, const float* t, float* h, const double* scal, int do_update) {
    float alpha = do_update ? (float)scal[S_ALPHA] : 0.f;
    if (do_update) x[i] = fmaf(alpha, p[i], x[i]);
    if (!finest[i]) {
        h[i] = 0.f;
        return;
    }
    float s = 0.f;
    for (uint32_t j = (uint32_t)i; j != ITS_NONE; j = up[j]) {
        float rj = r[j];
        if (do_update) {
            rj = fmaf(-alpha, t[j], rj);
            r[j] = rj;
        }
        s += rj;
    }
    h[i] = s;
}
// binary64 h_i = (D^T r)_i
__device__ inline void dt64_at(int64_t i, const uint8_t* finest, const uint32_t* up, const double* r, double* h) {
    if (!finest[i]) {
        h[i] = 0.0;
        return;
    }
    double s = 0.0;
    for (uint32_t j = (uint32_t)i; j != ITS_NONE; j = up[j]) s += r[j];
    h[i] = s;
}

// block sum of a blockDim.x-thread block; valid in thread 0
__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}
// Reductions write one partial per block; a one-workgroup kernel that
// follows sums them (kernel boundaries order the writes, so no fences or
// atomics are needed) in a fixed order: per-thread strided sums, then a
// fixed tree.  The total is valid in thread 0.
__device__ inline double sum_partials(const double* part, unsigned np, int stride, int o, double* sh) {
    double v = 0;
    for (unsigned b = threadIdx.x; b < np; b += blockDim.x) v += part[b * stride + o];
    return block_sum(v, sh);
}
// Block blk of nb: t = D g over its grid-stride share and the block's
// partial of ||t||^2 into part[blk]
__device__ inline void apply_down_block(int64_t n, unsigned blk, unsigned nb, const uint32_t* down, const float* g,
                                        float* t, double* part, double* sh) {
    double acc = 0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n; i += (int64_t)nb * blockDim.x) {
        float v = g[down[i]];
        t[i] = v;
        acc += (double)v * (double)v;
    }
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blk] = s;
}
// z = M^-1 nr and the block's partials of nr.nr and nr.z
__device__ inline void precond_block(int64_t n, unsigned blk, unsigned nb, const float* nr, const float* invd, float* z,
                                     double* part, double* sh) {
    double a = 0, c = 0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n; i += (int64_t)nb * blockDim.x) {
        float v = nr[i];
        float zi = invd[i] * v;
        z[i] = zi;
        a += (double)v * (double)v;
        c += (double)v * (double)zi;
    }
    double sa = block_sum(a, sh);
    double sc = block_sum(c, sh);
    if (threadIdx.x == 0) {
        part[2 * blk] = sa;
        part[2 * blk + 1] = sc;
    }
}
__device__ inline void dot64_block(int64_t n, unsigned blk, unsigned nb, const double* a, const double* b,
                                   double* part, double* sh) {
    double acc = 0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n; i += (int64_t)nb * blockDim.x)
        acc += a[i] * b[i];
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blk] = s;
}
// alpha = absNew / ||t||^2
__device__ inline void scalar_alpha_body(const double* part, unsigned np, double* scal, double* sh) {
    double tt = sum_partials(part, np, 1, 0, sh);
    if (threadIdx.x == 0) {
        scal[S_TT] = tt;
        scal[S_ALPHA] = tt > 0 ? scal[S_ABSNEW] / tt : 0.0;
    }
}
// rn2, absNew', the stopping test of Eigen's loop, beta.  ctl: [0] done,
// [1] iterations, [3] iteration limit
__device__ inline void scalar_beta_body(const double* part, unsigned np, double* scal, int* ctl, int init, double* sh) {
    double rn2 = sum_partials(part, np, 2, 0, sh);
    double zz = sum_partials(part, np, 2, 1, sh);
    if (threadIdx.x == 0) {
        scal[S_RN2] = rn2;
        if (init) {
            scal[S_RN2_0] = rn2;
            scal[S_ABSNEW] = zz;
            scal[S_BETA] = 0.0;
            if (rn2 < scal[S_THRESH] || zz == 0.0) ctl[0] = 1;
        } else {
            ctl[1] += 1;
            if (rn2 < scal[S_THRESH] || ctl[1] >= ctl[3]) {
                ctl[0] = 1;
            } else {
                double absOld = scal[S_ABSNEW];
                scal[S_ABSNEW] = zz;
                scal[S_BETA] = zz / absOld;
            }
        }
    }
}

// The coarse levels [lc, L) hold few vertices (at most kCoarse).  One
// workgroup runs all of them with the level values in LDS and a barrier
// between levels, replacing one launch per level.
constexpr int kCoarse = 12288;
constexpr int kCoarseT = 1024;
// The list offsets of the coarse range are staged in LDS first (one round of
// loads) when the range holds fewer than kCoarsePtr vertices, so a level
// waits for its list entries and values only.
constexpr int kCoarsePtr = 2049;
__device__ inline bool stage_ptr(const Coarse& C, const uint32_t* ptr, uint32_t* sptr) {
    const uint32_t base = C.off[C.lc], nptr = C.off[C.L] - base + 1;
    if (nptr > (uint32_t)kCoarsePtr) return false;
    for (uint32_t t = threadIdx.x; t < nptr; t += blockDim.x) sptr[t] = ptr[base + t];
    __syncthreads();
    return true;
}
// level of coarse vertex i (C.lc <= level < C.L)
__device__ inline int coarse_level(const Coarse& C, uint32_t i) {
    int d = C.lc;
    while (d + 1 < C.L && i >= C.off[d + 1]) ++d;
    return d;
}

// g = G p on the coarse levels (with the search direction update when upd).
// When the range holds at most one vertex per thread, every thread loads its
// vertex's list and values before the level loop, which then waits only for
// the LDS values of the level above; the sums keep the same order.
__device__ inline void prolong_coarse_body(const Coarse& C, const uint32_t* ptr, const uint32_t* idx, float* p,
                                           const float* z, float* g, const double* scal, int upd, float* sg,
                                           uint32_t* sptr) {
    const uint32_t base = C.off[C.lc];
    if (C.off[C.L] - base <= blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool own = i < C.off[C.L];
        const int lev = own ? coarse_level(C, i) : -1;
        uint32_t id[8];
        uint32_t k0 = 0, k1 = 0;
        float pi = 0.f;
        if (own) {
            k0 = ptr[i];
            k1 = ptr[i + 1];
#pragma unroll
            for (int q = 0; q < 8; ++q) id[q] = (k0 + q < k1) ? idx[k0 + q] : base;
            pi = p[i];
            if (upd) {
                pi = z[i] + (float)scal[S_BETA] * pi;
                p[i] = pi;
            }
        }
        for (int d = C.L - 1; d >= C.lc; --d) {
            if (lev == d) {
                float s = 0.f;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (k0 + q < k1) s += sg[id[q] - base];
                float w = (k1 > k0) ? 1.f / (float)(k1 - k0) : 0.f;
                float v = pi + w * s;
                sg[i - base] = v;
                g[i] = v;
            }
            __syncthreads();
        }
        return;
    }
    const bool staged = stage_ptr(C, ptr, sptr);
    const float beta = upd ? (float)scal[S_BETA] : 0.f;
    for (int d = C.L - 1; d >= C.lc; --d) {
        for (uint32_t i = C.off[d] + threadIdx.x; i < C.off[d + 1]; i += blockDim.x) {
            float pi = p[i];
            if (upd) {
                pi = z[i] + beta * pi;
                p[i] = pi;
            }
            const uint32_t k0 = staged ? sptr[i - base] : ptr[i], k1 = staged ? sptr[i + 1 - base] : ptr[i + 1];
            float s = 0.f;
            for (uint32_t kb = k0; kb < k1; kb += 8) {
                uint32_t id[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) id[q] = (kb + q < k1) ? idx[kb + q] : base;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (kb + q < k1) s += sg[id[q] - base];
            }
            float w = (k1 > k0) ? 1.f / (float)(k1 - k0) : 0.f;
            float v = pi + w * s;
            sg[i - base] = v;
            g[i] = v;
        }
        __syncthreads();
    }
}
// restriction over the coarse levels [C.lc, C.L) (C.lc >= 1); one vertex
// per thread when the range fits, as in prolong_coarse_body
__device__ inline void restrict_coarse_body(const Coarse& C, const uint32_t* ptr, const uint32_t* ridx, float* h,
                                            float* sh, uint32_t* sptr) {
    const uint32_t base = C.off[C.lc];
    if (C.off[C.L] - base <= blockDim.x) {
        const uint32_t c = base + threadIdx.x;
        const bool own = c < C.off[C.L];
        const int lev = own ? coarse_level(C, c) : -1;
        uint32_t v[32];   // child lists: at most 27 entries
        uint32_t k0 = 0, k1 = 0;
        float hc0 = 0.f, hlow[32];
        if (own) {
            k0 = ptr[c];
            k1 = ptr[c + 1];
#pragma unroll
            for (int q = 0; q < 32; ++q) v[q] = (k0 + q < k1) ? ridx[k0 + q] : base;
            hc0 = h[c];
            // children below the coarse range are final already: read them now
#pragma unroll
            for (int q = 0; q < 32; ++q) hlow[q] = (lev == C.lc && k0 + q < k1) ? h[v[q] & 0x0FFFFFFFu] : 0.f;
        }
        for (int d = C.lc; d < C.L; ++d) {
            if (lev == d) {
                float s = 0.f;
#pragma unroll
                for (int q = 0; q < 32; ++q) {
                    const float hy = (d == C.lc) ? hlow[q] : ((k0 + q < k1) ? sh[(v[q] & 0x0FFFFFFFu) - base] : 0.f);
                    if (k0 + q < k1) s += ldexpf(hy, -(int)(v[q] >> 28));
                }
                const float hc = hc0 + s;
                sh[c - base] = hc;
                h[c] = hc;
            }
            __syncthreads();
        }
        return;
    }
    const bool staged = stage_ptr(C, ptr, sptr);
    for (int d = C.lc; d < C.L; ++d) {
        for (uint32_t c = C.off[d] + threadIdx.x; c < C.off[d + 1]; c += blockDim.x) {
            float s = 0.f;
            const uint32_t k0 = staged ? sptr[c - base] : ptr[c], k1 = staged ? sptr[c + 1 - base] : ptr[c + 1];
            for (uint32_t kb = k0; kb < k1; kb += 32) {   // child lists: at most 27 entries
                uint32_t v[32];
                float hy[32];
#pragma unroll
                for (int q = 0; q < 32; ++q) v[q] = (kb + q < k1) ? ridx[kb + q] : base;   // base: in bounds for both reads
#pragma unroll
                for (int q = 0; q < 32; ++q) {
                    const uint32_t y = v[q] & 0x0FFFFFFFu;
                    hy[q] = (kb + q < k1) ? ((d == C.lc) ? h[y] : sh[y - base]) : 0.f;
                }
#pragma unroll
                for (int q = 0; q < 32; ++q)
                    if (kb + q < k1) s += ldexpf(hy[q], -(int)(v[q] >> 28));
            }
            float hc = h[c] + s;
            sh[c - base] = hc;
            h[c] = hc;
        }
        __syncthreads();
    }
}

struct Fit {
    int64_t n = 0;
    int L = 0;
    int64_t off[kMaxLevels + 1] = {0};
    DBuf<uint32_t> pro_ptr, pro_idx;   // prolongation: coarse parents of each vertex
    DBuf<uint32_t> res_ptr, res_idx;   // restriction: fine children of each vertex (| log2 count << 28)
    DBuf<uint32_t> down, up;           // finest coincident vertex; coincident vertex one level up
    DBuf<uint8_t> finest;
    DBuf<float> invd32;
    DBuf<double> invd64;
    // inner CGLS vectors (binary32)
    DBuf<float> x, r, p, z, t, g, h;
    // binary64 work vectors
    DBuf<double> g64, h64, r64, nr64, b64, tmp64;
    // reductions and device scalars
    DBuf<double> part;
    DBuf<double> scal;     // see Scal
    DBuf<int> ctl;         // [0] done, [1] iterations, [2] structure error
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_chunk = 0;
    int lc = 0;            // first level handled by the one-workgroup coarse kernels
    Coarse coarse{};
    ~Fit() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};
Fit* fit_get(its_svo s);

}  // namespace its
