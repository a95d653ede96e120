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
