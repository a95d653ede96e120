// Solver state of the matrix-free collocation operator (internal).
#pragma once
#include "internal.hpp"

namespace its {

struct Coarse {
    int L, lc;
    uint32_t off[kMaxLevels + 1];   // global vertex offsets
};

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
