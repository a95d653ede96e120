// Many octrees' coefficient fits at once (internal; lscg_batch.hip).
#pragma once
#include <memory>
#include <vector>

#include "internal.hpp"

namespace its {

// Device state of a many-octree solve, kept between calls: the descriptors,
// block tables and the captured iteration graph are rebuilt only when the
// set of octrees or one of their solver states changes.
struct ManySolve {
    struct Plans;
    std::unique_ptr<Plans> plan;
    DBuf<unsigned char> descs;
    DBuf<uint32_t> blocks;
    DBuf<int> ctl, act;
    DBuf<double> want2, out;
    std::vector<unsigned char> host;   // the descriptors the tables and graph were built from
    std::vector<int64_t> sizes;        // unknowns per octree
    int M = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    ManySolve();
    ~ManySolve();
    ManySolve(const ManySolve&) = delete;
    ManySolve& operator=(const ManySolve&) = delete;
};

// Fit every octree of svos to its right-hand side (s->rhs) into s->lambda64,
// with lscg_solve's algorithm and results; launches on st.  out[m] receives
// octree m's statistics (seconds: the wall time of the whole call).
void lscg_solve_many(ManySolve& ms, const std::vector<its_svo>& svos, double tol, int64_t max_iter, bool warm,
                     its_lscg_stats* out, hipStream_t st);

}  // namespace its
