// Many octrees' coefficient fits at once (gfx950).
//
// No reference counterpart: the reference fits one model per process
// (ThinShells::cpCoefficients, ThinShells.cpp:511-617, one Eigen solve).
//
// Runs the algorithm of lscg_solve (lscg.hip) for a set of octrees together.
// Every pass of the solver -- one level of the prolongation or restriction,
// the coincidence sums, a reduction, a scalar update -- is one launch over all
// octrees of the set; each block reads its octree and its block within that
// octree from a block table.  Per octree the blocks, the grid-stride shares of
// the reductions and the partial counts are those of the single-octree
// solver, and the per-element arithmetic is the same code (fit.hpp), so every
// octree's iterates, iteration count and coefficients are bit-identical to
// lscg_solve's.  Octrees that have finished skip their blocks.  The binary32
// iterations run as a graph of four, captured once per set; the binary64
// refinement steps are launched by the host, one launch per pass for all
// octrees still refining.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "fit.hpp"
#include "lscg_batch.hpp"

namespace its {
namespace {

constexpr int kT = kFitT;
constexpr int kChunk = 4;        // binary32 iterations per graph launch
constexpr int kMeshBits = 12;    // block table entry: octree << 20 | block
constexpr int kBlockBits = 20;

struct MeshDesc {
    int64_t n;
    int L, lc;
    int64_t off[kMaxLevels + 1];
    Coarse coarse;    // prolongation over the coarse levels (lc < L)
    Coarse rcoarse;   // restriction over the coarse levels (from level 1 on)
    const uint32_t *pro_ptr, *pro_idx, *res_ptr, *res_idx, *down, *up;
    const uint8_t* finest;
    const float* invd32;
    const double* b;
    double* lam;
    float *x, *r, *p, *z, *t, *g, *h;
    double *g64, *r64, *nr64, *h64;
    double *part, *scal;
    int* ctl;          // [0] done, [1] iterations, [3] iteration limit
    const int* act;    // the binary64 passes and the inner set-up include this octree
    const double* want2;
    double* out;       // two read-back slots
};

struct Blk {
    const MeshDesc* D;
    int64_t lb;
};
__device__ inline Blk resolve(const uint32_t* plan, const MeshDesc* descs) {
    const uint32_t e = plan[blockIdx.x];
    return Blk{descs + (e >> kBlockBits), (int64_t)(e & ((1u << kBlockBits) - 1))};
}

// ---- binary32 CGLS passes
__global__ __launch_bounds__(kT) void kb_prolong32(const uint32_t* plan, const MeshDesc* descs, int d) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (M.ctl[0]) return;
    const int64_t i = M.off[d] + B.lb * kT + threadIdx.x;
    if (i >= M.off[d + 1]) return;
    prolong_at<float, true>(i, M.pro_ptr, M.pro_idx, M.p, M.z, M.g, M.scal);
}
__global__ __launch_bounds__(kCoarseT) void kb_prolong_coarse(const uint32_t* plan, const MeshDesc* descs) {
    const MeshDesc& M = *resolve(plan, descs).D;
    if (M.ctl[0]) return;
    __shared__ float sg[kCoarse];
    __shared__ uint32_t sptr[kCoarsePtr];
    prolong_coarse_body(M.coarse, M.pro_ptr, M.pro_idx, M.p, M.z, M.g, M.scal, 1, sg, sptr);
}
__global__ __launch_bounds__(kT) void kb_apply_down(const uint32_t* plan, const MeshDesc* descs) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (M.ctl[0]) return;
    __shared__ double sh[kT / 64];
    apply_down_block(M.n, (unsigned)B.lb, fit_red_blocks(M.n), M.down, M.g, M.t, M.part, sh);
}
__global__ __launch_bounds__(kT) void kb_scalar_alpha(const uint32_t* plan, const MeshDesc* descs) {
    const MeshDesc& M = *resolve(plan, descs).D;
    if (M.ctl[0]) return;
    __shared__ double sh[kT / 64];
    scalar_alpha_body(M.part, fit_red_blocks(M.n), M.scal, sh);
}
__global__ __launch_bounds__(kT) void kb_update_dt(const uint32_t* plan, const MeshDesc* descs, int do_update) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (M.ctl[0]) return;
    const int64_t i = B.lb * kT + threadIdx.x;
    if (i >= M.n) return;
    update_dt_at(i, M.finest, M.up, M.x, M.p, M.r, M.t, M.h, M.scal, do_update);
}
__global__ __launch_bounds__(kT) void kb_restrict32(const uint32_t* plan, const MeshDesc* descs, int d) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (M.ctl[0]) return;
    const int64_t c = M.off[d] + B.lb * kT + threadIdx.x;
    if (c >= M.off[d + 1]) return;
    restrict_at<float>(c, M.res_ptr, M.res_idx, M.h);
}
__global__ __launch_bounds__(kCoarseT) void kb_restrict_coarse(const uint32_t* plan, const MeshDesc* descs) {
    const MeshDesc& M = *resolve(plan, descs).D;
    if (M.ctl[0]) return;
    __shared__ float sh[kCoarse];
    __shared__ uint32_t sptr[kCoarsePtr];
    restrict_coarse_body(M.rcoarse, M.res_ptr, M.res_idx, M.h, sh, sptr);
}
__global__ __launch_bounds__(kT) void kb_precond(const uint32_t* plan, const MeshDesc* descs) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (M.ctl[0]) return;
    __shared__ double sh[kT / 64];
    precond_block(M.n, (unsigned)B.lb, fit_red_blocks(M.n), M.h, M.invd32, M.z, M.part, sh);
}
__global__ __launch_bounds__(kT) void kb_scalar_beta(const uint32_t* plan, const MeshDesc* descs, int init) {
    const MeshDesc& M = *resolve(plan, descs).D;
    if (M.ctl[0]) return;
    __shared__ double sh[kT / 64];
    scalar_beta_body(M.part, fit_red_blocks(M.n), M.scal, M.ctl, init, sh);
}

// ---- inner set-up and the binary64 passes (octrees with act set)
// x = p = 0, r = (float)r64
__global__ __launch_bounds__(kT) void kb_inner_start(const uint32_t* plan, const MeshDesc* descs) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    const int64_t i = B.lb * kT + threadIdx.x;
    if (i >= M.n) return;
    M.x[i] = 0.f;
    M.p[i] = 0.f;
    M.r[i] = (float)M.r64[i];
}
// the stopping threshold is unknown until the first normal residual
__global__ void kb_thresh_zero(const uint32_t* plan, const MeshDesc* descs) {
    const MeshDesc& M = *resolve(plan, descs).D;
    if (*M.act && threadIdx.x == 0) M.scal[S_THRESH] = 0.0;
}
__global__ void kb_thresh_set(const uint32_t* plan, const MeshDesc* descs) {
    const MeshDesc& M = *resolve(plan, descs).D;
    if (*M.act && threadIdx.x == 0) M.scal[S_THRESH] = *M.want2 * M.scal[S_RN2_0];
}
// lambda += x
__global__ __launch_bounds__(kT) void kb_axpy(const uint32_t* plan, const MeshDesc* descs) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    const int64_t i = B.lb * kT + threadIdx.x;
    if (i < M.n) M.lam[i] += (double)M.x[i];
}
__global__ __launch_bounds__(kT) void kb_zero_lam(const uint32_t* plan, const MeshDesc* descs) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    const int64_t i = B.lb * kT + threadIdx.x;
    if (i < M.n) M.lam[i] = 0.0;
}
// g64 = G lambda on level d
__global__ __launch_bounds__(kT) void kb_prolong64(const uint32_t* plan, const MeshDesc* descs, int d) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    const int64_t i = M.off[d] + B.lb * kT + threadIdx.x;
    if (i >= M.off[d + 1]) return;
    prolong_at<double, false>(i, M.pro_ptr, M.pro_idx, M.lam, (const double*)nullptr, M.g64, (const double*)nullptr);
}
// r64 = b - D g64
__global__ __launch_bounds__(kT) void kb_residual64(const uint32_t* plan, const MeshDesc* descs) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    const int64_t i = B.lb * kT + threadIdx.x;
    if (i < M.n) M.r64[i] = M.b[i] - M.g64[M.down[i]];
}
// A^T of b (into h64, which = 0) or of r64 (into nr64, which = 1)
__device__ inline double* at64_dst(const MeshDesc& M, int which) { return which ? M.nr64 : M.h64; }
__global__ __launch_bounds__(kT) void kb_dt64(const uint32_t* plan, const MeshDesc* descs, int which) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    const int64_t i = B.lb * kT + threadIdx.x;
    if (i < M.n) dt64_at(i, M.finest, M.up, which ? M.r64 : M.b, at64_dst(M, which));
}
__global__ __launch_bounds__(kT) void kb_restrict64(const uint32_t* plan, const MeshDesc* descs, int d, int which) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    const int64_t c = M.off[d] + B.lb * kT + threadIdx.x;
    if (c >= M.off[d + 1]) return;
    restrict_at<double>(c, M.res_ptr, M.res_idx, at64_dst(M, which));
}
// binary64 dot products: 0 h64.h64, 1 b.b, 2 nr64.nr64, 3 r64.r64
__device__ inline const double* dot_vec(const MeshDesc& M, int pair) {
    return pair == 0 ? M.h64 : pair == 1 ? M.b : pair == 2 ? M.nr64 : M.r64;
}
__global__ __launch_bounds__(kT) void kb_dot64(const uint32_t* plan, const MeshDesc* descs, int pair) {
    const Blk B = resolve(plan, descs);
    const MeshDesc& M = *B.D;
    if (!*M.act) return;
    __shared__ double sh[kT / 64];
    const double* v = dot_vec(M, pair);
    dot64_block(M.n, (unsigned)B.lb, fit_red_blocks(M.n), v, v, M.part, sh);
}
__global__ __launch_bounds__(kT) void kb_dot_total(const uint32_t* plan, const MeshDesc* descs, int slot) {
    const MeshDesc& M = *resolve(plan, descs).D;
    if (!*M.act) return;
    __shared__ double sh[kT / 64];
    const double t = sum_partials(M.part, fit_red_blocks(M.n), 1, 0, sh);
    if (threadIdx.x == 0) M.out[slot] = t;
}

}  // namespace

// ---------------------------------------------------------------- host side
struct ManySolve::Plans {
    struct Ref {
        uint32_t off = 0, count = 0;
    };
    Ref all, red, mesh, pc, rc;
    Ref lev32[kMaxLevels], lev64[kMaxLevels];
    int max_lc = 0, max_L = 0;
};

ManySolve::ManySolve() : plan(new Plans()) {}
ManySolve::~ManySolve() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
}

namespace {
struct Launcher {
    const ManySolve& ms;
    hipStream_t st;
    const MeshDesc* descs() const { return (const MeshDesc*)ms.descs.p; }
    const uint32_t* at(const ManySolve::Plans::Ref& r) const { return ms.blocks.p + r.off; }
};
}  // namespace

// (Re)build the descriptors and block tables when the set or the solver
// state of one of its octrees changed.
static void many_prepare(ManySolve& ms, const std::vector<its_svo>& svos, hipStream_t st) {
    AllocStream as(st);
    const int M = (int)svos.size();
    ITS_REQUIRE(M > 0 && M < (1 << kMeshBits), ITS_E_RANGE, "too many octrees for one many-mesh solve");
    std::vector<MeshDesc> h(M);
    ms.ctl.ensure(4 * (size_t)M);
    ms.act.ensure(M);
    ms.want2.ensure(M);
    ms.out.ensure(2 * (size_t)M);
    for (int m = 0; m < M; ++m) {
        its_svo s = svos[m];
        Fit* f = fit_get(s);
        MeshDesc& D = h[m];
        std::memset(&D, 0, sizeof D);
        D.n = f->n;
        D.L = f->L;
        D.lc = f->lc;
        for (int d = 0; d <= f->L; ++d) D.off[d] = f->off[d];
        D.coarse = f->coarse;
        D.rcoarse = f->coarse;
        if (D.rcoarse.lc == 0) D.rcoarse.lc = 1;
        D.pro_ptr = f->pro_ptr.p;
        D.pro_idx = f->pro_idx.p;
        D.res_ptr = f->res_ptr.p;
        D.res_idx = f->res_idx.p;
        D.down = f->down.p;
        D.up = f->up.p;
        D.finest = f->finest.p;
        D.invd32 = f->invd32.p;
        D.b = s->rhs.p;
        D.lam = s->lambda64.p;
        D.x = f->x.p;
        D.r = f->r.p;
        D.p = f->p.p;
        D.z = f->z.p;
        D.t = f->t.p;
        D.g = f->g.p;
        D.h = f->h.p;
        D.g64 = f->g64.p;
        D.r64 = f->r64.p;
        D.nr64 = f->nr64.p;
        D.h64 = f->h64.p;
        D.part = f->part.p;
        D.scal = f->scal.p;
        D.ctl = ms.ctl.p + 4 * m;
        D.act = ms.act.p + m;
        D.want2 = ms.want2.p + m;
        D.out = ms.out.p + 2 * m;
    }
    const size_t bytes = h.size() * sizeof(MeshDesc);
    const bool same = ms.gexec && ms.host.size() == bytes && std::memcmp(ms.host.data(), h.data(), bytes) == 0;
    if (same) return;
    if (ms.gexec) (void)hipGraphExecDestroy(ms.gexec);
    if (ms.graph) (void)hipGraphDestroy(ms.graph);
    ms.gexec = nullptr;
    ms.graph = nullptr;
    // block tables
    std::vector<uint32_t> tab;
    ManySolve::Plans& P = *ms.plan;
    P = ManySolve::Plans{};
    auto add = [&](ManySolve::Plans::Ref& r, auto blocks_of) {
        r.off = (uint32_t)tab.size();
        for (int m = 0; m < M; ++m) {
            const int64_t nb = blocks_of(h[m]);
            ITS_REQUIRE(nb < (1 << kBlockBits), ITS_E_RANGE, "octree level too large for the many-mesh solve");
            for (int64_t k = 0; k < nb; ++k) tab.push_back((uint32_t)m << kBlockBits | (uint32_t)k);
        }
        r.count = (uint32_t)tab.size() - r.off;
    };
    auto level = [](int d, bool coarse_excluded) {
        return [d, coarse_excluded](const MeshDesc& D) -> int64_t {
            if (d >= (coarse_excluded ? D.lc : D.L)) return 0;
            const int64_t e = D.off[d + 1] - D.off[d];
            return e > 0 ? blocks_for(e, kT) : 0;
        };
    };
    add(P.all, [](const MeshDesc& D) -> int64_t { return blocks_for(D.n, kT); });
    add(P.red, [](const MeshDesc& D) -> int64_t { return fit_red_blocks(D.n); });
    add(P.mesh, [](const MeshDesc&) -> int64_t { return 1; });
    add(P.pc, [](const MeshDesc& D) -> int64_t { return D.lc < D.L ? 1 : 0; });
    add(P.rc, [](const MeshDesc& D) -> int64_t { return D.lc < D.L && D.rcoarse.lc < D.L ? 1 : 0; });
    for (int m = 0; m < M; ++m) {
        P.max_lc = std::max(P.max_lc, h[m].lc);
        P.max_L = std::max(P.max_L, h[m].L);
    }
    for (int d = 0; d < P.max_L; ++d) {
        add(P.lev32[d], level(d, true));
        add(P.lev64[d], level(d, false));
    }
    ms.blocks.alloc(std::max<size_t>(tab.size(), 1));
    ms.descs.alloc(M * sizeof(MeshDesc));
    ITS_HIP(hipMemcpyAsync(ms.blocks.p, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    ITS_HIP(hipMemcpyAsync(ms.descs.p, h.data(), M * sizeof(MeshDesc), hipMemcpyHostToDevice, st));
    ITS_HIP(hipStreamSynchronize(st));
    ms.host.assign((const unsigned char*)h.data(), (const unsigned char*)h.data() + bytes);
    ms.sizes.resize(M);
    for (int m = 0; m < M; ++m) ms.sizes[m] = h[m].n;
    ms.M = M;
}

namespace {
#define ITS_KB(kern, ref, threads, ...)                                                                          \
    do {                                                                                                         \
        if ((ref).count) {                                                                                       \
            hipLaunchKernelGGL(kern, dim3((ref).count), dim3(threads), 0, L.st, L.at(ref), L.descs(), ##__VA_ARGS__); \
            ITS_LAUNCH_CHECK();                                                                                  \
        }                                                                                                        \
    } while (0)

// one binary32 CGLS iteration of every octree not yet done
void many_iteration(const Launcher& L) {
    const ManySolve::Plans& P = *L.ms.plan;
    ITS_KB(kb_prolong_coarse, P.pc, kCoarseT);
    for (int d = P.max_lc - 1; d >= 0; --d) ITS_KB(kb_prolong32, P.lev32[d], kT, d);
    ITS_KB(kb_apply_down, P.red, kT);
    ITS_KB(kb_scalar_alpha, P.mesh, kT);
    ITS_KB(kb_update_dt, P.all, kT, 1);
    for (int d = 1; d < P.max_lc; ++d) ITS_KB(kb_restrict32, P.lev32[d], kT, d);
    ITS_KB(kb_restrict_coarse, P.rc, kCoarseT);
    ITS_KB(kb_precond, P.red, kT);
    ITS_KB(kb_scalar_beta, P.mesh, kT, 0);
}
// A^T (b or r64) into h64 or nr64
void many_AT64(const Launcher& L, int which) {
    const ManySolve::Plans& P = *L.ms.plan;
    ITS_KB(kb_dt64, P.all, kT, which);
    for (int d = 1; d < P.max_L; ++d) ITS_KB(kb_restrict64, P.lev64[d], kT, d, which);
}
// r64 = b - A lambda
void many_residual(const Launcher& L) {
    const ManySolve::Plans& P = *L.ms.plan;
    for (int d = P.max_L - 1; d >= 0; --d) ITS_KB(kb_prolong64, P.lev64[d], kT, d);
    ITS_KB(kb_residual64, P.all, kT);
}
void many_dot(const Launcher& L, int pair, int slot) {
    const ManySolve::Plans& P = *L.ms.plan;
    ITS_KB(kb_dot64, P.red, kT, pair);
    ITS_KB(kb_dot_total, P.mesh, kT, slot);
}
#undef ITS_KB

template <class T>
void upload(T* d, const std::vector<T>& h, hipStream_t st) {
    ITS_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}
}  // namespace

void lscg_solve_many(ManySolve& ms, const std::vector<its_svo>& svos, double tol, int64_t max_iter, bool warm,
                     its_lscg_stats* out, hipStream_t st) {
    auto t0 = std::chrono::steady_clock::now();
    for (its_svo s : svos) {
        // solver state and coefficients are allocated on the octree's stream
        AllocStream as(s->ctx->stream);
        (void)fit_get(s);
        lambda_changing(s);
        s->lambda64.ensure(s->num_verts);
        ITS_HIP(hipStreamSynchronize(s->ctx->stream));
    }
    many_prepare(ms, svos, st);
    const int M = ms.M;
    const Launcher L{ms, st};
    struct State {
        int64_t iters = 0, outer = 0, limit = 0;
        double rel = 0, rel0 = -1, rhs2 = 0, bb = 0;
        bool r_current = false, refining = false;
    };
    std::vector<State> S(M);
    std::vector<int> act(M, 1), ctl(4 * (size_t)M);
    std::vector<double> want2(M, 0.0), outv(2 * (size_t)M);
    auto read_out = [&]() {
        ITS_HIP(hipMemcpyAsync(outv.data(), ms.out.p, outv.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        ITS_HIP(hipStreamSynchronize(st));
    };
    for (int m = 0; m < M; ++m) S[m].limit = max_iter > 0 ? max_iter : 2 * ms.sizes[m];
    upload(ms.act.p, act, st);
    if (!warm) {
        hipLaunchKernelGGL(kb_zero_lam, dim3(ms.plan->all.count), dim3(kT), 0, st, L.at(ms.plan->all), L.descs());
        ITS_LAUNCH_CHECK();
    }
    // ||A^T b||^2 and ||b||^2
    many_AT64(L, 0);
    many_dot(L, 0, 0);
    many_dot(L, 1, 1);
    read_out();
    bool any_zero = false;
    for (int m = 0; m < M; ++m) {
        S[m].rhs2 = outv[2 * m];
        S[m].bb = outv[2 * m + 1];
        S[m].refining = S[m].rhs2 != 0;
        act[m] = S[m].refining ? 0 : 1;
        any_zero |= !S[m].refining;
    }
    if (any_zero) {   // a zero normal right-hand side: lambda = 0
        upload(ms.act.p, act, st);
        hipLaunchKernelGGL(kb_zero_lam, dim3(ms.plan->all.count), dim3(kT), 0, st, L.at(ms.plan->all), L.descs());
        ITS_LAUNCH_CHECK();
    }
    const double inner_tol = 1e-3;
    for (;;) {
        bool any = false;
        for (int m = 0; m < M; ++m) {
            act[m] = S[m].refining ? 1 : 0;
            any |= S[m].refining;
        }
        if (!any) break;
        // binary64 residual and normal residual of the refining octrees
        upload(ms.act.p, act, st);
        many_residual(L);
        many_AT64(L, 1);
        many_dot(L, 2, 0);
        read_out();
        any = false;
        for (int m = 0; m < M; ++m) {
            State& s = S[m];
            if (!s.refining) continue;
            s.r_current = true;
            s.rel = std::sqrt(outv[2 * m] / s.rhs2);
            if (s.rel0 < 0) s.rel0 = s.rel;
            if (s.rel < tol || s.iters >= s.limit) {
                s.refining = false;
                continue;
            }
            double want = std::max(inner_tol, 0.5 * tol / s.rel);
            if (want >= 1.0) want = 0.5;
            want2[m] = want * want;
            any = true;
        }
        if (!any) break;
        // inner binary32 CGLS of the octrees still refining
        int64_t most = 0;
        for (int m = 0; m < M; ++m) {
            const bool in = S[m].refining;
            act[m] = in ? 1 : 0;
            const int64_t lim = in ? std::min<int64_t>(S[m].limit - S[m].iters, 1 << 30) : 0;
            ctl[4 * m + 0] = in ? 0 : 1;
            ctl[4 * m + 1] = 0;
            ctl[4 * m + 2] = 0;
            ctl[4 * m + 3] = (int)lim;
            most = std::max(most, lim);
        }
        upload(ms.act.p, act, st);
        upload(ms.ctl.p, ctl, st);
        upload(ms.want2.p, want2, st);
        {
            const ManySolve::Plans& P = *ms.plan;
            hipLaunchKernelGGL(kb_inner_start, dim3(P.all.count), dim3(kT), 0, st, L.at(P.all), L.descs());
            ITS_LAUNCH_CHECK();
            hipLaunchKernelGGL(kb_update_dt, dim3(P.all.count), dim3(kT), 0, st, L.at(P.all), L.descs(), 0);
            ITS_LAUNCH_CHECK();
            for (int d = 1; d < P.max_lc; ++d)
                if (P.lev32[d].count) {
                    hipLaunchKernelGGL(kb_restrict32, dim3(P.lev32[d].count), dim3(kT), 0, st, L.at(P.lev32[d]),
                                       L.descs(), d);
                    ITS_LAUNCH_CHECK();
                }
            if (P.rc.count) {
                hipLaunchKernelGGL(kb_restrict_coarse, dim3(P.rc.count), dim3(kCoarseT), 0, st, L.at(P.rc), L.descs());
                ITS_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(kb_thresh_zero, dim3(P.mesh.count), dim3(64), 0, st, L.at(P.mesh), L.descs());
            ITS_LAUNCH_CHECK();
            hipLaunchKernelGGL(kb_precond, dim3(P.red.count), dim3(kT), 0, st, L.at(P.red), L.descs());
            ITS_LAUNCH_CHECK();
            hipLaunchKernelGGL(kb_scalar_beta, dim3(P.mesh.count), dim3(kT), 0, st, L.at(P.mesh), L.descs(), 1);
            ITS_LAUNCH_CHECK();
            hipLaunchKernelGGL(kb_thresh_set, dim3(P.mesh.count), dim3(64), 0, st, L.at(P.mesh), L.descs());
            ITS_LAUNCH_CHECK();
        }
        if (!ms.gexec) {
            ITS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < kChunk; ++k) many_iteration(L);
            ITS_HIP(hipStreamEndCapture(st, &ms.graph));
            ITS_HIP(hipGraphInstantiate(&ms.gexec, ms.graph, nullptr, nullptr, 0));
        }
        for (int64_t launched = 0; launched < most + kChunk;) {
            ITS_HIP(hipGraphLaunch(ms.gexec, st));
            launched += kChunk;
            ITS_HIP(hipMemcpyAsync(ctl.data(), ms.ctl.p, ctl.size() * sizeof(int), hipMemcpyDeviceToHost, st));
            ITS_HIP(hipStreamSynchronize(st));
            bool done = true;
            for (int m = 0; m < M; ++m) done &= ctl[4 * m] != 0;
            if (done) break;
        }
        hipLaunchKernelGGL(kb_axpy, dim3(ms.plan->all.count), dim3(kT), 0, st, L.at(ms.plan->all), L.descs());
        ITS_LAUNCH_CHECK();
        for (int m = 0; m < M; ++m) {
            State& s = S[m];
            if (!s.refining) continue;
            const int64_t it = ctl[4 * m + 1];
            s.iters += it;
            ++s.outer;
            // no inner iteration leaves x = 0, so lambda and r64 are unchanged
            s.r_current = (it == 0);
            if (it == 0) s.refining = false;
        }
    }
    // final residuals (current for every octree except after a zero rhs)
    bool stale = false;
    for (int m = 0; m < M; ++m) {
        act[m] = S[m].r_current ? 0 : 1;
        stale |= !S[m].r_current;
    }
    if (stale) {
        upload(ms.act.p, act, st);
        many_residual(L);
    }
    std::fill(act.begin(), act.end(), 1);
    upload(ms.act.p, act, st);
    many_dot(L, 3, 0);
    read_out();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int m = 0; m < M; ++m) {
        const State& s = S[m];
        its_lscg_stats R{};
        const double rr = outv[2 * m];
        R.iterations = s.iters;
        R.refinements = s.outer > 0 ? s.outer - 1 : 0;
        R.normal_rel = s.rel;
        R.initial_normal_rel = s.rel0 < 0 ? 0 : s.rel0;
        R.residual_abs = std::sqrt(rr);
        R.residual_rel = s.bb > 0 ? std::sqrt(rr / s.bb) : 0.0;
        R.seconds = secs;
        R.converged = s.rel < tol;
        if (out) out[m] = R;
    }
    for (its_svo s : svos) {
        s->has_lambda = true;
        s->field_ready = false;
    }
}

}  // namespace its
