// Many meshes in one launch (gfx950).
//
// No reference counterpart: the reference fits one model per process.
//
// Fit: one persistent kernel, one 1024-thread workgroup per mesh, runs each
// mesh's whole solve -- the binary64 refinement loop and the binary32 CGLS
// inside it, the same algorithm, stopping rule and operation order as
// its_lscg_solve -- with the level passes of the matrix-free operator
// separated by workgroup barriers instead of kernel boundaries.  No
// grid-wide synchronisation is needed, so every workgroup runs to its own
// end; the launch is captured once per batch as a hipGraph and replayed.
// Every scalar of the recurrence is computed by all threads from the same
// block sums, so all threads take the same branches and reach every barrier.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

#include "fit.hpp"
#include "lscg_batch.hpp"

namespace its {
namespace {
constexpr int kB = 1024;
constexpr int kBW = kB / 64;

struct FitDesc {
    int64_t n;
    int L;
    int warm;
    int64_t max_iter;
    double tol;
    int64_t off[kMaxLevels + 1];
    const uint32_t *pro_ptr, *pro_idx, *res_ptr, *res_idx, *down, *up;
    const uint8_t* finest;
    const float* invd32;
    const double* b;
    double* x64;
    float *x, *r, *p, *z, *t, *g, *h;
    double *g64, *r64, *h64;
    double* stats;   // 8 values, see its_batch_fit
};

// block sum, result in every thread (fixed order: waves, then wave 0..15)
__device__ inline double bsum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    for (int k = 0; k < kBW; ++k) s += sh[k];
    return s;
}

// g = G x: top-down prolongation, binary64
__device__ void prolong64(const FitDesc& D, const double* x, double* g) {
    for (int d = D.L - 1; d >= 0; --d) {
        for (int64_t i = D.off[d] + threadIdx.x; i < D.off[d + 1]; i += kB) {
            const uint32_t k0 = D.pro_ptr[i], k1 = D.pro_ptr[i + 1];
            const double s = gather_parents(D.pro_idx, k0, k1, (const double*)g);
            const double w = (k1 > k0) ? 1.0 / (double)(k1 - k0) : 0.0;
            g[i] = x[i] + w * s;
        }
        __syncthreads();
    }
}
// h = A^T r: sums along the coincidence chains, then bottom-up restriction
template <class T>
__device__ void apply_AT(const FitDesc& D, const T* r, T* h) {
    for (int64_t i = threadIdx.x; i < D.n; i += kB) {
        if (!D.finest[i]) {
            h[i] = 0;
            continue;
        }
        T s = 0;
        for (uint32_t j = (uint32_t)i; j != ITS_NONE; j = D.up[j]) s += r[j];
        h[i] = s;
    }
    __syncthreads();
    for (int d = 1; d < D.L; ++d) {
        for (int64_t c = D.off[d] + threadIdx.x; c < D.off[d + 1]; c += kB) {
            h[c] += gather_children(D.res_idx, D.res_ptr[c], D.res_ptr[c + 1], (const T*)h);
        }
        __syncthreads();
    }
}
// r64 = b - A x64; returns ||r64||^2
__device__ double residual64(const FitDesc& D, double* sh) {
    prolong64(D, D.x64, D.g64);
    double acc = 0;
    for (int64_t i = threadIdx.x; i < D.n; i += kB) {
        const double v = D.b[i] - D.g64[D.down[i]];
        D.r64[i] = v;
        acc += v * v;
    }
    __syncthreads();
    return bsum(acc, sh);
}
__device__ double dot_self(const double* a, int64_t n, double* sh) {
    double acc = 0;
    for (int64_t i = threadIdx.x; i < n; i += kB) acc += a[i] * a[i];
    return bsum(acc, sh);
}

// binary32 CGLS on A dx = r64 to a relative normal residual rel_tol;
// returns the iterations (the recurrence of the single-mesh solver)
__device__ int64_t inner_cgls(const FitDesc& D, double rel_tol, int64_t max_it, double* sh) {
    const int64_t n = D.n;
    for (int64_t i = threadIdx.x; i < n; i += kB) {
        D.x[i] = 0.f;
        D.p[i] = 0.f;
        D.r[i] = (float)D.r64[i];
    }
    __syncthreads();
    apply_AT<float>(D, D.r, D.h);
    double a = 0, c = 0;
    for (int64_t i = threadIdx.x; i < n; i += kB) {
        const float v = D.h[i], zi = D.invd32[i] * v;
        D.z[i] = zi;
        a += (double)v * (double)v;
        c += (double)v * (double)zi;
    }
    __syncthreads();
    const double rn2_0 = bsum(a, sh);
    double abs_new = bsum(c, sh);
    if (abs_new == 0.0) return 0;
    const double thresh = rel_tol * rel_tol * rn2_0;
    double beta = 0.0;
    int64_t it = 0;
    while (it < max_it) {
        // p = z + beta p; g = G p
        for (int d = D.L - 1; d >= 0; --d) {
            for (int64_t i = D.off[d] + threadIdx.x; i < D.off[d + 1]; i += kB) {
                const float pi = D.z[i] + (float)beta * D.p[i];
                D.p[i] = pi;
                const uint32_t k0 = D.pro_ptr[i], k1 = D.pro_ptr[i + 1];
                const float s = gather_parents(D.pro_idx, k0, k1, (const float*)D.g);
                const float w = (k1 > k0) ? 1.f / (float)(k1 - k0) : 0.f;
                D.g[i] = pi + w * s;
            }
            __syncthreads();
        }
        // t = D g, alpha
        double tt = 0;
        for (int64_t i = threadIdx.x; i < n; i += kB) {
            const float v = D.g[D.down[i]];
            D.t[i] = v;
            tt += (double)v * (double)v;
        }
        __syncthreads();
        tt = bsum(tt, sh);
        const float alpha = (float)(tt > 0 ? abs_new / tt : 0.0);
        // x += alpha p; r -= alpha t along each chain; h = D^T r
        for (int64_t i = threadIdx.x; i < n; i += kB) {
            D.x[i] = fmaf(alpha, D.p[i], D.x[i]);
            if (!D.finest[i]) {
                D.h[i] = 0.f;
                continue;
            }
            float s = 0.f;
            for (uint32_t j = (uint32_t)i; j != ITS_NONE; j = D.up[j]) {
                const float rj = fmaf(-alpha, D.t[j], D.r[j]);
                D.r[j] = rj;
                s += rj;
            }
            D.h[i] = s;
        }
        __syncthreads();
        for (int d = 1; d < D.L; ++d) {
            for (int64_t cc = D.off[d] + threadIdx.x; cc < D.off[d + 1]; cc += kB) {
                D.h[cc] += gather_children(D.res_idx, D.res_ptr[cc], D.res_ptr[cc + 1], (const float*)D.h);
            }
            __syncthreads();
        }
        // z = M h, stopping test, beta
        double a2 = 0, c2 = 0;
        for (int64_t i = threadIdx.x; i < n; i += kB) {
            const float v = D.h[i], zi = D.invd32[i] * v;
            D.z[i] = zi;
            a2 += (double)v * (double)v;
            c2 += (double)v * (double)zi;
        }
        __syncthreads();
        const double rn2 = bsum(a2, sh), zz = bsum(c2, sh);
        ++it;
        if (rn2 < thresh || it >= max_it) break;
        beta = zz / abs_new;
        abs_new = zz;
    }
    return it;
}

__global__ __launch_bounds__(kB) void k_fit_batch(const FitDesc* descs) {
    __shared__ double sh[kBW];
    const FitDesc& D = descs[blockIdx.x];
    const int64_t n = D.n;
    const int64_t max_iter = D.max_iter > 0 ? D.max_iter : 2 * n;
    if (!D.warm)
        for (int64_t i = threadIdx.x; i < n; i += kB) D.x64[i] = 0.0;
    __syncthreads();
    apply_AT<double>(D, D.b, D.h64);
    const double rhs2 = dot_self(D.h64, n, sh);
    int64_t iters = 0, outer = 0;
    double rel = 0, rel0 = -1;
    if (rhs2 == 0) {
        for (int64_t i = threadIdx.x; i < n; i += kB) D.x64[i] = 0.0;
        __syncthreads();
    } else {
        for (;;) {
            residual64(D, sh);
            apply_AT<double>(D, D.r64, D.h64);
            rel = sqrt(dot_self(D.h64, n, sh) / rhs2);
            if (rel0 < 0) rel0 = rel;
            if (rel < D.tol || iters >= max_iter) break;
            double want = fmax(1e-3, 0.5 * D.tol / rel);
            if (want >= 1.0) want = 0.5;
            const int64_t it = inner_cgls(D, want, max_iter - iters, sh);
            iters += it;
            for (int64_t i = threadIdx.x; i < n; i += kB) D.x64[i] += (double)D.x[i];
            __syncthreads();
            ++outer;
            if (it == 0) break;
        }
    }
    const double rr = residual64(D, sh);
    const double bb = dot_self(D.b, n, sh);
    if (threadIdx.x == 0) {
        D.stats[0] = (double)iters;
        D.stats[1] = (double)(outer > 0 ? outer - 1 : 0);
        D.stats[2] = rel;
        D.stats[3] = bb > 0 ? sqrt(rr / bb) : 0.0;
        D.stats[4] = sqrt(rr);
        D.stats[5] = rel0 < 0 ? 0.0 : rel0;
        D.stats[6] = rel < D.tol ? 1.0 : 0.0;
    }
}
}  // namespace
}  // namespace its

struct its_batch_s {
    its_ctx ctx = nullptr;
    std::vector<its_svo> svos;
    its::DBuf<its::FitDesc> descs;
    its::DBuf<double> stats;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t side = nullptr;       // the one-launch fit runs here, beside the large meshes
    std::vector<its::FitDesc> host;   // the descriptors the graph was captured with
    its::ManySolve large;             // the large meshes' solve, on the context stream
    ~its_batch_s() {
        // no stream may still use the buffers freed below
        if (side) (void)hipStreamSynchronize(side);
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace its {

its_batch batch_new(its_ctx ctx) {
    its_batch b = new its_batch_s();
    b->ctx = ctx;
    return b;
}
void batch_delete(its_batch b) { delete b; }
void batch_add(its_batch b, its_svo s) {
    ITS_REQUIRE(s->ctx->device == b->ctx->device, ITS_E_INVALID, "octree is on another device");
    ITS_REQUIRE(s->num_verts > 0, ITS_E_STATE, "empty octree");
    // two workgroups (or workers) solving one octree would race on its state
    ITS_REQUIRE(std::find(b->svos.begin(), b->svos.end(), s) == b->svos.end(), ITS_E_INVALID,
                "octree is already in the batch");
    b->svos.push_back(s);
}

// Octrees built on other contexts of the batch's device: their streams
// finish before the batch's streams read the octrees.
static void sync_octree_streams(its_batch b) {
    std::vector<hipStream_t> seen;
    for (its_svo s : b->svos) {
        if (s->ctx == b->ctx || std::find(seen.begin(), seen.end(), s->ctx->stream) != seen.end()) continue;
        seen.push_back(s->ctx->stream);
        ITS_HIP(hipStreamSynchronize(s->ctx->stream));
    }
}

// Fit every mesh of the batch.  Meshes with at most ITS_TUNE_BATCH_MAX_VERTS
// unknowns are solved together by the one-launch kernel (graph-captured,
// replayed on later calls) on the batch's own stream; larger meshes, whose
// solve on one workgroup would bound the launch, run at the same time on the
// context stream through the many-mesh solver (one launch per solver pass
// for all of them).  seconds: wall time of the whole fit.
void batch_fit(its_batch b, double tol, int64_t max_iter, bool warm, its_lscg_stats* out, double* seconds) {
    const int count = (int)b->svos.size();
    ITS_REQUIRE(count > 0, ITS_E_STATE, "empty batch");
    hipStream_t st = b->ctx->stream;
    const int64_t cap = tune_get(ITS_TUNE_BATCH_MAX_VERTS);
    std::vector<int> small, big;
    for (int m = 0; m < count; ++m) {
        its_svo s = b->svos[m];
        ITS_REQUIRE(s->has_rhs, ITS_E_STATE, "a mesh of the batch has no right-hand side");
        ITS_REQUIRE(!warm || s->has_lambda, ITS_E_STATE, "warm batch fit needs coefficients on every mesh");
        (s->num_verts <= cap ? small : big).push_back(m);
    }
    const int ns = (int)small.size();
    std::vector<FitDesc> h(ns);
    b->stats.ensure(8 * (int64_t)std::max(ns, 1));
    for (int q = 0; q < ns; ++q) {
        its_svo s = b->svos[small[q]];
        // the octree's solver state and coefficients live on its own
        // context's stream, where fit_get launches: allocate them there
        AllocStream as(s->ctx->stream);
        Fit* f = fit_get(s);
        lambda_changing(s);
        s->lambda64.ensure(f->n);
        FitDesc& D = h[q];
        D = FitDesc{};
        D.n = f->n;
        D.L = f->L;
        D.warm = warm ? 1 : 0;
        D.max_iter = max_iter;
        D.tol = tol;
        for (int d = 0; d <= f->L; ++d) D.off[d] = f->off[d];
        D.pro_ptr = f->pro_ptr.p;
        D.pro_idx = f->pro_idx.p;
        D.res_ptr = f->res_ptr.p;
        D.res_idx = f->res_idx.p;
        D.down = f->down.p;
        D.up = f->up.p;
        D.finest = f->finest.p;
        D.invd32 = f->invd32.p;
        D.b = s->rhs.p;
        D.x64 = s->lambda64.p;
        D.x = f->x.p;
        D.r = f->r.p;
        D.p = f->p.p;
        D.z = f->z.p;
        D.t = f->t.p;
        D.g = f->g.p;
        D.h = f->h.p;
        D.g64 = f->g64.p;
        D.r64 = f->r64.p;
        D.h64 = f->h64.p;
        D.stats = b->stats.p + 8 * q;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (ns > 0) {
        if (!b->side) ITS_HIP(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
        const bool same = b->gexec && b->host.size() == h.size() &&
                          std::equal(h.begin(), h.end(), b->host.begin(), [](const FitDesc& x, const FitDesc& y) {
                              return std::memcmp(&x, &y, sizeof x) == 0;
                          });
        if (!same) {
            if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->gexec = nullptr;
            b->graph = nullptr;
            b->descs.alloc(ns);
            ITS_HIP(hipStreamSynchronize(st));   // the allocation is ordered on the context stream
            ITS_HIP(hipMemcpyAsync(b->descs.p, h.data(), ns * sizeof(FitDesc), hipMemcpyHostToDevice, b->side));
            ITS_HIP(hipStreamSynchronize(b->side));
            ITS_HIP(hipStreamBeginCapture(b->side, hipStreamCaptureModeThreadLocal));
            hipLaunchKernelGGL(k_fit_batch, dim3(ns), dim3(kB), 0, b->side, (const FitDesc*)b->descs.p);
            ITS_HIP(hipStreamEndCapture(b->side, &b->graph));
            ITS_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
            b->host = h;
        }
        ITS_HIP(hipStreamSynchronize(st));   // inputs written on the context stream are complete
        sync_octree_streams(b);
        ITS_HIP(hipGraphLaunch(b->gexec, b->side));
    }
    // Large meshes: the many-mesh solver on the context stream, one launch
    // per solver pass for all of them, bit-identical per mesh to the
    // single-mesh solver.
    std::vector<its_lscg_stats> big_stats(big.size());
    if (!big.empty()) {
        std::vector<its_svo> set(big.size());
        for (size_t q = 0; q < big.size(); ++q) set[q] = b->svos[big[q]];
        ITS_HIP(hipStreamSynchronize(st));
        sync_octree_streams(b);
        lscg_solve_many(b->large, set, tol, max_iter, warm, big_stats.data(), st);
    }
    if (ns > 0) ITS_HIP(hipStreamSynchronize(b->side));
    ITS_HIP(hipStreamSynchronize(st));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (seconds) *seconds = secs;
    std::vector<double> hs(8 * (size_t)std::max(ns, 1));
    if (ns > 0) {
        ITS_HIP(hipMemcpyAsync(hs.data(), b->stats.p, 8 * (size_t)ns * sizeof(double), hipMemcpyDeviceToHost,
                               b->ctx->stream));
        ITS_HIP(hipStreamSynchronize(b->ctx->stream));
    }
    for (int q = 0; q < ns; ++q) {
        if (!out) break;
        its_lscg_stats& S = out[small[q]];
        S.iterations = (int64_t)hs[8 * q];
        S.refinements = (int64_t)hs[8 * q + 1];
        S.normal_rel = hs[8 * q + 2];
        S.residual_rel = hs[8 * q + 3];
        S.residual_abs = hs[8 * q + 4];
        S.initial_normal_rel = hs[8 * q + 5];
        S.converged = hs[8 * q + 6] != 0.0;
        S.seconds = secs;
    }
    for (size_t q = 0; q < big.size(); ++q)
        if (out) out[big[q]] = big_stats[q];
    for (int m = 0; m < count; ++m) {
        b->svos[m]->has_lambda = true;
        b->svos[m]->field_ready = false;
    }
}

// Rebuild the evaluation fields after a batch fit, each on its octree's
// stream, which it synchronises (its buffers and temporaries are allocated
// and freed in that stream's order).  Octrees of different contexts are
// independent: one host thread per context stream builds them side by side.
void batch_fields(its_batch b) {
    std::vector<hipStream_t> streams;
    std::vector<std::vector<its_svo>> groups;
    for (its_svo s : b->svos) {
        const size_t g = std::find(streams.begin(), streams.end(), s->ctx->stream) - streams.begin();
        if (g == streams.size()) {
            streams.push_back(s->ctx->stream);
            groups.emplace_back();
        }
        groups[g].push_back(s);
    }
    auto build = [](const std::vector<its_svo>& group) {
        for (its_svo s : group) {
            AllocStream as(s->ctx->stream);
            field_build(s);
        }
    };
    if (groups.size() == 1) {
        build(groups[0]);
        return;
    }
    const int device = b->ctx->device;
    std::vector<std::exception_ptr> errs(groups.size());
    std::vector<std::thread> threads;
    for (size_t g = 0; g < groups.size(); ++g)
        threads.emplace_back([&, g]() {
            try {
                ITS_HIP(hipSetDevice(device));
                build(groups[g]);
            } catch (...) {
                errs[g] = std::current_exception();
            }
        });
    for (std::thread& t : threads) t.join();
    for (const std::exception_ptr& e : errs)
        if (e) std::rethrow_exception(e);
}

int64_t batch_count(its_batch b) { return (int64_t)b->svos.size(); }
its_svo batch_svo(its_batch b, int64_t m) { return b->svos[(size_t)m]; }
its_ctx batch_ctx(its_batch b) { return b->ctx; }

}  // namespace its
