// Least-squares conjugate gradient fit of the coefficients for gfx950.
//
// Replaces the assembly and Eigen solve of ThinShells::cpCoefficients
// (ThinShells.cpp:511-617).  The reference assembles the n x n collocation
// matrix A_ij = B_j(v_i) from triplets and calls
// Eigen::LeastSquaresConjugateGradient with its defaults.  That is CGLS with a
// diagonal preconditioner 1/||A_:j||^2, starting from x = 0, and the stop test
// ||A^T(b - Ax)||^2 < tol^2 ||A^T b||^2.
//
// Here A is never assembled.  The basis is hierarchical degree-1 B-splines, so
// the two-scale relation gives A = D G:
//   - G = (I - P)^-1.  It maps coefficients to g, the value of the
//     coarser-or-equal part of f at every vertex.  P prolongates level d+1 to
//     level d by trilinear interpolation.  A vertex with k odd lattice
//     coordinates averages 2^k coarse vertices with weight 2^-k.
//   - D picks, for every vertex, g at the finest vertex at the same position.
//     Finer coincident basis functions are 1 there and the rest vanish.
// A p is one top-down pass over the levels.  A^T r is the transposed
// bottom-up pass: r summed along each coincidence chain, then restricted
// level by level.  Per iteration this moves about 2 x (4 + 4 + 3.4 x 4) bytes
// per vertex, against 16 bytes per non-zero (about 50 per row) for the
// assembled matrix.
//
// Precision: the inner CGLS keeps its vectors and matvec in binary32 and
// accumulates every dot product in binary64 (products of binary32 values are
// exact in binary64).  An outer loop recomputes the true residual b - A lambda
// and the normal residual in binary64.  It refines lambda in binary64 until
// the binary64 normal residual meets the tolerance.
// All reductions are per-block partials summed in fixed order, so results are
// deterministic.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "internal.hpp"
#include "fit.hpp"

namespace its {

void fit_release(Fit* f) { delete f; }

namespace {
constexpr int kT = kFitT;
constexpr int kRedBlocks = kFitRed;  // fixed partial count for deterministic sums

// ---------------------------------------------------------------- structure
__global__ void k_pro_count(int64_t n, const uint8_t* lev, int L, const uint16_t* vx, const uint16_t* vy,
                            const uint16_t* vz, uint32_t* cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (lev[i] == L - 1) { cnt[i] = 0; return; }
    int k = (vx[i] & 1) + (vy[i] & 1) + (vz[i] & 1);
    cnt[i] = 1u << k;
}

struct LatticeSet {
    const RankWord* lat[kMaxLevels];
    uint32_t D[kMaxLevels];
    uint32_t off[kMaxLevels];
};

__device__ inline uint32_t lat_find(const LatticeSet& S, int d, uint32_t X, uint32_t Y, uint32_t Z) {
    uint32_t D = S.D[d];
    if (X >= D || Y >= D || Z >= D) return ITS_NONE;
    uint32_t r = bit_find(S.lat[d], (X * D + Y) * D + Z);
    return r == ITS_NONE ? ITS_NONE : S.off[d] + r;
}

__global__ void k_pro_fill(int64_t n, const uint8_t* lev, int L, const uint16_t* vx, const uint16_t* vy,
                           const uint16_t* vz, LatticeSet S, const uint32_t* ptr, uint32_t* idx, int* err) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d = lev[i];
    if (d == L - 1) return;
    uint32_t X = vx[i], Y = vy[i], Z = vz[i];
    uint32_t xs[2] = {X >> 1, (X + 1) >> 1}, ys[2] = {Y >> 1, (Y + 1) >> 1}, zs[2] = {Z >> 1, (Z + 1) >> 1};
    int nx = (X & 1) ? 2 : 1, ny = (Y & 1) ? 2 : 1, nz = (Z & 1) ? 2 : 1;
    uint32_t o = ptr[i];
    for (int c = 0; c < nz; ++c)
        for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
                uint32_t j = lat_find(S, d + 1, xs[a], ys[b], zs[c]);
                if (j == ITS_NONE) { atomicOr(err, 1); j = (uint32_t)i; }
                idx[o++] = j;
            }
}

__global__ void k_res_count(int64_t n, const uint32_t* ptr, const uint32_t* idx, uint32_t* cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t k = ptr[i]; k < ptr[i + 1]; ++k) atomicAdd(&cnt[idx[k]], 1u);
}
__global__ void k_res_fill(int64_t n, const uint32_t* ptr, const uint32_t* idx, uint32_t* cursor, uint32_t* ridx) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t b = ptr[i], e = ptr[i + 1];
    uint32_t lg = __ffs(e - b) - 1;   // log2 of the count
    for (uint32_t k = b; k < e; ++k) {
        uint32_t pos = atomicAdd(&cursor[idx[k]], 1u);
        ridx[pos] = (uint32_t)i | (lg << 28);
    }
}
// deterministic order inside each restriction row
__global__ void k_res_sort(int64_t n, const uint32_t* ptr, uint32_t* ridx) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t b = ptr[i], e = ptr[i + 1];
    for (uint32_t k = b + 1; k < e; ++k) {
        uint32_t v = ridx[k];
        uint32_t key = v & 0x0FFFFFFFu;
        uint32_t j = k;
        while (j > b && (ridx[j - 1] & 0x0FFFFFFFu) > key) { ridx[j] = ridx[j - 1]; --j; }
        ridx[j] = v;
    }
}

// the list lengths into ctl[0] (parent entries) and ctl[1] (child entries);
// error bit 4 when they differ
__global__ void k_list_sizes(const uint32_t* npro, const uint32_t* nres, int* ctl) {
    ctl[0] = (int)*npro;
    ctl[1] = (int)*nres;
    if (*npro != *nres) ctl[2] |= 4;
}

__global__ void k_updown(int64_t n, const uint8_t* lev, int L, const uint16_t* vx, const uint16_t* vy,
                         const uint16_t* vz, LatticeSet S, uint32_t* up, uint32_t* dn1) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d = lev[i];
    uint32_t X = vx[i], Y = vy[i], Z = vz[i];
    uint32_t u = ITS_NONE, w = ITS_NONE;
    if (d < L - 1 && !(X & 1) && !(Y & 1) && !(Z & 1)) u = lat_find(S, d + 1, X >> 1, Y >> 1, Z >> 1);
    if (d > 0) w = lat_find(S, d - 1, 2 * X, 2 * Y, 2 * Z);
    up[i] = u;
    dn1[i] = w;
}
__global__ void k_down(int64_t n, const uint32_t* dn1, uint32_t* down, uint8_t* finest) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = (uint32_t)i;
    while (dn1[k] != ITS_NONE) k = dn1[k];
    down[i] = k;
    finest[i] = (k == (uint32_t)i);
}

// Row i of A: 1 at itself and at every finer coincident vertex, and the hat
// values of the corners of the cell containing v_i at every coarser level
// (the triplets of ThinShells.cpp:567-601).  Squares accumulate per column
// in 2^-32 fixed point, so the sum is exact integer arithmetic and the result
// is independent of the order of the atomics.  Columns of the coarse levels
// (index >= coarse_off, at most kColLds of them) receive a contribution from
// almost every row; a workgroup first accumulates those in LDS and adds each
// non-zero total to memory once.
constexpr int kColLds = 4096;
__global__ __launch_bounds__(kT) void k_colnorm(int64_t n, const uint8_t* lev, int L, const uint16_t* vx,
                                               const uint16_t* vy, const uint16_t* vz, LatticeSet S,
                                               const uint32_t* dn1, unsigned long long* acc, uint32_t coarse_off,
                                               uint32_t n_coarse, int* err) {
    __shared__ unsigned long long sh[kColLds];
    for (uint32_t k = threadIdx.x; k < n_coarse; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    const unsigned long long one = 1ull << 32;
    auto add = [&](uint32_t col, unsigned long long v) {
        if (col >= coarse_off) atomicAdd(&sh[col - coarse_off], v);
        else atomicAdd(&acc[col], v);
    };
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        add((uint32_t)i, one);
        for (uint32_t k = dn1[i]; k != ITS_NONE; k = dn1[k]) add(k, one);
        int d = lev[i];
        uint32_t X = vx[i], Y = vy[i], Z = vz[i];
        for (int l = d + 1; l < L; ++l) {
            int s = l - d;
            uint32_t mask = (1u << s) - 1u;
            double inv = 1.0 / (double)(1u << s);
            uint32_t q[3] = {X >> s, Y >> s, Z >> s};
            uint32_t rm[3] = {X & mask, Y & mask, Z & mask};
            double hv[3][2];
            int cnt[3];
            for (int a = 0; a < 3; ++a) {
                double fr = (double)rm[a] * inv;
                hv[a][0] = 1.0 - fr;
                hv[a][1] = fr;
                cnt[a] = rm[a] ? 2 : 1;
            }
            for (int c = 0; c < cnt[2]; ++c)
                for (int b = 0; b < cnt[1]; ++b)
                    for (int a = 0; a < cnt[0]; ++a) {
                        uint32_t j = lat_find(S, l, q[0] + a, q[1] + b, q[2] + c);
                        if (j == ITS_NONE) { atomicOr(err, 2); continue; }
                        double v = hv[0][a] * hv[1][b] * hv[2][c];
                        add(j, (unsigned long long)llrint(v * v * 4294967296.0));
                    }
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n_coarse; k += blockDim.x)
        if (sh[k]) atomicAdd(&acc[coarse_off + k], sh[k]);
}
__global__ void k_invdiag(int64_t n, const unsigned long long* acc, double* d64, float* d32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = (double)acc[i] * (1.0 / 4294967296.0);
    double v = s > 0.0 ? 1.0 / s : 1.0;
    d64[i] = v;
    d32[i] = (float)v;
}

// ---------------------------------------------------------------- operator
// g = G p on one level: g_i = p_i + 2^-k sum_{c in pro(i)} g_c.  When upd is
// set the first pass also forms the new search direction p = z + beta p.
template <class T, bool UPD>
__global__ __launch_bounds__(kT) void k_prolong(int64_t b, int64_t e, const uint32_t* ptr, const uint32_t* idx, T* p,
                                               const T* z, T* g, const double* scal, const int* ctl) {
    if (UPD && ctl[0]) return;
    int64_t i = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    prolong_at<T, UPD>(i, ptr, idx, p, z, g, scal);
}

// t = D g and the partials of ||t||^2
__global__ __launch_bounds__(kT) void k_apply_down(int64_t n, const uint32_t* down, const float* g, float* t,
                                                  double* part, const int* ctl) {
    if (ctl[0]) return;
    __shared__ double sh[kT / 64];
    apply_down_block(n, blockIdx.x, gridDim.x, down, g, t, part, sh);
}
// alpha = absNew / ||t||^2
__global__ __launch_bounds__(kT) void k_scalar_alpha(const double* part, unsigned np, double* scal, const int* ctl) {
    if (ctl[0]) return;
    __shared__ double sh[kT / 64];
    scalar_alpha_body(part, np, scal, sh);
}

// x += alpha p; r -= alpha t along each coincidence chain; h = D^T r
__global__ __launch_bounds__(kT) void k_update_dt(int64_t n, const uint8_t* finest, const uint32_t* up, float* x,
                                                 const float* p, float* r, const float* t, float* h, const double* scal,
                                                 const int* ctl, int do_update) {
    if (ctl[0]) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    update_dt_at(i, finest, up, x, p, r, t, h, scal, do_update);
}

// h_c += sum_{y in res(c)} 2^-k(y) h_y for the vertices of one level
template <class T>
__global__ __launch_bounds__(kT) void k_restrict(int64_t b, int64_t e, const uint32_t* ptr, const uint32_t* ridx, T* h,
                                                const int* ctl, int check) {
    if (check && ctl[0]) return;
    int64_t c = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= e) return;
    restrict_at<T>(c, ptr, ridx, h);
}

__global__ __launch_bounds__(kCoarseT) void k_prolong_coarse(Coarse C, const uint32_t* ptr, const uint32_t* idx,
                                                            float* p, const float* z, float* g, const double* scal,
                                                            const int* ctl, int upd) {
    if (upd && ctl[0]) return;
    __shared__ float sg[kCoarse];
    __shared__ uint32_t sptr[kCoarsePtr];
    prolong_coarse_body(C, ptr, idx, p, z, g, scal, upd, sg, sptr);
}
__global__ __launch_bounds__(kCoarseT) void k_restrict_coarse(Coarse C, const uint32_t* ptr, const uint32_t* ridx,
                                                             float* h, const int* ctl, int check) {
    if (check && ctl[0]) return;
    __shared__ float sh[kCoarse];
    __shared__ uint32_t sptr[kCoarsePtr];
    restrict_coarse_body(C, ptr, ridx, h, sh, sptr);
}

// z = M^-1 nr and the partials of nr.nr and nr.z
__global__ __launch_bounds__(kT) void k_precond(int64_t n, const float* nr, const float* invd, float* z, double* part,
                                               const int* ctl) {
    if (ctl[0]) return;
    __shared__ double sh[kT / 64];
    precond_block(n, blockIdx.x, gridDim.x, nr, invd, z, part, sh);
}
// rn2, absNew', the stopping test of Eigen's loop, beta
__global__ __launch_bounds__(kT) void k_scalar_beta(const double* part, unsigned np, double* scal, int* ctl, int init) {
    if (ctl[0]) return;
    __shared__ double sh[kT / 64];
    scalar_beta_body(part, np, scal, ctl, init, sh);
}

// binary64 helpers for the refinement loop
__global__ void k_apply_down64(int64_t n, const uint32_t* down, const double* g, double* t) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) t[i] = g[down[i]];
}
__global__ void k_dt64(int64_t n, const uint8_t* finest, const uint32_t* up, const double* r, double* h) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dt64_at(i, finest, up, r, h);
}
__global__ void k_sub64(int64_t n, const double* b, const double* a, double* r) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - a[i];
}
__global__ void k_to32(int64_t n, const double* a, float* b) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_axpy_f32_to64(int64_t n, const float* x, double* y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += (double)x[i];
}
__global__ __launch_bounds__(kT) void k_dot64(int64_t n, const double* a, const double* b, double* part) {
    __shared__ double sh[kT / 64];
    dot64_block(n, blockIdx.x, gridDim.x, a, b, part, sh);
}
__global__ __launch_bounds__(kT) void k_sum_partials(const double* part, unsigned np, double* out) {
    __shared__ double sh[kT / 64];
    double t = sum_partials(part, np, 1, 0, sh);
    if (threadIdx.x == 0) *out = t;
}
__global__ void k_set_thresh(double* scal, double rel2) { scal[S_THRESH] = rel2 * scal[S_RN2_0]; }

// ---------------------------------------------------------------- host side
unsigned red_blocks(int64_t n) { return fit_red_blocks(n); }

}  // namespace

Fit* fit_get(its_svo s) {
    if (s->fit) return s->fit.get();
    hipStream_t st = s->ctx->stream;
    auto f = std::unique_ptr<Fit>(new Fit());
    const int64_t n = s->num_verts;
    const int L = s->depth;
    f->n = n;
    f->L = L;
    for (int d = 0; d <= L; ++d) f->off[d] = s->esum_verts[d];
    // coarse levels for the one-workgroup kernels: as many top levels as fit
    f->lc = L;
    const int64_t coarse_max = std::min<int64_t>(kCoarse, tune_get(ITS_TUNE_COARSE_VERTS));
    while (f->lc > 0 && s->esum_verts[L] - s->esum_verts[f->lc - 1] <= coarse_max) --f->lc;
    f->coarse.L = L;
    f->coarse.lc = f->lc;
    for (int d = 0; d <= L; ++d) f->coarse.off[d] = (uint32_t)s->esum_verts[d];
    LatticeSet S{};
    for (int d = 0; d < L; ++d) {
        S.lat[d] = s->lv[d].lattice.p;
        S.D[d] = (uint32_t)s->lv[d].R + 1;
        S.off[d] = (uint32_t)s->esum_verts[d];
    }
    // ctl during the build: [0] parent entries, [1] child entries, [2] error
    // bits, read back together at the end, the one wait of the build
    f->ctl.alloc(4);
    ITS_HIP(hipMemsetAsync(f->ctl.p, 0, 4 * sizeof(int), st));
    int* err = f->ctl.p + 2;
    // prolongation lists: at most 8 parents per vertex; sized by that bound
    // first, cut to their length once it is known
    const int64_t cap = std::max<int64_t>(8 * n, 1);
    f->pro_ptr.alloc(n + 1);
    hipLaunchKernelGGL(k_pro_count, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, s->vlevel.p, L, s->vx.p, s->vy.p,
                       s->vz.p, f->pro_ptr.p);
    ITS_LAUNCH_CHECK();
    scan_exclusive_u32_async(f->pro_ptr.p, f->pro_ptr.p, n, st, f->pro_ptr.p + n);
    f->pro_idx.alloc(cap);
    hipLaunchKernelGGL(k_pro_fill, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, s->vlevel.p, L, s->vx.p, s->vy.p,
                       s->vz.p, S, f->pro_ptr.p, f->pro_idx.p, err);
    ITS_LAUNCH_CHECK();
    // restriction lists (transpose of prolongation)
    f->res_ptr.alloc(n + 1);
    ITS_HIP(hipMemsetAsync(f->res_ptr.p, 0, (n + 1) * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_res_count, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, f->pro_ptr.p, f->pro_idx.p,
                       f->res_ptr.p);
    ITS_LAUNCH_CHECK();
    scan_exclusive_u32_async(f->res_ptr.p, f->res_ptr.p, n, st, f->res_ptr.p + n);
    f->res_idx.alloc(cap);
    {
        DBuf<uint32_t> cursor(n + 1);
        ITS_HIP(hipMemcpyAsync(cursor.p, f->res_ptr.p, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_res_fill, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, f->pro_ptr.p, f->pro_idx.p,
                           cursor.p, f->res_idx.p);
        ITS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_res_sort, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, f->res_ptr.p, f->res_idx.p);
        ITS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_list_sizes, dim3(1), dim3(1), 0, st, (const uint32_t*)f->pro_ptr.p + n,
                       (const uint32_t*)f->res_ptr.p + n, f->ctl.p);
    ITS_LAUNCH_CHECK();
    // coincidence chains
    f->up.alloc(n);
    f->down.alloc(n);
    f->finest.alloc(n);
    {
        DBuf<uint32_t> dn1(n);
        hipLaunchKernelGGL(k_updown, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, s->vlevel.p, L, s->vx.p, s->vy.p,
                           s->vz.p, S, f->up.p, dn1.p);
        ITS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_down, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, dn1.p, f->down.p, f->finest.p);
        ITS_LAUNCH_CHECK();
        // preconditioner: inverse squared column norms
        DBuf<unsigned long long> acc(n);
        ITS_HIP(hipMemsetAsync(acc.p, 0, acc.bytes(), st));
        int lc = L;   // coarse levels whose columns fit the LDS accumulator
        while (lc > 0 && s->esum_verts[L] - s->esum_verts[lc - 1] <= kColLds) --lc;
        uint32_t coarse_off = (uint32_t)s->esum_verts[lc];
        uint32_t n_coarse = (uint32_t)(s->esum_verts[L] - s->esum_verts[lc]);
        unsigned nb = (unsigned)std::min<int64_t>(blocks_for(n, kT), 2048);
        hipLaunchKernelGGL(k_colnorm, dim3(nb), dim3(kT), 0, st, n, s->vlevel.p, L, s->vx.p, s->vy.p, s->vz.p, S,
                           dn1.p, acc.p, coarse_off, n_coarse, err);
        ITS_LAUNCH_CHECK();
        f->invd64.alloc(n);
        f->invd32.alloc(n);
        hipLaunchKernelGGL(k_invdiag, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, acc.p, f->invd64.p, f->invd32.p);
        ITS_LAUNCH_CHECK();
    }
    int info[3] = {0, 0, 0};
    ITS_HIP(hipMemcpyAsync(info, f->ctl.p, sizeof info, hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_REQUIRE(info[2] == 0, ITS_E_STATE,
                (info[2] & 4) ? "restriction size mismatch"
                              : "octree lattice is inconsistent (missing coarse corner vertex)");
    ITS_HIP(hipMemsetAsync(f->ctl.p, 0, 4 * sizeof(int), st));   // the solver's flags start cleared
    // the lists at their length (the bound-sized buffers go back to the stream's cache)
    for (auto* lst : {&f->pro_idx, &f->res_idx}) {
        DBuf<uint32_t> exact(std::max<int64_t>((uint32_t)info[0], 1));
        if (info[0] > 0)
            ITS_HIP(hipMemcpyAsync(exact.p, lst->p, (size_t)(uint32_t)info[0] * sizeof(uint32_t),
                                   hipMemcpyDeviceToDevice, st));
        std::swap(*lst, exact);
    }
    for (auto* v : {&f->x, &f->r, &f->p, &f->z, &f->t, &f->g, &f->h}) v->alloc(n);
    for (auto* v : {&f->g64, &f->h64, &f->r64, &f->nr64, &f->b64, &f->tmp64}) v->alloc(n);
    f->part.alloc(2 * kRedBlocks);
    f->scal.alloc(S_COUNT);
    ITS_HIP(hipMemsetAsync(f->scal.p, 0, S_COUNT * sizeof(double), st));
    s->fit = std::unique_ptr<Fit, void (*)(Fit*)>(f.release(), fit_release);
    return s->fit.get();
}

namespace {

// y = A x (binary64), via g64
void apply_A64(Fit* f, const double* x, double* y, hipStream_t st) {
    double* g = f->g64.p;
    for (int d = f->L - 1; d >= 0; --d) {
        int64_t b = f->off[d], e = f->off[d + 1];
        if (e <= b) continue;
        hipLaunchKernelGGL((k_prolong<double, false>), dim3(blocks_for(e - b, kT)), dim3(kT), 0, st, b, e, f->pro_ptr.p,
                           f->pro_idx.p, const_cast<double*>(x), (const double*)nullptr, g, (const double*)nullptr,
                           (const int*)nullptr);
        ITS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_apply_down64, dim3(blocks_for(f->n, kT)), dim3(kT), 0, st, f->n, f->down.p, g, y);
    ITS_LAUNCH_CHECK();
}
// y = A^T x (binary64), y must not alias x
void apply_AT64(Fit* f, const double* x, double* y, hipStream_t st) {
    hipLaunchKernelGGL(k_dt64, dim3(blocks_for(f->n, kT)), dim3(kT), 0, st, f->n, f->finest.p, f->up.p, x, y);
    ITS_LAUNCH_CHECK();
    for (int d = 1; d < f->L; ++d) {
        int64_t b = f->off[d], e = f->off[d + 1];
        if (e <= b) continue;
        hipLaunchKernelGGL((k_restrict<double>), dim3(blocks_for(e - b, kT)), dim3(kT), 0, st, b, e, f->res_ptr.p,
                           f->res_idx.p, y, (const int*)nullptr, 0);
        ITS_LAUNCH_CHECK();
    }
}
double dot64(Fit* f, const double* a, const double* b, hipStream_t st) {
    unsigned nb = red_blocks(f->n);
    hipLaunchKernelGGL(k_dot64, dim3(nb), dim3(kT), 0, st, f->n, a, b, f->part.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kT), 0, st, (const double*)f->part.p, nb, f->scal.p + S_ZZ);
    ITS_LAUNCH_CHECK();
    return read1(f->scal.p + S_ZZ, st);
}
// two dot products, read back together
void dot64_pair(Fit* f, const double* a1, const double* b1, const double* a2, const double* b2, hipStream_t st,
                double& r1, double& r2) {
    unsigned nb = red_blocks(f->n);
    hipLaunchKernelGGL(k_dot64, dim3(nb), dim3(kT), 0, st, f->n, a1, b1, f->part.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kT), 0, st, (const double*)f->part.p, nb, f->scal.p + S_ZZ);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dot64, dim3(nb), dim3(kT), 0, st, f->n, a2, b2, f->part.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kT), 0, st, (const double*)f->part.p, nb, f->scal.p + S_ZZ2);
    ITS_LAUNCH_CHECK();
    double v[2];
    ITS_HIP(hipMemcpyAsync(v, f->scal.p + S_ZZ, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    r1 = v[0];
    r2 = v[1];
}

// binary32 transposed pass on f->r into f->h (with optional fused update)
void launch_AT32(Fit* f, hipStream_t st, int do_update, int check) {
    hipLaunchKernelGGL(k_update_dt, dim3(blocks_for(f->n, kT)), dim3(kT), 0, st, f->n, f->finest.p, f->up.p, f->x.p,
                       f->p.p, f->r.p, f->t.p, f->h.p, f->scal.p, f->ctl.p, do_update);
    ITS_LAUNCH_CHECK();
    for (int d = 1; d < f->lc; ++d) {
        int64_t b = f->off[d], e = f->off[d + 1];
        if (e <= b) continue;
        hipLaunchKernelGGL((k_restrict<float>), dim3(blocks_for(e - b, kT)), dim3(kT), 0, st, b, e, f->res_ptr.p,
                           f->res_idx.p, f->h.p, f->ctl.p, check);
        ITS_LAUNCH_CHECK();
    }
    if (f->lc < f->L) {
        // level 0 has no restriction: a coarse range starting at 0 restricts from level 1 on
        Coarse C = f->coarse;
        if (C.lc == 0) C.lc = 1;
        if (C.lc < C.L) {
            hipLaunchKernelGGL(k_restrict_coarse, dim3(1), dim3(kCoarseT), 0, st, C, f->res_ptr.p, f->res_idx.p, f->h.p,
                               f->ctl.p, check);
            ITS_LAUNCH_CHECK();
        }
    }
}

// one CGLS iteration (all kernels early-out once ctl[0] is set)
void launch_iteration(Fit* f, hipStream_t st) {
    const int64_t n = f->n;
    const unsigned nb = red_blocks(n);
    if (f->lc < f->L) {
        hipLaunchKernelGGL(k_prolong_coarse, dim3(1), dim3(kCoarseT), 0, st, f->coarse, f->pro_ptr.p, f->pro_idx.p,
                           f->p.p, (const float*)f->z.p, f->g.p, (const double*)f->scal.p, (const int*)f->ctl.p, 1);
        ITS_LAUNCH_CHECK();
    }
    for (int d = f->lc - 1; d >= 0; --d) {
        int64_t b = f->off[d], e = f->off[d + 1];
        if (e <= b) continue;
        hipLaunchKernelGGL((k_prolong<float, true>), dim3(blocks_for(e - b, kT)), dim3(kT), 0, st, b, e, f->pro_ptr.p,
                           f->pro_idx.p, f->p.p, (const float*)f->z.p, f->g.p, (const double*)f->scal.p,
                           (const int*)f->ctl.p);
        ITS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_apply_down, dim3(nb), dim3(kT), 0, st, n, f->down.p, f->g.p, f->t.p, f->part.p, f->ctl.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scalar_alpha, dim3(1), dim3(kT), 0, st, (const double*)f->part.p, nb, f->scal.p, f->ctl.p);
    ITS_LAUNCH_CHECK();
    launch_AT32(f, st, 1, 1);
    hipLaunchKernelGGL(k_precond, dim3(nb), dim3(kT), 0, st, n, f->h.p, f->invd32.p, f->z.p, f->part.p, f->ctl.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scalar_beta, dim3(1), dim3(kT), 0, st, (const double*)f->part.p, nb, f->scal.p, f->ctl.p, 0);
    ITS_LAUNCH_CHECK();
}

constexpr int kChunk = 4;

// inner binary32 CGLS on rhs (device binary64); returns iterations used
int64_t inner_cgls(Fit* f, const double* rhs, double rel_tol, int64_t max_it, hipStream_t st) {
    const int64_t n = f->n;
    int ctl_init[4] = {0, 0, 0, (int)std::min<int64_t>(max_it, 1 << 30)};
    ITS_HIP(hipMemcpyAsync(f->ctl.p, ctl_init, sizeof ctl_init, hipMemcpyHostToDevice, st));
    ITS_HIP(hipMemsetAsync(f->x.p, 0, n * sizeof(float), st));
    ITS_HIP(hipMemsetAsync(f->p.p, 0, n * sizeof(float), st));
    hipLaunchKernelGGL(k_to32, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, rhs, f->r.p);
    ITS_LAUNCH_CHECK();
    launch_AT32(f, st, 0, 0);
    // z = M nr, absNew, rn2_0 (threshold not yet known: set +inf-free 0 first)
    ITS_HIP(hipMemsetAsync(f->scal.p + S_THRESH, 0, sizeof(double), st));
    hipLaunchKernelGGL(k_precond, dim3(red_blocks(n)), dim3(kT), 0, st, n, f->h.p, f->invd32.p, f->z.p, f->part.p,
                       f->ctl.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scalar_beta, dim3(1), dim3(kT), 0, st, (const double*)f->part.p, red_blocks(n), f->scal.p,
                       f->ctl.p, 1);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_set_thresh, dim3(1), dim3(1), 0, st, f->scal.p, rel_tol * rel_tol);
    ITS_LAUNCH_CHECK();
    // graph of kChunk iterations, captured once per octree
    if (!f->gexec) {
        ITS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kChunk; ++k) launch_iteration(f, st);
        ITS_HIP(hipStreamEndCapture(st, &f->graph));
        ITS_HIP(hipGraphInstantiate(&f->gexec, f->graph, nullptr, nullptr, 0));
    }
    int ctl[2] = {0, 0};
    int64_t launched = 0;
    while (launched < max_it + kChunk) {
        ITS_HIP(hipGraphLaunch(f->gexec, st));
        launched += kChunk;
        ITS_HIP(hipMemcpyAsync(ctl, f->ctl.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        ITS_HIP(hipStreamSynchronize(st));
        if (ctl[0]) break;
    }
    return (int64_t)ctl[1];
}
}  // namespace

// ---- warm start in the row space of A ----
//
// Jacobi-preconditioned CGLS keeps every iterate in x0 + range(D A^T), with
// D = diag(1/||A_:j||^2).  From x0 = 0 it converges to the one solution
// lambda* in range(D A^T); from any x0 = D A^T w0 it converges to the same
// lambda*, whatever w0 is.  A warm start of that form therefore changes the
// iteration count but not the fitted field.
//
// The dual coefficients of a solution lambda = D A^T w: with A = Dpick G and
// G = (I - P)^-1, A^T w = (I - P^T)^-1 Dpick^T w, so the coincidence-group
// sums W = Dpick^T w are W = (I - P^T)(lambda / D): one restriction pass
// without the cascade.  Only the group sums are determined; each vertex gets
// an equal share of its group's sum (the minimum-norm w).
namespace {
__global__ void k_dual_u(int64_t n, const double* lam, const double* invd, double* u) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = lam[i] / invd[i];
}
__global__ void k_dual_restrict(int64_t n, const uint32_t* ptr, const uint32_t* ridx, const double* u, double* W) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s = 0;
    for (uint32_t k = ptr[c]; k < ptr[c + 1]; ++k) {
        uint32_t v = ridx[k];
        s += u[v & 0x0FFFFFFFu] / (double)(1u << (v >> 28));
    }
    W[c] = u[c] - s;
}
__global__ void k_group_count(int64_t n, const uint32_t* down, uint32_t* cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[down[i]], 1u);
}
__global__ void k_group_split(int64_t n, const uint32_t* down, const double* W, const uint32_t* cnt, double* w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = down[i];
    w[i] = W[j] / (double)cnt[j];
}
__global__ void k_scale_invd(int64_t n, const double* invd, double* x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= invd[i];
}
}  // namespace

void warm_dual(its_svo s, const double* d_lam, double* d_w) {
    Fit* f = fit_get(s);
    hipStream_t st = s->ctx->stream;
    const int64_t n = f->n;
    DBuf<double> u(n), W(n);
    DBuf<uint32_t> cnt(n);
    ITS_HIP(hipMemsetAsync(cnt.p, 0, n * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_dual_u, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, d_lam, (const double*)f->invd64.p, u.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dual_restrict, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, (const uint32_t*)f->res_ptr.p,
                       (const uint32_t*)f->res_idx.p, (const double*)u.p, W.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_group_count, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, (const uint32_t*)f->down.p, cnt.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_group_split, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, (const uint32_t*)f->down.p,
                       (const double*)W.p, (const uint32_t*)cnt.p, d_w);
    ITS_LAUNCH_CHECK();
    ITS_HIP(hipStreamSynchronize(st));
}

// the Jacobi diagonal D = diag(1/||A_:j||^2) of the solver
void lscg_diag(its_svo s, double* d_out, hipStream_t st) {
    Fit* f = fit_get(s);
    ITS_HIP(hipMemcpyAsync(d_out, f->invd64.p, f->n * sizeof(double), hipMemcpyDeviceToDevice, st));
}

// lambda0 = D A^T w (binary64); d_lam0 must not alias d_w
void warm_apply(its_svo s, const double* d_w, double* d_lam0, hipStream_t st) {
    Fit* f = fit_get(s);
    apply_AT64(f, d_w, d_lam0, st);
    hipLaunchKernelGGL(k_scale_invd, dim3(blocks_for(f->n, kT)), dim3(kT), 0, st, f->n,
                       (const double*)f->invd64.p, d_lam0);
    ITS_LAUNCH_CHECK();
}

// Non-zeros of the collocation matrix the reference assembles
// (ThinShells.cpp:556-605): per row, the vertex itself, the finer vertices at
// the same position (a contiguous chain down to the finest coincident one),
// and at every coarser level the corners of the cell holding the vertex whose
// hat is non-zero there: one corner per axis where the vertex lies on that
// level's lattice plane, else two.
namespace {
__global__ void k_row_nnz(int64_t n, const uint8_t* lev, int L, const uint16_t* vx, const uint16_t* vy,
                          const uint16_t* vz, const uint32_t* down, unsigned long long* total) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c = 0;
    if (i < n) {
        const int d = lev[i];
        c = 1 + (unsigned long long)(d - lev[down[i]]);
        const uint32_t X = vx[i], Y = vy[i], Z = vz[i];
        for (int l = d + 1; l < L; ++l) {
            const uint32_t mask = (1u << (l - d)) - 1u;
            c += (unsigned long long)(((X & mask) ? 2 : 1) * ((Y & mask) ? 2 : 1) * ((Z & mask) ? 2 : 1));
        }
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, c);
}
}  // namespace

void lscg_structure(its_svo s, its_lscg_structure* out) {
    Fit* f = fit_get(s);
    hipStream_t st = s->ctx->stream;
    const int64_t n = f->n;
    DBuf<unsigned long long> tot(1);
    ITS_HIP(hipMemsetAsync(tot.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_row_nnz, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, s->vlevel.p, f->L, s->vx.p, s->vy.p,
                       s->vz.p, (const uint32_t*)f->down.p, tot.p);
    ITS_LAUNCH_CHECK();
    unsigned long long nnz = 0;
    ITS_HIP(hipMemcpyAsync(&nnz, tot.p, sizeof nnz, hipMemcpyDeviceToHost, st));
    uint32_t entries = 0;
    ITS_HIP(hipMemcpyAsync(&entries, f->pro_ptr.p + n, sizeof entries, hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    its_lscg_structure S{};
    S.n = n;
    S.nnz = (int64_t)nnz;
    S.parent_entries = (int64_t)entries;
    S.levels = f->L;
    S.block_levels = f->L - f->lc;
    S.bytes_assembled_per_iteration = 16.0 * (double)nnz + 72.0 * (double)n;
    // per vertex: prolongation p r/w, z, list offset, g (20 B); t = D g with
    // its index and gather (12 B); the update of x, r along the chains, h
    // (33 B); restriction offset and h r/w (12 B); preconditioner h, 1/d, z
    // (12 B).  Per parent entry: index and value, once down and once up (16 B).
    S.bytes_matrix_free_per_iteration = 89.0 * (double)n + 16.0 * (double)entries;
    *out = S;
}

void collocation_mul(its_svo s, const double* d_x, double* d_y, bool transpose) {
    Fit* f = fit_get(s);
    hipStream_t st = s->ctx->stream;
    if (transpose) apply_AT64(f, d_x, d_y, st);
    else apply_A64(f, d_x, d_y, st);
    ITS_HIP(hipStreamSynchronize(st));
}

void lscg_solve(its_svo s, const double* d_b, double* d_lambda, double tol, int64_t max_iter, bool warm,
                its_lscg_stats* out, hipStream_t stream) {
    auto t0 = std::chrono::steady_clock::now();
    Fit* f = fit_get(s);
    hipStream_t st = stream ? stream : s->ctx->stream;
    const int64_t n = f->n;
    if (max_iter <= 0) max_iter = 2 * n;
    if (!warm) ITS_HIP(hipMemsetAsync(d_lambda, 0, n * sizeof(double), st));
    // ||A^T b||^2
    apply_AT64(f, d_b, f->h64.p, st);
    double rhs2 = 0, bb = 0;
    dot64_pair(f, f->h64.p, f->h64.p, d_b, d_b, st, rhs2, bb);
    its_lscg_stats S{};
    int64_t iters = 0, outer = 0;
    double rel = 0, rel0 = -1;
    bool r_current = false;   // r64 = b - A lambda for the current lambda
    if (rhs2 == 0) {
        ITS_HIP(hipMemsetAsync(d_lambda, 0, n * sizeof(double), st));
    } else {
        const double inner_tol = 1e-3;
        for (;;) {
            apply_A64(f, d_lambda, f->tmp64.p, st);
            hipLaunchKernelGGL(k_sub64, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, d_b, f->tmp64.p, f->r64.p);
            ITS_LAUNCH_CHECK();
            r_current = true;
            apply_AT64(f, f->r64.p, f->nr64.p, st);
            double rn2 = dot64(f, f->nr64.p, f->nr64.p, st);
            rel = std::sqrt(rn2 / rhs2);
            if (rel0 < 0) rel0 = rel;
            if (rel < tol || iters >= max_iter) break;
            double want = std::max(inner_tol, 0.5 * tol / rel);
            if (want >= 1.0) want = 0.5;
            int64_t it = inner_cgls(f, f->r64.p, want, max_iter - iters, st);
            iters += it;
            hipLaunchKernelGGL(k_axpy_f32_to64, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, f->x.p, d_lambda);
            ITS_LAUNCH_CHECK();
            ++outer;
            // no inner iteration leaves x = 0, so lambda and r64 are unchanged
            r_current = (it == 0);
            if (it == 0) break;   // no progress possible
        }
    }
    // final residuals (the loop leaves r64 current except after a zero rhs)
    if (!r_current) {
        apply_A64(f, d_lambda, f->tmp64.p, st);
        hipLaunchKernelGGL(k_sub64, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, d_b, f->tmp64.p, f->r64.p);
        ITS_LAUNCH_CHECK();
    }
    double rr = dot64(f, f->r64.p, f->r64.p, st);
    ITS_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    S.iterations = iters;
    S.refinements = outer > 0 ? outer - 1 : 0;
    S.normal_rel = rel;
    S.initial_normal_rel = rel0 < 0 ? 0 : rel0;
    S.residual_abs = std::sqrt(rr);
    S.residual_rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    S.seconds = std::chrono::duration<double>(t1 - t0).count();
    S.converged = rel < tol;
    if (out) *out = S;
}

}  // namespace its
