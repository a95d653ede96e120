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
#include <algorithm>

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
    int L;
    int R;
    float ox, oy, oz, inv_w;
    double blo[3], bhi[3];   // accept box: open (SUPPORT) or closed (DENSE)
    int mode;
    float out_val;
};

struct SoaF32 {
    const float *x, *y, *z;
    __device__ void operator()(int64_t i, float& a, float& b, float& c, double& da, double& db, double& dc) const {
        a = x[i]; b = y[i]; c = z[i];
        da = a; db = b; dc = c;
    }
};
struct AosF64 {
    const double* p;
    __device__ void operator()(int64_t i, float& a, float& b, float& c, double& da, double& db, double& dc) const {
        da = p[3 * i]; db = p[3 * i + 1]; dc = p[3 * i + 2];
        a = (float)da; b = (float)db; c = (float)dc;
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
};

__device__ inline float hat0(float t) { return fmaxf(0.f, 1.f - fabsf(t)); }

template <class Loader>
__global__ __launch_bounds__(kT) void k_eval(Loader load, int64_t N, EvalParams P, float* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float x, y, z;
    double dx, dy, dz;
    load(i, x, y, z, dx, dy, dz);
    if (P.mode == ITS_EVAL_SUPPORT) {
        if (dx <= P.blo[0] || dy <= P.blo[1] || dz <= P.blo[2] || dx >= P.bhi[0] || dy >= P.bhi[1] || dz >= P.bhi[2]) {
            out[i] = P.out_val;
            return;
        }
    } else if (P.mode == ITS_EVAL_DENSE) {
        if (dx < P.blo[0] || dy < P.blo[1] || dz < P.blo[2] || dx > P.bhi[0] || dy > P.bhi[1] || dz > P.bhi[2]) {
            out[i] = P.out_val;
            return;
        }
    }
    const float ux = (x - P.ox) * P.inv_w, uy = (y - P.oy) * P.inv_w, uz = (z - P.oz) * P.inv_w;
    const int Rm = P.R - 1;
    const float fr = (float)Rm;   // clamp in float first: far or non-finite points stay defined
    const int cx = (int)fminf(fmaxf(floorf(ux), 0.f), fr), cy = (int)fminf(fmaxf(floorf(uy), 0.f), fr),
              cz = (int)fminf(fmaxf(floorf(uz), 0.f), fr);
    const uint32_t m0 = morton3(cx, cy, cz);
    float f = 0.f;
    for (int d = P.L - 1; d >= 0; --d) {
        const uint32_t m = m0 >> (3 * d);
        const RankWord w = P.cells[d][m >> 5];
        const uint32_t bit = 1u << (m & 31);
        if (!(w.bits & bit)) break;
        const uint32_t idx = w.rank + __popc(w.bits & (bit - 1u));
        const float4 b0 = P.blocks[d][2 * idx], b1 = P.blocks[d][2 * idx + 1];
        const float s = __uint_as_float((uint32_t)(127 - d) << 23);   // 2^-d
        const float tx = ux * s - (float)(cx >> d), ty = uy * s - (float)(cy >> d), tz = uz * s - (float)(cz >> d);
        const float wx0 = hat0(tx), wx1 = hat0(tx - 1.f);
        const float wy0 = hat0(ty), wy1 = hat0(ty - 1.f);
        const float wz0 = hat0(tz), wz1 = hat0(tz - 1.f);
        const float lo = fmaf(wy1, fmaf(wx1, b0.w, wx0 * b0.z), wy0 * fmaf(wx1, b0.y, wx0 * b0.x));
        const float hi = fmaf(wy1, fmaf(wx1, b1.w, wx0 * b1.z), wy0 * fmaf(wx1, b1.y, wx0 * b1.x));
        f = fmaf(wz1, hi, fmaf(wz0, lo, f));
    }
    out[i] = f;
}

EvalParams make_params(its_svo s, const double box_end[3], int mode, float out_val) {
    ITS_REQUIRE(s->field_ready, ITS_E_STATE, "no coefficient field: run a fit or set lambda first");
    EvalParams P{};
    for (int d = 0; d < s->depth; ++d) {
        P.cells[d] = s->lv[d].cells.p;
        P.blocks[d] = s->lv[d].blocks.p;
    }
    P.L = s->depth;
    P.R = s->R;
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
    }
    return P;
}
}  // namespace

void field_build(its_svo s) {
    ITS_REQUIRE(s->has_lambda, ITS_E_STATE, "no coefficients");
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
    ITS_HIP(hipStreamSynchronize(st));
    s->field_ready = true;
}

void eval_points(its_svo s, const float* x, const float* y, const float* z, int64_t N, const double box_end[3],
                 int mode, float out_val, float* f, hipStream_t st) {
    EvalParams P = make_params(s, box_end, mode, out_val);
    if (N <= 0) return;
    hipLaunchKernelGGL(k_eval<SoaF32>, dim3(blocks_for(N, kT)), dim3(kT), 0, st, SoaF32{x, y, z}, N, P, f);
    ITS_LAUNCH_CHECK();
}

void eval_points_aos64(its_svo s, const double* pts, int64_t N, const double box_end[3], int mode, float out_val,
                       float* f, hipStream_t st) {
    EvalParams P = make_params(s, box_end, mode, out_val);
    if (N <= 0) return;
    hipLaunchKernelGGL(k_eval<AosF64>, dim3(blocks_for(N, kT)), dim3(kT), 0, st, AosF64{pts}, N, P, f);
    ITS_LAUNCH_CHECK();
}

void eval_lattice(its_svo s, const double o[3], const double vs[3], const int64_t dims[3], float* out, hipStream_t st) {
    const double dummy[3] = {0, 0, 0};
    EvalParams P = make_params(s, dummy, ITS_EVAL_RAW, 0.f);
    int64_t N = dims[0] * dims[1] * dims[2];
    if (N <= 0) return;
    LatticeLoad ld{{o[0], o[1], o[2]}, {vs[0], vs[1], vs[2]}, dims[0], dims[1]};
    hipLaunchKernelGGL(k_eval<LatticeLoad>, dim3(blocks_for(N, kT)), dim3(kT), 0, st, ld, N, P, out);
    ITS_LAUNCH_CHECK();
}

}  // namespace its
