// Per-vertex input features of the warm-start network (gfx950).
//
// No reference counterpart.  One row of 16 bfloat16 values per octree vertex
// v of level d, whose owner node o is a level-d cell with v as a corner:
//   0-2   centre of o, mapped from the grid box to [-1, 1]
//   3     d / 8
//   4-11  signed distance at the 8 corners of o divided by the cell width
//         w_d.  Corner k sits at v + (s_x a, s_y b, s_z c) with
//         (a, b, c) the bits of k and s the direction from v into o, so
//         feature 4 is always v itself.
//   12-14 unit normal of the triangle nearest to v, in the same reflected
//         frame (component a multiplied by s_a)
//   15    0
// The reflection makes the features independent of which corner of its owner
// the vertex is.  The per-vertex scale w_d converts the network's output
// (dual coefficient / w_d) back to the dual coefficient.
#include "internal.hpp"

namespace its {
namespace {
constexpr int kT = 256;

struct LevelLattices {
    const RankWord* lat[kMaxLevels];
    uint32_t D[kMaxLevels];
    uint32_t off[kMaxLevels];
    double width[kMaxLevels];
};

__device__ inline uint16_t bf16_bits(float v) {
    const __bf16 h = (__bf16)v;
    return __builtin_bit_cast(uint16_t, h);
}

__global__ void k_features(int64_t n, const uint8_t* vlevel, const uint16_t* vx, const uint16_t* vy,
                           const uint16_t* vz, const uint32_t* vowner, const uint32_t* morton,
                           const double* node_origin, const double* node_width, const double* rhs,
                           const float4* nrm, LevelLattices S, double ox, double oy, double oz, double box_w,
                           uint16_t* X, float* scale, int* err) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = vlevel[i];
    const int v[3] = {vx[i], vy[i], vz[i]};
    const uint32_t o = vowner[i];
    uint32_t c[3];
    demorton3(morton[o], c[0], c[1], c[2]);
    int sdir[3];
    for (int a = 0; a < 3; ++a) sdir[a] = (v[a] == (int)c[a]) ? 1 : -1;
    const double w = S.width[d];
    float f[16];
    const double half = 0.5 * node_width[o];
    f[0] = (float)(2.0 * (node_origin[3 * (int64_t)o] + half - ox) / box_w - 1.0);
    f[1] = (float)(2.0 * (node_origin[3 * (int64_t)o + 1] + half - oy) / box_w - 1.0);
    f[2] = (float)(2.0 * (node_origin[3 * (int64_t)o + 2] + half - oz) / box_w - 1.0);
    f[3] = (float)d * 0.125f;
    const uint32_t D = S.D[d];
    for (int k = 0; k < 8; ++k) {
        const int p0 = v[0] + sdir[0] * (k & 1), p1 = v[1] + sdir[1] * ((k >> 1) & 1),
                  p2 = v[2] + sdir[2] * ((k >> 2) & 1);
        float val = 0.f;
        if (p0 >= 0 && p1 >= 0 && p2 >= 0 && (uint32_t)p0 < D && (uint32_t)p1 < D && (uint32_t)p2 < D) {
            const uint32_t r = bit_find(S.lat[d], ((uint32_t)p0 * D + (uint32_t)p1) * D + (uint32_t)p2);
            if (r != ITS_NONE) val = (float)(rhs[S.off[d] + r] / w);
            else atomicOr(err, 1);
        } else {
            atomicOr(err, 1);
        }
        f[4 + k] = val;
    }
    const float4 nv = nrm[i];
    f[12] = nv.x * (float)sdir[0];
    f[13] = nv.y * (float)sdir[1];
    f[14] = nv.z * (float)sdir[2];
    f[15] = 0.f;
    uint16_t* row = X + 16 * i;
    for (int k = 0; k < 16; ++k) row[k] = bf16_bits(f[k]);
    scale[i] = (float)w;
}
}  // namespace

void warm_features(its_svo s, uint16_t* d_x, float* d_scale, hipStream_t st, int* d_err) {
    ITS_REQUIRE(s->has_rhs, ITS_E_STATE, "no right-hand side: call its_sdf_at_vertices first");
    const int64_t n = s->num_verts;
    if (n <= 0) return;
    LevelLattices S{};
    for (int d = 0; d < s->depth; ++d) {
        S.lat[d] = s->lv[d].lattice.p;
        S.D[d] = (uint32_t)s->lv[d].R + 1;
        S.off[d] = (uint32_t)s->esum_verts[d];
        S.width[d] = s->voxel_width * (double)(1u << d);
    }
    AllocStream as(st);   // the flag lives on the stream that uses it
    DBuf<int> err;
    if (!d_err) {
        err.alloc(1);
        d_err = err.p;
    }
    ITS_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_features, dim3(blocks_for(n, kT)), dim3(kT), 0, st, n, s->vlevel.p, s->vx.p, s->vy.p, s->vz.p,
                       s->vowner.p, s->morton.p, s->node_origin.p, s->node_width.p, s->rhs.p, s->rhs_nrm.p, S,
                       s->origin[0], s->origin[1], s->origin[2], s->width[0], d_x, d_scale, d_err);
    ITS_LAUNCH_CHECK();
    if (err.p) ITS_REQUIRE(read1(err.p, st) == 0, ITS_E_STATE, "owner corner is not a vertex of its level");
}

}  // namespace its
