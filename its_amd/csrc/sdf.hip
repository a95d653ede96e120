// Signed distance at the octree vertices (right-hand side of the fit).
//
// Replaces ThinShells::cpSDFOfTreeNodes (reference ThinShells.cpp:481-509).
// The reference calls igl::signed_distance with the fast winding number sign,
// which is third-party code absent here.  Both this kernel and the CPU oracle
// use the same definition:
//   - |b| is the exact Euclidean distance to the nearest triangle (closest
//     point on a triangle, Ericson, Real-Time Collision Detection 5.1.5),
//     in binary64.
//   - The sign is negative where the generalised winding number exceeds 1/2.
//     The winding number is the sum of triangle solid angles over 4 pi (Van
//     Oosterom and Strackee 1983), accumulated in binary32.  Its only use is
//     the comparison with 1/2, and away from the surface the winding number is
//     within 1e-4 of 0 or 1.
// Brute force over triangles: each 256-thread block stages 256 triangles at
// a time in LDS, and every thread tests its point against the staged tile.
// The winding number needs every triangle; the binary64 closest point is
// computed only for triangles whose bounding sphere can beat the distance
// found so far (see k_sdf), which leaves the result unchanged.
#include <algorithm>

#include "internal.hpp"

namespace its {
namespace {
constexpr int kT = 256;

__device__ inline double d3dot(const double* a, const double* b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }

__device__ double point_tri_dist2(const double p[3], const double a[3], const double b[3], const double c[3]) {
    double ab[3], ac[3], ap[3], q[3];
    for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i]; }
    double d1 = d3dot(ab, ap), d2 = d3dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        for (int i = 0; i < 3; ++i) q[i] = a[i];
    } else {
        double bp[3];
        for (int i = 0; i < 3; ++i) bp[i] = p[i] - b[i];
        double d3 = d3dot(ab, bp), d4 = d3dot(ac, bp);
        double cp[3];
        for (int i = 0; i < 3; ++i) cp[i] = p[i] - c[i];
        double d5 = d3dot(ab, cp), d6 = d3dot(ac, cp);
        double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
        if (d3 >= 0.0 && d4 <= d3) {
            for (int i = 0; i < 3; ++i) q[i] = b[i];
        } else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
            double v = d1 / (d1 - d3);
            for (int i = 0; i < 3; ++i) q[i] = a[i] + v * ab[i];
        } else if (d6 >= 0.0 && d5 <= d6) {
            for (int i = 0; i < 3; ++i) q[i] = c[i];
        } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
            double w = d2 / (d2 - d6);
            for (int i = 0; i < 3; ++i) q[i] = a[i] + w * ac[i];
        } else if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
            double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            for (int i = 0; i < 3; ++i) q[i] = b[i] + w * (c[i] - b[i]);
        } else {
            double denom = 1.0 / (va + vb + vc);
            double v = vb * denom, w = vc * denom;
            for (int i = 0; i < 3; ++i) q[i] = a[i] + ab[i] * v + ac[i] * w;
        }
    }
    double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return d3dot(d, d);
}

__device__ inline float solid_angle_f(const float p[3], const double* A, const double* B, const double* C) {
    float a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) {
        a[i] = (float)(A[i] - (double)p[i]);
        b[i] = (float)(B[i] - (double)p[i]);
        c[i] = (float)(C[i] - (double)p[i]);
    }
    float la = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    float lb = sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    float lc = sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    float det = a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
    float ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    float bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    float ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
    float den = la * lb * lc + ab * lc + bc * la + ca * lb;
    return 2.0f * atan2f(det, den);
}

// Per-triangle bounding sphere in binary32: centre = vertex mean, radius =
// largest centre-vertex distance, both rounded outward by a margin that
// covers the binary32 rounding of every quantity derived from them.
__device__ inline float bound_margin(float scale) { return 1e-5f * scale + 1e-30f; }

__global__ void k_tri_spheres(const double* tris, int64_t nF, float4* sph) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nF) return;
    const double* T = tris + 9 * t;
    double c[3];
    for (int a = 0; a < 3; ++a) c[a] = (T[a] + T[3 + a] + T[6 + a]) / 3.0;
    double r2 = 0, cm = 0;
    for (int v = 0; v < 3; ++v) {
        double d2 = 0;
        for (int a = 0; a < 3; ++a) {
            const double d = T[3 * v + a] - c[a];
            d2 += d * d;
        }
        r2 = fmax(r2, d2);
    }
    for (int a = 0; a < 3; ++a) cm = fmax(cm, fabs(c[a]));
    const double r = sqrt(r2);
    sph[t] = make_float4((float)c[0], (float)c[1], (float)c[2], (float)(r * (1.0 + 1e-6) + 1e-6 * (cm + r)));
}

// Nearest triangle in binary64 (Ericson's closest point) and the winding
// number (binary32 solid angles), brute force over LDS tiles of triangles.
// A first pass over the triangles' bounding spheres gives an upper bound on
// the point's distance; the second pass computes the exact binary64 distance
// only for triangles whose sphere can come closer than the running best.  A
// triangle it skips is strictly farther than the nearest one, so the minimum
// and the index of its first occurrence are those of the full search.
__global__ __launch_bounds__(kT) void k_sdf(const double* tris, const float4* sph, int64_t nF, const double* pts,
                                           int64_t N, double* out, int32_t* tri_out) {
    __shared__ double sh[kT * 9];
    __shared__ float4 ss[kT];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double p[3] = {0, 0, 0};
    if (i < N) { p[0] = pts[3 * i]; p[1] = pts[3 * i + 1]; p[2] = pts[3 * i + 2]; }
    float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
    const float pm = fmaxf(fabsf(pf[0]), fmaxf(fabsf(pf[1]), fabsf(pf[2])));
    // pass 1: upper bound of the distance from the bounding spheres
    float ub = 3.0e38f;
    for (int64_t base = 0; base < nF; base += kT) {
        __syncthreads();
        int64_t t = base + threadIdx.x;
        if (t < nF) ss[threadIdx.x] = sph[t];
        __syncthreads();
        const int cnt = (int)std::min<int64_t>(kT, nF - base);
        if (i < N)
            for (int k = 0; k < cnt; ++k) {
                const float4 S = ss[k];
                const float dx = pf[0] - S.x, dy = pf[1] - S.y, dz = pf[2] - S.z;
                const float d = sqrtf(dx * dx + dy * dy + dz * dz);
                const float m = bound_margin(pm + fmaxf(fabsf(S.x), fmaxf(fabsf(S.y), fabsf(S.z))) + S.w);
                ub = fminf(ub, d + S.w + m);
            }
    }
    double best = (ub < 3.0e38f) ? (double)ub * (double)ub * (1.0 + 1e-6) : 1e300;
    int64_t best_t = -1;
    float wsum = 0.f;
    // pass 2: exact distance where it can matter, and the winding number
    for (int64_t base = 0; base < nF; base += kT) {
        __syncthreads();
        int64_t t = base + threadIdx.x;
        if (t < nF) {
            for (int k = 0; k < 9; ++k) sh[threadIdx.x * 9 + k] = tris[9 * t + k];
            ss[threadIdx.x] = sph[t];
        }
        __syncthreads();
        int cnt = (int)std::min<int64_t>(kT, nF - base);
        if (i < N) {
            for (int k = 0; k < cnt; ++k) {
                const double* T = &sh[9 * k];
                const float4 S = ss[k];
                const float dx = pf[0] - S.x, dy = pf[1] - S.y, dz = pf[2] - S.z;
                const float m = bound_margin(pm + fmaxf(fabsf(S.x), fmaxf(fabsf(S.y), fabsf(S.z))) + S.w);
                const float lb = sqrtf(dx * dx + dy * dy + dz * dz) - S.w - m;
                if (!(lb > 0.f && (double)lb * (double)lb > best)) {
                    const double d2 = point_tri_dist2(p, T, T + 3, T + 6);
                    if (d2 < best) { best = d2; best_t = base + k; }
                }
                wsum += solid_angle_f(pf, T, T + 3, T + 6);
            }
        }
    }
    if (i < N) {
        double d = sqrt(best);
        float w = wsum * (float)(0.25 / M_PI);
        out[i] = (w > 0.5f) ? -d : d;
        if (tri_out) tri_out[i] = (int32_t)best_t;
    }
}

__global__ void k_gather_tris(const double* V, const int32_t* F, int64_t nF, double* tris) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nF) return;
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) tris[9 * t + 3 * c + a] = V[3 * (int64_t)F[3 * t + c] + a];
}
}  // namespace

void signed_distance(hipStream_t st, const double* d_V, int64_t nV, const int32_t* d_F, int64_t nF,
                     const double* d_pts, int64_t N, double* d_out, int32_t* d_tri) {
    (void)nV;
    if (N <= 0) return;
    DBuf<double> tris(std::max<int64_t>(9 * nF, 9));
    DBuf<float4> sph(std::max<int64_t>(nF, 1));
    if (nF > 0) {
        hipLaunchKernelGGL(k_gather_tris, dim3(blocks_for(nF, kT)), dim3(kT), 0, st, d_V, d_F, nF, tris.p);
        ITS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_tri_spheres, dim3(blocks_for(nF, kT)), dim3(kT), 0, st, (const double*)tris.p, nF, sph.p);
        ITS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_sdf, dim3(blocks_for(N, kT)), dim3(kT), 0, st, tris.p, (const float4*)sph.p, nF, d_pts, N,
                       d_out, d_tri);
    ITS_LAUNCH_CHECK();
    ITS_HIP(hipStreamSynchronize(st));
}

namespace {
// unit normal of each point's nearest triangle (zero for a degenerate one)
__global__ void k_tri_normal(int64_t N, const double* V, const int32_t* F, const int32_t* tri, float4* nrm) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int32_t t = tri[i];
    const double* a = V + 3 * (int64_t)F[3 * (int64_t)t];
    const double* b = V + 3 * (int64_t)F[3 * (int64_t)t + 1];
    const double* c = V + 3 * (int64_t)F[3 * (int64_t)t + 2];
    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
    const double l = sqrt(x * x + y * y + z * z);
    nrm[i] = l > 0 ? make_float4((float)(x / l), (float)(y / l), (float)(z / l), 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}
}  // namespace

void nearest_normals(hipStream_t st, const double* d_V, const int32_t* d_F, const int32_t* d_tri, int64_t N,
                     float4* d_nrm) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_tri_normal, dim3(blocks_for(N, kT)), dim3(kT), 0, st, N, d_V, d_F, d_tri, d_nrm);
    ITS_LAUNCH_CHECK();
}

}  // namespace its
