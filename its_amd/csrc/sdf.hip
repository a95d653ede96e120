// Signed distance at the octree vertices (right-hand side of the fit).
//
// Replaces ThinShells::cpSDFOfTreeNodes (reference ThinShells.cpp:481-509).
// The reference calls igl::signed_distance with the fast winding number sign,
// which is third-party code absent here.  Both this file and the CPU oracle
// use the same definition:
//   - |b| is the exact Euclidean distance to the nearest triangle (closest
//     point on a triangle, Ericson, Real-Time Collision Detection 5.1.5),
//     in binary64.
//   - The sign is negative where the generalised winding number exceeds 1/2.
//     The winding number is the sum of triangle solid angles over 4 pi (Van
//     Oosterom and Strackee 1983), each solid angle in binary32 after a
//     binary64 subtraction of the query point.
//
// Triangle tree.  Triangles are sorted by the Morton code of their centroid
// and cut into buckets of kBucket; an implicit binary tree sits over the
// buckets (node k of level h covers sorted triangles [k*8*2^h, (k+1)*8*2^h)),
// with a binary32 box per node rounded outward, so it holds every vertex
// exactly.  Levels are built one launch each, without atomics.
//   - Distance: depth-first walk that skips a node whose box is farther than
//     the best binary64 distance so far, started from the bucket at the
//     query's own Morton position.  A skipped node holds no triangle whose
//     distance could equal or beat the best, so the minimum and the lowest
//     index among equal minima are those of the full search.
//   - Winding number, exactly (generalised winding numbers, Jacobson,
//     Kavan and Sorkine-Hornung 2013, section 3.2): for a node with triangle
//     set S and a query p outside the node's box, close S with the fan C of
//     its boundary edges around the box centre; S + C is a closed chain
//     inside the convex box, so its winding number at p is 0 and
//     w_S(p) = -w_C(p).  The boundary of S is the signed sum of its half
//     edges (a -> b counts +1 for the min -> max direction, -1 otherwise),
//     which is exact for open, non-manifold or inconsistently oriented
//     meshes too.  A node's fan is stored only when it has at most as many
//     edges as the node has triangles; otherwise the walk descends.
//   - Winding number of a closed mesh, by a ray: when every undirected edge
//     has exactly two half edges of opposite directions, the mesh is an
//     oriented closed surface, and its winding number at p is the signed
//     number of crossings of any ray from p (self-intersecting and nested
//     surfaces included).  Each point takes the axis direction (+-x, +-y,
//     +-z) that leaves the mesh's box soonest; the ray is tested against
//     the triangles whose boxes it meets: p is inside the projection of a
//     triangle across the ray when its three edge orientations agree, and
//     the crossing lies ahead of p when the orientation of the tetrahedron
//     (a, b, c, p) has the sign of the triangle's normal along the ray.  Each
//     orientation is a binary64 determinant whose sign is taken only when it
//     exceeds the rounding bound of its terms (the static filters of
//     Shewchuk, Adaptive Precision Floating-Point Arithmetic and Fast Robust
//     Geometric Predicates, 1997); a point with any undecided orientation
//     (on or within rounding of an edge line or a triangle plane) takes the
//     fan walk above.

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace its {
namespace {
constexpr int kT = 256;
constexpr int kBucket = 8;
// query walks treat nodes of this level (at most 64 triangles) as leaves
constexpr int kLeafLevel = 3;
static_assert((kBucket << kLeafLevel) == 64, "one triangle per lane in a leaf");
constexpr int kMaxLevels = 32;

// ------------------------------------------------------------ geometry ----
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

// atan2 for the solid-angle sums: octant reduction to a = min/max in
// [0, 1] and atan(a) = a P(a^2), P of degree 7 fitted here by weighted least
// squares on Chebyshev nodes (Lawson iterations toward the minimax error).
// Absolute error below 3.2e-7 over all quadrants in binary32, signed zeros
// and the +-pi cut as atan2.
__device__ inline float atan2_fit(float y, float x) {
#pragma clang fp contract(fast)
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 1e-30f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
    const float q = a * a;
    float r = -0.004054564982652664f;
    r = r * q + 0.021862955763936043f;
    r = r * q - 0.055912334471940994f;
    r = r * q + 0.0964219868183136f;
    r = r * q - 0.1390863060951233f;
    r = r * q + 0.19946566224098206f;
    r = r * q - 0.33329859375953674f;
    r = r * q + 0.9999993443489075f;
    r = r * a;
    if (ay > ax) r = 1.5707963267948966f - r;
    if (x < 0.f) r = 3.141592653589793f - r;
    return copysignf(r, y);
}

// solid angle of triangle ABC seen from p (signed by orientation): binary64
// differences, the rest in binary32 with fused multiply-adds
__device__ inline float solid_angle(const double p[3], const double* A, const double* B, const double* C) {
#pragma clang fp contract(fast)
    float a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) {
        a[i] = (float)(A[i] - p[i]);
        b[i] = (float)(B[i] - p[i]);
        c[i] = (float)(C[i] - p[i]);
    }
    float la = __builtin_amdgcn_sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    float lb = __builtin_amdgcn_sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    float lc = __builtin_amdgcn_sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    float det = a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
    float ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    float bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    float ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
    float den = la * lb * lc + ab * lc + bc * la + ca * lb;
    return 2.0f * atan2_fit(det, den);
}

// apex of a node's fan: the centre of its binary32 box, in binary64
__host__ __device__ inline double fan_apex(float lo, float hi) { return 0.5 * ((double)lo + (double)hi); }

// solid angle of the fan triangle (apex, a, b) seen from p, from the
// binary32 vectors z = apex - p (|z| = lz), u = a - p, v = b - p: the same
// formula as solid_angle with its first vertex fixed
__device__ inline float fan_angle(const float z[3], float lz, const float u[3], const float v[3]) {
#pragma clang fp contract(fast)
    const float lu = __builtin_amdgcn_sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const float lv = __builtin_amdgcn_sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const float det = z[0] * (u[1] * v[2] - u[2] * v[1]) + z[1] * (u[2] * v[0] - u[0] * v[2]) +
                      z[2] * (u[0] * v[1] - u[1] * v[0]);
    const float zu = z[0] * u[0] + z[1] * u[1] + z[2] * u[2];
    const float uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    const float vz = v[0] * z[0] + v[1] * z[1] + v[2] * z[2];
    const float den = lz * lu * lv + zu * lv + uv * lz + vz * lu;
    return 2.0f * atan2_fit(det, den);
}

// binary32 neighbours of a binary64 value, rounded down / up
__device__ inline float f_down(double x) {
    float f = (float)x;
    return ((double)f > x) ? nextafterf(f, -INFINITY) : f;
}
__device__ inline float f_up(double x) {
    float f = (float)x;
    return ((double)f < x) ? nextafterf(f, INFINITY) : f;
}

// Bounding sphere of a triangle in binary32: centre = vertex mean, radius =
// largest centre-vertex distance, both widened by a margin that covers the
// binary32 rounding of every quantity derived from them.
__device__ inline float bound_margin(float scale) { return 1e-5f * scale + 1e-30f; }

__device__ inline float4 tri_sphere(const double* T) {
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
    return make_float4((float)c[0], (float)c[1], (float)c[2], (float)(r * (1.0 + 1e-6) + 1e-6 * (cm + r)));
}

// ---------------------------------------------------------------- tree ----
struct Shape {
    int32_t T;                  // triangles
    int32_t H;                  // root level; level 0 holds the buckets
    int32_t count[kMaxLevels];  // nodes per level
    int32_t off[kMaxLevels];    // first node of each level
};

__host__ __device__ inline int64_t node_first(int h, int64_t k) { return k * ((int64_t)kBucket << h); }
__host__ __device__ inline int64_t node_last(const Shape& s, int h, int64_t k) {
    return std::min<int64_t>(s.T, (k + 1) * ((int64_t)kBucket << h));
}

// next node in depth-first order after the subtree at (h, k); false at the end
__device__ inline bool tree_next(const Shape& s, int& h, int& k) {
    while (h < s.H) {
        if (!(k & 1) && k + 1 < s.count[h]) {
            ++k;
            return true;
        }
        k >>= 1;
        ++h;
    }
    return false;
}

struct Quant {
    double lo[3], scale[3];
};

__device__ inline uint32_t quant_code(const Quant& q, const double x[3]) {
    uint32_t u[3];
    for (int a = 0; a < 3; ++a) {
        double v = (x[a] - q.lo[a]) * q.scale[a];
        v = fmin(fmax(v, 0.0), 1023.0);
        u[a] = (uint32_t)v;
    }
    return morton3(u[0], u[1], u[2]);
}

struct TreeView {
    Shape s;
    Quant q;
    const double* tris;      // 9 per sorted triangle
    const float4* sph;       // bounding sphere per sorted triangle
    const int32_t* tid;      // original index per sorted triangle
    const uint32_t* code;    // centroid Morton code per sorted triangle (ascending)
    const float4* node;      // 2 per node: box lo.xyz + first fan edge (int bits in w),
                             // box hi.xyz + fan edge count (int bits in w; -1: no fan)
    const float* fan;        // 6 per fan edge: a - centre, b - centre (centre of its node's box)
    const uint32_t* cell_lb; // per cell_bits-bit code prefix C: first sorted triangle with code >= C's range
    int cell_bits;
};

// -------------------------------------------------------------- build -----
__global__ void k_centroid_box(const double* V, const int32_t* F, int64_t nF, double* part) {
    __shared__ double sh[6][kT];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nF; t += (int64_t)gridDim.x * blockDim.x) {
        const double* a = V + 3 * (int64_t)F[3 * t];
        const double* b = V + 3 * (int64_t)F[3 * t + 1];
        const double* c = V + 3 * (int64_t)F[3 * t + 2];
        for (int x = 0; x < 3; ++x) {
            const double m = (a[x] + b[x] + c[x]) / 3.0;
            lo[x] = fmin(lo[x], m);
            hi[x] = fmax(hi[x], m);
        }
    }
    for (int x = 0; x < 3; ++x) { sh[x][threadIdx.x] = lo[x]; sh[3 + x][threadIdx.x] = hi[x]; }
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int x = 0; x < 3; ++x) {
                sh[x][threadIdx.x] = fmin(sh[x][threadIdx.x], sh[x][threadIdx.x + o]);
                sh[3 + x][threadIdx.x] = fmax(sh[3 + x][threadIdx.x], sh[3 + x][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// centroid code of every triangle, with its index as the sort's value
__global__ void k_tri_keys(const double* V, const int32_t* F, int64_t nF, Quant q, uint32_t* code, uint32_t* idx) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nF) return;
    const double* a = V + 3 * (int64_t)F[3 * t];
    const double* b = V + 3 * (int64_t)F[3 * t + 1];
    const double* c = V + 3 * (int64_t)F[3 * t + 2];
    double m[3];
    for (int x = 0; x < 3; ++x) m[x] = (a[x] + b[x] + c[x]) / 3.0;
    code[t] = quant_code(q, m);
    idx[t] = (uint32_t)t;
}

__global__ void k_tri_gather(const uint32_t* idx, int64_t nF, const double* V, const int32_t* F, double* tris,
                             float4* sph, int32_t* tid, int32_t* pos) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nF) return;
    const int64_t t = (int64_t)idx[i];
    double T[9];
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) T[3 * c + a] = V[3 * (int64_t)F[3 * t + c] + a];
    for (int j = 0; j < 9; ++j) tris[9 * i + j] = T[j];
    sph[i] = tri_sphere(T);
    tid[i] = (int32_t)t;
    pos[t] = (int32_t)i;
}

__global__ void k_bucket_box(const double* tris, Shape s, float4* node) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.count[0]) return;
    double l[3] = {INFINITY, INFINITY, INFINITY}, u[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = node_first(0, k); i < node_last(s, 0, k); ++i)
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) {
                l[a] = fmin(l[a], tris[9 * i + 3 * v + a]);
                u[a] = fmax(u[a], tris[9 * i + 3 * v + a]);
            }
    node[2 * (s.off[0] + k)] = make_float4(f_down(l[0]), f_down(l[1]), f_down(l[2]), 0.f);
    node[2 * (s.off[0] + k) + 1] = make_float4(f_up(u[0]), f_up(u[1]), f_up(u[2]), 0.f);
}

__global__ void k_level_box(Shape s, int h, float4* node) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.count[h]) return;
    const int64_t c0 = s.off[h - 1] + 2 * k;
    float4 l = node[2 * c0], u = node[2 * c0 + 1];
    if (2 * k + 1 < s.count[h - 1]) {
        const float4 l1 = node[2 * c0 + 2], u1 = node[2 * c0 + 3];
        l = make_float4(fminf(l.x, l1.x), fminf(l.y, l1.y), fminf(l.z, l1.z), 0.f);
        u = make_float4(fmaxf(u.x, u1.x), fmaxf(u.y, u1.y), fmaxf(u.z, u1.z), 0.f);
    }
    node[2 * (s.off[h] + k)] = l;
    node[2 * (s.off[h] + k) + 1] = u;
}

// Half edges keyed by their undirected edge (min << vb | max); the value is
// the sorted position of the triangle times 2, plus 1 for a max -> min edge.
// Degenerate edges (a == b) close nothing and get the key that sorts last.
__global__ void k_half_edges(const int32_t* F, int64_t nF, const int32_t* pos, int vb, uint64_t none, uint64_t* keys,
                             uint32_t* vals) {
    int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= 3 * nF) return;
    const int64_t t = h / 3;
    const int e = (int)(h - 3 * t);
    const uint32_t a = (uint32_t)F[3 * t + e], b = (uint32_t)F[3 * t + (e + 1) % 3];
    keys[h] = (a == b) ? none : (((uint64_t)std::min(a, b) << vb) | std::max(a, b));
    vals[h] = 2u * (uint32_t)pos[t] + (a > b ? 1u : 0u);
}

// Walks the boundary contributions of one undirected edge (the run of equal
// keys starting at head i): for every level h >= 1 and every node holding
// some of its half edges, coef = signed count of those half edges.  Calls
// sink(node, h, k, coef) for coef != 0.  Above the level where one node
// holds all half edges, coef stays at their total; the walk stops there when
// the total is 0 (an interior edge).
template <class Sink>
__device__ void edge_walk(const uint64_t* keys, const uint32_t* vals, int64_t i, int64_t M, const Shape& s, Sink&& sink) {
    int64_t end = i + 1;
    while (end < M && keys[end] == keys[i]) ++end;
    static_assert((kBucket & (kBucket - 1)) == 0, "node spans are powers of two");
    const int bucket_shift = __builtin_ctz(kBucket);
    for (int h = 1; h <= s.H; ++h) {
        // node of level h holding sorted triangle t: t / (kBucket << h)
        const int shift = bucket_shift + h;
        const uint32_t k0 = (vals[i] >> 1) >> shift;
        bool same = true;
        int tot = 0;
        for (int64_t j = i; j < end; ++j) {
            const uint32_t v = vals[j];
            if (((v >> 1) >> shift) != k0) same = false;
            tot += (v & 1u) ? -1 : 1;
        }
        if (same) {
            if (tot == 0) return;
            sink(s.off[h] + (int64_t)k0, h, (int64_t)k0, tot);
            continue;
        }
        for (int64_t j = i; j < end; ++j) {
            const uint32_t kj = (vals[j] >> 1) >> shift;
            bool seen = false;
            for (int64_t jj = i; jj < j && !seen; ++jj) seen = (((vals[jj] >> 1) >> shift) == kj);
            if (seen) continue;
            int coef = 0;
            for (int64_t jj = j; jj < end; ++jj)
                if (((vals[jj] >> 1) >> shift) == kj) coef += (vals[jj] & 1u) ? -1 : 1;
            if (coef) sink(s.off[h] + (int64_t)kj, h, (int64_t)kj, coef);
        }
    }
}

__device__ inline bool edge_head(const uint64_t* keys, int64_t i, uint64_t none) {
    return keys[i] != none && (i == 0 || keys[i] != keys[i - 1]);
}

// *open = 1 unless every undirected edge has exactly two half edges of
// opposite directions (a degenerate edge also opens the mesh)
__global__ void k_closed_check(const uint64_t* keys, const uint32_t* vals, int64_t M, uint64_t none, int* open) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    if (keys[i] == none) {
        *open = 1;
        return;
    }
    if (!edge_head(keys, i, none)) return;
    const bool pair = i + 1 < M && keys[i + 1] == keys[i] && (i + 2 >= M || keys[i + 2] != keys[i]);
    if (!pair || ((vals[i] ^ vals[i + 1]) & 1u) == 0u) *open = 1;
}

__global__ void k_fan_node_count(const uint64_t* keys, const uint32_t* vals, int64_t M, Shape s, uint64_t none,
                                 uint32_t* node_cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M || !edge_head(keys, i, none)) return;
    edge_walk(keys, vals, i, M, s, [&](int64_t node, int, int64_t, int coef) {
        atomicAdd(&node_cnt[node], (uint32_t)abs(coef));
    });
}

// keep[node] != 0: the winding walk uses the node's fan for points outside
// its box (chosen by fan_choice on the host)
__device__ inline bool fan_kept(const uint8_t* keep, int64_t node) { return keep[node] != 0; }

__global__ void k_fan_edge_count(const uint64_t* keys, const uint32_t* vals, int64_t M, Shape s, uint64_t none,
                                 const uint8_t* keep, uint32_t* cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    uint32_t c = 0;
    if (edge_head(keys, i, none))
        edge_walk(keys, vals, i, M, s, [&](int64_t node, int, int64_t, int coef) {
            if (fan_kept(keep, node)) c += (uint32_t)abs(coef);
        });
    cnt[i] = c;
}

__global__ void k_fan_emit(const uint64_t* keys, const uint32_t* vals, int64_t M, Shape s, uint64_t none, int vb,
                           const uint8_t* keep, const uint32_t* off, uint32_t* fan_node, uint64_t* fan_edge) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M || !edge_head(keys, i, none)) return;
    const uint64_t lo = keys[i] >> vb, hi = keys[i] & ((1ull << vb) - 1);
    uint32_t o = off[i];
    edge_walk(keys, vals, i, M, s, [&](int64_t node, int, int64_t, int coef) {
        if (!fan_kept(keep, node)) return;
        const uint64_t e = coef > 0 ? ((lo << 32) | hi) : ((hi << 32) | lo);
        for (int r = 0; r < abs(coef); ++r) {
            fan_node[o] = (uint32_t)node;
            fan_edge[o] = e;
            ++o;
        }
    });
}

__global__ void k_fan_table(int64_t nodes, const uint32_t* node_cnt, const uint8_t* keep, const uint32_t* fan_off,
                            float4* node) {
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    const bool kept = fan_kept(keep, n);
    node[2 * n].w = __int_as_float((int)fan_off[n]);
    node[2 * n + 1].w = __int_as_float(kept ? (int)node_cnt[n] : -1);
}

// fan edge ends relative to the centre of their node's box (the fan apex),
// rounded once to binary32
__global__ void k_fan_coords(const uint64_t* fan_edge, const uint32_t* fan_node, int64_t E, const double* V,
                             const float4* node, float* fan) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint64_t ab = fan_edge[e];
    const int64_t a = (int64_t)(ab >> 32), b = (int64_t)(ab & 0xffffffffull);
    const int64_t n = fan_node[e];
    const double c[3] = {fan_apex(node[2 * n].x, node[2 * n + 1].x), fan_apex(node[2 * n].y, node[2 * n + 1].y),
                         fan_apex(node[2 * n].z, node[2 * n + 1].z)};
    for (int x = 0; x < 3; ++x) {
        fan[6 * e + x] = (float)(V[3 * a + x] - c[x]);
        fan[6 * e + 3 + x] = (float)(V[3 * b + x] - c[x]);
    }
}

// kept fan edges per node (0 for dropped fans and buckets), for the scan
__global__ void k_fan_kept_count(int64_t nodes, const uint32_t* node_cnt, const uint8_t* keep, uint32_t* kept) {
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    kept[n] = fan_kept(keep, n) ? node_cnt[n] : 0u;
}

// Which nodes keep their fan, bottom-up with a cost per point outside the
// node: a visit is one node read; a fan adds ceil(edges / 64) rounds of
// solid angles, a leaf one round of its triangles, and descending the costs
// of the children.  A node keeps its fan when that is no dearer than
// descending.  Either way the winding number is exact; only the work varies.
// Nodes below the leaf level are never visited and keep nothing.
std::vector<uint8_t> fan_choice(const Shape& s, int leaf, const std::vector<uint32_t>& node_cnt) {
    std::vector<uint8_t> keep(node_cnt.size(), 0);
    std::vector<double> cost(node_cnt.size(), 0.0);
    for (int h = leaf; h <= s.H; ++h)
        for (int64_t k = 0; k < s.count[h]; ++k) {
            const int64_t n = s.off[h] + k;
            if (h == leaf) {
                cost[n] = 2.0;
                continue;
            }
            const int64_t c0 = s.off[h - 1] + 2 * k;
            const double down = 1.0 + cost[c0] + (2 * k + 1 < s.count[h - 1] ? cost[c0 + 1] : 0.0);
            const double fan = 1.0 + std::ceil(node_cnt[n] / 64.0);
            keep[n] = fan <= down;
            cost[n] = std::min(fan, down);
        }
    return keep;
}

int bits_for(uint64_t maxval) {
    int b = 1;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

// ---------------------------------------------------------- radix sort ----
// Stable LSD radix sort on the low `bits` bits of the keys, 8-bit digits,
// three launches per digit:
//   - k_rs_hist: the digit histogram of each tile of 2048 keys (LDS
//     atomics), stored digit-major, hist[d * tiles + tile];
//   - k_rs_offsets: one workgroup per digit scans that digit's row of tile
//     counts (exclusive) and stores the row total;
//   - k_rs_scatter: each tile starts every digit at the sum of the smaller
//     digits' totals plus its row offset, then, in 8 rounds of 256 keys in
//     input order, ranks every key among the keys of its digit before it (a
//     mask of the wavefront's lanes with the same digit from 8 ballots, and
//     per-wave counts in LDS) and writes key and value there.
// Keys with equal digits keep their input order within a tile, and tiles
// follow one another in the offsets, so every pass is stable and equal keys
// end in their input order, which the triangle order and the fan edge order
// rely on.
constexpr int kRsThreads = 256, kRsRounds = 8, kRsTile = kRsThreads * kRsRounds, kRsBins = 256;
static_assert(kRsThreads == kRsBins, "one histogram bin per thread");

template <class K>
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const K* keys, int64_t n, int shift, uint32_t* hist,
                                                       int64_t tiles) {
    __shared__ uint32_t h[kRsBins];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int j = 0; j < kRsRounds; ++j) {
        const int64_t i = base + (int64_t)j * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & (kRsBins - 1)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of 256 values of a 256-thread block (valid in every thread)
__device__ inline uint32_t block_excl_scan256(uint32_t v, uint32_t* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int q = 0; q < w; ++q) before += sh[q];
    __syncthreads();
    return before + incl - v;
}

// row d = blockIdx.x of the digit-major tile counts -> exclusive offsets
// within the digit; the row total to tot[d]
__global__ __launch_bounds__(kRsThreads) void k_rs_offsets(uint32_t* hist, int64_t tiles, uint32_t* tot) {
    __shared__ uint32_t sh[kRsThreads / 64 + 1];
    uint32_t* row = hist + (int64_t)blockIdx.x * tiles;
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < tiles; b0 += kRsThreads) {
        const int64_t b = b0 + threadIdx.x;
        const uint32_t v = b < tiles ? row[b] : 0u;
        const uint32_t ex = block_excl_scan256(v, sh);
        if (b < tiles) row[b] = carry + ex;
        if (threadIdx.x == kRsThreads - 1) sh[kRsThreads / 64] = ex + v;
        __syncthreads();
        carry += sh[kRsThreads / 64];
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

template <class K, class V>
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const K* kin, const V* vin, K* kout, V* vout, int64_t n,
                                                          int shift, const uint32_t* off, int64_t tiles,
                                                          const uint32_t* tot) {
    constexpr int kW = kRsThreads / 64;
    __shared__ uint32_t next[kRsBins];       // this tile's next output position per digit
    __shared__ uint32_t wcnt[kW][kRsBins];   // keys per digit in each wavefront of the current round
    __shared__ uint32_t sh[kW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t before = (1ull << lane) - 1ull;
    // digit d starts after every smaller digit's keys and this digit's keys of earlier tiles
    const uint32_t digit_base = block_excl_scan256(tot[threadIdx.x], sh);
    next[threadIdx.x] = digit_base + off[(int64_t)threadIdx.x * tiles + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int j = 0; j < kRsRounds; ++j) {
        const int64_t i = base + (int64_t)j * kRsThreads + threadIdx.x;
        const bool valid = i < n;
        const K k = valid ? kin[i] : K(0);
        const uint32_t d = (uint32_t)(k >> shift) & (kRsBins - 1);
        uint64_t same = __ballot(valid);   // lanes holding a key with this lane's digit
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool set = (d >> bit) & 1u;
            const uint64_t m = __ballot(set);
            same &= set ? m : ~m;
        }
        for (int q = threadIdx.x; q < kW * kRsBins; q += kRsThreads) (&wcnt[0][0])[q] = 0u;
        __syncthreads();
        const uint32_t rank = (uint32_t)__popcll(same & before);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(same);
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            pos = next[d] + rank;
            for (int q = 0; q < w; ++q) pos += wcnt[q][d];
        }
        __syncthreads();
        uint32_t round_total = 0;
        for (int q = 0; q < kW; ++q) round_total += wcnt[q][threadIdx.x];
        next[threadIdx.x] += round_total;
        if (valid) {
            kout[pos] = k;
            if (vout) vout[pos] = vin[i];
        }
        __syncthreads();
    }
}

// sorts keys (and vals, when given) in place; n below 2^32
template <class K, class V>
void radix_sort(DBuf<K>& keys, DBuf<V>* vals, int64_t n, int bits, hipStream_t st) {
    if (n <= 1 || bits <= 0) return;
    ITS_REQUIRE(n < (int64_t(1) << 32), ITS_E_RANGE, "radix sort of 2^32 keys or more");
    const int64_t tiles = (n + kRsTile - 1) / kRsTile;
    DBuf<K> k2(n);
    DBuf<V> v2(vals ? n : 0);
    DBuf<uint32_t> hist(kRsBins * tiles), tot(kRsBins);
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist<K>, dim3((unsigned)tiles), dim3(kRsThreads), 0, st, (const K*)keys.p, n, shift,
                           hist.p, tiles);
        ITS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_rs_offsets, dim3(kRsBins), dim3(kRsThreads), 0, st, hist.p, tiles, tot.p);
        ITS_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_rs_scatter<K, V>), dim3((unsigned)tiles), dim3(kRsThreads), 0, st, (const K*)keys.p,
                           vals ? (const V*)vals->p : (const V*)nullptr, k2.p, vals ? v2.p : (V*)nullptr, n, shift,
                           (const uint32_t*)hist.p, tiles, (const uint32_t*)tot.p);
        ITS_LAUNCH_CHECK();
        std::swap(keys, k2);
        if (vals) std::swap(*vals, v2);
    }
}

struct TriTree {
    Shape s{};
    Quant q{};
    int64_t nodes = 0, fans = 0;
    DBuf<double> tris;
    DBuf<float> fan;
    DBuf<float4> sph, node;
    DBuf<int32_t> tid;
    DBuf<uint32_t> code, cell_lb;
    int cell_bits = 0;
    bool closed = false;   // every undirected edge: exactly two half edges, opposite directions
    // half edges sorted by undirected edge (for the fans) and the open flag
    DBuf<uint64_t> hkeys;
    DBuf<uint32_t> hvals;
    DBuf<int> is_open;
    int64_t M = 0;
    uint64_t none = 0;
    int vb = 0;
    TreeView view() const {
        return TreeView{s, q, tris.p, sph.p, tid.p, code.p, node.p, fan.p, cell_lb.p, cell_bits};
    }
};

// lower bound of each code prefix's range in the sorted codes: entry C
// (0 <= C <= 2^bits) is the first position whose code is >= C << (30 - bits)
__global__ void k_cell_lb(const uint32_t* code, int64_t T, int bits, uint32_t* lb) {
    const int64_t C = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (C > ((int64_t)1 << bits)) return;
    const uint64_t target = (uint64_t)C << (30 - bits);
    int64_t a = 0, b = T;
    while (a < b) {
        const int64_t m = (a + b) >> 1;
        if ((uint64_t)code[m] < target) a = m + 1; else b = m;
    }
    lb[C] = (uint32_t)a;
}

void tree_build(TriTree& tr, hipStream_t st, const double* d_V, int64_t nV, const int32_t* d_F, int64_t nF) {
    ITS_REQUIRE(nF > 0 && nF < (1ll << 30) && nV < (1ll << 31), ITS_E_RANGE, "mesh too large for the triangle tree (2^30 triangles, 2^31 vertices)");
    Shape& s = tr.s;
    s.T = (int32_t)nF;
    s.count[0] = (int32_t)((nF + kBucket - 1) / kBucket);
    s.off[0] = 0;
    s.H = 0;
    while (s.count[s.H] > 1) {
        ITS_REQUIRE(s.H + 1 < kMaxLevels, ITS_E_INVALID, "triangle tree too deep");
        s.count[s.H + 1] = (s.count[s.H] + 1) / 2;
        s.off[s.H + 1] = s.off[s.H] + s.count[s.H];
        ++s.H;
    }
    tr.nodes = (int64_t)s.off[s.H] + 1;

    // Morton quantisation over the centroid box
    const int nb = (int)std::min<int64_t>(256, blocks_for(nF, kT));
    DBuf<double> part(6 * (size_t)nb);
    hipLaunchKernelGGL(k_centroid_box, dim3(nb), dim3(kT), 0, st, d_V, d_F, nF, part.p);
    ITS_LAUNCH_CHECK();
    std::vector<double> hp;
    to_host(hp, part.p, 6 * (size_t)nb, st);
    ITS_HIP(hipStreamSynchronize(st));
    for (int a = 0; a < 3; ++a) {
        double lo = INFINITY, hi = -INFINITY;
        for (int b = 0; b < nb; ++b) { lo = std::min(lo, hp[6 * b + a]); hi = std::max(hi, hp[6 * b + 3 + a]); }
        tr.q.lo[a] = lo;
        tr.q.scale[a] = (hi > lo) ? 1024.0 / (hi - lo) : 0.0;
    }

    // sort triangles by centroid code (stable: ties by index)
    tr.code.alloc(nF);
    DBuf<uint32_t> tidx(nF);
    hipLaunchKernelGGL(k_tri_keys, dim3(blocks_for(nF, kT)), dim3(kT), 0, st, d_V, d_F, nF, tr.q, tr.code.p, tidx.p);
    ITS_LAUNCH_CHECK();
    radix_sort(tr.code, &tidx, nF, 30, st);
    tr.tris.alloc(9 * (size_t)nF);
    tr.sph.alloc(nF);
    tr.tid.alloc(nF);
    DBuf<int32_t> pos(nF);
    hipLaunchKernelGGL(k_tri_gather, dim3(blocks_for(nF, kT)), dim3(kT), 0, st, (const uint32_t*)tidx.p, nF, d_V, d_F,
                       tr.tris.p, tr.sph.p, tr.tid.p, pos.p);
    ITS_LAUNCH_CHECK();
    // code-prefix table with about one prefix per triangle (3 to 21 bits)
    tr.cell_bits = 3;
    while (tr.cell_bits < 21 && ((int64_t)1 << tr.cell_bits) < nF) ++tr.cell_bits;
    const int64_t ncell = ((int64_t)1 << tr.cell_bits) + 1;
    tr.cell_lb.alloc(ncell);
    hipLaunchKernelGGL(k_cell_lb, dim3(blocks_for(ncell, kT)), dim3(kT), 0, st, (const uint32_t*)tr.code.p, nF,
                       tr.cell_bits, tr.cell_lb.p);
    ITS_LAUNCH_CHECK();

    // node boxes, one level per launch
    tr.node.alloc(2 * tr.nodes);
    hipLaunchKernelGGL(k_bucket_box, dim3(blocks_for(s.count[0], kT)), dim3(kT), 0, st, tr.tris.p, s, tr.node.p);
    ITS_LAUNCH_CHECK();
    for (int h = 1; h <= s.H; ++h) {
        hipLaunchKernelGGL(k_level_box, dim3(blocks_for(s.count[h], kT)), dim3(kT), 0, st, s, h, tr.node.p);
        ITS_LAUNCH_CHECK();
    }

    // half edges grouped by undirected edge: whether the mesh is closed, and
    // the boundaries of the fans
    const int64_t M = 3 * nF;
    const int vb = bits_for((uint64_t)(nV - 1));
    const uint64_t none = (2 * vb >= 64) ? ~0ull : ((1ull << (2 * vb)) - 1);
    tr.M = M;
    tr.vb = vb;
    tr.none = none;
    tr.hkeys.alloc(M);
    tr.hvals.alloc(M);
    hipLaunchKernelGGL(k_half_edges, dim3(blocks_for(M, kT)), dim3(kT), 0, st, d_F, nF, (const int32_t*)pos.p, vb, none,
                       tr.hkeys.p, tr.hvals.p);
    ITS_LAUNCH_CHECK();
    radix_sort(tr.hkeys, &tr.hvals, M, std::min(64, 2 * vb), st);
    tr.is_open.alloc(1);
    ITS_HIP(hipMemsetAsync(tr.is_open.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_closed_check, dim3(blocks_for(M, kT)), dim3(kT), 0, st, (const uint64_t*)tr.hkeys.p,
                       (const uint32_t*)tr.hvals.p, M, none, tr.is_open.p);
    ITS_LAUNCH_CHECK();
}

// Boundary fans of the tree's nodes, for the winding-number walk; built only
// when some point needs that walk.
void fans_build(TriTree& tr, hipStream_t st, const double* d_V) {
    const Shape& s = tr.s;
    const int64_t M = tr.M;
    const int vb = tr.vb;
    const uint64_t none = tr.none;
    const uint64_t* hkeys = tr.hkeys.p;
    const uint32_t* hvals = tr.hvals.p;
    DBuf<uint32_t> node_cnt(tr.nodes);
    ITS_HIP(hipMemsetAsync(node_cnt.p, 0, tr.nodes * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_fan_node_count, dim3(blocks_for(M, kT)), dim3(kT), 0, st, hkeys, hvals, M, s, none,
                       node_cnt.p);
    ITS_LAUNCH_CHECK();
    DBuf<uint8_t> keep(tr.nodes);
    {
        std::vector<uint32_t> h_cnt(tr.nodes);
        ITS_HIP(hipMemcpyAsync(h_cnt.data(), node_cnt.p, tr.nodes * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        ITS_HIP(hipStreamSynchronize(st));
        const std::vector<uint8_t> h_keep = fan_choice(s, std::min(kLeafLevel, (int)s.H), h_cnt);
        ITS_HIP(hipMemcpyAsync(keep.p, h_keep.data(), tr.nodes, hipMemcpyHostToDevice, st));
        ITS_HIP(hipStreamSynchronize(st));
    }
    DBuf<uint32_t> ecnt(M);
    hipLaunchKernelGGL(k_fan_edge_count, dim3(blocks_for(M, kT)), dim3(kT), 0, st, hkeys, hvals, M, s, none,
                       (const uint8_t*)keep.p, ecnt.p);
    ITS_LAUNCH_CHECK();
    const uint64_t E = scan_exclusive_u32(ecnt.p, ecnt.p, M, st);
    ITS_REQUIRE(E < (1ull << 31), ITS_E_RANGE, "triangle tree: more than 2^31 boundary fan edges");
    tr.fans = (int64_t)E;
    DBuf<uint32_t> fan_node(std::max<uint64_t>(E, 1));
    DBuf<uint64_t> fan_edge(std::max<uint64_t>(E, 1));
    if (E) {
        hipLaunchKernelGGL(k_fan_emit, dim3(blocks_for(M, kT)), dim3(kT), 0, st, hkeys, hvals, M, s, none, vb,
                           (const uint8_t*)keep.p, (const uint32_t*)ecnt.p, fan_node.p, fan_edge.p);
        ITS_LAUNCH_CHECK();
        // stable: within a node, fan edges stay in edge-key order
        radix_sort(fan_node, &fan_edge, (int64_t)E, bits_for((uint64_t)tr.nodes), st);
    }
    DBuf<uint32_t> fan_off(tr.nodes);
    hipLaunchKernelGGL(k_fan_kept_count, dim3(blocks_for(tr.nodes, kT)), dim3(kT), 0, st, tr.nodes,
                       (const uint32_t*)node_cnt.p, (const uint8_t*)keep.p, fan_off.p);
    ITS_LAUNCH_CHECK();
    const uint64_t E2 = scan_exclusive_u32(fan_off.p, fan_off.p, tr.nodes, st);
    ITS_REQUIRE(E2 == E, ITS_E_STATE, "triangle tree: fan edge counts disagree");
    hipLaunchKernelGGL(k_fan_table, dim3(blocks_for(tr.nodes, kT)), dim3(kT), 0, st, tr.nodes,
                       (const uint32_t*)node_cnt.p, (const uint8_t*)keep.p, (const uint32_t*)fan_off.p, tr.node.p);
    ITS_LAUNCH_CHECK();
    tr.fan.alloc(6 * std::max<uint64_t>(E, 1));
    if (E) {
        hipLaunchKernelGGL(k_fan_coords, dim3(blocks_for((int64_t)E, kT)), dim3(kT), 0, st, (const uint64_t*)fan_edge.p,
                           (const uint32_t*)fan_node.p, (int64_t)E, d_V, (const float4*)tr.node.p, tr.fan.p);
        ITS_LAUNCH_CHECK();
    }
    ITS_HIP(hipStreamSynchronize(st));
}

// -------------------------------------------------------------- query -----
// Both walks run one wavefront per point: the walk is the same for all
// lanes, and nodes of level kLeafLevel and below (at most 64 triangles) are
// leaves whose triangles the 64 lanes take one each.  Per-lane results are
// combined by a fixed butterfly, so nothing depends on timing.
constexpr int kWavesPerBlock = kT / 64;

struct NodeRec {
    float4 lo, hi;
    int first, count;
};
__device__ inline NodeRec load_node(const TreeView& t, int node) {
    const float4 lo = t.node[2 * node], hi = t.node[2 * node + 1];
    return NodeRec{lo, hi, __float_as_int(lo.w), __float_as_int(hi.w)};
}

// minimum of (d2, index) over the wavefront, ties to the lower index
__device__ inline void wave_argmin(double& d2, int32_t& id) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double od = __shfl_xor(d2, o, 64);
        const int32_t oi = __shfl_xor(id, o, 64);
        if (od < d2 || (od == d2 && oi < id)) {
            d2 = od;
            id = oi;
        }
    }
}

// Values that are equal in every lane (the point of the wavefront, the best
// distance after the butterfly) are moved to scalar registers, which the
// compiler cannot infer from threadIdx.x >> 6 or from shuffles.
__device__ inline int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double wave_uniform(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// false when the bounding sphere of sorted triangle i is farther than the
// best distance so far
__device__ inline bool sphere_within(const TreeView& t, int64_t i, const double p[3], float pm, double best) {
    const float4 S = t.sph[i];
    const float dx = (float)p[0] - S.x, dy = (float)p[1] - S.y, dz = (float)p[2] - S.z;
    const float m = bound_margin(pm + fmaxf(fabsf(S.x), fmaxf(fabsf(S.y), fabsf(S.z))) + S.w);
    const float lb = sqrtf(dx * dx + dy * dy + dz * dz) - S.w - m;
    return !(lb > 0.f && (double)lb * (double)lb > best);
}
// binary64 distance of sorted triangle i (when take), merged over the
// wavefront into (best, bt)
__device__ inline void tri_nearest(const TreeView& t, int64_t i, bool take, const double p[3], double& best,
                                   int32_t& bt) {
    double d2 = INFINITY;
    int32_t id = 0x7fffffff;
    if (take) {
        const double* T = t.tris + 9 * i;
        d2 = point_tri_dist2(p, T, T + 3, T + 6);
        id = t.tid[i];
        // a degenerate triangle can give 0/0; like a plain "d2 < best"
        // scan, a NaN never wins (and must not reach the butterfly)
        if (!(d2 <= INFINITY)) {
            d2 = INFINITY;
            id = 0x7fffffff;
        }
    }
    wave_argmin(d2, id);
    if (d2 < best || (d2 == best && id < bt)) {
        best = d2;
        bt = id;
    }
    best = wave_uniform(best);
    bt = wave_uniform(bt);
}
// nearest triangle of a leaf, one triangle per lane; a triangle whose
// bounding sphere is farther than the best so far is skipped
__device__ inline void leaf_nearest(const TreeView& t, int h, int64_t k, int lane, const double p[3], float pm,
                                    double& best, int32_t& bt) {
    const int64_t i = node_first(h, k) + lane;
    const bool take = i < node_last(t.s, h, k) && sphere_within(t, i, p, pm, best);
    tri_nearest(t, i, take, p, best, bt);
}

// true when no point of the box can be within the best distance found so
// far: binary32 distance to the box (boxes are rounded outward, so they hold
// every vertex), reduced by a margin that covers its rounding
__device__ inline float box_lb(const float pf[3], float pm, float4 lo, float4 hi) {
    const float l[3] = {lo.x, lo.y, lo.z}, u[3] = {hi.x, hi.y, hi.z};
    float d2 = 0.f, sc = pm;
    for (int a = 0; a < 3; ++a) {
        const float e = fmaxf(fmaxf(l[a] - pf[a], pf[a] - u[a]), 0.f);
        d2 += e * e;
        sc = fmaxf(sc, fmaxf(fabsf(l[a]), fabsf(u[a])));
    }
    return sqrtf(d2) - bound_margin(sc);
}
__device__ inline bool lb_beyond(float lb, double best) { return lb > 0.f && (double)lb * (double)lb > best; }
__device__ inline bool box_beyond(const float pf[3], float pm, float4 lo, float4 hi, double best) {
    return lb_beyond(box_lb(pf, pm, lo, hi), best);
}

// the writes of some lanes are read by other lanes of the same wavefront
__device__ inline void lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The frontier walks start at the lowest level of at most `most` nodes (not
// below the leaves), testing all its nodes at once, several per lane,
// instead of descending from the root one level per round.  A box holds its
// children's boxes, so a node that the walk from the root would have dropped
// with an ancestor only makes the frontier larger, never changes the result.
__device__ inline int start_level(const Shape& s, int leaf, int most) {
    int h = s.H;
    while (h > leaf && s.count[h - 1] <= most) --h;
    return h;
}
// the distance walk starts at up to 512 nodes (eight per lane); at most
// kFront of them survive
constexpr int kDistStartNodes = 512;

__device__ inline bool box_outside(const double p[3], float4 lo, float4 hi) {
    return p[0] < (double)lo.x || p[0] > (double)hi.x || p[1] < (double)lo.y || p[1] > (double)hi.y ||
           p[2] < (double)lo.z || p[2] > (double)hi.z;
}

// Nearest triangle: binary64 distance (unsigned, into out) and the original
// index of the nearest triangle.  The walk starts with the leaf at the
// point's Morton position, which usually holds the nearest triangle.
// Distance walk, level by level: the wavefront keeps the nodes of one level
// that can still hold the nearest triangle (the frontier, in LDS) and tests
// all their children at once, one child pair per lane, against the distance
// from the seed leaf.  The surviving leaves are then searched one by one,
// skipped when their box is beyond the best distance found so far.  A leaf
// that can hold the nearest triangle has every ancestor box within the final
// distance, which is no larger than the seed's, so it is never dropped: the
// result is the exact (d², index) minimum.  A frontier that would overflow
// LDS sends the point to the node-by-node walk.
constexpr int kFront = 512;
// candidate queue: below 64 entries between rounds, plus two leaves of 64
constexpr int kCandQueue = 192;
__device__ void dist_dfs(const TreeView& t, int leaf, int64_t seed, int lane, const double p[3], const float pf[3],
                         float pm, double& best, int32_t& bt) {
    int h = t.s.H, k = 0;
    while (true) {
        const NodeRec r = load_node(t, t.s.off[h] + k);
        if (!(h == leaf && k == seed) && !box_beyond(pf, pm, r.lo, r.hi, best)) {
            if (h > leaf) {
                --h;
                k <<= 1;
                continue;
            }
            leaf_nearest(t, h, k, lane, p, pm, best, bt);
        }
        if (!tree_next(t.s, h, k)) break;
    }
}

__global__ __launch_bounds__(kT) void k_sdf_dist(TreeView t, const double* pts, int64_t N, double* out,
                                                 int32_t* tri_out) {
    __shared__ uint32_t fk[2][kWavesPerBlock][kFront];   // frontier node indices within their level
    __shared__ float flb[kWavesPerBlock][kFront];        // box lower bounds of the leaf-level frontier
    __shared__ uint32_t cq[kWavesPerBlock][kCandQueue];  // candidate triangles (sorted positions)
    const int lane = threadIdx.x & 63, w = wave_uniform((int)(threadIdx.x >> 6));
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
    const int leaf = std::min(kLeafLevel, (int)t.s.H);
    const int start = start_level(t.s, leaf, kDistStartNodes), start_count = t.s.count[start];
    for (int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + w; i < N; i += nw) {
        const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
        const float pm = fmaxf(fabsf(pf[0]), fmaxf(fabsf(pf[1]), fabsf(pf[2])));
        double best = INFINITY;
        int32_t bt = 0x7fffffff;
        // the lower bound of c among the sorted codes lies within its prefix's range
        const uint32_t c = quant_code(t.q, p);
        const uint32_t C = c >> (30 - t.cell_bits);
        int64_t a = t.cell_lb[C], b = t.cell_lb[C + 1];
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if (t.code[m] < c) a = m + 1; else b = m;
        }
        const int64_t seed = std::min<int64_t>(a, t.s.T - 1) >> (__builtin_ctz(kBucket) + leaf);
        leaf_nearest(t, leaf, seed, lane, p, pm, best, bt);
        // frontier: the nodes of the start level whose boxes are within the
        // seed's distance
        int n = 0, cur = 0;
        {
            constexpr int kQ = kDistStartNodes / 64;
            static_assert(kDistStartNodes <= kFront && kDistStartNodes % 64 == 0, "start nodes: kQ per lane");
            float lb[kQ];
            bool keep[kQ];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                lb[q] = 0.f;
                keep[q] = false;
                if (lane + 64 * q < start_count) {
                    const NodeRec r = load_node(t, t.s.off[start] + lane + 64 * q);
                    lb[q] = box_lb(pf, pm, r.lo, r.hi);
                    keep[q] = !lb_beyond(lb[q], best);
                }
            }
            lds_fence();
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const uint64_t m = __ballot(keep[q]);
                if (keep[q]) {
                    const int pos = n + __popcll(m & below);
                    fk[0][w][pos] = (uint32_t)(lane + 64 * q);
                    flb[w][pos] = lb[q];
                }
                n += __popcll(m);
            }
            lds_fence();
        }
        bool overflow = false;
        for (int h = start; h > leaf && n > 0; --h) {
            const int64_t coff = t.s.off[h - 1];
            const int ccount = t.s.count[h - 1];
            const bool to_leaf = (h - 1 == leaf);
            int m = 0;   // next frontier size (uniform)
            for (int j0 = 0; j0 < n; j0 += 64) {
                const int j = j0 + lane;
                bool keep0 = false, keep1 = false;
                uint32_t c0 = 0;
                float lb0 = 0.f, lb1 = 0.f;
                if (j < n) {
                    c0 = 2u * fk[cur][w][j];
                    const NodeRec r0 = load_node(t, (int)(coff + c0));
                    lb0 = box_lb(pf, pm, r0.lo, r0.hi);
                    keep0 = !lb_beyond(lb0, best);
                    if ((int)c0 + 1 < ccount) {
                        const NodeRec r1 = load_node(t, (int)(coff + c0 + 1));
                        lb1 = box_lb(pf, pm, r1.lo, r1.hi);
                        keep1 = !lb_beyond(lb1, best);
                    }
                }
                const uint64_t m0 = __ballot(keep0), m1 = __ballot(keep1);
                const int tot = __popcll(m0) + __popcll(m1);
                if (m + tot > kFront) {
                    overflow = true;
                    break;
                }
                const int pos = m + __popcll(m0 & below) + __popcll(m1 & below);
                if (keep0) {
                    fk[cur ^ 1][w][pos] = c0;
                    if (to_leaf) flb[w][pos] = lb0;
                }
                if (keep1) {
                    fk[cur ^ 1][w][pos + (keep0 ? 1 : 0)] = c0 + 1u;
                    if (to_leaf) flb[w][pos + (keep0 ? 1 : 0)] = lb1;
                }
                m += tot;
            }
            lds_fence();
            if (overflow) break;
            cur ^= 1;
            n = m;
        }
        if (overflow) {
            dist_dfs(t, leaf, seed, lane, p, pf, pm, best, bt);
        } else {
            // Candidates: the triangles of the surviving leaves that pass
            // the sphere test are queued and their binary64 distances taken
            // 64 at a time, so the wavefront runs the exact distance once per
            // 64 candidates rather than once per leaf.  A test against an
            // older, larger best only queues more, so the minimum is that of
            // the full search.
            // Two leaves per round, so each lane has two sphere reads in
            // flight; both are culled against the best of the round's start.
            int qn = 0;   // queued candidates, below 64 between rounds
            for (int j = 0; j < n; j += 2) {
                const int64_t k0 = fk[cur][w][j];
                const int64_t k1 = (j + 1 < n) ? (int64_t)fk[cur][w][j + 1] : -1;
                const bool use0 = !(k0 == seed || lb_beyond(flb[w][j], best));
                const bool use1 = k1 >= 0 && !(k1 == seed || lb_beyond(flb[w][j + 1], best));
                const int64_t t0 = use0 ? node_first(leaf, k0) + lane : 0;
                const int64_t t1 = use1 ? node_first(leaf, k1) + lane : 0;
                const bool pass0 = use0 && t0 < node_last(t.s, leaf, k0) && sphere_within(t, t0, p, pm, best);
                const bool pass1 = use1 && t1 < node_last(t.s, leaf, k1) && sphere_within(t, t1, p, pm, best);
                const uint64_t m0 = __ballot(pass0), m1 = __ballot(pass1);
                if (pass0) cq[w][qn + __popcll(m0 & below)] = (uint32_t)t0;
                qn += __popcll(m0);
                if (pass1) cq[w][qn + __popcll(m1 & below)] = (uint32_t)t1;
                qn += __popcll(m1);
                while (qn >= 64) {
                    lds_fence();
                    tri_nearest(t, cq[w][lane], true, p, best, bt);
                    qn -= 64;
                    const uint32_t rest0 = lane < qn ? cq[w][64 + lane] : 0u;
                    const uint32_t rest1 = lane + 64 < qn ? cq[w][128 + lane] : 0u;
                    lds_fence();
                    if (lane < qn) cq[w][lane] = rest0;
                    if (lane + 64 < qn) cq[w][64 + lane] = rest1;
                }
            }
            if (qn > 0) {
                lds_fence();
                tri_nearest(t, lane < qn ? cq[w][lane] : 0u, lane < qn, p, best, bt);
            }
            lds_fence();
        }
        if (lane == 0) {
            out[i] = sqrt(best);
            if (tri_out) tri_out[i] = bt;
        }
    }
}

// Winding number: fans of the nodes the point is outside of, triangles of
// the leaves it reaches.  Level by level: the wavefront classifies all nodes
// of a level at once (one per lane) into leaves, fans (the point is outside
// the box) and nodes to descend, keeping the lists in LDS; fans carry their
// edge range and apex - p, so their sums start without another node read.
// Lists that would overflow send the point to the node-by-node walk.
constexpr int kWindFront = 128, kWindFans = 128, kWindLeaves = 64;
struct FanItem {
    int first, count;
    float z[3], lz;
};

__device__ inline float leaf_wind(const TreeView& t, int h, int64_t k, int lane, const double p[3]) {
    const int64_t j = node_first(h, k) + lane;
    if (j >= node_last(t.s, h, k)) return 0.f;
    const double* T = t.tris + 9 * j;
    return solid_angle(p, T, T + 3, T + 6);
}
__device__ inline float fan_wind(const TreeView& t, int first, int count, const float z[3], float lz, int lane) {
    float acc = 0.f;
    for (int e = lane; e < count; e += 64) {
        const float* ab = t.fan + 6 * ((int64_t)first + e);
        const float u[3] = {ab[0] + z[0], ab[1] + z[1], ab[2] + z[2]};
        const float v[3] = {ab[3] + z[0], ab[4] + z[1], ab[5] + z[2]};
        acc += fan_angle(z, lz, u, v);
    }
    return acc;
}
__device__ inline FanItem fan_item(const NodeRec& r, const double p[3]) {
    // apex - p once per node in binary64; the edge ends are stored relative
    // to the apex, so a - p = (a - apex) + (apex - p)
    FanItem f;
    f.first = r.first;
    f.count = r.count;
    f.z[0] = (float)(fan_apex(r.lo.x, r.hi.x) - p[0]);
    f.z[1] = (float)(fan_apex(r.lo.y, r.hi.y) - p[1]);
    f.z[2] = (float)(fan_apex(r.lo.z, r.hi.z) - p[2]);
    f.lz = __builtin_amdgcn_sqrtf(f.z[0] * f.z[0] + f.z[1] * f.z[1] + f.z[2] * f.z[2]);
    return f;
}

// the node-by-node walk (lists overflowed)
__device__ float wind_dfs(const TreeView& t, int leaf, int lane, const double p[3]) {
    float acc = 0.f;
    int h = t.s.H, k = 0;
    while (true) {
        const NodeRec r = load_node(t, t.s.off[h] + k);
        if (h <= leaf) {
            acc += leaf_wind(t, h, k, lane, p);
        } else if (r.count >= 0 && box_outside(p, r.lo, r.hi)) {
            const FanItem f = fan_item(r, p);
            acc += fan_wind(t, f.first, f.count, f.z, f.lz, lane);
        } else {
            --h;
            k <<= 1;
            continue;
        }
        if (!tree_next(t.s, h, k)) break;
    }
    return acc;
}

// Orientation signs with static error filters (Shewchuk 1997, the first
// stage of orient2d and orient3d): +1 or -1 when the binary64 determinant
// exceeds the bound on its rounding error, 0 when the sign is not certain.
// Nothing here may be contracted into fused multiply-adds: the bounds count
// every rounding of the products and sums as written.
constexpr double kO2dBound = 3.3306690738754716e-16;   // (3 + 16 eps) eps, eps = 2^-53
constexpr double kO3dBound = 7.7715611723761027e-16;   // (7 + 56 eps) eps
// (a - p) x (b - p), z component: > 0 when p lies left of a -> b in xy
__device__ inline int o2d_sign(double ax, double ay, double bx, double by, double px, double py) {
#pragma clang fp contract(off)
    const double l = (ax - px) * (by - py);
    const double r = (ay - py) * (bx - px);
    const double det = l - r;
    const double bound = kO2dBound * (fabs(l) + fabs(r));
    return det > bound ? 1 : (-det > bound ? -1 : 0);
}
// (a - p) . ((b - p) x (c - p))
__device__ inline int o3d_sign(const double* a, const double* b, const double* c, const double p[3]) {
#pragma clang fp contract(off)
    const double adx = a[0] - p[0], ady = a[1] - p[1], adz = a[2] - p[2];
    const double bdx = b[0] - p[0], bdy = b[1] - p[1], bdz = b[2] - p[2];
    const double cdx = c[0] - p[0], cdy = c[1] - p[1], cdz = c[2] - p[2];
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                        (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    const double bound = kO3dBound * perm;
    return det > bound ? 1 : (-det > bound ? -1 : 0);
}

// An axis ray from p: along axis w in direction sg (+1 or -1), with u, v the
// next two axes in cyclic order, so (u, v, w) is oriented like (x, y, z) and
// the (u, v) orientation of a triangle has the sign of its normal's w
// component.  Any ray gives the winding number; each point takes the axis
// direction that leaves the mesh's box soonest, which meets the fewest nodes.
struct Ray {
    int u, v, w, sg;
};
__device__ inline double sel3(const double* x, int i) { return i == 0 ? x[0] : (i == 1 ? x[1] : x[2]); }
__device__ inline float sel3(float4 x, int i) { return i == 0 ? x.x : (i == 1 ? x.y : x.z); }
__device__ inline Ray choose_ray(const double p[3], float4 lo, float4 hi) {
    Ray R{0, 1, 2, 1};
    float len = INFINITY;
    for (int a = 0; a < 3; ++a) {
        const float pa = (float)sel3(p, a);
        const float ahead = sel3(hi, a) - pa, behind = pa - sel3(lo, a);
        if (ahead < len) {
            len = ahead;
            R = Ray{(a + 1) % 3, (a + 2) % 3, a, 1};
        }
        if (behind < len) {
            len = behind;
            R = Ray{(a + 1) % 3, (a + 2) % 3, a, -1};
        }
    }
    return R;
}

// false when the ray surely misses sorted triangle j: its bounding sphere,
// which holds the triangle with a margin for the binary32 rounding here,
// lies wholly beside the ray or wholly behind p.  Such a triangle is
// rejected without reading its binary64 vertices; that p lies outside its
// projection is then certain, so the crossing count is unchanged.
__device__ inline bool sphere_on_ray(const TreeView& t, int64_t j, const double p[3], float pm, const Ray& R) {
    const float4 S = t.sph[j];
    const float du = (float)sel3(p, R.u) - sel3(S, R.u), dv = (float)sel3(p, R.v) - sel3(S, R.v);
    const float dw = (float)sel3(p, R.w) - sel3(S, R.w);
    const float m = bound_margin(pm + fmaxf(fabsf(S.x), fmaxf(fabsf(S.y), fabsf(S.z))) + S.w);
    const float beside = sqrtf(du * du + dv * dv) - S.w - m;
    const float behind = (R.sg > 0 ? dw : -dw) - S.w - m;
    return !(beside > 0.f) && !(behind > 0.f);
}

// Signed crossing of the ray from p with one triangle: +1 / -1 (the sign of
// n . d, its normal against the ray direction) when the ray crosses it ahead
// of p, else 0; unsure is set when an orientation that decides this is not
// certain.
__device__ inline int tri_ray(const double* T, const double p[3], const Ray& R, bool& unsure) {
    const double* a = T;
    const double* b = T + 3;
    const double* c = T + 6;
    const double pu = sel3(p, R.u), pv = sel3(p, R.v), pw = sel3(p, R.w);
    const double au = a[R.u], av = a[R.v], aw = a[R.w];
    const double bu = b[R.u], bv = b[R.v], bw = b[R.w];
    const double cu = c[R.u], cv = c[R.v], cw = c[R.w];
    const double wmin = fmin(aw, fmin(bw, cw)), wmax = fmax(aw, fmax(bw, cw));
    // outside the (u, v) bounding rectangle, or wholly behind p: no crossing
    if (pu < fmin(au, fmin(bu, cu)) || pu > fmax(au, fmax(bu, cu)) || pv < fmin(av, fmin(bv, cv)) ||
        pv > fmax(av, fmax(bv, cv)) || (R.sg > 0 ? wmax < pw : wmin > pw))
        return 0;
    const int s1 = o2d_sign(au, av, bu, bv, pu, pv);
    const int s2 = o2d_sign(bu, bv, cu, cv, pu, pv);
    const int s3 = o2d_sign(cu, cv, au, av, pu, pv);
    const bool pos = s1 > 0 || s2 > 0 || s3 > 0, neg = s1 < 0 || s2 < 0 || s3 < 0;
    if (pos && neg) return 0;   // two certain signs disagree: outside the projection
    if (s1 == 0 || s2 == 0 || s3 == 0) {
        unsure = true;
        return 0;
    }
    // inside the projection; s1 is the (u, v) orientation of (a, b, c), the
    // sign of n_w, so n . d has the sign dir.  The ray meets the plane at
    // t = ((a - p) . n) / (n . d), n = (b - a) x (c - a), and (a - p) . n is
    // the orientation of (a, b, c, p): ahead of p when it has the sign dir
    const int dir = R.sg > 0 ? s1 : -s1;
    if (R.sg > 0 ? wmin > pw : wmax < pw) return dir;   // wholly ahead
    const int s4 = o3d_sign(a, b, c, p);
    if (s4 == 0) {
        unsure = true;
        return 0;
    }
    return s4 == dir ? dir : 0;
}

// Winding numbers of a closed mesh by an axis ray (see the file comment):
// the wavefront walks the tree level by level, keeping the nodes whose
// boxes meet the ray, then tests the triangles of the leaves it reaches,
// one per lane.  decided[i] = 1 and the sign is applied when every
// orientation was certain and the lists fitted; otherwise decided[i] = 0
// and the fan walk takes the point.
constexpr int kRayFront = 256, kRayLeaves = 64;
// the ray walk starts at up to 256 nodes (four per lane), one round lower
constexpr int kRayStartNodes = 256;
__device__ inline bool box_on_ray(const double p[3], float4 lo, float4 hi, const Ray& R) {
    const double pu = sel3(p, R.u), pv = sel3(p, R.v), pw = sel3(p, R.w);
    return pu >= (double)sel3(lo, R.u) && pu <= (double)sel3(hi, R.u) && pv >= (double)sel3(lo, R.v) &&
           pv <= (double)sel3(hi, R.v) && (R.sg > 0 ? pw <= (double)sel3(hi, R.w) : pw >= (double)sel3(lo, R.w));
}
__global__ __launch_bounds__(kT) void k_sdf_ray(TreeView t, const double* pts, int64_t N, double* out,
                                               uint8_t* decided, unsigned* undecided) {
    __shared__ uint32_t fk[2][kWavesPerBlock][kRayFront];
    __shared__ uint32_t leaves[kWavesPerBlock][kRayLeaves];
    __shared__ uint32_t cq[kWavesPerBlock][kCandQueue];   // candidate triangles (sorted positions)
    const int lane = threadIdx.x & 63, w = wave_uniform((int)(threadIdx.x >> 6));
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
    const int leaf = std::min(kLeafLevel, (int)t.s.H);
    const int start = start_level(t.s, leaf, kRayStartNodes), start_count = t.s.count[start];
    const NodeRec root = load_node(t, t.s.off[t.s.H]);
    for (int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + w; i < N; i += nw) {
        const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        bool unsure = !(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]));
        const Ray R = choose_ray(p, root.lo, root.hi);
        int n = 0, cur = 0, nl = 0;
        {   // the nodes of the start level whose boxes meet the ray
            constexpr int kQ = kRayStartNodes / 64;
            static_assert(kRayStartNodes <= kRayFront && kRayStartNodes % 64 == 0, "start nodes: kQ per lane");
            bool keep[kQ];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                keep[q] = false;
                if (!unsure && lane + 64 * q < start_count) {
                    const NodeRec r = load_node(t, t.s.off[start] + lane + 64 * q);
                    keep[q] = box_on_ray(p, r.lo, r.hi, R);
                }
            }
            lds_fence();
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const uint64_t m = __ballot(keep[q]);
                if (keep[q]) fk[0][w][n + __popcll(m & below)] = (uint32_t)(lane + 64 * q);
                n += __popcll(m);
            }
            lds_fence();
        }
        for (int h = start; n > 0; --h) {
            if (h <= leaf) {   // every node of this level is a leaf
                if (nl + n > kRayLeaves) {
                    unsure = true;
                    break;
                }
                for (int j = lane; j < n; j += 64) leaves[w][nl + j] = fk[cur][w][j];
                nl += n;
                lds_fence();
                break;
            }
            const int64_t coff = t.s.off[h - 1];
            const int cnext = t.s.count[h - 1];
            int m = 0;
            for (int j0 = 0; j0 < n; j0 += 64) {
                const int j = j0 + lane;
                bool keep0 = false, keep1 = false;
                uint32_t c0 = 0;
                if (j < n) {
                    c0 = 2u * fk[cur][w][j];
                    const NodeRec r0 = load_node(t, (int)(coff + c0));
                    keep0 = box_on_ray(p, r0.lo, r0.hi, R);
                    if ((int)c0 + 1 < cnext) {
                        const NodeRec r1 = load_node(t, (int)(coff + c0 + 1));
                        keep1 = box_on_ray(p, r1.lo, r1.hi, R);
                    }
                }
                const uint64_t m0 = __ballot(keep0), m1 = __ballot(keep1);
                const int tot = __popcll(m0) + __popcll(m1);
                if (m + tot > kRayFront) {
                    unsure = true;
                    break;
                }
                const int pos = m + __popcll(m0 & below) + __popcll(m1 & below);
                if (keep0) fk[cur ^ 1][w][pos] = c0;
                if (keep1) fk[cur ^ 1][w][pos + (keep0 ? 1 : 0)] = c0 + 1u;
                m += tot;
            }
            lds_fence();
            if (unsure) break;
            cur ^= 1;
            n = m;
        }
        // Candidates: the triangles of the reached leaves whose bounding
        // spheres meet the ray are queued and tested 64 at a time, so the
        // wavefront reads binary64 vertices and runs the crossing test once
        // per 64 candidates rather than once per leaf.  The count is a sum of
        // integers, so the order does not change it.
        int cross = 0;
        bool lane_unsure = false;
        if (!unsure) {
            const float pm = fmaxf(fabsf((float)p[0]), fmaxf(fabsf((float)p[1]), fabsf((float)p[2])));
            int qn = 0;   // queued candidates, below 64 between rounds (two leaves a round)
            for (int l = 0; l < nl; l += 2) {
                const int64_t k0 = leaves[w][l];
                const int64_t k1 = (l + 1 < nl) ? (int64_t)leaves[w][l + 1] : -1;
                const int64_t j0 = node_first(leaf, k0) + lane;
                const int64_t j1 = k1 >= 0 ? node_first(leaf, k1) + lane : 0;
                const bool pass0 = j0 < node_last(t.s, leaf, k0) && sphere_on_ray(t, j0, p, pm, R);
                const bool pass1 = k1 >= 0 && j1 < node_last(t.s, leaf, k1) && sphere_on_ray(t, j1, p, pm, R);
                const uint64_t m0 = __ballot(pass0), m1 = __ballot(pass1);
                if (pass0) cq[w][qn + __popcll(m0 & below)] = (uint32_t)j0;
                qn += __popcll(m0);
                if (pass1) cq[w][qn + __popcll(m1 & below)] = (uint32_t)j1;
                qn += __popcll(m1);
                while (qn >= 64) {
                    lds_fence();
                    cross += tri_ray(t.tris + 9 * (int64_t)cq[w][lane], p, R, lane_unsure);
                    qn -= 64;
                    const uint32_t rest0 = lane < qn ? cq[w][64 + lane] : 0u;
                    const uint32_t rest1 = lane + 64 < qn ? cq[w][128 + lane] : 0u;
                    lds_fence();
                    if (lane < qn) cq[w][lane] = rest0;
                    if (lane + 64 < qn) cq[w][64 + lane] = rest1;
                }
            }
            if (qn > 0) {
                lds_fence();
                if (lane < qn) cross += tri_ray(t.tris + 9 * (int64_t)cq[w][lane], p, R, lane_unsure);
            }
            lds_fence();
        }
        unsure = unsure || __ballot(lane_unsure) != 0ull;
        const int wn = wave_sum(cross);
        if (lane == 0) {
            decided[i] = unsure ? 0 : 1;
            if (unsure) atomicAdd(undecided, 1u);
            if (!unsure && wn >= 1) out[i] = -out[i];   // winding number > 1/2
        }
    }
}

__global__ __launch_bounds__(kT) void k_sdf_wind(TreeView t, const double* pts, int64_t N, double* out,
                                                 const uint8_t* decided) {
    __shared__ uint32_t fk[2][kWavesPerBlock][kWindFront];   // nodes of the level being classified
    __shared__ FanItem fans[kWavesPerBlock][kWindFans];
    __shared__ uint32_t leaves[kWavesPerBlock][kWindLeaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
    const int leaf = std::min(kLeafLevel, (int)t.s.H);
    for (int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + w; i < N; i += nw) {
        if (decided && decided[i]) continue;   // the ray settled its sign
        const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        int n = 1, cur = 0, nf = 0, nl = 0;
        bool overflow = false;
        lds_fence();
        if (lane == 0) fk[0][w][0] = 0u;
        lds_fence();
        for (int h = t.s.H; n > 0; --h) {
            if (h <= leaf) {   // every node of this level is a leaf
                if (nl + n > kWindLeaves) {
                    overflow = true;
                    break;
                }
                for (int j = lane; j < n; j += 64) leaves[w][nl + j] = fk[cur][w][j];
                nl += n;
                lds_fence();
                break;
            }
            const int64_t off = t.s.off[h];
            const int cnext = t.s.count[h - 1];
            int m = 0;
            for (int j0 = 0; j0 < n; j0 += 64) {
                const int j = j0 + lane;
                bool is_fan = false, down = false;
                uint32_t k = 0;
                NodeRec r{};
                if (j < n) {
                    k = fk[cur][w][j];
                    r = load_node(t, (int)(off + k));
                    is_fan = r.count >= 0 && box_outside(p, r.lo, r.hi);
                    down = !is_fan;
                }
                const bool two = down && (int)(2 * k + 1) < cnext;
                const uint64_t mf = __ballot(is_fan), md = __ballot(down), m2 = __ballot(two);
                const int tf = __popcll(mf), td = __popcll(md) + __popcll(m2);
                if (nf + tf > kWindFans || m + td > kWindFront) {
                    overflow = true;
                    break;
                }
                if (is_fan) fans[w][nf + __popcll(mf & below)] = fan_item(r, p);
                if (down) {
                    const int pos = m + __popcll(md & below) + __popcll(m2 & below);
                    fk[cur ^ 1][w][pos] = 2 * k;
                    if (two) fk[cur ^ 1][w][pos + 1] = 2 * k + 1;
                }
                nf += tf;
                m += td;
            }
            lds_fence();
            if (overflow) break;
            cur ^= 1;
            n = m;
        }
        float acc = 0.f;
        if (overflow) {
            acc = wind_dfs(t, leaf, lane, p);
        } else {
            for (int f = 0; f < nf; ++f) {
                const FanItem it = fans[w][f];
                acc += fan_wind(t, it.first, it.count, it.z, it.lz, lane);
            }
            for (int l = 0; l < nl; ++l) acc += leaf_wind(t, leaf, leaves[w][l], lane, p);
        }
        const float wn = wave_sum(acc) * (float)(0.25 / M_PI);
        if (lane == 0) {
            const double d = out[i];
            out[i] = (wn > 0.5f) ? -d : d;
        }
    }
}
}  // namespace

void signed_distance(hipStream_t st, const double* d_V, int64_t nV, const int32_t* d_F, int64_t nF,
                     const double* d_pts, int64_t N, double* d_out, int32_t* d_tri) {
    if (N <= 0) return;
    TriTree tr;
    tree_build(tr, st, d_V, nV, d_F, nF);
    // One wavefront per point, grid-strided over as many blocks as the CUs
    // hold at once for that kernel: a block that waited for a free slot would
    // start its whole share after the resident blocks ended theirs.
    int dev = 0, cus = 256;
    ITS_HIP(hipGetDevice(&dev));
    ITS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    auto grid_of = [&](const void* kernel) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kT, 0) != hipSuccess || per_cu < 1) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        return (unsigned)std::min<int64_t>((N + kWavesPerBlock - 1) / kWavesPerBlock, (int64_t)per_cu * cus);
    };
    hipLaunchKernelGGL(k_sdf_dist, dim3(grid_of((const void*)k_sdf_dist)), dim3(kT), 0, st, tr.view(), d_pts, N, d_out,
                       d_tri);
    ITS_LAUNCH_CHECK();
    tr.closed = read1(tr.is_open.p, st) == 0;
    // closed mesh: signs by the ray where it decides them; the fans are built
    // and walked only for the points it leaves undecided
    DBuf<uint8_t> decided;
    int64_t undecided = N;
    if (tr.closed && tune_get(ITS_TUNE_SDF_RAY)) {
        decided.alloc(N);
        DBuf<unsigned> count(1);
        ITS_HIP(hipMemsetAsync(count.p, 0, sizeof(unsigned), st));
        hipLaunchKernelGGL(k_sdf_ray, dim3(grid_of((const void*)k_sdf_ray)), dim3(kT), 0, st, tr.view(), d_pts, N, d_out,
                           decided.p, count.p);
        ITS_LAUNCH_CHECK();
        undecided = read1(count.p, st);
    }
    if (undecided > 0) {
        fans_build(tr, st, d_V);
        hipLaunchKernelGGL(k_sdf_wind, dim3(grid_of((const void*)k_sdf_wind)), dim3(kT), 0, st, tr.view(), d_pts, N, d_out,
                           (const uint8_t*)decided.p);
        ITS_LAUNCH_CHECK();
    }
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
