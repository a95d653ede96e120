// Shell iso band for gfx950: the points where the mesh crosses the level-0
// node boundaries, and the inner / outer iso values of the thin shell.
//
// Replaces ThinShells::cpIntersectionPoints (reference ThinShells.cpp:38-479)
// and the min / max of cpLatentBSplineValue (655-755):
//   - edge points: every mesh edge (the ordered set of vertex pairs of the
//     faces) against the back (x), left (y) and bottom (z) face planes of the
//     level-0 nodes in its voxel range widened by one voxel, with the
//     reference's parameter and range tests in binary64;
//   - face points: every level-0 node edge against every triangle it can
//     cross: the plane parameter t in [0, 1] and the three edge-side tests of
//     the triangle (cross products against the unnormalised normal).
// The reference finds the candidate node edges by sweeps over edges sorted by
// their end point; a node edge outside a triangle's bounding box cannot pass
// the edge-side tests, so here every triangle enumerates the lattice edges of
// its bounding box widened by one voxel instead, and keeps those that are
// edges of a level-0 node (one of the 4 cells around the lattice edge is a
// node; the end points come from the node of smallest index).  Point sets
// are produced by count / scan / emit, so their order is deterministic.
// The iso values are the min and max of f over the model vertices and all
// points (binary32 evaluation of the fitted field).
#include <algorithm>
#include <vector>

#include "internal.hpp"

namespace its {
void eval_points_aos64(its_svo s, const double* pts, int64_t N, const double box_end[3], int mode, float out_val,
                       float* f, hipStream_t st);

namespace {
constexpr int kT = 256;

struct Fine {
    const uint32_t* morton;   // level-0 node codes, ascending
    const double* origin;     // node origins (global node index)
    const double* width;
    int64_t n0;
    int R;
    double o[3], vw;
};

__device__ inline int64_t find0(const Fine& S, uint32_t code) {
    int64_t lo = 0, hi = S.n0;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (S.morton[mid] < code) lo = mid + 1;
        else hi = mid;
    }
    return (lo < S.n0 && S.morton[lo] == code) ? lo : -1;
}
__device__ inline double dot3l(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// ---- mesh edges x node faces
__device__ inline void edge_range(const Fine& S, const double* p1, const double* p2, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        const int d1 = (int)((p1[a] - S.o[a]) / S.vw), d2 = (int)((p2[a] - S.o[a]) / S.vw);
        const int l = min(d1, d2) - 1, h = max(d1, d2) + 1;
        lo[a] = max(0, min(l, S.R - 1));
        hi[a] = max(0, min(h, S.R - 1));
    }
}
__global__ void k_edge_count(int64_t ne, const uint64_t* edges, const double* V, Fine S, uint64_t* cnt) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const double* p1 = V + 3 * (int64_t)(edges[e] >> 32);
    const double* p2 = V + 3 * (int64_t)(edges[e] & 0xFFFFFFFFu);
    int lo[3], hi[3];
    edge_range(S, p1, p2, lo, hi);
    cnt[e] = (uint64_t)(hi[0] - lo[0] + 1) * (uint64_t)(hi[1] - lo[1] + 1) * (uint64_t)(hi[2] - lo[2] + 1);
}
// one thread per (edge, voxel); pass 0 counts points per candidate, pass 1 writes them
template <int PASS>
__global__ void k_edge_points(int64_t ne, const uint64_t* edges, const double* V, Fine S, const uint64_t* coff,
                              uint64_t total, uint32_t* pcnt, const uint32_t* poff, double* out) {
    uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    int64_t lo_e = 0, hi_e = ne;   // owning edge: last e with coff[e] <= c
    while (hi_e - lo_e > 1) {
        int64_t mid = (lo_e + hi_e) >> 1;
        if (coff[mid] <= c) lo_e = mid;
        else hi_e = mid;
    }
    const int64_t e = lo_e;
    const double* p1 = V + 3 * (int64_t)(edges[e] >> 32);
    const double* p2 = V + 3 * (int64_t)(edges[e] & 0xFFFFFFFFu);
    int lo[3], hi[3];
    edge_range(S, p1, p2, lo, hi);
    const uint64_t k = c - coff[e];
    const uint64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    const int x = lo[0] + (int)(k % nx), y = lo[1] + (int)((k / nx) % ny), z = lo[2] + (int)(k / (nx * ny));
    uint32_t np = 0;
    const int64_t node = find0(S, morton3((uint32_t)x, (uint32_t)y, (uint32_t)z));
    if (node >= 0) {
        const double* o = S.origin + 3 * node;
        const double w = S.width[node];
        const double dir[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        for (int a = 0; a < 3; ++a) {
            double t = 1.7976931348623157e308;
            if (dir[a] != 0) t = (o[a] - p1[a]) / dir[a];
            if (!(0.0 <= t && t <= 1.0)) continue;
            const double q[3] = {p1[0] + t * dir[0], p1[1] + t * dir[1], p1[2] + t * dir[2]};
            const int b = a == 0 ? 1 : 0, d = a == 2 ? 1 : 2;
            if (!(o[b] <= q[b] && q[b] <= o[b] + w && o[d] <= q[d] && q[d] <= o[d] + w)) continue;
            if (PASS == 1) {
                double* dst = out + 3 * ((uint64_t)poff[c] + np);
                dst[0] = q[0];
                dst[1] = q[1];
                dst[2] = q[2];
            }
            ++np;
        }
    }
    if (PASS == 0) pcnt[c] = np;
}

// ---- node edges x mesh faces
__device__ inline void tri_range(const Fine& S, const double* T, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        const double mn = fmin(T[a], fmin(T[3 + a], T[6 + a])), mx = fmax(T[a], fmax(T[3 + a], T[6 + a]));
        lo[a] = max(0, (int)floor((mn - S.o[a]) / S.vw) - 1);
        hi[a] = min(S.R, (int)floor((mx - S.o[a]) / S.vw) + 1);
        if (hi[a] < lo[a]) hi[a] = lo[a] - 1;
    }
}
// lattice edges along axis a in the box: a-extent [lo, hi-1] (edge from i to i+1), others [lo, hi]
__device__ inline uint64_t tri_cands(const int lo[3], const int hi[3], int a) {
    uint64_t n = 1;
    for (int b = 0; b < 3; ++b) {
        const int64_t len = (b == a) ? (hi[b] - lo[b]) : (hi[b] - lo[b] + 1);
        if (len <= 0) return 0;
        n *= (uint64_t)len;
    }
    return n;
}
__global__ void k_face_count(int64_t nT, const double* tris, Fine S, uint64_t* cnt) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    int lo[3], hi[3];
    tri_range(S, tris + 9 * t, lo, hi);
    cnt[t] = tri_cands(lo, hi, 0) + tri_cands(lo, hi, 1) + tri_cands(lo, hi, 2);
}
template <int PASS>
__global__ void k_face_points(int64_t nT, const double* tris, Fine S, const uint64_t* coff, uint64_t total,
                              uint32_t* pcnt, const uint32_t* poff, double* out) {
    uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    int64_t lo_t = 0, hi_t = nT;
    while (hi_t - lo_t > 1) {
        int64_t mid = (lo_t + hi_t) >> 1;
        if (coff[mid] <= c) lo_t = mid;
        else hi_t = mid;
    }
    const int64_t t = lo_t;
    const double* T = tris + 9 * t;
    int lo[3], hi[3];
    tri_range(S, T, lo, hi);
    uint64_t k = c - coff[t];
    int a = 0;
    for (; a < 3; ++a) {
        const uint64_t n = tri_cands(lo, hi, a);
        if (k < n) break;
        k -= n;
    }
    int len[3];
    for (int b = 0; b < 3; ++b) len[b] = (b == a) ? (hi[b] - lo[b]) : (hi[b] - lo[b] + 1);
    int L[3];
    L[0] = lo[0] + (int)(k % len[0]);
    L[1] = lo[1] + (int)((k / len[0]) % len[1]);
    L[2] = lo[2] + (int)(k / ((uint64_t)len[0] * len[1]));
    // the cells around the lattice edge: offsets 0 / -1 in the two other axes
    const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
    int64_t best = -1;
    int bc[3] = {0, 0, 0};
    for (int s1 = 0; s1 < 2; ++s1)
        for (int s2 = 0; s2 < 2; ++s2) {
            int cell[3] = {L[0], L[1], L[2]};
            cell[b1] -= s1;
            cell[b2] -= s2;
            if (cell[0] < 0 || cell[1] < 0 || cell[2] < 0 || cell[0] >= S.R || cell[1] >= S.R || cell[2] >= S.R)
                continue;
            const int64_t node = find0(S, morton3((uint32_t)cell[0], (uint32_t)cell[1], (uint32_t)cell[2]));
            if (node >= 0 && (best < 0 || node < best)) {
                best = node;
                bc[0] = cell[0];
                bc[1] = cell[1];
                bc[2] = cell[2];
            }
        }
    uint32_t np = 0;
    if (best >= 0) {
        const double* o = S.origin + 3 * best;
        const double w = S.width[best];
        double q1[3], q2[3];
        for (int b = 0; b < 3; ++b) {
            const int off = L[b] - bc[b];   // 0 or 1 along the other axes, 0 along a
            q1[b] = o[b] + (off ? w : 0.0);
            q2[b] = o[b] + ((off || b == a) ? w : 0.0);
        }
        const double* p1 = T;
        const double* p2 = T + 3;
        const double* p3 = T + 6;
        double e1[3], e2[3], e3[3], u[3], v[3], n[3];
        for (int b = 0; b < 3; ++b) {
            e1[b] = p2[b] - p1[b];
            e2[b] = p3[b] - p2[b];
            e3[b] = p1[b] - p3[b];
            u[b] = p2[b] - p1[b];
            v[b] = p3[b] - p1[b];
        }
        cross3(u, v, n);
        const double mn[3] = {-n[0], -n[1], -n[2]};
        const double tdir = dot3l(mn, p1);
        const double ed[3] = {q2[0] - q1[0], q2[1] - q1[1], q2[2] - q1[2]};
        const double nd = dot3l(n, ed);
        if (fabsf((float)nd) >= 1e-9f) {
            const double tt = (-tdir - dot3l(n, q1)) / nd;
            if (!(tt < 0. || tt > 1.)) {
                const double ip[3] = {q1[0] + ed[0] * tt, q1[1] + ed[1] * tt, q1[2] + ed[2] * tt};
                double w1[3], w2[3], w3[3], cc[3];
                for (int b = 0; b < 3; ++b) {
                    w1[b] = ip[b] - p1[b];
                    w2[b] = ip[b] - p2[b];
                    w3[b] = ip[b] - p3[b];
                }
                bool in = true;
                cross3(e1, w1, cc);
                in = in && !(dot3l(cc, n) < 0);
                cross3(e2, w2, cc);
                in = in && !(dot3l(cc, n) < 0);
                cross3(e3, w3, cc);
                in = in && !(dot3l(cc, n) < 0);
                if (in) {
                    if (PASS == 1) {
                        double* dst = out + 3 * (uint64_t)poff[c];
                        dst[0] = ip[0];
                        dst[1] = ip[1];
                        dst[2] = ip[2];
                    }
                    np = 1;
                }
            }
        }
    }
    if (PASS == 0) pcnt[c] = np;
}
}  // namespace

// Intersection points on the device; returns the count of edge points and
// face points (edge points first).  d_V: nV x 3, d_tris: nT x 9 (binary64).
// ---- fine node edges (determineNodeEdge, sortEdge, uniqueEdge:
// reference SVO.cu:965-1008, 1033-1060, 1098-1110).  The reference sorts the
// 12 edges of every level-0 node by their end points in the fuzzy 1e-9
// lexicographic order, then by the smallest index of the nodes sharing the
// edge, and keeps one per group of equal edges.  End points are lattice
// points, so that order is the lattice order of the lower end (x, then y,
// then z) and, at a shared lower end, z-edge < y-edge < x-edge.  Here every
// edge is one bit of a lattice bitset in exactly that order,
//   bit = ((X (R+1) + Y) (R+1) + Z) * 3 + slot,   slot 0 z, 1 y, 2 x,
// its owner is the smallest level-0 node containing it (atomic minimum), and
// its end points are the owner's origin plus 0 or the width per axis.
namespace {
__device__ inline void fine_xyz(uint32_t m, int& x, int& y, int& z) {
    uint32_t ux, uy, uz;
    demorton3(m, ux, uy, uz);
    x = (int)ux;
    y = (int)uy;
    z = (int)uz;
}
__device__ inline uint64_t fedge_bit(int R, int X, int Y, int Z, int slot) {
    const uint64_t D = (uint64_t)R + 1;
    return (((uint64_t)X * D + (uint64_t)Y) * D + (uint64_t)Z) * 3u + (uint64_t)slot;
}
// the 12 edges of a cell: axis a (0 x, 1 y, 2 z) from the corner with
// offsets (dx, dy, dz), 0 along a
__device__ inline void cell_edge(int e, int& a, int off[3]) {
    a = e >> 2;
    const int u = e & 1, v = (e >> 1) & 1;
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    off[a] = 0;
    off[b] = u;
    off[c] = v;
}
__global__ void k_fedge_mark(int64_t n0, const uint32_t* morton, int R, RankWord* g) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 12 * n0) return;
    const int64_t i = t / 12;
    const int e = (int)(t - 12 * i);
    int x, y, z, a, off[3];
    fine_xyz(morton[i], x, y, z);
    cell_edge(e, a, off);
    const uint64_t bit = fedge_bit(R, x + off[0], y + off[1], z + off[2], 2 - a);
    atomicOr(&g[bit >> 5].bits, 1u << (bit & 31));
}
__global__ void k_fedge_owner(int64_t n0, const uint32_t* morton, int R, const RankWord* g, uint32_t* owner) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 12 * n0) return;
    const int64_t i = t / 12;
    const int e = (int)(t - 12 * i);
    int x, y, z, a, off[3];
    fine_xyz(morton[i], x, y, z);
    cell_edge(e, a, off);
    const uint64_t bit = fedge_bit(R, x + off[0], y + off[1], z + off[2], 2 - a);
    const RankWord w = g[bit >> 5];
    const uint32_t r = w.rank + __popc(w.bits & ((1u << (bit & 31)) - 1u));
    atomicMin(&owner[r], (uint32_t)i);
}
__global__ void k_fedge_emit(int64_t nwords, const RankWord* g, int R, const uint32_t* morton, const double* origin,
                             const double* width, const uint32_t* owner, double* ends, uint32_t* own_out) {
    const int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= nwords) return;
    const RankWord w = g[wi];
    uint32_t bits = w.bits, r = w.rank;
    const uint64_t D = (uint64_t)R + 1;
    while (bits) {
        const int k = __ffs(bits) - 1;
        bits &= bits - 1;
        const uint64_t bit = (uint64_t)wi * 32u + (uint64_t)k;
        const int slot = (int)(bit % 3u);
        const uint64_t cell = bit / 3u;
        const int Z = (int)(cell % D), Y = (int)((cell / D) % D), X = (int)(cell / (D * D));
        const int a = 2 - slot;
        const uint32_t node = owner[r];
        int x, y, z;
        fine_xyz(morton[node], x, y, z);
        const int lo[3] = {X - x, Y - y, Z - z};
        const double* o = origin + 3 * (int64_t)node;
        const double wd = width[node];
        for (int c = 0; c < 3; ++c) {
            const int hi = lo[c] + (c == a ? 1 : 0);
            ends[6 * (int64_t)r + c] = o[c] + (lo[c] ? wd : 0.0);
            ends[6 * (int64_t)r + 3 + c] = o[c] + (hi ? wd : 0.0);
        }
        own_out[r] = node;
        ++r;
    }
}
}  // namespace

int64_t fine_edges(its_svo s, DBuf<double>& ends, DBuf<uint32_t>& owner, hipStream_t st) {
    const int64_t n0 = s->esum_nodes[1] - s->esum_nodes[0];
    const uint32_t* morton = s->morton.p + s->esum_nodes[0];
    const int R = s->R;
    const uint64_t nbits = 3ull * ((uint64_t)R + 1) * ((uint64_t)R + 1) * ((uint64_t)R + 1);
    const int64_t nwords = (int64_t)((nbits + 31) / 32);
    DBuf<RankWord> g(nwords);
    ITS_HIP(hipMemsetAsync(g.p, 0, g.bytes(), st));
    if (n0 > 0) {
        hipLaunchKernelGGL(k_fedge_mark, dim3(blocks_for(12 * n0, kT)), dim3(kT), 0, st, n0, morton, R, g.p);
        ITS_LAUNCH_CHECK();
    }
    const int64_t n = (int64_t)rank_build(g.p, nwords, st);
    DBuf<uint32_t> own(std::max<int64_t>(n, 1));
    ITS_HIP(hipMemsetAsync(own.p, 0xFF, own.bytes(), st));
    if (n0 > 0) {
        hipLaunchKernelGGL(k_fedge_owner, dim3(blocks_for(12 * n0, kT)), dim3(kT), 0, st, n0, morton, R,
                           (const RankWord*)g.p, own.p);
        ITS_LAUNCH_CHECK();
    }
    ends.alloc(6 * std::max<int64_t>(n, 1));
    owner.alloc(std::max<int64_t>(n, 1));
    hipLaunchKernelGGL(k_fedge_emit, dim3(blocks_for(nwords, kT)), dim3(kT), 0, st, nwords, (const RankWord*)g.p, R,
                       morton, (const double*)s->node_origin.p + 3 * s->esum_nodes[0],
                       (const double*)s->node_width.p + s->esum_nodes[0], (const uint32_t*)own.p, ends.p, owner.p);
    ITS_LAUNCH_CHECK();
    ITS_HIP(hipStreamSynchronize(st));
    return n;
}

void shell_points(its_svo s, const double* d_V, const uint64_t* d_edges, int64_t ne, const double* d_tris, int64_t nT,
                  DBuf<double>& pts, int64_t* n_edge, int64_t* n_face, hipStream_t st) {
    Fine S{};
    S.morton = s->morton.p + s->esum_nodes[0];
    S.origin = s->node_origin.p;
    S.width = s->node_width.p;
    S.n0 = s->esum_nodes[1] - s->esum_nodes[0];
    S.R = s->R;
    for (int a = 0; a < 3; ++a) S.o[a] = s->origin[a];
    S.vw = s->voxel_width;
    // edge candidates
    DBuf<uint64_t> ec(std::max<int64_t>(ne, 1)), eo(std::max<int64_t>(ne, 1));
    uint64_t etot = 0;
    if (ne > 0) {
        hipLaunchKernelGGL(k_edge_count, dim3(blocks_for(ne, kT)), dim3(kT), 0, st, ne, d_edges, d_V, S, ec.p);
        ITS_LAUNCH_CHECK();
        etot = scan_exclusive_u64(ec.p, eo.p, ne, st);
    }
    DBuf<uint32_t> epc(std::max<uint64_t>(etot, 1)), epo(std::max<uint64_t>(etot, 1));
    uint64_t ecount = 0;
    if (etot > 0) {
        hipLaunchKernelGGL(k_edge_points<0>, dim3(blocks_for((int64_t)etot, kT)), dim3(kT), 0, st, ne, d_edges, d_V, S,
                           (const uint64_t*)eo.p, etot, epc.p, (const uint32_t*)nullptr, (double*)nullptr);
        ITS_LAUNCH_CHECK();
        ecount = scan_exclusive_u32(epc.p, epo.p, (int64_t)etot, st);
    }
    // face candidates
    DBuf<uint64_t> fc(std::max<int64_t>(nT, 1)), fo(std::max<int64_t>(nT, 1));
    uint64_t ftot = 0;
    if (nT > 0) {
        hipLaunchKernelGGL(k_face_count, dim3(blocks_for(nT, kT)), dim3(kT), 0, st, nT, d_tris, S, fc.p);
        ITS_LAUNCH_CHECK();
        ftot = scan_exclusive_u64(fc.p, fo.p, nT, st);
    }
    DBuf<uint32_t> fpc(std::max<uint64_t>(ftot, 1)), fpo(std::max<uint64_t>(ftot, 1));
    uint64_t fcount = 0;
    if (ftot > 0) {
        hipLaunchKernelGGL(k_face_points<0>, dim3(blocks_for((int64_t)ftot, kT)), dim3(kT), 0, st, nT, d_tris, S,
                           (const uint64_t*)fo.p, ftot, fpc.p, (const uint32_t*)nullptr, (double*)nullptr);
        ITS_LAUNCH_CHECK();
        fcount = scan_exclusive_u32(fpc.p, fpo.p, (int64_t)ftot, st);
    }
    pts.ensure(3 * std::max<uint64_t>(ecount + fcount, 1));
    if (ecount > 0) {
        hipLaunchKernelGGL(k_edge_points<1>, dim3(blocks_for((int64_t)etot, kT)), dim3(kT), 0, st, ne, d_edges, d_V, S,
                           (const uint64_t*)eo.p, etot, (uint32_t*)nullptr, (const uint32_t*)epo.p, pts.p);
        ITS_LAUNCH_CHECK();
    }
    if (fcount > 0) {
        hipLaunchKernelGGL(k_face_points<1>, dim3(blocks_for((int64_t)ftot, kT)), dim3(kT), 0, st, nT, d_tris, S,
                           (const uint64_t*)fo.p, ftot, (uint32_t*)nullptr, (const uint32_t*)fpo.p,
                           pts.p + 3 * ecount);
        ITS_LAUNCH_CHECK();
    }
    ITS_HIP(hipStreamSynchronize(st));
    *n_edge = (int64_t)ecount;
    *n_face = (int64_t)fcount;
}

// the ordered set of (min, max) vertex pairs of the faces (BaseModel.cpp:291-313)
std::vector<uint64_t> mesh_edges(const int32_t* F, int64_t nF) {
    std::vector<uint64_t> k(3 * (size_t)nF);
    for (int64_t f = 0; f < nF; ++f) {
        const int32_t i0 = F[3 * f], i1 = F[3 * f + 1], i2 = F[3 * f + 2];
        const int32_t mx = std::max(i0, std::max(i1, i2)), mn = std::min(i0, std::min(i1, i2));
        const int32_t md = i0 + i1 + i2 - mx - mn;
        k[3 * f] = ((uint64_t)(uint32_t)mn << 32) | (uint32_t)md;
        k[3 * f + 1] = ((uint64_t)(uint32_t)md << 32) | (uint32_t)mx;
        k[3 * f + 2] = ((uint64_t)(uint32_t)mn << 32) | (uint32_t)mx;
    }
    std::sort(k.begin(), k.end());
    k.erase(std::unique(k.begin(), k.end()), k.end());
    return k;
}

}  // namespace its
