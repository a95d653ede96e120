// Marching cubes over the fitted field for gfx950.
//
// Replaces MC::marching_cubes (reference MarchingCubes.cu:672-714) and its
// kernels prepareVoxelCornerKernel, determineVoxelKernel, compactVoxels and
// voxelToMeshKernel (124-377).
//
// The reference evaluates f at 8 corners per voxel.  Each corner is a brute
// force sum over all n basis functions, batched over 1024 streams.  Here f is
// evaluated once per lattice point ((res+1)^3 points), by the dense cascade on
// the octree's own lattice or by the octree field walk elsewhere.  A count
// pass writes every voxel's cube case and the vertex count of every step of
// 64 voxels along x, and lists the steps that have vertices; an exclusive
// scan of the step counts gives each step its offset; an emit pass, one wave
// per listed step, places each voxel's triangles with a wave prefix sum.  The
// soup order is the reference's: voxels x-fastest, triangles in table order.
// The 1e-6 snaps of vertexLerp (MarchingCubes.cu:72-90) are kept, and so is
// the lerp fma(t, v1, fma(-t, v0, v0)).
#include <algorithm>
#include <chrono>
#include <cmath>

#include <mutex>

#include "internal.hpp"
#include "mc_tables.inc"

namespace its {
void eval_lattice(its_svo s, const double o[3], const double vs[3], const int64_t dims[3], float* out, hipStream_t st);

namespace {
constexpr int kT = 256;
constexpr int kStepBits = 28;   // steps per grid < 2^28
constexpr unsigned long long kStepMask = (1ull << kStepBits) - 1;
__constant__ __align__(16) signed char c_tri[256][16];
__constant__ unsigned char c_nv[256];

// reference corner order (MarchingCubes.cu:142-149): (x,y,z) offsets
__constant__ unsigned char c_cx[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__constant__ unsigned char c_cy[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ unsigned char c_cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

// The voxel grid in the reference's voxel order (x fastest, then y, then z),
// cut into steps of 64 voxels along x: step s = (j + ry * k) * nch + chunk.
// The count pass gives each wave one 64-voxel column chunk over J rows of
// one z-slab: walking +y, the lattice rows at j+1 become those at j, so a
// step loads two new rows (x and x+1 for the planes k and k+1), prefetched
// one row ahead.  It writes each voxel's cube case and each step's vertex
// count; a scan of the step counts gives each step its place in the soup.
struct McGrid {
    int rx, ry, rz;
    int64_t lx, ly;   // lattice points per axis in x, y (res + 1)
    int nch;          // 64-voxel chunks per row
    int J;            // rows per count wave
    int njb;          // ceil(ry / J)
    int64_t nwaves;   // count waves: nch * njb * rz
    int64_t nsteps;   // ry * rz * nch
    float iso;      // the cube cases compare the binary32 lattice with this
    double iso64;   // the caller's iso value: vertexLerp's snaps and t
};

__device__ inline unsigned case8(const float (&v)[8], float iso) {
    unsigned c = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) c |= (unsigned)(v[q] < iso) << q;
    return c;
}

// pass 1: cube cases, vertices per step, voxels with triangles, and the
// list of steps with vertices.  One 64-bit counter holds both totals: listed
// steps in the low kStepBits bits, voxels with triangles above them, so a
// wave makes one atomic.  The list order depends on timing; each step's place
// in the soup comes from the scan of step_verts, so the output does not.
__global__ __launch_bounds__(kT) void k_mc_count(const float* __restrict__ lat, McGrid G, uint8_t* __restrict__ cases,
                                                 uint64_t* __restrict__ step_verts, unsigned long long* counts,
                                                 uint32_t* __restrict__ steps) {
    const int lane = threadIdx.x & 63;
    const int64_t w = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
    if (w >= G.nwaves) return;   // whole waves only
    const int ch = (int)(w % G.nch);
    const int64_t rest = w / G.nch;
    const int jb = (int)(rest % G.njb), k = (int)(rest / G.njb);
    const int j0 = jb * G.J, j1 = min(j0 + G.J, G.ry);
    const int i = ch * 64 + lane;
    const bool in = i < G.rx;
    const int ii = in ? i : G.rx - 1;   // lanes past the row read in bounds and count nothing
    const int64_t p0 = (int64_t)k * G.ly * G.lx + ii, p1 = p0 + G.ly * G.lx;
    // lattice rows at j (a: plane k, b: plane k + 1), each at x and x + 1
    float a0 = lat[p0 + j0 * G.lx], a1 = lat[p0 + j0 * G.lx + 1];
    float b0 = lat[p1 + j0 * G.lx], b1 = lat[p1 + j0 * G.lx + 1];
    float na0 = lat[p0 + (j0 + 1) * G.lx], na1 = lat[p0 + (j0 + 1) * G.lx + 1];
    float nb0 = lat[p1 + (j0 + 1) * G.lx], nb1 = lat[p1 + (j0 + 1) * G.lx + 1];
    unsigned nval = 0;
    uint32_t mask = 0;   // bit j - j0: step (k, j, ch) has vertices (J <= 32)
    for (int j = j0; j < j1; ++j) {
        const int64_t jn = min(j + 2, G.ry) * G.lx;
        const float f0 = lat[p0 + jn], f1 = lat[p0 + jn + 1], g0 = lat[p1 + jn], g1 = lat[p1 + jn + 1];
        // the reference's corner order: (x,y,z) = (0,0,0) (0,1,0) (1,1,0) (1,0,0) (0,0,1) (0,1,1) (1,1,1) (1,0,1)
        const float v[8] = {a0, na0, na1, a1, b0, nb0, nb1, b1};
        const unsigned c = case8(v, G.iso);
        const unsigned nv = in ? (unsigned)c_nv[c] : 0u;
        if (in) cases[((int64_t)k * G.ry + j) * G.rx + i] = (uint8_t)c;
        const unsigned sum = wave_sum(nv);
        nval += __popcll(__ballot(nv != 0));
        if (lane == 0) step_verts[((int64_t)k * G.ry + j) * G.nch + ch] = sum;
        if (sum) mask |= 1u << (j - j0);
        a0 = na0; a1 = na1; b0 = nb0; b1 = nb1;
        na0 = f0; na1 = f1; nb0 = g0; nb1 = g1;
    }
    if (!mask) return;
    unsigned long long slot = 0;
    if (lane == 0)
        slot = atomicAdd(counts, ((unsigned long long)nval << kStepBits) | (unsigned long long)__popc(mask)) &
               kStepMask;
    slot = __shfl(slot, 0, 64);
    if (lane < 32 && ((mask >> lane) & 1u))
        steps[slot + __popc(mask & ((1u << lane) - 1u))] = (uint32_t)(((int64_t)k * G.ry + j0 + lane) * G.nch + ch);
}

struct P3 {
    double x, y, z;
};
__device__ inline P3 vlerp(P3 p0, P3 p1, double s0, double s1, double iso) {
    if (fabs(iso - s0) < 1e-6) return p0;
    if (fabs(iso - s1) < 1e-6) return p1;
    if (fabs(s1 - s0) < 1e-6) return p0;
    double t = (iso - s0) / (s1 - s0);
    return P3{fma(t, p1.x, fma(-t, p0.x, p0.x)), fma(t, p1.y, fma(-t, p0.y, p0.y)), fma(t, p1.z, fma(-t, p0.z, p0.z))};
}

// The corner tables above as bit masks (bit q = offset of corner q), and the
// reference's edge numbering: edges 0-3 join corners e and (e+1) mod 4 on the
// z = 0 face, 4-7 the same on z = 1, 8-11 join corner e-8 to e-4.  Computed
// in registers, so no per-lane table lookups or indexed private arrays.
constexpr unsigned kCornerX = 0xCC, kCornerY = 0x66, kCornerZ = 0xF0;
__device__ inline int edge_a(int e) { return e < 8 ? e : e - 8; }
__device__ inline int edge_b(int e) { return e < 8 ? ((e & 4) | ((e + 1) & 3)) : e - 4; }
__device__ inline double pick8(const double (&v)[8], int q) {
    double r = v[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) r = q == k ? v[k] : r;
    return r;
}
__device__ inline P3 corner_pos(const P3& vp, double vx, double vy, double vz, int q) {
    // voxelToMeshKernel: corner = voxelPos + offset * voxelSize
    return P3{vp.x + (((kCornerX >> q) & 1) ? vx : 0.0), vp.y + (((kCornerY >> q) & 1) ? vy : 0.0),
              vp.z + (((kCornerZ >> q) & 1) ? vz : 0.0)};
}

// pass 2: one wave per listed step (64 voxels of a row).  A wave prefix sum
// over the 64 voxels places their triangles after those of all earlier
// voxels (step_off: exclusive scan of the step counts).
__device__ inline void emit_step(const float* __restrict__ lat, const McGrid& G, double ox, double oy, double oz,
                                 double vx, double vy, double vz, const uint8_t* __restrict__ cases, uint64_t at,
                                 int64_t s, uint64_t max_verts, float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int ch = (int)(s % G.nch);
    const int64_t row = s / G.nch;
    const int j = (int)(row % G.ry), k = (int)(row / G.ry);
    const int i = ch * 64 + lane;
    unsigned c = 0, nv = 0;
    if (i < G.rx) {
        c = cases[row * G.rx + i];
        nv = c_nv[c];
    }
    const uint32_t incl = wave_incl_scan(nv);
    if (!nv) return;   // after the wave scan: lanes part here
    const int64_t base = ((int64_t)k * G.ly + j) * G.lx;
    // voxelToMeshKernel: voxelPos = shift * voxelSize + origin
    const P3 vp{(double)i * vx + ox, (double)j * vy + oy, (double)k * vz + oz};
    double sv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sv[q] = (double)lat[base + (c_cz[q] * G.ly + c_cy[q]) * G.lx + i + c_cx[q]];
    const uint4 tri = *reinterpret_cast<const uint4*>(&c_tri[c][0]);
    const uint64_t first = at + incl - nv;
    const double iso = G.iso64;   // MarchingCubes.cu:72-90 works with the caller's binary64 iso
    for (unsigned t = 0; t < nv; t += 3) {
        const uint64_t pos = first + t;
        if (pos >= max_verts - 3) break;
        for (unsigned q = 0; q < 3; ++q) {
            const unsigned m = t + q;
            const unsigned word = (m >> 2) == 0 ? tri.x : (m >> 2) == 1 ? tri.y : (m >> 2) == 2 ? tri.z : tri.w;
            const int e = (int)(signed char)(word >> (8 * (m & 3)));
            const int qa = edge_a(e), qb = edge_b(e);
            const P3 p = vlerp(corner_pos(vp, vx, vy, vz, qa), corner_pos(vp, vx, vy, vz, qb), pick8(sv, qa),
                               pick8(sv, qb), iso);
            out[3 * (pos + q) + 0] = (float)p.x;
            out[3 * (pos + q) + 1] = (float)p.y;
            out[3 * (pos + q) + 2] = (float)p.z;
        }
    }
}

__global__ __launch_bounds__(kT) void k_mc_emit(const float* __restrict__ lat, McGrid G, double ox, double oy,
                                                double oz, double vx, double vy, double vz,
                                                const uint8_t* __restrict__ cases, const uint64_t* __restrict__ step_off,
                                                const uint32_t* __restrict__ steps,
                                                const unsigned long long* __restrict__ counts, uint64_t max_verts,
                                                float* __restrict__ out) {
    const int64_t nlist = (int64_t)min(counts[0] & kStepMask, (unsigned long long)G.nsteps);
    const int64_t e = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
    if (e >= nlist) return;   // whole waves
    const uint32_t s = steps[e];
    emit_step(lat, G, ox, oy, oz, vx, vy, vz, cases, step_off[s], (int64_t)s, max_verts, out);
}

// once per device, also when several host threads make their first call
void upload_tables() {
    static std::mutex mu;
    static uint64_t done_mask = 0;   // devices holding the tables
    int dev;
    ITS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (dev < 64 && ((done_mask >> dev) & 1u)) return;
    unsigned char nv[256];
    for (int c = 0; c < 256; ++c) {
        int n = 0;
        while (n < 16 && kMcTri[c][n] >= 0) ++n;
        nv[c] = (unsigned char)n;
    }
    ITS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tri), kMcTri, sizeof kMcTri));
    ITS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_nv), nv, sizeof nv));
    if (dev < 64) done_mask |= 1ull << dev;
}
}  // namespace

// The MC grid is the octree's level-0 lattice when it has R cells per axis
// over the octree box (the case of ThinShells::mcVisualization at the octree
// resolution).  Its lattice values are then the dense prolongation cascade
// down to level 0.  Otherwise every lattice point is evaluated on its own.
// The octree lattice spacing is width.x / R on all axes, and the MC spacing is
// width[a] / res; the float casts of the box padding make these differ by
// O(1e-8) relative, a difference in f far below binary32 resolution.
static void lattice_compute(its_svo s, const double origin[3], const double width[3], const int res[3],
                            DBuf<float>& lat, hipStream_t st);

// The last lattice stays in the octree's work space and is reused while the
// grid and the coefficients are the same: the count and fill calls of one
// extraction, and extractions at several iso values (mcVisualization).
static void lattice_values(its_svo s, const double origin[3], const double width[3], const int res[3],
                           DBuf<float>& lat, hipStream_t st) {
    auto& key = s->mc_key;
    bool same = key.valid && key.version == s->lambda_version && &lat == &s->mc_lat;
    for (int a = 0; a < 3; ++a)
        same = same && key.origin[a] == origin[a] && key.width[a] == width[a] && key.res[a] == res[a];
    if (same) return;
    key.valid = false;
    lattice_compute(s, origin, width, res, lat, st);
    if (&lat == &s->mc_lat) {
        key.valid = true;
        key.version = s->lambda_version;
        for (int a = 0; a < 3; ++a) {
            key.origin[a] = origin[a];
            key.width[a] = width[a];
            key.res[a] = res[a];
        }
    }
}

static void lattice_compute(its_svo s, const double origin[3], const double width[3], const int res[3],
                            DBuf<float>& lat, hipStream_t st) {
    double vs[3] = {width[0] / res[0], width[1] / res[1], width[2] / res[2]};
    int64_t dims[3] = {res[0] + 1, res[1] + 1, res[2] + 1};
    bool aligned = s->depth > 0;
    for (int a = 0; a < 3; ++a)
        aligned = aligned && res[a] == s->R && origin[a] == s->origin[a] &&
                  std::fabs(vs[a] - s->voxel_width) <= 1e-6 * s->voxel_width;
    if (aligned) {
        dense_cascade(s, 0, lat);
        return;
    }
    lat.ensure(dims[0] * dims[1] * dims[2]);
    eval_lattice(s, origin, vs, dims, lat.p, st);
}

void mc_lattice(its_svo s, const double origin[3], const double width[3], const int res[3], float* h_vals) {
    hipStream_t st = s->ctx->stream;
    int64_t dims[3] = {res[0] + 1, res[1] + 1, res[2] + 1};
    int64_t np = dims[0] * dims[1] * dims[2];
    lattice_values(s, origin, width, res, s->mc_lat, st);
    ITS_HIP(hipMemcpyAsync(h_vals, s->mc_lat.p, np * sizeof(float), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
}

void mc_extract(its_svo s, const double origin[3], const double width[3], const int res[3], double iso, float* h_tri,
                int64_t capacity, its_mc_stats* stats) {
    hipStream_t st = s->ctx->stream;
    hipEvent_t* ev = s->ctx->mc_ev;
    upload_tables();
    ITS_REQUIRE(res[0] > 0 && res[1] > 0 && res[2] > 0, ITS_E_INVALID, "resolution must be positive");
    its_mc_stats S{};
    auto t0 = std::chrono::steady_clock::now();
    double vs[3] = {width[0] / res[0], width[1] / res[1], width[2] / res[2]};   // MarchingCubes.cu:682
    int64_t nvox = (int64_t)res[0] * res[1] * res[2];
    ITS_REQUIRE(nvox <= ((int64_t)1 << 33), ITS_E_RANGE, "marching-cubes grid too large");
    ITS_HIP(hipEventRecord(ev[0], st));
    lattice_values(s, origin, width, res, s->mc_lat, st);
    ITS_HIP(hipEventRecord(ev[1], st));
    ITS_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    McGrid G{};
    G.rx = res[0];
    G.ry = res[1];
    G.rz = res[2];
    G.lx = res[0] + 1;
    G.ly = res[1] + 1;
    G.nch = (res[0] + 63) / 64;
    G.J = 32;
    G.njb = (res[1] + G.J - 1) / G.J;
    G.nwaves = (int64_t)G.nch * G.njb * res[2];
    G.nsteps = (int64_t)res[1] * res[2] * G.nch;
    G.iso = (float)iso;
    G.iso64 = iso;
    s->mc_case.ensure(nvox);
    s->mc_step.ensure(G.nsteps);
    ITS_REQUIRE(G.nsteps < (int64_t)kStepMask, ITS_E_RANGE, "marching-cubes grid has too many rows");
    s->mc_list.ensure(G.nsteps);
    s->mc_valid.ensure(2);
    ITS_HIP(hipMemsetAsync(s->mc_valid.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_mc_count, dim3((unsigned)((G.nwaves * 64 + kT - 1) / kT)), dim3(kT), 0, st,
                       (const float*)s->mc_lat.p, G, s->mc_case.p, s->mc_step.p, s->mc_valid.p, s->mc_list.p);
    ITS_LAUNCH_CHECK();
    // the packed counter and the scan's total come back in one read
    static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "counter width");
    scan_exclusive_u64_async(s->mc_step.p, s->mc_step.p, G.nsteps, st, (uint64_t*)s->mc_valid.p + 1);
    unsigned long long back[2];
    ITS_HIP(hipMemcpyAsync(back, s->mc_valid.p, sizeof back, hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    const unsigned long long counts = back[0];
    const uint64_t total = back[1];
    const int64_t nlist = (int64_t)(counts & kStepMask);
    S.num_verts = (int64_t)total;
    const uint64_t max_verts = 18ull * (uint64_t)nvox;   // maxVerts = nVoxels * 18
    if (h_tri && total) {
        ITS_REQUIRE((int64_t)total <= capacity, ITS_E_RANGE, "triangle buffer too small");
        s->mc_soup.ensure(3 * total);
        hipLaunchKernelGGL(k_mc_emit, dim3((unsigned)((nlist * 64 + kT - 1) / kT)), dim3(kT), 0, st,
                           (const float*)s->mc_lat.p, G, origin[0], origin[1], origin[2], vs[0], vs[1], vs[2],
                           (const uint8_t*)s->mc_case.p, (const uint64_t*)s->mc_step.p, (const uint32_t*)s->mc_list.p,
                           (const unsigned long long*)s->mc_valid.p, max_verts, s->mc_soup.p);
        ITS_LAUNCH_CHECK();
        ITS_HIP(hipEventRecord(ev[2], st));
        ITS_HIP(hipMemcpyAsync(h_tri, s->mc_soup.p, 3 * total * sizeof(float), hipMemcpyDeviceToHost, st));
    } else {
        ITS_HIP(hipEventRecord(ev[2], st));
    }
    ITS_HIP(hipStreamSynchronize(st));
    S.num_valid_voxels = (int64_t)(counts >> kStepBits);
    auto t2 = std::chrono::steady_clock::now();
    float ms_lat = 0.f, ms_ext = 0.f;
    ITS_HIP(hipEventElapsedTime(&ms_lat, ev[0], ev[1]));
    ITS_HIP(hipEventElapsedTime(&ms_ext, ev[1], ev[2]));
    S.seconds_eval = std::chrono::duration<double>(t1 - t0).count();
    S.seconds_extract = std::chrono::duration<double>(t2 - t1).count();
    S.seconds_eval_device = 1e-3 * ms_lat;
    S.seconds_extract_device = 1e-3 * ms_ext;
    if (stats) *stats = S;
}

}  // namespace its
