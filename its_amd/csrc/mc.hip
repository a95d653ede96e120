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
// pass writes the vertex count of every step of 64 voxels along x and lists
// the voxels that have triangles, each with its case and its offset within
// its step; an exclusive scan of the step counts gives each step its offset;
// an emit pass, one thread per listed voxel, writes the voxel's triangles.
// The soup order is the reference's: voxels x-fastest, triangles in table
// order.
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
// The count pass writes each step's vertex count; a scan of the step counts
// gives each step its place in the soup.
struct McGrid {
    int rx, ry, rz;
    int64_t lx, ly;   // lattice points per axis in x, y (res + 1)
    int nch;          // 64-voxel chunks per row
    int njb;          // ceil(ry / kJ): row blocks of the count waves
    int nkb;          // ceil(rz / kK): slab blocks of the count waves
    int64_t nwaves;   // count waves: nch * njb * nkb
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

// pass 1: vertices per step and the list of voxels with triangles.  A wave
// takes one 64-voxel chunk of x over kJ rows (y) and kK slabs (z) and walks
// +z: each lattice plane of its kJ + 1 rows is loaded once (the x + 1
// neighbour comes from the next lane; the last lane loads its own), and
// plane z + 1 becomes plane z of the next slab.  A voxel has triangles
// exactly when its case is neither 0 nor 255 (every other case of the table
// has some), so only rows with such a voxel look up vertex counts and sum
// them.  The cases wait in LDS while the wave walks; the wave then reserves
// one list slot per voxel with triangles (one atomic) and, for the rows that
// have them, forms each voxel's offset within its step (a wave prefix sum)
// and writes its entry: step (kStepBits bits), lane (6), case (8), offset
// within the step (10).  The list order depends on timing; each voxel's
// place in the soup is its step's place (the scan of step_verts) plus its
// offset, so the output does not.
constexpr int kJ = 4;              // rows (y) per count wave
constexpr int kK = 8;              // slabs (z) per count wave
constexpr int kRows = kJ * kK;     // step rows per count wave (at most 32: a bit mask)
constexpr int kWaves = kT / 64;
__device__ inline uint64_t pack_entry(int64_t s, int lane, unsigned c, unsigned off) {
    return (uint64_t)s | ((uint64_t)lane << kStepBits) | ((uint64_t)c << (kStepBits + 6)) |
           ((uint64_t)off << (kStepBits + 14));
}
__device__ inline bool has_triangles(unsigned c) { return c != 0u && c != 255u; }
// lattice rows y = j0 .. j0 + kJ (clamped to the lattice) of plane z at x
// (P0) and x + 1 (P1)
// (all loads are issued before the first shuffle waits for one)
__device__ inline void load_plane(const float* __restrict__ lat, const McGrid& G, int z, int j0, int xl, int xlast,
                                  int lane, float (&P0)[kJ + 1], float (&P1)[kJ + 1]) {
    float E[kJ + 1];
#pragma unroll
    for (int r = 0; r <= kJ; ++r) {
        const int64_t base = ((int64_t)z * G.ly + min(j0 + r, G.ry)) * G.lx;
        P0[r] = lat[base + xl];
        E[r] = lat[base + xlast];   // the same point for every lane
    }
#pragma unroll
    for (int r = 0; r <= kJ; ++r) {
        const float nb = __shfl_down(P0[r], 1, 64);
        P1[r] = lane == 63 ? E[r] : nb;
    }
}
__global__ __launch_bounds__(kT) void k_mc_count(const float* __restrict__ lat, McGrid G,
                                                 uint64_t* __restrict__ step_verts, unsigned long long* counts,
                                                 uint64_t* __restrict__ entries) {
    __shared__ uint8_t s_case[kWaves][kRows][64];
    __shared__ uint8_t s_nv[256];   // vertex counts per case, read per lane
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t w = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
    static_assert(kT == 256, "one table entry per thread");
    s_nv[threadIdx.x] = c_nv[threadIdx.x];
    __syncthreads();
    if (w >= G.nwaves) return;   // whole waves only
    const int ch = (int)(w % G.nch);
    const int64_t rest = w / G.nch;
    const int jb = (int)(rest % G.njb), kb = (int)(rest / G.njb);
    const int j0 = jb * kJ, j1 = min(j0 + kJ, G.ry);
    const int k0 = kb * kK, k1 = min(k0 + kK, G.rz);
    const int i = ch * 64 + lane;
    const bool in = i < G.rx;
    // lattice x of this lane (lanes past the row read in bounds and count
    // nothing), and x + 1 of the last lane
    const int xl = min(i, G.rx), xlast = min(ch * 64 + 64, G.rx);
    float C0[kJ + 1], C1[kJ + 1], N0[kJ + 1], N1[kJ + 1];
    load_plane(lat, G, k0, j0, xl, xlast, lane, C0, C1);
    unsigned nval = 0;
    uint32_t rows = 0;   // bit (k - k0) kJ + (j - j0): that step row has a voxel with triangles
    for (int k = k0; k < k1; ++k) {
        load_plane(lat, G, k + 1, j0, xl, xlast, lane, N0, N1);
#pragma unroll
        for (int r = 0; r < kJ; ++r) {
            const int j = j0 + r;
            if (j >= j1) break;   // uniform
            // the reference's corner order: (x,y,z) = (0,0,0) (0,1,0) (1,1,0) (1,0,0) (0,0,1) (0,1,1) (1,1,1) (1,0,1)
            const float v[8] = {C0[r], C0[r + 1], C1[r + 1], C1[r], N0[r], N0[r + 1], N1[r + 1], N1[r]};
            const unsigned c = in ? case8(v, G.iso) : 0u;
            const int row = (k - k0) * kJ + r;
            s_case[wv][row][lane] = (uint8_t)c;
            const uint64_t bal = __ballot(has_triangles(c));
            unsigned sum = 0;
            if (bal) {
                sum = wave_sum(has_triangles(c) ? (unsigned)s_nv[c] : 0u);
                nval += __popcll(bal);
                rows |= 1u << row;
            }
            if (lane == 0) step_verts[((int64_t)k * G.ry + j) * G.nch + ch] = sum;
        }
#pragma unroll
        for (int r = 0; r <= kJ; ++r) {
            C0[r] = N0[r];
            C1[r] = N1[r];
        }
    }
    if (!nval) return;
    unsigned long long slot = 0;
    if (lane == 0) slot = atomicAdd(counts, (unsigned long long)nval);
    slot = __shfl(slot, 0, 64);
    const uint64_t below = (1ull << lane) - 1ull;
    // every lane reads back only the cases it wrote itself
    while (rows) {
        const int row = __builtin_ctz(rows);
        rows &= rows - 1u;
        const unsigned c = s_case[wv][row][lane];
        const bool act = has_triangles(c);
        const unsigned nv = act ? (unsigned)s_nv[c] : 0u;
        const unsigned incl = wave_incl_scan(nv);
        const uint64_t bal = __ballot(act);
        const int k = k0 + row / kJ, j = j0 + row % kJ;
        if (act) entries[slot + __popcll(bal & below)] = pack_entry(((int64_t)k * G.ry + j) * G.nch + ch, lane, c,
                                                                    incl - nv);
        slot += __popcll(bal);
    }
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

// pass 2: one thread per voxel with triangles (a fixed grid walks the list,
// whose length it reads on the device, so the launch needs nothing from the
// host).  A voxel's triangles start at its step's place in the soup
// (step_off: exclusive scan of the step counts) plus its offset within the
// step; no vertex goes at or beyond max_verts, the smaller of the
// reference's bound and the soup buffer.
__device__ inline void emit_voxel(const float* __restrict__ lat, const McGrid& G, double ox, double oy, double oz,
                                  double vx, double vy, double vz, const uint64_t* __restrict__ step_off, uint64_t E,
                                  uint64_t max_verts, float* __restrict__ out) {
    const int64_t s = (int64_t)(E & kStepMask);
    const int lane = (int)((E >> kStepBits) & 63u);
    const unsigned c = (unsigned)((E >> (kStepBits + 6)) & 255u);
    const uint64_t loc = (E >> (kStepBits + 14)) & 1023u;
    const int ch = (int)(s % G.nch);
    const int64_t row = s / G.nch;
    const int j = (int)(row % G.ry), k = (int)(row / G.ry);
    const int i = ch * 64 + lane;
    const unsigned nv = c_nv[c];
    const int64_t base = ((int64_t)k * G.ly + j) * G.lx;
    // voxelToMeshKernel: voxelPos = shift * voxelSize + origin
    const P3 vp{(double)i * vx + ox, (double)j * vy + oy, (double)k * vz + oz};
    double sv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sv[q] = (double)lat[base + (c_cz[q] * G.ly + c_cy[q]) * G.lx + i + c_cx[q]];
    const uint4 tri = *reinterpret_cast<const uint4*>(&c_tri[c][0]);
    const uint64_t first = step_off[s] + loc;
    const double iso = G.iso64;   // MarchingCubes.cu:72-90 works with the caller's binary64 iso
    for (unsigned t = 0; t < nv; t += 3) {
        const uint64_t pos = first + t;
        if (pos + 3 > max_verts) break;
        for (unsigned q = 0; q < 3; ++q) {
            const unsigned m = t + q;
            const unsigned word = (m >> 2) == 0 ? tri.x : (m >> 2) == 1 ? tri.y : (m >> 2) == 2 ? tri.z : tri.w;
            const int ed = (int)(signed char)(word >> (8 * (m & 3)));
            const int qa = edge_a(ed), qb = edge_b(ed);
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
                                                const uint64_t* __restrict__ step_off,
                                                const uint64_t* __restrict__ entries,
                                                const unsigned long long* __restrict__ counts, uint64_t max_verts,
                                                float* __restrict__ out) {
    const int64_t nval = (int64_t)counts[0];
    for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < nval; e += (int64_t)gridDim.x * kT)
        emit_voxel(lat, G, ox, oy, oz, vx, vy, vz, step_off, entries[e], max_verts, out);
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
    G.njb = (res[1] + kJ - 1) / kJ;
    G.nkb = (res[2] + kK - 1) / kK;
    G.nwaves = (int64_t)G.nch * G.njb * G.nkb;
    G.nsteps = (int64_t)res[1] * res[2] * G.nch;
    G.iso = (float)iso;
    G.iso64 = iso;
    s->mc_step.ensure(G.nsteps);
    ITS_REQUIRE(G.nsteps < (int64_t)kStepMask, ITS_E_RANGE, "marching-cubes grid has too many rows");
    s->mc_list.ensure(nvox);   // at most one entry per voxel
    s->mc_valid.ensure(2);
    ITS_HIP(hipMemsetAsync(s->mc_valid.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_mc_count, dim3((unsigned)((G.nwaves * 64 + kT - 1) / kT)), dim3(kT), 0, st,
                       (const float*)s->mc_lat.p, G, s->mc_step.p, s->mc_valid.p, s->mc_list.p);
    ITS_LAUNCH_CHECK();
    // the scan's total lands beside the list length
    static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "counter width");
    scan_exclusive_u64_async(s->mc_step.p, s->mc_step.p, G.nsteps, st, (uint64_t*)s->mc_valid.p + 1);
    const uint64_t max_verts = 18ull * (uint64_t)nvox;   // maxVerts = nVoxels * 18
    const unsigned emit_grid =
        (unsigned)std::min<int64_t>((nvox + kT - 1) / kT, (int64_t)s->ctx->prop.multiProcessorCount * 16);
    auto emit = [&]() {
        hipLaunchKernelGGL(k_mc_emit, dim3(emit_grid), dim3(kT), 0, st, (const float*)s->mc_lat.p, G, origin[0],
                           origin[1], origin[2], vs[0], vs[1], vs[2], (const uint64_t*)s->mc_step.p,
                           (const uint64_t*)s->mc_list.p, (const unsigned long long*)s->mc_valid.p,
                           std::min<uint64_t>(max_verts, s->mc_soup.n / 3), s->mc_soup.p);
        ITS_LAUNCH_CHECK();
    };
    // the soup buffer only grows: when it already holds the last soup, the
    // emit pass follows the scan with no host round trip, and it is run again
    // only when this soup needs more room
    const bool want = h_tri != nullptr;
    if (want && s->mc_soup.n >= 3) emit();
    ITS_HIP(hipEventRecord(ev[2], st));
    unsigned long long back[2];
    ITS_HIP(hipMemcpyAsync(back, s->mc_valid.p, sizeof back, hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    const int64_t nval = (int64_t)back[0];
    const uint64_t total = back[1];
    S.num_verts = (int64_t)total;
    if (want && total) {
        ITS_REQUIRE((int64_t)total <= capacity, ITS_E_RANGE, "triangle buffer too small");
        if (3 * total > s->mc_soup.n) {
            s->mc_soup.ensure(3 * total);
            emit();
            ITS_HIP(hipEventRecord(ev[2], st));
        }
        ITS_HIP(hipMemcpyAsync(h_tri, s->mc_soup.p, 3 * total * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    ITS_HIP(hipStreamSynchronize(st));
    S.num_valid_voxels = nval;
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
