// Marching cubes over the fitted field for gfx950.
//
// Replaces MC::marching_cubes (reference MarchingCubes.cu:672-714) and its
// kernels prepareVoxelCornerKernel, determineVoxelKernel, compactVoxels and
// voxelToMeshKernel (124-377).
//
// The reference evaluates f at 8 corners per voxel.  Each corner is a brute
// force sum over all n basis functions, batched over 1024 streams.  Here f is
// evaluated once per lattice point ((res+1)^3 points) with the octree field
// walk.  A classify pass writes the cube case and vertex count, an exclusive
// scan gives output offsets, and an emit pass writes the triangle soup.  The
// soup order is the reference's: voxels x-fastest, triangles in table order.
// The 1e-6 snaps of vertexLerp (MarchingCubes.cu:72-90) are kept, and so is
// the lerp fma(t, v1, fma(-t, v0, v0)).
#include <algorithm>
#include <chrono>
#include <cmath>

#include "internal.hpp"
#include "mc_tables.inc"

namespace its {
void eval_lattice(its_svo s, const double o[3], const double vs[3], const int64_t dims[3], float* out, hipStream_t st);

namespace {
constexpr int kT = 256;
__constant__ signed char c_tri[256][16];
__constant__ unsigned char c_nv[256];

// reference corner order (MarchingCubes.cu:142-149): (x,y,z) offsets
__constant__ unsigned char c_cx[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__constant__ unsigned char c_cy[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ unsigned char c_cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

struct McDims {
    int rx, ry, rz;
    int64_t lx, ly;   // lattice dims in x, y (res+1)
    float iso;
};

__device__ inline float corner_val(const float* lat, const McDims& D, int64_t i, int64_t j, int64_t k, int c) {
    return lat[((k + c_cz[c]) * D.ly + (j + c_cy[c])) * D.lx + (i + c_cx[c])];
}

__global__ void k_classify(int64_t nvox, const float* lat, McDims D, uint8_t* cube, uint32_t* cnt,
                           unsigned long long* nvalid) {
    __shared__ unsigned int blk_valid;
    if (threadIdx.x == 0) blk_valid = 0;
    __syncthreads();
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nvox) {
        int64_t i = v % D.rx, j = (v / D.rx) % D.ry, k = v / ((int64_t)D.rx * D.ry);
        unsigned c = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) c |= (unsigned)(corner_val(lat, D, i, j, k, q) < D.iso) << q;
        cube[v] = (uint8_t)c;
        unsigned n = c_nv[c];
        cnt[v] = n;
        if (n) atomicAdd(&blk_valid, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && blk_valid) atomicAdd(nvalid, (unsigned long long)blk_valid);
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

__constant__ unsigned char c_e0[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
__constant__ unsigned char c_e1[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

__global__ void k_emit(int64_t nvox, const float* lat, McDims D, double ox, double oy, double oz, double vx, double vy,
                       double vz, const uint8_t* cube, const uint32_t* offs, uint64_t max_verts, float* out) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    unsigned c = cube[v];
    unsigned nv = c_nv[c];
    if (!nv) return;
    int64_t i = v % D.rx, j = (v / D.rx) % D.ry, k = v / ((int64_t)D.rx * D.ry);
    // voxelToMeshKernel: voxelPos = shift * voxelSize + origin
    P3 vp{(double)i * vx + ox, (double)j * vy + oy, (double)k * vz + oz};
    P3 cr[8];
    double sv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        cr[q] = P3{vp.x + (c_cx[q] ? vx : 0.0), vp.y + (c_cy[q] ? vy : 0.0), vp.z + (c_cz[q] ? vz : 0.0)};
        sv[q] = (double)corner_val(lat, D, i, j, k, q);
    }
    const double iso = (double)D.iso;
    uint64_t base = offs[v];
    for (unsigned t = 0; t < nv; t += 3) {
        uint64_t pos = base + t;
        if (pos >= max_verts - 3) break;
        for (int q = 0; q < 3; ++q) {
            int e = c_tri[c][t + q];
            P3 p = vlerp(cr[c_e0[e]], cr[c_e1[e]], sv[c_e0[e]], sv[c_e1[e]], iso);
            out[3 * (pos + q) + 0] = (float)p.x;
            out[3 * (pos + q) + 1] = (float)p.y;
            out[3 * (pos + q) + 2] = (float)p.z;
        }
    }
}

void upload_tables() {
    static int dev_done = -1;
    int dev;
    ITS_HIP(hipGetDevice(&dev));
    if (dev_done == dev) return;
    unsigned char nv[256];
    for (int c = 0; c < 256; ++c) {
        int n = 0;
        while (n < 16 && kMcTri[c][n] >= 0) ++n;
        nv[c] = (unsigned char)n;
    }
    ITS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tri), kMcTri, sizeof kMcTri));
    ITS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_nv), nv, sizeof nv));
    dev_done = dev;
}
}  // namespace

// The MC grid is the octree's level-0 lattice when it has R cells per axis
// over the octree box (the case of ThinShells::mcVisualization at the octree
// resolution).  Its lattice values are then the dense prolongation cascade
// down to level 0.  Otherwise every lattice point is evaluated on its own.
// The octree lattice spacing is width.x / R on all axes, and the MC spacing is
// width[a] / res; the float casts of the box padding make these differ by
// O(1e-8) relative, a difference in f far below binary32 resolution.
static void lattice_values(its_svo s, const double origin[3], const double width[3], const int res[3],
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
    DBuf<float> lat;
    lattice_values(s, origin, width, res, lat, st);
    ITS_HIP(hipMemcpyAsync(h_vals, lat.p, np * sizeof(float), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
}

void mc_extract(its_svo s, const double origin[3], const double width[3], const int res[3], double iso, float* h_tri,
                int64_t capacity, its_mc_stats* stats) {
    hipStream_t st = s->ctx->stream;
    upload_tables();
    ITS_REQUIRE(res[0] > 0 && res[1] > 0 && res[2] > 0, ITS_E_INVALID, "resolution must be positive");
    its_mc_stats S{};
    auto t0 = std::chrono::steady_clock::now();
    double vs[3] = {width[0] / res[0], width[1] / res[1], width[2] / res[2]};   // MarchingCubes.cu:682
    int64_t dims[3] = {res[0] + 1, res[1] + 1, res[2] + 1};
    int64_t nvox = (int64_t)res[0] * res[1] * res[2];
    DBuf<float> lat;
    lattice_values(s, origin, width, res, lat, st);
    ITS_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    McDims D{res[0], res[1], res[2], dims[0], dims[1], (float)iso};
    DBuf<uint8_t> cube(nvox);
    DBuf<uint32_t> cnt(nvox);
    DBuf<unsigned long long> nvalid(1);
    ITS_HIP(hipMemsetAsync(nvalid.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_classify, dim3(blocks_for(nvox, kT)), dim3(kT), 0, st, nvox, lat.p, D, cube.p, cnt.p,
                       nvalid.p);
    ITS_LAUNCH_CHECK();
    uint64_t total = scan_exclusive_u32(cnt.p, cnt.p, nvox, st);
    S.num_verts = (int64_t)total;
    S.num_valid_voxels = (int64_t)read1(nvalid.p, st);
    uint64_t max_verts = 18ull * (uint64_t)nvox;   // maxVerts = nVoxels * 18
    if (h_tri && total) {
        ITS_REQUIRE((int64_t)total <= capacity, ITS_E_RANGE, "triangle buffer too small");
        DBuf<float> out(3 * total);
        hipLaunchKernelGGL(k_emit, dim3(blocks_for(nvox, kT)), dim3(kT), 0, st, nvox, lat.p, D, origin[0], origin[1],
                           origin[2], vs[0], vs[1], vs[2], cube.p, cnt.p, max_verts, out.p);
        ITS_LAUNCH_CHECK();
        ITS_HIP(hipMemcpyAsync(h_tri, out.p, 3 * total * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    ITS_HIP(hipStreamSynchronize(st));
    auto t2 = std::chrono::steady_clock::now();
    S.seconds_eval = std::chrono::duration<double>(t1 - t0).count();
    S.seconds_extract = std::chrono::duration<double>(t2 - t1).count();
    if (stats) *stats = S;
}

}  // namespace its
