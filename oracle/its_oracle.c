/*
 * its_oracle.c -- FP64 CPU restatement of the ITS hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library; the product never does.
 *
 * Citations are to the reference sources (src/...).  The restatement follows
 * the reference's algorithms step by step; where the reference relies on
 * undefined behaviour or on ordering that cannot be reproduced (thrust sort
 * stability, Eigen SIMD reduction order) the convention chosen here is stated
 * in the comment of the function concerned.
 */
#include "its_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Morton codec: x -> bit 3i, y -> bit 3i+1, z -> bit 3i+2                    */
/* (src/MortonLUT.h:7-11 masks, 261-312 encode, 314-343 decode)               */
/* ------------------------------------------------------------------------ */
static uint32_t spread3(uint32_t v) {
    /* bit i of v -> bit 3i (v < 2^10) */
    uint32_t r = 0;
    for (int i = 0; i < 10; ++i) r |= ((v >> i) & 1u) << (3 * i);
    return r;
}
static uint32_t compact3(uint32_t m) {
    uint32_t r = 0;
    for (int i = 0; i < 11; ++i) r |= ((m >> (3 * i)) & 1u) << i;
    return r;
}

uint32_t ora_morton_encode(uint32_t x, uint32_t y, uint32_t z) {
    /* mortonEncode_LUT(uint16,uint16,uint16) (MortonLUT.h:292-312): two
     * byte-LUT rounds, hi byte shifted by 24, truncated to 32 bits. */
    uint32_t hi = spread3((x >> 8) & 0xFF) | (spread3((y >> 8) & 0xFF) << 1) | (spread3((z >> 8) & 0xFF) << 2);
    uint32_t lo = spread3(x & 0xFF) | (spread3(y & 0xFF) << 1) | (spread3(z & 0xFF) << 2);
    return (uint32_t)((hi << 24) | lo);
}

void ora_morton_decode(uint32_t m, uint32_t* x, uint32_t* y, uint32_t* z) {
    /* morton3D_32_decode (MortonLUT.h:314-343): 4 rounds of 9-bit chunks */
    *x = compact3(m) & 0xFFFF;
    *y = compact3(m >> 1) & 0xFFFF;
    *z = compact3(m >> 2) & 0xFFFF;
}

void ora_morton_tables(uint32_t enc_x[256], uint32_t enc_y[256], uint32_t enc_z[256],
                       uint8_t dec_x[512], uint8_t dec_y[512], uint8_t dec_z[512]) {
    for (uint32_t v = 0; v < 256; ++v) {
        enc_x[v] = spread3(v);
        enc_y[v] = spread3(v) << 1;
        enc_z[v] = spread3(v) << 2;
    }
    for (uint32_t v = 0; v < 512; ++v) {
        dec_x[v] = (uint8_t)(compact3(v) & 7);
        dec_y[v] = (uint8_t)(compact3(v >> 1) & 7);
        dec_z[v] = (uint8_t)(compact3(v >> 2) & 7);
    }
}

/* neighbour index k = (dx+1) + 3(dy+1) + 9(dz+1); child key = x | y<<1 | z<<2.
 * Closed forms of neighbor_LUTparent / neighbor_LUTchild / d_vertSharedLUT /
 * d_edgeSharedLUT (src/cuAcc/SVO.cu:242-348). */
void ora_neighbor_luts(int16_t lut_parent[8 * 27], int16_t lut_child[8 * 27],
                       int16_t vert_shared[64], int16_t edge_shared[48]) {
    for (int key = 0; key < 8; ++key) {
        int c[3] = {key & 1, (key >> 1) & 1, (key >> 2) & 1};
        for (int k = 0; k < 27; ++k) {
            int d[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
            int pk = 0, ck = 0, mul = 1;
            for (int a = 0; a < 3; ++a) {
                int t = c[a] + d[a];             /* -1..2 */
                int pd = (t < 0) ? -1 : (t > 1 ? 1 : 0);
                pk += (pd + 1) * mul;
                ck |= (t & 1) << a;
                mul *= 3;
            }
            lut_parent[key * 27 + k] = (int16_t)pk;
            lut_child[key * 27 + k] = (int16_t)ck;
        }
    }
    for (int i = 0; i < 8; ++i) {
        int c[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
        for (int j = 0; j < 8; ++j) {
            int s[3] = {j & 1, (j >> 1) & 1, (j >> 2) & 1};
            int k = 0, mul = 1;
            for (int a = 0; a < 3; ++a) { k += (c[a] - 1 + s[a] + 1) * mul; mul *= 3; }
            vert_shared[i * 8 + j] = (int16_t)k;
        }
    }
    /* edges 0-2 2-3 1-3 0-1; 4-6 6-7 5-7 4-5; 0-4 2-6 3-7 1-5 (SVO.cu:326-329):
     * the 4 nodes sharing an edge, in ascending neighbour index. */
    static const int edge_corners[12][2] = {{0, 2}, {2, 3}, {1, 3}, {0, 1}, {4, 6}, {6, 7},
                                            {5, 7}, {4, 5}, {0, 4}, {2, 6}, {3, 7}, {1, 5}};
    for (int e = 0; e < 12; ++e) {
        int a = edge_corners[e][0], b = edge_corners[e][1];
        int ca[3] = {a & 1, (a >> 1) & 1, (a >> 2) & 1};
        int cb[3] = {b & 1, (b >> 1) & 1, (b >> 2) & 1};
        int ks[8], nk = 0;
        /* nodes sharing the edge: along the edge axis offset 0; across the
         * other two axes offset in {c-1, c} */
        for (int j = 0; j < 8; ++j) {
            int s[3] = {j & 1, (j >> 1) & 1, (j >> 2) & 1};
            int ok = 1, k = 0, mul = 1;
            for (int ax = 0; ax < 3; ++ax) {
                int d;
                if (ca[ax] != cb[ax]) { if (s[ax]) { ok = 0; break; } d = 0; }
                else d = ca[ax] - 1 + s[ax];
                k += (d + 1) * mul;
                mul *= 3;
            }
            if (ok) ks[nk++] = k;
        }
        for (int i = 0; i < nk; ++i)
            for (int j = i + 1; j < nk; ++j)
                if (ks[j] < ks[i]) { int t = ks[i]; ks[i] = ks[j]; ks[j] = t; }
        for (int i = 0; i < 4; ++i) edge_shared[e * 4 + i] = (int16_t)ks[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Normalisation (src/BaseModel.cpp:54-92) and grid box (24-51)              */
/* ------------------------------------------------------------------------ */
void ora_model_to_unit_cube(double* V, int64_t nV, double scale_factor) {
    double bmin[3], bmax[3];
    for (int a = 0; a < 3; ++a) { bmin[a] = DBL_MAX; bmax[a] = -DBL_MAX; }
    for (int64_t i = 0; i < nV; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = V[3 * i + a];
            if (v < bmin[a]) bmin[a] = v;
            if (v > bmax[a]) bmax[a] = v;
        }
    /* calcUnitCubeTransformMatrix: scale = max extent */
    double scale = bmax[0] - bmin[0];
    for (int d = 1; d < 3; ++d) { double e = bmax[d] - bmin[d]; if (e > scale) scale = e; }
    scale *= scale_factor;
    double center[3];
    for (int a = 0; a < 3; ++a) center[a] = 0.5 * bmax[a] + 0.5 * bmin[a];
    for (int a = 0; a < 3; ++a) center[a] -= scale / 2;
    double inv = 1. / scale;
    /* row += (-center); row = row * diag(1/scale) */
    for (int64_t i = 0; i < nV; ++i)
        for (int a = 0; a < 3; ++a) V[3 * i + a] = (V[3 * i + a] + (-center[a])) * inv;
}

void ora_uniform_bbox(const double* V, int64_t nV, double origin[3], double end[3], double width[3]) {
    double minV[3], maxV[3];
    for (int a = 0; a < 3; ++a) { minV[a] = DBL_MAX; maxV[a] = -DBL_MAX; }
    for (int64_t i = 0; i < nV; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = V[3 * i + a];
            if (v < minV[a]) minV[a] = v;
            if (v > maxV[a]) maxV[a] = v;
        }
    double lengths[3];
    for (int a = 0; a < 3; ++a) { origin[a] = minV[a]; end[a] = maxV[a]; lengths[a] = maxV[a] - minV[a]; }
    /* float max_length = fmaxf(lx, fmaxf(ly, lz)) with the C float fmaxf */
    float max_length = fmaxf((float)lengths[0], fmaxf((float)lengths[1], (float)lengths[2]));
    for (int i = 0; i < 3; ++i) {
        if ((double)max_length == lengths[i]) continue;
        float delta = (float)((double)max_length - lengths[i]);
        origin[i] = minV[i] - (double)(delta / 2.0f);
        end[i] = maxV[i] + (double)(delta / 2.0f);
    }
    double eps[3];
    for (int a = 0; a < 3; ++a) eps[a] = (end[a] - origin[a]) / 11;
    for (int a = 0; a < 3; ++a) { origin[a] -= eps[a]; end[a] += eps[a]; width[a] = end[a] - origin[a]; }
}

/* ------------------------------------------------------------------------ */
/* Voxelisation: surfaceVoxelize (src/cuAcc/SVO.cu:366-493)                  */
/* ------------------------------------------------------------------------ */
static inline double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
static inline double dot2(double a0, double a1, double b0, double b1) { return a0 * b0 + a1 * b1; }
static inline double dmin(double a, double b) { return a < b ? a : b; } /* CUDAMath.hpp:131 */
static inline double dmax(double a, double b) { return a > b ? a : b; } /* CUDAMath.hpp:136 */
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

typedef void (*voxel_sink)(void* ctx, uint32_t morton);

static void voxelize_tri(const double* tri, const double go[3], const double unit[3], int R,
                         voxel_sink sink, void* ctx) {
    double v0[3], v1[3], v2[3], e0[3], e1[3], e2[3], n[3];
    for (int a = 0; a < 3; ++a) {
        v0[a] = tri[a] - go[a];
        v1[a] = tri[3 + a] - go[a];
        v2[a] = tri[6 + a] - go[a];
    }
    for (int a = 0; a < 3; ++a) { e0[a] = v1[a] - v0[a]; e1[a] = v2[a] - v1[a]; e2[a] = v0[a] - v2[a]; }
    /* n = e0.cross(e1).normalized() (Eigen: n / sqrt(squaredNorm) if > 0) */
    n[0] = e0[1] * e1[2] - e0[2] * e1[1];
    n[1] = e0[2] * e1[0] - e0[0] * e1[2];
    n[2] = e0[0] * e1[1] - e0[1] * e1[0];
    double z2 = dot3(n, n);
    if (z2 > 0.0) {
        double s = sqrt(z2);
        n[0] = n[0] / s; n[1] = n[1] / s; n[2] = n[2] / s;
    }
    double bo[3], be[3];
    for (int a = 0; a < 3; ++a) {
        bo[a] = dmin(v0[a], dmin(v1[a], v2[a]));
        be[a] = dmax(v0[a], dmax(v1[a], v2[a]));
    }
    int gmax = R - 1;
    int go_i[3], ge_i[3];
    for (int a = 0; a < 3; ++a) {
        go_i[a] = imax(0, imin((int)(bo[a] / unit[a]), gmax));   /* Vector3i(double) truncates */
        ge_i[a] = imax(0, imin((int)(be[a] / unit[a]), gmax));
    }
    double c[3] = {0.0, 0.0, 0.0};
    if (n[0] > 0.0) c[0] = unit[0];
    if (n[1] > 0.0) c[1] = unit[1];
    if (n[2] > 0.0) c[2] = unit[2];
    double cv[3], dv[3];
    for (int a = 0; a < 3; ++a) { cv[a] = c[a] - v0[a]; dv[a] = (unit[a] - c[a]) - v0[a]; }
    double d1 = dot3(n, cv);
    double d2 = dot3(n, dv);

    /* XY */
    double nxy[3][2] = {{-1.0 * e0[1], e0[0]}, {-1.0 * e1[1], e1[0]}, {-1.0 * e2[1], e2[0]}};
    if (n[2] < 0.0)
        for (int i = 0; i < 3; ++i) { nxy[i][0] = -nxy[i][0]; nxy[i][1] = -nxy[i][1]; }
    const double* vs[3] = {v0, v1, v2};
    double dxy[3], dyz[3], dzx[3];
    for (int i = 0; i < 3; ++i)
        dxy[i] = (-1.0 * dot2(nxy[i][0], nxy[i][1], vs[i][0], vs[i][1])) + dmax(0.0, unit[0] * nxy[i][0]) +
                 dmax(0.0, unit[1] * nxy[i][1]);
    /* YZ */
    double nyz[3][2] = {{-1.0 * e0[2], e0[1]}, {-1.0 * e1[2], e1[1]}, {-1.0 * e2[2], e2[1]}};
    if (n[0] < 0.0)
        for (int i = 0; i < 3; ++i) { nyz[i][0] = -nyz[i][0]; nyz[i][1] = -nyz[i][1]; }
    for (int i = 0; i < 3; ++i)
        dyz[i] = (-1.0 * dot2(nyz[i][0], nyz[i][1], vs[i][1], vs[i][2])) + dmax(0.0, unit[1] * nyz[i][0]) +
                 dmax(0.0, unit[2] * nyz[i][1]);
    /* ZX */
    double nzx[3][2] = {{-1.0 * e0[0], e0[2]}, {-1.0 * e1[0], e1[2]}, {-1.0 * e2[0], e2[2]}};
    if (n[1] < 0.0)
        for (int i = 0; i < 3; ++i) { nzx[i][0] = -nzx[i][0]; nzx[i][1] = -nzx[i][1]; }
    for (int i = 0; i < 3; ++i)
        dzx[i] = (-1.0 * dot2(nzx[i][0], nzx[i][1], vs[i][2], vs[i][0])) + dmax(0.0, unit[2] * nzx[i][0]) +
                 dmax(0.0, unit[0] * nzx[i][1]);

    for (int z = go_i[2]; z <= ge_i[2]; ++z)
        for (int y = go_i[1]; y <= ge_i[1]; ++y)
            for (int x = go_i[0]; x <= ge_i[0]; ++x) {
                double p[3] = {(double)x * unit[0], (double)y * unit[1], (double)z * unit[2]};
                double nDOTp = dot3(n, p);
                if ((nDOTp + d1) * (nDOTp + d2) > 0.0) continue;
                int ok = 1;
                for (int i = 0; i < 3 && ok; ++i) if ((dot2(nxy[i][0], nxy[i][1], p[0], p[1]) + dxy[i]) < 0.0) ok = 0;
                for (int i = 0; i < 3 && ok; ++i) if ((dot2(nyz[i][0], nyz[i][1], p[1], p[2]) + dyz[i]) < 0.0) ok = 0;
                for (int i = 0; i < 3 && ok; ++i) if ((dot2(nzx[i][0], nzx[i][1], p[2], p[0]) + dzx[i]) < 0.0) ok = 0;
                if (!ok) continue;
                sink(ctx, ora_morton_encode((uint32_t)x, (uint32_t)y, (uint32_t)z));
            }
}

typedef struct { uint8_t* dense; } dense_ctx;
static void dense_sink(void* c, uint32_t m) { ((dense_ctx*)c)->dense[m] = 1; }

/* Dense occupancy array indexed by Morton code, as the reference (SVO.cu:697,
 * gridCNodeSize = morton(R-1,R-1,R-1)+1), compacted in Morton order. */
static uint32_t* voxelize_sorted(const double* tris, int64_t T, const double origin[3], const double width[3],
                                 int R, int64_t* count) {
    double unit[3] = {width[0] / R, width[1] / R, width[2] / R};
    uint64_t grid = (uint64_t)ora_morton_encode(R - 1, R - 1, R - 1) + 1;
    uint8_t* dense = (uint8_t*)calloc(grid, 1);
    dense_ctx ctx = {dense};
    /* threads write 1 to the same bytes: a benign race with identical values,
     * as atomicExch in the reference (SVO.cu:487) */
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < T; ++t) voxelize_tri(tris + 9 * t, origin, unit, R, dense_sink, &ctx);
    int64_t cnt = 0;
    for (uint64_t m = 0; m < grid; ++m) cnt += dense[m];
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * (cnt ? cnt : 1));
    int64_t k = 0;
    for (uint64_t m = 0; m < grid; ++m)
        if (dense[m]) keys[k++] = (uint32_t)m;
    free(dense);
    *count = cnt;
    return keys;
}

int64_t ora_voxelize(const double* tris, int64_t T, const double origin[3], const double width[3], int R,
                     uint32_t* keys_out, int64_t cap) {
    int64_t cnt;
    uint32_t* keys = voxelize_sorted(tris, T, origin, width, R, &cnt);
    if (cnt > cap) { free(keys); return -cnt; }
    memcpy(keys_out, keys, sizeof(uint32_t) * cnt);
    free(keys);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Sparse voxel octree: createOctree (SVO.cu:674-856), findNeighbors (858-911),
 * determineNodeVertex + sort/unique (913-963, 1010-1096)                    */
/* ------------------------------------------------------------------------ */
struct ora_svo {
    int depth;
    int64_t num_voxels;
    int64_t num_nodes;
    int64_t depth_num_nodes[ORA_MAX_LEVELS];
    int64_t esum_nodes[ORA_MAX_LEVELS + 1];
    uint32_t* morton;
    uint8_t* is_leaf;
    double* origin;
    double* width;
    uint32_t* parent;
    uint32_t* childs;
    uint32_t* neighbors;
    int64_t num_verts;
    int64_t esum_verts[ORA_MAX_LEVELS + 1];
    double* vxyz;
    uint32_t* vowner;
    double grid_origin[3];
    double voxel_width;
    int R;
};

static int64_t lower_bound_u32(const uint32_t* a, int64_t n, uint32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) { int64_t mid = (lo + hi) / 2; if (a[mid] < v) lo = mid + 1; else hi = mid; }
    return lo;
}

/* fuzzy 1e-9 lexicographic order of std::less<V3d> (BasicDataType.h:44-56) */
static inline int less_point(const double* a, const double* b) {
    for (int i = 0; i < 3; ++i) {
        if (fabs(a[i] - b[i]) < 1e-9) continue;
        if (a[i] < b[i]) return 1;
        else if (a[i] > b[i]) return -1;
    }
    return 0;
}

typedef struct { double p[3]; uint32_t owner; } vrec;

static int vrec_less(const vrec* a, const vrec* b) { /* sortVert (SVO.cu:1024-1031) */
    int t = less_point(a->p, b->p);
    if (t == 0) return a->owner < b->owner;
    return t == 1;
}

/* stable merge sort (thrust's merge sort for user comparators is stable; the
 * survivor of each duplicate group is then the copy of the lowest-indexed
 * node, i.e. the owner) */
static void vrec_sort(vrec* a, int64_t n) {
    vrec* tmp = (vrec*)malloc(sizeof(vrec) * (n ? n : 1));
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (vrec_less(&a[j], &a[i])) tmp[k++] = a[j++];
                else tmp[k++] = a[i++];
            }
            while (i < mid) tmp[k++] = a[i++];
            while (j < hi) tmp[k++] = a[j++];
        }
        memcpy(a, tmp, sizeof(vrec) * n);
    }
    free(tmp);
}

/* Eigen isApprox(b, 1e-9): ||a-b||^2 <= prec^2 * min(||a||^2, ||b||^2) */
static int approx_eq(const double* a, const double* b) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    double na = dot3(a, a), nb = dot3(b, b);
    return dot3(d, d) <= 1e-9 * 1e-9 * (na < nb ? na : nb);
}

static inline void node_origin(const double go[3], double w, uint32_t morton, double out[3]) {
    uint32_t x, y, z;
    ora_morton_decode(morton, &x, &y, &z);
    /* gridOrigin + width * Vector3d(x,y,z) (SVO.cu:587,625,668), not fused */
    out[0] = go[0] + w * (double)x;
    out[1] = go[1] + w * (double)y;
    out[2] = go[2] + w * (double)z;
}

ora_svo* ora_svo_build(const double* tris, int64_t T, const double origin[3], const double width[3], int R) {
    ora_svo* s = (ora_svo*)calloc(1, sizeof(ora_svo));
    s->R = R;
    memcpy(s->grid_origin, origin, sizeof(double) * 3);
    double unit_node_width = width[0] / R;   /* unitVoxelSize.x() (SVO.cu:683-686) */
    s->voxel_width = unit_node_width;

    int64_t nk;
    uint32_t* K = voxelize_sorted(tris, T, origin, width, R, &nk);
    s->num_voxels = nk;

    /* level loop (SVO.cu:717-815) building level arrays, then concatenation */
    uint32_t* lv_morton[ORA_MAX_LEVELS];
    uint8_t* lv_leaf[ORA_MAX_LEVELS];
    uint32_t* lv_childbase[ORA_MAX_LEVELS];  /* local index (in level-1) of first child or NONE */
    uint32_t* lv_parent_local[ORA_MAX_LEVELS];
    double lv_width[ORA_MAX_LEVELS];
    int64_t lv_n[ORA_MAX_LEVELS];
    int depth = 0;
    double w = unit_node_width;
    uint32_t* prevK = NULL;   /* occupied keys of the level below */
    int64_t prevnk = 0;
    while (nk > 0 && depth < ORA_MAX_LEVELS) {
        int d = depth++;
        int64_t numNodes;
        uint32_t* P = NULL;
        int64_t np = 0;
        if (nk > 1) {
            P = (uint32_t*)malloc(sizeof(uint32_t) * nk);
            for (int64_t i = 0; i < nk; ++i) {
                uint32_t p = (K[i] >> 3) & 0xfffffff;
                if (np == 0 || P[np - 1] != p) P[np++] = p;
            }
            numNodes = 8 * np;
        } else {
            /* single occupied cell: the reference creates one node whose
             * address (stale scan value + key) is out of bounds unless the key
             * is 0; this restatement places it at index 0 with its own code. */
            numNodes = 1;
        }
        lv_n[d] = numNodes;
        lv_width[d] = w;
        uint32_t* mort = (uint32_t*)malloc(sizeof(uint32_t) * numNodes);
        uint8_t* leaf = (uint8_t*)malloc(numNodes);
        uint32_t* cb = (uint32_t*)malloc(sizeof(uint32_t) * numNodes);
        uint32_t* pl = (uint32_t*)malloc(sizeof(uint32_t) * numNodes);
        for (int64_t i = 0; i < numNodes; ++i) { leaf[i] = 1; cb[i] = ORA_NONE; pl[i] = ORA_NONE; }
        if (nk > 1) {
            /* createNode_1/2 + createRemainNode: octet r holds the children
             * of occupied parent P[r], node 8r+key has Morton 8*P[r]+key */
            for (int64_t r = 0; r < np; ++r)
                for (int i = 0; i < 8; ++i) mort[8 * r + i] = P[r] * 8 + (uint32_t)i;
            int64_t r = 0;
            for (int64_t i = 0; i < nk; ++i) {
                uint32_t p = K[i] >> 3;
                while (P[r] != p) ++r;
                int64_t addr = 8 * r + (K[i] & 7);
                if (d > 0) { leaf[addr] = 0; cb[addr] = (uint32_t)(8 * i); }
            }
        } else {
            mort[0] = K[0];
            if (d > 0) { leaf[0] = 0; cb[0] = 0; }
        }
        /* children's parent (local index at this level) */
        if (d > 0) {
            for (int64_t i = 0; i < numNodes; ++i)
                if (cb[i] != ORA_NONE)
                    for (int c = 0; c < 8; ++c) lv_parent_local[d - 1][cb[i] + c] = (uint32_t)i;
        }
        lv_morton[d] = mort;
        lv_leaf[d] = leaf;
        lv_childbase[d] = cb;
        lv_parent_local[d] = pl;
        free(prevK);
        prevK = K;
        prevnk = nk;
        if (nk == 1) { free(P); break; }
        K = P;
        nk = np;
        w *= 2.0;
    }
    free(prevK);
    (void)prevnk;
    s->depth = depth;

    /* concatenate finest -> root */
    int64_t total = 0;
    for (int d = 0; d < depth; ++d) {
        s->depth_num_nodes[d] = lv_n[d];
        s->esum_nodes[d] = total;
        total += lv_n[d];
    }
    s->esum_nodes[depth] = total;
    s->num_nodes = total;
    s->morton = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
    s->is_leaf = (uint8_t*)malloc(total ? total : 1);
    s->origin = (double*)malloc(sizeof(double) * 3 * (total ? total : 1));
    s->width = (double*)malloc(sizeof(double) * (total ? total : 1));
    s->parent = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
    s->childs = (uint32_t*)malloc(sizeof(uint32_t) * 8 * (total ? total : 1));
    s->neighbors = (uint32_t*)malloc(sizeof(uint32_t) * 27 * (total ? total : 1));
    for (int d = 0; d < depth; ++d) {
        int64_t off = s->esum_nodes[d];
        for (int64_t i = 0; i < lv_n[d]; ++i) {
            int64_t g = off + i;
            s->morton[g] = lv_morton[d][i];
            s->is_leaf[g] = lv_leaf[d][i];
            s->width[g] = lv_width[d];
            node_origin(origin, lv_width[d], lv_morton[d][i], &s->origin[3 * g]);
            s->parent[g] = (lv_parent_local[d][i] == ORA_NONE) ? ORA_NONE
                                                                : (uint32_t)(s->esum_nodes[d + 1] + lv_parent_local[d][i]);
            for (int c = 0; c < 8; ++c)
                s->childs[8 * g + c] =
                    (lv_childbase[d][i] == ORA_NONE) ? ORA_NONE : (uint32_t)(s->esum_nodes[d - 1] + lv_childbase[d][i] + c);
            for (int k = 0; k < 27; ++k) s->neighbors[27 * g + k] = ORA_NONE;
        }
        free(lv_morton[d]); free(lv_leaf[d]); free(lv_childbase[d]); free(lv_parent_local[d]);
    }

    /* neighbours top-down (SVO.cu:892-911) */
    if (depth >= 1) {
        int16_t lp[8 * 27], lc[8 * 27], vs[64], es[48];
        ora_neighbor_luts(lp, lc, vs, es);
        int64_t top = total - 1;
        s->neighbors[27 * top + 13] = (uint32_t)top;
        for (int d = depth - 2; d >= 0; --d) {
            int64_t off = s->esum_nodes[d];
#pragma omp parallel for
            for (int64_t i = 0; i < lv_n[d]; ++i) {
                int64_t g = off + i;
                uint32_t p = s->parent[g];
                int key = s->morton[g] & 7;
                for (int k = 0; k < 27; ++k) {
                    uint32_t pn = s->neighbors[27 * (int64_t)p + lp[key * 27 + k]];
                    s->neighbors[27 * g + k] = (pn != ORA_NONE) ? s->childs[8 * (int64_t)pn + lc[key * 27 + k]] : ORA_NONE;
                }
            }
        }

        /* vertices per level (constructNodeVertexAndEdge, SVO.cu:1064-1096) */
        s->vxyz = NULL;
        int64_t nv_total = 0;
        vrec* lv_verts[ORA_MAX_LEVELS];
        int64_t lv_nv[ORA_MAX_LEVELS];
        for (int d = 0; d < depth; ++d) {
            int64_t nn = lv_n[d], off = s->esum_nodes[d];
            vrec* a = (vrec*)malloc(sizeof(vrec) * 8 * nn);
#pragma omp parallel for
            for (int64_t i = 0; i < nn; ++i) {
                int64_t g = off + i;
                const double* o = &s->origin[3 * g];
                double wd = s->width[g];
                for (int c = 0; c < 8; ++c) {
                    /* corner = origin + width * Vector3d(x,y,z), x,y,z in {0,1} */
                    vrec* r = &a[8 * i + c];
                    r->p[0] = o[0] + wd * (double)(c & 1);
                    r->p[1] = o[1] + wd * (double)((c >> 1) & 1);
                    r->p[2] = o[2] + wd * (double)((c >> 2) & 1);
                    uint64_t idx = (uint64_t)g;
                    for (int j = 0; j < 8; ++j) {
                        uint32_t nb = s->neighbors[27 * g + vs[c * 8 + j]];
                        if ((uint64_t)nb < idx) idx = nb;
                    }
                    r->owner = (uint32_t)idx;
                }
            }
            vrec_sort(a, 8 * nn);
            int64_t m = 0;
            for (int64_t i = 0; i < 8 * nn; ++i) {
                if (m > 0 && approx_eq(a[i - 1].p, a[i].p)) continue;  /* thrust::unique vs previous */
                a[m++] = a[i];
            }
            lv_verts[d] = a;
            lv_nv[d] = m;
            s->esum_verts[d] = nv_total;
            nv_total += m;
        }
        s->esum_verts[depth] = nv_total;
        s->num_verts = nv_total;
        s->vxyz = (double*)malloc(sizeof(double) * 3 * (nv_total ? nv_total : 1));
        s->vowner = (uint32_t*)malloc(sizeof(uint32_t) * (nv_total ? nv_total : 1));
        for (int d = 0; d < depth; ++d) {
            for (int64_t i = 0; i < lv_nv[d]; ++i) {
                int64_t g = s->esum_verts[d] + i;
                memcpy(&s->vxyz[3 * g], lv_verts[d][i].p, sizeof(double) * 3);
                s->vowner[g] = lv_verts[d][i].owner;
            }
            free(lv_verts[d]);
        }
    } else {
        s->vxyz = (double*)malloc(sizeof(double) * 3);
        s->vowner = (uint32_t*)malloc(sizeof(uint32_t));
    }
    return s;
}

void ora_svo_free(ora_svo* s) {
    if (!s) return;
    free(s->morton); free(s->is_leaf); free(s->origin); free(s->width); free(s->parent);
    free(s->childs); free(s->neighbors); free(s->vxyz); free(s->vowner);
    free(s);
}
int ora_svo_depth(const ora_svo* s) { return s->depth; }
int64_t ora_svo_num_nodes(const ora_svo* s) { return s->num_nodes; }
int64_t ora_svo_num_verts(const ora_svo* s) { return s->num_verts; }
int64_t ora_svo_num_voxels(const ora_svo* s) { return s->num_voxels; }
void ora_svo_level_info(const ora_svo* s, int64_t* dn, int64_t* ev) {
    for (int d = 0; d < s->depth; ++d) dn[d] = s->depth_num_nodes[d];
    for (int d = 0; d <= s->depth; ++d) ev[d] = s->esum_verts[d];
}
void ora_svo_nodes(const ora_svo* s, uint32_t* morton, uint8_t* is_leaf, double* origin, double* width,
                   uint32_t* parent, uint32_t* childs, uint32_t* neighbors) {
    int64_t n = s->num_nodes;
    if (morton) memcpy(morton, s->morton, sizeof(uint32_t) * n);
    if (is_leaf) memcpy(is_leaf, s->is_leaf, n);
    if (origin) memcpy(origin, s->origin, sizeof(double) * 3 * n);
    if (width) memcpy(width, s->width, sizeof(double) * n);
    if (parent) memcpy(parent, s->parent, sizeof(uint32_t) * n);
    if (childs) memcpy(childs, s->childs, sizeof(uint32_t) * 8 * n);
    if (neighbors) memcpy(neighbors, s->neighbors, sizeof(uint32_t) * 27 * n);
}
void ora_svo_verts(const ora_svo* s, double* xyz, uint32_t* owner) {
    if (xyz) memcpy(xyz, s->vxyz, sizeof(double) * 3 * s->num_verts);
    if (owner) memcpy(owner, s->vowner, sizeof(uint32_t) * s->num_verts);
}

/* fuzzy std::map<V3d,size_t>::find over one level's sorted vertex array */
static int64_t find_vert(const ora_svo* s, int d, const double p[3]) {
    int64_t lo = s->esum_verts[d], hi = s->esum_verts[d + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (less_point(&s->vxyz[3 * mid], p) == 1) lo = mid + 1;
        else hi = mid;
    }
    if (lo < s->esum_verts[d + 1] && less_point(p, &s->vxyz[3 * lo]) != 1) return lo;
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Hat basis (src/BSpline.hpp:5-27)                                           */
/* ------------------------------------------------------------------------ */
double ora_base_function(double x, double c, double w) {
    if (fabs(x - (c - w)) < 1e-9 || x < c - w || fabs(x - (c + w)) < 1e-9 || x > c + w) return 0.0;
    if (x <= c) return 1 + (x - c) / w;
    return 1 - (x - c) / w;
}
double ora_base_function3(const double c[3], double w, const double q[3]) {
    double x = ora_base_function(q[0], c[0], w);
    if (x <= 0.0) return 0.0;
    double y = ora_base_function(q[1], c[1], w);
    if (y <= 0.0) return 0.0;
    double z = ora_base_function(q[2], c[2], w);
    if (z <= 0.0) return 0.0;
    return x * y * z;
}

/* ------------------------------------------------------------------------ */
/* Assembly (src/ThinShells.cpp:511-605)                                      */
/* ------------------------------------------------------------------------ */
struct ora_mat {
    int64_t n;
    int64_t* rowptr;
    int32_t* cols;
    double* vals;
    /* transpose, for A^T x */
    int64_t* trowptr;
    int32_t* tcols;
    double* tvals;
};

typedef struct { int32_t col; double val; } entry;
static int entry_cmp(const void* a, const void* b) {
    int32_t x = ((const entry*)a)->col, y = ((const entry*)b)->col;
    return (x > y) - (x < y);
}

/* setInDomainPoints (SVO.cu:1160-1193): corners of the node and its
 * ancestors with width and global vertex index */
static int set_in_domain(const ora_svo* s, uint32_t node, int depth, double (*pts)[3], double* wid, int64_t* idx) {
    int cnt = 0;
    while (depth < s->depth && node != ORA_NONE) {
        const double* o = &s->origin[3 * (int64_t)node];
        double wd = s->width[node];
        for (int k = 0; k < 8; ++k) {
            double c[3] = {o[0] + wd * (double)(k & 1), o[1] + wd * (double)((k >> 1) & 1), o[2] + wd * (double)((k >> 2) & 1)};
            int64_t v = find_vert(s, depth, c);
            if (v < 0) { fprintf(stderr, "oracle: corner not found (map::at)\n"); abort(); }
            memcpy(pts[cnt], c, sizeof c);
            wid[cnt] = wd;
            idx[cnt] = v;
            ++cnt;
        }
        node = s->parent[node];
        ++depth;
    }
    return cnt;
}

ora_mat* ora_assemble(const ora_svo* s) {
    int64_t n = s->num_verts;
    ora_mat* A = (ora_mat*)calloc(1, sizeof(ora_mat));
    A->n = n;
    int L = s->depth;
    int64_t* rowcnt = (int64_t*)calloc(n + 1, sizeof(int64_t));
    entry** rows = (entry**)calloc(n ? n : 1, sizeof(entry*));
    for (int d = 0; d < L; ++d) {
        int64_t b = s->esum_verts[d], e = s->esum_verts[d + 1];
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t row = b; row < e; ++row) {
            const double* v = &s->vxyz[3 * row];
            uint32_t from = s->vowner[row];
            entry buf[1 + ORA_MAX_LEVELS + 8 * ORA_MAX_LEVELS];
            int m = 0;
            buf[m].col = (int32_t)row; buf[m].val = 1; ++m;      /* self */
            for (int j = d - 1; j >= 0; --j) {                      /* coincident finer */
                int64_t c = find_vert(s, j, v);
                if (c < 0) break;
                buf[m].col = (int32_t)c; buf[m].val = 1; ++m;
            }
            double pts[8 * ORA_MAX_LEVELS][3], wid[8 * ORA_MAX_LEVELS];
            int64_t idx[8 * ORA_MAX_LEVELS];
            int np = set_in_domain(s, s->parent[from], d + 1, pts, wid, idx);
            for (int k = 0; k < np; ++k) {
                double val = ora_base_function3(pts[k], wid[k], v);
                if (val != 0) { buf[m].col = (int32_t)idx[k]; buf[m].val = val; ++m; }
            }
            qsort(buf, m, sizeof(entry), entry_cmp);
            rows[row] = (entry*)malloc(sizeof(entry) * m);
            memcpy(rows[row], buf, sizeof(entry) * m);
            rowcnt[row] = m;
        }
    }
    A->rowptr = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    A->rowptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) A->rowptr[i + 1] = A->rowptr[i] + rowcnt[i];
    int64_t nnz = A->rowptr[n];
    A->cols = (int32_t*)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    A->vals = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t k = 0; k < rowcnt[i]; ++k) {
            A->cols[A->rowptr[i] + k] = rows[i][k].col;
            A->vals[A->rowptr[i] + k] = rows[i][k].val;
        }
        free(rows[i]);
    }
    free(rows);
    free(rowcnt);
    /* transpose (rows of A^T in increasing row order of A) */
    A->trowptr = (int64_t*)calloc(n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < nnz; ++k) A->trowptr[A->cols[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) A->trowptr[i + 1] += A->trowptr[i];
    A->tcols = (int32_t*)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    A->tvals = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1));
    for (int64_t i = 0; i < n; ++i) fill[i] = A->trowptr[i];
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k) {
            int64_t p = fill[A->cols[k]]++;
            A->tcols[p] = (int32_t)i;
            A->tvals[p] = A->vals[k];
        }
    free(fill);
    return A;
}
void ora_mat_free(ora_mat* A) {
    if (!A) return;
    free(A->rowptr); free(A->cols); free(A->vals); free(A->trowptr); free(A->tcols); free(A->tvals);
    free(A);
}
int64_t ora_mat_nnz(const ora_mat* A) { return A->rowptr[A->n]; }
int64_t ora_mat_n(const ora_mat* A) { return A->n; }
void ora_mat_csr(const ora_mat* A, int64_t* rowptr, int32_t* cols, double* vals) {
    if (rowptr) memcpy(rowptr, A->rowptr, sizeof(int64_t) * (A->n + 1));
    if (cols) memcpy(cols, A->cols, sizeof(int32_t) * A->rowptr[A->n]);
    if (vals) memcpy(vals, A->vals, sizeof(double) * A->rowptr[A->n]);
}
static void csr_mul(int64_t n, const int64_t* rp, const int32_t* c, const double* v, const double* x, double* y) {
#pragma omp parallel for schedule(static, 1024)
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) acc += v[k] * x[c[k]];
        y[i] = acc;
    }
}
void ora_mat_mul(const ora_mat* A, const double* x, double* y) { csr_mul(A->n, A->rowptr, A->cols, A->vals, x, y); }
void ora_mat_tmul(const ora_mat* A, const double* x, double* y) { csr_mul(A->n, A->trowptr, A->tcols, A->tvals, x, y); }

/* ------------------------------------------------------------------------ */
/* LSCG: Eigen's least_square_conjugate_gradient with the least-squares
 * diagonal preconditioner (inverse squared column norms, 1 for empty columns),
 * as instantiated by LeastSquaresConjugateGradient<SparseMatrix<double>>
 * (src/ThinShells.cpp:615-617).  Reductions use a fixed blocked order so the
 * result does not depend on the thread count.                              */
/* ------------------------------------------------------------------------ */
#define RED_BLOCK 4096
static double det_dot(const double* a, const double* b, int64_t n) {
    int64_t nb = (n + RED_BLOCK - 1) / RED_BLOCK;
    double* part = (double*)malloc(sizeof(double) * (nb ? nb : 1));
#pragma omp parallel for
    for (int64_t blk = 0; blk < nb; ++blk) {
        double acc = 0.0;
        int64_t e = (blk + 1) * RED_BLOCK < n ? (blk + 1) * RED_BLOCK : n;
        for (int64_t i = blk * RED_BLOCK; i < e; ++i) acc += a[i] * b[i];
        part[blk] = acc;
    }
    double s = 0.0;
    for (int64_t blk = 0; blk < nb; ++blk) s += part[blk];
    free(part);
    return s;
}

int ora_lscg(const ora_mat* A, const double* b, double* x, double tol, int64_t max_iter, ora_lscg_stats* st) {
    int64_t n = A->n;
    double* invdiag = (double*)malloc(sizeof(double) * n);
    for (int64_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (int64_t k = A->trowptr[j]; k < A->trowptr[j + 1]; ++k) s += A->tvals[k] * A->tvals[k];
        invdiag[j] = s > 0.0 ? 1.0 / s : 1.0;
    }
    double* residual = (double*)malloc(sizeof(double) * n);
    double* nres = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* tmp = (double*)malloc(sizeof(double) * n);

    ora_mat_mul(A, x, tmp);
    for (int64_t i = 0; i < n; ++i) residual[i] = b[i] - tmp[i];
    ora_mat_tmul(A, residual, nres);
    ora_mat_tmul(A, b, tmp);
    double rhsNorm2 = det_dot(tmp, tmp, n);
    int64_t iters = 0;
    double residualNorm2;
    if (rhsNorm2 == 0) {
        for (int64_t i = 0; i < n; ++i) x[i] = 0;
        residualNorm2 = 0;
        goto done;
    }
    {
        double threshold = tol * tol * rhsNorm2;
        residualNorm2 = det_dot(nres, nres, n);
        if (residualNorm2 < threshold) goto done;
        for (int64_t i = 0; i < n; ++i) p[i] = invdiag[i] * nres[i];
        double absNew = det_dot(nres, p, n);
        int64_t i = 0;
        while (i < max_iter) {
            ora_mat_mul(A, p, tmp);
            double alpha = absNew / det_dot(tmp, tmp, n);
            for (int64_t k = 0; k < n; ++k) x[k] += alpha * p[k];
            for (int64_t k = 0; k < n; ++k) residual[k] -= alpha * tmp[k];
            ora_mat_tmul(A, residual, nres);
            residualNorm2 = det_dot(nres, nres, n);
            if (residualNorm2 < threshold) break;
            for (int64_t k = 0; k < n; ++k) z[k] = invdiag[k] * nres[k];
            double absOld = absNew;
            absNew = det_dot(nres, z, n);
            double beta = absNew / absOld;
            for (int64_t k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
            i++;
        }
        iters = i;
    }
done:
    if (st) {
        st->iterations = iters;
        st->normal_rel = rhsNorm2 > 0 ? sqrt(residualNorm2 / rhsNorm2) : 0.0;
        ora_mat_mul(A, x, tmp);
        for (int64_t k = 0; k < n; ++k) residual[k] = tmp[k] - b[k];
        double rr = det_dot(residual, residual, n), bb = det_dot(b, b, n);
        st->residual_abs = sqrt(rr);
        st->residual_rel = bb > 0 ? sqrt(rr / bb) : 0.0;
        /* recompute the normal residual from scratch */
        ora_mat_tmul(A, residual, nres);
        if (rhsNorm2 > 0) st->normal_rel = sqrt(det_dot(nres, nres, n) / rhsNorm2);
    }
    free(invdiag); free(residual); free(nres); free(p); free(z); free(tmp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Evaluation of f                                                           */
/* ------------------------------------------------------------------------ */
static int64_t find_node(const ora_svo* s, int d, uint32_t morton) {
    int64_t off = s->esum_nodes[d], n = s->depth_num_nodes[d];
    int64_t i = lower_bound_u32(s->morton + off, n, morton);
    if (i < n && s->morton[off + i] == morton) return off + i;
    return -1;
}

static inline uint32_t next_pow2(uint32_t v) { /* common.hpp nextPow2 */
    if (v == 0) return 0;
    --v;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}


/* mq_setInDomainPoints (SVO.cu:1195-1261): corners of virtual cells that are
 * vertices, then all corners of the first existing node and its ancestors */
static double mq_sum(const ora_svo* s, const double* lambda, uint32_t morton, double node_w, int depth,
                     const double p[3]) {
    const double* mo = s->grid_origin;
    double sum = 0.0;
    while (depth < s->depth && find_node(s, depth, morton) < 0) {
        uint32_t x, y, z;
        ora_morton_decode(morton, &x, &y, &z);
        /* modelOrigin + V3d(x,y,z) * virtualNodeWidth */
        double vo[3] = {mo[0] + (double)(int)x * node_w, mo[1] + (double)(int)y * node_w, mo[2] + (double)(int)z * node_w};
        for (int k = 0; k < 8; ++k) {
            double c[3] = {vo[0] + node_w * (double)(k & 1), vo[1] + node_w * (double)((k >> 1) & 1),
                           vo[2] + node_w * (double)((k >> 2) & 1)};
            int64_t v = find_vert(s, depth, c);
            if (v >= 0) sum += lambda[v] * ora_base_function3(c, node_w, p);
        }
        morton /= 8;
        node_w *= 2;
        ++depth;
    }
    if (depth == s->depth) return sum;
    int64_t node = find_node(s, depth, morton);
    while (node >= 0 && (uint32_t)node != ORA_NONE) {
        const double* o = &s->origin[3 * node];
        double wd = s->width[node];
        for (int k = 0; k < 8; ++k) {
            double c[3] = {o[0] + wd * (double)(k & 1), o[1] + wd * (double)((k >> 1) & 1), o[2] + wd * (double)((k >> 2) & 1)};
            int64_t v = find_vert(s, depth, c);
            if (v < 0) { fprintf(stderr, "oracle: node corner not found (map::at)\n"); abort(); }
            sum += lambda[v] * ora_base_function3(c, wd, p);
        }
        uint32_t par = s->parent[node];
        node = (par == ORA_NONE) ? -1 : (int64_t)par;
        ++depth;
    }
    return sum;
}

/* multiPointQuery CPU path mt_cpuTest_2 (ThinShells.cpp:1249-1294) */
void ora_eval_support(const ora_svo* s, const double* lambda, const double* pts, int64_t N, const double box_end[3],
                      double out_val, double* f) {
    const double* bo = s->grid_origin;
    double vw = s->voxel_width;
    int R = s->R;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < N; ++q) {
        const double* p = &pts[3 * q];
        if (p[0] <= bo[0] || p[1] <= bo[1] || p[2] <= bo[2] || p[0] >= box_end[0] || p[1] >= box_end[1] ||
            p[2] >= box_end[2]) {
            f[q] = out_val;
            continue;
        }
        int dis[3];
        for (int a = 0; a < 3; ++a) dis[a] = (int)((p[a] - bo[a]) / vw);
        int maxOffset = 0, searchDepth = 0;
        double w = vw;
        for (int a = 0; a < 3; ++a) {
            if (dis[a] <= -1) { int o = abs(dis[a]); if (o > maxOffset) maxOffset = o; dis[a] = 0; }
            else if (dis[a] >= R) { int o = dis[a] - R + 1; if (o > maxOffset) maxOffset = o; dis[a] = R - 1; }
        }
        uint32_t m = ora_morton_encode((uint16_t)dis[0], (uint16_t)dis[1], (uint16_t)dis[2]);
        uint32_t mo = next_pow2((uint32_t)maxOffset);
        while (mo >= 2) { m /= 8; ++searchDepth; w *= 2; mo >>= 1; }
        f[q] = (s->depth > 0) ? mq_sum(s, lambda, m, w, searchDepth, p) : 0.0;
    }
}

/* cpLatentBSplineValue (ThinShells.cpp:655-701): f through the level-0 node
 * containing the point; the reference throws (map::at) if there is none. */
int ora_eval_latent(const ora_svo* s, const double* lambda, const double* pts, int64_t N, double* f) {
    const double* bo = s->grid_origin;
    double vw = s->voxel_width;   /* svoNodeArray[0].width */
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
    for (int64_t q = 0; q < N; ++q) {
        const double* p = &pts[3 * q];
        int dis[3];
        for (int a = 0; a < 3; ++a) dis[a] = (int)((p[a] - bo[a]) / vw);
        uint32_t m = ora_morton_encode((uint16_t)dis[0], (uint16_t)dis[1], (uint16_t)dis[2]);
        int64_t node = find_node(s, 0, m);
        if (node < 0) { f[q] = NAN; bad = 1; continue; }
        double pts8[8 * ORA_MAX_LEVELS][3], wid[8 * ORA_MAX_LEVELS];
        int64_t idx[8 * ORA_MAX_LEVELS];
        int np = set_in_domain(s, (uint32_t)node, 0, pts8, wid, idx);
        double sum = 0.0;
        for (int j = 0; j < np; ++j) sum += lambda[idx[j]] * ora_base_function3(pts8[j], wid[j], p);
        f[q] = sum;
    }
    return bad ? -1 : 0;
}

/* GPU dense formulation (CUDACompute.cu:468): sum over all vertices with the
 * width of the owner node */
void ora_eval_dense(const ora_svo* s, const double* lambda, const double* pts, int64_t N, double* f) {
    int64_t n = s->num_verts;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < N; ++q) {
        double sum = 0.0;
        for (int64_t j = 0; j < n; ++j)
            sum += lambda[j] * ora_base_function3(&s->vxyz[3 * j], s->width[s->vowner[j]], &pts[3 * q]);
        f[q] = sum;
    }
}

/* The dense value through per-level support: at level d the only basis
 * functions that can be non-zero at p are the corners of the level-d cell
 * containing p (cells clamped into the grid, so points on or beyond the box
 * faces see the boundary vertices).  Equal to ora_eval_dense. */
void ora_eval_dense_fast(const ora_svo* s, const double* lambda, const double* pts, int64_t N, double* f) {
    const double* go = s->grid_origin;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < N; ++q) {
        const double* p = &pts[3 * q];
        double sum = 0.0;
        for (int d = 0; d < s->depth; ++d) {
            double w = s->voxel_width * (double)(1u << d);
            int Rd = s->R >> d;
            if (Rd < 1) Rd = 1;
            int cell[3];
            for (int a = 0; a < 3; ++a) {
                double u = floor((p[a] - go[a]) / w);
                int c = (u < 0) ? 0 : (u > Rd - 1 ? Rd - 1 : (int)u);
                cell[a] = c;
            }
            /* corner coordinates as the octree computes them: origin of the
             * cell, plus width */
            double o[3] = {go[0] + w * (double)cell[0], go[1] + w * (double)cell[1], go[2] + w * (double)cell[2]};
            for (int k = 0; k < 8; ++k) {
                double c[3] = {o[0] + w * (double)(k & 1), o[1] + w * (double)((k >> 1) & 1), o[2] + w * (double)((k >> 2) & 1)};
                int64_t v = find_vert(s, d, c);
                if (v >= 0) sum += lambda[v] * ora_base_function3(c, w, p);
            }
        }
        f[q] = sum;
    }
}

/* ------------------------------------------------------------------------ */
/* Marching cubes (src/cuAcc/MarchingCubes/MarchingCubes.cu)                 */
/* ------------------------------------------------------------------------ */
/* Bourke's public-domain triangle table: row c lists the cut edges of cube
 * case c, three per triangle, as hexadecimal digits (edge e joins corners
 * (0,1) (1,2) (2,3) (3,0) (4,5) (5,6) (6,7) (7,4) (0,4) (1,5) (2,6) (3,7)). */
static const char* const kTriRows[256] = {
    "", "083", "019", "183981", "12a", "08312a", "92a029", "2832a8a98", "3b2", "0b28b0", "19023b",
    "1b219b98b", "3a1ba3", "0a108a8ba", "3903b9ba9", "98aa8b", "478", "430734", "019847", "419471731",
    "12a847", "34730412a", "92a902847", "2a9297273794", "8473b2", "b47b24204", "90184723b", "47b94b9b2921",
    "3a13ba784", "1ba14b1047b4", "47890b9bab03", "47b4b99ba", "954", "954083", "054150", "854835315",
    "12a954", "30812a495", "52a542402", "2a5325354348", "95423b", "0b208b495", "05401523b", "21525828b485",
    "a3ba13954", "4950818a18ba", "54050b5bab03", "54858aa8b", "978579", "930953573", "078017157", "153357",
    "978957a12", "a12950530573", "802825857a52", "2a5253357", "7957893b2", "95797292027b", "23b018178157",
    "b21b17715", "958857a13a3b", "5705097b010aba0", "ba0b03a50807570", "ba57b5", "a65", "0835a6", "9015a6",
    "1831985a6", "165261", "165126308", "965906026", "598582526328", "23ba65", "b08b20a65", "01923b5a6",
    "5a61929b298b", "63b653513", "08b0b50515b6", "3b6036065059", "65969bb98", "5a6478", "43047365a",
    "1905a6847", "a65197173794", "612651478", "125526304347", "847905065026", "739794329596269", "3b2784a65",
    "5a647242027b", "01947823b5a6", "9219b294b7b45a6", "8473b53515b6", "51b5b610b7b404b", "059065036b63847",
    "65969b4797b9", "a4964a", "4a649a083", "a01a60640", "83181686461a", "149124264", "308129249264",
    "024426", "832824426", "a49a64b23", "08228b49a4a6", "3b201606461a", "64161a48121b8b1", "964936913b63",
    "8b1810b61914641", "3b6360064", "648b68", "7a678a89a", "0730a709a67a", "a671a7178180", "a67a71173",
    "126168189867", "269291679093739", "780706602", "732672", "23ba68a89867", "20727b09767a9a7",
    "1801781a767a23b", "b21b17a61671", "896867916b63136", "091b67", "7807063b0b60", "7b6", "76b", "308b76",
    "019b76", "819831b76", "a126b7", "12a3086b7", "2902a96b7", "6b72a3a83a98", "723627", "708760620",
    "276237019", "162186198876", "a76a17137", "a7617a187108", "03707a0a96a7", "76a7a88a9", "684b86",
    "36b306046", "86b846901", "946963931b36", "6846b82a1", "12a30b06b046", "4b846b0292a9", "a93a32943b36463",
    "823842462", "042462", "190234246438", "194142246", "8138618466a1", "a10a06604", "4634386a3039a93",
    "a946a4", "49576b", "083495b76", "50154076b", "b76834354315", "954a1276b", "6b712a083495",
    "76b54a42a402", "348354325a52b76", "723762549", "954086062687", "362376150540", "628687218485158",
    "954a16176137", "16a176107870954", "40a4a503a6a737a", "76a7a854a48a", "6956b9b89", "36b063056095",
    "0b805b01556b", "6b3635531", "12a95b9b8b56", "0b306b09656912a", "b85b56805a52025", "6b36352a3a53",
    "589528562382", "956960062", "158180568382628", "156216", "13616a386569896", "a10a06950560", "03856a",
    "a56", "b5a75b", "b5ab75830", "5b75ab190", "a75ab7981831", "b12b71751", "08312717572b", "9759279022b7",
    "75272b592328982", "25a235375", "820852875a25", "9015a35373a2", "982921872a25752", "135375", "087071175",
    "903935537", "987597", "5845a8ab8", "5045b05abb30", "01984a8aba45", "ab4a45b34941314", "2512852b8458",
    "04b0b345b2b151b", "0250592b5458b85", "9452b3", "25a352345384", "5a2524420", "3a235a385458019",
    "5a2524192942", "845853351", "045105", "845853905035", "945", "4b749b9ab", "0834979b79ab",
    "1ab1b414074b", "3143481a474bab4", "4b79b492b912", "9749b791b2b1083", "b74b42240", "b74b42834324",
    "29a279237749", "9a7974a27870207", "37a3a274a1a040a", "1a2874", "491417713", "491417081871", "403743",
    "487", "9a8ab8", "30939bb9a", "01a0a88ab", "31ab3a", "12b1b99b8", "30939b1292b9", "02b80b", "32b",
    "23828aa89", "9a2092", "23828a0181a8", "1a2", "138918", "091", "038", ""
};
static int g_tri[256 * 16];
static int g_nverts[256];
static int g_tables_ready = 0;
static void mc_tables(void) {
    if (g_tables_ready) return;
    for (int c = 0; c < 256; ++c) {
        const char* r = kTriRows[c];
        int n = (int)strlen(r);
        for (int i = 0; i < 16; ++i) {
            int v = -1;
            if (i < n) { char ch = r[i]; v = (ch >= 'a') ? ch - 'a' + 10 : ch - '0'; }
            g_tri[c * 16 + i] = v;
        }
        g_nverts[c] = n;
    }
    g_tables_ready = 1;
}
const int* ora_mc_tri_table(void) { mc_tables(); return g_tri; }
const int* ora_mc_nverts_table(void) { mc_tables(); return g_nverts; }

struct ora_mc {
    int64_t nverts;          /* allTriVertices */
    int64_t nvalid;          /* nValidVoxels */
    int64_t nvox;
    double* soup;            /* nverts x 3 */
    uint8_t* cube;           /* nvox */
};

typedef struct { double x, y, z; } d3;

/* vertexLerp with the reference's 1e-6 snaps and lerp = fma(t, v1, fma(-t, v0, v0))
 * (MarchingCubes.cu:72-90, MarchingCubes.h:17-20) */
static d3 vertex_lerp(d3 p0, d3 p1, double s0, double s1, double iso) {
    if (fabs(iso - s0) < 1e-6) return p0;
    if (fabs(iso - s1) < 1e-6) return p1;
    if (fabs(s1 - s0) < 1e-6) return p0;
    double t = (iso - s0) / (s1 - s0);
    d3 r;
    r.x = fma(t, p1.x, fma(-t, p0.x, p0.x));
    r.y = fma(t, p1.y, fma(-t, p0.y, p0.y));
    r.z = fma(t, p1.z, fma(-t, p0.z, p0.z));
    return r;
}

/* reference corner order (MarchingCubes.cu:142-149) */
static const int kCorner[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}};
static const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

static void voxel_corners(const double o[3], const double vs[3], int64_t ix, int64_t iy, int64_t iz, d3 c[8]) {
    /* prepareVoxelCornerKernel: voxelPos = origin + shift * voxelSize, then
     * voxelPos + V3d(0, vs.y, 0) etc. */
    double vp[3] = {o[0] + (double)ix * vs[0], o[1] + (double)iy * vs[1], o[2] + (double)iz * vs[2]};
    for (int k = 0; k < 8; ++k) {
        c[k].x = vp[0] + (kCorner[k][0] ? vs[0] : 0.0);
        c[k].y = vp[1] + (kCorner[k][1] ? vs[1] : 0.0);
        c[k].z = vp[2] + (kCorner[k][2] ? vs[2] : 0.0);
    }
}

/* MC::marching_cubes (MarchingCubes.cu:672-714): corner values by the dense
 * evaluation of every voxel's 8 corners, cube index bit i <=> f_i < iso,
 * soup in voxel-linear order (x fastest), triangles in table order.
 * The work is restricted to the voxels of the z-slab [z0, z1) of the grid;
 * positions are those of the whole grid, so a slab's soup is the contiguous
 * piece of the whole grid's soup that those voxels emit. */
static ora_mc* mc_slab(const ora_svo* s, const double* lambda, const double origin[3], const double width[3],
                       const int res[3], double iso, int z0, int z1) {
    mc_tables();
    ora_mc* m = (ora_mc*)calloc(1, sizeof(ora_mc));
    const int64_t plane = (int64_t)res[0] * res[1];
    const int64_t v0 = plane * z0;
    int64_t nvox = plane * (z1 - z0);
    m->nvox = nvox;
    m->cube = (uint8_t*)calloc(nvox ? nvox : 1, 1);
    double vs[3] = {width[0] / res[0], width[1] / res[1], width[2] / res[2]};
    double* sdf = (double*)malloc(sizeof(double) * 8 * (nvox ? nvox : 1));
    /* corner positions, evaluated through the dense formula */
    const int64_t chunk = 1 << 16;
    double* pts = (double*)malloc(sizeof(double) * 3 * 8 * chunk);
    for (int64_t b = 0; b < nvox; b += chunk) {
        int64_t e = b + chunk < nvox ? b + chunk : nvox;
        for (int64_t v = b; v < e; ++v) {
            int64_t g = v + v0;
            int64_t ix = g % res[0], iy = g % plane / res[0], iz = g / plane;
            d3 c[8];
            voxel_corners(origin, vs, ix, iy, iz, c);
            for (int k = 0; k < 8; ++k) {
                pts[3 * (8 * (v - b) + k) + 0] = c[k].x;
                pts[3 * (8 * (v - b) + k) + 1] = c[k].y;
                pts[3 * (8 * (v - b) + k) + 2] = c[k].z;
            }
        }
        ora_eval_dense_fast(s, lambda, pts, 8 * (e - b), sdf + 8 * b);
    }
    free(pts);
    int64_t* vscan = (int64_t*)malloc(sizeof(int64_t) * (nvox + 1));
    vscan[0] = 0;
    int64_t nvalid = 0;
    for (int64_t v = 0; v < nvox; ++v) {
        int c = 0;
        for (int k = 0; k < 8; ++k) c |= (sdf[8 * v + k] < iso) << k;
        m->cube[v] = (uint8_t)c;
        vscan[v + 1] = vscan[v] + g_nverts[c];
        nvalid += g_nverts[c] > 0;
    }
    m->nvalid = nvalid;
    m->nverts = vscan[nvox];
    m->soup = (double*)malloc(sizeof(double) * 3 * (m->nverts ? m->nverts : 1));
    int64_t maxVerts = 18 * nvox;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < nvox; ++v) {
        int c = m->cube[v];
        int nv = g_nverts[c];
        if (!nv) continue;
        int64_t g = v + v0;
        int64_t ix = g % res[0], iy = g % plane / res[0], iz = g / plane;
        /* voxelToMeshKernel recomputes voxelPos = shift * voxelSize + origin */
        double vp[3] = {(double)ix * vs[0] + origin[0], (double)iy * vs[1] + origin[1], (double)iz * vs[2] + origin[2]};
        d3 cr[8];
        for (int k = 0; k < 8; ++k) {
            cr[k].x = vp[0] + (kCorner[k][0] ? vs[0] : 0.0);
            cr[k].y = vp[1] + (kCorner[k][1] ? vs[1] : 0.0);
            cr[k].z = vp[2] + (kCorner[k][2] ? vs[2] : 0.0);
        }
        d3 tv[12];
        for (int e = 0; e < 12; ++e)
            tv[e] = vertex_lerp(cr[kEdge[e][0]], cr[kEdge[e][1]], sdf[8 * v + kEdge[e][0]], sdf[8 * v + kEdge[e][1]], iso);
        for (int i = 0; i < nv; i += 3) {
            int64_t pos = vscan[v] + i;
            if (pos < maxVerts - 3) {
                for (int t = 0; t < 3; ++t) {
                    d3 q = tv[g_tri[c * 16 + i + t]];
                    m->soup[3 * (pos + t) + 0] = q.x;
                    m->soup[3 * (pos + t) + 1] = q.y;
                    m->soup[3 * (pos + t) + 2] = q.z;
                }
            }
        }
    }
    free(vscan);
    free(sdf);
    return m;
}
ora_mc* ora_marching_cubes(const ora_svo* s, const double* lambda, const double origin[3], const double width[3],
                           const int res[3], double iso) {
    return mc_slab(s, lambda, origin, width, res, iso, 0, res[2]);
}
ora_mc* ora_marching_cubes_slab(const ora_svo* s, const double* lambda, const double origin[3],
                                const double width[3], const int res[3], double iso, int z0, int z1) {
    if (z0 < 0) z0 = 0;
    if (z1 > res[2]) z1 = res[2];
    if (z1 < z0) z1 = z0;
    return mc_slab(s, lambda, origin, width, res, iso, z0, z1);
}
void ora_mc_free(ora_mc* m) { if (!m) return; free(m->soup); free(m->cube); free(m); }
int64_t ora_mc_num_verts(const ora_mc* m) { return m->nverts; }
int64_t ora_mc_num_valid_voxels(const ora_mc* m) { return m->nvalid; }
void ora_mc_soup(const ora_mc* m, double* xyz) { memcpy(xyz, m->soup, sizeof(double) * 3 * m->nverts); }
void ora_mc_cube_index(const ora_mc* m, uint8_t* cube) { memcpy(cube, m->cube, m->nvox); }

/* ------------------------------------------------------------------------ */
/* Signed distance for the fit's right-hand side (ThinShells.cpp:481-509).
 * Unsigned part: closest point on each triangle (Ericson, Real-Time
 * Collision Detection, 5.1.5), minimum over triangles.  Sign: negative where
 * the generalised winding number (sum of solid angles, Van Oosterom and
 * Strackee 1983, over 4 pi) exceeds 1/2.                                    */
/* ------------------------------------------------------------------------ */
static double point_tri_dist2(const double p[3], const double a[3], const double b[3], const double c[3]) {
    double ab[3], ac[3], ap[3];
    for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i]; }
    double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    double q[3];
    if (d1 <= 0.0 && d2 <= 0.0) { memcpy(q, a, sizeof q); goto out; }
    double bp[3];
    for (int i = 0; i < 3; ++i) bp[i] = p[i] - b[i];
    double d3_ = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3_ >= 0.0 && d4 <= d3_) { memcpy(q, b, sizeof q); goto out; }
    double vc = d1 * d4 - d3_ * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) {
        double v = d1 / (d1 - d3_);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + v * ab[i];
        goto out;
    }
    double cp[3];
    for (int i = 0; i < 3; ++i) cp[i] = p[i] - c[i];
    double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { memcpy(q, c, sizeof q); goto out; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + w * ac[i];
        goto out;
    }
    double va = d3_ * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3_) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3_) / ((d4 - d3_) + (d5 - d6));
        for (int i = 0; i < 3; ++i) q[i] = b[i] + w * (c[i] - b[i]);
        goto out;
    }
    {
        double denom = 1.0 / (va + vb + vc);
        double v = vb * denom, w = vc * denom;
        for (int i = 0; i < 3; ++i) q[i] = a[i] + ab[i] * v + ac[i] * w;
    }
out:;
    double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return dot3(d, d);
}

static double solid_angle(const double p[3], const double A[3], const double B[3], const double C[3]) {
    double a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) { a[i] = A[i] - p[i]; b[i] = B[i] - p[i]; c[i] = C[i] - p[i]; }
    double la = sqrt(dot3(a, a)), lb = sqrt(dot3(b, b)), lc = sqrt(dot3(c, c));
    double bxc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    double det = dot3(a, bxc);
    double den = la * lb * lc + dot3(a, b) * lc + dot3(b, c) * la + dot3(c, a) * lb;
    return 2.0 * atan2(det, den);
}

double ora_winding_number(const double* V, const int32_t* F, int64_t nF, const double p[3]) {
    double w = 0.0;
    for (int64_t t = 0; t < nF; ++t)
        w += solid_angle(p, &V[3 * F[3 * t]], &V[3 * F[3 * t + 1]], &V[3 * F[3 * t + 2]]);
    return w / (4.0 * M_PI);
}

void ora_signed_distance(const double* V, int64_t nV, const int32_t* F, int64_t nF, const double* pts, int64_t n,
                         double* out) {
    (void)nV;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        const double* p = &pts[3 * i];
        double best = DBL_MAX;
        for (int64_t t = 0; t < nF; ++t) {
            double d2 = point_tri_dist2(p, &V[3 * F[3 * t]], &V[3 * F[3 * t + 1]], &V[3 * F[3 * t + 2]]);
            if (d2 < best) best = d2;
        }
        double d = sqrt(best);
        double w = ora_winding_number(V, F, nF, p);
        out[i] = (w > 0.5) ? -d : d;
    }
}

/* ------------------------------------------------------------------------ */
/* Shell iso band: intersection points (ThinShells.cpp:38-479)               */
/* ------------------------------------------------------------------------ */
typedef struct { double a[3], b[3]; } seg;

static int cmp_u64(const void* x, const void* y) {
    uint64_t a = *(const uint64_t*)x, b = *(const uint64_t*)y;
    return a < b ? -1 : (a > b ? 1 : 0);
}
/* qsort form of the fuzzy std::less<V3d> */
static int cmp_point(const void* x, const void* y) {
    const int l = less_point((const double*)x, (const double*)y);
    return l == 1 ? -1 : (l == -1 ? 1 : 0);
}
/* std::less<V3d> of two segments: first endpoints, then second (SVO.cu:1033-1052) */
static int cmp_seg(const void* x, const void* y) {
    const seg* a = (const seg*)x;
    const seg* b = (const seg*)y;
    if (less_point(a->a, b->a) == 1) return -1;
    if (less_point(b->a, a->a) == 1) return 1;
    if (less_point(a->b, b->b) == 1) return -1;
    if (less_point(b->b, a->b) == 1) return 1;
    return 0;
}
static int g_axis; /* sort key axis of the edge sweeps */
/* x_sortEdge / y_sortEdge / z_sortEdge (ThinShells.cpp:170-209): by the end
 * point's sweep coordinate, then the other two in the order of the reference */
static int cmp_end(const void* x, const void* y) {
    const double* a = ((const seg*)x)->b;
    const double* b = ((const seg*)y)->b;
    static const int order[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
    for (int k = 0; k < 3; ++k) {
        int i = order[g_axis][k];
        if (fabs(a[i] - b[i]) < 1e-9) continue;
        return a[i] < b[i] ? -1 : 1;
    }
    return 0;
}
static inline double dot3l(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

typedef struct { double* p; int64_t n, cap; } pbuf;
static void push3(pbuf* b, const double* q) {
    if (b->n == b->cap) { b->cap = b->cap ? 2 * b->cap : 1024; b->p = (double*)realloc(b->p, 3 * sizeof(double) * b->cap); }
    memcpy(b->p + 3 * b->n, q, 3 * sizeof(double));
    b->n++;
}

/* Fine node edges (determineNodeEdge, src/cuAcc/SVO.cu:965-1008): 12 edges
 * per level-0 node, each tagged with the smallest index among the node and
 * its neighbours sharing the edge (d_edgeSharedLUT).  Sorted by the first
 * end, then the second end, in the fuzzy order of lessPoint, then by that
 * index (sortEdge, SVO.cu:1033-1046); of each run of approximately equal
 * edges (uniqueEdge, SVO.cu:1054-1060, Eigen isApprox at 1e-12) the first
 * is kept.  Entries equal in that order come from different nodes and differ
 * by at most an ulp; the reference's unstable sort leaves their order open,
 * and here the owner node's own entry comes first. */
typedef struct { double a[3], b[3]; uint32_t idx, node; } oedge;
static int cmp_oedge(const void* x, const void* y) {
    const oedge* p = (const oedge*)x;
    const oedge* q = (const oedge*)y;
    int t = -less_point(p->a, q->a);   /* lessPoint: 1 less, -1 greater, 0 equal */
    if (t) return t;
    t = -less_point(p->b, q->b);
    if (t) return t;
    if (p->idx != q->idx) return p->idx < q->idx ? -1 : 1;
    const int po = p->node == p->idx, qo = q->node == q->idx;
    if (po != qo) return po ? -1 : 1;
    return p->node < q->node ? -1 : (p->node > q->node ? 1 : 0);
}
static int approx3(const double* a, const double* b) {
    double d = 0, na = 0, nb = 0;
    for (int i = 0; i < 3; ++i) { d += (a[i] - b[i]) * (a[i] - b[i]); na += a[i] * a[i]; nb += b[i] * b[i]; }
    return d <= 1e-24 * (na < nb ? na : nb);
}
int64_t ora_fine_edges(const ora_svo* s, double* ends, uint32_t* owner, int64_t cap) {
    static const int eo[12][2][3] = {
        {{0, 0, 0}, {0, 1, 0}}, {{0, 1, 0}, {1, 1, 0}}, {{1, 0, 0}, {1, 1, 0}}, {{0, 0, 0}, {1, 0, 0}},
        {{0, 0, 1}, {0, 1, 1}}, {{0, 1, 1}, {1, 1, 1}}, {{1, 0, 1}, {1, 1, 1}}, {{0, 0, 1}, {1, 0, 1}},
        {{0, 0, 0}, {0, 0, 1}}, {{0, 1, 0}, {0, 1, 1}}, {{1, 1, 0}, {1, 1, 1}}, {{1, 0, 0}, {1, 0, 1}}};
    int16_t lp[8 * 27], lc[8 * 27], vs[64], es[48];
    ora_neighbor_luts(lp, lc, vs, es);
    const int64_t n0 = s->depth_num_nodes[0];
    oedge* E = (oedge*)malloc(sizeof(oedge) * 12 * (n0 > 0 ? n0 : 1));
    for (int64_t i = 0; i < n0; ++i) {
        const int64_t g = s->esum_nodes[0] + i;
        const double* o = s->origin + 3 * g;
        const double w = s->width[g];
        for (int e = 0; e < 12; ++e) {
            oedge* q = &E[12 * i + e];
            for (int a = 0; a < 3; ++a) {
                q->a[a] = o[a] + (eo[e][0][a] ? w : 0.0);
                q->b[a] = o[a] + (eo[e][1][a] ? w : 0.0);
            }
            uint32_t idx = (uint32_t)g;
            for (int j = 0; j < 4; ++j) {
                const uint32_t nb = s->neighbors[27 * g + es[4 * e + j]];
                if (nb < idx) idx = nb;
            }
            q->idx = idx;
            q->node = (uint32_t)g;
        }
    }
    qsort(E, 12 * n0, sizeof(oedge), cmp_oedge);
    int64_t n = 0;
    for (int64_t k = 0; k < 12 * n0; ++k) {
        if (k > 0) {
            const oedge* p = &E[k - 1];
            const oedge* q = &E[k];
            if ((approx3(p->a, q->a) && approx3(p->b, q->b)) || (approx3(p->a, q->b) && approx3(p->b, q->a))) continue;
        }
        if (n < cap) {
            if (ends) {
                memcpy(ends + 6 * n, E[k].a, 3 * sizeof(double));
                memcpy(ends + 6 * n + 3, E[k].b, 3 * sizeof(double));
            }
            if (owner) owner[n] = E[k].idx;
        }
        ++n;
    }
    free(E);
    return n;
}

int64_t ora_intersections(const ora_svo* s, const double* V, int64_t nV, const int32_t* F, int64_t nF, double* out,
                          int64_t cap, int64_t* n_edge) {
    (void)nV;
    const int R = s->R;
    const double* mo = s->grid_origin;
    const double vw = s->voxel_width;
    const int64_t n0 = s->depth_num_nodes[0];
    pbuf E = {0, 0, 0}, Fp = {0, 0, 0};
    /* model edges: the ordered set of (min, max) vertex pairs (BaseModel.cpp:291-313) */
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * 3 * (nF > 0 ? nF : 1));
    for (int64_t f = 0; f < nF; ++f) {
        int32_t i0 = F[3 * f], i1 = F[3 * f + 1], i2 = F[3 * f + 2];
        int32_t mx = i0 > i1 ? (i0 > i2 ? i0 : i2) : (i1 > i2 ? i1 : i2);
        int32_t mn = i0 < i1 ? (i0 < i2 ? i0 : i2) : (i1 < i2 ? i1 : i2);
        int32_t md = i0 + i1 + i2 - mx - mn;
        keys[3 * f] = ((uint64_t)(uint32_t)mn << 32) | (uint32_t)md;
        keys[3 * f + 1] = ((uint64_t)(uint32_t)md << 32) | (uint32_t)mx;
        keys[3 * f + 2] = ((uint64_t)(uint32_t)mn << 32) | (uint32_t)mx;
    }
    qsort(keys, 3 * nF, sizeof(uint64_t), cmp_u64);
    int64_t ne = 0;
    for (int64_t k = 0; k < 3 * nF; ++k)
        if (k == 0 || keys[k] != keys[k - 1]) keys[ne++] = keys[k];
    /* mesh edges x back/left/bottom faces of level-0 nodes (ThinShells.cpp:65-138) */
    for (int64_t k = 0; k < ne; ++k) {
        const double* p1 = V + 3 * (int64_t)(keys[k] >> 32);
        const double* p2 = V + 3 * (int64_t)(keys[k] & 0xFFFFFFFFu);
        double dir[3];
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            dir[a] = p2[a] - p1[a];
            int d1 = (int)((p1[a] - mo[a]) / vw), d2 = (int)((p2[a] - mo[a]) / vw);
            int l = (d1 < d2 ? d1 : d2) - 1, h = (d1 > d2 ? d1 : d2) + 1;
            lo[a] = l < 0 ? 0 : (l > R - 1 ? R - 1 : l);
            hi[a] = h < 0 ? 0 : (h > R - 1 ? R - 1 : h);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    int64_t node = find_node(s, 0, ora_morton_encode((uint32_t)x, (uint32_t)y, (uint32_t)z));
                    if (node < 0) continue;
                    const double* o = s->origin + 3 * node;
                    const double w = s->width[node];
                    for (int a = 0; a < 3; ++a) {   /* back (x), left (y), bottom (z) planes */
                        double t = DBL_MAX;
                        if (dir[a] != 0) t = (o[a] - p1[a]) / dir[a];
                        if (!(0.0 <= t && t <= 1.0)) continue;
                        double q[3] = {p1[0] + t * dir[0], p1[1] + t * dir[1], p1[2] + t * dir[2]};
                        int b = (a + 1) % 3, c = (a + 2) % 3;
                        if (b > c) { int tmp = b; b = c; c = tmp; }
                        if (o[b] <= q[b] && q[b] <= o[b] + w && o[c] <= q[c] && q[c] <= o[c] + w) push3(&E, q);
                    }
                }
    }
    free(keys);
    qsort(E.p, E.n, 3 * sizeof(double), cmp_point);
    int64_t nu = 0;
    for (int64_t k = 0; k < E.n; ++k)
        if (nu == 0 || memcmp(E.p + 3 * k, E.p + 3 * (nu - 1), 3 * sizeof(double)) != 0) {
            memmove(E.p + 3 * nu, E.p + 3 * k, 3 * sizeof(double));
            ++nu;
        }
    E.n = nu;
    /* fine node edges (determineNodeEdge, SVO.cu:965-1008), sorted and unique */
    static const int eo[12][2][3] = {
        {{0, 0, 0}, {0, 1, 0}}, {{0, 1, 0}, {1, 1, 0}}, {{1, 0, 0}, {1, 1, 0}}, {{0, 0, 0}, {1, 0, 0}},
        {{0, 0, 1}, {0, 1, 1}}, {{0, 1, 1}, {1, 1, 1}}, {{1, 0, 1}, {1, 1, 1}}, {{0, 0, 1}, {1, 0, 1}},
        {{0, 0, 0}, {0, 0, 1}}, {{0, 1, 0}, {0, 1, 1}}, {{1, 1, 0}, {1, 1, 1}}, {{1, 0, 0}, {1, 0, 1}}};
    seg* S = (seg*)malloc(sizeof(seg) * 12 * (n0 > 0 ? n0 : 1));
    for (int64_t i = 0; i < n0; ++i) {
        const double* o = s->origin + 3 * (s->esum_nodes[0] + i);
        const double w = s->width[s->esum_nodes[0] + i];
        for (int e = 0; e < 12; ++e)
            for (int a = 0; a < 3; ++a) {
                S[12 * i + e].a[a] = o[a] + (eo[e][0][a] ? w : 0.0);
                S[12 * i + e].b[a] = o[a] + (eo[e][1][a] ? w : 0.0);
            }
    }
    qsort(S, 12 * n0, sizeof(seg), cmp_seg);
    int64_t ns = 0;
    for (int64_t k = 0; k < 12 * n0; ++k)
        if (ns == 0 || cmp_seg(&S[k], &S[ns - 1]) != 0) S[ns++] = S[k];
    /* per axis, sorted by the end point (ThinShells.cpp:211-224) */
    seg* A[3];
    int64_t na[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) {
        A[a] = (seg*)malloc(sizeof(seg) * (ns > 0 ? ns : 1));
        for (int64_t k = 0; k < ns; ++k) {
            int axis = fabs(S[k].a[0] - S[k].b[0]) > 1e-12 ? 0 : (fabs(S[k].a[1] - S[k].b[1]) > 1e-12 ? 1 : 2);
            if (axis == a) A[a][na[a]++] = S[k];
        }
        g_axis = a;
        qsort(A[a], na[a], sizeof(seg), cmp_end);
    }
    free(S);
    /* mesh faces x node edges, swept per axis (ThinShells.cpp:242-338) */
    for (int64_t f = 0; f < nF; ++f) {
        const double* p1 = V + 3 * (int64_t)F[3 * f];
        const double* p2 = V + 3 * (int64_t)F[3 * f + 1];
        const double* p3 = V + 3 * (int64_t)F[3 * f + 2];
        double e1[3], e2[3], e3[3], u[3], v[3], n[3], bo[3], be[3];
        for (int a = 0; a < 3; ++a) {
            e1[a] = p2[a] - p1[a];
            e2[a] = p3[a] - p2[a];
            e3[a] = p1[a] - p3[a];
            u[a] = p2[a] - p1[a];
            v[a] = p3[a] - p1[a];
            bo[a] = fmin(p1[a], fmin(p2[a], p3[a]));
            be[a] = fmax(p1[a], fmax(p2[a], p3[a]));
        }
        cross3(u, v, n);
        const double mn[3] = {-n[0], -n[1], -n[2]};
        const double tdir = dot3l(mn, p1);
        for (int a = 0; a < 3; ++a) {
            /* lower_bound: first edge whose end coordinate is not below bo[a] by more than 1e-9 */
            int64_t lo = 0, hi = na[a];
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                double ea = A[a][mid].b[a];
                int less = (fabs(bo[a] - ea) >= 1e-9 && bo[a] > ea);
                if (less) lo = mid + 1; else hi = mid;
            }
            for (int64_t k = lo; k < na[a]; ++k) {
                const double* q1 = A[a][k].a;
                const double* q2 = A[a][k].b;
                if (fabs(q1[a] - be[a]) >= 1e-9 && q1[a] > be[a]) break;
                double ed[3] = {q2[0] - q1[0], q2[1] - q1[1], q2[2] - q1[2]};
                const double nd = dot3l(n, ed);
                if (fabsf((float)nd) < 1e-9f) continue;
                const double t = (-tdir - dot3l(n, q1)) / nd;
                if (t < 0. || t > 1.) continue;
                double ip[3] = {q1[0] + ed[0] * t, q1[1] + ed[1] * t, q1[2] + ed[2] * t};
                double w1[3], w2[3], w3[3], c[3];
                for (int b = 0; b < 3; ++b) { w1[b] = ip[b] - p1[b]; w2[b] = ip[b] - p2[b]; w3[b] = ip[b] - p3[b]; }
                cross3(e1, w1, c); if (dot3l(c, n) < 0) continue;
                cross3(e2, w2, c); if (dot3l(c, n) < 0) continue;
                cross3(e3, w3, c); if (dot3l(c, n) < 0) continue;
                push3(&Fp, ip);
            }
        }
    }
    for (int a = 0; a < 3; ++a) free(A[a]);
    const int64_t total = E.n + Fp.n;
    if (n_edge) *n_edge = E.n;
    if (out && total <= cap) {
        memcpy(out, E.p, 3 * sizeof(double) * E.n);
        memcpy(out + 3 * E.n, Fp.p, 3 * sizeof(double) * Fp.n);
    }
    free(E.p);
    free(Fp.p);
    return total;
}
