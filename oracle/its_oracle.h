/*
 * its_oracle.h -- FP64 CPU restatement of the ITS hot path (octree build,
 * hat B-spline evaluation, collocation assembly, LSCG, marching cubes, SDF).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in its_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * Every function names the reference file:line it restates.  Floating-point
 * convention (applies to the whole file and to the HIP kernels that must agree
 * bit for bit): IEEE binary64, no FMA contraction (built with
 * -ffp-contract=off), 3-vector dot products and squared norms evaluated as
 * a0*b0 + (a1*b1 + a2*b2) (the unrolled Eigen reduction order of device code),
 * except where the reference writes fma() explicitly (marching-cubes lerp).
 */
#ifndef ITS_ORACLE_H
#define ITS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_NONE 0xFFFFFFFFu
#define ORA_MAX_LEVELS 16

/* ---- Morton codec (src/MortonLUT.h:261-343) ---- */
uint32_t ora_morton_encode(uint32_t x, uint32_t y, uint32_t z);
void ora_morton_decode(uint32_t m, uint32_t* x, uint32_t* y, uint32_t* z);
/* closed-form regeneration of the reference tables, for pinning */
void ora_morton_tables(uint32_t enc_x[256], uint32_t enc_y[256], uint32_t enc_z[256],
                       uint8_t dec_x[512], uint8_t dec_y[512], uint8_t dec_z[512]);
void ora_neighbor_luts(int16_t lut_parent[8 * 27], int16_t lut_child[8 * 27],
                       int16_t vert_shared[64], int16_t edge_shared[48]);

/* ---- model normalisation and grid box (src/BaseModel.cpp:24-51, 54-92) ---- */
void ora_model_to_unit_cube(double* V, int64_t nV, double scale_factor);
void ora_uniform_bbox(const double* V, int64_t nV, double origin[3], double end[3], double width[3]);

/* ---- voxelisation (src/cuAcc/SVO.cu:366-515) ----
 * tris: T x 9 (p1,p2,p3).  Writes sorted unique Morton codes of occupied
 * voxels; returns the count (or -needed if cap is too small). */
int64_t ora_voxelize(const double* tris, int64_t T, const double origin[3], const double width[3],
                     int R, uint32_t* keys_out, int64_t cap);

/* ---- sparse voxel octree (src/cuAcc/SVO.cu:674-1113) ---- */
typedef struct ora_svo ora_svo;
ora_svo* ora_svo_build(const double* tris, int64_t T, const double origin[3], const double width[3], int R);
void ora_svo_free(ora_svo* s);
int ora_svo_depth(const ora_svo* s);
int64_t ora_svo_num_nodes(const ora_svo* s);
int64_t ora_svo_num_verts(const ora_svo* s);
int64_t ora_svo_num_voxels(const ora_svo* s);
/* depth_num_nodes[L], esum_verts[L+1] */
void ora_svo_level_info(const ora_svo* s, int64_t* depth_num_nodes, int64_t* esum_verts);
void ora_svo_nodes(const ora_svo* s, uint32_t* morton, uint8_t* is_leaf, double* origin, double* width,
                   uint32_t* parent, uint32_t* childs, uint32_t* neighbors);
void ora_svo_verts(const ora_svo* s, double* xyz, uint32_t* owner);

/* fine node edges with owners, in the order of the reference's
 * fineNodeEdgeArray (src/cuAcc/SVO.cu:965-1008,1033-1060,1098-1110): 6
 * doubles and an owner index per edge; writes up to cap, returns the count */
int64_t ora_fine_edges(const ora_svo* s, double* ends, uint32_t* owner, int64_t cap);

/* ---- hat basis (src/BSpline.hpp:5-27) ---- */
double ora_base_function(double x, double c, double w);
double ora_base_function3(const double c[3], double w, const double q[3]);

/* ---- collocation assembly (src/ThinShells.cpp:511-605) ----
 * CSR with rows sorted by column.  Returns nnz; arrays may be NULL to query. */
typedef struct ora_mat ora_mat;
ora_mat* ora_assemble(const ora_svo* s);
void ora_mat_free(ora_mat* A);
int64_t ora_mat_nnz(const ora_mat* A);
int64_t ora_mat_n(const ora_mat* A);
void ora_mat_csr(const ora_mat* A, int64_t* rowptr, int32_t* cols, double* vals);
void ora_mat_mul(const ora_mat* A, const double* x, double* y);   /* y = A x  */
void ora_mat_tmul(const ora_mat* A, const double* x, double* y);  /* y = A^T x */

/* ---- LSCG (Eigen LeastSquaresConjugateGradient as called at ThinShells.cpp:615-617) ---- */
typedef struct {
    int64_t iterations;      /* Eigen's iteration count */
    double normal_rel;       /* ||A^T(b-Ax)|| / ||A^T b|| */
    double residual_rel;     /* ||b-Ax|| / ||b|| */
    double residual_abs;     /* ||Ax-b|| (printed by the reference) */
} ora_lscg_stats;
/* x holds the initial guess on entry (zero = Eigen's default solve()). */
int ora_lscg(const ora_mat* A, const double* b, double* x, double tol, int64_t max_iter, ora_lscg_stats* st);

/* ---- evaluation of f ----
 * support: multiPointQuery CPU path mt_cpuTest_2 (ThinShells.cpp:1249-1294,
 *          SVO.cu:1195-1261); points outside the open box get out_val.
 * latent:  cpLatentBSplineValue (ThinShells.cpp:655-701); returns -1 if a
 *          point's fine voxel is not a level-0 node (the reference throws).
 * dense:   sum over every vertex, the GPU formulation CUDACompute.cu:468 and
 *          the marching-cubes corner evaluation; O(N n), small inputs only.
 * dense_fast: the same dense value computed through per-level support
 *          (clamped cells); equal to dense for every point. */
void ora_eval_support(const ora_svo* s, const double* lambda, const double* pts, int64_t N, const double box_end[3],
                      double out_val, double* f);
int ora_eval_latent(const ora_svo* s, const double* lambda, const double* pts, int64_t N, double* f);
void ora_eval_dense(const ora_svo* s, const double* lambda, const double* pts, int64_t N, double* f);
void ora_eval_dense_fast(const ora_svo* s, const double* lambda, const double* pts, int64_t N, double* f);

/* ---- marching cubes (src/cuAcc/MarchingCubes/MarchingCubes.cu:72-377, 672-714) ---- */
typedef struct ora_mc ora_mc;
ora_mc* ora_marching_cubes(const ora_svo* s, const double* lambda, const double origin[3], const double width[3],
                           const int res[3], double iso);
/* the voxels of the z-slab [z0, z1) of the same grid only: their cube
 * indices and the piece of the whole grid's soup they emit */
ora_mc* ora_marching_cubes_slab(const ora_svo* s, const double* lambda, const double origin[3],
                                const double width[3], const int res[3], double iso, int z0, int z1);
void ora_mc_free(ora_mc* m);
int64_t ora_mc_num_verts(const ora_mc* m);           /* soup vertices (3 per triangle) */
int64_t ora_mc_num_valid_voxels(const ora_mc* m);
void ora_mc_soup(const ora_mc* m, double* xyz);
void ora_mc_cube_index(const ora_mc* m, uint8_t* cube);  /* res^3 cube indices, x fastest */
const int* ora_mc_tri_table(void);                       /* 256 x 16 */
const int* ora_mc_nverts_table(void);                    /* 256 */

/* ---- signed distance (RHS of the fit, ThinShells.cpp:481-509) ----
 * unsigned distance by brute force over triangles, sign negative where the
 * solid-angle winding number exceeds 1/2. */
void ora_signed_distance(const double* V, int64_t nV, const int32_t* F, int64_t nF,
                         const double* pts, int64_t n, double* out);
double ora_winding_number(const double* V, const int32_t* F, int64_t nF, const double p[3]);

/* ---- shell iso band (ThinShells::cpIntersectionPoints, ThinShells.cpp:38-479) ----
 * Points where mesh edges cross the back/left/bottom faces of level-0 nodes
 * (sorted with the fuzzy order and exact duplicates removed), then points
 * where level-0 node edges cross mesh triangles.  Writes up to cap points
 * (x, y, z) and returns the total count; n_edge receives the edge points. */
int64_t ora_intersections(const ora_svo* s, const double* V, int64_t nV, const int32_t* F, int64_t nF, double* out,
                          int64_t cap, int64_t* n_edge);

int ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
