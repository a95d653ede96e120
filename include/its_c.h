/*
 * its_c.h -- C ABI of the MI355X (gfx950) implementation of the ITS hot path.
 *
 * Conventions
 *   - Every function returns ITS_OK (0) or a negative ITS_E_* code; the text of
 *     the last error on the calling thread is returned by its_last_error().
 *     Nothing crosses this boundary as an exception or exit().
 *   - Host pointers (suffix-free arguments) are caller owned; the library copies
 *     what it needs.  Entry points with the suffix _device take device pointers
 *     and a hipStream_t (passed as void*; NULL means the context stream) and do
 *     not synchronise.
 *   - Opaque handles own device memory; one context per host thread.
 *   - Floating point: octree geometry is binary64 and bit-identical to the CPU
 *     restatement of the reference; B-spline evaluation and the LSCG matvec run
 *     in binary32 with binary64 reductions, residuals and refinement.
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef ITS_C_H
#define ITS_C_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ITS_OK 0
#define ITS_E_INVALID (-1)   /* bad argument */
#define ITS_E_HIP (-2)       /* HIP runtime error */
#define ITS_E_NOMEM (-3)     /* device or host allocation failed */
#define ITS_E_STATE (-4)     /* call out of order (e.g. eval before a fit) */
#define ITS_E_RANGE (-5)     /* size beyond a documented limit */
#define ITS_E_NOTFOUND (-6)  /* a point's fine voxel is not an octree node */

#define ITS_NONE 0xFFFFFFFFu /* "no node" (UINT_MAX in the reference) */
#define ITS_MAX_LEVELS 12

typedef struct its_ctx_s* its_ctx;
typedef struct its_svo_s* its_svo;

const char* its_last_error(void);
const char* its_version(void);

/* ---- context (replaces the per-call device selection and cudaMalloc/free
 *      of src/cuAcc/CUDACompute.cu:950-982 and DeviceQuery.cuh:41-56) ---- */
int its_ctx_create(int device, its_ctx* out);
int its_ctx_destroy(its_ctx ctx);
/* make the library launch on an external stream (e.g. torch's current stream) */
int its_ctx_set_stream(its_ctx ctx, void* hip_stream);
int its_ctx_synchronize(its_ctx ctx);
int its_device_count(int* count);

/* ---- model normalisation and grid box (BaseModel::model2UnitCube,
 *      BaseModel.cpp:84-92; BaseModel::setUniformBoundingBox, 24-51) ---- */
int its_model_to_unit_cube(double* V, int64_t nV, double scale_factor);
int its_uniform_bbox(const double* V, int64_t nV, double origin[3], double end[3], double width[3]);

/* ---- sparse voxel octree (SparseVoxelOctree::createOctree, SVO.h:104,
 *      SVO.cu:674-856, 858-1113) ----
 * tris: T x 9 binary64 (p1, p2, p3), as the Triangle<V3d> list.
 * origin/width: the grid box (AABox boxOrigin / boxWidth).  R: voxels per axis,
 * a power of two in [2, 1024] (the reference's surfaceVoxelGridSize, cubic). */
int its_svo_build(its_ctx ctx, const double* tris, int64_t T, const double origin[3], const double width[3], int R,
                  its_svo* out);
int its_svo_build_device(its_ctx ctx, const double* d_tris, int64_t T, const double origin[3], const double width[3],
                         int R, its_svo* out);
int its_svo_destroy(its_svo svo);
/* An octree object that holds only a vertex set, from the arrays the
 * reference passes to MC::marching_cubes (MarchingCubes.h:136-139):
 * nodeVertexArray as xyz (n_verts x 3) and owner node (n_verts), and
 * nodeWidthArray (n_nodes x 3, the x component is used).  A vertex's level
 * follows from its owner's width (the smallest width times 2^level), the grid
 * origin is the smallest coordinate of each axis (the root's first corner).
 * Such an octree supports its_svo_get_info, its_svo_export_verts and
 * its_svo_export_vert_lattice (level / lattice order), coefficients
 * (its_svo_set_lambda / get_lambda, in the caller's vertex order),
 * evaluation, marching cubes and the lattice entry points; the others
 * return ITS_E_STATE.  Every vertex must be a corner of its level's lattice,
 * and no two vertices may share a level and a lattice point (ITS_E_INVALID). */
int its_svo_from_vertices(its_ctx ctx, const double* xyz, const uint32_t* owner, int64_t n_verts,
                          const double* node_width, int64_t n_nodes, its_svo* out);

typedef struct {
    int tree_depth;                               /* treeDepth (levels), finest = 0 */
    int R;
    int64_t num_voxels;                           /* numVoxels */
    int64_t num_nodes;                            /* numTreeNodes */
    int64_t num_verts;                            /* numNodeVerts (= unknowns n) */
    int64_t depth_num_nodes[ITS_MAX_LEVELS];      /* depthNumNodes */
    int64_t esum_nodes[ITS_MAX_LEVELS + 1];
    int64_t esum_verts[ITS_MAX_LEVELS + 1];       /* esumDepthNodeVerts */
    double grid_origin[3];
    double voxel_width;                           /* svoNodeArray[0].width */
} its_svo_info;
int its_svo_get_info(its_svo svo, its_svo_info* info);

/* SVONode fields (SVO.h:15-32) as SoA, global order finest -> root.  Any
 * pointer may be NULL.  childs: 8 per node, neighbors: 27 per node. */
int its_svo_export_nodes(its_svo svo, uint32_t* morton, uint8_t* is_leaf, double* origin, double* width,
                         uint32_t* parent, uint32_t* childs, uint32_t* neighbors);
/* nodeVertexArray (SVO.h:53): corner position and owner node, level by level */
int its_svo_export_verts(its_svo svo, double* xyz, uint32_t* owner);
/* sorted occupied-voxel Morton codes (the compacted level-0 array) */
int its_svo_export_voxels(its_svo svo, uint32_t* morton);
/* per vertex (global order): its lattice coordinates at its own level, the
 * key the vertex order of depthNodeVertexArray sorts by (x, then y, then z;
 * SVO.cu:1064-1096), and that level.  Either array may be NULL. */
int its_svo_export_vert_lattice(its_svo svo, uint16_t* ijk, uint8_t* level);
/* level-0 node edges, the reference's fineNodeEdgeArray (determineNodeEdge,
 * sortEdge, uniqueEdge: src/cuAcc/SVO.cu:965-1008, 1033-1060, 1098-1110):
 * end points (6 doubles per edge, lower end first) and the smallest index of
 * the level-0 nodes containing the edge, in the reference's order.  ends and
 * owner may be NULL to count; capacity is in edges. */
int its_svo_export_fine_edges(its_svo svo, double* ends, uint32_t* owner, int64_t capacity, int64_t* count);

/* ---- right-hand side: signed distance at every node vertex
 *      (ThinShells::cpSDFOfTreeNodes, ThinShells.cpp:481-509) ----
 * V: nV x 3, F: nF x 3.  b: n = num_verts values (host).  Also keeps the
 * result and the nearest-triangle index on the device for the fit. */
int its_sdf_at_vertices(its_ctx ctx, its_svo svo, const double* V, int64_t nV, const int32_t* F, int64_t nF,
                        double* b);
/* general points (N x 3) */
int its_signed_distance(its_ctx ctx, const double* V, int64_t nV, const int32_t* F, int64_t nF, const double* pts,
                        int64_t N, double* out);

/* ---- coefficient fit (ThinShells::cpCoefficients solve step,
 *      Eigen LeastSquaresConjugateGradient at ThinShells.cpp:615-617) ---- */
typedef struct {
    int64_t iterations;       /* inner CGLS iterations (all refinement steps) */
    int64_t refinements;      /* binary64 residual evaluations after the first */
    double normal_rel;        /* ||A^T(b - A lambda)|| / ||A^T b||, binary64 */
    double residual_rel;      /* ||b - A lambda|| / ||b||, binary64 */
    double residual_abs;      /* ||A lambda - b|| (printed by the reference) */
    double initial_normal_rel;/* same quantity at the starting guess */
    double seconds;           /* wall time of the solve on the device stream */
    int converged;
} its_lscg_stats;
/* b: host RHS (NULL = use the device RHS from its_sdf_at_vertices).
 * lambda_inout: host n-vector; read as the initial guess when warm != 0,
 * otherwise the solve starts from 0 (Eigen's solve()); always written.
 * tol: stopping tolerance on normal_rel; max_iter <= 0 means 2n (Eigen's
 * default).  The solution also becomes the octree's coefficient field. */
int its_lscg_solve(its_ctx ctx, its_svo svo, const double* b, double* lambda_inout, double tol, int64_t max_iter,
                   int warm, its_lscg_stats* stats);

/* Matrix-free products with the collocation matrix A_ij = B_j(v_i)
 * (the assembled matrix of ThinShells.cpp:556-605), binary64, host vectors. */
int its_collocation_mul(its_ctx ctx, its_svo svo, const double* x, double* y, int transpose);
/* Sizes of the operator: n, the non-zeros of the matrix the reference
 * assembles (ThinShells.cpp:556-605, counted on the device), the entries of
 * the matrix-free operator's parent lists, the levels (and how many of them
 * the one-workgroup kernels take), and the bytes one CGLS iteration moves:
 * 16 B per non-zero + 72 B per row for the assembled matrix, and this
 * library's matrix-free count (89 B per vertex + 16 B per parent entry). */
typedef struct {
    int64_t n;
    int64_t nnz;
    int64_t parent_entries;
    int levels;
    int block_levels;
    double bytes_assembled_per_iteration;
    double bytes_matrix_free_per_iteration;
} its_lscg_structure;
int its_lscg_structure_get(its_ctx ctx, its_svo svo, its_lscg_structure* out);
/* The solver's Jacobi diagonal 1/||A_:j||^2 (1 for an empty column), the
 * LeastSquareDiagonalPreconditioner of the Eigen solve; host n-vector. */
int its_lscg_diag(its_ctx ctx, its_svo svo, double* invd);

/* ---- coefficient field used by evaluation and marching cubes ---- */
int its_svo_set_lambda(its_ctx ctx, its_svo svo, const double* lambda);
int its_svo_get_lambda(its_svo svo, double* lambda);

/* ---- evaluation of f(x) = sum_i lambda_i B_i(x) ----
 * mode ITS_EVAL_SUPPORT: the CPU query path mt_cpuTest_2
 *   (ThinShells.cpp:1249-1294): points outside the open box (origin, end)
 *   get out_val, others the octree-support sum.
 * mode ITS_EVAL_DENSE: cuAcc::cpPointQuery (CUDACompute.cu:1070-1156):
 *   points outside [origin - width, end + width] get out_val, others the sum
 *   over every basis function (equal to the support sum).
 * mode ITS_EVAL_RAW: the sum over every basis function everywhere
 *   (cuAcc::cpBSplineVal, CUDACompute.cu:946-1068; marching-cubes corners).
 * box_end: AABox boxEnd. */
#define ITS_EVAL_SUPPORT 0
#define ITS_EVAL_DENSE 1
#define ITS_EVAL_RAW 2
int its_bspline_eval(its_ctx ctx, its_svo svo, const float* pts, int64_t N, const double box_end[3], int mode,
                     float out_val, float* f);
int its_bspline_eval_f64pts(its_ctx ctx, its_svo svo, const double* pts, int64_t N, const double box_end[3], int mode,
                            float out_val, float* f);
/* device SoA coordinates x[N], y[N], z[N]; f[N] */
int its_bspline_eval_device(its_ctx ctx, its_svo svo, const float* d_x, const float* d_y, const float* d_z, int64_t N,
                            const double box_end[3], int mode, float out_val, float* d_f, void* stream);
/* ---- shell iso band (ThinShells::cpIntersectionPoints, ThinShells.cpp:38-479,
 *      and the iso values of cpLatentBSplineValue, 655-755) ----
 * Points where the mesh edges cross the back/left/bottom faces of level-0
 * nodes, then points where level-0 node edges cross the triangles (edge
 * points first; duplicates are kept, order is deterministic).  pts may be NULL
 * to count; capacity is in points. */
int its_shell_points(its_ctx ctx, its_svo svo, const double* V, int64_t nV, const int32_t* F, int64_t nF,
                     double* pts, int64_t capacity, int64_t* n_edge, int64_t* n_face);
/* inner / outer iso values: min and max of f over the model vertices and
 * those points (needs the fitted field); n_points: points evaluated */
int its_shell_iso(its_ctx ctx, its_svo svo, const double* V, int64_t nV, const int32_t* F, int64_t nF,
                  double iso[2], int64_t* n_points);

/* {-1, 0, 1} classification of f (ThinShells.cpp:1412-1417) */
int its_classify(const float* f, int64_t N, double inner_iso, double outer_iso, int8_t* cls);

/* ---- marching cubes (MC::marching_cubes, MarchingCubes.h:136-139,
 *      MarchingCubes.cu:672-714; the three iso values of mcVisualization,
 *      ThinShells.cpp:1086-1129) ----
 * Grid of res[0] x res[1] x res[2] voxels over origin + [0, width).
 * lambda: new coefficients (host n-vector, as its_svo_set_lambda) or NULL
 * for the octree's current ones.  iso: n_iso values (1..ITS_MC_MAX_ISO); one
 * soup per value, concatenated in that order, each as float xyz, 3 vertices
 * per triangle, in the reference order (voxels x-fastest, triangles in table
 * order).  num_verts (may be NULL): n_iso counts.  tri_xyz may be NULL to
 * query the counts; capacity is in vertices (all soups together).  stats:
 * sums over the iso values.  The octree keeps the last lattice, so a fill
 * call after its count call, and every iso value after the first on one
 * grid, evaluate no lattice. */
#define ITS_MC_MAX_ISO 16
typedef struct {
    int64_t num_verts;        /* allTriVertices */
    int64_t num_valid_voxels; /* nValidVoxels */
    double seconds_eval;      /* lattice evaluation (host clock); near 0 when the
                                 lattice of the previous call on this octree was
                                 reused (same grid, coefficients unchanged) */
    double seconds_extract;   /* count, scan and emission, with the copy of the
                                 soup to the caller's buffer (host clock) */
    double seconds_eval_device;      /* lattice evaluation, between events on the
                                        context stream */
    double seconds_extract_device;   /* count, scan and emission on the stream,
                                        without the copy to the host */
} its_mc_stats;
int its_mc_extract(its_ctx ctx, its_svo svo, const double* lambda, const double origin[3], const double width[3],
                   const int res[3], const double* iso, int n_iso, float* tri_xyz, int64_t capacity,
                   int64_t* num_verts, its_mc_stats* stats);
/* lattice values f at the (res+1)^3 grid points, x fastest (for checking) */
int its_mc_lattice(its_ctx ctx, its_svo svo, const double origin[3], const double width[3], const int res[3],
                   float* values);
/* binary64 values of the field on the octree's own level-0 lattice, (R+1)^3
 * points over the grid box, x fastest: the lattice of mcVisualization at the
 * octree resolution, without rounding to binary32 (for precision reports). */
int its_lattice_f64(its_ctx ctx, its_svo svo, double* values);

/* ---- learned warm start (no reference counterpart: the reference starts
 *      its solve from zero, ThinShells.cpp:615-617) ----
 * A network of 4 dense layers (16 -> 64 -> 64 -> 64 -> 1, GELU) predicts, for
 * every vertex, a row-space coefficient w_i of the fit, and the solve starts
 * from lambda0 = D A^T w with D = diag(1/||A_:j||^2).  Jacobi-preconditioned
 * CGLS keeps its iterates in lambda0 + range(D A^T), so from such a start it
 * converges to the same lambda as from zero: the prediction changes the
 * iteration count, never the fitted field.
 * Parameters (binary32 master copy, ITS_MLP_NPARAM values): W1[16][64], b1[64],
 * W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4; W[k][n] maps input k to
 * output n.  BF16 MFMA operands, binary32 accumulation. */
#define ITS_MLP_IN 16
#define ITS_MLP_HIDDEN 64
#define ITS_MLP_NPARAM 9473
typedef struct its_mlp_s* its_mlp;
int its_mlp_create(its_ctx ctx, uint64_t seed, its_mlp* out);   /* Glorot-uniform weights, zero biases */
int its_mlp_destroy(its_mlp net);
int its_mlp_set_params(its_mlp net, const float* params);
int its_mlp_get_params(its_mlp net, float* params);
/* d_x: rows x 16 bfloat16 (raw 16-bit patterns), d_y: rows outputs */
int its_mlp_forward_device(its_mlp net, const uint16_t* d_x, int64_t rows, float* d_y, void* stream);
/* gradient of mean((y - t)^2) over the rows: grad (host, ITS_MLP_NPARAM), loss = the mean */
int its_mlp_grad_device(its_mlp net, const uint16_t* d_x, const float* d_t, int64_t rows, float* grad, double* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that gradient; loss before the step */
int its_mlp_train_step(its_mlp net, const uint16_t* d_x, const float* d_t, int64_t rows, float lr, double* loss);
/* features of every vertex (n x 16 bfloat16) and the cell width of its level;
 * needs its_sdf_at_vertices */
int its_warm_features(its_ctx ctx, its_svo svo, uint16_t* d_x, float* d_scale, void* stream);
/* dual coefficients w of a solution lambda (lambda = D A^T w), equal shares
 * within each group of coincident vertices; host n-vectors */
int its_warm_dual(its_ctx ctx, its_svo svo, const double* lambda, double* w);
/* lambda0 = D A^T w; host n-vectors */
int its_warm_apply(its_ctx ctx, its_svo svo, const double* w, double* lambda0);
/* features -> network -> lambda0 -> the unmodified LSCG of its_lscg_solve
 * (b: NULL = device RHS).  seconds_predict: device time of the prediction;
 * stats->seconds: the solve alone. */
int its_lscg_solve_mlp(its_ctx ctx, its_svo svo, its_mlp net, const double* b, double* lambda, double tol,
                       int64_t max_iter, its_lscg_stats* stats, double* seconds_predict);

/* ---- many meshes per launch (no reference counterpart: the reference fits
 *      one model per process) ----
 * A batch holds octrees, each with its right-hand side
 * (its_sdf_at_vertices).  its_batch_fit solves the meshes with at most
 * ITS_TUNE_BATCH_MAX_VERTS unknowns in one persistent launch (one workgroup
 * per mesh runs that mesh's whole mixed-precision LSCG, the algorithm and
 * stopping rule of its_lscg_solve), captured once as a hipGraph and replayed
 * on later calls; the larger meshes run at the same time through the
 * many-mesh solver (one launch per solver pass for all of them; per mesh the
 * coefficients and statistics of its_lscg_solve, bit for bit).
 * stats: one per mesh; seconds: wall time of the fit.  warm != 0 starts
 * every mesh from its current coefficients.  The coefficients stay on the
 * octrees; their evaluation fields are rebuilt by its_batch_fields (octrees
 * of different contexts side by side).  its_batch_eval_device evaluates f for all meshes in one
 * launch: queries [offsets[m], offsets[m+1]) of the device SoA arrays belong
 * to mesh m (offsets: host, count + 1 values), support semantics inside each
 * octree's box, out_val outside. */
typedef struct its_batch_s* its_batch;
int its_batch_create(its_ctx ctx, its_batch* out);
int its_batch_destroy(its_batch b);
/* svo may come from any context on the batch's device (for instance octrees
 * built by several host threads, one context each); the fit synchronises
 * their streams before reading them. */
int its_batch_add(its_batch b, its_svo svo);
int its_batch_fit(its_batch b, double tol, int64_t max_iter, int warm, its_lscg_stats* stats, double* seconds);
int its_batch_fields(its_batch b);
int its_batch_eval_device(its_batch b, const float* d_x, const float* d_y, const float* d_z, const int64_t* offsets,
                          float out_val, float* d_f, double* seconds);

/* its_batch_eval_device without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-mesh parameters and work items (a mesh's queries cut into chunks of
 * ITS_NOALLOC_CHUNK) live in a fixed device array of the library, one per
 * device, about 1.9 MB, set up by its_ctx_create and held until the process
 * ends; host-side staging is static too.  Limits: at most
 * ITS_NOALLOC_MAX_MESHES meshes and ITS_NOALLOC_MAX_ITEMS chunks per call,
 * else ITS_E_RANGE before anything is launched.  Results are bit-identical to
 * its_batch_eval_device.  Calls share the array of their device: calls from
 * one process are serialised by the library and each waits for its launch;
 * calls from several processes on one GPU each have their own array.  The
 * fields must already exist (its_batch_fields). */
#define ITS_NOALLOC_MAX_MESHES 1024
#define ITS_NOALLOC_MAX_ITEMS 65536
#define ITS_NOALLOC_CHUNK 16384
int its_batch_eval_device_noalloc(its_batch b, const float* d_x, const float* d_y, const float* d_z,
                                  const int64_t* offsets, float out_val, float* d_f, double* seconds);

/* ---- tunables (process-wide; set by the autotuner, its_amd/autotune.py) ----
 * ITS_TUNE_DENSE_CELLS: the evaluation field collapses every level whose
 *   cell count is at most this into one dense block grid (default 32768).
 * ITS_TUNE_EVAL_BLOCKS_PER_CU: persistent evaluation grid, workgroups per CU
 *   (default 64, at most 64).
 * ITS_TUNE_COARSE_VERTS: the LSCG cascade runs the top levels holding at most
 *   this many vertices in one workgroup (default 2048, at most 12288).
 * ITS_TUNE_EVAL_ILP: queries carried per evaluation thread, 1, 2 or 4
 *   (default 1), when the queue is off.
 * ITS_TUNE_EVAL_QUEUE: 1 = queries that need the fine octree levels are
 *   gathered per wavefront in LDS and walked 64 at a time; 0 = every thread
 *   walks its own query (default 1).
 * ITS_TUNE_BATCH_MAX_VERTS: meshes of a batch with at most this many
 *   unknowns share the one-launch fit (default 65536).
 * ITS_TUNE_SDF_RAY: 1 = for a closed, consistently oriented mesh the sign of
 *   the signed distance comes from the crossings of a ray where exact
 *   orientation tests decide it (the winding-number fan sum for the other
 *   points); 0 = the fan sum for every point (default 1; same results).
 * No reference counterpart (the reference hard-codes its launch shapes). */
#define ITS_TUNE_DENSE_CELLS 0
#define ITS_TUNE_EVAL_BLOCKS_PER_CU 1
#define ITS_TUNE_COARSE_VERTS 2
#define ITS_TUNE_EVAL_ILP 3
#define ITS_TUNE_EVAL_QUEUE 4
#define ITS_TUNE_BATCH_MAX_VERTS 5
#define ITS_TUNE_SDF_RAY 6
#define ITS_TUNE_COUNT 7
int its_tune_set(int key, int64_t value);
int its_tune_get(int key, int64_t* value);

/* ---- timing helpers on the context stream (hipEvents) ---- */
int its_timer_start(its_ctx ctx);
int its_timer_stop(its_ctx ctx, double* ms);

/* ---- achievable HBM bandwidth of the context's device ----
 * Copies `bytes` (rounded down to a multiple of 16; at least 1 MiB, at most
 * 16 GiB) device to device with the library's own streaming kernel: two
 * untimed launches, then `reps` (1..100) launches timed one by one with
 * events on the context stream.  *gbps = 2 * bytes / median launch time
 * (bytes read plus bytes written), in 1e9 B/s.  The call allocates and frees
 * its two buffers.  Kernels are reported against this as well as against the
 * published peak.  No reference counterpart. */
int its_hbm_bandwidth(its_ctx ctx, int64_t bytes, int reps, double* gbps);

#ifdef __cplusplus
}
#endif
#endif
