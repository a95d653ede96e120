// Internal object layout of the ITS gfx950 library (not part of the C ABI).
#pragma once
#include <memory>
#include <vector>

#include "common.hpp"

namespace its {

constexpr int kMaxLevels = ITS_MAX_LEVELS;

// One octree level (0 = finest voxels).
struct Level {
    int64_t n_nodes = 0;    // nodes (octets of children of occupied parents)
    int64_t n_occ = 0;      // occupied cells (sorted Morton codes)
    int64_t n_verts = 0;    // distinct node corners
    int R = 1;              // cells per axis at this level
    double width = 0;       // node width
    DBuf<uint32_t> occ;     // occupied cell codes, ascending
    DBuf<uint32_t> prank;   // rank of each occupied cell's parent among the parents
    DBuf<RankWord> lattice; // corner lattice (R+1)^3 bits, index (x*(R+1)+y)*(R+1)+z
    // evaluation field: cells whose corners include a vertex
    DBuf<RankWord> cells;   // R^3 bits by Morton code
    DBuf<float4> blocks;    // 2 float4 (8 corner coefficients) per marked cell; on the
                            // linked levels the 5 low bits of each carry the child link
    int64_t n_cells = 0;
};

struct Fit;   // matrix-free collocation operator and solver state (lscg.hip)
struct McScratch;

}  // namespace its

struct its_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;            // its_timer_start / its_timer_stop
    hipEvent_t mc_ev[3] = {nullptr, nullptr, nullptr};   // marching-cubes phase times
    hipDeviceProp_t prop{};
};

struct its_svo_s {
    its_ctx ctx = nullptr;
    // built from a vertex set (its_svo_from_vertices): vertices, levels and
    // the coefficient field only, no nodes, voxels or solver
    bool vertices_only = false;
    // vertices-only octrees: internal index of each caller vertex when the
    // caller's order is not the level / lattice order (empty: the same order)
    std::vector<int64_t> vperm;
    int depth = 0;
    int R = 0;
    double origin[3] = {0, 0, 0};
    double width[3] = {0, 0, 0};
    double voxel_width = 0;
    int64_t num_voxels = 0, num_nodes = 0, num_verts = 0;
    int64_t esum_nodes[its::kMaxLevels + 1] = {0};
    int64_t esum_verts[its::kMaxLevels + 1] = {0};
    its::Level lv[its::kMaxLevels];

    // nodes, global order finest -> root
    its::DBuf<uint32_t> morton, parent, childbase, neighbors, node_corner;
    its::DBuf<uint8_t> is_leaf;
    its::DBuf<double> node_origin, node_width;
    // vertices, global order level by level, lexicographic within a level
    its::DBuf<double> vxyz;
    its::DBuf<uint32_t> vowner;
    its::DBuf<uint16_t> vx, vy, vz;   // lattice coordinates at the vertex's level
    its::DBuf<uint8_t> vlevel;

    // right-hand side and coefficient field
    its::DBuf<double> rhs;
    its::DBuf<int32_t> rhs_tri;     // nearest triangle per vertex
    its::DBuf<float4> rhs_nrm;      // unit normal of that triangle
    bool has_rhs = false;
    its::DBuf<double> lambda64;
    bool has_lambda = false;
    bool field_ready = false;
    its::DBuf<float4> dense_blocks;   // per dense_level cell: corner values of levels >= dense_level,
                                      // child link in the 5 low bits of each
    int dense_level = 0;
    // bumped before anything may change lambda64 or the evaluation field;
    // a kept marching-cubes lattice is reused only at the same version
    uint64_t lambda_version = 0;
    struct {
        bool valid;
        uint64_t version;
        double origin[3], width[3];
        int res[3];
    } mc_key{};
    // grow-only work space: two binary64 levels of the dense cascade, and the
    // marching-cubes lattice, cube cases, per-step vertex counts and soup
    its::DBuf<double> casc[2];
    its::DBuf<float> mc_lat, mc_soup;
    its::DBuf<uint64_t> mc_step;
    its::DBuf<unsigned long long> mc_valid;   // [0] voxels with triangles; [1] vertices
    its::DBuf<uint64_t> mc_list;              // voxels with triangles: step, lane, case, offset in the step

    std::unique_ptr<its::Fit, void (*)(its::Fit*)> fit{nullptr, nullptr};
};

namespace its {
// call before writing lambda64 or rebuilding the field
inline void lambda_changing(its_svo s) { ++s->lambda_version; }
// tunables (capi.hip)
int64_t tune_get(int key);
// svo.hip
void svo_build(its_svo s, const double* d_tris, int64_t T);
void svo_from_vertices(its_svo s, const double* xyz, const uint32_t* owner, int64_t n, const double* node_width,
                       int64_t n_nodes);
// eval.hip
void field_build(its_svo s);   // cells + blocks from lambda64
void dense_cascade(its_svo s, int stop, DBuf<float>& out);
void dense_cascade64(its_svo s, int stop, DBuf<double>& out);
void eval_points(its_svo s, const float* x, const float* y, const float* z, int64_t N, const double box_end[3],
                 int mode, float out_val, float* f, hipStream_t st);
void eval_points_aos32(its_svo s, const float* pts, int64_t N, const double box_end[3], int mode, float out_val,
                       float* f, hipStream_t st);
void eval_points_aos64(its_svo s, const double* pts, int64_t N, const double box_end[3], int mode, float out_val,
                       float* f, hipStream_t st);
// lscg.hip
void fit_release(Fit* f);
// on `stream` when given (the octree's solver state must exist already,
// fit_get), else on the context stream
void lscg_solve(its_svo s, const double* d_b, double* d_lambda, double tol, int64_t max_iter, bool warm,
                its_lscg_stats* st, hipStream_t stream = nullptr);
void collocation_mul(its_svo s, const double* d_x, double* d_y, bool transpose);
void lscg_structure(its_svo s, its_lscg_structure* out);
// mc.hip
void mc_extract(its_svo s, const double origin[3], const double width[3], const int res[3], double iso,
                float* h_tri, int64_t capacity, its_mc_stats* st);
void mc_lattice(its_svo s, const double origin[3], const double width[3], const int res[3], float* h_vals);
// sdf.hip
void signed_distance(hipStream_t st, const double* d_V, int64_t nV, const int32_t* d_F, int64_t nF,
                     const double* d_pts, int64_t N, double* d_out, int32_t* d_tri);
void nearest_normals(hipStream_t st, const double* d_V, const int32_t* d_F, const int32_t* d_tri, int64_t N,
                     float4* d_nrm);
void eval_batch(its_svo const* svos, int64_t count, const int64_t* off, const float* x, const float* y,
                const float* z, float out_val, float* f, hipStream_t st);
// the same launch with its tables in the file-scope scratch array of eval.hip
constexpr int64_t kNoAllocMeshes = ITS_NOALLOC_MAX_MESHES;
constexpr int64_t kNoAllocItems = ITS_NOALLOC_MAX_ITEMS;
void eval_scratch_init(int device);   // once per device, with that device current
void eval_batch_noalloc(its_svo const* svos, int64_t count, const int64_t* off, const float* x, const float* y,
                        const float* z, float out_val, float* f, hipStream_t st);
// shell.hip
// level-0 node edges with owners in the reference's fineNodeEdgeArray order;
// returns the count
int64_t fine_edges(its_svo s, DBuf<double>& ends, DBuf<uint32_t>& owner, hipStream_t st);
void shell_points(its_svo s, const double* d_V, const uint64_t* d_edges, int64_t ne, const double* d_tris, int64_t nT,
                  DBuf<double>& pts, int64_t* n_edge, int64_t* n_face, hipStream_t st);
std::vector<uint64_t> mesh_edges(const int32_t* F, int64_t nF);
// batch.hip
its_batch batch_new(its_ctx ctx);
void batch_delete(its_batch b);
void batch_add(its_batch b, its_svo s);
void batch_fit(its_batch b, double tol, int64_t max_iter, bool warm, its_lscg_stats* out, double* seconds);
void batch_fields(its_batch b);
int64_t batch_count(its_batch b);
its_svo batch_svo(its_batch b, int64_t m);
its_ctx batch_ctx(its_batch b);
// lscg.hip: warm start in the row space of A
void warm_dual(its_svo s, const double* d_lam, double* d_w);
void warm_apply(its_svo s, const double* d_w, double* d_lam0, hipStream_t st);
void lscg_diag(its_svo s, double* d_out, hipStream_t st);
// warm.hip
// d_err: when given, the structure check's flag is left there (cleared
// first) for the caller to read later; otherwise it is read before returning
void warm_features(its_svo s, uint16_t* d_x, float* d_scale, hipStream_t st, int* d_err = nullptr);
// mlp.hip
struct Mlp;
void mlp_init(Mlp& m, its_ctx ctx, uint64_t seed);
void mlp_forward(Mlp& m, const uint16_t* d_x, int64_t n, float* d_y, hipStream_t st);
void mlp_grad(Mlp& m, const uint16_t* d_x, const float* d_t, int64_t n, hipStream_t st);
void mlp_adam(Mlp& m, float lr, hipStream_t st);
Mlp& mlp_of(its_mlp net);
its_ctx mlp_ctx(its_mlp net);
its_mlp mlp_new();
void mlp_delete(its_mlp net);
void mlp_set(its_mlp net, const float* h, hipStream_t st);
void mlp_get(its_mlp net, float* h, hipStream_t st);
void mlp_read_grad(its_mlp net, float* grad, double* loss, int64_t n, hipStream_t st);
void scale_to_f64(int64_t n, const float* y, const float* s, double* w, hipStream_t st);
}  // namespace its
