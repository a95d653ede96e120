// extern "C" boundary of the ITS gfx950 library (declared in include/its_c.h).
// Every entry point converts exceptions into a status code and records the
// message for its_last_error(); nothing escapes as an exception or exit().
#include <cfloat>
#include <cmath>
#include <algorithm>
#include <atomic>
#include <vector>
#include <cstring>
#include <new>
#include <string>

#include "internal.hpp"

namespace its {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

static int64_t g_tune[ITS_TUNE_COUNT] = {32768, 64, 2048, 1, 1, 65536, 1};
int64_t tune_get(int key) { return g_tune[key]; }
}  // namespace its

using namespace its;

#define ITS_API_BEGIN try {
#define ITS_API_END                                      \
    return ITS_OK;                                       \
    }                                                    \
    catch (const its::Error& e) {                        \
        set_last_error(e.what());                        \
        return e.code;                                   \
    }                                                    \
    catch (const std::bad_alloc&) {                      \
        set_last_error("host allocation failed");        \
        return ITS_E_NOMEM;                              \
    }                                                    \
    catch (const std::exception& e) {                    \
        set_last_error(e.what());                        \
        return ITS_E_INVALID;                            \
    }

namespace {
// live contexts per device: the last one to go trims the device's pool
std::atomic<int> g_live_ctx[64];

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        ITS_HIP(hipGetDevice(&prev));
        if (prev != dev) ITS_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

template <class T>
DBuf<T> upload(const T* h, size_t n, hipStream_t st) {
    DBuf<T> d(n ? n : 1);
    if (n) ITS_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
    return d;
}

void check_svo(its_svo s) { ITS_REQUIRE(s != nullptr && s->ctx != nullptr, ITS_E_INVALID, "null octree handle"); }
// entry points that need nodes, voxels or the solver
void check_full(its_svo s) {
    check_svo(s);
    ITS_REQUIRE(!s->vertices_only, ITS_E_STATE, "octree built from vertices has no nodes, voxels or solver");
}
// caller order <-> internal (level, lattice) order of per-vertex vectors
void to_internal(its_svo s, const double* in, std::vector<double>& out) {
    out.resize(s->num_verts);
    for (int64_t i = 0; i < s->num_verts; ++i) out[s->vperm[i]] = in[i];
}
}  // namespace

extern "C" {

const char* its_last_error(void) { return g_last_error.c_str(); }
const char* its_version(void) { return "its_amd 0.1 (gfx950)"; }

int its_device_count(int* count) {
    ITS_API_BEGIN
    ITS_REQUIRE(count, ITS_E_INVALID, "null pointer");
    ITS_HIP(hipGetDeviceCount(count));
    ITS_API_END
}

int its_ctx_create(int device, its_ctx* out) {
    ITS_API_BEGIN
    ITS_REQUIRE(out, ITS_E_INVALID, "null output pointer");
    int n = 0;
    ITS_HIP(hipGetDeviceCount(&n));
    ITS_REQUIRE(device >= 0 && device < n && device < 64, ITS_E_INVALID, "no such device");
    DeviceGuard g(device);
    // device address of the allocation-free evaluation's scratch array
    eval_scratch_init(device);
    auto* c = new its_ctx_s();
    c->device = device;
    ITS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
    ITS_HIP(hipEventCreate(&c->ev0));
    ITS_HIP(hipEventCreate(&c->ev1));
    for (hipEvent_t& e : c->mc_ev) ITS_HIP(hipEventCreate(&e));
    ITS_HIP(hipGetDeviceProperties(&c->prop, device));
    // stream-ordered allocations come from the library's own pool on this
    // device, which keeps freed blocks; the default pool is left as it is
    (void)device_pool();
    g_live_ctx[device].fetch_add(1);
    *out = c;
    ITS_API_END
}

int its_ctx_destroy(its_ctx ctx) {
    ITS_API_BEGIN
    if (!ctx) return ITS_OK;
    DeviceGuard g(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream) block_cache_trim(ctx->stream);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->mc_ev)
        if (e) (void)hipEventDestroy(e);
    const int dev = ctx->device;
    delete ctx;
    // the last context of a device gives the pool's unused blocks back
    if (g_live_ctx[dev].fetch_sub(1) == 1) device_pool_trim();
    ITS_API_END
}

int its_ctx_set_stream(its_ctx ctx, void* hip_stream) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx, ITS_E_INVALID, "null context");
    DeviceGuard g(ctx->device);
    ITS_HIP(hipStreamSynchronize(ctx->stream));
    block_cache_trim(ctx->stream);   // blocks kept for the old stream
    if (ctx->own_stream) ITS_HIP(hipStreamDestroy(ctx->stream));
    if (hip_stream) {
        ctx->stream = (hipStream_t)hip_stream;
        ctx->own_stream = false;
    } else {
        ITS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->own_stream = true;
    }
    ITS_API_END
}

int its_ctx_synchronize(its_ctx ctx) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx, ITS_E_INVALID, "null context");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    ITS_HIP(hipStreamSynchronize(ctx->stream));
    ITS_API_END
}

// ---- BaseModel::model2UnitCube (BaseModel.cpp:54-92)
int its_model_to_unit_cube(double* V, int64_t nV, double scale_factor) {
    ITS_API_BEGIN
    ITS_REQUIRE(V && nV > 0, ITS_E_INVALID, "empty vertex array");
    double bmin[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, bmax[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = 0; i < nV; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = V[3 * i + a];
            if (v < bmin[a]) bmin[a] = v;
            if (v > bmax[a]) bmax[a] = v;
        }
    double scale = bmax[0] - bmin[0];
    for (int d = 1; d < 3; ++d) scale = std::max(scale, bmax[d] - bmin[d]);
    scale *= scale_factor;
    double center[3];
    for (int a = 0; a < 3; ++a) center[a] = 0.5 * bmax[a] + 0.5 * bmin[a];
    for (int a = 0; a < 3; ++a) center[a] -= scale / 2;
    const double inv = 1. / scale;
    for (int64_t i = 0; i < nV; ++i)
        for (int a = 0; a < 3; ++a) V[3 * i + a] = (V[3 * i + a] + (-center[a])) * inv;
    ITS_API_END
}

// ---- BaseModel::setUniformBoundingBox (BaseModel.cpp:24-51), float casts kept
int its_uniform_bbox(const double* V, int64_t nV, double origin[3], double end[3], double width[3]) {
    ITS_API_BEGIN
    ITS_REQUIRE(V && nV > 0 && origin && end && width, ITS_E_INVALID, "bad arguments");
    double minV[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, maxV[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = 0; i < nV; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = V[3 * i + a];
            if (v < minV[a]) minV[a] = v;
            if (v > maxV[a]) maxV[a] = v;
        }
    double lengths[3];
    for (int a = 0; a < 3; ++a) {
        origin[a] = minV[a];
        end[a] = maxV[a];
        lengths[a] = maxV[a] - minV[a];
    }
    float max_length = fmaxf((float)lengths[0], fmaxf((float)lengths[1], (float)lengths[2]));
    for (int i = 0; i < 3; ++i) {
        if ((double)max_length == lengths[i]) continue;
        float delta = (float)((double)max_length - lengths[i]);
        origin[i] = minV[i] - (double)(delta / 2.0f);
        end[i] = maxV[i] + (double)(delta / 2.0f);
    }
    double eps[3];
    for (int a = 0; a < 3; ++a) eps[a] = (end[a] - origin[a]) / 11;
    for (int a = 0; a < 3; ++a) {
        origin[a] -= eps[a];
        end[a] += eps[a];
        width[a] = end[a] - origin[a];
    }
    ITS_API_END
}

// ---- octree
static int svo_build_common(its_ctx ctx, const double* tris, bool device_ptr, int64_t T, const double origin[3],
                            const double width[3], int R, its_svo* out) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx && out && origin && width, ITS_E_INVALID, "bad arguments");
    ITS_REQUIRE(T >= 0 && (T == 0 || tris), ITS_E_INVALID, "bad triangle array");
    ITS_REQUIRE(R >= 2 && R <= 1024 && (R & (R - 1)) == 0, ITS_E_RANGE, "R must be a power of two in [2, 1024]");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    std::unique_ptr<its_svo_s> s(new its_svo_s());
    s->ctx = ctx;
    s->R = R;
    for (int a = 0; a < 3; ++a) {
        s->origin[a] = origin[a];
        s->width[a] = width[a];
    }
    if (device_ptr || T == 0) {
        svo_build(s.get(), tris, T);
    } else {
        DBuf<double> d = upload(tris, 9 * (size_t)T, ctx->stream);
        svo_build(s.get(), d.p, T);
    }
    *out = s.release();
    ITS_API_END
}

int its_svo_build(its_ctx ctx, const double* tris, int64_t T, const double origin[3], const double width[3], int R,
                  its_svo* out) {
    return svo_build_common(ctx, tris, false, T, origin, width, R, out);
}
int its_svo_build_device(its_ctx ctx, const double* d_tris, int64_t T, const double origin[3], const double width[3],
                         int R, its_svo* out) {
    return svo_build_common(ctx, d_tris, true, T, origin, width, R, out);
}

int its_svo_from_vertices(its_ctx ctx, const double* xyz, const uint32_t* owner, int64_t n_verts,
                          const double* node_width, int64_t n_nodes, its_svo* out) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx && out && xyz && owner && node_width && n_verts > 0 && n_nodes > 0, ITS_E_INVALID,
                "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    std::unique_ptr<its_svo_s> s(new its_svo_s());
    s->ctx = ctx;
    svo_from_vertices(s.get(), xyz, owner, n_verts, node_width, n_nodes);
    *out = s.release();
    ITS_API_END
}

int its_svo_destroy(its_svo svo) {
    ITS_API_BEGIN
    if (!svo) return ITS_OK;
    DeviceGuard g(svo->ctx->device);
    AllocStream as(svo->ctx->stream);
    (void)hipStreamSynchronize(svo->ctx->stream);
    delete svo;
    ITS_API_END
}

int its_svo_get_info(its_svo s, its_svo_info* info) {
    ITS_API_BEGIN
    check_svo(s);
    ITS_REQUIRE(info, ITS_E_INVALID, "null info");
    std::memset(info, 0, sizeof *info);
    info->tree_depth = s->depth;
    info->R = s->R;
    info->num_voxels = s->num_voxels;
    info->num_nodes = s->num_nodes;
    info->num_verts = s->num_verts;
    for (int d = 0; d < s->depth; ++d) info->depth_num_nodes[d] = s->lv[d].n_nodes;
    for (int d = 0; d <= s->depth; ++d) {
        info->esum_nodes[d] = s->esum_nodes[d];
        info->esum_verts[d] = s->esum_verts[d];
    }
    for (int a = 0; a < 3; ++a) info->grid_origin[a] = s->origin[a];
    info->voxel_width = s->voxel_width;
    ITS_API_END
}

int its_svo_export_nodes(its_svo s, uint32_t* morton, uint8_t* is_leaf, double* origin, double* width,
                         uint32_t* parent, uint32_t* childs, uint32_t* neighbors) {
    ITS_API_BEGIN
    check_full(s);
    DeviceGuard g(s->ctx->device);
    AllocStream as(s->ctx->stream);
    hipStream_t st = s->ctx->stream;
    const int64_t n = s->num_nodes;
    if (n == 0) return ITS_OK;
    auto cp = [&](void* h, const void* d, size_t bytes) {
        if (h) ITS_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st));
    };
    cp(morton, s->morton.p, n * sizeof(uint32_t));
    cp(is_leaf, s->is_leaf.p, n);
    cp(origin, s->node_origin.p, 3 * n * sizeof(double));
    cp(width, s->node_width.p, n * sizeof(double));
    cp(parent, s->parent.p, n * sizeof(uint32_t));
    cp(neighbors, s->neighbors.p, 27 * n * sizeof(uint32_t));
    if (childs) {
        std::vector<uint32_t> cb;
        to_host(cb, s->childbase.p, n, st);
        for (int64_t i = 0; i < n; ++i)
            for (int c = 0; c < 8; ++c) childs[8 * i + c] = (cb[i] == ITS_NONE) ? ITS_NONE : cb[i] + (uint32_t)c;
    }
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

int its_svo_export_vert_lattice(its_svo s, uint16_t* ijk, uint8_t* level) {
    ITS_API_BEGIN
    check_svo(s);
    DeviceGuard g(s->ctx->device);
    AllocStream as(s->ctx->stream);
    hipStream_t st = s->ctx->stream;
    const int64_t n = s->num_verts;
    if (n == 0) return ITS_OK;
    if (ijk) {
        std::vector<uint16_t> x, y, z;
        to_host(x, s->vx.p, n, st);
        to_host(y, s->vy.p, n, st);
        to_host(z, s->vz.p, n, st);
        ITS_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < n; ++i) {
            ijk[3 * i] = x[i];
            ijk[3 * i + 1] = y[i];
            ijk[3 * i + 2] = z[i];
        }
    }
    if (level) {
        ITS_HIP(hipMemcpyAsync(level, s->vlevel.p, n, hipMemcpyDeviceToHost, st));
        ITS_HIP(hipStreamSynchronize(st));
    }
    ITS_API_END
}

int its_svo_export_fine_edges(its_svo s, double* ends, uint32_t* owner, int64_t capacity, int64_t* count) {
    ITS_API_BEGIN
    check_full(s);
    DeviceGuard g(s->ctx->device);
    AllocStream as(s->ctx->stream);
    hipStream_t st = s->ctx->stream;
    DBuf<double> E;
    DBuf<uint32_t> O;
    const int64_t n = fine_edges(s, E, O, st);
    if (count) *count = n;
    if ((ends || owner) && n > 0) {
        ITS_REQUIRE(n <= capacity, ITS_E_RANGE, "edge buffer too small");
        if (ends) ITS_HIP(hipMemcpyAsync(ends, E.p, 6 * n * sizeof(double), hipMemcpyDeviceToHost, st));
        if (owner) ITS_HIP(hipMemcpyAsync(owner, O.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        ITS_HIP(hipStreamSynchronize(st));
    }
    ITS_API_END
}

int its_svo_export_verts(its_svo s, double* xyz, uint32_t* owner) {
    ITS_API_BEGIN
    check_svo(s);
    DeviceGuard g(s->ctx->device);
    AllocStream as(s->ctx->stream);
    hipStream_t st = s->ctx->stream;
    const int64_t n = s->num_verts;
    if (n == 0) return ITS_OK;
    if (xyz) ITS_HIP(hipMemcpyAsync(xyz, s->vxyz.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (owner) ITS_HIP(hipMemcpyAsync(owner, s->vowner.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

int its_svo_export_voxels(its_svo s, uint32_t* morton) {
    ITS_API_BEGIN
    check_full(s);
    DeviceGuard g(s->ctx->device);
    AllocStream as(s->ctx->stream);
    if (s->num_voxels && morton) {
        ITS_HIP(hipMemcpyAsync(morton, s->lv[0].occ.p, s->num_voxels * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               s->ctx->stream));
        ITS_HIP(hipStreamSynchronize(s->ctx->stream));
    }
    ITS_API_END
}

// ---- right-hand side
int its_sdf_at_vertices(its_ctx ctx, its_svo s, const double* V, int64_t nV, const int32_t* F, int64_t nF, double* b) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx, ITS_E_INVALID, "octree belongs to another context");
    ITS_REQUIRE(V && F && nV > 0 && nF > 0, ITS_E_INVALID, "empty mesh");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    for (int64_t i = 0; i < 3 * nF; ++i)
        ITS_REQUIRE(F[i] >= 0 && F[i] < nV, ITS_E_INVALID, "face index out of range");
    DBuf<double> dV = upload(V, 3 * (size_t)nV, st);
    DBuf<int32_t> dF = upload(F, 3 * (size_t)nF, st);
    const int64_t n = s->num_verts;
    s->rhs.ensure(std::max<int64_t>(n, 1));
    s->rhs_tri.ensure(std::max<int64_t>(n, 1));
    signed_distance(st, dV.p, nV, dF.p, nF, s->vxyz.p, n, s->rhs.p, s->rhs_tri.p);
    s->rhs_nrm.ensure(std::max<int64_t>(n, 1));
    nearest_normals(st, dV.p, dF.p, s->rhs_tri.p, n, s->rhs_nrm.p);
    ITS_HIP(hipStreamSynchronize(st));
    s->has_rhs = true;
    if (b && n) {
        ITS_HIP(hipMemcpyAsync(b, s->rhs.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
        ITS_HIP(hipStreamSynchronize(st));
    }
    ITS_API_END
}

int its_signed_distance(its_ctx ctx, const double* V, int64_t nV, const int32_t* F, int64_t nF, const double* pts,
                        int64_t N, double* out) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx && V && F && pts && out && nV > 0 && nF > 0 && N >= 0, ITS_E_INVALID, "bad arguments");
    for (int64_t i = 0; i < 3 * nF; ++i)
        ITS_REQUIRE(F[i] >= 0 && F[i] < nV, ITS_E_INVALID, "face index out of range");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    if (N == 0) return ITS_OK;
    DBuf<double> dV = upload(V, 3 * (size_t)nV, st);
    DBuf<int32_t> dF = upload(F, 3 * (size_t)nF, st);
    DBuf<double> dP = upload(pts, 3 * (size_t)N, st);
    DBuf<double> dO(N);
    signed_distance(st, dV.p, nV, dF.p, nF, dP.p, N, dO.p, nullptr);
    ITS_HIP(hipMemcpyAsync(out, dO.p, N * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

// ---- fit
int its_lscg_solve(its_ctx ctx, its_svo s, const double* b, double* lambda_inout, double tol, int64_t max_iter,
                   int warm, its_lscg_stats* stats) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx, ITS_E_INVALID, "octree belongs to another context");
    ITS_REQUIRE(lambda_inout, ITS_E_INVALID, "null lambda");
    ITS_REQUIRE(tol >= 0, ITS_E_INVALID, "negative tolerance");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    const int64_t n = s->num_verts;
    ITS_REQUIRE(n > 0, ITS_E_STATE, "octree has no vertices");
    DBuf<double> db;
    const double* d_b;
    if (b) {
        db = upload(b, n, st);
        d_b = db.p;
    } else {
        ITS_REQUIRE(s->has_rhs, ITS_E_STATE, "no right-hand side: pass b or call its_sdf_at_vertices");
        d_b = s->rhs.p;
    }
    lambda_changing(s);
    s->lambda64.ensure(n);
    if (warm) ITS_HIP(hipMemcpyAsync(s->lambda64.p, lambda_inout, n * sizeof(double), hipMemcpyHostToDevice, st));
    lscg_solve(s, d_b, s->lambda64.p, tol, max_iter, warm != 0, stats);
    ITS_HIP(hipMemcpyAsync(lambda_inout, s->lambda64.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    s->has_lambda = true;
    field_build(s);
    ITS_API_END
}

int its_collocation_mul(its_ctx ctx, its_svo s, const double* x, double* y, int transpose) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx && x && y, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    const int64_t n = s->num_verts;
    DBuf<double> dx = upload(x, n, st), dy(n);
    collocation_mul(s, dx.p, dy.p, transpose != 0);
    ITS_HIP(hipMemcpyAsync(y, dy.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

int its_svo_set_lambda(its_ctx ctx, its_svo s, const double* lambda) {
    ITS_API_BEGIN
    check_svo(s);
    ITS_REQUIRE(ctx == s->ctx && lambda, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    const int64_t n = s->num_verts;
    lambda_changing(s);
    s->lambda64.ensure(std::max<int64_t>(n, 1));
    std::vector<double> perm;
    if (!s->vperm.empty()) {
        to_internal(s, lambda, perm);
        lambda = perm.data();
    }
    if (n) ITS_HIP(hipMemcpyAsync(s->lambda64.p, lambda, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    s->has_lambda = true;
    field_build(s);   // synchronises the stream before perm goes out of scope
    ITS_API_END
}

int its_svo_get_lambda(its_svo s, double* lambda) {
    ITS_API_BEGIN
    check_svo(s);
    ITS_REQUIRE(lambda, ITS_E_INVALID, "null lambda");
    ITS_REQUIRE(s->has_lambda, ITS_E_STATE, "no coefficients");
    DeviceGuard g(s->ctx->device);
    AllocStream as(s->ctx->stream);
    if (!s->vperm.empty()) {
        std::vector<double> tmp(s->num_verts);
        ITS_HIP(hipMemcpyAsync(tmp.data(), s->lambda64.p, s->num_verts * sizeof(double), hipMemcpyDeviceToHost,
                               s->ctx->stream));
        ITS_HIP(hipStreamSynchronize(s->ctx->stream));
        for (int64_t i = 0; i < s->num_verts; ++i) lambda[i] = tmp[s->vperm[i]];
        return ITS_OK;
    }
    ITS_HIP(hipMemcpyAsync(lambda, s->lambda64.p, s->num_verts * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
    ITS_HIP(hipStreamSynchronize(s->ctx->stream));
    ITS_API_END
}

// ---- evaluation
static void check_mode(int mode) {
    ITS_REQUIRE(mode == ITS_EVAL_SUPPORT || mode == ITS_EVAL_DENSE || mode == ITS_EVAL_RAW, ITS_E_INVALID, "bad mode");
}

int its_bspline_eval(its_ctx ctx, its_svo s, const float* pts, int64_t N, const double box_end[3], int mode,
                     float out_val, float* f) {
    ITS_API_BEGIN
    check_svo(s);
    check_mode(mode);
    ITS_REQUIRE(ctx == s->ctx && (N == 0 || (pts && f)) && box_end, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    if (N == 0) return ITS_OK;
    // the caller's interleaved points go to the device as they are; the
    // kernel reads them interleaved (no reordering pass on the host)
    DBuf<float> d = upload(pts, 3 * (size_t)N, st), out(N);
    eval_points_aos32(s, d.p, N, box_end, mode, out_val, out.p, st);
    ITS_HIP(hipMemcpyAsync(f, out.p, N * sizeof(float), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

int its_bspline_eval_f64pts(its_ctx ctx, its_svo s, const double* pts, int64_t N, const double box_end[3], int mode,
                            float out_val, float* f) {
    ITS_API_BEGIN
    check_svo(s);
    check_mode(mode);
    ITS_REQUIRE(ctx == s->ctx && (N == 0 || (pts && f)) && box_end, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    if (N == 0) return ITS_OK;
    DBuf<double> d = upload(pts, 3 * (size_t)N, st);
    DBuf<float> out(N);
    eval_points_aos64(s, d.p, N, box_end, mode, out_val, out.p, st);
    ITS_HIP(hipMemcpyAsync(f, out.p, N * sizeof(float), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

int its_bspline_eval_device(its_ctx ctx, its_svo s, const float* d_x, const float* d_y, const float* d_z, int64_t N,
                            const double box_end[3], int mode, float out_val, float* d_f, void* stream) {
    ITS_API_BEGIN
    check_svo(s);
    check_mode(mode);
    ITS_REQUIRE(ctx == s->ctx && box_end && (N == 0 || (d_x && d_y && d_z && d_f)), ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    eval_points(s, d_x, d_y, d_z, N, box_end, mode, out_val, d_f, stream ? (hipStream_t)stream : ctx->stream);
    ITS_API_END
}

int its_classify(const float* f, int64_t N, double inner_iso, double outer_iso, int8_t* cls) {
    ITS_API_BEGIN
    ITS_REQUIRE(N == 0 || (f && cls), ITS_E_INVALID, "bad arguments");
    for (int64_t i = 0; i < N; ++i) {
        double v = f[i];
        cls[i] = (v >= outer_iso) ? 1 : (v <= inner_iso ? -1 : 0);
    }
    ITS_API_END
}

// ---- marching cubes
int its_mc_extract(its_ctx ctx, its_svo s, const double* lambda, const double origin[3], const double width[3],
                   const int res[3], const double* iso, int n_iso, float* tri_xyz, int64_t capacity,
                   int64_t* num_verts, its_mc_stats* stats) {
    ITS_API_BEGIN
    check_svo(s);
    ITS_REQUIRE(ctx == s->ctx && origin && width && res && iso, ITS_E_INVALID, "bad arguments");
    ITS_REQUIRE(n_iso >= 1 && n_iso <= ITS_MC_MAX_ISO, ITS_E_INVALID, "n_iso must be in [1, ITS_MC_MAX_ISO]");
    ITS_REQUIRE(capacity >= 0, ITS_E_INVALID, "negative capacity");
    if (lambda) {
        const int rc = its_svo_set_lambda(ctx, s, lambda);
        if (rc != ITS_OK) return rc;
    }
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    ITS_REQUIRE(s->has_lambda, ITS_E_STATE, "no coefficients");
    its_mc_stats total{};
    int64_t at = 0;
    for (int k = 0; k < n_iso; ++k) {
        its_mc_stats S{};
        // the lattice of the first value serves the others (same grid, same coefficients)
        mc_extract(s, origin, width, res, iso[k], tri_xyz ? tri_xyz + 3 * at : nullptr,
                   tri_xyz ? capacity - at : 0, &S);
        if (num_verts) num_verts[k] = S.num_verts;
        at += S.num_verts;
        total.num_verts += S.num_verts;
        total.num_valid_voxels += S.num_valid_voxels;
        total.seconds_eval += S.seconds_eval;
        total.seconds_extract += S.seconds_extract;
        total.seconds_eval_device += S.seconds_eval_device;
        total.seconds_extract_device += S.seconds_extract_device;
    }
    if (stats) *stats = total;
    ITS_API_END
}

int its_mc_lattice(its_ctx ctx, its_svo s, const double origin[3], const double width[3], const int res[3],
                   float* values) {
    ITS_API_BEGIN
    check_svo(s);
    ITS_REQUIRE(ctx == s->ctx && origin && width && res && values, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    mc_lattice(s, origin, width, res, values);
    ITS_API_END
}

int its_lscg_structure_get(its_ctx ctx, its_svo s, its_lscg_structure* out) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx && out, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    lscg_structure(s, out);
    ITS_API_END
}

int its_lscg_diag(its_ctx ctx, its_svo s, double* invd) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx && invd, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    const int64_t n = s->num_verts;
    if (n == 0) return ITS_OK;
    DBuf<double> d(n);
    lscg_diag(s, d.p, st);
    ITS_HIP(hipMemcpyAsync(invd, d.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

int its_lattice_f64(its_ctx ctx, its_svo s, double* values) {
    ITS_API_BEGIN
    check_svo(s);
    ITS_REQUIRE(ctx == s->ctx && values, ITS_E_INVALID, "bad arguments");
    ITS_REQUIRE(s->has_lambda && s->depth > 0, ITS_E_STATE, "no coefficients");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    DBuf<double> lat;
    dense_cascade64(s, 0, lat);
    const int64_t D = s->R + 1;
    ITS_HIP(hipMemcpyAsync(values, lat.p, D * D * D * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}

// ---- shell iso band
extern "C++" {   // a C++ helper inside the extern "C" block: C++ linkage
namespace {
struct ShellInput {
    DBuf<double> V, tris;
    DBuf<uint64_t> edges;
    int64_t ne = 0;
};
ShellInput shell_upload(const double* V, int64_t nV, const int32_t* F, int64_t nF, hipStream_t st) {
    ITS_REQUIRE(V && F && nV > 0 && nF > 0, ITS_E_INVALID, "empty mesh");
    for (int64_t i = 0; i < 3 * nF; ++i)
        ITS_REQUIRE(F[i] >= 0 && F[i] < nV, ITS_E_INVALID, "face index out of range");
    ShellInput in;
    std::vector<uint64_t> e = mesh_edges(F, nF);
    std::vector<double> t(9 * (size_t)nF);
    for (int64_t f = 0; f < nF; ++f)
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) t[9 * f + 3 * c + a] = V[3 * (int64_t)F[3 * f + c] + a];
    in.V = upload(V, 3 * (size_t)nV, st);
    in.tris = upload(t.data(), t.size(), st);
    in.edges = upload(e.data(), e.size(), st);
    in.ne = (int64_t)e.size();
    return in;
}
}  // namespace
}  // extern "C++"

int its_shell_points(its_ctx ctx, its_svo s, const double* V, int64_t nV, const int32_t* F, int64_t nF, double* pts,
                     int64_t capacity, int64_t* n_edge, int64_t* n_face) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx, ITS_E_INVALID, "octree belongs to another context");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    ShellInput in = shell_upload(V, nV, F, nF, st);
    DBuf<double> P;
    int64_t ne = 0, nf = 0;
    shell_points(s, in.V.p, in.edges.p, in.ne, in.tris.p, nF, P, &ne, &nf, st);
    if (n_edge) *n_edge = ne;
    if (n_face) *n_face = nf;
    if (pts && ne + nf > 0) {
        ITS_REQUIRE(ne + nf <= capacity, ITS_E_RANGE, "point buffer too small");
        ITS_HIP(hipMemcpyAsync(pts, P.p, 3 * (ne + nf) * sizeof(double), hipMemcpyDeviceToHost, st));
        ITS_HIP(hipStreamSynchronize(st));
    }
    ITS_API_END
}

int its_shell_iso(its_ctx ctx, its_svo s, const double* V, int64_t nV, const int32_t* F, int64_t nF, double iso[2],
                  int64_t* n_points) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx && iso, ITS_E_INVALID, "bad arguments");
    ITS_REQUIRE(s->field_ready, ITS_E_STATE, "no coefficient field: run a fit or set lambda first");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    ShellInput in = shell_upload(V, nV, F, nF, st);
    DBuf<double> P;
    int64_t ne = 0, nf = 0;
    shell_points(s, in.V.p, in.edges.p, in.ne, in.tris.p, nF, P, &ne, &nf, st);
    const int64_t np = ne + nf, nall = nV + np;
    DBuf<float> f(nall);
    const double dummy[3] = {0, 0, 0};
    eval_points_aos64(s, in.V.p, nV, dummy, ITS_EVAL_RAW, 0.f, f.p, st);
    if (np > 0) eval_points_aos64(s, P.p, np, dummy, ITS_EVAL_RAW, 0.f, f.p + nV, st);
    std::vector<float> h(nall);
    ITS_HIP(hipMemcpyAsync(h.data(), f.p, nall * sizeof(float), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    const auto mm = std::minmax_element(h.begin(), h.end());
    iso[0] = *mm.first;
    iso[1] = *mm.second;
    if (n_points) *n_points = np;
    ITS_API_END
}

// ---- batches
int its_batch_create(its_ctx ctx, its_batch* out) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx && out, ITS_E_INVALID, "bad arguments");
    *out = batch_new(ctx);
    ITS_API_END
}
int its_batch_destroy(its_batch b) {
    ITS_API_BEGIN
    if (b) {
        DeviceGuard g(batch_ctx(b)->device);
    AllocStream as(batch_ctx(b)->stream);
        batch_delete(b);
    }
    ITS_API_END
}
int its_batch_add(its_batch b, its_svo s) {
    ITS_API_BEGIN
    ITS_REQUIRE(b, ITS_E_INVALID, "null batch");
    check_full(s);
    batch_add(b, s);
    ITS_API_END
}
int its_batch_fit(its_batch b, double tol, int64_t max_iter, int warm, its_lscg_stats* stats, double* seconds) {
    ITS_API_BEGIN
    ITS_REQUIRE(b, ITS_E_INVALID, "null batch");
    ITS_REQUIRE(tol > 0, ITS_E_INVALID, "tolerance must be positive");
    DeviceGuard g(batch_ctx(b)->device);
    AllocStream as(batch_ctx(b)->stream);
    batch_fit(b, tol, max_iter, warm != 0, stats, seconds);
    ITS_API_END
}
int its_batch_fields(its_batch b) {
    ITS_API_BEGIN
    ITS_REQUIRE(b, ITS_E_INVALID, "null batch");
    DeviceGuard g(batch_ctx(b)->device);
    AllocStream as(batch_ctx(b)->stream);
    batch_fields(b);
    ITS_API_END
}
// argument checks shared by both batched evaluation entry points
static void check_batch_eval(its_batch b, const float* d_x, const float* d_y, const float* d_z,
                             const int64_t* offsets, float* d_f) {
    ITS_REQUIRE(b && offsets, ITS_E_INVALID, "bad arguments");
    const int64_t count = batch_count(b);
    ITS_REQUIRE(count > 0, ITS_E_STATE, "empty batch");
    ITS_REQUIRE(offsets[0] == 0, ITS_E_INVALID, "offsets must start at 0");
    for (int64_t m = 0; m < count; ++m) {
        ITS_REQUIRE(offsets[m + 1] >= offsets[m], ITS_E_INVALID, "offsets must not decrease");
        ITS_REQUIRE(batch_svo(b, m)->field_ready, ITS_E_STATE, "a mesh has no evaluation field (its_batch_fields)");
    }
    ITS_REQUIRE(offsets[count] == 0 || (d_x && d_y && d_z && d_f), ITS_E_INVALID, "null query arrays");
}

int its_batch_eval_device(its_batch b, const float* d_x, const float* d_y, const float* d_z, const int64_t* offsets,
                          float out_val, float* d_f, double* seconds) {
    ITS_API_BEGIN
    check_batch_eval(b, d_x, d_y, d_z, offsets, d_f);
    const int64_t count = batch_count(b);
    its_ctx ctx = batch_ctx(b);
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    std::vector<its_svo> v(count);
    for (int64_t m = 0; m < count; ++m) v[m] = batch_svo(b, m);
    ITS_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    eval_batch(v.data(), count, offsets, d_x, d_y, d_z, out_val, d_f, ctx->stream);
    ITS_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ITS_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    ITS_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (seconds) *seconds = ms * 1e-3;
    ITS_API_END
}

int its_batch_eval_device_noalloc(its_batch b, const float* d_x, const float* d_y, const float* d_z,
                                  const int64_t* offsets, float out_val, float* d_f, double* seconds) {
    ITS_API_BEGIN
    check_batch_eval(b, d_x, d_y, d_z, offsets, d_f);
    const int64_t count = batch_count(b);
    ITS_REQUIRE(count <= ITS_NOALLOC_MAX_MESHES, ITS_E_RANGE,
                "batch holds more meshes than the allocation-free evaluation takes (ITS_NOALLOC_MAX_MESHES)");
    its_ctx ctx = batch_ctx(b);
    DeviceGuard g(ctx->device);
    its_svo v[ITS_NOALLOC_MAX_MESHES];
    for (int64_t m = 0; m < count; ++m) v[m] = batch_svo(b, m);
    ITS_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    eval_batch_noalloc(v, count, offsets, d_x, d_y, d_z, out_val, d_f, ctx->stream);
    ITS_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ITS_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    ITS_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (seconds) *seconds = ms * 1e-3;
    ITS_API_END
}

// ---- learned warm start
static void check_mlp(its_mlp net) { ITS_REQUIRE(net, ITS_E_INVALID, "null network"); }

int its_mlp_create(its_ctx ctx, uint64_t seed, its_mlp* out) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx && out, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    its_mlp net = mlp_new();
    try {
        mlp_init(mlp_of(net), ctx, seed);
    } catch (...) {
        mlp_delete(net);
        throw;
    }
    *out = net;
    ITS_API_END
}
int its_mlp_destroy(its_mlp net) {
    ITS_API_BEGIN
    if (net) {
        DeviceGuard g(mlp_ctx(net)->device);
    AllocStream as(mlp_ctx(net)->stream);
        mlp_delete(net);
    }
    ITS_API_END
}
int its_mlp_set_params(its_mlp net, const float* params) {
    ITS_API_BEGIN
    check_mlp(net);
    ITS_REQUIRE(params, ITS_E_INVALID, "null parameters");
    its_ctx ctx = mlp_ctx(net);
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    mlp_set(net, params, ctx->stream);
    ITS_API_END
}
int its_mlp_get_params(its_mlp net, float* params) {
    ITS_API_BEGIN
    check_mlp(net);
    ITS_REQUIRE(params, ITS_E_INVALID, "null parameters");
    its_ctx ctx = mlp_ctx(net);
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    mlp_get(net, params, ctx->stream);
    ITS_API_END
}
int its_mlp_forward_device(its_mlp net, const uint16_t* d_x, int64_t rows, float* d_y, void* stream) {
    ITS_API_BEGIN
    check_mlp(net);
    ITS_REQUIRE(rows >= 0 && (rows == 0 || (d_x && d_y)), ITS_E_INVALID, "bad arguments");
    its_ctx ctx = mlp_ctx(net);
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    mlp_forward(mlp_of(net), d_x, rows, d_y, stream ? (hipStream_t)stream : ctx->stream);
    ITS_API_END
}
int its_mlp_grad_device(its_mlp net, const uint16_t* d_x, const float* d_t, int64_t rows, float* grad, double* loss) {
    ITS_API_BEGIN
    check_mlp(net);
    ITS_REQUIRE(rows > 0 && d_x && d_t, ITS_E_INVALID, "bad arguments");
    its_ctx ctx = mlp_ctx(net);
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    mlp_grad(mlp_of(net), d_x, d_t, rows, ctx->stream);
    mlp_read_grad(net, grad, loss, rows, ctx->stream);
    ITS_API_END
}
int its_mlp_train_step(its_mlp net, const uint16_t* d_x, const float* d_t, int64_t rows, float lr, double* loss) {
    ITS_API_BEGIN
    check_mlp(net);
    ITS_REQUIRE(rows > 0 && d_x && d_t && lr > 0, ITS_E_INVALID, "bad arguments");
    its_ctx ctx = mlp_ctx(net);
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    mlp_grad(mlp_of(net), d_x, d_t, rows, ctx->stream);
    mlp_adam(mlp_of(net), lr, ctx->stream);
    if (loss) mlp_read_grad(net, nullptr, loss, rows, ctx->stream);
    ITS_API_END
}
int its_warm_features(its_ctx ctx, its_svo s, uint16_t* d_x, float* d_scale, void* stream) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx && d_x && d_scale, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    warm_features(s, d_x, d_scale, stream ? (hipStream_t)stream : ctx->stream);
    ITS_API_END
}
int its_warm_dual(its_ctx ctx, its_svo s, const double* lambda, double* w) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx && lambda && w, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    const int64_t n = s->num_verts;
    if (n == 0) return ITS_OK;
    DBuf<double> dl = upload(lambda, n, st);
    DBuf<double> dw(n);
    warm_dual(s, dl.p, dw.p);
    ITS_HIP(hipMemcpyAsync(w, dw.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}
int its_warm_apply(its_ctx ctx, its_svo s, const double* w, double* lambda0) {
    ITS_API_BEGIN
    check_full(s);
    ITS_REQUIRE(ctx == s->ctx && w && lambda0, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    const int64_t n = s->num_verts;
    if (n == 0) return ITS_OK;
    DBuf<double> dw = upload(w, n, st);
    DBuf<double> dl(n);
    warm_apply(s, dw.p, dl.p, st);
    ITS_HIP(hipMemcpyAsync(lambda0, dl.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    ITS_API_END
}
int its_lscg_solve_mlp(its_ctx ctx, its_svo s, its_mlp net, const double* b, double* lambda, double tol,
                       int64_t max_iter, its_lscg_stats* stats, double* seconds_predict) {
    ITS_API_BEGIN
    check_full(s);
    check_mlp(net);
    ITS_REQUIRE(ctx == s->ctx && mlp_ctx(net) == ctx && lambda, ITS_E_INVALID, "bad arguments");
    ITS_REQUIRE(tol > 0, ITS_E_INVALID, "tolerance must be positive");
    ITS_REQUIRE(s->has_rhs, ITS_E_STATE, "the features need its_sdf_at_vertices");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    const int64_t n = s->num_verts;
    ITS_REQUIRE(n > 0, ITS_E_STATE, "empty octree");
    DBuf<double> db;
    const double* d_b = s->rhs.p;
    if (b) {
        db = upload(b, n, st);
        d_b = db.p;
    }
    ITS_HIP(hipEventRecord(ctx->ev0, st));
    lambda_changing(s);
    s->lambda64.ensure(n);
    DBuf<uint16_t> x(16 * n);
    DBuf<float> scale(n), y(n);
    DBuf<double> w(n);
    DBuf<int> ferr(1);
    // the features' structure check is read after the solve, so the
    // prediction runs without a host round trip
    warm_features(s, x.p, scale.p, st, ferr.p);
    mlp_forward(mlp_of(net), x.p, n, y.p, st);
    scale_to_f64(n, y.p, scale.p, w.p, st);
    warm_apply(s, w.p, s->lambda64.p, st);
    ITS_HIP(hipEventRecord(ctx->ev1, st));
    ITS_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    ITS_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (seconds_predict) *seconds_predict = ms * 1e-3;
    lscg_solve(s, d_b, s->lambda64.p, tol, max_iter, true, stats);
    ITS_REQUIRE(read1(ferr.p, st) == 0, ITS_E_STATE, "owner corner is not a vertex of its level");
    ITS_HIP(hipMemcpyAsync(lambda, s->lambda64.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    s->has_lambda = true;
    field_build(s);
    ITS_API_END
}

// ---- tunables
int its_tune_set(int key, int64_t value) {
    ITS_API_BEGIN
    ITS_REQUIRE(key >= 0 && key < ITS_TUNE_COUNT, ITS_E_INVALID, "unknown tunable");
    const bool is_switch = key == ITS_TUNE_EVAL_QUEUE || key == ITS_TUNE_SDF_RAY;
    ITS_REQUIRE(value >= (is_switch ? 0 : 1), ITS_E_RANGE, "tunable out of range");
    if (is_switch) ITS_REQUIRE(value <= 1, ITS_E_RANGE, "a switch is 0 or 1");
    if (key == ITS_TUNE_COARSE_VERTS) ITS_REQUIRE(value <= 12288, ITS_E_RANGE, "coarse vertex bound above 12288");
    if (key == ITS_TUNE_EVAL_BLOCKS_PER_CU) ITS_REQUIRE(value <= 64, ITS_E_RANGE, "at most 64 workgroups per CU");
    if (key == ITS_TUNE_EVAL_ILP) ITS_REQUIRE(value == 1 || value == 2 || value == 4, ITS_E_RANGE, "ILP must be 1, 2 or 4");
    g_tune[key] = value;
    ITS_API_END
}
int its_tune_get(int key, int64_t* value) {
    ITS_API_BEGIN
    ITS_REQUIRE(key >= 0 && key < ITS_TUNE_COUNT && value, ITS_E_INVALID, "bad arguments");
    *value = g_tune[key];
    ITS_API_END
}

// ---- achievable bandwidth: a plain streaming copy, one 16-byte element per
// thread over a grid that covers the buffer.  On MI355X (1 GiB) this gave
// 6.1 TB/s, against 4.0-4.9 TB/s for grid-stride loops with 4 or 8 loads in
// flight (nontemporal stores or not, 4-32 workgroups per CU), 5.3-5.7 TB/s
// for 1 KiB chunks per workgroup and 5.0 TB/s for hipMemcpyAsync.
namespace {
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ a, uint4* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}
}  // namespace

int its_hbm_bandwidth(its_ctx ctx, int64_t bytes, int reps, double* gbps) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx && gbps, ITS_E_INVALID, "bad arguments");
    ITS_REQUIRE(bytes >= (1 << 20) && bytes <= ((int64_t)16 << 30), ITS_E_RANGE, "bytes must be in [1 MiB, 16 GiB]");
    ITS_REQUIRE(reps >= 1 && reps <= 100, ITS_E_RANGE, "reps must be in [1, 100]");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    hipStream_t st = ctx->stream;
    const int64_t n = bytes / 16;
    DBuf<uint4> a(n), b(n);
    ITS_HIP(hipMemsetAsync(a.p, 1, n * sizeof(uint4), st));
    const unsigned grid = (unsigned)((n + 255) / 256);   // n <= 2^30: grid < 2^22
    // events of its own, so a timer running on the context is left alone
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    ITS_HIP(hipEventCreate(&ev.e[0]));
    ITS_HIP(hipEventCreate(&ev.e[1]));
    std::vector<float> ms(reps);
    for (int r = -2; r < reps; ++r) {
        ITS_HIP(hipEventRecord(ev.e[0], st));
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, st, (const uint4*)a.p, b.p, n);
        ITS_LAUNCH_CHECK();
        ITS_HIP(hipEventRecord(ev.e[1], st));
        ITS_HIP(hipEventSynchronize(ev.e[1]));
        if (r >= 0) ITS_HIP(hipEventElapsedTime(&ms[r], ev.e[0], ev.e[1]));
    }
    std::sort(ms.begin(), ms.end());
    const double t = 1e-3 * (double)ms[reps / 2];
    *gbps = t > 0 ? 2.0 * (double)(n * 16) / t * 1e-9 : 0.0;
    ITS_API_END
}

// ---- timing
int its_timer_start(its_ctx ctx) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx, ITS_E_INVALID, "null context");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    ITS_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    ITS_API_END
}
int its_timer_stop(its_ctx ctx, double* ms) {
    ITS_API_BEGIN
    ITS_REQUIRE(ctx && ms, ITS_E_INVALID, "bad arguments");
    DeviceGuard g(ctx->device);
    AllocStream as(ctx->stream);
    ITS_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ITS_HIP(hipEventSynchronize(ctx->ev1));
    float t = 0;
    ITS_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t;
    ITS_API_END
}

}  // extern "C"
