// its_host.hpp -- C++ host classes over the C ABI (include/its_c.h).
//
// Keeps the reference's host-side names and call shapes for the hot path so
// that main.cpp-style callers drop in:
//   SparseVoxelOctree::createOctree      (reference src/SVO.h:104)
//   ThinShells(filename, gx, gy, gz, is2UnitCube, scaleFactor)
//                                         (reference src/ThinShells.h:67-77)
//   ThinShells::creatShell               (reference src/ThinShells.cpp:757-815,
//                                         cpIntersectionPoints 38-479)
//   ThinShells::getShellIsoVal           (reference src/ThinShells.h:126)
//   ThinShells::multiPointQuery          (reference src/ThinShells.cpp:1183-1420)
//   ThinShells::mcVisualization          (reference src/ThinShells.cpp:1086-1129)
//   MC::marching_cubes                   (reference MarchingCubes.h:136-139),
//                                         the reference's argument list, and an
//                                         overload on the octree object
//   ThinShells(filename, V3i), ThinShells(filename, V3i, is2UnitCube,
//     scaleFactor)                       (reference src/ThinShells.h:52-60, 88-97)
//   ThinShells::multiPointQuery(out_file, points)
//                                         (reference src/ThinShells.cpp:932-974)
// Eigen types are replaced by small value types with the same element
// access (V3d::operator[], x(), y(), z()); VXd is std::vector<double>, and
// double3 / uint3 / make_double3 / make_uint3 are small structs of this
// namespace.  Errors from the C ABI are thrown as its_host::Error.  Header
// only; link with its_amd/libits_hip.so.
//
// Differences, by design:
//   - creatShell computes the right-hand side, the fit, f at the model
//     vertices and the shell iso values over the model vertices and the
//     intersection points (its_shell_iso); the intersection points
//     themselves stay on the device.
//   - mcVisualization extracts from the octree's own device arrays through
//     the octree overload of MC::marching_cubes; the reference-shaped
//     overload uploads the host arrays it is given (its_svo_from_vertices).
#pragma once
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cmath>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "its_c.h"
#include "its_io.hpp"

namespace its_host {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc) {
    if (rc != ITS_OK) {
        const char* m = its_last_error();
        throw Error(rc, m ? m : "its error");
    }
}

template <class T>
struct Vec3 {
    T v[3] = {T(0), T(0), T(0)};
    Vec3() = default;
    Vec3(T a, T b, T c) : v{a, b, c} {}
    T& operator[](int i) { return v[i]; }
    const T& operator[](int i) const { return v[i]; }
    T& x() { return v[0]; }
    T& y() { return v[1]; }
    T& z() { return v[2]; }
    const T& x() const { return v[0]; }
    const T& y() const { return v[1]; }
    const T& z() const { return v[2]; }
};
using V3d = Vec3<double>;
using V3i = Vec3<int>;

template <class Real>
struct Triangle {
    Real p1, p2, p3;
    Triangle() = default;
    Triangle(const Real& a, const Real& b, const Real& c) : p1(a), p2(b), p3(c) {}
};
template <class Real>
struct AABox {
    Real boxOrigin, boxEnd, boxWidth;
};

// SVONode fields (reference src/SVO.h:15-32)
struct SVONode {
    uint32_t mortonCode = 0;
    bool isLeaf = true;
    V3d origin;
    double width = 0;
    unsigned int parent = ITS_NONE;
    unsigned int childs[8];
    unsigned int neighbors[27];
};
using node_vertex_type = std::pair<V3d, uint32_t>;
using VXd = std::vector<double>;
// the CUDA vector types of MC::marching_cubes's signature
struct double3 {
    double x, y, z;
};
struct uint3 {
    unsigned x, y, z;
};
inline double3 make_double3(double x, double y, double z) { return double3{x, y, z}; }
inline double3 make_double3(const V3d& v) { return double3{v[0], v[1], v[2]}; }
inline uint3 make_uint3(unsigned x, unsigned y, unsigned z) { return uint3{x, y, z}; }
inline uint3 make_uint3(const V3i& v) { return uint3{(unsigned)v[0], (unsigned)v[1], (unsigned)v[2]}; }

// one context per host thread, created on first use
inline its_ctx default_ctx() {
    thread_local struct Holder {
        its_ctx c = nullptr;
        ~Holder() {
            if (c) its_ctx_destroy(c);
        }
    } h;
    if (!h.c) check(its_ctx_create(0, &h.c));
    return h.c;
}

class SparseVoxelOctree {
public:
    int treeDepth = 0;
    size_t numVoxels = 0, numTreeNodes = 0, numNodeVerts = 0;
    V3i surfaceVoxelGridSize;
    std::vector<size_t> depthNumNodes, esumDepthNodeVerts;

    SparseVoxelOctree() = default;
    SparseVoxelOctree(int gx, int gy, int gz) : surfaceVoxelGridSize(gx, gy, gz) {}
    explicit SparseVoxelOctree(const V3i& g) : surfaceVoxelGridSize(g) {}
    SparseVoxelOctree(const SparseVoxelOctree&) = delete;
    SparseVoxelOctree& operator=(const SparseVoxelOctree&) = delete;
    ~SparseVoxelOctree() {
        if (svo_) its_svo_destroy(svo_);
    }

    // base_filename is kept for the signature; nothing is written
    void createOctree(size_t nModelTris, const std::vector<Triangle<V3d>>& modelTris,
                      const AABox<V3d>& modelBBox, const std::string& base_filename = "") {
        (void)base_filename;
        const int R = surfaceVoxelGridSize[0];
        if (R != surfaceVoxelGridSize[1] || R != surfaceVoxelGridSize[2])
            throw Error(ITS_E_INVALID, "the grid must be cubic");
        std::vector<double> t(9 * nModelTris);
        for (size_t i = 0; i < nModelTris; ++i) {
            const Triangle<V3d>& tr = modelTris[i];
            for (int a = 0; a < 3; ++a) {
                t[9 * i + a] = tr.p1[a];
                t[9 * i + 3 + a] = tr.p2[a];
                t[9 * i + 6 + a] = tr.p3[a];
            }
        }
        if (svo_) {
            its_svo_destroy(svo_);
            svo_ = nullptr;
        }
        const double o[3] = {modelBBox.boxOrigin[0], modelBBox.boxOrigin[1], modelBBox.boxOrigin[2]};
        const double w[3] = {modelBBox.boxWidth[0], modelBBox.boxWidth[1], modelBBox.boxWidth[2]};
        check(its_svo_build(ctx_(), t.data(), (int64_t)nModelTris, o, w, R, &svo_));
        its_svo_info info;
        check(its_svo_get_info(svo_, &info));
        treeDepth = info.tree_depth;
        numVoxels = (size_t)info.num_voxels;
        numTreeNodes = (size_t)info.num_nodes;
        numNodeVerts = (size_t)info.num_verts;
        depthNumNodes.assign(info.depth_num_nodes, info.depth_num_nodes + treeDepth);
        esumDepthNodeVerts.assign(info.esum_verts, info.esum_verts + treeDepth + 1);
        voxel_width_ = info.voxel_width;
    }

    // svoNodeArray / nodeVertexArray of the reference, copied to the host on demand
    std::vector<SVONode> svoNodeArray() const {
        const size_t n = numTreeNodes;
        std::vector<uint32_t> mc(n), par(n), ch(8 * n), nb(27 * n);
        std::vector<uint8_t> leaf(n);
        std::vector<double> org(3 * n), wd(n);
        check(its_svo_export_nodes(svo_, mc.data(), leaf.data(), org.data(), wd.data(), par.data(), ch.data(),
                                   nb.data()));
        std::vector<SVONode> out(n);
        for (size_t i = 0; i < n; ++i) {
            SVONode& s = out[i];
            s.mortonCode = mc[i];
            s.isLeaf = leaf[i] != 0;
            s.origin = V3d(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
            s.width = wd[i];
            s.parent = par[i];
            std::copy(ch.begin() + 8 * i, ch.begin() + 8 * i + 8, s.childs);
            std::copy(nb.begin() + 27 * i, nb.begin() + 27 * i + 27, s.neighbors);
        }
        return out;
    }
    std::vector<node_vertex_type> nodeVertexArray() const {
        std::vector<double> xyz(3 * numNodeVerts);
        std::vector<uint32_t> own(numNodeVerts);
        check(its_svo_export_verts(svo_, xyz.data(), own.data()));
        std::vector<node_vertex_type> out(numNodeVerts);
        for (size_t i = 0; i < numNodeVerts; ++i)
            out[i] = {V3d(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]), own[i]};
        return out;
    }

    // fineNodeEdgeArray: level-0 node edges (lower end, upper end) and the
    // smallest index of the level-0 nodes containing each (SVO.cu:965-1110)
    std::vector<std::pair<std::pair<V3d, V3d>, uint32_t>> fineNodeEdgeArray() const {
        int64_t n = 0;
        check(its_svo_export_fine_edges(svo_, nullptr, nullptr, 0, &n));
        std::vector<double> e(6 * (size_t)n);
        std::vector<uint32_t> o((size_t)n);
        if (n) check(its_svo_export_fine_edges(svo_, e.data(), o.data(), n, &n));
        std::vector<std::pair<std::pair<V3d, V3d>, uint32_t>> out((size_t)n);
        for (int64_t i = 0; i < n; ++i)
            out[i] = {{V3d(e[6 * i], e[6 * i + 1], e[6 * i + 2]), V3d(e[6 * i + 3], e[6 * i + 4], e[6 * i + 5])}, o[i]};
        return out;
    }

    // cubes of every level, one OBJ per level: <dir>/svo_1.obj (finest) ..
    // svo_<treeDepth>.obj (SVO.cu:1266-1318, saveSVO)
    void saveSVO(const std::string& dir) const {
        const size_t n = numTreeNodes;
        std::vector<uint32_t> mc(n), par(n), ch(8 * n), nb(27 * n);
        std::vector<uint8_t> leaf(n);
        std::vector<double> org(3 * n), wd(n);
        check(its_svo_export_nodes(svo_, mc.data(), leaf.data(), org.data(), wd.data(), par.data(), ch.data(),
                                   nb.data()));
        size_t first = 0;
        for (int d = 0; d < treeDepth; ++d) {
            std::vector<io::P3> o;
            std::vector<double> w;
            for (size_t i = first; i < first + depthNumNodes[d]; ++i) {
                o.push_back({org[3 * i], org[3 * i + 1], org[3 * i + 2]});
                w.push_back(wd[i]);
            }
            io::save_cubes(dir + "/svo_" + std::to_string(d + 1) + ".obj", o, w);
            first += depthNumNodes[d];
        }
    }

    its_svo handle() const { return svo_; }
    its_ctx ctx_() const { return default_ctx(); }
    double voxelWidth() const { return voxel_width_; }

private:
    its_svo svo_ = nullptr;
    double voxel_width_ = 0;
};

namespace Test {
enum type { CPU = 0, CPU_SIMD = 1, CUDA = 2 };
}

class ThinShells {
public:
    int treeDepth = 0;
    std::string uniformDir = "non-uniform";
    std::string noiseDir = "non-noise";
    std::string modelName = "model";

    // BaseModel(vector<V3d> verts, vector<V3i> faces): no normalisation
    ThinShells(const std::vector<V3d>& verts, const std::vector<V3i>& faces, int gx, int gy, int gz,
               bool is2UnitCube = false, double scaleFactor = 1.0)
        : svo(gx, gy, gz) {
        init(verts, faces, is2UnitCube, scaleFactor);
    }
    // ThinShells(filename, gx, gy, gz, is2UnitCube, scaleFactor): OFF files
    ThinShells(const std::string& filename, int gx, int gy, int gz, bool is2UnitCube = false,
               double scaleFactor = 1.0)
        : svo(gx, gy, gz) {
        std::vector<V3d> V;
        std::vector<V3i> F;
        readOFF(filename, V, F);
        const size_t slash = filename.find_last_of("/\\");
        modelName = filename.substr(slash == std::string::npos ? 0 : slash + 1);
        modelName = modelName.substr(0, modelName.find_last_of('.'));
        init(V, F, is2UnitCube, scaleFactor);
    }

    // ThinShells(filename, V3i) and ThinShells(filename, V3i, is2UnitCube,
    // scaleFactor) (ThinShells.h:52-60, 88-97)
    ThinShells(const std::string& filename, const V3i& grid) : ThinShells(filename, grid[0], grid[1], grid[2]) {}
    ThinShells(const std::string& filename, const V3i& grid, bool is2UnitCube, double scaleFactor)
        : ThinShells(filename, grid[0], grid[1], grid[2], is2UnitCube, scaleFactor) {}

    void creatShell() {
        its_ctx c = svo.ctx_();
        const int64_t n = (int64_t)svo.numNodeVerts;
        sdfVal.assign(n, 0.0);
        check(its_sdf_at_vertices(c, svo.handle(), mV.data(), (int64_t)(mV.size() / 3), mF.data(),
                                  (int64_t)(mF.size() / 3), sdfVal.data()));
        lambda.assign(n, 0.0);
        check(its_lscg_solve(c, svo.handle(), nullptr, lambda.data(), tol, 0, 0, &lastStats));
        const int64_t nv = (int64_t)(mV.size() / 3), nf = (int64_t)(mF.size() / 3);
        double iso[2];
        int64_t npts = 0;
        check(its_shell_iso(c, svo.handle(), mV.data(), nv, mF.data(), nf, iso, &npts));
        innerShellIsoVal = iso[0];
        outerShellIsoVal = iso[1];
        numInterPoints = npts;
        // the latent points (cpIntersectionPoints, ThinShells.cpp:38-479):
        // edge points sorted by the reference's fuzzy lexicographic order with
        // exact duplicates removed, then the face points
        int64_t ne = 0, nfc = 0;
        check(its_shell_points(c, svo.handle(), mV.data(), nv, mF.data(), nf, nullptr, 0, &ne, &nfc));
        std::vector<double> P(3 * (size_t)(ne + nfc));
        if (ne + nfc)
            check(its_shell_points(c, svo.handle(), mV.data(), nv, mF.data(), nf, P.data(), ne + nfc, &ne, &nfc));
        std::vector<io::P3> edge(ne);
        for (int64_t i = 0; i < ne; ++i) edge[i] = {P[3 * i], P[3 * i + 1], P[3 * i + 2]};
        std::sort(edge.begin(), edge.end(), [](const io::P3& a, const io::P3& b) {
            for (int k = 0; k < 3; ++k) {
                if (std::fabs(a[k] - b[k]) < 1e-9) continue;
                return a[k] < b[k];
            }
            return false;
        });
        edge.erase(std::unique(edge.begin(), edge.end()), edge.end());
        allInterPoints.assign(edge.begin(), edge.end());
        for (int64_t i = ne; i < ne + nfc; ++i) allInterPoints.push_back({P[3 * i], P[3 * i + 1], P[3 * i + 2]});
        // f at the model vertices and the latent points (cpLatentBSplineValue)
        std::vector<double> pts(mV);
        for (const io::P3& q : allInterPoints) pts.insert(pts.end(), q.begin(), q.end());
        const int64_t np = (int64_t)(pts.size() / 3);
        std::vector<float> f(np);
        const double end[3] = {bbox.boxEnd[0], bbox.boxEnd[1], bbox.boxEnd[2]};
        check(its_bspline_eval_f64pts(c, svo.handle(), pts.data(), np, end, ITS_EVAL_SUPPORT, 0.f, f.data()));
        bSplineVal.assign(f.begin(), f.end());
    }

    // ---- text outputs (ThinShells.cpp:1017-1081); an empty name gives
    // output/<model>/<uniformDir>/<noiseDir>/<treeDepth>/<file>
    void saveSDFValue(const std::string& filename = "") const { io::save_values(outPath(filename, "SDFValue.txt"), sdfVal); }
    void saveCoefficients(const std::string& filename = "") const {
        io::save_values(outPath(filename, "Coefficients.txt"), lambda);
    }
    void saveLatentPoint(const std::string& filename = "") const {
        std::vector<io::P3> pts;
        for (size_t i = 0; i < mV.size() / 3; ++i) pts.push_back({mV[3 * i], mV[3 * i + 1], mV[3 * i + 2]});
        pts.insert(pts.end(), allInterPoints.begin(), allInterPoints.end());
        std::ofstream out = io::open_out(outPath(filename, "latent_point.xyz"));
        io::write_point_cloud(out, pts, ".xyz");
    }
    void saveBSplineValue(const std::string& filename = "") const {
        io::save_values_fixed9(outPath(filename, "BSplineValue.txt"), bSplineVal);
    }

    // Halton points in the model box scaled by scaleFactor about its centre
    // and moved by dis, also written to filename (".obj" or ".xyz")
    // (BaseModel::generateUniformRandomPoints, BaseModel.cpp:246-264)
    std::vector<V3d> generateUniformRandomPoints(const std::string& filename, size_t numPoints, double scaleFactor,
                                                 const V3d& dis) const {
        const std::vector<io::P3> p =
            io::uniform_points({bbox.boxOrigin[0], bbox.boxOrigin[1], bbox.boxOrigin[2]},
                               {bbox.boxEnd[0], bbox.boxEnd[1], bbox.boxEnd[2]}, numPoints, scaleFactor,
                               {dis[0], dis[1], dis[2]});
        if (!filename.empty()) io::save_point_cloud(filename, p);
        std::vector<V3d> out(p.size());
        for (size_t i = 0; i < p.size(); ++i) out[i] = V3d(p[i][0], p[i][1], p[i][2]);
        return out;
    }

    std::vector<V3d> getModelVerts() const {
        std::vector<V3d> v(mV.size() / 3);
        for (size_t i = 0; i < v.size(); ++i) v[i] = V3d(mV[3 * i], mV[3 * i + 1], mV[3 * i + 2]);
        return v;
    }
    std::vector<V3i> getModelFaces() const {
        std::vector<V3i> f(mF.size() / 3);
        for (size_t i = 0; i < f.size(); ++i) f[i] = V3i(mF[3 * i], mF[3 * i + 1], mF[3 * i + 2]);
        return f;
    }
    const std::vector<io::P3>& latentPoints() const { return allInterPoints; }

    std::array<double, 2> getShellIsoVal() const { return {innerShellIsoVal, outerShellIsoVal}; }

    // {-1, 0, 1} per point; points outside the box get outerShellIsoVal
    std::vector<int> multiPointQuery(const std::vector<V3d>& points, double& time, const int& session = 1,
                                     const Test::type& choice = Test::CUDA) {
        (void)choice;
        std::vector<int> result;
        if (innerShellIsoVal == -kInf || outerShellIsoVal == -kInf) return result;
        const int64_t N = (int64_t)points.size();
        std::vector<double> p(3 * N);
        for (int64_t i = 0; i < N; ++i)
            for (int a = 0; a < 3; ++a) p[3 * i + a] = points[i][a];
        std::vector<float> f(N);
        std::vector<int8_t> cls(N);
        const double end[3] = {bbox.boxEnd[0], bbox.boxEnd[1], bbox.boxEnd[2]};
        auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < std::max(session, 1); ++s)
            check(its_bspline_eval_f64pts(svo.ctx_(), svo.handle(), p.data(), N, end, ITS_EVAL_SUPPORT,
                                          (float)outerShellIsoVal, f.data()));
        auto t1 = std::chrono::steady_clock::now();
        time = std::chrono::duration<double>(t1 - t0).count() / std::max(session, 1);
        check(its_classify(f.data(), N, innerShellIsoVal, outerShellIsoVal, cls.data()));
        result.assign(cls.begin(), cls.end());
        return result;
    }
    int singlePointQuery(const V3d& point) {
        double t;
        std::vector<int> r = multiPointQuery({point}, t, 1);
        return r.empty() ? 0 : r[0];
    }
    // multiPointQuery(out_file, points) (ThinShells.cpp:932-966): f of every
    // point over all basis functions (cuAcc::cpBSplineVal, the RAW mode), as
    // an OBJ point cloud "v x y z r g b": green (0.56471, 0.93333, 0.56471)
    // where inner < f < outer, red (1, 0.27059, 0) elsewhere.  A name without
    // the .obj extension gets it.
    void multiPointQuery(const std::string& out_file, const std::vector<V3d>& points) const {
        if (innerShellIsoVal == -kInf || outerShellIsoVal == -kInf) {
            std::printf("Error: You must create shells first!");
            return;
        }
        const int64_t N = (int64_t)points.size();
        std::vector<double> p(3 * N);
        for (int64_t i = 0; i < N; ++i)
            for (int a = 0; a < 3; ++a) p[3 * i + a] = points[i][a];
        std::vector<float> f(N);
        const double end[3] = {bbox.boxEnd[0], bbox.boxEnd[1], bbox.boxEnd[2]};
        if (N) check(its_bspline_eval_f64pts(svo.ctx_(), svo.handle(), p.data(), N, end, ITS_EVAL_RAW, 0.f, f.data()));
        std::string name = out_file;
        if (io::extension(name) != ".obj") {
            const size_t dot = name.find_last_of('.'), slash = name.find_last_of('/');
            if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) name = name.substr(0, dot);
            name += ".obj";
        }
        std::ofstream out = io::open_out(name);
        std::printf("-- Save query result to \"%s\"-- [RED] points not on the surface, [GREEN] points lie on the "
                    "surface\n", name.c_str());
        for (int64_t i = 0; i < N; ++i) {
            const double v = f[i];
            const bool on = innerShellIsoVal < v && v < outerShellIsoVal;
            out << "v " << points[i][0] << " " << points[i][1] << " " << points[i][2] << " "
                << (on ? 0.56471 : 1.0) << " " << (on ? 0.93333 : 0.27059) << " " << (on ? 0.56471 : 0.0) << '\n';
        }
    }
    // singlePointQuery(out_file, point) (ThinShells.cpp:901-930): the same for one point
    void singlePointQuery(const std::string& out_file, const V3d& point) const {
        multiPointQuery(out_file, std::vector<V3d>{point});
    }

    void mcVisualization(const std::string& innerFilename, const V3i& innerResolution, const std::string& outerFilename,
                         const V3i& outerResolution, const std::string& isoFilename, const V3i& isoResolution) const;

    const std::vector<double>& coefficients() const { return lambda; }
    const std::vector<double>& sdfValues() const { return sdfVal; }
    const std::vector<double>& latentValues() const { return bSplineVal; }
    const its_lscg_stats& solveStats() const { return lastStats; }
    const AABox<V3d>& boundingBox() const { return bbox; }
    const SparseVoxelOctree& octree() const { return svo; }
    double tol = 1e-6;   // stopping tolerance of the fit (normal-equation residual)
    int64_t numInterPoints = 0;   // intersection points behind the iso values

private:
    static constexpr double kInf = std::numeric_limits<double>::max();
    SparseVoxelOctree svo;
    std::vector<double> mV;
    std::vector<int32_t> mF;
    AABox<V3d> bbox;
    std::vector<double> sdfVal, lambda, bSplineVal;
    std::vector<io::P3> allInterPoints;
    double innerShellIsoVal = -kInf, outerShellIsoVal = -kInf;

    std::string outPath(const std::string& filename, const std::string& name) const {
        if (!filename.empty()) return filename;
        return "output/" + modelName + "/" + uniformDir + "/" + noiseDir + "/" + std::to_string(treeDepth) + "/" + name;
    }
    its_lscg_stats lastStats{};

    void init(const std::vector<V3d>& verts, const std::vector<V3i>& faces, bool unit, double scale) {
        mV.resize(3 * verts.size());
        for (size_t i = 0; i < verts.size(); ++i)
            for (int a = 0; a < 3; ++a) mV[3 * i + a] = verts[i][a];
        mF.resize(3 * faces.size());
        for (size_t i = 0; i < faces.size(); ++i)
            for (int a = 0; a < 3; ++a) mF[3 * i + a] = faces[i][a];
        if (unit) {
            uniformDir = "uniform";
            check(its_model_to_unit_cube(mV.data(), (int64_t)verts.size(), scale));
        }
        double o[3], e[3], w[3];
        check(its_uniform_bbox(mV.data(), (int64_t)verts.size(), o, e, w));
        bbox.boxOrigin = V3d(o[0], o[1], o[2]);
        bbox.boxEnd = V3d(e[0], e[1], e[2]);
        bbox.boxWidth = V3d(w[0], w[1], w[2]);
        std::vector<Triangle<V3d>> tris(faces.size());
        for (size_t i = 0; i < faces.size(); ++i) {
            auto vert = [&](int k) {
                const int j = mF[3 * i + k];
                return V3d(mV[3 * j], mV[3 * j + 1], mV[3 * j + 2]);
            };
            tris[i] = Triangle<V3d>(vert(0), vert(1), vert(2));
        }
        svo.createOctree(tris.size(), tris, bbox, "");
        treeDepth = svo.treeDepth;
    }

    static void readOFF(const std::string& fn, std::vector<V3d>& V, std::vector<V3i>& F) {
        std::ifstream in(fn);
        if (!in) throw Error(ITS_E_INVALID, "cannot open " + fn);
        std::string line;
        auto next = [&](std::string& l) {
            while (std::getline(in, l)) {
                const size_t p = l.find_first_not_of(" \t\r");
                if (p == std::string::npos || l[p] == '#') continue;
                return true;
            }
            return false;
        };
        if (!next(line)) throw Error(ITS_E_INVALID, "empty OFF file");
        std::istringstream hs(line);
        std::string tag;
        hs >> tag;
        if (tag != "OFF") throw Error(ITS_E_INVALID, "not an OFF file");
        long nv = -1, nf = -1, ne = 0;
        if (!(hs >> nv >> nf >> ne)) {
            if (!next(line)) throw Error(ITS_E_INVALID, "truncated OFF header");
            std::istringstream cs(line);
            cs >> nv >> nf >> ne;
        }
        V.resize(nv);
        for (long i = 0; i < nv; ++i) {
            if (!next(line)) throw Error(ITS_E_INVALID, "truncated OFF vertices");
            std::istringstream vs(line);
            vs >> V[i][0] >> V[i][1] >> V[i][2];
        }
        F.clear();
        for (long i = 0; i < nf; ++i) {
            if (!next(line)) throw Error(ITS_E_INVALID, "truncated OFF faces");
            std::istringstream fs(line);
            int k;
            fs >> k;
            std::vector<int> idx(k);
            for (int j = 0; j < k; ++j) fs >> idx[j];
            for (int j = 1; j + 1 < k; ++j) F.push_back(V3i(idx[0], idx[j], idx[j + 1]));   // fan
        }
    }
};

namespace MC {
namespace detail {
// soup of one octree object at one iso value (lambda: NULL = its current
// coefficients), written as OBJ when filename is not empty
inline int64_t extract(its_ctx ctx, its_svo svo, const double* lambda, const double o[3], const double w[3],
                       const int res[3], double isoVal, const std::string& filename) {
    its_mc_stats st;
    int64_t nv = 0;
    check(its_mc_extract(ctx, svo, lambda, o, w, res, &isoVal, 1, nullptr, 0, &nv, &st));
    std::vector<float> tri(3 * (size_t)nv);
    if (nv) check(its_mc_extract(ctx, svo, nullptr, o, w, res, &isoVal, 1, tri.data(), nv, &nv, &st));
    if (!filename.empty()) {
        std::vector<io::P3> soup(nv);
        for (int64_t i = 0; i < nv; ++i) soup[i] = {tri[3 * i], tri[3 * i + 1], tri[3 * i + 2]};
        io::save_soup_obj(filename, soup);
    }
    return nv;
}
}  // namespace detail

// Triangle soup of the iso-surface over gridOrigin + [0, gridWidth) at
// `resolution` voxels, written as OBJ (one vertex per corner, faces 3 by 3,
// the soup order of MarchingCubes.cu:635-661), from the octree's own arrays
// and current coefficients.  Returns the vertex count.
inline int64_t marching_cubes(const SparseVoxelOctree& svo, const V3d& gridOrigin, const V3d& gridWidth,
                              const V3i& resolution, double isoVal, const std::string& filename) {
    const double o[3] = {gridOrigin[0], gridOrigin[1], gridOrigin[2]};
    const double w[3] = {gridWidth[0], gridWidth[1], gridWidth[2]};
    const int res[3] = {resolution[0], resolution[1], resolution[2]};
    return detail::extract(svo.ctx_(), svo.handle(), nullptr, o, w, res, isoVal, filename);
}

// The reference's MC::marching_cubes (MarchingCubes.h:136-139): the field
// f(x) = sum_i lambda_i B_i(x) of the vertices nodeVertexArray (position,
// owner node) with the widths nodeWidthArray of their owners, on the grid
// gridOrigin + [0, gridWidth) of `resolution` voxels, iso value isoVal; the
// soup is written to `filename` as OBJ.  The arrays are uploaded on each call
// (its_svo_from_vertices) and released after it.
inline void marching_cubes(const std::vector<node_vertex_type>& nodeVertexArray, const size_t& numNodes,
                           const std::vector<V3d>& nodeWidthArray, const size_t& numNodeVerts, const VXd& lambda,
                           const double3& gridOrigin, const double3& gridWidth, const uint3& resolution,
                           const double& isoVal, const std::string& filename) {
    if (nodeVertexArray.size() < numNodeVerts || nodeWidthArray.size() < numNodes || lambda.size() < numNodeVerts)
        throw Error(ITS_E_INVALID, "MC::marching_cubes: arrays shorter than numNodeVerts / numNodes");
    std::vector<double> xyz(3 * numNodeVerts), nw(3 * numNodes);
    std::vector<uint32_t> owner(numNodeVerts);
    for (size_t i = 0; i < numNodeVerts; ++i) {
        for (int a = 0; a < 3; ++a) xyz[3 * i + a] = nodeVertexArray[i].first[a];
        owner[i] = nodeVertexArray[i].second;
    }
    for (size_t j = 0; j < numNodes; ++j)
        for (int a = 0; a < 3; ++a) nw[3 * j + a] = nodeWidthArray[j][a];
    its_ctx ctx = default_ctx();
    its_svo s = nullptr;
    check(its_svo_from_vertices(ctx, xyz.data(), owner.data(), (int64_t)numNodeVerts, nw.data(), (int64_t)numNodes,
                                &s));
    struct Release {
        its_svo s;
        ~Release() { its_svo_destroy(s); }
    } release{s};
    const double o[3] = {gridOrigin.x, gridOrigin.y, gridOrigin.z};
    const double w[3] = {gridWidth.x, gridWidth.y, gridWidth.z};
    const int res[3] = {(int)resolution.x, (int)resolution.y, (int)resolution.z};
    const int64_t nv = detail::extract(ctx, s, lambda.data(), o, w, res, isoVal, filename);
    std::printf("-- [MC] %lld soup vertices at iso %.9g\n", (long long)nv, isoVal);
}
}  // namespace MC

// Point in/out validation (reference main.cpp:44-134, testPointInOut): the
// shell classification of Halton points against the sign of the exact
// signed distance (its_signed_distance: binary64 nearest triangle, winding
// number sign), which stands in for the reference's fcpw scene.  Prints the
// reference's report lines and returns the figures.  queryResFile, unused by
// the reference, receives the points whose classes differ (".xyz" / ".obj")
// when it is not empty.
struct PointInOutReport {
    size_t numPoints = 0, correct = 0, numErrorPoints = 0;
    double correctRate = 0;         // percent
    double oursSeconds = 0;         // one multiPointQuery over all points
    double exactSeconds = 0;        // exact signed distance over all points
    double oursPlusExactSeconds = 0;  // ours + exact distance of the differing points
};

inline PointInOutReport testPointInOut(ThinShells& thinShell, size_t numPoints, const std::string& queryFile,
                                       const std::string& queryResFile, int query_distance) {
    PointInOutReport r;
    r.numPoints = numPoints;
    std::printf("\n[Test] Point INSIDE or OUTSIDE surface\n");
    const std::vector<V3d> pts = thinShell.generateUniformRandomPoints(queryFile, numPoints, query_distance, V3d(0, 0, 0));
    const int session = 3;
    double time_1 = 0;
    const std::vector<int> ours = thinShell.multiPointQuery(pts, time_1, session, Test::CPU);
    if (ours.empty()) return r;
    r.oursSeconds = time_1;
    std::printf("-- [Ours]: Multi points query spent %lf us.\n", time_1 * 1e6 / (double)numPoints);

    std::vector<double> V, P(3 * numPoints);
    std::vector<int32_t> F;
    for (const V3d& v : thinShell.getModelVerts()) V.insert(V.end(), {v[0], v[1], v[2]});
    for (const V3i& f : thinShell.getModelFaces()) F.insert(F.end(), {f[0], f[1], f[2]});
    for (size_t i = 0; i < numPoints; ++i)
        for (int a = 0; a < 3; ++a) P[3 * i + a] = pts[i][a];
    std::vector<double> sdf(numPoints);
    auto exact = [&](const double* p, int64_t n, double* out) {
        check(its_signed_distance(thinShell.octree().ctx_(), V.data(), (int64_t)(V.size() / 3), F.data(),
                                  (int64_t)(F.size() / 3), p, n, out));
    };
    auto t0 = std::chrono::steady_clock::now();
    exact(P.data(), (int64_t)numPoints, sdf.data());
    r.exactSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("-- [Exact SDF]: Multi points query spent %lf us.\n", r.exactSeconds * 1e6 / (double)numPoints);

    std::vector<double> errorPts;
    std::vector<io::P3> errorOut;
    for (size_t i = 0; i < numPoints; ++i) {
        const int e = sdf[i] > 0 ? 1 : (sdf[i] < 0 ? -1 : 0);
        if (ours[i] == e) {
            ++r.correct;
        } else {
            errorPts.insert(errorPts.end(), {pts[i][0], pts[i][1], pts[i][2]});
            errorOut.push_back({pts[i][0], pts[i][1], pts[i][2]});
        }
    }
    r.numErrorPoints = errorOut.size();
    r.correctRate = numPoints ? (double)r.correct * 100.0 / (double)numPoints : 0.0;
    std::printf("-- Correct rate = %lf%%\n", r.correctRate);
    if (!queryResFile.empty()) io::save_point_cloud(queryResFile, errorOut);

    // the differing points through the exact distance, averaged over sessions
    double time_3 = 0;
    if (r.numErrorPoints) {
        std::vector<double> d(r.numErrorPoints);
        for (int k = 0; k < session; ++k) {
            auto t1 = std::chrono::steady_clock::now();
            exact(errorPts.data(), (int64_t)r.numErrorPoints, d.data());
            time_3 += std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        }
        time_3 /= session;
    }
    r.oursPlusExactSeconds = time_1 + time_3;
    std::printf("-- [Ours+Exact]: Multi points query spent %lf us.\n", r.oursPlusExactSeconds * 1e6 / (double)numPoints);
    return r;
}

inline void ThinShells::mcVisualization(const std::string& innerFilename, const V3i& innerResolution,
                                        const std::string& outerFilename, const V3i& outerResolution,
                                        const std::string& isoFilename, const V3i& isoResolution) const {
    if (!outerFilename.empty() && outerShellIsoVal != -kInf)
        MC::marching_cubes(svo, bbox.boxOrigin, bbox.boxWidth, outerResolution, outerShellIsoVal, outerFilename);
    if (!innerFilename.empty() && innerShellIsoVal != -kInf)
        MC::marching_cubes(svo, bbox.boxOrigin, bbox.boxWidth, innerResolution, innerShellIsoVal, innerFilename);
    if (!isoFilename.empty()) MC::marching_cubes(svo, bbox.boxOrigin, bbox.boxWidth, isoResolution, 0.0, isoFilename);
}

}  // namespace its_host
