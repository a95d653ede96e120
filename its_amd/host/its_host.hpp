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
//   MC::marching_cubes                   (reference MarchingCubes.h:136-139)
// Eigen types are replaced by small value types with the same element
// access (V3d::operator[], x(), y(), z()).  Errors from the C ABI are thrown
// as its_host::Error.  Header only; link with its_amd/libits_hip.so.
//
// Differences, by design:
//   - creatShell computes the right-hand side, the fit, f at the model
//     vertices and the shell iso values over the model vertices and the
//     intersection points (its_shell_iso); the intersection points
//     themselves stay on the device.
//   - MC::marching_cubes takes the octree object (which holds the vertex
//     arrays and the coefficients on the device) instead of host copies of
//     nodeVertexArray, nodeWidthArray and lambda.
#pragma once
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "its_c.h"

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
        init(V, F, is2UnitCube, scaleFactor);
    }

    void creatShell() {
        its_ctx c = svo.ctx_();
        const int64_t n = (int64_t)svo.numNodeVerts;
        sdfVal.assign(n, 0.0);
        check(its_sdf_at_vertices(c, svo.handle(), mV.data(), (int64_t)(mV.size() / 3), mF.data(),
                                  (int64_t)(mF.size() / 3), sdfVal.data()));
        lambda.assign(n, 0.0);
        check(its_lscg_solve(c, svo.handle(), nullptr, lambda.data(), tol, 0, 0, &lastStats));
        const int64_t nv = (int64_t)(mV.size() / 3);
        std::vector<float> f(nv);
        const double end[3] = {bbox.boxEnd[0], bbox.boxEnd[1], bbox.boxEnd[2]};
        check(its_bspline_eval_f64pts(c, svo.handle(), mV.data(), nv, end, ITS_EVAL_SUPPORT, 0.f, f.data()));
        bSplineVal.assign(f.begin(), f.end());
        double iso[2];
        int64_t npts = 0;
        check(its_shell_iso(c, svo.handle(), mV.data(), nv, mF.data(), (int64_t)(mF.size() / 3), iso, &npts));
        innerShellIsoVal = iso[0];
        outerShellIsoVal = iso[1];
        numInterPoints = npts;
    }

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
    double innerShellIsoVal = -kInf, outerShellIsoVal = -kInf;
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
// Triangle soup of the iso-surface over gridOrigin + [0, gridWidth) at
// `resolution` voxels, written as OBJ (one vertex per corner, faces 3 by 3,
// the soup order of MarchingCubes.cu:635-661).  Returns the vertex count.
inline int64_t marching_cubes(const SparseVoxelOctree& svo, const V3d& gridOrigin, const V3d& gridWidth,
                              const V3i& resolution, double isoVal, const std::string& filename) {
    const double o[3] = {gridOrigin[0], gridOrigin[1], gridOrigin[2]};
    const double w[3] = {gridWidth[0], gridWidth[1], gridWidth[2]};
    const int res[3] = {resolution[0], resolution[1], resolution[2]};
    its_mc_stats st;
    check(its_mc_extract(svo.ctx_(), svo.handle(), o, w, res, isoVal, nullptr, 0, &st));
    std::vector<float> tri(3 * (size_t)st.num_verts);
    if (st.num_verts)
        check(its_mc_extract(svo.ctx_(), svo.handle(), o, w, res, isoVal, tri.data(), st.num_verts, &st));
    if (!filename.empty()) {
        std::ofstream out(filename);
        if (!out) throw Error(ITS_E_INVALID, "cannot write " + filename);
        for (int64_t i = 0; i < st.num_verts; ++i)
            out << "v " << tri[3 * i] << ' ' << tri[3 * i + 1] << ' ' << tri[3 * i + 2] << '\n';
        for (int64_t i = 0; i + 2 < st.num_verts; i += 3) out << "f " << i + 1 << ' ' << i + 2 << ' ' << i + 3 << '\n';
    }
    return st.num_verts;
}
}  // namespace MC

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
