// A main.cpp-style caller of the host classes (reference src/main.cpp:165-190):
// build the shells of an OFF model, classify query points, extract the
// iso-surface, and print a one-line JSON summary.  With a dump directory it
// also writes the reference's text outputs there (SDFValue.txt,
// Coefficients.txt, latent_point.xyz, BSplineValue.txt, svo/svo_<d>.obj) and
// runs the point in/out validation of main.cpp:44-134 on nq Halton points
// (query.xyz, query_err.xyz).
//
//   shell_demo model.off R [nq] [out.obj] [dump_dir]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "its_host.hpp"

using namespace its_host;

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s model.off R [nq] [out.obj] [dump_dir]\n", argv[0]);
        return 2;
    }
    const int R = std::atoi(argv[2]);
    const int nq = argc > 3 ? std::atoi(argv[3]) : 1000;
    const std::string obj = argc > 4 ? argv[4] : "";
    const std::string dump = argc > 5 ? argv[5] : "";
    try {
        ThinShells shell(argv[1], R, R, R, true, 1.0);
        shell.creatShell();
        const auto iso = shell.getShellIsoVal();
        // deterministic query points over the padded box
        const AABox<V3d>& b = shell.boundingBox();
        std::vector<V3d> pts(nq);
        for (int i = 0; i < nq; ++i)
            for (int a = 0; a < 3; ++a) {
                const double u = std::fmod(0.6180339887498949 * (i + 1) * (a + 1) + 0.1 * a, 1.0);
                pts[i][a] = b.boxOrigin[a] - 0.05 * b.boxWidth[a] + 1.1 * u * b.boxWidth[a];
            }
        double t = 0;
        const std::vector<int> cls = shell.multiPointQuery(pts, t, 1);
        int counts[3] = {0, 0, 0};
        for (int c : cls) counts[c + 1]++;
        int64_t mc_verts = MC::marching_cubes(shell.octree(), b.boxOrigin, b.boxWidth, V3i(R, R, R), 0.0, obj);
        double lsum = 0, labs = 0;
        for (double l : shell.coefficients()) {
            lsum += l;
            labs += std::fabs(l);
        }
        const SparseVoxelOctree& s = shell.octree();
        PointInOutReport rep;
        if (!dump.empty()) {
            shell.saveSDFValue(dump + "/SDFValue.txt");
            shell.saveCoefficients(dump + "/Coefficients.txt");
            shell.saveLatentPoint(dump + "/latent_point.xyz");
            shell.saveBSplineValue(dump + "/BSplineValue.txt");
            s.saveSVO(dump + "/svo");
            rep = testPointInOut(shell, (size_t)nq, dump + "/query.xyz", dump + "/query_err.xyz", 1);
        }
        std::printf("{\"tree_depth\": %d, \"num_voxels\": %zu, \"num_nodes\": %zu, \"num_verts\": %zu, "
                    "\"lambda_sum\": %.17g, \"lambda_abs\": %.17g, \"inner_iso\": %.9g, \"outer_iso\": %.9g, "
                    "\"inside\": %d, \"shell\": %d, \"outside\": %d, \"mc_verts\": %lld, \"iterations\": %lld, "
                    "\"normal_rel\": %.3e, \"latent_points\": %zu, \"check_correct\": %zu, "
                    "\"check_errors\": %zu, \"check_rate\": %.6f}\n",
                    shell.treeDepth, s.numVoxels, s.numTreeNodes, s.numNodeVerts, lsum, labs, iso[0], iso[1],
                    counts[0], counts[1], counts[2], (long long)mc_verts, (long long)shell.solveStats().iterations,
                    shell.solveStats().normal_rel, shell.latentPoints().size(), rep.correct, rep.numErrorPoints,
                    rep.correctRate);
    } catch (const Error& e) {
        std::fprintf(stderr, "its error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
