// A caller written against the reference's own call shapes (src/main.cpp:
// 165-190, ThinShells::mcVisualization at src/ThinShells.cpp:1086-1129):
//   ThinShells thinShell(file, V3i(R, R, R), is2Cube, 1.0);
//   thinShell.creatShell();
//   thinShell.mcVisualization(inner, V3i(res), outer, V3i(res), iso, V3i(res));
//   MC::marching_cubes(svo.nodeVertexArray, svo.numTreeNodes, nodeWidthArray,
//                      svo.numNodeVerts, lambda, make_double3(gridOrigin),
//                      make_double3(gridWidth), make_uint3(res), isoVal, file);
//   thinShell.multiPointQuery(out_file, points);
// The only edits against the reference text: svo.nodeVertexArray is a call
// here, and the CUDA vector helpers come from its_host.  Prints a JSON line.
//
//   ref_caller model.off R res out_dir
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "its_host.hpp"

using namespace its_host;

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s model.off R res out_dir\n", argv[0]);
        return 2;
    }
    const int svo_res = std::atoi(argv[2]);
    const int res = std::atoi(argv[3]);
    const std::string dir = argv[4];
    try {
        bool is2Cube = true;
        ThinShells thinShell(argv[1], V3i(svo_res, svo_res, svo_res), is2Cube, 1.0);
        thinShell.creatShell();
        thinShell.mcVisualization(dir + "/mc_innerShell.obj", V3i(res, res, res), dir + "/mc_outerShell.obj",
                                  V3i(res, res, res), dir + "/mc_isosurface.obj", V3i(res, res, res));

        // the reference's MC::marching_cubes argument list, from host arrays
        const SparseVoxelOctree& svo = thinShell.octree();
        std::vector<V3d> nodeWidthArray;
        for (const SVONode& node : svo.svoNodeArray()) nodeWidthArray.emplace_back(node.width, node.width, node.width);
        const V3d gridOrigin = thinShell.boundingBox().boxOrigin;
        const V3d gridWidth = thinShell.boundingBox().boxWidth;
        MC::marching_cubes(svo.nodeVertexArray(), svo.numTreeNodes, nodeWidthArray, svo.numNodeVerts,
                           thinShell.coefficients(), make_double3(gridOrigin), make_double3(gridWidth),
                           make_uint3(res, res, res), .0, dir + "/mc_reference_signature.obj");

        // queries written as a coloured point cloud (name without .obj gets it)
        std::vector<V3d> points;
        for (int i = 0; i < 200; ++i) {
            V3d p;
            for (int a = 0; a < 3; ++a)
                p[a] = gridOrigin[a] + gridWidth[a] * std::fmod(0.7548776662466927 * (i + 1) * (a + 1) + 0.2 * a, 1.0);
            points.push_back(p);
        }
        thinShell.multiPointQuery(dir + "/query_result.xyz", points);
        thinShell.singlePointQuery(dir + "/single_query", points[0]);

        const auto iso = thinShell.getShellIsoVal();
        std::printf("{\"tree_depth\": %d, \"num_nodes\": %zu, \"num_verts\": %zu, \"inner_iso\": %.17g, "
                    "\"outer_iso\": %.17g}\n",
                    thinShell.treeDepth, svo.numTreeNodes, svo.numNodeVerts, iso[0], iso[1]);
    } catch (const Error& e) {
        std::fprintf(stderr, "its error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
