// Text formats of the ITS outputs, as the reference writes them (plain C++,
// no GPU): the host classes in its_host.hpp call these.
//
//   - value dumps, one value per line at the stream's default precision:
//     ThinShells::saveSDFValue / saveCoefficients (reference
//     ThinShells.cpp:1017-1046);
//   - f at the latent points, fixed with 9 decimals, right-aligned to the
//     widest value as Eigen prints a column vector: saveBSplineValue
//     (ThinShells.cpp:1068-1081);
//   - point clouds, ".xyz" ("x y z") or ".obj" ("v x y z"): saveLatentPoint,
//     generateUniformRandomPoints (ThinShells.cpp:1048-1066,
//     BaseModel.cpp:246-264, utils/IO.hpp:7-23,126-154);
//   - octree cubes as OBJ quads, 8 vertices and 6 faces per node, one file
//     per level: SparseVoxelOctree::saveSVO with CUBE_WRITE, the reference's
//     x64 build setting (SVO.cu:1266-1318, utils/IO.hpp:44-124,
//     3DThinShell.vcxproj:159,185);
//   - marching-cubes soup as OBJ, three vertices then their face:
//     MC::writeToOBJFile (MarchingCubes.cu:635-661);
//   - the Halton points of getUniformRandomMatrix (utils/common.hpp:320-370)
//     in the scaled and translated box of AABox::scaleAndTranslate
//     (utils/Geometry.hpp:18-30).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <sys/stat.h>

namespace its_host {
namespace io {

using P3 = std::array<double, 3>;

// create the directories on the way to a file (mkdir -p of its parent)
inline void check_dir(const std::string& filename) {
    size_t pos = 0;
    while ((pos = filename.find('/', pos + 1)) != std::string::npos) {
        const std::string dir = filename.substr(0, pos);
        if (!dir.empty()) ::mkdir(dir.c_str(), 0755);
    }
}

inline std::ofstream open_out(const std::string& filename) {
    check_dir(filename);
    std::ofstream out(filename, std::ofstream::out | std::ofstream::trunc);
    if (!out) throw std::runtime_error("cannot write " + filename);
    return out;
}

inline std::string extension(const std::string& filename) {
    const size_t dot = filename.find_last_of('.');
    const size_t slash = filename.find_last_of('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    return filename.substr(dot);
}

// one value per line, default stream format (SDFValue.txt, Coefficients.txt)
inline void save_values(const std::string& filename, const std::vector<double>& v) {
    std::ofstream out = open_out(filename);
    for (double x : v) out << x << '\n';
}

// fixed, 9 decimals, every value right-aligned to the widest one, the last
// line ended too (BSplineValue.txt)
inline void save_values_fixed9(const std::string& filename, const std::vector<double>& v) {
    std::vector<std::string> s(v.size());
    size_t width = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        std::ostringstream o;
        o << std::setiosflags(std::ios::fixed) << std::setprecision(9) << v[i];
        s[i] = o.str();
        width = std::max(width, s[i].size());
    }
    std::ofstream out = open_out(filename);
    for (size_t i = 0; i < s.size(); ++i) {
        if (i) out << '\n';
        out << std::string(width - s[i].size(), ' ') << s[i];
    }
    out << '\n';
}

inline void write_vertex(std::ostream& out, const P3& p) { out << "v " << p[0] << " " << p[1] << " " << p[2] << '\n'; }
inline void write_xyz(std::ostream& out, const P3& p) { out << p[0] << " " << p[1] << " " << p[2] << '\n'; }

// point cloud by extension: ".obj" -> "v x y z", ".xyz" -> "x y z"; other
// extensions leave the file empty, as the reference does
inline void write_point_cloud(std::ostream& out, const std::vector<P3>& pts, const std::string& ext) {
    if (ext == ".obj")
        for (const P3& p : pts) write_vertex(out, p);
    else if (ext == ".xyz")
        for (const P3& p : pts) write_xyz(out, p);
}
inline void save_point_cloud(const std::string& filename, const std::vector<P3>& pts) {
    std::ofstream out = open_out(filename);
    write_point_cloud(out, pts, extension(filename));
}

// one cube: vertices v1..v8 of the reference's corner numbering, then its 6
// quad faces; faceBegIdx counts the vertices written before it
inline void write_cube(std::ostream& out, const P3& o, double w, size_t& faceBegIdx) {
    // corner offsets (x, y, z) of v1..v8; every coordinate is origin + offset
    static const int off[8][3] = {{0, 1, 1}, {0, 0, 1}, {0, 1, 0}, {0, 0, 0},
                                  {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    for (const auto& d : off)
        write_vertex(out, P3{o[0] + (d[0] ? w : 0.0), o[1] + (d[1] ? w : 0.0), o[2] + (d[2] ? w : 0.0)});
    static const int quads[6][4] = {{3, 4, 2, 1}, {6, 5, 4, 3}, {1, 8, 6, 3}, {1, 2, 7, 8}, {4, 5, 7, 2}, {8, 7, 5, 6}};
    for (const auto& q : quads)
        out << "f " << faceBegIdx + q[0] << " " << faceBegIdx + q[1] << " " << faceBegIdx + q[2] << " "
            << faceBegIdx + q[3] << '\n';
    faceBegIdx += 8;
}

// cubes of one octree level (origins xyz, width per node)
inline void save_cubes(const std::string& filename, const std::vector<P3>& origins, const std::vector<double>& widths) {
    std::ofstream out = open_out(filename);
    size_t base = 0;
    for (size_t i = 0; i < origins.size(); ++i) write_cube(out, origins[i], widths[i], base);
}

// marching-cubes soup: each triangle as three vertices and its face
inline void save_soup_obj(const std::string& filename, const std::vector<P3>& soup) {
    std::ofstream out = open_out(filename);
    for (size_t i = 0; i + 2 < soup.size(); i += 3) {
        write_vertex(out, soup[i]);
        write_vertex(out, soup[i + 1]);
        write_vertex(out, soup[i + 2]);
        out << "f " << i + 1 << ' ' << i + 2 << ' ' << i + 3 << '\n';
    }
}

// radical inverse of index in the given base (the Halton coordinate)
inline double halton(int index, int base) {
    double result = 0.0, f = 1.0 / base;
    int i = index;
    while (i > 0) {
        result += f * (i % base);
        i = i / base;
        f /= base;
    }
    return result;
}

// num Halton points (bases 2, 3, 5, indices 0 .. num-1) in the box
// [lo, hi] scaled by s about its centre and moved by t
inline std::vector<P3> uniform_points(const P3& lo, const P3& hi, size_t num, double s, const P3& t) {
    std::vector<P3> pts;
    if (s <= 0.0) return pts;
    P3 a, b;
    for (int k = 0; k < 3; ++k) {
        const double c = (lo[k] + hi[k]) / 2.0;
        a[k] = (lo[k] - c) * s + c + t[k];
        b[k] = (hi[k] - c) * s + c + t[k];
    }
    pts.reserve(num);
    static const int base[3] = {2, 3, 5};
    for (size_t i = 0; i < num; ++i) {
        P3 p;
        for (int k = 0; k < 3; ++k) p[k] = a[k] + halton((int)i, base[k]) * (b[k] - a[k]);
        pts.push_back(p);
    }
    return pts;
}

}  // namespace io
}  // namespace its_host
