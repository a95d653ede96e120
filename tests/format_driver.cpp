// Writes fixed inputs through the library's text writers (its_io.hpp) into
// the directory given as argv[1]; tests/test_formats.py compares the files
// with the formats restated in oracle/formats.py.
#include <cstdio>
#include <limits>

#include "its_io.hpp"

using namespace its_host::io;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s outdir\n", argv[0]);
        return 2;
    }
    const std::string d = argv[1];
    const std::vector<double> vals = {0.0,  -0.0, 1.0,     0.1,       1e-7,   -123.456789, 1234567.0,
                                      3.14159265358979, 1e20, -2.5e-5, 0.5000004, 99999.95, 1e-300};
    save_values(d + "/values.txt", vals);
    save_values_fixed9(d + "/fixed9.txt", {0.5, -12.25, 1e-10, 123456.789, -0.000000001, 0.0});
    save_values_fixed9(d + "/fixed9_empty.txt", {});
    save_cubes(d + "/cubes.obj", {{0.0, 0.0, 0.0}, {-0.25, 0.125, 1.0 / 3.0}}, {1.0, 0.0625});
    save_soup_obj(d + "/soup.obj", {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0.1, 0.2, 0.3}, {1e-8, -2.5, 7}, {3, 2, 1}});
    const std::vector<P3> pts = uniform_points({-0.1, 0.0, 0.2}, {1.1, 0.9, 1.3}, 40, 1.5, {0.01, -0.02, 0.0});
    save_point_cloud(d + "/points.xyz", pts);
    save_point_cloud(d + "/points.obj", pts);
    save_point_cloud(d + "/points.txt", pts);
    std::FILE* h = std::fopen((d + "/halton.txt").c_str(), "w");
    for (int i = 0; i < 64; ++i)
        std::fprintf(h, "%.17g %.17g %.17g\n", halton(i, 2), halton(i, 3), halton(i, 5));
    std::fclose(h);
    return 0;
}
