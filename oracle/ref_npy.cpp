// TEST INFRASTRUCTURE ONLY: writes .npy files with the Nimbus reference's own
// NPY writer (src/core/output/npy_writer.cpp, compiled where it lies by the
// `ref` target of oracle/Makefile) for golden byte comparisons.
//
// Usage: ref_npy <out.npy> f4|f8 <d0> <d1> [<d2>]
// The payload is element c (C order) = 0.25 * c - 1.5 (exact in f4 and f8).
#include "core/output/npy_writer.hpp"

#include <cstdlib>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 5) return 2;
    const std::string path = argv[1], dt = argv[2];
    const int d0 = std::atoi(argv[3]), d1 = std::atoi(argv[4]), d2 = argc > 5 ? std::atoi(argv[5]) : 0;
    const std::size_t n = std::size_t(d0) * d1 * (d2 > 0 ? d2 : 1);
    bool ok = false;
    if (dt == "f4")
    {
        std::vector<float> v(n);
        for (std::size_t c = 0; c < n; ++c) v[c] = float(0.25 * double(c) - 1.5);
        ok = d2 > 0 ? npy::write_3d(v.data(), d0, d1, d2, path) : npy::write_2d(v.data(), d0, d1, path);
    }
    else if (dt == "f8" && d2 > 0)
    {
        std::vector<double> v(n);
        for (std::size_t c = 0; c < n; ++c) v[c] = 0.25 * double(c) - 1.5;
        ok = npy::write_3d(v.data(), d0, d1, d2, path);
    }
    return ok ? 0 : 1;
}
