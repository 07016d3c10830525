// The reference README's usage (README.md:8-13) against the drop-in header:
//   result(i, j, k) = a(i, c0, j, c1) * b(k, c0, c1)
// Writes two small FROSTT-style files, reads them back, contracts on the GPU
// and prints "nnz <count> sum <value sum>" for the caller to check.
#include <cstdio>
#include <fstream>
#include <random>
#include <set>
#include <tuple>

#include "fastcc.hpp"

static void write_tns(const char *path, const std::vector<std::vector<uint32_t>> &c, const std::vector<double> &v) {
    std::ofstream f(path);
    for (size_t n = 0; n < v.size(); n++) {
        for (auto x : c[n]) f << x << ' ';
        char buf[64];
        snprintf(buf, sizeof buf, "%.17g", v[n]);
        f << buf << '\n';
    }
}

int main(int argc, char **argv) {
    const char *dir = argc > 1 ? argv[1] : ".";
    std::mt19937_64 rng(20261203);
    auto make = [&](std::vector<uint32_t> dims, size_t nnz, const std::string &path) {
        std::set<std::vector<uint32_t>> seen;
        std::vector<std::vector<uint32_t>> c;
        std::vector<double> v;
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        while (c.size() < nnz) {
            std::vector<uint32_t> x;
            for (auto d : dims) x.push_back((uint32_t)(rng() % d));
            if (!seen.insert(x).second) continue;
            c.push_back(x);
            v.push_back(U(rng));
        }
        write_tns(path.c_str(), c, v);
    };
    std::string pa = std::string(dir) + "/a.tns", pb = std::string(dir) + "/b.tns";
    make({40, 16, 30, 8}, 3000, pa);
    make({25, 16, 8}, 1500, pb);

    Tensor<double> a(pa);
    Tensor<double> b(pb);
    ListTensor<double> result = a.fastcc_multiply<TileAccumulator<double>, double>(b, {1, 3}, {1, 2});
    double s = 0.0;
    for (int64_t n = 0; n < result.compute_nnz_count(); n++) s += result.value(n);
    printf("nnz %lld sum %.17g\n", (long long)result.compute_nnz_count(), s);
    return 0;
}
