// Test driver of the C++ host mirror (granite_amd/host/granite_hip/ccz4_hip.hpp).
//
// Stands in for a GRANITE translation unit: a block type with GridBlock's API (a test double
// holding the same storage) and a params struct with CCZ4Params' members. Reads a block
// from a file, runs computeRHSVacuum, computeRHS with zero sources, computeConstraints and
// two Level::sspRK3Step through the mirror, and writes the results for the caller to check;
// then steps a two-level Hierarchy and checks a Checkpoint against the levels' state.
//   host_mirror_driver <in.bin> <out.bin>  (in: int n[3], int ng, double lo[3], double hi[3], data)
#include "granite_hip/ccz4_hip.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

struct TestBlock {
    std::array<int, 3> n;
    int ng;
    std::array<double, 3> lo, hi, dx;
    std::vector<double> data;
    TestBlock(std::array<int, 3> n_, int ng_, std::array<double, 3> lo_, std::array<double, 3> hi_)
        : n(n_), ng(ng_), lo(lo_), hi(hi_) {
        for (int d = 0; d < 3; ++d) dx[d] = (hi[d] - lo[d]) / n[d];
        data.assign(22 * totalSize(), 0.0);
    }
    int getNumVars() const { return 22; }
    int getNumGhost() const { return ng; }
    const std::array<int, 3>& numCells() const { return n; }
    const std::array<double, 3>& lowerCorner() const { return lo; }
    const std::array<double, 3>& cellSpacing() const { return dx; }
    std::size_t totalSize() const { return (std::size_t)(n[0] + 2 * ng) * (n[1] + 2 * ng) * (n[2] + 2 * ng); }
    double* flatBuffer() { return data.data(); }
    const double* flatBuffer() const { return data.data(); }
};

struct Params {
    double kappa1 = 0.02, kappa2 = 0.0, eta = 2.0, ko_sigma = 0.1;
    int fd_order = 4;
    bool use_moving_puncture = true;
    double chi_blend_center = 0.05, chi_blend_width = 0.02;
};

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[4];
    double box[6];
    if (std::fread(hdr, sizeof(int), 4, f) != 4 || std::fread(box, sizeof(double), 6, f) != 6) return 2;
    TestBlock u({hdr[0], hdr[1], hdr[2]}, hdr[3], {box[0], box[1], box[2]}, {box[3], box[4], box[5]});
    if (std::fread(u.data.data(), sizeof(double), u.data.size(), f) != u.data.size()) return 2;
    std::fclose(f);
    try {
        granite_hip::Engine eng(0);
        Params prm;
        granite_hip::CCZ4Evolution<Params> ev(prm, eng);
        TestBlock r1 = u, r2 = u;
        std::fill(r1.data.begin(), r1.data.end(), 0.0);
        std::fill(r2.data.begin(), r2.data.end(), 0.0);
        ev.computeRHSVacuum(u, r1);
        const std::size_t N = u.totalSize();
        std::vector<double> rho(N, 0.0), S(N, 0.0);
        std::vector<std::array<double, 3>> Si(N, {0, 0, 0});
        std::vector<std::array<double, 6>> Sij(N, {0, 0, 0, 0, 0, 0});
        ev.computeRHS(u, r2, rho, Si, Sij, S);
        std::vector<double> ham;
        std::vector<std::array<double, 3>> mom;
        ev.computeConstraints(u, ham, mom);
        TestBlock s = u;
        std::vector<TestBlock*> blocks{&s};
        granite_hip::Level lvl(eng, blocks);
        lvl.upload(0, s);
        lvl.sspRK3Step(prm, 0.0625, GR_PREC_FP64, GR_BC_SOMMERFELD, GR_BC_NONE, false);
        lvl.sspRK3Step(prm, 0.0625, GR_PREC_FP64, GR_BC_SOMMERFELD, GR_BC_NONE, false);
        lvl.sync();
        lvl.download(0, s);
        FILE* o = std::fopen(argv[2], "wb");
        std::fwrite(r1.data.data(), sizeof(double), r1.data.size(), o);
        std::fwrite(r2.data.data(), sizeof(double), r2.data.size(), o);
        std::fwrite(ham.data(), sizeof(double), ham.size(), o);
        std::fwrite(mom.data()->data(), sizeof(double), 3 * mom.size(), o);
        std::fwrite(s.data.data(), sizeof(double), s.data.size(), o);
        std::fclose(o);
        // a two-level hierarchy through the mirror (a child at half the spacing inside the
        // block), one coarse step, then a checkpoint that must equal the levels' state
        TestBlock c0 = u;
        TestBlock c1({16, 16, 8}, 4, {-1.0, -1.0, -0.5}, {1.0, 1.0, 0.5});
        granite_hip::Level L0(eng, std::vector<TestBlock*>{&c0}), L1(eng, std::vector<TestBlock*>{&c1});
        L0.upload(0, c0);
        granite_hip::Hierarchy H(eng, {&L0, &L1});
        H.prolongate(1);
        H.fillGhostZones(0);
        H.step(prm, 0.0625);
        const double h2 = H.hamiltonianL2();
        granite_hip::Checkpoint ck(eng, {&L0, &L1});
        ck.begin();
        ck.wait();
        TestBlock k0 = c0, k1 = c1;
        ck.copyTo(0, k0);
        ck.copyTo(1, k1);
        L0.download(0, c0);
        L1.download(0, c1);
        if (ck.numBlocks() != 2 || k0.data != c0.data || k1.data != c1.data || !std::isfinite(h2)) {
            std::printf("hierarchy/checkpoint mismatch\n");
            return 4;
        }
        std::printf("hierarchy ok: ||H||2 = %.6e\n", h2);
        // an invalid call must surface as an exception (ng = 3)
        TestBlock bad({4, 4, 4}, 3, {0, 0, 0}, {1, 1, 1}), badr = bad;
        try {
            ev.computeRHSVacuum(bad, badr);
            std::printf("no exception for ng=3\n");
            return 3;
        } catch (const std::runtime_error& e) {
            std::printf("expected error: %s\n", e.what());
        }
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
