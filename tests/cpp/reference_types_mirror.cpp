// The C++ host mirror (granite_hip/ccz4_hip.hpp) instantiated with GRANITE's own types:
// granite::GridBlock (include/granite/core/grid.hpp, src/core/grid.cpp) and
// granite::spacetime::CCZ4Params (include/granite/spacetime/ccz4.hpp:25-34), compiled from the
// reference's headers and grid.cpp where they lie. Every method of the mirror is instantiated;
// without a gfx950 device the engine constructor throws and the program reports it.
#include "granite/core/grid.hpp"
#include "granite/spacetime/ccz4.hpp"
#include "granite_hip/ccz4_hip.hpp"

#include <iostream>
#include <stdexcept>
#include <vector>

using granite::GridBlock;
using granite::spacetime::CCZ4Params;

int main() {
    GridBlock u(0, 0, {8, 6, 4}, {-1.0, -0.75, -0.5}, {1.0, 0.75, 0.5}, 4, 22);
    GridBlock rhs(1, 0, {8, 6, 4}, {-1.0, -0.75, -0.5}, {1.0, 0.75, 0.5}, 4, 22);
    GridBlock fine(2, 1, {8, 8, 8}, {-0.5, -0.5, -0.25}, {0.5, 0.5, 0.25}, 4, 22);
    // the geometry the mirror hands the engine is the block's own
    const gr_geom g = granite_hip::geom_of(u);
    if (g.n[1] != 6 || g.ng != 4 || g.dx[0] != u.dx(0) || g.lo[2] != -0.5) {
        std::cout << "geometry mismatch\n";
        return 3;
    }
    try {
        granite_hip::Engine eng(0);
        CCZ4Params prm;
        granite_hip::CCZ4Evolution<CCZ4Params> ev(prm, eng);
        const std::size_t N = u.totalSize();
        std::vector<double> rho(N), S(N), ham;
        std::vector<std::array<double, 3>> Si(N), mom;
        std::vector<std::array<double, 6>> Sij(N);
        ev.computeRHS(u, rhs, rho, Si, Sij, S);
        ev.computeRHSVacuum(u, rhs);
        ev.applyDissipation(u, rhs);
        ev.computeConstraints(u, ham, mom);
        granite_hip::Level coarse(eng, std::vector<GridBlock*>{&u});
        granite_hip::Level lf(eng, std::vector<GridBlock*>{&fine});
        coarse.upload(0, u);
        coarse.sspRK3Step(ev.params(), 0.01);
        coarse.download(0, u);
        granite_hip::Hierarchy h(eng, {&coarse, &lf});
        h.prolongate(1);
        h.step(prm, 0.01);
        granite_hip::Checkpoint ck(eng, {&coarse, &lf});
        ck.begin();
        ck.wait();
        ck.copyTo(0, u);
        std::cout << "ran on the device: H2 = " << h.hamiltonianL2() << "\n";
    } catch (const std::runtime_error& ex) {
        std::cout << "runtime_error: " << ex.what() << "\n";
        return 1;
    }
    return 0;
}
