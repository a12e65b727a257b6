// Host build of the HIP engine's per-cell CCZ4 math (granite_amd/csrc/ccz4_cell.hpp).
//
// Test-only: the same header the gfx950 stage kernel includes is compiled here with g++
// and driven with a Stager that reads the global array directly, so that the kernel's
// algebra can be compared with the oracle on a CPU. The product never runs this.
#include "../granite_amd/csrc/ccz4_cell.hpp"

#include <cstring>

namespace {

struct HostStager {
    const double* u;
    long stride, c;
    double centre(int v) const { return u[v * stride + c]; }
    gr::RuntimeStrides st;
    const double* field(int v) const { return u + v * stride + c; }
    const gr::RuntimeStrides& strides() const { return st; }
    void metric64(double g[6]) const {
        for (int a = 0; a < 6; ++a) g[a] = u[(gr::GXX + a) * stride + c];
    }
};

struct HostSink {
    double* rhs;
    long stride, c;
    bool ko_on;
    void put(int v, double r, double ko, double) { rhs[v * stride + c] = ko_on ? r + ko : r; }
    void finish() {}
};

} // namespace

template <bool KO_EXACT>
static void run(const int* n, int ng, const double* dx, const double* prm, const double* u, double* rhs,
                const double* rho, const double* Si, const double* Sij, const double* S) {
    const int nt0 = n[0] + 2 * ng, nt1 = n[1] + 2 * ng, nt2 = n[2] + 2 * ng;
    const long stride = (long)nt0 * nt1 * nt2;
    gr::Stencil<double> K;
    for (int d = 0; d < 3; ++d) {
        K.ih[d] = 1.0 / (12.0 * dx[d]);
        K.ih2[d] = 1.0 / (12.0 * dx[d] * dx[d]);
        K.ko[d] = prm[3] / (64.0 * dx[d]);
        const double wts[4] = {1.0, -6.0, 15.0, -20.0};
        for (int s = 0; s < 4; ++s) K.kw[d][s] = K.ko[d] * wts[s];
    }
    K.ihh[0] = 1.0 / (4.0 * dx[0] * dx[1]);
    K.ihh[1] = 1.0 / (4.0 * dx[0] * dx[2]);
    K.ihh[2] = 1.0 / (4.0 * dx[1] * dx[2]);
    gr::CellParams P{prm[0], prm[1], prm[2], prm[4], prm[5]};
    for (int k = ng; k < ng + n[2]; ++k)
        for (int j = ng; j < ng + n[1]; ++j)
            for (int i = ng; i < ng + n[0]; ++i) {
                const long c = i + (long)nt0 * (j + (long)nt1 * k);
                HostStager st{u, stride, c, {{1, (long)nt0, (long)nt0 * nt1}}};
                HostSink sk{rhs, stride, c, prm[3] > 0.0};
                gr::CellSources src{};
                if (rho) {
                    src.rho = rho[c];
                    src.S = S[c];
                    for (int a = 0; a < 3; ++a) src.Si[a] = Si[3 * c + a];
                    for (int a = 0; a < 6; ++a) src.Sij[a] = Sij[6 * c + a];
                    gr::ccz4_cell<double, true, KO_EXACT>(st, sk, K, P, src);
                } else {
                    gr::ccz4_cell<double, false, KO_EXACT>(st, sk, K, P, src);
                }
            }
}

extern "C" void hcell_rhs(const int* n, int ng, const double* dx, const double* prm, const double* u, double* rhs,
                          const double* rho, const double* Si, const double* Sij, const double* S, int ko_exact) {
    if (ko_exact)
        run<true>(n, ng, dx, prm, u, rhs, rho, Si, Sij, S);
    else
        run<false>(n, ng, dx, prm, u, rhs, rho, Si, Sij, S);
}
