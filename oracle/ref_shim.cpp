// C entry points over the compiled GRANITE reference (test infrastructure only).
//
// This file is ours; it includes the reference's public headers and calls its own
// classes, so that tests can run the reference CPU path on the same inputs as the
// oracle and the HIP engine. It is compiled together with the reference sources
// (src/spacetime/ccz4.cpp, src/core/grid.cpp, src/amr/amr.cpp,
// src/initial_data/initial_data.cpp) by oracle/Makefile into oracle/_ref/.
//
// Every block is described by n[3], ng, lo[3], hi[3]: the reference GridBlock derives
// dx = (hi - lo) / n itself (src/core/grid.cpp:22-24).
#include "granite/amr/amr.hpp"
#include "granite/core/grid.hpp"
#include "granite/core/types.hpp"
#include "granite/initial_data/initial_data.hpp"
#include "granite/spacetime/ccz4.hpp"

#include <cstring>
#include <memory>
#include <vector>

using namespace granite;

namespace {

std::unique_ptr<GridBlock> make_block(const int* n, int ng, const double* lo, const double* hi, int level = 0,
                                      int nvars = NUM_SPACETIME_VARS) {
    return std::make_unique<GridBlock>(0, level, std::array<int, 3>{n[0], n[1], n[2]},
                                       std::array<Real, 3>{lo[0], lo[1], lo[2]},
                                       std::array<Real, 3>{hi[0], hi[1], hi[2]}, ng, nvars);
}

size_t total(const GridBlock& b) { return b.totalSize() * (size_t)b.getNumVars(); }

spacetime::CCZ4Params make_params(const double* p) {
    spacetime::CCZ4Params P;
    P.kappa1 = p[0];
    P.kappa2 = p[1];
    P.eta = p[2];
    P.ko_sigma = p[3];
    P.chi_blend_center = p[4];
    P.chi_blend_width = p[5];
    return P;
}

} // namespace

extern "C" {

// rhs is read first (ghosts keep their input values, as GridBlock storage would),
// then overwritten by computeRHS on the interior. Null sources = computeRHSVacuum.
void ref_rhs(const int* n, int ng, const double* lo, const double* hi, const double* prm, const double* u,
             double* rhs, const double* rho, const double* Si, const double* Sij, const double* S) {
    auto g = make_block(n, ng, lo, hi);
    auto r = make_block(n, ng, lo, hi);
    std::memcpy(g->flatBuffer(), u, total(*g) * sizeof(double));
    std::memcpy(r->flatBuffer(), rhs, total(*r) * sizeof(double));
    spacetime::CCZ4Evolution ev(make_params(prm));
    if (!rho) {
        ev.computeRHSVacuum(*g, *r);
    } else {
        const size_t N = g->totalSize();
        std::vector<Real> vr(rho, rho + N), vs(S, S + N);
        std::vector<std::array<Real, 3>> vsi(N);
        std::vector<std::array<Real, 6>> vsij(N);
        for (size_t p = 0; p < N; ++p) {
            for (int a = 0; a < 3; ++a) vsi[p][a] = Si[3 * p + a];
            for (int a = 0; a < 6; ++a) vsij[p][a] = Sij[6 * p + a];
        }
        ev.computeRHS(*g, *r, vr, vsi, vsij, vs);
    }
    std::memcpy(rhs, r->flatBuffer(), total(*r) * sizeof(double));
}

void ref_ko(const int* n, int ng, const double* lo, const double* hi, const double* prm, const double* u,
            double* rhs) {
    auto g = make_block(n, ng, lo, hi);
    auto r = make_block(n, ng, lo, hi);
    std::memcpy(g->flatBuffer(), u, total(*g) * sizeof(double));
    std::memcpy(r->flatBuffer(), rhs, total(*r) * sizeof(double));
    spacetime::CCZ4Evolution ev(make_params(prm));
    ev.applyDissipation(*g, *r);
    std::memcpy(rhs, r->flatBuffer(), total(*r) * sizeof(double));
}

void ref_constraints(const int* n, int ng, const double* lo, const double* hi, const double* u, double* ham,
                     double* mom) {
    auto g = make_block(n, ng, lo, hi);
    std::memcpy(g->flatBuffer(), u, total(*g) * sizeof(double));
    spacetime::CCZ4Evolution ev(spacetime::CCZ4Params{});
    std::vector<Real> h;
    std::vector<std::array<Real, 3>> m;
    ev.computeConstraints(*g, h, m);
    for (size_t p = 0; p < h.size(); ++p) {
        ham[p] = h[p];
        for (int a = 0; a < 3; ++a) mom[3 * p + a] = m[p][a];
    }
}

void ref_init_bl(const int* n, int ng, const double* lo, const double* hi, int nbh, const double* mass,
                 const double* pos, double* u) {
    auto g = make_block(n, ng, lo, hi);
    std::vector<initial_data::BlackHoleParams> bhs(nbh);
    for (int b = 0; b < nbh; ++b) {
        bhs[b].mass = mass[b];
        bhs[b].position = {pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]};
        bhs[b].momentum = {0, 0, 0};
        bhs[b].spin = {0, 0, 0};
    }
    initial_data::BrillLindquist bl(bhs);
    bl.apply(*g);
    std::memcpy(u, g->flatBuffer(), total(*g) * sizeof(double));
}

double ref_bl_psi(int nbh, const double* mass, const double* pos, double x, double y, double z) {
    std::vector<initial_data::BlackHoleParams> bhs(nbh);
    for (int b = 0; b < nbh; ++b) {
        bhs[b].mass = mass[b];
        bhs[b].position = {pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]};
        bhs[b].momentum = {0, 0, 0};
        bhs[b].spin = {0, 0, 0};
    }
    return initial_data::BrillLindquist(bhs).conformalFactor(x, y, z);
}

void ref_init_gauge_wave(const int* n, int ng, const double* lo, const double* hi, double A, double L, double* u) {
    auto g = make_block(n, ng, lo, hi);
    spacetime::setGaugeWaveData(*g, A, L);
    std::memcpy(u, g->flatBuffer(), total(*g) * sizeof(double));
}

void ref_prolongate(const int* cn, int cng, const double* clo, const double* chi, const double* cu, const int* fn,
                    int fng, const double* flo, const double* fhi, double* fu, int nvars) {
    auto c = make_block(cn, cng, clo, chi, 0, nvars);
    auto f = make_block(fn, fng, flo, fhi, 1, nvars);
    std::memcpy(c->flatBuffer(), cu, total(*c) * sizeof(double));
    std::memcpy(f->flatBuffer(), fu, total(*f) * sizeof(double));
    amr::AMRHierarchy h(amr::AMRParams{}, SimulationParams{});
    h.prolongate(*c, *f);
    std::memcpy(fu, f->flatBuffer(), total(*f) * sizeof(double));
}

void ref_restrict(const int* fn, int fng, const double* flo, const double* fhi, const double* fu, const int* cn,
                  int cng, const double* clo, const double* chi, double* cu, int nvars) {
    auto f = make_block(fn, fng, flo, fhi, 1, nvars);
    auto c = make_block(cn, cng, clo, chi, 0, nvars);
    std::memcpy(f->flatBuffer(), fu, total(*f) * sizeof(double));
    std::memcpy(c->flatBuffer(), cu, total(*c) * sizeof(double));
    amr::AMRHierarchy h(amr::AMRParams{}, SimulationParams{});
    h.restrict_data(*f, *c);
    std::memcpy(cu, c->flatBuffer(), total(*c) * sizeof(double));
}

// ---- the reference's AMRHierarchy, driven as src/main.cpp drives it ---------------------

struct ref_hier {
    std::unique_ptr<amr::AMRHierarchy> h;
};

// AMRHierarchy over ncells on [dlo, dhi] with ng ghosts, max_levels levels and ratio 2,
// regridding never during subcycle (regrid_interval 0); one tracking sphere per center with
// the given radius and min_level, then initialize() with a tagger that tags nothing, so the
// levels come from the tracking spheres alone (src/amr/amr.cpp:34-78, 80-301).
ref_hier* ref_hier_build(const int* ncells, const double* dlo, const double* dhi, int ng, int max_levels,
                         int nspheres, const double* centers, const double* radii, const int* min_levels) {
    amr::AMRParams ap;
    ap.max_levels = max_levels;
    ap.refinement_ratio = 2;
    ap.regrid_interval = 0;
    SimulationParams sp;
    for (int d = 0; d < 3; ++d) {
        sp.ncells[d] = ncells[d];
        sp.domain_lo[d] = dlo[d];
        sp.domain_hi[d] = dhi[d];
    }
    sp.ghost_cells = ng;
    auto* r = new ref_hier;
    r->h = std::make_unique<amr::AMRHierarchy>(ap, sp);
    for (int s = 0; s < nspheres; ++s)
        r->h->addTrackingSphere({centers[3 * s], centers[3 * s + 1], centers[3 * s + 2]}, radii[s], min_levels[s]);
    r->h->initialize([](const GridBlock&, int, int, int) { return false; });
    return r;
}

void ref_hier_free(ref_hier* r) { delete r; }

int ref_hier_nblocks(ref_hier* r) { return (int)r->h->getAllBlocks().size(); }

// geometry of block i of getAllBlocks() (coarsest level first)
void ref_hier_block(ref_hier* r, int i, int* level, int* n, double* lo, double* hi, double* dx) {
    GridBlock* b = r->h->getAllBlocks()[i];
    *level = b->getLevel();
    for (int d = 0; d < 3; ++d) {
        n[d] = b->numCells()[d];
        lo[d] = b->lowerCorner()[d];
        hi[d] = b->upperCorner()[d];
        dx[d] = b->cellSpacing()[d];
    }
}

double* ref_hier_data(ref_hier* r, int i) { return r->h->getAllBlocks()[i]->flatBuffer(); }

void ref_hier_fill_ghosts(ref_hier* r, int level) { r->h->fillGhostZones(level); }

// per-block evolve callback: the level's SSP-RK3 step on one block, in place
typedef void (*ref_evolve_fn)(int level, const int* n, int ng, const double* lo, const double* dx, double* u,
                              double dt);

// one coarse step as src/main.cpp:1184-1187 takes it: setLevelDt(0, dt0), subcycle(0)
void ref_hier_step(ref_hier* r, double dt0, ref_evolve_fn evolve) {
    r->h->setLevelDt(0, dt0);
    r->h->subcycle(
        0,
        [evolve](std::vector<GridBlock*>& blocks, Real dt) {
            for (GridBlock* b : blocks) {
                const int n[3] = {b->numCells()[0], b->numCells()[1], b->numCells()[2]};
                evolve(b->getLevel(), n, b->getNumGhost(), b->lowerCorner().data(), b->cellSpacing().data(),
                       b->flatBuffer(), dt);
            }
        },
        [](const GridBlock&, int, int, int) { return false; });
}

} // extern "C"
