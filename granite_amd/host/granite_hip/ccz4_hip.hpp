// C++17 host mirror of GRANITE's CCZ4Evolution over the MI355X engine's C ABI.
//
// Header-only. A GRANITE build includes it next to include/granite/spacetime/ccz4.hpp and
// links libgranite_ccz4.so; the classes keep the reference's method names, arguments and
// effects (include/granite/spacetime/ccz4.hpp:49-111) and take the reference's own types:
//   Block  - anything with GridBlock's API (include/granite/core/grid.hpp:44-162):
//            getNumVars, getNumGhost, numCells, lowerCorner, cellSpacing, totalSize,
//            flatBuffer
//   Params - anything with CCZ4Params' members (include/granite/spacetime/ccz4.hpp:25-34)
// Failures of the C ABI are thrown as std::runtime_error, as the reference reports its only
// error condition (src/core/grid.cpp:133-139). Nothing here computes: every call is a
// device call through include/granite_ccz4.h.
#pragma once

#include "granite_ccz4.h"

#include <array>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace granite_hip {

class Engine {
public:
    explicit Engine(int device = 0) {
        const int rc = gr_engine_create(device, &e_);
        if (rc != GR_OK) throw std::runtime_error("granite_hip: no usable gfx950 device (status " + std::to_string(rc) + ")");
    }
    ~Engine() {
        if (e_) gr_engine_destroy(e_);
    }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    gr_engine* get() const { return e_; }
    void check(int rc) const {
        if (rc != GR_OK) throw std::runtime_error("granite_hip: " + std::string(gr_last_error(e_)) + " (status " + std::to_string(rc) + ")");
    }

private:
    gr_engine* e_ = nullptr;
};

template <class Block>
gr_geom geom_of(const Block& b) {
    gr_geom g;
    const auto& n = b.numCells();
    const auto& lo = b.lowerCorner();
    const auto& dx = b.cellSpacing();
    for (int d = 0; d < 3; ++d) {
        g.n[d] = n[d];
        g.lo[d] = lo[d];
        g.dx[d] = dx[d];
    }
    g.ng = b.getNumGhost();
    return g;
}

template <class Params>
gr_ccz4_params params_of(const Params& p) {
    return gr_ccz4_params{p.kappa1, p.kappa2, p.eta, p.ko_sigma, p.chi_blend_center, p.chi_blend_width};
}

// Drop-in for spacetime::CCZ4Evolution (compatibility mode: host blocks in, host blocks out).
template <class Params>
class CCZ4Evolution {
public:
    CCZ4Evolution(const Params& params, Engine& engine) : params_(params), eng_(&engine) {}

    // src/spacetime/ccz4.cpp:251-865 (+ dissipation, as the reference calls it at the end)
    template <class Block>
    void computeRHS(const Block& grid, Block& rhs, const std::vector<double>& rho_matter,
                    const std::vector<std::array<double, 3>>& Si_matter,
                    const std::vector<std::array<double, 6>>& Sij_matter, const std::vector<double>& S_trace) const {
        const std::size_t N = grid.totalSize();
        if (rho_matter.size() < N || Si_matter.size() < N || Sij_matter.size() < N || S_trace.size() < N)
            throw std::runtime_error("granite_hip: matter source arrays smaller than the block");
        const gr_geom g = geom_of(grid);
        const gr_ccz4_params p = params_of(params_);
        eng_->check(gr_ccz4_rhs(eng_->get(), &g, &p, grid.getNumVars(), grid.flatBuffer(), rhs.flatBuffer(),
                                rho_matter.data(), Si_matter.data()->data(), Sij_matter.data()->data(),
                                S_trace.data()));
    }

    // src/spacetime/ccz4.cpp:229-245
    template <class Block>
    void computeRHSVacuum(const Block& grid, Block& rhs) const {
        const gr_geom g = geom_of(grid);
        const gr_ccz4_params p = params_of(params_);
        eng_->check(gr_ccz4_rhs(eng_->get(), &g, &p, grid.getNumVars(), grid.flatBuffer(), rhs.flatBuffer(), nullptr,
                                nullptr, nullptr, nullptr));
    }

    // src/spacetime/ccz4.cpp:941-1159
    template <class Block>
    void computeConstraints(const Block& grid, std::vector<double>& ham, std::vector<std::array<double, 3>>& mom) const {
        const std::size_t N = grid.totalSize();
        ham.assign(N, 0.0);
        mom.assign(N, {0.0, 0.0, 0.0});
        const gr_geom g = geom_of(grid);
        eng_->check(gr_ccz4_constraints(eng_->get(), &g, grid.flatBuffer(), ham.data(), mom.data()->data()));
    }

    // src/spacetime/ccz4.cpp:871-935
    template <class Block>
    void applyDissipation(const Block& grid, Block& rhs) const {
        const gr_geom g = geom_of(grid);
        const gr_ccz4_params p = params_of(params_);
        eng_->check(gr_ccz4_dissipation(eng_->get(), &g, &p, grid.getNumVars(), grid.flatBuffer(), rhs.flatBuffer()));
    }

    const Params& params() const { return params_; }

private:
    Params params_;
    Engine* eng_;
};

// Device-resident blocks of one refinement level: the body of the reference's per-level
// callback (src/main.cpp:1142-1177, a TimeIntegrator::sspRK3Step over the level's blocks)
// as one call, with the state kept in HBM between steps.
class Level {
public:
    template <class Block>
    Level(Engine& engine, const std::vector<Block*>& blocks) : eng_(&engine) {
        std::vector<gr_geom> g;
        for (const Block* b : blocks) g.push_back(geom_of(*b));
        eng_->check(gr_level_create(eng_->get(), (int)g.size(), g.data(), &l_));
    }
    ~Level() {
        if (l_) gr_level_destroy(l_);
    }
    Level(const Level&) = delete;
    Level& operator=(const Level&) = delete;

    template <class Block>
    void upload(int i, const Block& b) { eng_->check(gr_level_upload(l_, i, b.flatBuffer())); }
    template <class Block>
    void download(int i, Block& b) { eng_->check(gr_level_download(l_, i, b.flatBuffer())); }

    // one SSP-RK3 step with the reference's level-0 ghost semantics by default
    template <class Params>
    void sspRK3Step(const Params& prm, double dt, int precision = GR_PREC_FP64, int bc_mode = GR_BC_SOMMERFELD,
                    int pre_fill = GR_BC_OUTFLOW, bool use_graph = true) {
        const gr_ccz4_params p = params_of(prm);
        eng_->check(gr_level_rk3_step(l_, &p, dt, precision, bc_mode, pre_fill, use_graph ? 1 : 0));
    }
    void sync() { eng_->check(gr_level_sync(l_)); }
    // sqrt(sum H^2 / count), the reference's ||H||_2 telemetry (src/main.cpp:1216-1250)
    double hamiltonianL2() {
        double s = 0.0;
        long long n = 0;
        eng_->check(gr_level_constraints_l2(l_, &s, &n));
        return n > 0 ? std::sqrt(s / (double)n) : 0.0;
    }
    gr_level* get() const { return l_; }

private:
    Engine* eng_;
    gr_level* l_ = nullptr;
};

namespace detail {
inline std::vector<gr_level*> handles(const std::vector<Level*>& levels) {
    std::vector<gr_level*> h;
    for (const Level* l : levels) h.push_back(l->get());
    return h;
}
} // namespace detail

// Device-resident Berger-Oliger hierarchy of fixed geometry, coarsest level first: the
// operations of amr::AMRHierarchy (include/granite/amr/amr.hpp:49-117) on the device. The
// levels must outlive it; after a regrid, re-create levels and hierarchy.
class Hierarchy {
public:
    Hierarchy(Engine& engine, const std::vector<Level*>& levels) : eng_(&engine) {
        const std::vector<gr_level*> h = detail::handles(levels);
        eng_->check(gr_hierarchy_create(eng_->get(), (int)h.size(), h.data(), &h_));
    }
    ~Hierarchy() {
        if (h_) gr_hierarchy_destroy(h_);
    }
    Hierarchy(const Hierarchy&) = delete;
    Hierarchy& operator=(const Hierarchy&) = delete;

    // setLevelDt(0, dt0) + subcycle(0, sspRK3Step, tagger) without regrid (src/amr/amr.cpp:303-350)
    template <class Params>
    void step(const Params& prm, double dt0, int precision = GR_PREC_FP64, bool use_graph = true) {
        const gr_ccz4_params p = params_of(prm);
        eng_->check(gr_hierarchy_step(h_, &p, dt0, precision, use_graph ? 1 : 0));
    }
    // src/amr/amr.cpp:400-482
    void fillGhostZones(int level) { eng_->check(gr_hierarchy_fill_ghosts(h_, level)); }
    // parent level-1 into every cell of level (src/amr/amr.cpp:484-564), as regrid fills a new level
    void prolongate(int level) { eng_->check(gr_hierarchy_prolongate(h_, level)); }
    // level into the covered interior of level-1 (src/amr/amr.cpp:566-626)
    void restrict_data(int level) { eng_->check(gr_hierarchy_restrict(h_, level)); }
    // ||H||_2 over the interiors of every block of every level (src/main.cpp:1216-1250)
    double hamiltonianL2() {
        double s = 0.0;
        long long n = 0;
        eng_->check(gr_hierarchy_constraints_l2(h_, &s, nullptr, &n));
        return n > 0 ? std::sqrt(s / (double)n) : 0.0;
    }

private:
    Engine* eng_;
    gr_hierarchy* h_ = nullptr;
};

// Checkpoint staging for HDF5Writer::writeCheckpoint (src/io/hdf5_io.cpp:166-232): begin()
// snapshots every block of the levels (coarsest first) and copies it to pinned host memory
// while the evolution goes on; after wait(), copyTo() fills a host block (rawData layout).
class Checkpoint {
public:
    Checkpoint(Engine& engine, const std::vector<Level*>& levels) : eng_(&engine) {
        const std::vector<gr_level*> h = detail::handles(levels);
        eng_->check(gr_checkpoint_create(eng_->get(), (int)h.size(), h.data(), &c_));
    }
    ~Checkpoint() {
        if (c_) gr_checkpoint_destroy(c_);
    }
    Checkpoint(const Checkpoint&) = delete;
    Checkpoint& operator=(const Checkpoint&) = delete;

    void begin() { eng_->check(gr_checkpoint_begin(c_)); }
    bool done() {
        int d = 0;
        eng_->check(gr_checkpoint_query(c_, &d));
        return d != 0;
    }
    void wait() { eng_->check(gr_checkpoint_wait(c_)); }
    int numBlocks() {
        int n = 0;
        eng_->check(gr_checkpoint_nblocks(c_, &n));
        return n;
    }
    // block i of the last completed snapshot into a host block of the same geometry
    template <class Block>
    void copyTo(int i, Block& b) {
        const double* data = nullptr;
        long stride = 0;
        eng_->check(gr_checkpoint_block(c_, i, &data, &stride));
        if ((std::size_t)stride != b.totalSize()) throw std::runtime_error("granite_hip: checkpoint block size mismatch");
        std::memcpy(b.flatBuffer(), data, (std::size_t)b.getNumVars() * (std::size_t)stride * sizeof(double));
    }

private:
    Engine* eng_;
    gr_checkpoint* c_ = nullptr;
};

} // namespace granite_hip
