// granite_main: GRANITE's simulation driver (src/main.cpp:699-1262) for the vacuum CCZ4
// path, with the whole AMR hierarchy device-resident on the MI355X engine.
//
//     granite_main params.yaml [--precision fp64|mixed] [--dry-run]
//
// Reads the same parameter file keys as the reference driver (grid, time, ccz4, io,
// initial_data, black_holes; src/main.cpp:730-830) and prints the same telemetry lines
// (banner, grid, dt, tracking spheres, AMR levels, "[CFL-GUARD] Advection CFL=..." after a
// level step whose advective CFL exceeds 0.95, "[NaN@step=N] ST var=v (i,j,k) = val" or
// "[NaN@step=N] all finite" for the first 20 steps, "step= t= α_center= ||H||₂= [Blocks: n]"
// every output_interval steps), which is what scripts/run_granite.py and
// python/granite_analysis/core/parsers.py:24-49 consume. The NaN scan and the CFL guard run
// on the device (gr_level_first_nonfinite, the hierarchy's CFL monitor).
//
// Differences from the reference driver, all outside the CCZ4 + RK3 hot path:
//   * vacuum only: no GRMHD (matter sources are zero; the reference's atmosphere density is
//     1e-12 and enters the sources at that level);
//   * initial data: brill_lindquist, gauge_wave (flat metric, lapse 1 + A sin(2 pi x / L),
//     src/main.cpp:1042-1064) and flat; two_punctures and bowen_york need the reference's
//     elliptic solvers and are rejected;
//   * every tracking-sphere level exists from the first step (the reference creates level 1
//     before the loop and the deeper ones at its first regrids; every fine level is
//     re-prolongated from its parent at each substep either way) and the hierarchy is not
//     regridded (the punctures of Brill-Lindquist data do not move);
//   * no HDF5: a checkpoint (step % checkpoint_interval == 0 or t >= t_final, as
//     src/main.cpp:1256-1261) holds what HDF5Writer::writeCheckpoint stores
//     (src/io/hdf5_io.cpp:166-232) in a flat little-endian file <output_dir>/checkpoint_<step>.gck:
//       "GRCKPT01", int32 step, int32 num_blocks, float64 time, then per block (coarsest level
//       first) int32 level, ng, n[3], nvars, float64 lo[3], dx[3], and nvars * stride float64
//       (var v = GridBlock::rawData(v), ghosts included). The state is staged through
//       gr_checkpoint (device snapshot, then a copy stream to pinned memory), so the file is
//       written while the evolution goes on;
//   * the CFL-guard lines of a coarse step are printed once the step has run on the device
//     (the reference prints each inside its level step; the order of the lines is the same).
// --dry-run prints the geometry and exits without touching a GPU; --no-checkpoints skips
// the checkpoint files.
#include "granite_ccz4.h"
#include "granite_hip/params_yaml.hpp"
#include "granite_hip/tracking.hpp"

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

using granite_hip::BoxGeom;
using granite_hip::LevelGeoms;
using granite_hip::TrackingSphere;
using granite_hip::YNode;

namespace {

struct BlackHole {
    double mass = 1.0;
    std::array<double, 3> position{0.0, 0.0, 0.0};
};

// SimulationParams defaults (include/granite/core/types.hpp:181-215)
struct Sim {
    std::array<int, 3> ncells{64, 64, 64};
    std::array<double, 3> lo{-100.0, -100.0, -100.0}, hi{100.0, 100.0, 100.0};
    int ng = 4;
    double cfl = 0.25, t_final = 1000.0;
    int max_steps = 1000000;
    double kappa1 = 0.02, kappa2 = 0.0, eta = 2.0, ko_sigma = 0.1;
    int output_interval = 10, checkpoint_interval = 1000;
    std::string output_dir = "output";
    std::string id_type = "flat";
    double amplitude = 0.1, wavelength = 1.0;
    std::vector<BlackHole> bhs;
};

template <class T, class F>
void read3(const YNode& n, std::array<T, 3>& out, F conv) {
    if (!n) return;
    if (n.kind != YNode::Seq || n.size() != 3) throw std::runtime_error("expected a list of 3 values");
    for (int d = 0; d < 3; ++d) out[d] = conv(n[d]);
}

Sim load(const std::string& path) {
    Sim p;
    const YNode cfg = granite_hip::load_yaml_file(path);
    auto dbl = [](const YNode& n) { return n.as_double(); };
    auto itg = [](const YNode& n) { return n.as_int(); };
    if (const YNode& g = cfg["grid"]) {
        read3(g["ncells"], p.ncells, itg);
        read3(g["domain_lo"], p.lo, dbl);
        read3(g["domain_hi"], p.hi, dbl);
        if (g["ghost_cells"]) p.ng = g["ghost_cells"].as_int();
    }
    if (const YNode& t = cfg["time"]) {
        if (t["cfl"]) p.cfl = t["cfl"].as_double();
        if (t["t_final"]) p.t_final = t["t_final"].as_double();
        if (t["max_steps"]) p.max_steps = t["max_steps"].as_int();
    }
    if (const YNode& c = cfg["ccz4"]) {
        if (c["kappa1"]) p.kappa1 = c["kappa1"].as_double();
        if (c["kappa2"]) p.kappa2 = c["kappa2"].as_double();
        if (c["eta"]) p.eta = c["eta"].as_double();
        if (c["ko_sigma"]) p.ko_sigma = c["ko_sigma"].as_double();
    }
    if (const YNode& io = cfg["io"]) {
        if (io["output_dir"]) p.output_dir = io["output_dir"].as_string();
        if (io["output_interval"]) p.output_interval = io["output_interval"].as_int();
        if (io["checkpoint_interval"]) p.checkpoint_interval = io["checkpoint_interval"].as_int();
    }
    if (const YNode& id = cfg["initial_data"]) {
        if (id["type"]) p.id_type = id["type"].as_string();
        if (id["amplitude"]) p.amplitude = id["amplitude"].as_double();
        if (id["wavelength"]) p.wavelength = id["wavelength"].as_double();
    }
    const YNode& bh = cfg["black_holes"];
    if (bh && bh.kind == YNode::Seq)
        for (size_t i = 0; i < bh.size(); ++i) {
            BlackHole b;
            if (bh[i]["mass"]) b.mass = bh[i]["mass"].as_double();
            read3(bh[i]["position"], b.position, dbl);
            p.bhs.push_back(b);
        }
    return p;
}

struct Check {
    gr_engine* e;
    void operator()(int rc) const {
        if (rc != GR_OK) throw std::runtime_error(std::string("engine: ") + (e ? gr_last_error(e) : "no engine"));
    }
};

gr_geom geom_of(const BoxGeom& b, int ng) {
    gr_geom g;
    const auto dx = b.dx();
    for (int d = 0; d < 3; ++d) {
        g.n[d] = b.n[d];
        g.lo[d] = b.lo[d];
        g.dx[d] = dx[d];
    }
    g.ng = ng;
    return g;
}

// the flat state of setFlatSpacetime (src/spacetime/ccz4.cpp:1165-1188) and, for the gauge
// wave, the lapse 1 + A sin(2 pi x / L) of src/main.cpp:1042-1064, on every cell
std::vector<double> host_initial(const gr_geom& g, const Sim& p) {
    const long nt[3] = {g.n[0] + 2L * g.ng, g.n[1] + 2L * g.ng, g.n[2] + 2L * g.ng};
    const long stride = nt[0] * nt[1] * nt[2];
    std::vector<double> u(22 * stride, 0.0);
    for (long c = 0; c < stride; ++c) {
        u[0 * stride + c] = 1.0;
        u[1 * stride + c] = 1.0;
        u[4 * stride + c] = 1.0;
        u[6 * stride + c] = 1.0;
        u[18 * stride + c] = 1.0;
    }
    if (p.id_type == "gauge_wave")
        for (long k = 0; k < nt[2]; ++k)
            for (long j = 0; j < nt[1]; ++j)
                for (long i = 0; i < nt[0]; ++i) {
                    const double x = g.lo[0] + ((double)i - g.ng + 0.5) * g.dx[0];
                    u[18 * stride + i + nt[0] * (j + nt[1] * k)] =
                        1.0 + p.amplitude * std::sin(2.0 * M_PI * x / p.wavelength);
                }
    return u;
}

struct CheckpointFile {
    std::string path;
    int step = 0;
    double time = 0.0;
};

// the file layout described at the top of this file, from the host views of gr_checkpoint
void write_checkpoint(gr_checkpoint* c, const CheckpointFile& f, const LevelGeoms& levels, int ng) {
    std::ofstream out(f.path, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("cannot create " + f.path);
    auto put = [&](const void* p, size_t n) { out.write(static_cast<const char*>(p), (std::streamsize)n); };
    int nblocks = 0;
    if (gr_checkpoint_nblocks(c, &nblocks) != GR_OK) throw std::runtime_error("checkpoint: block count");
    put("GRCKPT01", 8);
    const int32_t head[2] = {f.step, nblocks};
    put(head, sizeof(head));
    put(&f.time, sizeof(double));
    int index = 0;
    for (size_t l = 0; l < levels.size(); ++l)
        for (const BoxGeom& g : levels[l]) {
            const double* data = nullptr;
            long stride = 0;
            if (gr_checkpoint_block(c, index++, &data, &stride) != GR_OK) throw std::runtime_error("checkpoint: block");
            const auto dx = g.dx();
            const int32_t meta[6] = {(int32_t)l, ng, g.n[0], g.n[1], g.n[2], 22};
            put(meta, sizeof(meta));
            put(g.lo.data(), 3 * sizeof(double));
            put(dx.data(), 3 * sizeof(double));
            put(data, 22 * (size_t)stride * sizeof(double));
        }
    if (!out) throw std::runtime_error("write failed: " + f.path);
}

} // namespace

int main(int argc, char** argv) {
    std::cout << std::unitbuf;
    std::string param_file, precision = "fp64";
    bool dry = false, checkpoints = true;
    for (int a = 1; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "--dry-run") dry = true;
        else if (s == "--no-checkpoints") checkpoints = false;
        else if (s == "--precision" && a + 1 < argc) precision = argv[++a];
        else param_file = s;
    }
    std::cout << "================================================================\n";
    std::cout << " GRANITE v0.6.7 (vacuum CCZ4 on the MI355X engine, " << gr_version() << ")\n";
    std::cout << "================================================================\n\n";
    Sim p;
    try {
        if (!param_file.empty()) {
            std::cout << "Loading parameter file: " << param_file << "\n";
            p = load(param_file);
        }
    } catch (const std::exception& ex) {
        std::cerr << "YAML parsing error: " << ex.what() << "\n";
        return 1;
    }
    const bool punctures = p.id_type == "brill_lindquist";
    if (p.id_type == "two_punctures" || p.id_type == "bowen_york") {
        std::cerr << "initial_data type '" << p.id_type
                  << "' needs the reference's elliptic solver; this driver supports brill_lindquist, gauge_wave, flat\n";
        return 2;
    }
    if (!punctures && p.id_type != "gauge_wave" && p.id_type != "flat")
        std::cerr << "Warning: unknown initial_data type '" << p.id_type << "', falling back to 'flat'.\n";
    if (punctures && p.bhs.empty()) p.bhs.push_back(BlackHole{});
    if (p.ng < 4) {
        std::cerr << "ghost_cells must be >= 4 (the upwind stencil reaches 4 cells)\n";
        return 2;
    }
    // the reference's grid phase shift (src/main.cpp:835-841)
    for (int d = 0; d < 3; ++d) {
        p.lo[d] += 1.3621415e-6;
        p.hi[d] += 1.3621415e-6;
    }
    // levels: the domain block, then the tracking spheres of src/main.cpp:1119-1126 (the
    // reference's default max_levels is 15, so min_level = 3)
    std::vector<TrackingSphere> spheres;
    if (punctures)
        for (const auto& b : p.bhs) spheres.push_back({b.position, std::max(2.0 * b.mass, 0.5), 3});
    const LevelGeoms levels = granite_hip::tracking_levels(p.ncells, p.lo, p.hi, spheres, 15);
    std::cout << "AMR: Level 0  (" << p.ncells[0] << "x" << p.ncells[1] << "x" << p.ncells[2] << ")  domain ["
              << p.lo[0] << ", " << p.hi[0] << "]\n";
    if (punctures) {
        double m = 0.0;
        for (const auto& b : p.bhs) m += b.mass;
        std::cout << "  Brill-Lindquist: " << p.bhs.size() << " BH(s), ADM mass = " << m << " M_sun\n";
    } else if (p.id_type == "gauge_wave") {
        std::cout << "  Gauge wave: A=" << p.amplitude << ", L=" << p.wavelength << "\n";
    } else {
        std::cout << "  Flat Minkowski spacetime\n";
    }
    const BoxGeom& base = levels[0][0];
    const auto bdx = base.dx();
    const double dx_min = std::min({bdx[0], bdx[1], bdx[2]});
    double dt = p.cfl * dx_min;
    std::cout << "Grid: " << p.ncells[0] << " x " << p.ncells[1] << " x " << p.ncells[2] << "\n";
    std::cout << "dx = " << dx_min << ", dt = " << dt << "\n";
    std::cout << "t_final = " << p.t_final << "\n\n";
    for (const auto& s : spheres)
        std::cout << "  [AMR] Tracking sphere @ (" << s.center[0] << "," << s.center[1] << "," << s.center[2]
                  << ") R=" << s.radius << " min_level=" << s.min_level << "\n";
    int nblocks = 0;
    for (size_t l = 0; l < levels.size(); ++l) {
        nblocks += (int)levels[l].size();
        if (l == 0) continue;
        for (size_t b = 0; b < levels[l].size(); ++b) {
            const BoxGeom& g = levels[l][b];
            std::cout << "AMR: Level " << l << "  block_id=" << 1000 * l + b << "  ncells=(" << g.n[0] << "x" << g.n[1]
                      << "x" << g.n[2] << ")  lo=(" << g.lo[0] << "," << g.lo[1] << "," << g.lo[2] << ")  hi=("
                      << g.hi[0] << "," << g.hi[1] << "," << g.hi[2] << ")  dx=" << g.dx()[0]
                      << (levels[l].size() < spheres.size() ? "  [MERGED]" : "") << "\n";
        }
    }
    if (dry) return 0;

    gr_engine* eng = nullptr;
    if (gr_engine_create(0, &eng) != GR_OK) {
        std::cerr << "no gfx950 device\n";
        return 3;
    }
    const Check ok{eng};
    const int prec = precision == "mixed" ? GR_PREC_MIXED : GR_PREC_FP64;
    std::vector<gr_level*> lv;
    gr_hierarchy* H = nullptr;
    gr_checkpoint* ck = nullptr;
    try {
        for (const auto& level : levels) {
            std::vector<gr_geom> g;
            for (const auto& b : level) g.push_back(geom_of(b, p.ng));
            gr_level* L = nullptr;
            ok(gr_level_create(eng, (int)g.size(), g.data(), &L));
            lv.push_back(L);
        }
        ok(gr_hierarchy_create(eng, (int)lv.size(), lv.data(), &H));
        ok(gr_hierarchy_set_cfl_monitor(H, 1));
        std::vector<double> cfl(1u << lv.size());
        std::vector<int> cfl_level(cfl.size());
        // initial data on level 0 (src/main.cpp:962-1071), finer levels prolongated from
        // their parents as regrid does (src/amr/amr.cpp:290-293), then fillGhostZones(0)
        if (punctures) {
            std::vector<double> m, pos;
            for (const auto& b : p.bhs) {
                m.push_back(b.mass);
                pos.insert(pos.end(), b.position.begin(), b.position.end());
            }
            ok(gr_level_init_brill_lindquist(lv[0], (int)m.size(), m.data(), pos.data()));
        } else {
            for (size_t b = 0; b < levels[0].size(); ++b) {
                const std::vector<double> u = host_initial(geom_of(levels[0][b], p.ng), p);
                ok(gr_level_upload(lv[0], (int)b, u.data()));
            }
        }
        for (size_t l = 1; l < lv.size(); ++l) ok(gr_hierarchy_prolongate(H, (int)l));
        ok(gr_hierarchy_fill_ghosts(H, 0));

        if (checkpoints) {
            std::filesystem::create_directories(p.output_dir); // as HDF5Writer's constructor
            ok(gr_checkpoint_create(eng, (int)lv.size(), lv.data(), &ck));
        }
        bool ck_pending = false;
        CheckpointFile ck_file;
        auto flush_checkpoint = [&](bool block) {
            if (!ck_pending) return;
            int done = 0;
            if (block) ok(gr_checkpoint_wait(ck));
            else ok(gr_checkpoint_query(ck, &done));
            if (!block && !done) return;
            write_checkpoint(ck, ck_file, levels, p.ng);
            ck_pending = false;
        };

        const gr_ccz4_params prm{p.kappa1, p.kappa2, p.eta, p.ko_sigma, 0.05, 0.02};
        const int ic = p.ncells[0] / 2 + p.ng; // the reference's centre probe (src/main.cpp:1238-1244)
        double t = 0.0;
        int step = 0;
        while (t < p.t_final && step < p.max_steps) {
            if (t + dt > p.t_final) dt = p.t_final - t;
            ok(gr_hierarchy_step(H, &prm, dt, prec, 1));
            // the advective CFL guard of every level step of this coarse step
            // (src/main.cpp:1155-1176), in subcycle order
            int ncfl = 0;
            ok(gr_hierarchy_cfl(H, cfl.data(), cfl_level.data(), (int)cfl.size(), &ncfl));
            for (int q = 0; q < ncfl; ++q)
                if (cfl[q] > 0.95) std::cout << "  [CFL-GUARD] Advection CFL=" << cfl[q] << " > 0.95 at sub-step!\n";
            t += dt;
            ++step;
            // per-step NaN scan of the first 20 steps (src/main.cpp:1192-1214): the first
            // non-finite interior value, blocks coarsest level first, then var, k, j, i
            if (step <= 20) {
                bool found_nan = false;
                for (gr_level* L : lv) {
                    int found = 0, blk = 0, var = 0, ci = 0, cj = 0, ck = 0;
                    double val = 0.0;
                    ok(gr_level_first_nonfinite(L, &found, &blk, &var, &ci, &cj, &ck, &val));
                    if (!found) continue;
                    std::ostringstream msg;
                    msg << "  [NaN@step=" << step << "] ST var=" << var << " (" << ci << "," << cj << "," << ck << ")"
                        << " = " << val;
                    std::cout << msg.str() << std::endl;
                    found_nan = true;
                    break;
                }
                if (!found_nan) std::cout << "  [NaN@step=" << step << "] all finite\n";
            }
            if (step % p.output_interval == 0) {
                double h2 = 0.0;
                long long count = 0;
                ok(gr_hierarchy_constraints_l2(H, &h2, nullptr, &count));
                double alpha_c = 1.0;
                if (ic < base.n[1] + 2 * p.ng && ic < base.n[2] + 2 * p.ng) // ncells[0] is used on all axes
                    ok(gr_level_read_cell(lv[0], 0, 18, ic, ic, ic, &alpha_c));
                const double hl2 = count > 0 ? std::sqrt(h2 / (double)count) : 0.0;
                std::cout << "  step=" << step << "  t=" << t << "  α_center=" << alpha_c << "  ||H||₂=" << hl2
                          << "  [Blocks: " << nblocks << "]\n";
            }
            flush_checkpoint(false);
            if (ck && (step % p.checkpoint_interval == 0 || t >= p.t_final)) {
                flush_checkpoint(true); // the previous file first: begin reuses the host buffers
                char name[32];
                std::snprintf(name, sizeof(name), "checkpoint_%06d.gck", step);
                ck_file = CheckpointFile{p.output_dir + "/" + name, step, t};
                std::cout << "Writing checkpoint: " << ck_file.path << " (t=" << t << ")\n";
                ok(gr_checkpoint_begin(ck));
                ck_pending = true;
            }
        }
        flush_checkpoint(true);
        ok(gr_engine_sync(eng));
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << "\n";
        if (ck) gr_checkpoint_destroy(ck);
        if (H) gr_hierarchy_destroy(H);
        for (gr_level* L : lv) gr_level_destroy(L);
        gr_engine_destroy(eng);
        return 4;
    }
    std::cout << "Evolution complete.\n";
    if (ck) gr_checkpoint_destroy(ck);
    gr_hierarchy_destroy(H);
    for (gr_level* L : lv) gr_level_destroy(L);
    gr_engine_destroy(eng);
    return 0;
}
