// C ABI of the CCZ4 RHS + SSP-RK3 engine (declared in include/granite_ccz4.h).
//
// Host pointers are borrowed for the duration of a call. Device memory is owned by the
// engine (compatibility-mode staging buffers) and by levels (state + two stage buffers per
// block). No exception crosses the boundary: every failure becomes a status code and a
// message readable through gr_last_error().
#include "../../include/granite_ccz4.h"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <tuple>

// run the FP64 tiles of the mixed form on a second stream, beside the FP32 kernel
#ifndef GR_MIXED_OVERLAP
#define GR_MIXED_OVERLAP 1
#endif
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

// Reduction partials of the allocation-free constraint monitor
// (gr_ccz4_constraints_l2_noalloc): a file-scope device array, one copy per GPU, held until
// the process ends. In host code `scratch` names a host-side stand-in, never device memory:
// its device address is resolved per device with hipGetSymbolAddress when an engine is
// created on that device (noalloc_scratch below) and only that address is passed to kernels.
// Every call on one GPU shares the array; calls that overlap on one GPU race.
constexpr int kNoAllocPartials = 2048;
static __device__ __align__(256) double scratch[kNoAllocPartials];

struct gr_engine {
    int device = -1;
    hipStream_t stream = nullptr;
    double* scratch = nullptr; // device address of `scratch` on this engine's device
    gr::AuxStream aux{nullptr, nullptr, nullptr}; // runs the FP64 tiles beside the FP32 ones
    std::string err;
    // multi-level step: a second stream for the ghost fills, events ordering each level's
    // stages and fills, and the captured graph of the last level set
    hipStream_t fill_stream = nullptr;
    std::vector<hipEvent_t> ls_events;
    hipGraph_t ls_graph = nullptr;
    hipGraphExec_t ls_exec = nullptr;
    std::vector<double> ls_key;
    // compatibility-mode staging buffers (grown on demand, doubles)
    double* d_u = nullptr;
    double* d_r = nullptr;
    size_t cap_state = 0;
    size_t cap_r = 0;
    double* d_src = nullptr;
    size_t cap_src = 0;
    double* d_aux = nullptr;
    size_t cap_aux = 0;
};

// Blocks of a level that share one grid spacing bit for bit: one stage launch (the stencil
// constants are launch arguments). Blocks whose dx differ (by an ulp, as separately built
// boxes of one level may) run in separate launches with their own constants, as the
// reference evaluates each block with its own dx.
struct gr_group {
    std::vector<int> blocks;         // level block indices, in level order
    std::vector<gr::BatchBlock> table;
    gr::BatchBlock* d_table = nullptr;
    int* d_tile_block = nullptr;
    int* d_ghost_wg_block = nullptr;
    unsigned char* flags = nullptr;  // mixed precision: per-tile FP64 flags of the group's tiles
    gr::Batch batch;
};

struct gr_level {
    gr_engine* eng = nullptr;
    std::vector<gr::BlockDesc> blk;
    std::vector<double*> u, s1, s2;
    std::vector<gr::Sources> src;   // per block (device pointers borrowed from the caller)
    bool has_src = false;
    double* ham = nullptr;     // monitor scratch (max stride)
    double* mom = nullptr;     // momentum-constraint scratch (3 * max stride, on first use)
    long maxstride = 0;
    double* partial = nullptr; // reduction partials
    double* red = nullptr;     // reduction result (1 double)
    double* cfl_part = nullptr; // CFL-guard partials and result (on first use)
    double* bh = nullptr;      // puncture masses/positions for initial data
    double chi_fp64_below = 0.2;
    unsigned generation = 0;   // bumped whenever captured launches of the level go stale
    // batched launches over all blocks (ghost fills, which use each block's own dx): device
    // block table, tile -> block map, ghost workgroup -> block map
    std::vector<gr::BatchBlock> table;
    gr::BatchBlock* d_table = nullptr;
    int* d_tile_block = nullptr;
    int* d_ghost_wg_block = nullptr;
    gr::Batch batch;
    // stage launches: one per group of blocks with equal dx
    std::vector<gr_group> groups;
    int partial_cap = 1024;
    // captured RK3 step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool have_graph = false;
    double gkey[11] = {0};
    // optional per-kernel event timing of eager steps: kind -> list of (start, stop)
    bool profile = false;
    std::vector<hipEvent_t> ev[3];
    size_t ev_used[3] = {0, 0, 0};
};

namespace {

constexpr const char* kVersion = "granite-ccz4-mi355x 0.1";

int fail(gr_engine* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}

int hipfail(gr_engine* e, hipError_t st, const char* what) {
    if (e) e->err = std::string(what) + ": " + hipGetErrorString(st);
    return (int)st;
}

#define GR_HIP(e, call)                                                                                            \
    do {                                                                                                           \
        hipError_t st_ = (call);                                                                                   \
        if (st_ != hipSuccess) return hipfail((e), st_, #call);                                                    \
    } while (0)

int check_geom(gr_engine* e, const gr_geom* g, gr::BlockDesc* out) {
    if (!g) return fail(e, GR_ERR_ARG, "null geometry");
    for (int d = 0; d < 3; ++d) {
        if (g->n[d] < 1) return fail(e, GR_ERR_GEOM, "n must be >= 1 on every axis");
        if (!(g->dx[d] > 0.0) || !std::isfinite(g->dx[d])) return fail(e, GR_ERR_GEOM, "dx must be positive");
    }
    // the one-sided advective stencil reaches 4 cells (src/spacetime/ccz4.cpp:107-152)
    if (g->ng < 4) return fail(e, GR_ERR_GEOM, "ng must be >= 4 (the upwind stencil reaches 4 cells)");
    const long nt0 = g->n[0] + 2L * g->ng, nt1 = g->n[1] + 2L * g->ng, nt2 = g->n[2] + 2L * g->ng;
    const long stride = nt0 * nt1 * nt2;
    // the stage kernels address a field with 32-bit byte offsets
    if (stride >= (1L << 29)) return fail(e, GR_ERR_GEOM, "block too large (over 2^29 cells with ghosts)");
    for (int d = 0; d < 3; ++d) {
        out->n[d] = g->n[d];
        out->nt[d] = g->n[d] + 2 * g->ng;
        out->lo[d] = g->lo[d];
        out->dx[d] = g->dx[d];
    }
    out->ng = g->ng;
    out->stride = stride;
    return GR_OK;
}

int check_params(gr_engine* e, const gr_ccz4_params* p) {
    if (!p) return fail(e, GR_ERR_ARG, "null parameters");
    if (!(p->chi_blend_width != 0.0)) return fail(e, GR_ERR_ARG, "chi_blend_width must be non-zero");
    return GR_OK;
}

gr::CellParams cell_params(const gr_ccz4_params* p) {
    return gr::CellParams{p->kappa1, p->kappa2, p->eta, p->chi_blend_center, p->chi_blend_width};
}

int grow(gr_engine* e, double** buf, size_t* cap, size_t n) {
    if (*cap >= n) return GR_OK;
    if (*buf) GR_HIP(e, hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    GR_HIP(e, hipMalloc((void**)buf, n * sizeof(double)));
    *cap = n;
    return GR_OK;
}

int use_device(gr_engine* e) {
    if (!e) return GR_ERR_ARG;
    GR_HIP(e, hipSetDevice(e->device));
    return GR_OK;
}

// Device address of `scratch` on `device` (which must be current), looked up once per device
// and kept for the life of the process.
hipError_t noalloc_scratch(int device, double** out) {
    static std::mutex mu;
    static std::vector<double*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess) return st;
        addr[device] = static_cast<double*>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

} // namespace

extern "C" {

const char* gr_version(void) { return kVersion; }

int gr_stage_tile(int precision, int* t3) {
    if (!t3) return GR_ERR_ARG;
    if (precision != GR_PREC_FP64 && precision != GR_PREC_MIXED) return GR_ERR_ARG;
    if (precision == GR_PREC_FP64) gr::stage_tile_fp64(t3);
    else gr::stage_tile_mixed(t3);
    return GR_OK;
}

int gr_engine_create(int device, gr_engine** out) {
    if (!out) return GR_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return GR_ERR_NODEV;
    if (device < 0 || device >= count) return GR_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return GR_ERR_NODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GR_ERR_NODEV;
    gr_engine* e = new gr_engine();
    e->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&e->stream) != hipSuccess) {
        delete e;
        return GR_ERR_NODEV;
    }
    if (noalloc_scratch(device, &e->scratch) != hipSuccess) {
        gr_engine_destroy(e);
        return GR_ERR_NODEV;
    }
#if GR_MIXED_OVERLAP
    if (hipStreamCreate(&e->aux.stream) != hipSuccess ||
        hipEventCreateWithFlags(&e->aux.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->aux.join, hipEventDisableTiming) != hipSuccess) {
        gr_engine_destroy(e);
        return GR_ERR_NODEV;
    }
#endif
    *out = e;
    return GR_OK;
}

int gr_engine_destroy(gr_engine* e) {
    if (!e) return GR_ERR_ARG;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->d_u) (void)hipFree(e->d_u);
    if (e->d_r) (void)hipFree(e->d_r);
    if (e->d_src) (void)hipFree(e->d_src);
    if (e->d_aux) (void)hipFree(e->d_aux);
    if (e->aux.stream) (void)hipStreamSynchronize(e->aux.stream);
    if (e->aux.fork) (void)hipEventDestroy(e->aux.fork);
    if (e->aux.join) (void)hipEventDestroy(e->aux.join);
    if (e->aux.stream) (void)hipStreamDestroy(e->aux.stream);
    if (e->ls_exec) (void)hipGraphExecDestroy(e->ls_exec);
    if (e->ls_graph) (void)hipGraphDestroy(e->ls_graph);
    for (hipEvent_t ev : e->ls_events) (void)hipEventDestroy(ev);
    if (e->fill_stream) {
        (void)hipStreamSynchronize(e->fill_stream);
        (void)hipStreamDestroy(e->fill_stream);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return GR_OK;
}

const char* gr_last_error(const gr_engine* e) { return e ? e->err.c_str() : "null engine"; }

int gr_engine_sync(gr_engine* e) {
    int rc = use_device(e);
    if (rc) return rc;
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_ccz4_rhs(gr_engine* e, const gr_geom* geom, const gr_ccz4_params* prm, int nvars, const double* u,
                double* rhs, const double* rho, const double* Si, const double* Sij, const double* S) {
    int rc = use_device(e);
    if (rc) return rc;
    gr::BlockDesc b;
    if ((rc = check_geom(e, geom, &b))) return rc;
    if ((rc = check_params(e, prm))) return rc;
    if (nvars != GR_NUM_VARS) return fail(e, GR_ERR_NVARS, "nvars must be 22");
    if (!u || !rhs) return fail(e, GR_ERR_ARG, "null state or rhs");
    const int nsrc = (rho != nullptr) + (Si != nullptr) + (Sij != nullptr) + (S != nullptr);
    if (nsrc != 0 && nsrc != 4) return fail(e, GR_ERR_ARG, "matter sources must be all null or all non-null");
    const size_t n = 22 * (size_t)b.stride;
    if ((rc = grow(e, &e->d_u, &e->cap_state, n))) return rc;
    if ((rc = grow(e, &e->d_r, &e->cap_r, n))) return rc;
    gr::Sources src{nullptr, nullptr, nullptr, nullptr};
    if (nsrc == 4) {
        if ((rc = grow(e, &e->d_src, &e->cap_src, 11 * (size_t)b.stride))) return rc;
        double* d = e->d_src;
        GR_HIP(e, hipMemcpyAsync(d, rho, b.stride * sizeof(double), hipMemcpyHostToDevice, e->stream));
        GR_HIP(e, hipMemcpyAsync(d + b.stride, Si, 3 * b.stride * sizeof(double), hipMemcpyHostToDevice, e->stream));
        GR_HIP(e, hipMemcpyAsync(d + 4 * b.stride, Sij, 6 * b.stride * sizeof(double), hipMemcpyHostToDevice, e->stream));
        GR_HIP(e, hipMemcpyAsync(d + 10 * b.stride, S, b.stride * sizeof(double), hipMemcpyHostToDevice, e->stream));
        src = gr::Sources{d, d + b.stride, d + 4 * b.stride, d + 10 * b.stride};
    }
    GR_HIP(e, hipMemcpyAsync(e->d_u, u, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    GR_HIP(e, hipMemcpyAsync(e->d_r, rhs, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    static const int sel[3] = {0, 1, 2};
    GR_HIP(e, gr::launch_stage_fp64(gr::single_batch(b, e->d_u, e->d_u, e->d_r, src), sel, src.rho != nullptr,
                                    cell_params(prm), prm->ko_sigma, gr::MODE_RHS, gr::StageCoef{0.0, 0.0, 0.0},
                                    e->stream));
    GR_HIP(e, hipMemcpyAsync(rhs, e->d_r, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_ccz4_rhs_device(gr_engine* e, const gr_geom* geom, const gr_ccz4_params* prm, const double* d_u,
                       double* d_rhs, const double* d_rho, const double* d_Si, const double* d_Sij,
                       const double* d_S) {
    int rc = use_device(e);
    if (rc) return rc;
    gr::BlockDesc b;
    if ((rc = check_geom(e, geom, &b))) return rc;
    if ((rc = check_params(e, prm))) return rc;
    if (!d_u || !d_rhs) return fail(e, GR_ERR_ARG, "null state or rhs");
    const int nsrc = (d_rho != nullptr) + (d_Si != nullptr) + (d_Sij != nullptr) + (d_S != nullptr);
    if (nsrc != 0 && nsrc != 4) return fail(e, GR_ERR_ARG, "matter sources must be all null or all non-null");
    gr::Sources src{d_rho, d_Si, d_Sij, d_S};
    static const int sel[3] = {0, 1, 2};
    GR_HIP(e, gr::launch_stage_fp64(gr::single_batch(b, d_u, d_u, d_rhs, src), sel, src.rho != nullptr,
                                    cell_params(prm), prm->ko_sigma, gr::MODE_RHS, gr::StageCoef{0.0, 0.0, 0.0},
                                    e->stream));
    return GR_OK;
}

int gr_ccz4_dissipation(gr_engine* e, const gr_geom* geom, const gr_ccz4_params* prm, int nvars, const double* u,
                        double* rhs) {
    int rc = use_device(e);
    if (rc) return rc;
    gr::BlockDesc b;
    if ((rc = check_geom(e, geom, &b))) return rc;
    if (!prm) return fail(e, GR_ERR_ARG, "null parameters");
    if (nvars != GR_NUM_VARS) return fail(e, GR_ERR_NVARS, "nvars must be 22");
    if (!u || !rhs) return fail(e, GR_ERR_ARG, "null state or rhs");
    if (prm->ko_sigma <= 0.0) return GR_OK; // src/spacetime/ccz4.cpp:873-874 (a NaN sigma dissipates)
    const size_t n = 22 * (size_t)b.stride;
    if ((rc = grow(e, &e->d_aux, &e->cap_aux, 2 * n))) return rc;
    double* du = e->d_aux;
    double* dr = e->d_aux + n;
    GR_HIP(e, hipMemcpyAsync(du, u, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    GR_HIP(e, hipMemcpyAsync(dr, rhs, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    GR_HIP(e, gr::launch_dissipation(b, du, dr, prm->ko_sigma, e->stream));
    GR_HIP(e, hipMemcpyAsync(rhs, dr, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_ccz4_constraints(gr_engine* e, const gr_geom* geom, const double* u, double* ham, double* mom) {
    int rc = use_device(e);
    if (rc) return rc;
    gr::BlockDesc b;
    if ((rc = check_geom(e, geom, &b))) return rc;
    if (!u || !ham || !mom) return fail(e, GR_ERR_ARG, "null state, ham or mom");
    const size_t n = 22 * (size_t)b.stride;
    if ((rc = grow(e, &e->d_aux, &e->cap_aux, n + 4 * (size_t)b.stride))) return rc;
    double* du = e->d_aux;
    double* dh = e->d_aux + n;
    double* dm = dh + b.stride;
    GR_HIP(e, hipMemcpyAsync(du, u, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    GR_HIP(e, gr::launch_constraints(b, du, dh, dm, e->stream));
    GR_HIP(e, hipMemcpyAsync(ham, dh, b.stride * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    GR_HIP(e, hipMemcpyAsync(mom, dm, 3 * b.stride * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

// ---------------------------------------------------------------------------------------
// levels
// ---------------------------------------------------------------------------------------

} // extern "C"

// ---------------------------------------------------------------------------------------
// Tuning table: the kernel variant chosen for each (choice, block-size class, precision).
// Choices: the FP64 stage kernel's staging variant (0 registers, 1 direct LDS), the ghost
// fill's field groups (1, 2, 4, 11) and the AMR transfers' field groups (1, 2, 4, 11). The
// class is the mean interior cells per block: < 24^3, < 48^3, < 96^3, larger. Levels and
// hierarchies read the table when they are created; gr_level_autotune / gr_hierarchy_autotune
// time the variants on their own blocks and write the winners; gr_tune_save / gr_tune_load
// keep the table in a text file between processes.
// ---------------------------------------------------------------------------------------
namespace {

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, int> g_tune;

int size_class(double cells) {
    if (cells < 24.0 * 24 * 24) return 0;
    if (cells < 48.0 * 48 * 48) return 1;
    if (cells < 96.0 * 96 * 96) return 2;
    return 3;
}

template <class V>
int blocks_class(const gr_level* L, const V& which) {
    double c = 0.0;
    for (int i : which) c += (double)L->blk[i].n[0] * L->blk[i].n[1] * L->blk[i].n[2];
    return size_class(which.empty() ? 0.0 : c / (double)which.size());
}

int level_class(const gr_level* L) {
    std::vector<int> all(L->blk.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
    return blocks_class(L, all);
}

bool tune_get(int kind, int cls, int prec, int* v) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune.find(std::make_tuple(kind, cls, prec));
    if (it == g_tune.end()) return false;
    *v = it->second;
    return true;
}

void tune_put(int kind, int cls, int prec, int v) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[std::make_tuple(kind, cls, prec)] = v;
}

bool tune_valid(int kind, int v) {
    if (kind == GR_TUNE_STAGE) return v >= 0 && v < gr::FP64_STAGE_VARIANTS;
    return gr::ghost_groups_valid(v); // GR_TUNE_GHOST, GR_TUNE_TRANSFER
}

} // namespace

// The batched launch description of blocks `which` of a level: host table (geometry, the
// three buffers, sources), its device copy, the tile -> block and ghost workgroup -> block
// maps; one block needs no table.
static hipError_t build_batch(gr_level* L, const std::vector<int>& which, std::vector<gr::BatchBlock>& table,
                              gr::BatchBlock** d_table, int** d_tile_block, int** d_ghost_wg_block, gr::Batch* B) {
    table.resize(which.size());
    for (size_t n = 0; n < which.size(); ++n) {
        const int i = which[n];
        table[n].b = L->blk[i];
        table[n].buf[0] = L->u[i];
        table[n].buf[1] = L->s1[i];
        table[n].buf[2] = L->s2[i];
        table[n].src = L->src[i];
    }
    std::vector<int> tile_block, ghost_wg_block;
    const int ntiles = gr::batch_layout(table, tile_block, ghost_wg_block);
    if (which.size() == 1) {
        *B = gr::single_batch(table[0].b, table[0].buf[0], table[0].buf[1], table[0].buf[2], table[0].src);
        return hipSuccess;
    }
    hipError_t st = hipMalloc((void**)d_table, table.size() * sizeof(gr::BatchBlock));
    if (st == hipSuccess) st = hipMalloc((void**)d_tile_block, ntiles * sizeof(int));
    if (st == hipSuccess) st = hipMalloc((void**)d_ghost_wg_block, ghost_wg_block.size() * sizeof(int));
    if (st == hipSuccess)
        st = hipMemcpy(*d_table, table.data(), table.size() * sizeof(gr::BatchBlock), hipMemcpyHostToDevice);
    if (st == hipSuccess) st = hipMemcpy(*d_tile_block, tile_block.data(), ntiles * sizeof(int), hipMemcpyHostToDevice);
    if (st == hipSuccess)
        st = hipMemcpy(*d_ghost_wg_block, ghost_wg_block.data(), ghost_wg_block.size() * sizeof(int),
                       hipMemcpyHostToDevice);
    B->table = *d_table;
    B->tile_block = *d_tile_block;
    B->ghost_wg_block = *d_ghost_wg_block;
    B->nblocks = (int)table.size();
    B->one = table[0];
    B->ntiles = ntiles;
    B->ghost_wgs = (int)ghost_wg_block.size();
    for (int d = 0; d < 3; ++d) B->dx[d] = table[0].b.dx[d];
    B->pairs_ok = 1;
    for (const gr::BatchBlock& T : table)
        for (int k = 0; k < 3; ++k) B->pairs_ok = B->pairs_ok && gr::pairs_aligned(T.b, T.buf[k]);
    return st;
}

extern "C" {

int gr_level_create(gr_engine* e, int nblocks, const gr_geom* geoms, gr_level** out) {
    int rc = use_device(e);
    if (rc) return rc;
    if (!out || !geoms || nblocks < 1) return fail(e, GR_ERR_ARG, "bad level arguments");
    *out = nullptr;
    gr_level* L = new gr_level();
    L->eng = e;
    long maxstride = 0;
    for (int i = 0; i < nblocks; ++i) {
        gr::BlockDesc b;
        if ((rc = check_geom(e, &geoms[i], &b))) {
            gr_level_destroy(L);
            return rc;
        }
        L->blk.push_back(b);
        if (b.stride > maxstride) maxstride = b.stride;
        double *u = nullptr, *a = nullptr, *c = nullptr;
        const size_t bytes = 22 * (size_t)b.stride * sizeof(double);
        hipError_t st = hipMalloc((void**)&u, bytes);
        if (st == hipSuccess) st = hipMalloc((void**)&a, bytes);
        if (st == hipSuccess) st = hipMalloc((void**)&c, bytes);
        L->u.push_back(u);
        L->s1.push_back(a);
        L->s2.push_back(c);
        if (st != hipSuccess) {
            gr_level_destroy(L);
            return hipfail(e, st, "hipMalloc(level block)");
        }
    }
    L->src.assign(nblocks, gr::Sources{nullptr, nullptr, nullptr, nullptr});
    // the whole level (ghost fills), then one group per distinct dx (stage launches)
    std::vector<int> all(nblocks);
    for (int i = 0; i < nblocks; ++i) all[i] = i;
    hipError_t st = build_batch(L, all, L->table, &L->d_table, &L->d_tile_block, &L->d_ghost_wg_block, &L->batch);
    for (int i = 0; st == hipSuccess && i < nblocks; ++i) {
        bool placed = false;
        for (gr_group& G : L->groups)
            if (std::memcmp(L->blk[G.blocks[0]].dx, L->blk[i].dx, sizeof(L->blk[i].dx)) == 0) {
                G.blocks.push_back(i);
                placed = true;
                break;
            }
        if (!placed) {
            L->groups.emplace_back();
            L->groups.back().blocks.push_back(i);
        }
    }
    for (gr_group& G : L->groups) {
        if (st != hipSuccess) break;
        st = build_batch(L, G.blocks, G.table, &G.d_table, &G.d_tile_block, &G.d_ghost_wg_block, &G.batch);
        if (st == hipSuccess) st = hipMalloc((void**)&G.flags, G.batch.ntiles);
        // without a table entry: the direct-to-LDS staging in 8x4x8 parts where the rows allow
        // it and the launch is large enough to fill the machine (at least 40^3 cells), else
        // the register staging (below that a stage launch is latency-bound: measured 10% faster
        // in variant 0 for single blocks up to 32^3, 22% slower from 48^3)
        double cells = 0.0;
        for (int i : G.blocks) cells += (double)L->blk[i].n[0] * L->blk[i].n[1] * L->blk[i].n[2];
        int v = (G.batch.pairs_ok && cells >= 40.0 * 40 * 40) ? 2 : 0;
        tune_get(GR_TUNE_STAGE, blocks_class(L, G.blocks), GR_PREC_FP64, &v);
        G.batch.variant = v;
    }
    {
        int g = 0;
        if (tune_get(GR_TUNE_GHOST, level_class(L), GR_PREC_FP64, &g)) L->batch.ghost_groups = g;
    }
    L->maxstride = maxstride;
    if (st == hipSuccess) st = hipMalloc((void**)&L->ham, maxstride * sizeof(double));
    if (st == hipSuccess) st = hipMalloc((void**)&L->partial, L->partial_cap * sizeof(double));
    if (st == hipSuccess) st = hipMalloc((void**)&L->red, sizeof(double));
    if (st != hipSuccess) {
        gr_level_destroy(L);
        return hipfail(e, st, "hipMalloc(level scratch)");
    }
    *out = L;
    return GR_OK;
}

int gr_level_destroy(gr_level* L) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    if (e) {
        (void)hipSetDevice(e->device);
        (void)hipStreamSynchronize(e->stream);
        if (e->fill_stream) (void)hipStreamSynchronize(e->fill_stream);
        // a captured level-set step may hold this level's buffers
        if (e->ls_exec) (void)hipGraphExecDestroy(e->ls_exec);
        if (e->ls_graph) (void)hipGraphDestroy(e->ls_graph);
        e->ls_exec = nullptr;
        e->ls_graph = nullptr;
        e->ls_key.clear();
    }
    if (L->exec) (void)hipGraphExecDestroy(L->exec);
    if (L->graph) (void)hipGraphDestroy(L->graph);
    for (auto& v : L->ev)
        for (hipEvent_t ev : v) (void)hipEventDestroy(ev);
    for (double* p : L->u)
        if (p) (void)hipFree(p);
    for (double* p : L->s1)
        if (p) (void)hipFree(p);
    for (double* p : L->s2)
        if (p) (void)hipFree(p);
    for (gr_group& G : L->groups) {
        if (G.flags) (void)hipFree(G.flags);
        if (G.d_table) (void)hipFree(G.d_table);
        if (G.d_tile_block) (void)hipFree(G.d_tile_block);
        if (G.d_ghost_wg_block) (void)hipFree(G.d_ghost_wg_block);
    }
    if (L->d_table) (void)hipFree(L->d_table);
    if (L->d_tile_block) (void)hipFree(L->d_tile_block);
    if (L->d_ghost_wg_block) (void)hipFree(L->d_ghost_wg_block);
    if (L->ham) (void)hipFree(L->ham);
    if (L->mom) (void)hipFree(L->mom);
    if (L->partial) (void)hipFree(L->partial);
    if (L->red) (void)hipFree(L->red);
    if (L->cfl_part) (void)hipFree(L->cfl_part);
    if (L->bh) (void)hipFree(L->bh);
    delete L;
    return GR_OK;
}

static int level_block(gr_level* L, int b) {
    if (!L) return GR_ERR_ARG;
    if (b < 0 || b >= (int)L->blk.size()) return fail(L->eng, GR_ERR_ARG, "block index out of range");
    return use_device(L->eng);
}

int gr_level_upload(gr_level* L, int b, const double* host_u) {
    int rc = level_block(L, b);
    if (rc) return rc;
    if (!host_u) return fail(L->eng, GR_ERR_ARG, "null host state");
    gr_engine* e = L->eng;
    GR_HIP(e, hipMemcpyAsync(L->u[b], host_u, 22 * L->blk[b].stride * sizeof(double), hipMemcpyHostToDevice,
                             e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_level_download(gr_level* L, int b, double* host_u) {
    int rc = level_block(L, b);
    if (rc) return rc;
    if (!host_u) return fail(L->eng, GR_ERR_ARG, "null host state");
    gr_engine* e = L->eng;
    GR_HIP(e, hipMemcpyAsync(host_u, L->u[b], 22 * L->blk[b].stride * sizeof(double), hipMemcpyDeviceToHost,
                             e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_level_state(gr_level* L, int b, double** d_u) {
    int rc = level_block(L, b);
    if (rc) return rc;
    if (!d_u) return fail(L->eng, GR_ERR_ARG, "null output");
    *d_u = L->u[b];
    return GR_OK;
}

int gr_level_init_brill_lindquist(gr_level* L, int nbh, const double* mass, const double* pos) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (nbh < 0 || nbh > 64 || (nbh > 0 && (!mass || !pos))) return fail(e, GR_ERR_ARG, "bad puncture list");
    if (!L->bh) GR_HIP(e, hipMalloc((void**)&L->bh, 4 * 64 * sizeof(double)));
    std::vector<double> h(4 * 64, 0.0);
    for (int i = 0; i < nbh; ++i) {
        h[i] = mass[i];
        for (int d = 0; d < 3; ++d) h[64 + 3 * i + d] = pos[3 * i + d];
    }
    GR_HIP(e, hipMemcpyAsync(L->bh, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    for (size_t b = 0; b < L->blk.size(); ++b)
        GR_HIP(e, gr::launch_init_bl(L->blk[b], L->u[b], nbh, L->bh, L->bh + 64, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_level_add_noise(gr_level* L, uint64_t seed, double amp, uint32_t mask) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    for (size_t b = 0; b < L->blk.size(); ++b)
        GR_HIP(e, gr::launch_add_noise(L->blk[b], L->u[b], seed, amp, mask, e->stream));
    return GR_OK;
}

int gr_level_fill_ghosts(gr_level* L, int mode) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (mode < 0 || mode > 2) return fail(e, GR_ERR_ARG, "ghost mode must be 0, 1 or 2");
    GR_HIP(e, gr::launch_fill_ghosts_batch(L->batch, 0, mode, e->stream));
    return GR_OK;
}

} // extern "C"

// records a start/stop event pair around `launch` when profiling is on
template <class F>
static hipError_t timed(gr_level* L, int kind, F&& launch) {
    if (!L->profile) return launch();
    auto& v = L->ev[kind];
    size_t& used = L->ev_used[kind];
    while (v.size() < used + 2) {
        hipEvent_t ev;
        hipError_t st = hipEventCreate(&ev);
        if (st != hipSuccess) return st;
        v.push_back(ev);
    }
    hipError_t st = hipEventRecord(v[used], L->eng->stream);
    if (st != hipSuccess) return st;
    if ((st = launch()) != hipSuccess) return st;
    if ((st = hipEventRecord(v[used + 1], L->eng->stream)) != hipSuccess) return st;
    used += 2;
    return hipSuccess;
}

// one stage over the blocks of batch B in the requested precision; sel = {input, u^n,
// output} buffer indices; flags: the mixed form's per-tile bytes (B.ntiles of them)
static hipError_t stage_batch(gr_level* L, const gr::Batch& B, unsigned char* flags, const int sel[3],
                              const gr_ccz4_params* prm, int mode, const gr::StageCoef& c, int precision) {
    const gr::CellParams P = cell_params(prm);
    if (precision == GR_PREC_MIXED)
        return gr::launch_stage_mixed(B, sel, L->has_src, P, prm->ko_sigma, mode, c, L->eng->stream, flags,
                                      L->chi_fp64_below, L->eng->aux.stream ? &L->eng->aux : nullptr);
    return gr::launch_stage_fp64(B, sel, L->has_src, P, prm->ko_sigma, mode, c, L->eng->stream);
}

// one stage over every block of a level: one launch per group of equal spacing
static hipError_t stage(gr_level* L, const int sel[3], const gr_ccz4_params* prm, int mode, const gr::StageCoef& c,
                        int precision) {
    for (gr_group& G : L->groups) {
        hipError_t st = stage_batch(L, G.batch, G.flags, sel, prm, mode, c, precision);
        if (st != hipSuccess) return st;
    }
    return hipSuccess;
}

// stage weights of TimeIntegrator::sspRK3Step, src/main.cpp:673,679,685, and the buffers
// each stage reads and writes (0 = u, 1 = s1, 2 = s2)
static gr::StageCoef stage_coef(int k, double dt) {
    if (k == 1) return gr::StageCoef{1.0, 0.0, 1.0 * dt};
    if (k == 2) return gr::StageCoef{0.75, 0.25, 0.25 * dt};
    return gr::StageCoef{1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt};
}
static const int kStageSel[4][3] = {{0, 0, 0}, {0, 0, 1}, {1, 0, 2}, {2, 0, 0}};

// RK3 stage k (1..3) of every block: the stage kernel, then the ghost fill `bc` of the
// buffer it wrote; pre (>= 0) fills the state's ghosts first
static hipError_t enqueue_stage(gr_level* L, const gr_ccz4_params* prm, double dt, int k, int bc, int pre,
                                int precision) {
    hipStream_t s = L->eng->stream;
    const gr::Batch& B = L->batch;
    hipError_t st;
    if (pre >= 0 && (st = timed(L, 2, [&] { return gr::launch_fill_ghosts_batch(B, 0, pre, s); })) != hipSuccess)
        return st;
    const int mode = k == 3 ? gr::MODE_FINAL : gr::MODE_STAGE;
    if ((st = timed(L, k == 3 ? 1 : 0,
                    [&] { return stage(L, kStageSel[k], prm, mode, stage_coef(k, dt), precision); })) != hipSuccess)
        return st;
    return timed(L, 2, [&] { return gr::launch_fill_ghosts_batch(B, kStageSel[k][2], bc, s); });
}

static hipError_t enqueue_rk3(gr_level* L, const gr_ccz4_params* prm, double dt, int bc, int pre, int precision) {
    hipError_t st = enqueue_stage(L, prm, dt, 1, bc, pre, precision);
    if (st == hipSuccess) st = enqueue_stage(L, prm, dt, 2, bc, GR_BC_NONE, precision);
    if (st == hipSuccess) st = enqueue_stage(L, prm, dt, 3, bc, GR_BC_NONE, precision);
    return st;
}

// One RK3 step of n independent levels as one schedule over two streams: the stage launches
// go to the engine stream in level order, stage by stage; each level's ghost fill goes to
// the fill stream once its stage has run (event), and the level's next stage waits for that
// fill (event). A level's fill thus overlaps the following levels' stage kernels, while each
// level sees exactly the launches of its own step in the same order.
static hipError_t enqueue_levels_rk3(gr_engine* e, int n, gr_level* const* lv, const double* dts,
                                     const gr_ccz4_params* prm, int bc, int pre, int precision) {
    const size_t need = 2 * 3 * (size_t)n + 1;
    while (e->ls_events.size() < need) {
        hipEvent_t ev = nullptr;
        hipError_t st = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (st != hipSuccess) return st;
        e->ls_events.push_back(ev);
    }
    hipEvent_t* staged = e->ls_events.data();          // [level][stage]
    hipEvent_t* filled = e->ls_events.data() + 3 * n;  // [level][stage]
    hipEvent_t fork = e->ls_events[6 * (size_t)n];
    hipStream_t A = e->stream, B = e->fill_stream;
    hipError_t st = hipEventRecord(fork, A);
    if (st == hipSuccess) st = hipStreamWaitEvent(B, fork, 0); // B joins A's order (and a capture)
    for (int k = 1; k <= 3 && st == hipSuccess; ++k)
        for (int i = 0; i < n && st == hipSuccess; ++i) {
            gr_level* L = lv[i];
            if (k > 1) st = hipStreamWaitEvent(A, filled[3 * i + k - 2], 0);
            if (st == hipSuccess && k == 1 && pre >= 0) st = gr::launch_fill_ghosts_batch(L->batch, 0, pre, A);
            const int mode = k == 3 ? gr::MODE_FINAL : gr::MODE_STAGE;
            if (st == hipSuccess) st = stage(L, kStageSel[k], prm, mode, stage_coef(k, dts[i]), precision);
            if (st == hipSuccess) st = hipEventRecord(staged[3 * i + k - 1], A);
            if (st == hipSuccess) st = hipStreamWaitEvent(B, staged[3 * i + k - 1], 0);
            if (st == hipSuccess) st = gr::launch_fill_ghosts_batch(L->batch, kStageSel[k][2], bc, B);
            if (st == hipSuccess) st = hipEventRecord(filled[3 * i + k - 1], B);
        }
    for (int i = 0; i < n && st == hipSuccess; ++i) st = hipStreamWaitEvent(A, filled[3 * i + 2], 0); // join
    return st;
}

extern "C" {

int gr_level_rk3_step(gr_level* L, const gr_ccz4_params* prm, double dt, int precision, int bc_mode, int pre_fill,
                      int use_graph) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if ((rc = check_params(e, prm))) return rc;
    if (!std::isfinite(dt)) return fail(e, GR_ERR_ARG, "dt must be finite");
    if (precision != GR_PREC_FP64 && precision != GR_PREC_MIXED) return fail(e, GR_ERR_ARG, "unknown precision");
    if (bc_mode < 0 || bc_mode > 2) return fail(e, GR_ERR_ARG, "bc_mode must be 0, 1 or 2");
    if (pre_fill < -1 || pre_fill > 2) return fail(e, GR_ERR_ARG, "pre_fill must be -1, 0, 1 or 2");
    if (!use_graph || L->profile) {
        GR_HIP(e, enqueue_rk3(L, prm, dt, bc_mode, pre_fill, precision));
        return GR_OK;
    }
    const double key[11] = {dt, prm->kappa1, prm->kappa2, prm->eta, prm->ko_sigma, prm->chi_blend_center,
                            prm->chi_blend_width, (double)precision, (double)bc_mode, (double)pre_fill,
                            (double)L->generation};
    if (!L->have_graph || std::memcmp(key, L->gkey, sizeof(key)) != 0) {
        if (L->exec) GR_HIP(e, hipGraphExecDestroy(L->exec));
        if (L->graph) GR_HIP(e, hipGraphDestroy(L->graph));
        L->exec = nullptr;
        L->graph = nullptr;
        L->have_graph = false;
        GR_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        hipError_t st = enqueue_rk3(L, prm, dt, bc_mode, pre_fill, precision);
        hipGraph_t g = nullptr;
        hipError_t st2 = hipStreamEndCapture(e->stream, &g);
        if (st != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hipfail(e, st, "capture of the RK3 step");
        }
        if (st2 != hipSuccess) return hipfail(e, st2, "hipStreamEndCapture");
        L->graph = g;
        GR_HIP(e, hipGraphInstantiate(&L->exec, L->graph, nullptr, nullptr, 0));
        std::memcpy(L->gkey, key, sizeof(key));
        L->have_graph = true;
    }
    GR_HIP(e, hipGraphLaunch(L->exec, e->stream));
    return GR_OK;
}

int gr_levels_rk3_step(int n, gr_level* const* levels, const double* dts, const gr_ccz4_params* prm, int precision,
                       int bc_mode, int pre_fill, int use_graph) {
    if (n < 1 || !levels || !dts || !levels[0]) return GR_ERR_ARG;
    gr_engine* e = levels[0]->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if ((rc = check_params(e, prm))) return rc;
    for (int i = 0; i < n; ++i) {
        if (!levels[i] || levels[i]->eng != e) return fail(e, GR_ERR_ARG, "every level must belong to one engine");
        for (int j = 0; j < i; ++j)
            if (levels[j] == levels[i]) return fail(e, GR_ERR_ARG, "a level appears twice");
        if (!std::isfinite(dts[i])) return fail(e, GR_ERR_ARG, "dt must be finite");
        if (levels[i]->profile) return fail(e, GR_ERR_STATE, "profiled levels step one at a time");
    }
    if (precision != GR_PREC_FP64 && precision != GR_PREC_MIXED) return fail(e, GR_ERR_ARG, "unknown precision");
    if (bc_mode < 0 || bc_mode > 2) return fail(e, GR_ERR_ARG, "bc_mode must be 0, 1 or 2");
    if (pre_fill < -1 || pre_fill > 2) return fail(e, GR_ERR_ARG, "pre_fill must be -1, 0, 1 or 2");
    if (!e->fill_stream) GR_HIP(e, hipStreamCreate(&e->fill_stream)); // default priority
    if (!use_graph) {
        GR_HIP(e, enqueue_levels_rk3(e, n, levels, dts, prm, bc_mode, pre_fill, precision));
        return GR_OK;
    }
    std::vector<double> key = {(double)n, prm->kappa1, prm->kappa2, prm->eta, prm->ko_sigma, prm->chi_blend_center,
                               prm->chi_blend_width, (double)precision, (double)bc_mode, (double)pre_fill};
    for (int i = 0; i < n; ++i) {
        key.push_back((double)reinterpret_cast<uintptr_t>(levels[i]));
        key.push_back((double)levels[i]->generation);
        key.push_back(dts[i]);
    }
    if (!e->ls_exec || key != e->ls_key) {
        if (e->ls_exec) GR_HIP(e, hipGraphExecDestroy(e->ls_exec));
        if (e->ls_graph) GR_HIP(e, hipGraphDestroy(e->ls_graph));
        e->ls_exec = nullptr;
        e->ls_graph = nullptr;
        e->ls_key.clear();
        GR_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        hipError_t st = enqueue_levels_rk3(e, n, levels, dts, prm, bc_mode, pre_fill, precision);
        hipGraph_t g = nullptr;
        hipError_t st2 = hipStreamEndCapture(e->stream, &g);
        if (st != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hipfail(e, st, "capture of the level-set step");
        }
        if (st2 != hipSuccess) return hipfail(e, st2, "hipStreamEndCapture");
        e->ls_graph = g;
        GR_HIP(e, hipGraphInstantiate(&e->ls_exec, e->ls_graph, nullptr, nullptr, 0));
        e->ls_key = key;
    }
    GR_HIP(e, hipGraphLaunch(e->ls_exec, e->stream));
    return GR_OK;
}

int gr_level_set_mixed_policy(gr_level* L, double chi_fp64_below) {
    if (!L) return GR_ERR_ARG;
    if (!std::isfinite(chi_fp64_below)) return fail(L->eng, GR_ERR_ARG, "threshold must be finite");
    L->chi_fp64_below = chi_fp64_below;
    ++L->generation; // captured steps bake the threshold in
    return GR_OK;
}

int gr_level_rk3_stage(gr_level* L, const gr_ccz4_params* prm, double dt, int stage_index, int precision,
                       int bc_mode, int pre_fill) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if ((rc = check_params(e, prm))) return rc;
    if (!std::isfinite(dt)) return fail(e, GR_ERR_ARG, "dt must be finite");
    if (stage_index < 1 || stage_index > 3) return fail(e, GR_ERR_ARG, "stage must be 1, 2 or 3");
    if (precision != GR_PREC_FP64 && precision != GR_PREC_MIXED) return fail(e, GR_ERR_ARG, "unknown precision");
    if (bc_mode < 0 || bc_mode > 2) return fail(e, GR_ERR_ARG, "bc_mode must be 0, 1 or 2");
    if (pre_fill < -1 || pre_fill > 2) return fail(e, GR_ERR_ARG, "pre_fill must be -1, 0, 1 or 2");
    GR_HIP(e, enqueue_stage(L, prm, dt, stage_index, bc_mode, pre_fill, precision));
    return GR_OK;
}

int gr_level_set_sources(gr_level* L, int b, const double* d_rho, const double* d_Si, const double* d_Sij,
                         const double* d_S) {
    int rc = level_block(L, b);
    if (rc) return rc;
    gr_engine* e = L->eng;
    const int nsrc = (d_rho != nullptr) + (d_Si != nullptr) + (d_Sij != nullptr) + (d_S != nullptr);
    if (nsrc != 0 && nsrc != 4) return fail(e, GR_ERR_ARG, "matter sources must be all null or all non-null");
    // the block tables are read by launches already queued: change them in stream order
    GR_HIP(e, hipStreamSynchronize(e->stream));
    L->src[b] = gr::Sources{d_rho, d_Si, d_Sij, d_S};
    L->has_src = false;
    for (const gr::Sources& S : L->src) L->has_src = L->has_src || S.rho != nullptr;
    auto refresh = [&](std::vector<gr::BatchBlock>& table, const std::vector<int>& which, gr::BatchBlock* d_table,
                       gr::Batch& B) -> hipError_t {
        for (size_t n = 0; n < which.size(); ++n) table[n].src = L->src[which[n]];
        if (which.size() == 1) {
            B.one.src = table[0].src;
            return hipSuccess;
        }
        B.one.src = table[0].src;
        return hipMemcpy(d_table, table.data(), table.size() * sizeof(gr::BatchBlock), hipMemcpyHostToDevice);
    };
    std::vector<int> all(L->blk.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
    GR_HIP(e, refresh(L->table, all, L->d_table, L->batch));
    for (gr_group& G : L->groups) GR_HIP(e, refresh(G.table, G.blocks, G.d_table, G.batch));
    ++L->generation; // captured steps hold the launch form and the single-block arguments
    return GR_OK;
}

int gr_level_rhs(gr_level* L, int b, const gr_ccz4_params* prm, int precision, double* host_rhs) {
    int rc = level_block(L, b);
    if (rc) return rc;
    gr_engine* e = L->eng;
    if ((rc = check_params(e, prm))) return rc;
    if (!host_rhs) return fail(e, GR_ERR_ARG, "null output");
    if (precision != GR_PREC_FP64 && precision != GR_PREC_MIXED) return fail(e, GR_ERR_ARG, "unknown precision");
    const size_t n = 22 * (size_t)L->blk[b].stride;
    GR_HIP(e, hipMemsetAsync(L->s1[b], 0, n * sizeof(double), e->stream));
    static const int sel[3] = {0, 1, 2};
    unsigned char* flags = nullptr;
    gr::Batch B = gr::single_batch(L->blk[b], L->u[b], L->u[b], L->s1[b], L->src[b]);
    for (gr_group& G : L->groups)
        for (int i : G.blocks)
            if (i == b) {
                flags = G.flags;
                B.variant = G.batch.variant; // the staging variant of the block's group
            }
    GR_HIP(e, stage_batch(L, B, flags, sel, prm, gr::MODE_RHS, gr::StageCoef{0.0, 0.0, 0.0}, precision));
    GR_HIP(e, hipMemcpyAsync(host_rhs, L->s1[b], n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_level_profile(gr_level* L, int enable) {
    if (!L) return GR_ERR_ARG;
    L->profile = enable != 0;
    return GR_OK;
}

int gr_level_profile_read(gr_level* L, int kind, double* total_ms, int* launches) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (kind < 0 || kind > 2 || !total_ms || !launches) return fail(e, GR_ERR_ARG, "bad profile query");
    GR_HIP(e, hipStreamSynchronize(e->stream));
    double t = 0.0;
    for (size_t i = 0; i + 1 < L->ev_used[kind]; i += 2) {
        float ms = 0.f;
        GR_HIP(e, hipEventElapsedTime(&ms, L->ev[kind][i], L->ev[kind][i + 1]));
        t += ms;
    }
    *total_ms = t;
    *launches = (int)(L->ev_used[kind] / 2);
    L->ev_used[kind] = 0;
    return GR_OK;
}

int gr_level_sync(gr_level* L) {
    if (!L) return GR_ERR_ARG;
    return gr_engine_sync(L->eng);
}

int gr_level_constraints_l2(gr_level* L, double* sumsq, long long* count) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!sumsq || !count) return fail(e, GR_ERR_ARG, "null output");
    double total = 0.0;
    long long n = 0;
    for (size_t b = 0; b < L->blk.size(); ++b) {
        const gr::BlockDesc& B = L->blk[b];
        int nparts = 0;
        GR_HIP(e, gr::launch_constraints(B, L->u[b], L->ham, nullptr, e->stream));
        GR_HIP(e, gr::launch_sumsq_interior(B, L->ham, 1, L->partial, L->partial_cap, &nparts, e->stream));
        GR_HIP(e, gr::launch_sum_partials(L->partial, nparts, L->red, e->stream));
        double h = 0.0;
        GR_HIP(e, hipMemcpyAsync(&h, L->red, sizeof(double), hipMemcpyDeviceToHost, e->stream));
        GR_HIP(e, hipStreamSynchronize(e->stream));
        total += h;
        n += (long long)B.n[0] * B.n[1] * B.n[2];
    }
    *sumsq = total;
    *count = n;
    return GR_OK;
}

int gr_ccz4_constraints_l2_noalloc(gr_engine* e, const gr_geom* geom, const double* d_u, double* d_sumsq,
                                   long long* count) {
    int rc = use_device(e);
    if (rc) return rc;
    gr::BlockDesc b;
    if ((rc = check_geom(e, geom, &b))) return rc;
    if (!d_u || !d_sumsq || !count) return fail(e, GR_ERR_ARG, "null state or output");
    if (!e->scratch) return fail(e, GR_ERR_NODEV, "no scratch address on this device");
    // the partition gives each workgroup at least ceil(cells / kNoAllocPartials) cells, so
    // every block the geometry check accepts fits the scratch; checked before any launch
    if (gr::reduce_parts(b, kNoAllocPartials) > kNoAllocPartials)
        return fail(e, GR_ERR_GEOM, "block needs more reduction partials than the scratch holds (2048)");
    int nparts = 0;
    GR_HIP(e, gr::launch_ham_sumsq(b, d_u, e->scratch, kNoAllocPartials, &nparts, e->stream));
    GR_HIP(e, gr::launch_sum_partials(e->scratch, nparts, d_sumsq, e->stream));
    *count = (long long)b.n[0] * b.n[1] * b.n[2];
    return GR_OK;
}

int gr_level_read_cell(gr_level* L, int b, int var, int i, int j, int k, double* value) {
    int rc = level_block(L, b);
    if (rc) return rc;
    gr_engine* e = L->eng;
    const gr::BlockDesc& B = L->blk[b];
    if (!value || var < 0 || var >= 22 || i < 0 || j < 0 || k < 0 || i >= B.nt[0] || j >= B.nt[1] || k >= B.nt[2])
        return fail(e, GR_ERR_ARG, "cell or field index out of range");
    const long off = var * B.stride + i + (long)B.nt[0] * (j + (long)B.nt[1] * k);
    GR_HIP(e, hipMemcpyAsync(value, L->u[b] + off, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

int gr_level_nonfinite(gr_level* L, long long* count) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!count) return fail(e, GR_ERR_ARG, "null output");
    double total = 0.0;
    for (size_t b = 0; b < L->blk.size(); ++b) {
        int nparts = 0;
        GR_HIP(e, gr::launch_count_nonfinite(L->blk[b], L->u[b], L->partial, L->partial_cap, &nparts, e->stream));
        GR_HIP(e, gr::launch_sum_partials(L->partial, nparts, L->red, e->stream));
        double h = 0.0;
        GR_HIP(e, hipMemcpyAsync(&h, L->red, sizeof(double), hipMemcpyDeviceToHost, e->stream));
        GR_HIP(e, hipStreamSynchronize(e->stream));
        total += h;
    }
    *count = (long long)total;
    return GR_OK;
}

int gr_level_first_nonfinite(gr_level* L, int* found, int* block, int* var, int* i, int* j, int* k, double* value) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!found || !block || !var || !i || !j || !k || !value) return fail(e, GR_ERR_ARG, "null output");
    *found = 0;
    unsigned long long* part = reinterpret_cast<unsigned long long*>(L->partial);
    unsigned long long* res = reinterpret_cast<unsigned long long*>(L->red);
    for (size_t b = 0; b < L->blk.size(); ++b) {
        const gr::BlockDesc& B = L->blk[b];
        int nparts = 0;
        GR_HIP(e, gr::launch_first_nonfinite(B, L->u[b], part, L->partial_cap, &nparts, e->stream));
        GR_HIP(e, gr::launch_min_partials(part, nparts, res, e->stream));
        unsigned long long q = 0;
        GR_HIP(e, hipMemcpyAsync(&q, res, sizeof(q), hipMemcpyDeviceToHost, e->stream));
        GR_HIP(e, hipStreamSynchronize(e->stream));
        if (q == ~0ull) continue;
        const long long ncell = (long long)B.n[0] * B.n[1] * B.n[2];
        const long long t = (long long)q % ncell;
        *found = 1;
        *block = (int)b;
        *var = (int)((long long)q / ncell);
        *i = B.ng + (int)(t % B.n[0]);
        *j = B.ng + (int)((t / B.n[0]) % B.n[1]);
        *k = B.ng + (int)(t / ((long long)B.n[0] * B.n[1]));
        const long off = *var * B.stride + *i + (long)B.nt[0] * (*j + (long)B.nt[1] * *k);
        GR_HIP(e, hipMemcpyAsync(value, L->u[b] + off, sizeof(double), hipMemcpyDeviceToHost, e->stream));
        GR_HIP(e, hipStreamSynchronize(e->stream));
        return GR_OK;
    }
    return GR_OK;
}

int gr_level_init_gauge_wave(gr_level* L, double amplitude, double wavelength) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!std::isfinite(amplitude) || !std::isfinite(wavelength) || wavelength == 0.0)
        return fail(e, GR_ERR_ARG, "gauge wave needs a finite amplitude and a finite non-zero wavelength");
    for (size_t b = 0; b < L->blk.size(); ++b)
        GR_HIP(e, gr::launch_init_gauge_wave(L->blk[b], L->u[b], amplitude, wavelength, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    return GR_OK;
}

} // extern "C"

// per-block CFL partials of a level reduced into out[0] (one device double), in stream order
static hipError_t enqueue_cfl(gr_level* L, double dt, double* partial, int per_block, double* out) {
    hipStream_t s = L->eng->stream;
    int used = 0;
    for (size_t b = 0; b < L->blk.size(); ++b) {
        int nparts = 0;
        hipError_t st = gr::launch_cfl_max(L->blk[b], L->u[b], dt, partial + used, per_block, &nparts, s);
        if (st != hipSuccess) return st;
        used += nparts;
    }
    return gr::launch_max_partials(partial, used, out, s);
}

constexpr int kCflPartsPerBlock = 256;

extern "C" {

int gr_level_cfl(gr_level* L, double dt, double* max_cfl) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!max_cfl || !std::isfinite(dt)) return fail(e, GR_ERR_ARG, "null output or non-finite dt");
    if (!L->cfl_part)
        GR_HIP(e, hipMalloc((void**)&L->cfl_part, (L->blk.size() * kCflPartsPerBlock + 1) * sizeof(double)));
    double* part = L->cfl_part;
    hipError_t st = enqueue_cfl(L, dt, part, kCflPartsPerBlock, part + L->blk.size() * kCflPartsPerBlock);
    if (st == hipSuccess)
        st = hipMemcpyAsync(max_cfl, part + L->blk.size() * kCflPartsPerBlock, sizeof(double), hipMemcpyDeviceToHost,
                            e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) return hipfail(e, st, "advective CFL reduction");
    return GR_OK;
}

} // extern "C"

namespace {

// mean device time (ms) of `reps` launches on the engine stream, after one untimed launch
template <class F>
hipError_t time_launches(gr_engine* e, int reps, F&& launch, double* ms) {
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t st = hipEventCreate(&a);
    if (st == hipSuccess) st = hipEventCreate(&b);
    if (st == hipSuccess) st = launch();
    if (st == hipSuccess) st = hipEventRecord(a, e->stream);
    for (int r = 0; r < reps && st == hipSuccess; ++r) st = launch();
    if (st == hipSuccess) st = hipEventRecord(b, e->stream);
    if (st == hipSuccess) st = hipEventSynchronize(b);
    float t = 0.0f;
    if (st == hipSuccess) st = hipEventElapsedTime(&t, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    *ms = (double)t / reps;
    return st;
}

constexpr int kGroupChoices[4] = {1, 2, 4, 11};

} // namespace

extern "C" {

int gr_level_autotune(gr_level* L, const gr_ccz4_params* prm, double dt, double* times, int cap, int* n) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if ((rc = check_params(e, prm))) return rc;
    if (!std::isfinite(dt)) return fail(e, GR_ERR_ARG, "dt must be finite");
    if (cap > 0 && !times) return fail(e, GR_ERR_ARG, "null times");
    std::vector<double> out;
    // stage 1 reads the state and writes the first stage buffer only
    static const int sel1[3] = {0, 0, 1};
    const gr::StageCoef c1 = stage_coef(1, dt);
    for (gr_group& G : L->groups) {
        int best = 0;
        double tbest = INFINITY;
        for (int v = 0; v < gr::FP64_STAGE_VARIANTS; ++v) {
            if (v != 0 && !G.batch.pairs_ok) {
                out.push_back(-1.0);
                continue;
            }
            gr::Batch B = G.batch;
            B.variant = v;
            double ms = 0.0;
            GR_HIP(e, time_launches(
                          e, 3, [&] { return stage_batch(L, B, nullptr, sel1, prm, gr::MODE_STAGE, c1, GR_PREC_FP64); },
                          &ms));
            out.push_back(ms);
            if (ms < tbest) {
                tbest = ms;
                best = v;
            }
        }
        G.batch.variant = best;
        tune_put(GR_TUNE_STAGE, blocks_class(L, G.blocks), GR_PREC_FP64, best);
    }
    int gbest = L->batch.ghost_groups;
    double tbest = INFINITY;
    for (int g : kGroupChoices) {
        gr::Batch B = L->batch;
        B.ghost_groups = g;
        double ms = 0.0;
        // the outflow fill of the first stage buffer's ghosts
        GR_HIP(e, time_launches(e, 5, [&] { return gr::launch_fill_ghosts_batch(B, 1, GR_BC_OUTFLOW, e->stream); },
                                &ms));
        out.push_back(ms);
        if (ms < tbest) {
            tbest = ms;
            gbest = g;
        }
    }
    L->batch.ghost_groups = gbest;
    tune_put(GR_TUNE_GHOST, level_class(L), GR_PREC_FP64, gbest);
    ++L->generation; // captured steps hold the launch forms
    for (int i = 0; i < cap && i < (int)out.size(); ++i) times[i] = out[i];
    if (n) *n = (int)std::min<size_t>(out.size(), (size_t)std::max(cap, 0));
    return GR_OK;
}

int gr_level_get_tuning(gr_level* L, int kind, int group, int* value) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    if (!value) return fail(e, GR_ERR_ARG, "null output");
    if (kind == GR_TUNE_STAGE) {
        if (group < 0 || group >= (int)L->groups.size()) return fail(e, GR_ERR_ARG, "group index out of range");
        const gr::Batch& B = L->groups[group].batch;
        *value = (B.variant != 0 && B.pairs_ok) ? B.variant : 0;
        return GR_OK;
    }
    if (kind == GR_TUNE_GHOST) {
        *value = L->batch.ghost_groups;
        return GR_OK;
    }
    return fail(e, GR_ERR_ARG, "unknown tuning kind for a level");
}

int gr_level_set_tuning(gr_level* L, int kind, int group, int value) {
    if (!L) return GR_ERR_ARG;
    gr_engine* e = L->eng;
    if (kind == GR_TUNE_STAGE) {
        if (group < 0 || group >= (int)L->groups.size()) return fail(e, GR_ERR_ARG, "group index out of range");
        if (!tune_valid(kind, value)) return fail(e, GR_ERR_ARG, "unknown staging variant");
        if (value != 0 && !L->groups[group].batch.pairs_ok)
            return fail(e, GR_ERR_ARG, "direct LDS staging needs even row lengths and 16-byte aligned buffers");
        L->groups[group].batch.variant = value;
    } else if (kind == GR_TUNE_GHOST) {
        if (!tune_valid(kind, value)) return fail(e, GR_ERR_ARG, "ghost field groups must be 1, 2, 4 or 11");
        L->batch.ghost_groups = value;
    } else {
        return fail(e, GR_ERR_ARG, "unknown tuning kind for a level");
    }
    ++L->generation;
    return GR_OK;
}

int gr_tune_save(const char* path) {
    if (!path) return GR_ERR_ARG;
    std::ofstream f(path);
    if (!f) return GR_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    f << "# granite-ccz4 tuning table: kind class precision value\n";
    for (const auto& kv : g_tune)
        f << std::get<0>(kv.first) << ' ' << std::get<1>(kv.first) << ' ' << std::get<2>(kv.first) << ' ' << kv.second
          << '\n';
    return f ? GR_OK : GR_ERR_ARG;
}

int gr_tune_load(const char* path) {
    if (!path) return GR_ERR_ARG;
    std::ifstream f(path);
    if (!f) return GR_ERR_ARG;
    std::vector<std::pair<std::tuple<int, int, int>, int>> rows;
    std::string line;
    while (std::getline(f, line)) {
        const size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '#') continue;
        std::istringstream in(line);
        int kind = -1, cls = -1, prec = -1, value = -1;
        std::string rest;
        if (!(in >> kind >> cls >> prec >> value) || (in >> rest)) return GR_ERR_ARG;
        if (kind < GR_TUNE_STAGE || kind > GR_TUNE_TRANSFER || cls < 0 || cls > 3 || prec < 0 || prec > 1 ||
            !tune_valid(kind, value))
            return GR_ERR_ARG;
        rows.push_back({std::make_tuple(kind, cls, prec), value});
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (const auto& r : rows) g_tune[r.first] = r.second;
    return GR_OK;
}

int gr_tune_clear(void) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.clear();
    return GR_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------------------
// Berger-Oliger hierarchy
// ---------------------------------------------------------------------------------------

struct gr_hierarchy {
    gr_engine* eng = nullptr;
    std::vector<gr_level*> lv;
    // per level, per block: outflow faces (bit 2d low, 2d+1 high side of axis d) that no
    // block of the parent level covers; all six on level 0
    std::vector<std::vector<int>> faces;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool have_graph = false;
    double gkey[8] = {0};
    std::vector<double> policy; // each level's mixed-precision threshold and generation at capture
    // advective CFL guard after every level step (src/main.cpp:1155-1176): one device value
    // per evolve call of a coarse step, in subcycle order, and the level of each
    bool cfl_on = false;
    double* cfl = nullptr;
    double* cfl_part = nullptr;
    std::vector<int> cfl_level;
    // per level l > 0: prolongation stencil tables of every (fine block f, parent block c)
    // pair, pair f * nparents + c at ptab_ofs[l][f * nparents + c] entries into ptab[l]
    std::vector<gr::ProlongAxis*> ptab;
    std::vector<std::vector<size_t>> ptab_ofs;
    std::vector<int> tgroups; // per level l > 0: field groups of the transfers into / from it
};

namespace {

// GridBlock::x (include/granite/core/grid.hpp:92-94)
double cell_x(const gr::BlockDesc& b, int d, int idx) { return b.lo[d] + ((double)idx - b.ng + 0.5) * b.dx[d]; }

// faces of fine block f left to the outflow copy by fillGhostZones (src/amr/amr.cpp:412-428):
// a side counts as covered when any parent block's first (last) interior cell centre lies
// at or below (above) the fine block's first (last) interior centre -/+ half a fine cell
int uncovered_faces(const gr::BlockDesc& f, const std::vector<gr::BlockDesc>& parents) {
    bool lo[3] = {false, false, false}, hi[3] = {false, false, false};
    for (const gr::BlockDesc& c : parents)
        for (int d = 0; d < 3; ++d) {
            const double fine_lo = cell_x(f, d, f.ng), fine_hi = cell_x(f, d, f.ng + f.n[d] - 1);
            const double c_lo = cell_x(c, d, c.ng), c_hi = cell_x(c, d, c.ng + c.n[d] - 1);
            if (c_lo <= fine_lo + 0.5 * f.dx[d]) lo[d] = true;
            if (c_hi >= fine_hi - 0.5 * f.dx[d]) hi[d] = true;
        }
    int m = 0;
    for (int d = 0; d < 3; ++d) m |= (lo[d] ? 0 : 1 << (2 * d)) | (hi[d] ? 0 : 1 << (2 * d + 1));
    return m;
}

// prolongation into level l's states (into its first stage buffers when `to_stage`, which
// the autotuner times without touching the state), with `groups` field groups (0: the level's)
hipError_t h_prolongate(gr_hierarchy* H, int l, bool to_stage = false, int groups = 0) {
    gr_level* F = H->lv[l];
    gr_level* C = H->lv[l - 1];
    const int g = groups > 0 ? groups : H->tgroups[l];
    for (size_t f = 0; f < F->blk.size(); ++f)
        for (size_t c = 0; c < C->blk.size(); ++c) {
            const gr::ProlongAxis* tab = H->ptab[l] + H->ptab_ofs[l][f * C->blk.size() + c];
            hipError_t st = gr::launch_prolongate(C->blk[c], C->u[c], F->blk[f], to_stage ? F->s1[f] : F->u[f], tab,
                                                  H->eng->stream, g);
            if (st != hipSuccess) return st;
        }
    return hipSuccess;
}

// restriction of level l into level l - 1 (into its first stage buffers when `to_stage`)
hipError_t h_restrict(gr_hierarchy* H, int l, bool to_stage = false, int groups = 0) {
    gr_level* F = H->lv[l];
    gr_level* C = H->lv[l - 1];
    const int g = groups > 0 ? groups : H->tgroups[l];
    for (size_t f = 0; f < F->blk.size(); ++f)
        for (size_t c = 0; c < C->blk.size(); ++c) {
            hipError_t st = gr::launch_restrict(F->blk[f], F->u[f], C->blk[c], to_stage ? C->s1[c] : C->u[c],
                                                H->eng->stream, g);
            if (st != hipSuccess) return st;
        }
    return hipSuccess;
}

// fillGhostZones(l), src/amr/amr.cpp:400-482
hipError_t h_fill(gr_hierarchy* H, int l) {
    gr_level* L = H->lv[l];
    hipStream_t s = H->eng->stream;
    if (l == 0) return gr::launch_fill_ghosts_batch(L->batch, 0, GR_BC_OUTFLOW, s);
    hipError_t st = h_prolongate(H, l);
    for (size_t b = 0; st == hipSuccess && b < L->blk.size(); ++b)
        if (H->faces[l][b])
            st = gr::launch_fill_ghosts(L->blk[b], L->u[b], GR_BC_OUTFLOW, s, H->faces[l][b], L->batch.ghost_groups);
    return st;
}

// subcycle(l), src/amr/amr.cpp:303-350, without the regrid of its step 6. Its step 1
// (prolongation into level l > 0) is the first half of fillGhostZones(l), which follows
// at once and reads the same parent data, so the prolongation runs once.
hipError_t h_subcycle(gr_hierarchy* H, int l, double dt, const gr_ccz4_params* prm, int precision, int& slot) {
    hipError_t st = h_fill(H, l);
    if (st == hipSuccess) st = enqueue_rk3(H->lv[l], prm, dt, GR_BC_SOMMERFELD, GR_BC_NONE, precision);
    // the reference's evolve callback ends with the CFL guard over the level's blocks
    if (st == hipSuccess && H->cfl_on) st = enqueue_cfl(H->lv[l], dt, H->cfl_part, kCflPartsPerBlock, H->cfl + slot++);
    if (st == hipSuccess) st = h_fill(H, l);
    if (st == hipSuccess && l + 1 < (int)H->lv.size()) {
        const double dt_fine = dt / 2.0;
        for (int r = 0; r < 2 && st == hipSuccess; ++r) st = h_subcycle(H, l + 1, dt_fine, prm, precision, slot);
        if (st == hipSuccess) st = h_restrict(H, l + 1);
    }
    return st;
}

hipError_t h_step(gr_hierarchy* H, double dt0, const gr_ccz4_params* prm, int precision) {
    int slot = 0;
    return h_subcycle(H, 0, dt0, prm, precision, slot);
}

int check_hier_level(gr_hierarchy* H, int level, int lo) {
    if (!H) return GR_ERR_ARG;
    if (level < lo || level >= (int)H->lv.size()) return fail(H->eng, GR_ERR_ARG, "level index out of range");
    return use_device(H->eng);
}

} // namespace

extern "C" {

int gr_hierarchy_create(gr_engine* e, int nlevels, gr_level* const* levels, gr_hierarchy** out) {
    int rc = use_device(e);
    if (rc) return rc;
    if (!out || !levels || nlevels < 1) return fail(e, GR_ERR_ARG, "bad hierarchy arguments");
    *out = nullptr;
    for (int l = 0; l < nlevels; ++l) {
        if (!levels[l] || levels[l]->eng != e) return fail(e, GR_ERR_ARG, "every level must belong to this engine");
        if (l > 0)
            for (int d = 0; d < 3; ++d) {
                const double want = levels[l - 1]->blk[0].dx[d] / 2.0;
                if (std::fabs(levels[l]->blk[0].dx[d] - want) > 1e-6 * want)
                    return fail(e, GR_ERR_GEOM, "dx of each level must be half of its parent's (refinement ratio 2)");
            }
    }
    gr_hierarchy* H = new gr_hierarchy();
    H->eng = e;
    H->lv.assign(levels, levels + nlevels);
    H->faces.resize(nlevels);
    H->faces[0].assign(levels[0]->blk.size(), 0x3f);
    for (int l = 1; l < nlevels; ++l)
        for (const gr::BlockDesc& f : levels[l]->blk) H->faces[l].push_back(uncovered_faces(f, levels[l - 1]->blk));
    H->ptab.assign(nlevels, nullptr);
    H->ptab_ofs.resize(nlevels);
    H->tgroups.assign(nlevels, 4);
    for (int l = 1; l < nlevels; ++l) {
        int g = 0;
        if (tune_get(GR_TUNE_TRANSFER, level_class(levels[l]), GR_PREC_FP64, &g)) H->tgroups[l] = g;
    }
    for (int l = 1; l < nlevels; ++l) {
        size_t n = 0;
        for (const gr::BlockDesc& f : levels[l]->blk)
            for (size_t c = 0; c < levels[l - 1]->blk.size(); ++c) {
                H->ptab_ofs[l].push_back(n);
                n += (size_t)f.nt[0] + f.nt[1] + f.nt[2];
            }
        hipError_t st = hipMalloc((void**)&H->ptab[l], n * sizeof(gr::ProlongAxis));
        for (size_t f = 0; st == hipSuccess && f < levels[l]->blk.size(); ++f)
            for (size_t c = 0; st == hipSuccess && c < levels[l - 1]->blk.size(); ++c)
                st = gr::launch_prolong_axes(levels[l - 1]->blk[c], levels[l]->blk[f],
                                             H->ptab[l] + H->ptab_ofs[l][f * levels[l - 1]->blk.size() + c], e->stream);
        if (st != hipSuccess) {
            gr_hierarchy_destroy(H);
            return hipfail(e, st, "prolongation stencil tables");
        }
    }
    *out = H;
    return GR_OK;
}

int gr_hierarchy_destroy(gr_hierarchy* H) {
    if (!H) return GR_ERR_ARG;
    (void)hipSetDevice(H->eng->device);
    (void)hipStreamSynchronize(H->eng->stream);
    if (H->exec) (void)hipGraphExecDestroy(H->exec);
    if (H->graph) (void)hipGraphDestroy(H->graph);
    for (gr::ProlongAxis* p : H->ptab)
        if (p) (void)hipFree(p);
    if (H->cfl) (void)hipFree(H->cfl);
    if (H->cfl_part) (void)hipFree(H->cfl_part);
    delete H;
    return GR_OK;
}

int gr_hierarchy_set_cfl_monitor(gr_hierarchy* H, int enable) {
    if (!H) return GR_ERR_ARG;
    gr_engine* e = H->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (enable && !H->cfl) {
        // one slot per evolve call of a coarse step: level l steps 2^l times
        H->cfl_level.clear();
        size_t maxblocks = 0;
        for (size_t l = 0; l < H->lv.size(); ++l) {
            for (size_t r = 0; r < ((size_t)1 << l); ++r) H->cfl_level.push_back((int)l);
            maxblocks = std::max(maxblocks, H->lv[l]->blk.size());
        }
        // subcycle order: l, then two passes of l + 1 (recursively)
        std::vector<int> order;
        struct Walk {
            static void run(std::vector<int>& o, int l, int n) {
                o.push_back(l);
                if (l + 1 < n)
                    for (int r = 0; r < 2; ++r) run(o, l + 1, n);
            }
        };
        Walk::run(order, 0, (int)H->lv.size());
        H->cfl_level = order;
        GR_HIP(e, hipMalloc((void**)&H->cfl, order.size() * sizeof(double)));
        GR_HIP(e, hipMalloc((void**)&H->cfl_part, maxblocks * kCflPartsPerBlock * sizeof(double)));
    }
    H->cfl_on = enable != 0;
    return GR_OK;
}

int gr_hierarchy_cfl(gr_hierarchy* H, double* values, int* levels, int cap, int* n) {
    if (!H) return GR_ERR_ARG;
    gr_engine* e = H->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!n || (cap > 0 && (!values || !levels))) return fail(e, GR_ERR_ARG, "null output");
    if (!H->cfl_on || !H->cfl) return fail(e, GR_ERR_STATE, "the CFL monitor is off (gr_hierarchy_set_cfl_monitor)");
    const int m = (int)H->cfl_level.size();
    *n = m;
    if (cap < m) return fail(e, GR_ERR_ARG, "output capacity below the number of level steps per coarse step");
    GR_HIP(e, hipMemcpyAsync(values, H->cfl, m * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    GR_HIP(e, hipStreamSynchronize(e->stream));
    for (int i = 0; i < m; ++i) levels[i] = H->cfl_level[i];
    return GR_OK;
}

int gr_hierarchy_prolongate(gr_hierarchy* H, int level) {
    int rc = check_hier_level(H, level, 1);
    if (rc) return rc;
    GR_HIP(H->eng, h_prolongate(H, level));
    return GR_OK;
}

int gr_hierarchy_restrict(gr_hierarchy* H, int level) {
    int rc = check_hier_level(H, level, 1);
    if (rc) return rc;
    GR_HIP(H->eng, h_restrict(H, level));
    return GR_OK;
}

int gr_hierarchy_fill_ghosts(gr_hierarchy* H, int level) {
    int rc = check_hier_level(H, level, 0);
    if (rc) return rc;
    GR_HIP(H->eng, h_fill(H, level));
    return GR_OK;
}

int gr_hierarchy_step(gr_hierarchy* H, const gr_ccz4_params* prm, double dt0, int precision, int use_graph) {
    if (!H) return GR_ERR_ARG;
    gr_engine* e = H->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if ((rc = check_params(e, prm))) return rc;
    if (!std::isfinite(dt0)) return fail(e, GR_ERR_ARG, "dt must be finite");
    if (precision != GR_PREC_FP64 && precision != GR_PREC_MIXED) return fail(e, GR_ERR_ARG, "unknown precision");
    bool profiling = false;
    for (gr_level* L : H->lv) profiling = profiling || L->profile;
    if (!use_graph || profiling) {
        GR_HIP(e, h_step(H, dt0, prm, precision));
        return GR_OK;
    }
    const double key[8] = {dt0, prm->kappa1, prm->kappa2, prm->eta, prm->ko_sigma, prm->chi_blend_center,
                           prm->chi_blend_width, (double)precision};
    // the mixed-precision threshold of every level is baked into the captured kernels
    std::vector<double> policy;
    for (gr_level* L : H->lv) {
        policy.push_back(L->chi_fp64_below);
        policy.push_back((double)L->generation);
    }
    policy.push_back(H->cfl_on ? 1.0 : 0.0);
    const bool fresh = H->have_graph && std::memcmp(key, H->gkey, sizeof(key)) == 0 && policy == H->policy;
    if (!fresh) {
        if (H->exec) GR_HIP(e, hipGraphExecDestroy(H->exec));
        if (H->graph) GR_HIP(e, hipGraphDestroy(H->graph));
        H->exec = nullptr;
        H->graph = nullptr;
        H->have_graph = false;
        GR_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        hipError_t st = h_step(H, dt0, prm, precision);
        hipGraph_t g = nullptr;
        hipError_t st2 = hipStreamEndCapture(e->stream, &g);
        if (st != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hipfail(e, st, "capture of the coarse step");
        }
        if (st2 != hipSuccess) return hipfail(e, st2, "hipStreamEndCapture");
        H->graph = g;
        GR_HIP(e, hipGraphInstantiate(&H->exec, H->graph, nullptr, nullptr, 0));
        std::memcpy(H->gkey, key, sizeof(key));
        H->policy = policy;
        H->have_graph = true;
    }
    GR_HIP(e, hipGraphLaunch(H->exec, e->stream));
    return GR_OK;
}

int gr_hierarchy_constraints_l2(gr_hierarchy* H, double* ham_sumsq, double* mom_sumsq, long long* count) {
    if (!H) return GR_ERR_ARG;
    gr_engine* e = H->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!count) return fail(e, GR_ERR_ARG, "null output");
    double th = 0.0, tm = 0.0;
    long long n = 0;
    for (gr_level* L : H->lv) {
        if (mom_sumsq && !L->mom) GR_HIP(e, hipMalloc((void**)&L->mom, 3 * L->maxstride * sizeof(double)));
        for (size_t b = 0; b < L->blk.size(); ++b) {
            const gr::BlockDesc& B = L->blk[b];
            GR_HIP(e, gr::launch_constraints(B, L->u[b], L->ham, mom_sumsq ? L->mom : nullptr, e->stream));
            double h[2] = {0.0, 0.0};
            for (int q = 0; q < 2; ++q) {
                if (q == 0 ? !ham_sumsq : !mom_sumsq) continue;
                int nparts = 0;
                GR_HIP(e, gr::launch_sumsq_interior(B, q == 0 ? L->ham : L->mom, q == 0 ? 1 : 3, L->partial,
                                                    L->partial_cap, &nparts, e->stream));
                GR_HIP(e, gr::launch_sum_partials(L->partial, nparts, L->red, e->stream));
                GR_HIP(e, hipMemcpyAsync(&h[q], L->red, sizeof(double), hipMemcpyDeviceToHost, e->stream));
                GR_HIP(e, hipStreamSynchronize(e->stream));
            }
            th += h[0];
            tm += h[1];
            n += (long long)B.n[0] * B.n[1] * B.n[2];
        }
    }
    if (ham_sumsq) *ham_sumsq = th;
    if (mom_sumsq) *mom_sumsq = tm;
    *count = n;
    return GR_OK;
}

int gr_hierarchy_autotune(gr_hierarchy* H, const gr_ccz4_params* prm, double dt0, double* times, int cap, int* n) {
    if (!H) return GR_ERR_ARG;
    gr_engine* e = H->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (cap > 0 && !times) return fail(e, GR_ERR_ARG, "null times");
    std::vector<double> out;
    double dt = dt0;
    for (gr_level* L : H->lv) {
        std::vector<double> t(2 * L->groups.size() + 4);
        int m = 0;
        if ((rc = gr_level_autotune(L, prm, dt, t.data(), (int)t.size(), &m))) return rc;
        out.insert(out.end(), t.begin(), t.begin() + m);
        dt = dt / 2.0;
    }
    for (int l = 1; l < (int)H->lv.size(); ++l) {
        int best = H->tgroups[l];
        double tbest = INFINITY;
        for (int g : kGroupChoices) {
            double ms = 0.0;
            // into the stage buffers: the states are untouched
            GR_HIP(e, time_launches(e, 3, [&] {
                hipError_t st = h_prolongate(H, l, true, g);
                return st == hipSuccess ? h_restrict(H, l, true, g) : st;
            }, &ms));
            out.push_back(ms);
            if (ms < tbest) {
                tbest = ms;
                best = g;
            }
        }
        H->tgroups[l] = best;
        tune_put(GR_TUNE_TRANSFER, level_class(H->lv[l]), GR_PREC_FP64, best);
    }
    H->have_graph = false; // the captured coarse step holds the launch forms
    for (int i = 0; i < cap && i < (int)out.size(); ++i) times[i] = out[i];
    if (n) *n = (int)std::min<size_t>(out.size(), (size_t)std::max(cap, 0));
    return GR_OK;
}

int gr_hierarchy_get_tuning(gr_hierarchy* H, int level, int* groups) {
    if (!H) return GR_ERR_ARG;
    if (!groups || level < 1 || level >= (int)H->lv.size()) return fail(H->eng, GR_ERR_ARG, "level must be 1 .. n-1");
    *groups = H->tgroups[level];
    return GR_OK;
}

int gr_hierarchy_set_tuning(gr_hierarchy* H, int level, int groups) {
    if (!H) return GR_ERR_ARG;
    if (level < 1 || level >= (int)H->lv.size()) return fail(H->eng, GR_ERR_ARG, "level must be 1 .. n-1");
    if (!tune_valid(GR_TUNE_TRANSFER, groups)) return fail(H->eng, GR_ERR_ARG, "transfer field groups must be 1, 2, 4 or 11");
    H->tgroups[level] = groups;
    H->have_graph = false;
    return GR_OK;
}

} // extern "C"

// ------------------------------------------------------------------------------------
// checkpoint staging: device snapshot, then device -> pinned host on a copy stream
// ------------------------------------------------------------------------------------

struct gr_checkpoint {
    gr_engine* eng = nullptr;
    std::vector<gr_level*> lv;
    std::vector<std::pair<gr_level*, int>> blocks; // (level, block), coarsest level first
    std::vector<size_t> ofs;                       // first double of each block in the buffers
    size_t total = 0;                              // doubles
    double* dev = nullptr;                         // device snapshot
    double* host = nullptr;                        // pinned host copy
    hipStream_t copy = nullptr;                    // default priority, like every stream here
    hipEvent_t taken = nullptr, landed = nullptr;  // snapshot on the engine stream / host copy done
    bool pending = false;
};

namespace {

void checkpoint_free(gr_checkpoint* c) {
    if (c->copy) (void)hipStreamSynchronize(c->copy);
    if (c->taken) (void)hipEventDestroy(c->taken);
    if (c->landed) (void)hipEventDestroy(c->landed);
    if (c->copy) (void)hipStreamDestroy(c->copy);
    if (c->dev) (void)hipFree(c->dev);
    if (c->host) (void)hipHostFree(c->host);
    delete c;
}

} // namespace

extern "C" {

int gr_checkpoint_create(gr_engine* e, int nlevels, gr_level* const* levels, gr_checkpoint** out) {
    int rc = use_device(e);
    if (rc) return rc;
    if (!out || !levels || nlevels < 1) return fail(e, GR_ERR_ARG, "bad checkpoint arguments");
    *out = nullptr;
    for (int l = 0; l < nlevels; ++l)
        if (!levels[l] || levels[l]->eng != e) return fail(e, GR_ERR_ARG, "every level must belong to this engine");
    gr_checkpoint* c = new gr_checkpoint();
    c->eng = e;
    c->lv.assign(levels, levels + nlevels);
    for (gr_level* L : c->lv)
        for (size_t b = 0; b < L->blk.size(); ++b) {
            c->blocks.emplace_back(L, (int)b);
            c->ofs.push_back(c->total);
            c->total += 22 * (size_t)L->blk[b].stride;
        }
    hipError_t st = hipMalloc((void**)&c->dev, c->total * sizeof(double));
    if (st == hipSuccess) st = hipHostMalloc((void**)&c->host, c->total * sizeof(double), hipHostMallocDefault);
    if (st == hipSuccess) st = hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking);
    if (st == hipSuccess) st = hipEventCreateWithFlags(&c->taken, hipEventDisableTiming);
    if (st == hipSuccess) st = hipEventCreateWithFlags(&c->landed, hipEventDisableTiming);
    if (st != hipSuccess) {
        checkpoint_free(c);
        return hipfail(e, st, "checkpoint buffers");
    }
    *out = c;
    return GR_OK;
}

int gr_checkpoint_destroy(gr_checkpoint* c) {
    if (!c) return GR_ERR_ARG;
    (void)hipSetDevice(c->eng->device);
    checkpoint_free(c);
    return GR_OK;
}

int gr_checkpoint_begin(gr_checkpoint* c) {
    if (!c) return GR_ERR_ARG;
    gr_engine* e = c->eng;
    int rc = use_device(e);
    if (rc) return rc;
    // the snapshot may be overwritten only once the previous transfer has read it
    if (c->pending) GR_HIP(e, hipStreamWaitEvent(e->stream, c->landed, 0));
    for (size_t i = 0; i < c->blocks.size(); ++i) {
        gr_level* L = c->blocks[i].first;
        const int b = c->blocks[i].second;
        GR_HIP(e, hipMemcpyAsync(c->dev + c->ofs[i], L->u[b], 22 * L->blk[b].stride * sizeof(double),
                                 hipMemcpyDeviceToDevice, e->stream));
    }
    GR_HIP(e, hipEventRecord(c->taken, e->stream));
    GR_HIP(e, hipStreamWaitEvent(c->copy, c->taken, 0));
    GR_HIP(e, hipMemcpyAsync(c->host, c->dev, c->total * sizeof(double), hipMemcpyDeviceToHost, c->copy));
    GR_HIP(e, hipEventRecord(c->landed, c->copy));
    c->pending = true;
    return GR_OK;
}

int gr_checkpoint_query(gr_checkpoint* c, int* done) {
    if (!c) return GR_ERR_ARG;
    gr_engine* e = c->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (!done) return fail(e, GR_ERR_ARG, "null output");
    if (!c->pending) {
        *done = 1;
        return GR_OK;
    }
    const hipError_t st = hipEventQuery(c->landed);
    if (st == hipErrorNotReady) {
        *done = 0;
        return GR_OK;
    }
    if (st != hipSuccess) return hipfail(e, st, "hipEventQuery");
    *done = 1;
    return GR_OK;
}

int gr_checkpoint_wait(gr_checkpoint* c) {
    if (!c) return GR_ERR_ARG;
    gr_engine* e = c->eng;
    int rc = use_device(e);
    if (rc) return rc;
    if (c->pending) GR_HIP(e, hipEventSynchronize(c->landed));
    return GR_OK;
}

int gr_checkpoint_nblocks(gr_checkpoint* c, int* nblocks) {
    if (!c) return GR_ERR_ARG;
    if (!nblocks) return fail(c->eng, GR_ERR_ARG, "null output");
    *nblocks = (int)c->blocks.size();
    return GR_OK;
}

int gr_checkpoint_block(gr_checkpoint* c, int index, const double** data, long* stride) {
    if (!c) return GR_ERR_ARG;
    if (!data || !stride) return fail(c->eng, GR_ERR_ARG, "null output");
    if (index < 0 || index >= (int)c->blocks.size()) return fail(c->eng, GR_ERR_ARG, "block index out of range");
    if (c->pending) {
        const hipError_t st = hipEventQuery(c->landed);
        if (st == hipErrorNotReady)
            return fail(c->eng, GR_ERR_STATE, "checkpoint transfer pending (gr_checkpoint_wait or _query first)");
        if (st != hipSuccess) return hipfail(c->eng, st, "hipEventQuery");
        c->pending = false;
    }
    *data = c->host + c->ofs[index];
    *stride = c->blocks[index].first->blk[c->blocks[index].second].stride;
    return GR_OK;
}

} // extern "C"
