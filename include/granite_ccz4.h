/*
 * granite_ccz4.h - C ABI of the MI355X CCZ4 right-hand-side and SSP-RK3 stage engine.
 *
 * This is the drop-in boundary for GRANITE's CPU path. Every entry point names the
 * reference interface it replaces. Conventions:
 *   - Storage of one block follows GridBlock (include/granite/core/grid.hpp:28-42,157-161):
 *     u[v*stride + i + nx*(j + ny*k)], nx = n[0]+2*ng, ny = n[1]+2*ng, nz = n[2]+2*ng,
 *     stride = nx*ny*nz, v in 0..21 in SpacetimeVar order (include/granite/core/types.hpp:104-127).
 *   - Host pointers passed in are borrowed for the duration of the call only.
 *   - Every call returns int: 0 = GR_OK, > 0 = a hipError_t, < 0 = an engine error (GR_ERR_*).
 *     gr_last_error() gives the text of the last failure on that engine.
 *   - One host thread per engine; all work of an engine runs in order on one HIP stream.
 *   - No torch, HIP or C++ types cross this boundary.
 */
#ifndef GRANITE_CCZ4_H
#define GRANITE_CCZ4_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_NUM_VARS 22

enum gr_status {
    GR_OK = 0,
    GR_ERR_ARG = -1,      /* null pointer or out-of-range argument */
    GR_ERR_GEOM = -2,     /* bad block geometry (n < 1, ng < 4, non-positive dx) */
    GR_ERR_NVARS = -3,    /* a variable count other than 22 */
    GR_ERR_NODEV = -4,    /* no usable gfx950 device / extension not initialised */
    GR_ERR_STATE = -5,    /* call not valid in the current state */
    GR_ERR_NONFINITE = -6 /* (reported by gr_level_nonfinite only) */
};

/* Ghost-fill modes. SOMMERFELD is the reference's outer BC after every RK stage
 * (src/main.cpp:351-521); OUTFLOW is the level-0 zero-gradient fill (src/amr/amr.cpp:436-480);
 * PERIODIC has no reference counterpart (used for the periodic gauge-wave box). */
enum gr_bc_mode { GR_BC_NONE = -1, GR_BC_SOMMERFELD = 0, GR_BC_PERIODIC = 1, GR_BC_OUTFLOW = 2 };

/* Arithmetic of the RK3 stage kernels. FP64 is the parity form. MIXED evaluates the
 * stencils and RHS algebra in FP32 and keeps FP64 in the RK3 state and combine, in
 * chi and the lapse, in the conformal-metric determinant/inverse and in the constraint
 * monitor. */
enum gr_precision { GR_PREC_FP64 = 0, GR_PREC_MIXED = 1 };

/* Block geometry as GridBlock holds it: n interior cells per axis, ng ghosts,
 * lo = lower corner of the interior, dx = (hi - lo) / n per axis (src/core/grid.cpp:22-24). */
typedef struct gr_geom {
    int n[3];
    int ng;
    double lo[3];
    double dx[3];
} gr_geom;

/* CCZ4Params (include/granite/spacetime/ccz4.hpp:25-34); fd_order and
 * use_moving_puncture are not read by the reference RHS and are not carried. */
typedef struct gr_ccz4_params {
    double kappa1;
    double kappa2;
    double eta;
    double ko_sigma;
    double chi_blend_center;
    double chi_blend_width;
} gr_ccz4_params;

typedef struct gr_engine gr_engine;
typedef struct gr_level gr_level;

/* ---- engine ------------------------------------------------------------------- */

/* Create an engine on HIP device `device` (its own stream, its kernels loaded).
 * Replaces: construction of spacetime::CCZ4Evolution (include/granite/spacetime/ccz4.hpp:51)
 * as far as device state is concerned; parameters travel with each call. */
int gr_engine_create(int device, gr_engine** out);
int gr_engine_destroy(gr_engine* eng);
/* Text of the last error on this engine (static storage owned by the engine). */
const char* gr_last_error(const gr_engine* eng);
/* Wait for all work queued on the engine's stream. */
int gr_engine_sync(gr_engine* eng);

/* ---- per-block calls on host memory (compatibility mode) ---------------------------- */

/* CCZ4 RHS plus Kreiss-Oliger dissipation on the interior of one block.
 * Replaces: CCZ4Evolution::computeRHS (src/spacetime/ccz4.cpp:251-865, declared
 * include/granite/spacetime/ccz4.hpp:69-74) and, with all four source pointers null,
 * CCZ4Evolution::computeRHSVacuum (src/spacetime/ccz4.cpp:229-245).
 * u: 22*stride doubles (ghosts filled). rhs: 22*stride doubles; only interior cells are
 * written, ghost cells keep their values. Sources (each totalSize = stride cells, all
 * null or all non-null): rho[stride], Si[3*stride] (AoS), Sij[6*stride] (AoS, xx,xy,xz,yy,yz,zz),
 * S[stride]. nvars must be 22. */
int gr_ccz4_rhs(gr_engine* eng, const gr_geom* geom, const gr_ccz4_params* prm, int nvars, const double* u,
                double* rhs, const double* rho, const double* Si, const double* Sij, const double* S);

/* Adds the Kreiss-Oliger term to rhs (interior); no-op when ko_sigma <= 0.
 * Replaces: CCZ4Evolution::applyDissipation (src/spacetime/ccz4.cpp:871-935). */
int gr_ccz4_dissipation(gr_engine* eng, const gr_geom* geom, const gr_ccz4_params* prm, int nvars,
                        const double* u, double* rhs);

/* Hamiltonian and momentum constraints on interior cells; ghost cells of ham/mom are 0.
 * ham: stride doubles; mom: 3*stride doubles (AoS).
 * Replaces: CCZ4Evolution::computeConstraints (src/spacetime/ccz4.cpp:941-1159). */
int gr_ccz4_constraints(gr_engine* eng, const gr_geom* geom, const double* u, double* ham, double* mom);

/* ---- per-block calls on device memory ------------------------------------------------ */

/* As gr_ccz4_rhs, but every pointer is device memory of this engine's device and the
 * call is asynchronous on the engine's stream. */
int gr_ccz4_rhs_device(gr_engine* eng, const gr_geom* geom, const gr_ccz4_params* prm, const double* d_u,
                       double* d_rhs, const double* d_rho, const double* d_Si, const double* d_Sij,
                       const double* d_S);

/* ---- a refinement level held on the device -------------------------------------------- */

/* A set of blocks evolved together (the blocks the reference passes to its per-level
 * callback EvolutionStepFunc, include/granite/amr/amr.hpp:43; body src/main.cpp:1142-1177).
 * Device memory per block: state u plus two stage buffers, 3 * 22 * stride doubles.
 * Blocks whose dx are bit-identical advance in one stage launch; blocks with another dx
 * (separately built boxes of one level can differ by an ulp) in a launch of their own, each
 * block with the stencil constants of its own dx, as the reference evaluates it. */
int gr_level_create(gr_engine* eng, int nblocks, const gr_geom* geoms, gr_level** out);
int gr_level_destroy(gr_level* lvl);
int gr_level_upload(gr_level* lvl, int block, const double* host_u);
int gr_level_download(gr_level* lvl, int block, double* host_u);
/* Device address of block `block`'s state (22*stride doubles) for device-resident callers. */
int gr_level_state(gr_level* lvl, int block, double** d_u);

/* Brill-Lindquist data on every cell of every block, ghosts included
 * (src/initial_data/initial_data.cpp:97-158). pos: 3*nbh doubles. */
int gr_level_init_brill_lindquist(gr_level* lvl, int nbh, const double* mass, const double* pos);
/* Adds amp * N(0,1) to each field whose bit is set in field_mask, on every cell. The
 * normal deviate for field v at flat cell p is a function of (seed, v*stride + p) only:
 * splitmix64 at counters 2q and 2q+1, then Box-Muller (cosine branch). */
int gr_level_add_noise(gr_level* lvl, uint64_t seed, double amp, uint32_t field_mask);

/* setGaugeWaveData (src/spacetime/ccz4.cpp:1190-1221) on every cell of every block:
 * H = A sin(2 pi x / L), chi = (1+H)^(-1/3), gt_xx = (1+H) chi, gt_yy = gt_zz = chi,
 * alpha = sqrt(1+H); all other fields 0. */
int gr_level_init_gauge_wave(gr_level* lvl, double amplitude, double wavelength);

/* Fill the ghost cells of every block's state with `mode`. */
int gr_level_fill_ghosts(gr_level* lvl, int mode);

/* Matter sources of block `block` for the stage kernels of this level (device memory of the
 * engine's device, borrowed until replaced): rho[stride], Si[3*stride] (AoS), Sij[6*stride]
 * (AoS, xx,xy,xz,yy,yz,zz), S[stride], the arrays grmhd.computeMatterSources fills and
 * sspRK3Step hands to computeRHS (src/main.cpp:295-301). All four null = vacuum (the default).
 * The arrays are read by every stage launched after this call; a caller that recomputes
 * the sources from each stage's state, as the reference does, writes them between
 * gr_level_rk3_stage calls. Waits for the level's queued work. */
int gr_level_set_sources(gr_level* lvl, int block, const double* d_rho, const double* d_Si, const double* d_Sij,
                         const double* d_S);

/* One SSP-RK3 step of every block of the level, in the reference order of
 * TimeIntegrator::sspRK3Step (src/main.cpp:523-689): three (RHS + dissipation + stage
 * combine + floors) stages, the algebraic constraints after the last, and the ghost
 * fill `bc_mode` after each stage (GR_BC_SOMMERFELD reproduces the reference).
 * pre_fill (a gr_bc_mode or GR_BC_NONE) is applied to the state before the step, as the
 * reference's AMRHierarchy::subcycle does on level 0 (src/amr/amr.cpp:319).
 * use_graph != 0 replays the step from a captured HIP graph (captured on first use for a
 * given (dt, params, modes)). Asynchronous: ordered on the engine stream. */
int gr_level_rk3_step(gr_level* lvl, const gr_ccz4_params* prm, double dt, int precision, int bc_mode,
                      int pre_fill, int use_graph);

/* One stage (stage_index 1, 2 or 3) of the step gr_level_rk3_step runs, with the same
 * buffers, weights, floors, algebraic constraints (stage 3) and ghost fill `bc_mode` of the
 * buffer it writes; pre_fill as there (stage 1 only, normally). The seam for callers that
 * recompute the matter sources from each stage's state between stages
 * (TimeIntegrator::sspRK3Step, src/main.cpp:243-304,673-689). Asynchronous, not captured. */
int gr_level_rk3_stage(gr_level* lvl, const gr_ccz4_params* prm, double dt, int stage_index, int precision,
                       int bc_mode, int pre_fill);

/* Mixed-precision policy: tiles whose smallest chi is below chi_fp64_below (or not
 * finite) run the FP64 stage kernel; all other tiles the FP32 one. Default 0.2. */
int gr_level_set_mixed_policy(gr_level* lvl, double chi_fp64_below);

/* One SSP-RK3 step of n independent levels of one engine, level i with dts[i] (the batched
 * uneven-block workload: every refinement level advanced by its own step): the same
 * launches as gr_level_rk3_step on each level, scheduled over two streams so that each
 * level's ghost fill (memory-bound) overlaps the following levels' stage kernels
 * (compute-bound); each level's own stages and fills keep their order (events). Results are
 * bit-identical to stepping the levels one by one. use_graph: the whole schedule as one
 * captured HIP graph, re-captured when the level set, a dt, the parameters or a level's
 * tuning change. */
int gr_levels_rk3_step(int n, gr_level* const* levels, const double* dts, const gr_ccz4_params* prm, int precision,
                       int bc_mode, int pre_fill, int use_graph);

/* RHS (with dissipation) of block `block`'s current state in the given precision, written
 * to host_rhs (22*stride doubles, interior cells; ghost cells set to 0). The FP64 form is
 * CCZ4Evolution::computeRHSVacuum (src/spacetime/ccz4.cpp:229-245) on device-resident data;
 * the MIXED form is the arithmetic the mixed-precision RK3 stages use. */
int gr_level_rhs(gr_level* lvl, int block, const gr_ccz4_params* prm, int precision, double* host_rhs);

/* Wait for the level's queued work. */
int gr_level_sync(gr_level* lvl);

/* Sum of H^2 over the interiors of all blocks and the number of cells summed, with the
 * FP64 monitor (the reference prints sqrt(sumsq/count), src/main.cpp:1216-1250). */
int gr_level_constraints_l2(gr_level* lvl, double* sumsq, long long* count);

/* Allocation-free constraint monitor: sum of H^2 over the interior of one block whose state
 * the caller holds in device memory (GridBlock layout), written to the device double
 * *d_sumsq; *count = interior cells. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops). It
 * calls no hipMalloc/hipFree: the reduction partials live in a fixed device array of 2048
 * doubles per GPU, shared by every call on that GPU (calls that overlap on one GPU race;
 * each call rewrites what it reads). Enqueued on the engine stream, not synchronised
 * (gr_engine_sync). Blocks over 2^29 cells with ghosts are rejected (GR_ERR_GEOM). */
int gr_ccz4_constraints_l2_noalloc(gr_engine* eng, const gr_geom* geom, const double* d_u, double* d_sumsq,
                                   long long* count);

/* Number of non-finite values in the interiors of all blocks' state (NaN scan of
 * src/main.cpp:1192-1214). */
int gr_level_nonfinite(gr_level* lvl, long long* count);

/* The first non-finite interior value in the reference's scan order (src/main.cpp:1192-1214:
 * blocks in order, then field v, then k, j, i): *found = 1 and its block, field, cell
 * (counted with ghosts, as the reference prints it) and value; *found = 0 when every interior
 * value is finite. Synchronous. */
int gr_level_first_nonfinite(gr_level* lvl, int* found, int* block, int* var, int* i, int* j, int* k,
                             double* value);

/* Advective CFL of the state, max over the interiors of all blocks of
 * |beta^x| dt/dx + |beta^y| dt/dy + |beta^z| dt/dz (NaN cells ignored), the quantity of the
 * reference's per-substep guard (src/main.cpp:1155-1176). Synchronous. */
int gr_level_cfl(gr_level* lvl, double dt, double* max_cfl);

/* One value of a block's device state: field var at cell (i, j, k) counted with ghosts, as
 * GridBlock::data(var, i, j, k) (include/granite/core/grid.hpp:107-114), e.g. the lapse at
 * the grid centre that the telemetry line prints (src/main.cpp:1238-1244). Synchronous. */
int gr_level_read_cell(gr_level* lvl, int block, int var, int i, int j, int k, double* value);

/* ---- Berger-Oliger hierarchy of device levels ---------------------------------------------
 *
 * A hierarchy is an ordered list of levels (coarsest first, refinement ratio 2, every level's
 * blocks created with gr_level_create on one engine), playing the part of
 * amr::AMRHierarchy (include/granite/amr/amr.hpp:49-117) for a fixed geometry: regridding
 * stays with the caller, who re-creates the levels and the hierarchy after it. The hierarchy
 * borrows the levels; destroy it before them. Level 0 is the domain block(s). */
typedef struct gr_hierarchy gr_hierarchy;

/* Fails with GR_ERR_GEOM unless dx of level l+1 is dx of level l / 2 (relative 1e-6, the
 * reference's regrid check at src/amr/amr.cpp:280-287) and GR_ERR_ARG unless all levels
 * share the engine. */
int gr_hierarchy_create(gr_engine* eng, int nlevels, gr_level* const* levels, gr_hierarchy** out);
int gr_hierarchy_destroy(gr_hierarchy* h);

/* AMRHierarchy::prolongate of every block of level-1 into every cell (ghosts included) of
 * every block of `level` >= 1, in the reference's block order (src/amr/amr.cpp:484-564);
 * what regrid does to a new level (src/amr/amr.cpp:290-293). */
int gr_hierarchy_prolongate(gr_hierarchy* h, int level);
/* AMRHierarchy::restrict_data of every block of `level` >= 1 into level-1
 * (src/amr/amr.cpp:566-626). */
int gr_hierarchy_restrict(gr_hierarchy* h, int level);
/* AMRHierarchy::fillGhostZones(level) (src/amr/amr.cpp:400-482): level 0 outflow copy on
 * all faces; level >= 1 prolongation from the parent level into all cells, then the outflow
 * copy on the faces no parent block covers. */
int gr_hierarchy_fill_ghosts(gr_hierarchy* h, int level);

/* One coarse step: AMRHierarchy::subcycle(0, evolve) (src/amr/amr.cpp:303-350) with dt0 on
 * level 0 and dt0 / 2^l on level l (setLevelDt, src/amr/amr.cpp:389-398), where evolve is
 * the per-level SSP-RK3 step with 1/r ghosts after every stage (gr_level_rk3_step with
 * GR_BC_SOMMERFELD and no pre-fill). Reproduces the reference's order exactly, including
 * the re-prolongation of all fine cells after each fine step; no regrid. use_graph != 0
 * replays the whole coarse step from one HIP graph captured on first use for the given
 * (dt0, parameters, precision). */
int gr_hierarchy_step(gr_hierarchy* h, const gr_ccz4_params* prm, double dt0, int precision, int use_graph);

/* Constraint telemetry over the interiors of all blocks of all levels
 * (src/main.cpp:1216-1250): sum of H^2, sum of |M|^2 (M_i the momentum constraint of
 * computeConstraints) and the number of cells; ||H||_2 = sqrt(ham_sumsq / count). Either
 * sum pointer may be null. */
int gr_hierarchy_constraints_l2(gr_hierarchy* h, double* ham_sumsq, double* mom_sumsq, long long* count);

/* enable != 0: every coarse step also evaluates, after each level's RK3 step, the advective
 * CFL guard of the reference's evolve callback (src/main.cpp:1155-1176) on the device (inside
 * the captured graph). gr_hierarchy_cfl returns the values of the last coarse step, one per
 * level step in subcycle order, with each one's level; *n = the number of level steps per
 * coarse step (sum of 2^l). The reference prints "[CFL-GUARD] Advection CFL=v > 0.95 at
 * sub-step!" for each value above 0.95. */
int gr_hierarchy_set_cfl_monitor(gr_hierarchy* h, int enable);
int gr_hierarchy_cfl(gr_hierarchy* h, double* values, int* levels, int cap, int* n);

/* ---- checkpoint staging ---------------------------------------------------------------------
 * The data HDF5Writer::writeCheckpoint (src/io/hdf5_io.cpp:166-232) writes at
 * step % checkpoint_interval == 0 or t >= t_final (src/main.cpp:1256-1261): every block of the
 * hierarchy, coarsest level first (AMRHierarchy::getAllBlocks), each var v as rawData(v), the
 * whole stride ghosts included. gr_checkpoint_begin snapshots the device state into a device
 * copy, ordered on the engine stream after the work already enqueued (a few hundred us for
 * a GB), and moves that copy into pinned host memory on a copy stream of its own, so the
 * evolution goes on while the transfer runs. */
typedef struct gr_checkpoint gr_checkpoint;

/* Snapshot buffers (device and pinned host, 22 * stride doubles per block) for every block of
 * levels[0..nlevels), in that order. The levels must outlive the checkpoint. */
int gr_checkpoint_create(gr_engine* eng, int nlevels, gr_level* const* levels, gr_checkpoint** out);
/* Waits for a pending transfer, then frees the buffers. */
int gr_checkpoint_destroy(gr_checkpoint* c);
/* Snapshot every block now (in engine-stream order) and start the transfer; returns without
 * waiting. A begin while the previous transfer runs is ordered after it on the device; the
 * caller must be done reading the previous host views. */
int gr_checkpoint_begin(gr_checkpoint* c);
/* *done = 1 when the last begin's transfer has completed (0 while it runs). */
int gr_checkpoint_query(gr_checkpoint* c, int* done);
/* Blocks until the last begin's transfer has completed. */
int gr_checkpoint_wait(gr_checkpoint* c);
/* Number of blocks, and the host view of block `index` (valid after gr_checkpoint_wait until
 * the next begin): var v is data[v * stride .. v * stride + stride), GridBlock::rawData(v).
 * Fails with GR_ERR_STATE while the last begin's transfer is still running. */
int gr_checkpoint_nblocks(gr_checkpoint* c, int* nblocks);
int gr_checkpoint_block(gr_checkpoint* c, int index, const double** data, long* stride);

/* ---- measurement (no reference counterpart) ----------------------------------------------- */

/* enable != 0: every RK3 step run without a graph records HIP events around each of its
 * kernels on the engine stream (graph replay is not instrumented). */
int gr_level_profile(gr_level* lvl, int enable);
/* Waits for the stream, then returns and clears the accumulated device time of one kernel
 * kind since the last read: kind 0 = stage kernels of RK3 stages 1-2, 1 = the last-stage
 * kernel (with the algebraic constraints), 2 = ghost fills. */
int gr_level_profile_read(gr_level* lvl, int kind, double* total_ms, int* launches);

/* ---- autotuning (no reference counterpart: the reference's only schedule is a fixed OpenMP
 * collapse(3) schedule(static), src/spacetime/ccz4.cpp:266-268) ---------------------------- */

/* Tuned choices. GR_TUNE_STAGE: the FP64 stage kernel's halo staging, 0 = prefetch through
 * registers into a padded LDS box (8x4x4-cell parts, one wave per SIMD), 1 = direct
 * global-to-LDS loads into an unpadded box (8x4x4, two waves per SIMD), 2 = as 1 with
 * 8x4x8-cell parts (less halo per cell); 1 and 2 need even row lengths (others run 0).
 * GR_TUNE_GHOST / GR_TUNE_TRANSFER: field groups (1, 2, 4 or 11) of the ghost fill and of the
 * prolongation and restriction launches. */
enum gr_tune_kind { GR_TUNE_STAGE = 0, GR_TUNE_GHOST = 1, GR_TUNE_TRANSFER = 2 };

/* Times every variant of the stage kernel (per group of equal spacing: stage 1 into the
 * first stage buffer, the state untouched) and of the ghost fill (into that buffer's ghosts)
 * on this level's own blocks, keeps the fastest in the level and in the process-wide tuning
 * table under the level's block-size class (mean interior cells per block: < 24^3, < 48^3,
 * < 96^3, larger). times (optional, cap entries) receives the measured ms per launch in the
 * order: for each group, each stage variant 0, 1, 2 (-1 where not applicable); then each ghost group
 * count (1, 2, 4, 11); *n = entries written. Captured graphs of the level are re-recorded. */
int gr_level_autotune(gr_level* lvl, const gr_ccz4_params* prm, double dt, double* times, int cap, int* n);
/* gr_level_autotune of every level (dt0 / 2^l), then the transfer field groups of every
 * level l > 0 (prolongation into the level's first stage buffer, restriction into the
 * parent's, the states untouched). times: the levels' entries, then per level l > 0 the
 * four group counts. */
int gr_hierarchy_autotune(gr_hierarchy* h, const gr_ccz4_params* prm, double dt0, double* times, int cap, int* n);
/* The choice in force for group `group` of the level (GR_TUNE_STAGE; group < number of
 * groups of equal spacing) or for the level (GR_TUNE_GHOST), and forcing one (tests and
 * experiments; an unknown value fails with GR_ERR_ARG). */
int gr_level_get_tuning(gr_level* lvl, int kind, int group, int* value);
int gr_level_set_tuning(gr_level* lvl, int kind, int group, int value);
/* Transfer field groups of level l > 0 of a hierarchy (get / force). */
int gr_hierarchy_get_tuning(gr_hierarchy* h, int level, int* groups);
int gr_hierarchy_set_tuning(gr_hierarchy* h, int level, int groups);
/* The process-wide tuning table: write to / read from a text file (one line per entry:
 * kind class precision value), or empty it. Levels and hierarchies created afterwards use
 * the entries of their class; absent entries keep the defaults (staging 2 for a group of
 * blocks with even row lengths and at least 40^3 interior cells in all, else 0; 4 groups). */
int gr_tune_save(const char* path);
int gr_tune_load(const char* path);
int gr_tune_clear(void);

/* ---- introspection ------------------------------------------------------------------- */

/* Stage-kernel tile shape in cells (x, y, z) used for the given precision: FP64, the part of
 * a tile one workgroup runs; mixed, the whole tile (the unit of the precision choice). */
int gr_stage_tile(int precision, int* tile3);
/* Library version string. */
const char* gr_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GRANITE_CCZ4_H */
