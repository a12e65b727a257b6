// Internal launch interface between the C ABI (capi.hip) and the gfx950 kernels
// (kernels.hip). Not part of the public boundary.
#pragma once

#include <hip/hip_runtime.h>

#include "ccz4_cell.hpp"

#include <vector>

namespace gr {

// one block as the kernels see it
struct BlockDesc {
    int n[3];
    int nt[3];
    int ng;
    long stride;
    double lo[3];
    double dx[3];
};

enum StageMode : int { MODE_RHS = 0, MODE_STAGE = 1, MODE_FINAL = 2 };

// Matter sources of one block (device arrays over all cells: rho[stride], S_i[3*stride]
// and S_ij[6*stride] cell-major, S[stride]); all null = vacuum.
struct Sources {
    const double* rho;
    const double* Si;
    const double* Sij;
    const double* S;
};

// One block of a batched launch: its geometry, its three state buffers (state u and the
// two RK3 stage buffers, or for a single-block call: input, u^n, output), its matter
// sources and the index of its first stage tile in the launch.
struct BatchBlock {
    BlockDesc b;
    double* buf[3];
    Sources src;
    int tile0;
    int tiles[3];
    int ghost_wg0; // first workgroup of the block in a batched ghost launch
};

// workgroup size of the ghost launches
constexpr int GHOST_WG = 256;

// What a stage / ghost launch runs over: a device table of blocks (table != nullptr,
// with tile_block[t] = block of global tile t and ghost_wg_block[w] = block of ghost
// workgroup w; each block's ghost cells start on a workgroup of their own) or the single
// block `one`. sel_* pick the buffer of
// each block that is the stage input, u^n and the output (0, 1, 2 = buf[0..2]).
// FP64 staging variants of the stage kernel (chosen per launch, by the autotuner): 0 = halo
// prefetched into registers and written to a padded LDS box (parts of 8x4x4 cells, one wave
// per SIMD); 1 = halo loaded straight into an unpadded LDS box (8x4x4, two waves per SIMD);
// 2 = as 1 with parts of 8x4x8 (less halo per cell). 1 and 2 need even row lengths.
constexpr int FP64_STAGE_VARIANTS = 3;

struct Batch {
    const BatchBlock* table;
    const int* tile_block;
    const int* ghost_wg_block;
    int nblocks;
    BatchBlock one;
    int ntiles;
    int ghost_wgs; // workgroups of a ghost launch
    double dx[3]; // grid spacing shared by all blocks of the launch
    int pairs_ok; // every block: even row length (nt0) and 16-byte aligned buffers
    int variant;  // FP64 staging variant (0 registers, 1 direct LDS; 1 needs pairs_ok)
    int ghost_groups; // field groups of the ghost fill launch (1, 2, 4 or 11)
};
// whether block b's rows can be read as 16-byte pairs from buffer p
bool pairs_aligned(const BlockDesc& b, const void* p);

// builds the tile bookkeeping of a block list (host side): tile0/tiles and ghost_wg0 of
// each block, the tile->block map and the ghost-workgroup->block map. Returns the total
// number of tiles.
int batch_layout(std::vector<BatchBlock>& blocks, std::vector<int>& tile_block, std::vector<int>& ghost_wg_block);
// a one-block batch over (in, un, out), with optional matter sources
Batch single_batch(const BlockDesc& b, const double* in, const double* un, double* out,
                   const Sources& src = Sources{nullptr, nullptr, nullptr, nullptr});

struct StageCoef {
    double a1;   // weight of u^n (read pointwise from `un`)
    double a2;   // weight of the stage input (centre of `in`)
    double a3dt; // a3 * dt, weight of the RHS
};

// stencil constants for a block, with the reference's operation order
Stencil<double> make_stencil(const BlockDesc& b, double ko_sigma);

// RHS / RK3 stage kernel (FP64) over every tile of a batch. sel = {input, u^n, output}
// buffer indices; input ghosts must be filled; u^n is ignored for MODE_RHS; the output
// receives the RHS (MODE_RHS) or the new stage state. has_src: some block of the batch
// carries matter sources (blocks whose sources are null are vacuum).
hipError_t launch_stage_fp64(const Batch& B, const int sel[3], bool has_src, const CellParams& P,
                             double ko_sigma, int mode, const StageCoef& cf, hipStream_t s,
                             const unsigned char* flags = nullptr);

// A second stream (default priority) with the two events that fork work onto it from the
// main stream and join it back.
struct AuxStream {
    hipStream_t stream;
    hipEvent_t fork, join;
};

// Mixed-precision stage: classifies the tiles of the input (FP64 where the smallest chi of
// the tile is below chi_fp64_below), then runs the FP32 kernel on the other tiles and the
// FP64 kernel on those (on `aux` beside the FP32 kernel when given, else after it on s).
// flags: device buffer of B.ntiles bytes.
hipError_t launch_stage_mixed(const Batch& B, const int sel[3], bool has_src, const CellParams& P,
                              double ko_sigma, int mode, const StageCoef& cf, hipStream_t s, unsigned char* flags,
                              double chi_fp64_below, const AuxStream* aux = nullptr);
// number of stage tiles of a block
int stage_tiles(const BlockDesc& b);
// ghost fill of buffer `sel` of every block of a batch (mode as launch_fill_ghosts)
hipError_t launch_fill_ghosts_batch(const Batch& B, int sel, int mode, hipStream_t s);

// Ghost fill of one block: mode 0 Sommerfeld (chained X,Y,Z), 1 periodic, 2 outflow copy
// of the faces set in `faces` (bit 2d low side, 2d+1 high side of axis d).
hipError_t launch_fill_ghosts(const BlockDesc& b, double* u, int mode, hipStream_t s, int faces = 0x3f,
                              int groups = 4);
// field-group counts the ghost fill is compiled for
bool ghost_groups_valid(int g);

// one axis of the trilinear coarse -> fine stencil at one fine index: the two coarse
// neighbours and the weight of c1
struct ProlongAxis {
    int c0, c1;
    double ax;
};

// stencil table of fine block f inside coarse block c: f.nt[0] + f.nt[1] + f.nt[2] entries
// (the x indices, then y, then z)
hipError_t launch_prolong_axes(const BlockDesc& c, const BlockDesc& f, ProlongAxis* tab, hipStream_t s);
// trilinear coarse -> fine into every fine cell (ghosts included) from that table, and 2x2x2
// average of the fine block into the covered coarse interior cells
hipError_t launch_prolongate(const BlockDesc& c, const double* cu, const BlockDesc& f, double* fu,
                             const ProlongAxis* tab, hipStream_t s, int groups = 0);
hipError_t launch_restrict(const BlockDesc& f, const double* fu, const BlockDesc& c, double* cu, hipStream_t s,
                           int groups = 0);

// rhs += Kreiss-Oliger(u) on the interior (reference accumulation order)
hipError_t launch_dissipation(const BlockDesc& b, const double* u, double* rhs, double ko_sigma, hipStream_t s);

// H and M_i on interior cells; ghosts of ham/mom are zeroed. mom may be null.
hipError_t launch_constraints(const BlockDesc& b, const double* u, double* ham, double* mom, hipStream_t s);

// Brill-Lindquist initial data on all cells. mass/pos: device arrays (nbh, 3*nbh).
hipError_t launch_init_bl(const BlockDesc& b, double* u, int nbh, const double* mass, const double* pos,
                          hipStream_t s);

hipError_t launch_add_noise(const BlockDesc& b, double* u, unsigned long long seed, double amp, unsigned mask,
                            hipStream_t s);

// partial[blockIdx] = sum over interior cells of the squares of the ncomp values per cell
// (ncomp 1: ham, 3: mom); count is analytic. Returns the number of partials written in
// *nparts (<= cap).
hipError_t launch_sumsq_interior(const BlockDesc& b, const double* ham, int ncomp, double* partial, int cap,
                                 int* nparts, hipStream_t s);
// sums `n` doubles into out[0] with one workgroup (fixed order)
hipError_t launch_sum_partials(const double* partial, int n, double* out, hipStream_t s);
// number of partials the interior reductions split a block into (at most `cap`)
int reduce_parts(const BlockDesc& b, int cap);
// partial sums of H^2 over the interior (H formed on the fly), at most `cap` partials
hipError_t launch_ham_sumsq(const BlockDesc& b, const double* u, double* partial, int cap, int* nparts,
                            hipStream_t s);
// partial[blockIdx] = number of non-finite interior values over all 22 fields
hipError_t launch_count_nonfinite(const BlockDesc& b, const double* u, double* partial, int cap, int* nparts,
                                  hipStream_t s);
// First non-finite interior value of a block in the reference's scan order (field, then k, j,
// i): partial[blockIdx] = smallest position q = v * ncell + (k * n1 + j) * n0 + i (interior-
// relative) of a non-finite value in the workgroup's range, or ~0 if none; then
// launch_min_partials reduces them into out[0].
hipError_t launch_first_nonfinite(const BlockDesc& b, const double* u, unsigned long long* partial, int cap,
                                  int* nparts, hipStream_t s);
hipError_t launch_min_partials(const unsigned long long* partial, int n, unsigned long long* out, hipStream_t s);
// Advective CFL of the shift over the interior, sum_d |beta^d| * dt / dx_d, as the reference's
// per-substep guard computes it (max over cells, NaN ignored): partial[blockIdx] = the
// workgroup's maximum; launch_max_partials reduces n partials into out[0].
hipError_t launch_cfl_max(const BlockDesc& b, const double* u, double dt, double* partial, int cap, int* nparts,
                          hipStream_t s);
hipError_t launch_max_partials(const double* partial, int n, double* out, hipStream_t s);
// setGaugeWaveData (src/spacetime/ccz4.cpp:1190-1221) on every cell; other fields set to 0
hipError_t launch_init_gauge_wave(const BlockDesc& b, double* u, double A, double L, hipStream_t s);

// tile shape of the FP64 stage kernel
void stage_tile_fp64(int* t3);
// tile of the launch bookkeeping and of the mixed-precision kernel (cells x, y, z)
void stage_tile_mixed(int* t3);

} // namespace gr
