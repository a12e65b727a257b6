// gfx950 kernels of the CCZ4 RHS + SSP-RK3 stage engine (FP64 form).
//
// Stage kernel layout: one workgroup per tile of TX x TY x TZ interior cells, one thread
// per cell. The 22 fields are streamed one at a time through an LDS box holding the tile
// plus the stencil halo (4 cells along each axis, 1 cell on the edges used by the mixed
// second derivatives); each field is read from HBM once per tile in coalesced rows along x.
// The per-cell arithmetic is ccz4_cell.hpp; the RK3 combine, the floors and (last stage)
// the algebraic constraints are applied to each value before it is stored, so one launch
// per stage reads the stage input and u^n and writes the new stage state.
#include "kernels.hpp"

#include <cstdint>

// tile shape (overridable at build time): a tile of GR_TILE_X x Y x Z cells is the unit of
// the launch bookkeeping and of the mixed-precision classification and kernel (one thread
// per cell); the FP64 kernel, whose registers allow one wave per SIMD, runs each tile as
// GR_FP64_SPLIT x GR_FP64_ZSPLIT workgroups of GR_TILE_X / GR_FP64_SPLIT x Y x
// GR_TILE_Z / GR_FP64_ZSPLIT cells
#ifndef GR_TILE_X
#define GR_TILE_X 16
#endif
#ifndef GR_TILE_Y
#define GR_TILE_Y 4
#endif
#ifndef GR_TILE_Z
#define GR_TILE_Z 8
#endif
#ifndef GR_FP64_ZSPLIT
#define GR_FP64_ZSPLIT 2
#endif
#ifndef GR_FP64_SPLIT
#define GR_FP64_SPLIT 2
#endif
#ifndef GR_FP64_PREFETCH
#define GR_FP64_PREFETCH 1
#endif
#ifndef GR_FP64_KO_EXACT
#define GR_FP64_KO_EXACT 1
#endif
#ifndef GR_FP64_MIN_WAVES
#define GR_FP64_MIN_WAVES 1
#endif
// parts of a tile along x the mixed kernel runs as separate workgroups (1 or 2)
#ifndef GR_MIXED_SPLIT
#define GR_MIXED_SPLIT 1
#endif
#ifndef GR_MIXED_MIN_WAVES
#define GR_MIXED_MIN_WAVES 1
#endif
#ifndef GR_LDS_PAD
#define GR_LDS_PAD 1
#endif
// LDS layout of the FP64 parts: 0 = rows padded to a conflict-free pitch, lanes x-fastest
// (55 KB for two buffers of an 8x4x4 part: two workgroups per CU); 1 = unpadded rows,
// planes padded, lanes interleaved over y and z (38 KB: room for four workgroups per CU,
// i.e. two waves per SIMD when the registers allow it)
// field buffers of the direct-to-LDS staging, at most (2 or 3; fewer where more would not
// leave two waves per SIMD room in the CU's LDS); the last stage's own limit
#ifndef GR_DMA_BUFFERS
#define GR_DMA_BUFFERS 3
#endif
#ifndef GR_DMA_BUFFERS_FINAL
#define GR_DMA_BUFFERS_FINAL 2
#endif
#ifndef GR_MIXED_PREFETCH
#define GR_MIXED_PREFETCH 1
#endif
#ifndef GR_LDS_PAD32
#define GR_LDS_PAD32 1
#endif
// default field groups (second grid dimension) of the ghost fill (1, 2, 4 or 11)
#ifndef GR_GHOST_GROUPS
#define GR_GHOST_GROUPS 4
#endif
// field groups (second grid dimension) of the prolongation and restriction kernels
#ifndef GR_TRANSFER_GROUPS
#define GR_TRANSFER_GROUPS 4
#endif

namespace gr {

// Device buffers reached through tables (block descriptors, stage selectors) are generic
// pointers to the compiler, which would access them with flat instructions; flat accesses
// also count against the LDS wait counter, so every LDS wait would wait for the in-flight
// prefetch of the next field as well. The stage kernels access HBM through this type instead.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) char gchar;

// element at byte offset `off` (< 4 GiB) from a workgroup-uniform base
__device__ __forceinline__ double gload(const gdouble* base, unsigned off) {
    return *(const gdouble*)((const gchar*)base + off);
}
__device__ __forceinline__ void gstore(gdouble* base, unsigned off, double v) {
    *(gdouble*)((gchar*)base + off) = v;
}
// the same load with the offset made opaque at the point of use, so that the 32-bit offset
// reaches the load itself (scalar base + 32-bit vector offset, no 64-bit address add per
// load) instead of a 64-bit copy of it hoisted out of the loop over fields
__device__ __forceinline__ double gload_sv(const gdouble* base, unsigned off) {
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+v"(off));
#endif
    return *(const gdouble*)((const gchar*)base + off);
}

namespace {

constexpr int HALO = 4;

// ------------------------------------------------------------------------------------
// exact (contraction-free) pointwise pieces shared with the reference
// ------------------------------------------------------------------------------------

// d = a1*s1 + a2*s2 + a3*dt*r, src/main.cpp:328-329
__device__ __forceinline__ double combine_exact(double a1, double s1, double a2, double s2, double a3dt, double r) {
#pragma clang fp contract(off)
    return a1 * s1 + a2 * s2 + a3dt * r;
}

__device__ __forceinline__ double floor_chi(double c) {
    // src/main.cpp:598-602
    if (!__builtin_isfinite(c) || c < 1.0e-4) c = 1.0e-4;
    if (c > 1.5) c = 1.5;
    return c;
}

__device__ __forceinline__ double floor_alpha(double a) {
    // src/main.cpp:603-605
    if (!__builtin_isfinite(a) || a < 1.0e-6) a = 1.0e-6;
    return a;
}

// det(gt) -> 1 by a cube-root rescale (or reset to identity), then remove the trace of A~
// with the cofactors of the rescaled metric, src/main.cpp:622-668
__device__ __forceinline__ void algebraic_exact(double g[6], double A[6]) {
#pragma clang fp contract(off)
    double &xx = g[0], &xy = g[1], &xz = g[2], &yy = g[3], &yz = g[4], &zz = g[5];
    const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    if (__builtin_isfinite(det) && det > 1.0e-10) {
        const double s = cbrt(1.0 / det);
        for (int a = 0; a < 6; ++a) g[a] *= s;
    } else {
        xx = 1.0; xy = 0.0; xz = 0.0; yy = 1.0; yz = 0.0; zz = 1.0;
    }
    const double cxx = yy * zz - yz * yz;
    const double cxy = xz * yz - xy * zz;
    const double cxz = xy * yz - xz * yy;
    const double cyy = xx * zz - xz * xz;
    const double cyz = xz * xy - xx * yz;
    const double czz = xx * yy - xy * xy;
    const double tr = cxx * A[0] + cyy * A[3] + czz * A[5] + 2.0 * (cxy * A[1] + cxz * A[2] + cyz * A[4]);
    const double t3 = tr / 3.0;
    for (int a = 0; a < 6; ++a) A[a] -= t3 * g[a];
}

// ------------------------------------------------------------------------------------
// LDS staging of one field of a tile
// ------------------------------------------------------------------------------------
// Smallest LDS row pitch >= bx0 (elements of esize bytes, rows of a box by rows high) at
// which the 64 lanes of a wave, laid over the tile x fastest, then y, then z, read their
// cells without bank conflicts. The LDS serves 64 four-byte banks per pass, so a wave's
// read of esize-byte elements takes esize / 4 passes of 256 / esize consecutive lanes, each
// conflict-free when its words fall on distinct banks. The stencil taps shift all lanes
// alike, so they inherit the pattern of the centre read.
template <int TX, int TY>
constexpr int conflict_free_pitch(int bx0, int by, int esize) {
    const int words = esize / 4, lanes = 64 / words;
    for (int p = bx0; p < bx0 + 64; ++p) {
        bool clean = true;
        for (int pass = 0; pass < words && clean; ++pass) {
            bool used[64] = {};
            for (int t = pass * lanes; t < (pass + 1) * lanes && clean; ++t) {
                const int lx = t % TX, ly = (t / TX) % TY, lz = t / (TX * TY);
                const int a = (lz * by + ly) * p + lx;
                for (int w = 0; w < words; ++w) {
                    const int bank = (a * words + w) % 64;
                    clean = clean && !used[bank];
                    used[bank] = true;
                }
            }
        }
        if (clean) return p;
    }
    return bx0;
}

// Lane -> cell maps of a tile's threads. Layout 0: x fastest, then y, then z. Layout 1
// (TX = 8, TY even, TZ a multiple of 4): lane bits 0-2 are x, bit 3 the low bit of z, bit 4
// the low bit of y, bit 5 the second bit of z, and the wave index supplies the rest of y,
// then of z. Unpadded rows (pitch TX + 8 = 16 doubles) then read conflict-free when a
// plane spans 8 doubles modulo 16 (a pad of 8 doubles per plane of 192): the 16 lanes of a
// 128-byte group of a 64-bit read are 8 x-neighbours in 2 planes (banks 0-15 and 16-31 of 32),
// and the 32 lanes of a 256-byte group add 2 rows (banks 32-63 of 64).
template <int TX, int TY, int TZ, int LAYOUT>
struct LaneMap {
    static_assert(LAYOUT == 0 || (TX == 8 && TY % 2 == 0 && TZ % 4 == 0), "layout 1 needs 8 x even x 4k parts");
    __host__ __device__ static constexpr void cell(int tid, int& x, int& y, int& z) {
        if (LAYOUT == 0) {
            x = tid % TX;
            y = (tid / TX) % TY;
            z = tid / (TX * TY);
        } else {
            const int w = tid >> 6;
            x = tid & 7;
            z = ((tid >> 3) & 1) + 2 * ((tid >> 5) & 1) + 4 * (w / (TY / 2));
            y = ((tid >> 4) & 1) + 2 * (w % (TY / 2));
        }
    }
};

// whether a wave's read of its cells (elements of esize bytes, row pitch P, plane stride Q)
// is free of bank conflicts, modelled both as 16-lane groups of 32 four-byte banks and as
// 32-lane groups of 64 banks for 8-byte elements (64-lane groups of 64 banks for 4-byte ones)
template <int TX, int TY, int TZ, int LAYOUT>
constexpr bool lds_conflict_free(int P, int Q, int esize) {
    const int words = esize / 4, NT = TX * TY * TZ;
    for (int banks = 32; banks <= 64; banks *= 2) {
        const int lanes = banks / words;
        for (int base = 0; base < NT; base += lanes) {
            bool used[64] = {};
            for (int t = base; t < base + lanes && t < NT; ++t) {
                int x = 0, y = 0, z = 0;
                LaneMap<TX, TY, TZ, LAYOUT>::cell(t, x, y, z);
                const int a = z * Q + y * P + x;
                for (int w = 0; w < words; ++w) {
                    const int bank = (a * words + w) % banks;
                    if (used[bank]) return false;
                    used[bank] = true;
                }
            }
        }
    }
    return true;
}

// smallest plane stride >= P * BY that is conflict-free at row pitch P
template <int TX, int TY, int TZ, int LAYOUT>
constexpr int conflict_free_plane(int P, int BY, int esize) {
    for (int q = P * BY; q < P * BY + 64; ++q)
        if (lds_conflict_free<TX, TY, TZ, LAYOUT>(P, q, esize)) return q;
    return P * BY;
}

template <int TX, int TY, int TZ, bool PAD = true, int ESIZE = 8, int LAYOUT = 0>
struct TileGeom {
    static constexpr int NT = TX * TY * TZ;
    static constexpr int BX0 = TX + 2 * HALO, BY = TY + 2 * HALO, BZ = TZ + 2 * HALO;
    using Lanes = LaneMap<TX, TY, TZ, LAYOUT>;
    // LDS row pitch and plane stride (elements): conflict-free centre reads
    static constexpr int BX = (LAYOUT == 0 && GR_LDS_PAD && PAD) ? conflict_free_pitch<TX, TY>(BX0, BY, ESIZE) : BX0;
    static constexpr int PLANE = LAYOUT == 0 ? BX * BY : conflict_free_plane<TX, TY, TZ, LAYOUT>(BX, BY, ESIZE);
    static_assert(LAYOUT == 0 || lds_conflict_free<TX, TY, TZ, LAYOUT>(BX, PLANE, ESIZE), "no conflict-free plane");
    static constexpr int NLDS = PLANE * BZ;
    // sub-boxes of the halo region that some stencil reads (x0,x1,y0,y1,z0,z1)
    static constexpr int NBOX = 7;
    static constexpr int box[NBOX][6] = {
        {-HALO, TX + HALO, 0, TY, 0, TZ},          // tile + x halo
        {-1, TX + 1, -HALO, 0, 0, TZ},             // y halo (low) with the xy edge column
        {-1, TX + 1, TY, TY + HALO, 0, TZ},        // y halo (high)
        {-1, TX + 1, -1, TY + 1, -1, 0},           // z = -1 layer with xz / yz edges
        {-1, TX + 1, -1, TY + 1, TZ, TZ + 1},      // z = TZ layer
        {0, TX, 0, TY, -HALO, -1},                 // z halo (low), axis taps only
        {0, TX, 0, TY, TZ + 1, TZ + HALO},         // z halo (high)
    };
    static constexpr int count(int b) {
        return (box[b][1] - box[b][0]) * (box[b][3] - box[b][2]) * (box[b][5] - box[b][4]);
    }
    static constexpr int start(int b) { return b == 0 ? 0 : start(b - 1) + count(b - 1); }
    static constexpr int TOTAL = start(NBOX - 1) + count(NBOX - 1);
    __host__ __device__ static constexpr int lds(int x, int y, int z) {
        return (z + HALO) * PLANE + (y + HALO) * BX + (x + HALO);
    }
    // tile-relative coordinates of halo entry p (boxes in order, x fastest inside a box);
    // unchanged for p >= TOTAL. The boxes are visited at compile time, so their extents
    // are immediates and the divisions constant.
    template <int K = 0>
    __device__ static void entry(int p, int& x, int& y, int& z) {
        if constexpr (K < NBOX) {
            constexpr int nx = box[K][1] - box[K][0], ny = box[K][3] - box[K][2];
            const int q = p - start(K);
            if (q >= 0 && q < count(K)) {
                x = box[K][0] + q % nx;
                y = box[K][2] + (q / nx) % ny;
                z = box[K][4] + q / (nx * ny);
            }
            entry<K + 1>(p, x, y, z);
        }
    }
};

// Per-thread input/output of one cell of a tile: stages the fields into LDS for the
// stencils (Stager interface of ccz4_cell) and combines/stores the results (Sink interface).
//
// Every thread owns the same NL entries of the tile's halo box for all fields, so the
// LDS slot and the global offset of each entry are computed once. Field loads are issued
// PF fields ahead of their use (in the fixed FIELD_ORDER), and the LDS tile is double
// buffered: while field p is evaluated from one buffer, field p+1 is written into the
// other, so each field costs one barrier and its HBM latency overlaps the work on the
// fields before it. The pointwise u^n value needed by the RK3 combine of a field travels
// with that field's prefetch.
template <int TX, int TY, int TZ, int MODE, int PF>
struct TileIO {
    using T = TileGeom<TX, TY, TZ, true, 8, 0>;
    static constexpr int NL = (T::TOTAL + T::NT - 1) / T::NT;
    static constexpr int LBUF = T::NLDS + 1; // one LDS buffer incl. a discard slot

    double* lds; // 2 * LBUF doubles
    const gdouble* in; // global address space: loads/stores wait on vmcnt only, never on LDS waits
    const gdouble* un;
    gdouble* out;
    long stride;
    long cidx;
    unsigned cb; // byte offset of the thread's cell inside one field
    int cl;
    bool active;
    bool ko_on;
    StageCoef cf;
    int loff[NL]; // LDS slot of each owned entry (T::NLDS = discard slot)
    unsigned goff[NL]; // byte offset of the entry inside one field of the block
    double buf[PF][NL];
    double unb[PF];
    double unv[NV];
    // u^n is the stage input itself (the first stage): the u^n value of a cell is then the
    // centre value the combine already receives, and is not loaded again (uniform)
    bool un_in;
    double hg[6], hA[6]; // last stage: new metric and A~ held for the algebraic constraints

    __device__ void init(int tid, int ox, int oy, int oz, int nt0, int nt1, int nt2) {
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int p = tid + i * T::NT;
            int x = 0, y = 0, z = 0;
            T::entry(p, x, y, z);
            const int gx = ox + x, gy = oy + y, gz = oz + z;
            const bool ok = p < T::TOTAL && gx < nt0 && gy < nt1 && gz < nt2;
            loff[i] = ok ? T::lds(x, y, z) : T::NLDS;
            goff[i] = ok ? 8u * (unsigned)(gx + nt0 * (gy + nt1 * gz)) : 0u;
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) prefetch(k);
        stage_to_lds(0);
        prefetch(PF);
    }

    __device__ __forceinline__ void prefetch(int pos) {
        if (pos >= NV) return;
        const int v = FIELD_ORDER[pos];
        const int slot = pos % PF;
        const gdouble* f = in + v * stride;
#pragma unroll
        for (int i = 0; i < NL; ++i) buf[slot][i] = gload_sv(f, goff[i]);
        if (MODE != MODE_RHS && !un_in) unb[slot] = gload(un + v * stride, cb);
    }

    // writes the prefetched field at position pos into its LDS buffer
    __device__ __forceinline__ void stage_to_lds(int pos) {
        if (pos >= NV) return;
        const int slot = pos % PF;
        double* L = lds + (pos & 1) * LBUF;
#pragma unroll
        for (int i = 0; i < NL; ++i) L[loff[i]] = buf[slot][i];
        if (MODE != MODE_RHS && !un_in) unv[FIELD_ORDER[pos]] = unb[slot];
    }

    // ---- Stager
    __device__ double centre(int v) const { return gload(in + v * stride, cb); }
    __device__ void metric64(double g[6]) const {
#pragma unroll
        for (int a = 0; a < 6; ++a) g[a] = gload(in + (GXX + a) * stride, cb);
    }
    __device__ FixedStrides<1, T::BX, T::PLANE> strides() const { return {}; }
    __device__ const double* field(int v) {
        const int pos = field_pos(v);
        __syncthreads(); // field pos is in its buffer; every thread is done with field pos-1
        stage_to_lds(pos + 1);
        prefetch(pos + 1 + PF);
        return lds + (pos & 1) * LBUF + cl;
    }

    // ---- Sink
    __device__ __forceinline__ void put(int v, double r, double ko, double uin) {
        const double rr = ko_on ? r + ko : r;
        if (MODE == MODE_RHS) {
            if (active) gstore(out + v * stride, cb, rr);
            return;
        }
        double val = combine_exact(cf.a1, un_in ? uin : unv[v], cf.a2, uin, cf.a3dt, rr);
        if (v == CHI) val = floor_chi(val);
        if (v == ALPHA) val = floor_alpha(val);
        // (one wave per SIMD: the held metric costs no occupancy here, and reading it back at
        // the end would expose a load latency)
        if (MODE == MODE_FINAL && v >= GXX && v < GXX + 6) {
            hg[v - GXX] = val;
            return;
        }
        if (MODE == MODE_FINAL && v >= AXX && v < AXX + 6) {
            hA[v - AXX] = val;
            return;
        }
        if (active) gstore(out + v * stride, cb, val);
    }

    __device__ __forceinline__ void finish() {
        if (MODE != MODE_FINAL) return;
        algebraic_exact(hg, hA);
        if (!active) return;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            gstore(out + (GXX + a) * stride, cb, hg[a]);
            gstore(out + (AXX + a) * stride, cb, hA[a]);
        }
    }
};

// TileIO whose halo staging goes straight from global memory to LDS (LDS layout 1). The
// buffer of a field is the whole box of the part (planes of rows of TX + 8 doubles), filled
// in 1 KiB chunks: lane l of a wave's load of chunk c writes doubles 128 c + 2 l, + 1 of the
// buffer, fetched from the pair of x-neighbours the layout puts there (pad slots and slots
// beyond the block fetch the field's first pair and are never read by an active cell).
// The waves take the chunks in turn. With NBUF field buffers, field p + NBUF - 1 is loaded
// while field p is evaluated. A load's LDS data is complete once the wave's vector memory
// counter has counted it back; the counter drains in issue order (loads and stores alike),
// so the barrier that starts field p waits until no more of this wave's operations are
// outstanding than the chunk loads it has issued since those of field p: the loads of the
// fields after p stay in flight across it.
// Needs 16-byte aligned pairs: an even row length (nt0) and 16-byte aligned buffers.
template <int TX, int TY, int TZ, int MODE>
struct TileIODma {
    typedef __attribute__((address_space(3))) double* LdsPtr;
    using T = TileGeom<TX, TY, TZ, true, 8, 1>;
    static constexpr int CH = 128; // doubles per chunk (64 lanes x 16 bytes)
    static constexpr int NCH = (T::NLDS + CH - 1) / CH;
    static constexpr int LBUF = NCH * CH; // doubles per buffer
    static constexpr int NW = T::NT / 64; // waves
    static constexpr int KW = (NCH + NW - 1) / NW; // chunk loads per wave
    static_assert(T::NT % 64 == 0, "whole waves");
    // field buffers: as many as the limit allows while two waves per SIMD (512 / NT
    // workgroups) still share the CU's 160 KB
    static constexpr int WG_PER_CU = 512 / T::NT;
    static constexpr int NBUF_MAX = MODE == MODE_FINAL ? GR_DMA_BUFFERS_FINAL : GR_DMA_BUFFERS;
    static constexpr int NBUF = (NBUF_MAX >= 3 && 3 * LBUF * 8 * WG_PER_CU <= 163840) ? 3 : 2;
    // in the last stage also the new metric of every cell, from its combine to the
    // algebraic constraints at the end, where it still fits beside the buffers; otherwise
    // the metric is stored at once and read back at the end
    static constexpr bool METRIC_LDS = MODE == MODE_FINAL && (NBUF * LBUF + 6 * T::NT) * 8 * WG_PER_CU <= 163840;
    static constexpr int LDS_DOUBLES = NBUF * LBUF + (METRIC_LDS ? 6 * T::NT : 0);
    // vector memory operations that every wave issues after a field's chunk loads and
    // before the barrier of that field: the unconditional chunk loads of the next field
    // (the last chunk of a wave may be skipped)
    static constexpr int YOUNGER = NBUF == 3 ? KW - 1 : 0;
    static_assert(YOUNGER < 16, "vmcnt field of s_waitcnt");

    double* lds; // LDS_DOUBLES doubles
    const gdouble* in;
    const gdouble* un;
    gdouble* out;
    long stride;
    long cidx;
    unsigned cb;
    int cl;
    int wave;
    bool active;
    bool ko_on;
    StageCoef cf;
    unsigned doff[KW]; // byte offset of this lane's pair in each of its wave's chunks
    int tid;
    double unv[NV];
    bool un_in; // u^n is the stage input (the first stage): not loaded again, as in TileIO
    double hA[6];
    // the barrier that starts a field waits for every vector memory operation issued before
    // the chunk loads of the field after it, stores included, so a field's result is stored
    // after the next barrier instead of just before it: the store then has a whole field to
    // complete
    double pend_val;
    int pend_v;

    // whether a stencil reads element x or x + 1 (x even) of row (y, z) of the box
    __host__ __device__ static constexpr bool pair_read(int x, int y, int z) {
        const bool zin = z >= 0 && z < TZ, zedge = z == -1 || z == TZ;
        const bool yin = y >= 0 && y < TY, yedge = y >= -1 && y <= TY;
        const bool xmid = x >= -2 && x <= TX, xin = x >= 0 && x < TX;
        return zin ? (yin || xmid) : (zedge ? (yedge && xmid) : (yin && xin));
    }
    // pair_read against the sub-box table itself, over the whole box
    static constexpr bool pair_read_matches_boxes() {
        for (int z = -HALO; z < TZ + HALO; ++z)
            for (int y = -HALO; y < TY + HALO; ++y)
                for (int x = -HALO; x < TX + HALO; x += 2) {
                    bool in = false;
                    for (int b = 0; b < T::NBOX; ++b)
                        for (int xx = x; xx <= x + 1; ++xx)
                            in = in || (xx >= T::box[b][0] && xx < T::box[b][1] && y >= T::box[b][2] &&
                                        y < T::box[b][3] && z >= T::box[b][4] && z < T::box[b][5]);
                    if (in != pair_read(x, y, z)) return false;
                }
        return true;
    }
    static_assert(pair_read_matches_boxes(), "fetched rows = rows some stencil reads");

    __device__ void init(int tid_, int ox, int oy, int oz, int nt0, int nt1, int nt2) {
        tid = tid_;
        wave = __builtin_amdgcn_readfirstlane(tid / 64); // uniform: LDS bases go to M0 from scalars
        const int lane = tid % 64;
#pragma unroll
        for (int k = 0; k < KW; ++k) {
            const int c = wave + k * NW;
            const int e = c * CH + 2 * lane;
            const int pl = e / T::PLANE, r = e % T::PLANE;
            const int x = r % T::BX - HALO, y = r / T::BX - HALO, z = pl - HALO;
            const int gx = ox + x, gy = oy + y, gz = oz + z;
            const bool ok = c < NCH && pl < T::BZ && r < T::BX * T::BY && gx + 1 < nt0 && gy < nt1 && gz < nt2;
            // whether a stencil reads the pair (x, x + 1) (x even): the union of the halo
            // sub-boxes of TileGeom::box, row by row. The box corners no stencil reads (about
            // a third of its rows) are not fetched: those lanes, and the lanes past the
            // block, load the pair at the part's origin, a line the same field loads anyway.
            const bool need = pair_read(x, y, z);
            doff[k] = 8u * (unsigned)(ok && need ? gx + nt0 * (gy + nt1 * gz) : ox + nt0 * (oy + nt1 * oz));
        }
        pend_v = -1;
#pragma unroll
        for (int p = 0; p < NBUF - 1; ++p) prefetch(p);
    }

    __device__ __forceinline__ void flush() {
        if (pend_v >= 0 && active) gstore(out + pend_v * stride, cb, pend_val);
        pend_v = -1;
    }

    __device__ __forceinline__ void prefetch(int pos) {
        if (pos >= NV) return;
        const int v = FIELD_ORDER[pos];
        const gchar* f = (const gchar*)(in + v * stride);
        // the LDS base is made opaque at every field, so the chunk addresses (M0 values) are
        // recomputed from it with one scalar add each instead of being kept (and spilled) as
        // 2 x KW live scalars across the whole kernel
        LdsPtr L = (LdsPtr)lds;
#ifdef __HIP_DEVICE_COMPILE__
        asm volatile("" : "+s"(L));
#endif
        L += (pos % NBUF) * LBUF;
#pragma unroll
        for (int k = 0; k < KW; ++k) {
            const int c = wave + k * NW; // uniform over the wave
            // every wave has chunks 0 .. KW-2; only the last may run past the box (a test on
            // each chunk costs a lane-mask register pair per chunk, live across the kernel)
            if (k < KW - 1 || c < NCH) {
#ifdef __HIP_DEVICE_COMPILE__
                // keeps the 32-bit offset at the load (scalar base + offset), in place
                asm volatile("" : "+v"(doff[k]));
#endif
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(f + doff[k]),
                                                 (__attribute__((address_space(3))) void*)(L + c * CH), 16, 0, 0);
            }
        }
        if (MODE != MODE_RHS && !un_in) unv[v] = gload(un + v * stride, cb);
    }

    // ---- Stager
    __device__ double centre(int v) const { return gload(in + v * stride, cb); }
    __device__ void metric64(double g[6]) const {
#pragma unroll
        for (int a = 0; a < 6; ++a) g[a] = gload(in + (GXX + a) * stride, cb);
    }
    __device__ FixedStrides<1, T::BX, T::PLANE> strides() const { return {}; }
    __device__ const double* field(int v) {
        const int pos = field_pos(v);
        if (NBUF == 2) {
            __syncthreads(); // field pos has landed in its buffer; every thread is done with pos - 1
        } else {
            // this wave's loads of field pos have landed (only those of pos + 1 may still be
            // outstanding: none after the last field) and its LDS accesses are done; then
            // the barrier: every wave's loads of pos have landed and every thread is done
            // with field pos - 1, whose buffer the prefetch below refills. The barrier is
            // written as a statement that may touch any memory, so the compiler keeps every
            // access on its side of it; a workgroup fence would also drain the loads of
            // pos + 1 (they write LDS)
            if (pos + 1 < NV)
                __builtin_amdgcn_s_waitcnt(0x0070 | YOUNGER); // vmcnt(YOUNGER) lgkmcnt(0)
            else
                __builtin_amdgcn_s_waitcnt(0x0070);
#ifdef __HIP_DEVICE_COMPILE__
            asm volatile("s_barrier" ::: "memory");
#endif
        }
        flush();
        prefetch(pos + NBUF - 1);
        return lds + (pos % NBUF) * LBUF + cl;
    }

    // ---- Sink
    __device__ __forceinline__ void put(int v, double r, double ko, double uin) {
        const double rr = ko_on ? r + ko : r;
        flush(); // at most one result waits (several per field only after the last barrier)
        if (MODE == MODE_RHS) {
            pend_v = v;
            pend_val = rr;
            return;
        }
        double val = combine_exact(cf.a1, un_in ? uin : unv[v], cf.a2, uin, cf.a3dt, rr);
        if (v == CHI) val = floor_chi(val);
        if (v == ALPHA) val = floor_alpha(val);
        // last stage: the new metric waits in LDS (this thread's own slots; or in the output,
        // read back at the end) for the algebraic constraints, instead of in registers across
        // the rest of the RHS; A~ comes last and is held in registers
        if (METRIC_LDS && v >= GXX && v < GXX + 6) {
            lds[NBUF * LBUF + (v - GXX) * T::NT + tid] = val;
            return;
        }
        if (MODE == MODE_FINAL && v >= AXX && v < AXX + 6) {
            hA[v - AXX] = val;
            return;
        }
        pend_v = v;
        pend_val = val;
    }

    __device__ __forceinline__ void finish() {
        flush();
        if (MODE != MODE_FINAL) return;
        double hg[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) hg[a] = METRIC_LDS ? lds[NBUF * LBUF + a * T::NT + tid] : gload(out + (GXX + a) * stride, cb);
        algebraic_exact(hg, hA);
        if (!active) return;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            gstore(out + (GXX + a) * stride, cb, hg[a]);
            gstore(out + (AXX + a) * stride, cb, hA[a]);
        }
    }
};

// Mixed-precision counterpart of TileIO: the stencils and the RHS algebra see FP32
// values, the RK3 combine, the floors and the algebraic constraints are done in FP64 on
// the FP64 state. Each field is staged as FP32 deltas from one FP64 reference value per
// tile (the field at the tile's first cell): every stencil's weights sum to zero, so
// derivatives of the deltas are derivatives of the field, while small structure on top of
// an O(1) field (1e-8 noise on the metric) keeps FP32 relative precision of the delta.
// UN_IN: u^n is the stage input itself (the first stage, chosen at launch): its value at the
// cell is the stage input's centre value, so it is not loaded separately
template <int TX, int TY, int TZ, int MODE, int PF, bool UN_IN = false>
struct TileIO32 {
    using T = TileGeom<TX, TY, TZ, GR_LDS_PAD32 != 0, 4>;
    static constexpr int NL = (T::TOTAL + T::NT - 1) / T::NT;
    static constexpr int LBUF = T::NLDS + 1;

    float* lds; // 2 * LBUF floats
    const gdouble* in; // global address space: loads/stores wait on vmcnt only, never on LDS waits
    const gdouble* un;
    gdouble* out;
    long stride;
    long cidx;
    unsigned cb; // byte offset of the thread's cell inside one field
    long ridx; // the tile's reference cell
    int cl;
    bool active;
    bool ko_on;
    StageCoef cf;
    int loff[NL];
    unsigned goff[NL];
    double buf[PF][NL];
    double refb[PF], unb[PF], inb[PF];
    double unv[NV], inv[NV];
    double hg[6], hA[6];

    __device__ void init(int tid, int ox, int oy, int oz, int nt0, int nt1, int nt2) {
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int p = tid + i * T::NT;
            int x = 0, y = 0, z = 0;
            T::entry(p, x, y, z);
            const int gx = ox + x, gy = oy + y, gz = oz + z;
            const bool ok = p < T::TOTAL && gx < nt0 && gy < nt1 && gz < nt2;
            loff[i] = ok ? T::lds(x, y, z) : T::NLDS;
            goff[i] = ok ? 8u * (unsigned)(gx + nt0 * (gy + nt1 * gz)) : 0u;
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) prefetch(k);
        stage_to_lds(0);
        prefetch(PF);
    }

    __device__ __forceinline__ void prefetch(int pos) {
        if (pos >= NV) return;
        const int v = FIELD_ORDER[pos];
        const int slot = pos % PF;
        const gdouble* f = in + v * stride;
#pragma unroll
        for (int i = 0; i < NL; ++i) buf[slot][i] = gload(f, goff[i]);
        refb[slot] = f[ridx];
        inb[slot] = gload(f, cb);
        if (MODE != MODE_RHS && !UN_IN) unb[slot] = gload(un + v * stride, cb);
    }

    __device__ __forceinline__ void stage_to_lds(int pos) {
        if (pos >= NV) return;
        const int slot = pos % PF;
        float* L = lds + (pos & 1) * LBUF;
        const double r = refb[slot];
#pragma unroll
        for (int i = 0; i < NL; ++i) L[loff[i]] = (float)(buf[slot][i] - r);
        const int v = FIELD_ORDER[pos];
        inv[v] = inb[slot];
        if (MODE != MODE_RHS) unv[v] = UN_IN ? inb[slot] : unb[slot];
    }

    // ---- Stager
    __device__ float centre(int v) const { return (float)gload(in + v * stride, cb); }
    __device__ void metric64(double g[6]) const {
#pragma unroll
        for (int a = 0; a < 6; ++a) g[a] = gload(in + (GXX + a) * stride, cb);
    }
    __device__ FixedStrides<1, T::BX, T::PLANE> strides() const { return {}; }
    __device__ const float* field(int v) {
        const int pos = field_pos(v);
        __syncthreads();
        stage_to_lds(pos + 1);
        prefetch(pos + 1 + PF);
        return lds + (pos & 1) * LBUF + cl;
    }

    // ---- Sink (FP64 combine on the FP64 stage input and u^n)
    __device__ __forceinline__ void put(int v, float r, float ko, float) {
        const double rr = ko_on ? (double)r + (double)ko : (double)r;
        if (MODE == MODE_RHS) {
            if (active) gstore(out + v * stride, cb, rr);
            return;
        }
        double val = combine_exact(cf.a1, unv[v], cf.a2, inv[v], cf.a3dt, rr);
        if (v == CHI) val = floor_chi(val);
        if (v == ALPHA) val = floor_alpha(val);
        if (MODE == MODE_FINAL && v >= GXX && v < GXX + 6) {
            hg[v - GXX] = val;
            return;
        }
        if (MODE == MODE_FINAL && v >= AXX && v < AXX + 6) {
            hA[v - AXX] = val;
            return;
        }
        if (active) gstore(out + v * stride, cb, val);
    }

    __device__ __forceinline__ void finish() {
        if (MODE != MODE_FINAL) return;
        algebraic_exact(hg, hA);
        if (!active) return;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            gstore(out + (GXX + a) * stride, cb, hg[a]);
            gstore(out + (AXX + a) * stride, cb, hA[a]);
        }
    }
};

struct StageArgs {
    Batch batch;
    int sel_in, sel_un, sel_out;
    int ntiles;
    int xcd_remap;
    Stencil<double> K;
    Stencil<float> K32;
    CellParams P;
    StageCoef cf;
    int ko_on;
    // tiles whose flag differs from `want` are left to the other precision's kernel
    const unsigned char* flags;
    int want;
    // scan != 0: launched as stage_kernel_scan over a small grid
    int scan;
};

// workgroup id -> tile id: consecutive tiles go to the same XCD (workgroups are dealt
// round-robin over the 8 XCDs), so that halo rows shared by neighbouring tiles are
// found in that XCD's L2.
__device__ __forceinline__ int tile_of_block(int bid, int ntiles, int remap) {
    if (!remap) return bid;
    // XCD x = bid % 8 takes ntiles / 8 consecutive slots, one more for the first ntiles % 8
    const int q = ntiles >> 3, r = ntiles & 7, x = bid & 7;
    return x * q + (x < r ? x : r) + (bid >> 3);
}

// a value the whole workgroup holds alike, moved to scalar registers
__device__ __forceinline__ unsigned long long wg_uniform(unsigned long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
template <class P>
__device__ __forceinline__ P* wg_uniform(P* p) {
    return (P*)wg_uniform((unsigned long long)p);
}

// the block and local tile coordinates of global tile t (uniform per workgroup)
struct TileRef {
    BlockDesc b;
    const gdouble* in; // global address space: loads/stores wait on vmcnt only, never on LDS waits
    const gdouble* un;
    gdouble* out;
    Sources src;
    int tix, tiy, tiz;
};

__device__ __forceinline__ TileRef resolve_tile(const StageArgs& A, int t) {
    TileRef r;
    int tile0, tiles0, tiles1;
    const double *b0, *b1, *b2;
    if (A.batch.table) {
        const BatchBlock& BB = A.batch.table[A.batch.tile_block[t]];
        r.b = BB.b;
        r.src = BB.src;
        b0 = BB.buf[0];
        b1 = BB.buf[1];
        b2 = BB.buf[2];
        tile0 = BB.tile0;
        tiles0 = BB.tiles[0];
        tiles1 = BB.tiles[1];
    } else {
        r.b = A.batch.one.b;
        r.src = A.batch.one.src;
        b0 = A.batch.one.buf[0];
        b1 = A.batch.one.buf[1];
        b2 = A.batch.one.buf[2];
        tile0 = 0;
        tiles0 = A.batch.one.tiles[0];
        tiles1 = A.batch.one.tiles[1];
    }
    const double* bufs[3] = {b0, b1, b2};
    // t is the same for the whole workgroup, and so is everything read for it
    r.in = wg_uniform((const gdouble*)bufs[A.sel_in]);
    r.un = wg_uniform((const gdouble*)bufs[A.sel_un]);
    r.out = wg_uniform((gdouble*)const_cast<double*>(bufs[A.sel_out]));
    r.b.stride = (long)wg_uniform((unsigned long long)r.b.stride);
    const int lt = t - tile0;
    // Tile order: x fastest, then z inside groups of 4 layers, then y, then the groups. The
    // workgroups that run together on one XCD (consecutive tiles) then cover a brick of
    // tiles whose y and z halo rows they share in that XCD's L2 (x, y, z order: 5% slower
    // at 128^3, where the stage input does not fit the memory-side cache).
    {
        constexpr int G = 4;
        const int tiles2 = (r.b.n[2] + GR_TILE_Z - 1) / GR_TILE_Z;
        const int row = lt / tiles0; // (y, z) row of tiles
        const int grp = row / (G * tiles1), in = row % (G * tiles1);
        const int gz = min(G, tiles2 - grp * G); // z layers in this group
        r.tix = lt % tiles0;
        r.tiz = grp * G + in % gz;
        r.tiy = in / gz;
    }
    return r;
}

// matter sources of the cell (zero for a block without sources: a vacuum block of a batch
// that has sources elsewhere; zero sources give the vacuum RHS bit for bit)
__device__ __forceinline__ void load_sources(const Sources& S, long cidx, CellSources& src) {
    if (!S.rho) { // uniform over the workgroup
        src.rho = src.S = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) src.Si[a] = 0.0;
#pragma unroll
        for (int a = 0; a < 6; ++a) src.Sij[a] = 0.0;
        return;
    }
    src.rho = S.rho[cidx];
    src.S = S.S[cidx];
#pragma unroll
    for (int a = 0; a < 3; ++a) src.Si[a] = S.Si[3 * cidx + a];
#pragma unroll
    for (int a = 0; a < 6; ++a) src.Sij[a] = S.Sij[6 * cidx + a];
}

// staging variant SV: 0 = registers into a padded box (layout 0); 1, 2 = direct LDS loads
// into an unpadded box (layout 1); 2 runs parts twice as deep in z
constexpr int sv_layout(int SV) { return SV == 0 ? 0 : 1; }

// part `part` of a tile: TX x TY x TZ cells at x = (part % (TILE_X / TX)) * TX,
// z = (part / (TILE_X / TX)) * TZ inside the tile
template <int TX, int TY, int TZ, int MODE, bool HAS_SRC, bool KO_EXACT, int SV>
__device__ __forceinline__ void stage_tile(const StageArgs& A, const TileRef& R, int part, double* lds) {
    using T = TileGeom<TX, TY, TZ, true, 8, sv_layout(SV)>;
    constexpr int XS = GR_TILE_X / TX;
    const BlockDesc& b = R.b;
    const int tid = threadIdx.x;
    int lx = 0, ly = 0, lz = 0;
    T::Lanes::cell(tid, lx, ly, lz);
    const int x0 = R.tix * GR_TILE_X + (part % XS) * TX;
    const int z0 = R.tiz * GR_TILE_Z + (part / XS) * TZ;
    const int ix = x0 + lx, iy = R.tiy * TY + ly, iz = z0 + lz; // interior-relative
    const bool active = ix < b.n[0] && iy < b.n[1] && iz < b.n[2];
    const int gi = b.ng + (active ? ix : 0), gj = b.ng + (active ? iy : 0), gk = b.ng + (active ? iz : 0);
    const long cidx = gi + (long)b.nt[0] * (gj + (long)b.nt[1] * gk);

    typename std::conditional<SV != 0, TileIODma<TX, TY, TZ, MODE>, TileIO<TX, TY, TZ, MODE, GR_FP64_PREFETCH>>::type io;
    io.lds = lds;
    io.in = R.in;
    io.un = R.un;
    io.un_in = R.un == R.in;
    io.out = R.out;
    io.stride = b.stride;
    io.cidx = cidx;
    io.cb = 8u * (unsigned)cidx;
    io.cl = T::lds(lx, ly, lz);
    io.active = active;
    io.ko_on = A.ko_on != 0;
    io.cf = A.cf;
    io.init(tid, b.ng + x0, b.ng + R.tiy * TY, b.ng + z0, b.nt[0], b.nt[1], b.nt[2]);

    CellSources src{};
    if (HAS_SRC) load_sources(R.src, cidx, src);
    ccz4_cell<double, HAS_SRC, KO_EXACT>(io, io, A.K, A.P, src);
}

// parts of a tile of TX x . x TZ cells
template <int TX, int TZ>
constexpr int parts_of() { return (GR_TILE_X / TX) * (GR_TILE_Z / TZ); }

// LDS doubles of one FP64 part's two field buffers under staging variant SV
template <int TX, int TY, int TZ, int SV, int MODE>
constexpr int fp64_lds_doubles() {
    return SV != 0 ? TileIODma<TX, TY, TZ, MODE>::LDS_DOUBLES : 2 * (TileGeom<TX, TY, TZ, true, 8, 0>::NLDS + 1);
}

// whether part `part` of tile R holds no cell of the block (uniform over the workgroup)
template <int TX, int TZ>
__device__ __forceinline__ bool part_outside(const TileRef& R, int part) {
    constexpr int XS = GR_TILE_X / TX;
    return R.tix * GR_TILE_X + (part % XS) * TX >= R.b.n[0] || R.tiz * GR_TILE_Z + (part / XS) * TZ >= R.b.n[2];
}

// one part of one tile per workgroup (parts_of consecutive workgroups per tile)
template <int TX, int TY, int TZ, int MODE, bool HAS_SRC, bool KO_EXACT, int SV>
__global__ __launch_bounds__(TX* TY* TZ, GR_FP64_MIN_WAVES) void stage_kernel(StageArgs A) {
    constexpr int NP = parts_of<TX, TZ>();
    __shared__ double lds[fp64_lds_doubles<TX, TY, TZ, SV, MODE>()];
    const int slot = tile_of_block(blockIdx.x, A.ntiles * NP, A.xcd_remap);
    const int tile = slot / NP, part = slot % NP;
    if (A.flags && A.flags[tile] != A.want) return; // whole workgroup leaves before any barrier
    const TileRef R = resolve_tile(A, tile);
    if (part_outside<TX, TZ>(R, part)) return;
    stage_tile<TX, TY, TZ, MODE, HAS_SRC, KO_EXACT, SV>(A, R, part, lds);
}

// the flagged tiles only, from a small grid: each workgroup gathers the tiles with
// flags == want among tiles blockIdx.x + k * gridDim.x and runs them one after the other
// (every thread takes the same trip counts, so the barriers are uniform)
template <int TX, int TY, int TZ, int MODE, bool HAS_SRC, bool KO_EXACT, int SV>
__global__ __launch_bounds__(TX* TY* TZ, GR_FP64_MIN_WAVES) void stage_kernel_scan(StageArgs A) {
    constexpr int NT = TX * TY * TZ;
    __shared__ double lds[fp64_lds_doubles<TX, TY, TZ, SV, MODE>()];
    __shared__ int list[NT];
    __shared__ int nlist;
    const int per = (int)gridDim.x * NT;
    for (int base = 0; base < A.ntiles; base += per) {
        if (threadIdx.x == 0) nlist = 0;
        __syncthreads();
        const int t = base + (int)blockIdx.x + (int)threadIdx.x * (int)gridDim.x;
        if (t < A.ntiles && A.flags[t] == A.want) list[atomicAdd(&nlist, 1)] = t;
        __syncthreads();
        const int n = nlist;
        __syncthreads();
        for (int item = 0; item < n; ++item) {
            const TileRef R = resolve_tile(A, list[item]);
            for (int part = 0; part < parts_of<TX, TZ>(); ++part) {
                if (part_outside<TX, TZ>(R, part)) continue; // uniform over the workgroup
                stage_tile<TX, TY, TZ, MODE, HAS_SRC, KO_EXACT, SV>(A, R, part, lds);
                __syncthreads(); // the next part reuses the LDS buffers
            }
        }
    }
}

static_assert(GR_TILE_X % GR_FP64_SPLIT == 0 && GR_TILE_Z % GR_FP64_ZSPLIT == 0,
              "the FP64 split must divide the tile");
constexpr int TILE_X = GR_TILE_X, TILE_Y = GR_TILE_Y, TILE_Z = GR_TILE_Z;
constexpr int FP64_TX = GR_TILE_X / GR_FP64_SPLIT, FP64_TY = GR_TILE_Y, FP64_TZ = GR_TILE_Z / GR_FP64_ZSPLIT;
static_assert(!GR_LDS_PAD || FP64_TX != 8 || TILE_X != 16 || TILE_Y != 4 ||
                  (TileGeom<FP64_TX, FP64_TY, FP64_TZ, true, 8, 1>::BX == 16 &&
                   TileGeom<FP64_TX, FP64_TY, FP64_TZ, true, 8, 1>::PLANE == 200 &&
                   TileGeom<FP64_TX, FP64_TY, FP64_TZ>::BX == 24 && TileGeom<TILE_X, TILE_Y, TILE_Z, true, 4>::BX == 48),
              "LDS layouts of the default tiles: FP64 rows of 24 (layout 0), rows of 16 in planes of 200 (layout "
              "1); FP32 rows of 48");
// part shapes of the staging variants: 0 and 1 the build's part, 2 twice as deep in z
template <int SV>
struct SvShape {
    static constexpr int TX = SV == 0 ? FP64_TX : 8, TY = SV == 0 ? FP64_TY : 4, TZ = SV == 0 ? FP64_TZ : (SV == 1 ? 4 : 8);
    static constexpr int NT = TX * TY * TZ, PARTS = parts_of<TX, TZ>();
};
static_assert(TILE_X == 16 && TILE_Y == 4 && TILE_Z == 8, "the direct-staging parts (8x4x4, 8x4x8) tile 16x4x8 tiles");
static_assert(FP64_STAGE_VARIANTS == 3, "staging variants 0, 1, 2");
// workgroups of the FP64 pass of the mixed form: 256 threads per CU
constexpr int FP64_SCAN_GRID = 256 * 256 / (FP64_TX * FP64_TY * FP64_TZ);

// one part of TX x TY x TZ cells (TILE_X / TX parts along x) of a launch tile per workgroup;
// the tile is the unit of the FP32 / FP64 choice
template <int TX, int TY, int TZ, int MODE, bool HAS_SRC, bool UN_IN = false>
__global__ __launch_bounds__(TX* TY* TZ, GR_MIXED_MIN_WAVES) void stage_kernel_mixed(StageArgs A) {
    using T = TileGeom<TX, TY, TZ, GR_LDS_PAD32 != 0, 4>;
    constexpr int NP = GR_TILE_X / TX;
    static_assert(GR_TILE_X % TX == 0 && TY == GR_TILE_Y && TZ == GR_TILE_Z, "parts split the tile along x");
    __shared__ float lds[2 * (T::NLDS + 1)];

    const int slot = tile_of_block(blockIdx.x, A.ntiles * NP, A.xcd_remap);
    const int tile = slot / NP, part = slot % NP;
    if (A.flags && A.flags[tile] != A.want) return;
    const TileRef R = resolve_tile(A, tile);
    const BlockDesc& b = R.b;
    const int x0 = R.tix * GR_TILE_X + part * TX;
    if (x0 >= b.n[0]) return; // whole workgroup, before any barrier
    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = (tid / TX) % TY, lz = tid / (TX * TY);
    const int ix = x0 + lx, iy = R.tiy * TY + ly, iz = R.tiz * TZ + lz;
    const bool active = ix < b.n[0] && iy < b.n[1] && iz < b.n[2];
    const int gi = b.ng + (active ? ix : 0), gj = b.ng + (active ? iy : 0), gk = b.ng + (active ? iz : 0);
    const long cidx = gi + (long)b.nt[0] * (gj + (long)b.nt[1] * gk);

    TileIO32<TX, TY, TZ, MODE, GR_MIXED_PREFETCH, UN_IN> io;
    io.lds = lds;
    io.in = R.in;
    io.un = R.un;
    io.out = R.out;
    io.stride = b.stride;
    io.cidx = cidx;
    io.cb = 8u * (unsigned)cidx;
    io.ridx = (b.ng + x0) + (long)b.nt[0] * ((b.ng + R.tiy * TY) + (long)b.nt[1] * (b.ng + R.tiz * TZ));
    io.cl = T::lds(lx, ly, lz);
    io.active = active;
    io.ko_on = A.ko_on != 0;
    io.cf = A.cf;
    io.init(tid, b.ng + x0, b.ng + R.tiy * TY, b.ng + R.tiz * TZ, b.nt[0], b.nt[1], b.nt[2]);

    CellSources src{};
    if (HAS_SRC) load_sources(R.src, cidx, src);
    ccz4_cell<float, HAS_SRC, false>(io, io, A.K32, A.P, src);
}

// flags[tile] = 1 where the tile needs the FP64 kernel in the mixed form: its smallest chi
// (interior cells of the stage input) is below chi_min or not finite.
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(TX* TY* TZ) void classify_tiles_kernel(StageArgs A, double chi_min, unsigned char* flags) {
    __shared__ int bad;
    const int tile = blockIdx.x;
    const TileRef R = resolve_tile(A, tile);
    const BlockDesc& b = R.b;
    const int tid = threadIdx.x;
    const int ix = R.tix * TX + tid % TX, iy = R.tiy * TY + (tid / TX) % TY, iz = R.tiz * TZ + tid / (TX * TY);
    if (tid == 0) bad = 0;
    __syncthreads();
    if (ix < b.n[0] && iy < b.n[1] && iz < b.n[2]) {
        const double c =
            R.in[CHI * b.stride + (b.ng + ix) + (long)b.nt[0] * ((b.ng + iy) + (long)b.nt[1] * (b.ng + iz))];
        if (!(c >= chi_min)) bad = 1; // benign race: every writer stores 1
    }
    __syncthreads();
    if (tid == 0) flags[tile] = (unsigned char)bad;
}

// scan: nblk workgroups gather the flagged tiles; otherwise one workgroup per part
template <int MODE, bool HAS_SRC, int SV>
hipError_t launch_stage_sv(const StageArgs& A, int scan_grid, hipStream_t s) {
    using S = SvShape<SV>;
    if (A.scan)
        stage_kernel_scan<S::TX, S::TY, S::TZ, MODE, HAS_SRC, GR_FP64_KO_EXACT != 0, SV>
            <<<dim3(scan_grid), dim3(S::NT), 0, s>>>(A);
    else if (A.ntiles > 0)
        stage_kernel<S::TX, S::TY, S::TZ, MODE, HAS_SRC, GR_FP64_KO_EXACT != 0, SV>
            <<<dim3(A.ntiles * S::PARTS), dim3(S::NT), 0, s>>>(A);
    return hipGetLastError();
}

template <int MODE, bool HAS_SRC>
hipError_t launch_stage_t(const StageArgs& A, int scan_grid, int variant, hipStream_t s) {
    if (variant == 1) return launch_stage_sv<MODE, HAS_SRC, 1>(A, scan_grid, s);
    if (variant == 2) return launch_stage_sv<MODE, HAS_SRC, 2>(A, scan_grid, s);
    return launch_stage_sv<MODE, HAS_SRC, 0>(A, scan_grid, s);
}

// ------------------------------------------------------------------------------------
// ghost fill
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double coord(const BlockDesc& b, int d, int idx) {
#pragma clang fp contract(off)
    // GridBlock::x, include/granite/core/grid.hpp:92-94
    return b.lo[d] + ((double)idx - (double)b.ng + 0.5) * b.dx[d];
}

__device__ __forceinline__ double radius(double x, double y, double z) {
#pragma clang fp contract(off)
    return sqrt(x * x + y * y + z * z) + 1e-30;
}

__device__ __forceinline__ double f_inf(int v) {
    return (v == 0 || v == 1 || v == 4 || v == 6 || v == 18) ? 1.0 : 0.0;
}

// Final value of a ghost cell after the reference's three passes: each pass over axis d
// rewrites the ghost layers of d over the full transverse extent from the boundary layer,
// so a cell that is a ghost along a set of axes ends with the chain of 1/r fall-offs from
// its interior anchor, applied in X, Y, Z order at the intermediate positions
// (src/main.cpp:395-483).
__device__ __forceinline__ long ghost_count(const BlockDesc& b) {
    return 2L * b.ng * b.nt[1] * b.nt[0] + (long)b.n[2] * 2 * b.ng * b.nt[0] + (long)b.n[2] * b.n[1] * 2 * b.ng;
}

// ghost cell number t of block b (z slabs, then y rows of interior z, then x columns).
// For the outflow copy (mode 2) only the faces whose bit is set in `faces` are filled
// (bit 2d: low side of axis d, bit 2d+1: high side); the others keep their values, as
// fillGhostZones leaves the faces covered by a parent level to prolongation
// (src/amr/amr.cpp:412-480).
// Fields g, g + G, ... of the cell: the G field groups run as separate workgroups (second
// grid dimension), so a small ghost region still fills the machine.
template <int G>
__device__ void fill_ghost_cell(const BlockDesc& b, gdouble* u, long t_, int mode, int faces, int g) {
#pragma clang fp contract(off)
    // ghost cells number fewer than 2^29 (the block limit): 32-bit index arithmetic
    const int t = (int)t_;
    const int nt0 = b.nt[0], nt1 = b.nt[1];
    const int nA = 2 * b.ng * nt1 * nt0;
    const int nB = b.n[2] * 2 * b.ng * nt0;
    const int ng = b.ng, g2 = 2 * ng;
    int i, j, k;
    if (t < nA) {
        i = t % nt0;
        const int r = t / nt0;
        j = r % nt1;
        const int kk = r / nt1;
        k = kk < ng ? kk : b.n[2] + kk;
    } else if (t < nA + nB) {
        const int q = t - nA;
        i = q % nt0;
        const int r = q / nt0;
        const int jj = r % g2;
        k = ng + r / g2;
        j = jj < ng ? jj : b.n[1] + jj;
    } else {
        const int q = t - nA - nB;
        const int ii = q % g2;
        const int r = q / g2;
        j = ng + r % b.n[1];
        k = ng + r / b.n[1];
        i = ii < ng ? ii : b.n[0] + ii;
    }
    const int I[3] = {i, j, k};
    int A[3];
    bool gh[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        gh[d] = I[d] < ng || I[d] >= ng + b.n[d];
        if (mode == 1) {
            int r = (I[d] - ng) % b.n[d];
            if (r < 0) r += b.n[d];
            A[d] = ng + r;
        } else {
            const bool lo = I[d] < ng, hi = I[d] >= ng + b.n[d];
            const bool fill = mode != 2 || (lo && ((faces >> (2 * d)) & 1)) || (hi && ((faces >> (2 * d + 1)) & 1));
            A[d] = (lo && fill) ? ng : ((hi && fill) ? ng + b.n[d] - 1 : I[d]);
        }
    }
    const long dst = I[0] + (long)nt0 * (I[1] + (long)nt1 * I[2]);
    const long src = A[0] + (long)nt0 * (A[1] + (long)nt1 * A[2]);
    // the anchor is never a cell this launch writes, so all loads go out before any store
    constexpr int NF = (NV + G - 1) / G;
    double val[NF];
#pragma unroll
    for (int m = 0; m < NF; ++m) {
        const int v = g + G * m;
        if (v < NV) val[m] = u[v * b.stride + src];
    }
    if (mode == 0) {
        // (r_int, r_gh) of each step of the chain: start at the anchor, move one ghost axis
        // at a time in X, Y, Z order
        double Q[3] = {coord(b, 0, A[0]), coord(b, 1, A[1]), coord(b, 2, A[2])};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (!gh[d]) continue;
            const double rin = radius(Q[0], Q[1], Q[2]);
            Q[d] = coord(b, d, I[d]);
            const double rgh = radius(Q[0], Q[1], Q[2]);
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                const double fi = f_inf(g + G * m);
                val[m] = fi + (val[m] - fi) * rin / rgh;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < NF; ++m) {
        const int v = g + G * m;
        if (v < NV) u[v * b.stride + dst] = val[m];
    }
}

template <int G>
__global__ void fill_ghosts_kernel(BlockDesc b, double* u, int mode, int faces) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ghost_count(b)) return;
    fill_ghost_cell<G>(b, (gdouble*)u, t, mode, faces, blockIdx.y);
}

// ------------------------------------------------------------------------------------
// AMR transfer between a coarse and a fine block (one block per level)
// ------------------------------------------------------------------------------------

// Per-axis part of the trilinear coarse -> fine stencil, with the reference's index rounding
// and clamping (src/amr/amr.cpp:484-564): entry e < f.nt[0] is x index e, then the y indices,
// then the z indices. The stencil is separable, so the prolongation itself reads three entries
// per fine cell instead of repeating three divisions for every field group.
__global__ void prolong_axes_kernel(BlockDesc c, BlockDesc f, ProlongAxis* __restrict__ tab) {
#pragma clang fp contract(off)
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= f.nt[0] + f.nt[1] + f.nt[2]) return;
    const int d = t < f.nt[0] ? 0 : (t < f.nt[0] + f.nt[1] ? 1 : 2);
    const int i = t - (d > 0 ? f.nt[0] : 0) - (d > 1 ? f.nt[1] : 0);
    const double x = coord(f, d, i);
    const double cf = (x - coord(c, d, c.ng)) / c.dx[d] + (double)c.ng;
    int ci = (int)round(cf);
    ci = ci < 0 ? 0 : (ci > c.nt[d] - 1 ? c.nt[d] - 1 : ci);
    const double fr = cf - (double)ci;
    int cj = ci + (fr >= 0.0 ? 1 : -1);
    cj = cj > c.nt[d] - 1 ? c.nt[d] - 1 : cj;
    cj = cj < 0 ? 0 : cj;
    tab[t] = ProlongAxis{ci, cj, fabs(fr)};
}

// Trilinear coarse -> fine into every fine cell, ghosts included (src/amr/amr.cpp:484-564),
// from the per-axis stencil table of this (coarse, fine) pair. One workgroup per fine tile of
// PT_X x PT_Y x PT_Z cells (a thread per (x, y) column of PT_Z cells): per field, the coarse
// box the tile's stencils reach is staged in LDS (double-buffered, one barrier per field) and
// the eight taps are read from there, so the coarse field is read from memory once per tile
// instead of eight times per fine cell. A tile whose stencils reach beyond the box capacity
// (a fine block not at half the parent's spacing) reads the taps from memory directly.
#ifndef GR_PT_X
#define GR_PT_X 32
#endif
#ifndef GR_PT_Y
#define GR_PT_Y 8
#endif
#ifndef GR_PT_Z
#define GR_PT_Z 8
#endif
#ifndef GR_PROLONG_PREFETCH
#define GR_PROLONG_PREFETCH 1
#endif
constexpr int PT_X = GR_PT_X, PT_Y = GR_PT_Y, PT_Z = GR_PT_Z, PT_N = PT_X * PT_Y;
constexpr int PB_X = PT_X / 2 + 2, PB_Y = PT_Y / 2 + 2, PB_Z = PT_Z / 2 + 2, PB_N = PB_X * PB_Y * PB_Z;
constexpr int PB_PER_THREAD = (PB_N + PT_N - 1) / PT_N;

__global__ __launch_bounds__(PT_N) void prolong_kernel(BlockDesc c, const double* __restrict__ cu, BlockDesc f,
                                                             double* __restrict__ fu,
                                                             const ProlongAxis* __restrict__ tab, int ntx, int nty) {
#pragma clang fp contract(off)
    __shared__ double box[2][PB_N];
    __shared__ int lim[6]; // per axis: lowest, highest coarse index any stencil of the tile reads
    const int tile = blockIdx.x;
    const int i0 = (tile % ntx) * PT_X, j0 = ((tile / ntx) % nty) * PT_Y, k0 = (tile / (ntx * nty)) * PT_Z;
    const int tid = threadIdx.x;
    if (tid < 6) lim[tid] = (tid & 1) ? -1 : 0x7fffffff;
    __syncthreads();
    if (tid < PT_X + PT_Y + PT_Z) {
        const int d = tid < PT_X ? 0 : (tid < PT_X + PT_Y ? 1 : 2);
        const int e = d == 0 ? i0 + tid : (d == 1 ? j0 + tid - PT_X : k0 + tid - PT_X - PT_Y);
        if (e < f.nt[d]) {
            const ProlongAxis A = tab[(d > 0 ? f.nt[0] : 0) + (d > 1 ? f.nt[1] : 0) + e];
            atomicMin(&lim[2 * d], A.c0 < A.c1 ? A.c0 : A.c1);
            atomicMax(&lim[2 * d + 1], A.c0 > A.c1 ? A.c0 : A.c1);
        }
    }
    __syncthreads();
    const int bx = lim[0], by = lim[2], bz = lim[4];
    const int ex = lim[1] - bx + 1, ey = lim[3] - by + 1, ez = lim[5] - bz + 1;
    const bool staged = ex <= PB_X && ey <= PB_Y && ez <= PB_Z; // uniform over the workgroup

    const int i = i0 + tid % PT_X, j = j0 + tid / PT_X;
    const bool live = i < f.nt[0] && j < f.nt[1];
    const ProlongAxis X = tab[live ? i : 0], Y = tab[f.nt[0] + (live ? j : 0)];
    // leading (x, y) factors of the reference's weight products, (wx * wy) * wz
    const double p00 = (1 - X.ax) * (1 - Y.ax), p10 = X.ax * (1 - Y.ax), p01 = (1 - X.ax) * Y.ax, p11 = X.ax * Y.ax;
    const ProlongAxis* Ztab = tab + f.nt[0] + f.nt[1];
    const int kn = f.nt[2] - k0 < PT_Z ? f.nt[2] - k0 : PT_Z;
    const int fplane = f.nt[0] * f.nt[1], fcol = i + f.nt[0] * j;
    if (staged) {
        const int q00 = (X.c0 - bx) + PB_X * (Y.c0 - by), q10 = (X.c1 - bx) + PB_X * (Y.c0 - by),
                  q01 = (X.c0 - bx) + PB_X * (Y.c1 - by), q11 = (X.c1 - bx) + PB_X * (Y.c1 - by);
        // coarse offsets of this thread's box elements (-1: none)
        int src[PB_PER_THREAD];
#pragma unroll
        for (int r = 0; r < PB_PER_THREAD; ++r) {
            const int e = tid + r * PT_N;
            const int a = e % PB_X, b = (e / PB_X) % PB_Y, z = e / (PB_X * PB_Y);
            src[r] = (e < PB_N && a < ex && b < ey && z < ez) ? (bx + a) + c.nt[0] * ((by + b) + c.nt[1] * (bz + z)) : -1;
        }
        double next[PB_PER_THREAD];
#if GR_PROLONG_PREFETCH
#pragma unroll
        for (int r = 0; r < PB_PER_THREAD; ++r)
            if (src[r] >= 0) next[r] = cu[(long)blockIdx.y * c.stride + src[r]];
#endif
        int buf = 0;
        for (int v = blockIdx.y; v < NV; v += gridDim.y) {
#if GR_PROLONG_PREFETCH
            // stage field v, then start loading field v + gridDim.y
#pragma unroll
            for (int r = 0; r < PB_PER_THREAD; ++r)
                if (src[r] >= 0) box[buf][tid + r * PT_N] = next[r];
            if (v + (int)gridDim.y < NV) {
                const double* Un = cu + (long)(v + gridDim.y) * c.stride;
#pragma unroll
                for (int r = 0; r < PB_PER_THREAD; ++r)
                    if (src[r] >= 0) next[r] = Un[src[r]];
            }
#else
            const double* U = cu + (long)v * c.stride;
#pragma unroll
            for (int r = 0; r < PB_PER_THREAD; ++r)
                if (src[r] >= 0) next[r] = U[src[r]];
#pragma unroll
            for (int r = 0; r < PB_PER_THREAD; ++r)
                if (src[r] >= 0) box[buf][tid + r * PT_N] = next[r];
#endif
            __syncthreads();
            if (live) {
                const double* B = box[buf];
                double* out = fu + (long)v * f.stride + fcol;
                for (int kk = 0; kk < kn; ++kk) {
                    const ProlongAxis Z = Ztab[k0 + kk];
                    const int z0 = PB_X * PB_Y * (Z.c0 - bz), z1 = PB_X * PB_Y * (Z.c1 - bz);
                    const double az = Z.ax, bz1 = 1 - Z.ax;
                    // the reference's term order: 000, 100, 010, 001, 110, 101, 011, 111
                    out[(long)fplane * (k0 + kk)] = p00 * bz1 * B[q00 + z0] + p10 * bz1 * B[q10 + z0] +
                        p01 * bz1 * B[q01 + z0] + p00 * az * B[q00 + z1] + p11 * bz1 * B[q11 + z0] +
                        p10 * az * B[q10 + z1] + p01 * az * B[q01 + z1] + p11 * az * B[q11 + z1];
                }
            }
            buf ^= 1;
        }
    } else if (live) {
        const int r00 = X.c0 + c.nt[0] * Y.c0, r10 = X.c1 + c.nt[0] * Y.c0, r01 = X.c0 + c.nt[0] * Y.c1,
                  r11 = X.c1 + c.nt[0] * Y.c1;
        const int cplane = c.nt[0] * c.nt[1];
        for (int v = blockIdx.y; v < NV; v += gridDim.y) {
            const double* U = cu + (long)v * c.stride;
            double* out = fu + (long)v * f.stride + fcol;
            for (int kk = 0; kk < kn; ++kk) {
                const ProlongAxis Z = Ztab[k0 + kk];
                const int z0 = cplane * Z.c0, z1 = cplane * Z.c1;
                const double az = Z.ax, bz1 = 1 - Z.ax;
                out[(long)fplane * (k0 + kk)] = p00 * bz1 * U[r00 + z0] + p10 * bz1 * U[r10 + z0] +
                    p01 * bz1 * U[r01 + z0] + p00 * az * U[r00 + z1] + p11 * bz1 * U[r11 + z0] +
                    p10 * az * U[r10 + z1] + p01 * az * U[r01 + z1] + p11 * az * U[r11 + z1];
            }
        }
    }
}

// 2x2x2 average of fine cells into the covered coarse interior cells
// (src/amr/amr.cpp:566-626, refinement ratio 2).
__global__ void restrict_kernel(BlockDesc f, const double* __restrict__ fu, BlockDesc c, double* __restrict__ cu) {
#pragma clang fp contract(off)
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x); // interior cells < 2^29
    const int ncell = c.n[0] * c.n[1] * c.n[2];
    if (t >= ncell) return;
    const int C[3] = {c.ng + t % c.n[0], c.ng + (t / c.n[0]) % c.n[1], c.ng + t / (c.n[0] * c.n[1])};
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        b[d] = (int)round((coord(c, d, C[d]) - coord(f, d, 0)) / f.dx[d]) - 1;
        if (!(b[d] >= f.ng && b[d] + 2 <= f.ng + f.n[d] + f.ng)) return;
    }
    const long dst = C[0] + (long)c.nt[0] * (C[1] + (long)c.nt[1] * C[2]);
    for (int v = blockIdx.y; v < NV; v += gridDim.y) {
        const double* F = fu + v * f.stride;
        double s = 0.0;
        for (int rk = 0; rk < 2; ++rk)
            for (int rj = 0; rj < 2; ++rj)
                for (int ri = 0; ri < 2; ++ri) s += F[(b[0] + ri) + (long)f.nt[0] * ((b[1] + rj) + (long)f.nt[1] * (b[2] + rk))];
        cu[v * c.stride + dst] = s / 8;
    }
}

// every ghost cell of buffer `sel` of every block of a table; ofs[i] = first ghost cell of
// block i (nblocks + 1 entries)
template <int G>
__global__ void fill_ghosts_batch_kernel(const BatchBlock* __restrict__ table, const int* __restrict__ wg_block,
                                         int sel, int mode) {
    const BatchBlock& B = table[wg_block[blockIdx.x]]; // one block per workgroup
    const long t = (long)(blockIdx.x - B.ghost_wg0) * GHOST_WG + threadIdx.x;
    if (t >= ghost_count(B.b)) return;
    fill_ghost_cell<G>(B.b, (gdouble*)B.buf[sel], t, mode, 0x3f, blockIdx.y);
}

// ------------------------------------------------------------------------------------
// Kreiss-Oliger only (applyDissipation), global-memory stencils, reference order
// ------------------------------------------------------------------------------------
__global__ void dissipation_kernel(BlockDesc b, const double* __restrict__ u, double* rhs, Stencil<double> K) {
#pragma clang fp contract(off)
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ncell = (long)b.n[0] * b.n[1] * b.n[2];
    if (t >= ncell) return;
    const int i = b.ng + (int)(t % b.n[0]);
    const int j = b.ng + (int)((t / b.n[0]) % b.n[1]);
    const int k = b.ng + (int)(t / ((long)b.n[0] * b.n[1]));
    const long c = i + (long)b.nt[0] * (j + (long)b.nt[1] * k);
    const long st[3] = {1, b.nt[0], (long)b.nt[0] * b.nt[1]};
    for (int v = 0; v < NV; ++v) {
        const double* f = u + v * b.stride + c;
        double acc = 0.0;
        for (int d = 0; d < 3; ++d)
            acc = ko_axis_ordered(acc, K.kw[d], f[-3 * st[d]], f[-2 * st[d]], f[-st[d]], f[0], f[st[d]], f[2 * st[d]],
                                  f[3 * st[d]]);
        rhs[v * b.stride + c] += acc;
    }
}

// ------------------------------------------------------------------------------------
// constraint monitor (src/spacetime/ccz4.cpp:941-1159), global-memory stencils
// ------------------------------------------------------------------------------------
struct GlobalCell {
    const double* u;
    long stride, c, st[3];
    double ih[3], ih2[3], ihh[3];
    __device__ double f(int v, long off) const { return u[v * stride + c + off]; }
    __device__ double d1(int v, int d) const {
        const long q = st[d];
        return ih[d] * (-f(v, 2 * q) + 8.0 * f(v, q) - 8.0 * f(v, -q) + f(v, -2 * q));
    }
    __device__ double d2(int v, int a, int b) const {
        if (a == b) {
            const long q = st[a];
            return ih2[a] * (-f(v, 2 * q) + 16.0 * f(v, q) - 30.0 * f(v, 0) + 16.0 * f(v, -q) - f(v, -2 * q));
        }
        const long pa = st[a], pb = st[b];
        const double h = (a == 0) ? ihh[b == 1 ? 0 : 1] : ihh[2];
        return h * (f(v, pa + pb) - f(v, pa - pb) - f(v, -pa + pb) + f(v, -pa - pb));
    }
};

// H at interior cell t of the block and, when mom is not null, M_i into mom[0..2]
__device__ __forceinline__ double constraint_cell(const BlockDesc& b, const double* __restrict__ u, long t,
                                                  const Stencil<double>& K, double* mom) {
    GlobalCell C;
    C.u = u;
    C.stride = b.stride;
    C.c = t;
    C.st[0] = 1;
    C.st[1] = b.nt[0];
    C.st[2] = (long)b.nt[0] * b.nt[1];
    for (int d = 0; d < 3; ++d) {
        C.ih[d] = K.ih[d];
        C.ih2[d] = K.ih2[d];
        C.ihh[d] = K.ihh[d];
    }
    const double chi = C.f(CHI, 0);
    const double Ktr = C.f(KTR, 0);
    double gt[6], At[6], Gh[3];
    for (int a = 0; a < 6; ++a) {
        gt[a] = C.f(GXX + a, 0);
        At[a] = C.f(AXX + a, 0);
    }
    for (int a = 0; a < 3; ++a) Gh[a] = C.f(GHX + a, 0);
    const double det = gt[0] * (gt[3] * gt[5] - gt[4] * gt[4]) - gt[1] * (gt[1] * gt[5] - gt[4] * gt[2]) +
        gt[2] * (gt[1] * gt[4] - gt[3] * gt[2]);
    const double idet = 1.0 / (det + 1.0e-30);
    const double gu[6] = {(gt[3] * gt[5] - gt[4] * gt[4]) * idet, (gt[2] * gt[4] - gt[1] * gt[5]) * idet,
                          (gt[1] * gt[4] - gt[2] * gt[3]) * idet, (gt[0] * gt[5] - gt[2] * gt[2]) * idet,
                          (gt[1] * gt[2] - gt[0] * gt[4]) * idet, (gt[0] * gt[3] - gt[1] * gt[1]) * idet};
    double Au[3][3];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c)
            Au[a][c] = gu[sym(a, 0)] * At[sym(0, c)] + gu[sym(a, 1)] * At[sym(1, c)] + gu[sym(a, 2)] * At[sym(2, c)];
    double AA = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) AA += Au[a][c] * Au[c][a];
    double dgt[6][3], ddgt[6][6], dchi[3], ddchi[6], dGh[3][3];
    for (int v = 0; v < 6; ++v)
        for (int d = 0; d < 3; ++d) {
            dgt[v][d] = C.d1(GXX + v, d);
            for (int e = d; e < 3; ++e) ddgt[v][sym(d, e)] = C.d2(GXX + v, d, e);
        }
    for (int d = 0; d < 3; ++d) {
        dchi[d] = C.d1(CHI, d);
        for (int e = d; e < 3; ++e) ddchi[sym(d, e)] = C.d2(CHI, d, e);
        for (int a = 0; a < 3; ++a) dGh[a][d] = C.d1(GHX + a, d);
    }
    double chr[3][6];
    for (int q = 0; q < 3; ++q)
        for (int p = 0; p < 6; ++p) {
            const int a = sym_a(p), c = sym_b(p);
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s += gu[sym(q, m)] * (dgt[sym(m, a)][c] + dgt[sym(m, c)][a] - dgt[p][m]);
            chr[q][p] = 0.5 * s;
        }
    double G[3][6];
    for (int q = 0; q < 3; ++q)
        for (int p = 0; p < 6; ++p) G[q][p] = gt[sym(q, 0)] * chr[0][p] + gt[sym(q, 1)] * chr[1][p] + gt[sym(q, 2)] * chr[2][p];
    double Gd[3];
    for (int q = 0; q < 3; ++q) {
        double s = 0.0;
        for (int p = 0; p < 6; ++p) s += ((p == 0 || p == 3 || p == 5) ? 1.0 : 2.0) * gu[p] * chr[q][p];
        Gd[q] = s;
    }
    double lapchi = 0.0, gradchi2 = 0.0;
    for (int p = 0; p < 6; ++p) {
        const int l = sym_a(p), m = sym_b(p);
        const double f = (l == m) ? 1.0 : 2.0;
        lapchi += f * gu[p] * (ddchi[p] - chr[0][p] * dchi[0] - chr[1][p] * dchi[1] - chr[2][p] * dchi[2]);
        gradchi2 += f * gu[p] * dchi[l] * dchi[m];
    }
    double Rs = 0.0;
    for (int p = 0; p < 6; ++p) {
        const int a = sym_a(p), c = sym_b(p);
        double R = 0.0;
        for (int q = 0; q < 6; ++q) R += ((q == 0 || q == 3 || q == 5) ? -0.5 : -1.0) * gu[q] * ddgt[p][q];
        for (int q = 0; q < 3; ++q) {
            R += 0.5 * (gt[sym(q, a)] * dGh[q][c] + gt[sym(q, c)] * dGh[q][a]);
            R += (Gh[q] - Gd[q]) * G[q][p];
            for (int l = 0; l < 3; ++l) {
                const double Hc = gu[sym(l, 0)] * G[q][sym(c, 0)] + gu[sym(l, 1)] * G[q][sym(c, 1)] + gu[sym(l, 2)] * G[q][sym(c, 2)];
                const double Ha = gu[sym(l, 0)] * G[q][sym(a, 0)] + gu[sym(l, 1)] * G[q][sym(a, 1)] + gu[sym(l, 2)] * G[q][sym(a, 2)];
                R += chr[q][sym(l, a)] * Hc + chr[q][sym(l, c)] * Ha;
            }
        }
        const double Dab = ddchi[p] - chr[0][p] * dchi[0] - chr[1][p] * dchi[1] - chr[2][p] * dchi[2];
        const double Rc = (0.5 / (chi + 1e-30)) * (Dab + gt[p] * lapchi) -
            (0.25 / ((chi * chi) + 1e-30)) * (dchi[a] * dchi[c] - 3.0 * gt[p] * gradchi2);
        Rs += ((p == 0 || p == 3 || p == 5) ? 1.0 : 2.0) * gu[p] * (R + Rc);
    }
    Rs *= chi;
    const double H = Rs + (2.0 / 3.0) * Ktr * Ktr - AA;
    if (!mom) return H;
    double dAt[6][3], dK[3];
    for (int v = 0; v < 6; ++v)
        for (int d = 0; d < 3; ++d) dAt[v][d] = C.d1(AXX + v, d);
    for (int d = 0; d < 3; ++d) dK[d] = C.d1(KTR, d);
    for (int a = 0; a < 3; ++a) {
        double M = 0.0;
        for (int c = 0; c < 3; ++c)
            for (int q = 0; q < 3; ++q) {
                const double gcq = gu[sym(c, q)];
                double DA = dAt[sym(q, a)][c];
                for (int m = 0; m < 3; ++m) DA -= chr[m][sym(c, q)] * At[sym(m, a)] + chr[m][sym(c, a)] * At[sym(q, m)];
                M += gcq * DA;
                M -= (3.0 / (2.0 * chi + 1e-30)) * At[sym(a, c)] * gcq * dchi[q];
            }
        M -= (2.0 / 3.0) * dK[a];
        mom[a] = M;
    }
    return H;
}

__global__ void constraints_kernel(BlockDesc b, const double* __restrict__ u, double* ham, double* mom,
                                   Stencil<double> K) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.stride) return;
    const int i = (int)(t % b.nt[0]);
    const int j = (int)((t / b.nt[0]) % b.nt[1]);
    const int k = (int)(t / ((long)b.nt[0] * b.nt[1]));
    const bool inner = i >= b.ng && i < b.ng + b.n[0] && j >= b.ng && j < b.ng + b.n[1] && k >= b.ng &&
        k < b.ng + b.n[2];
    if (!inner) {
        ham[t] = 0.0;
        if (mom) mom[3 * t] = mom[3 * t + 1] = mom[3 * t + 2] = 0.0;
        return;
    }
    ham[t] = constraint_cell(b, u, t, K, mom ? mom + 3 * t : nullptr);
}


// ------------------------------------------------------------------------------------
// initial data and noise
// ------------------------------------------------------------------------------------
__global__ void init_bl_kernel(BlockDesc b, double* u, int nbh, const double* mass, const double* pos) {
#pragma clang fp contract(off)
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.stride) return;
    const int i = (int)(t % b.nt[0]);
    const int j = (int)((t / b.nt[0]) % b.nt[1]);
    const int k = (int)(t / ((long)b.nt[0] * b.nt[1]));
    const double x = coord(b, 0, i), y = coord(b, 1, j), z = coord(b, 2, k);
    // flat background (src/spacetime/ccz4.cpp:1165-1188) then the BL conformal factor
    // (src/initial_data/initial_data.cpp:79-158)
    for (int v = 0; v < NV; ++v) u[v * b.stride + t] = 0.0;
    u[(GXX + 0) * b.stride + t] = 1.0;
    u[(GXX + 3) * b.stride + t] = 1.0;
    u[(GXX + 5) * b.stride + t] = 1.0;
    double psi = 1.0;
    for (int h = 0; h < nbh; ++h) {
        const double ax = x - pos[3 * h], ay = y - pos[3 * h + 1], az = z - pos[3 * h + 2];
        const double r = sqrt(ax * ax + ay * ay + az * az + 1.0e-30);
        psi += mass[h] / (2.0 * r);
    }
    const double p2 = psi * psi;
    const double chi = 1.0 / (p2 * p2);
    u[CHI * b.stride + t] = chi;
    u[ALPHA * b.stride + t] = sqrt(chi);
    double dp[3] = {0.0, 0.0, 0.0};
    for (int h = 0; h < nbh; ++h) {
        const double rx = x - pos[3 * h], ry = y - pos[3 * h + 1], rz = z - pos[3 * h + 2];
        const double r2 = rx * rx + ry * ry + rz * rz + 1.0e-20;
        const double r3 = r2 * sqrt(r2);
        const double c = -mass[h] / (2.0 * r3);
        dp[0] += c * rx;
        dp[1] += c * ry;
        dp[2] += c * rz;
    }
    const double gf = 8.0 / (3.0 * psi);
    for (int a = 0; a < 3; ++a) u[(GHX + a) * b.stride + t] = gf * dp[a];
}

__device__ __forceinline__ uint64_t sm64(uint64_t s) {
    s += 0x9E3779B97F4A7C15ull;
    s = (s ^ (s >> 30)) * 0xBF58476D1CE4E5B9ull;
    s = (s ^ (s >> 27)) * 0x94D049BB133111EBull;
    return s ^ (s >> 31);
}

__global__ void add_noise_kernel(BlockDesc b, double* u, uint64_t seed, double amp, unsigned mask) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.stride) return;
    for (int v = 0; v < NV; ++v) {
        if (!((mask >> v) & 1u)) continue;
        const uint64_t q = (uint64_t)v * (uint64_t)b.stride + (uint64_t)t;
        const uint64_t z1 = sm64(seed + (2 * q) * 0x9E3779B97F4A7C15ull);
        const uint64_t z2 = sm64(seed + (2 * q + 1) * 0x9E3779B97F4A7C15ull);
        const double u1 = ((double)(z1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(z2 >> 11) * (1.0 / 9007199254740992.0);
        u[v * b.stride + t] += amp * (sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2));
    }
}

// ------------------------------------------------------------------------------------
// reductions (fixed order: per-workgroup tree, then one workgroup over the partials)
// ------------------------------------------------------------------------------------
constexpr int RED_T = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = RED_T / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    return sh[0];
}

// sum over the interior cells of the squares of an array of ncomp values per cell
// (cell-major, as the reference's ham and mom arrays)
__global__ __launch_bounds__(RED_T) void sumsq_kernel(BlockDesc b, const double* ham, int ncomp, double* partial,
                                                      long per) {
    __shared__ double sh[RED_T];
    const long ncell = (long)b.n[0] * b.n[1] * b.n[2];
    const long beg = (long)blockIdx.x * per;
    const long end = beg + per < ncell ? beg + per : ncell;
    double acc = 0.0;
    for (long t = beg + threadIdx.x; t < end; t += RED_T) {
        const int i = b.ng + (int)(t % b.n[0]);
        const int j = b.ng + (int)((t / b.n[0]) % b.n[1]);
        const int k = b.ng + (int)(t / ((long)b.n[0] * b.n[1]));
        const long c = i + (long)b.nt[0] * (j + (long)b.nt[1] * k);
        for (int a = 0; a < ncomp; ++a) {
            const double h = ham[ncomp * c + a];
            acc += h * h;
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// sum over the interior of H^2 with H formed on the fly (no per-cell H array): the partition
// and the visiting order of sumsq_kernel over constraints_kernel's output
__global__ __launch_bounds__(RED_T) void ham_sumsq_kernel(BlockDesc b, const double* __restrict__ u,
                                                          Stencil<double> K, double* partial, long per) {
    __shared__ double sh[RED_T];
    const long ncell = (long)b.n[0] * b.n[1] * b.n[2];
    const long beg = (long)blockIdx.x * per;
    const long end = beg + per < ncell ? beg + per : ncell;
    double acc = 0.0;
    for (long t = beg + threadIdx.x; t < end; t += RED_T) {
        const int i = b.ng + (int)(t % b.n[0]);
        const int j = b.ng + (int)((t / b.n[0]) % b.n[1]);
        const int k = b.ng + (int)(t / ((long)b.n[0] * b.n[1]));
        const long c = i + (long)b.nt[0] * (j + (long)b.nt[1] * k);
        const double h = constraint_cell(b, u, c, K, nullptr);
        acc += h * h;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(RED_T) void nonfinite_kernel(BlockDesc b, const double* u, double* partial, long per) {
    __shared__ double sh[RED_T];
    const long ncell = (long)b.n[0] * b.n[1] * b.n[2];
    const long beg = (long)blockIdx.x * per;
    const long end = beg + per < ncell ? beg + per : ncell;
    double acc = 0.0;
    for (long t = beg + threadIdx.x; t < end; t += RED_T) {
        const int i = b.ng + (int)(t % b.n[0]);
        const int j = b.ng + (int)((t / b.n[0]) % b.n[1]);
        const int k = b.ng + (int)(t / ((long)b.n[0] * b.n[1]));
        const long c = i + (long)b.nt[0] * (j + (long)b.nt[1] * k);
        for (int v = 0; v < NV; ++v) acc += __builtin_isfinite(u[v * b.stride + c]) ? 0.0 : 1.0;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(RED_T) void sum_partials_kernel(const double* partial, int n, double* out) {
    __shared__ double sh[RED_T];
    double acc = 0.0;
    for (int t = threadIdx.x; t < n; t += RED_T) acc += partial[t];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[0] = s;
}

// First non-finite interior value in the order of the reference's NaN scan
// (src/main.cpp:1192-1214: field v, then k, j, i over the interior). Position q of field v
// at interior cell (i', j', k') is v * ncell + (k' * n1 + j') * n0 + i'; each workgroup
// takes the contiguous range [blockIdx.x * per, ...) of positions (consecutive threads on
// consecutive x: coalesced) and writes the smallest non-finite position in it, or ~0.
__global__ __launch_bounds__(RED_T) void first_nonfinite_kernel(BlockDesc b, const double* u,
                                                                unsigned long long* partial, long per) {
    __shared__ unsigned long long sh[RED_T];
    const long ncell = (long)b.n[0] * b.n[1] * b.n[2];
    const long total = NV * ncell;
    const long beg = (long)blockIdx.x * per;
    const long end = beg + per < total ? beg + per : total;
    unsigned long long best = ~0ull;
    for (long q = beg + threadIdx.x; q < end; q += RED_T) {
        const int v = (int)(q / ncell);
        const long t = q - v * ncell;
        const int i = b.ng + (int)(t % b.n[0]);
        const int j = b.ng + (int)((t / b.n[0]) % b.n[1]);
        const int k = b.ng + (int)(t / ((long)b.n[0] * b.n[1]));
        const double x = u[v * b.stride + i + (long)b.nt[0] * (j + (long)b.nt[1] * k)];
        if (!__builtin_isfinite(x)) {
            best = (unsigned long long)q; // the thread's positions rise: the first is its smallest
            break;
        }
    }
    sh[threadIdx.x] = best;
    __syncthreads();
    for (int o = RED_T / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && sh[threadIdx.x + o] < sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(RED_T) void min_partials_kernel(const unsigned long long* partial, int n,
                                                             unsigned long long* out) {
    __shared__ unsigned long long sh[RED_T];
    unsigned long long m = ~0ull;
    for (int t = threadIdx.x; t < n; t += RED_T) m = partial[t] < m ? partial[t] : m;
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = RED_T / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && sh[threadIdx.x + o] < sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// the reference's max_adv_cfl = std::max(max_adv_cfl, local) keeps the running value when
// local is NaN (the comparison is false), so NaN never enters the maximum
__device__ __forceinline__ double cfl_max2(double a, double b) { return (a < b) ? b : a; }

// per-substep advective CFL guard (src/main.cpp:1155-1176): over the interior,
// |bx| * dt / dx0 + |by| * dt / dx1 + |bz| * dt / dx2, maximum per workgroup
__global__ __launch_bounds__(RED_T) void cfl_max_kernel(BlockDesc b, const double* u, double dt, double* partial,
                                                        long per) {
#pragma clang fp contract(off)
    __shared__ double sh[RED_T];
    const long ncell = (long)b.n[0] * b.n[1] * b.n[2];
    const long beg = (long)blockIdx.x * per;
    const long end = beg + per < ncell ? beg + per : ncell;
    double m = 0.0;
    for (long t = beg + threadIdx.x; t < end; t += RED_T) {
        const int i = b.ng + (int)(t % b.n[0]);
        const int j = b.ng + (int)((t / b.n[0]) % b.n[1]);
        const int k = b.ng + (int)(t / ((long)b.n[0] * b.n[1]));
        const long c = i + (long)b.nt[0] * (j + (long)b.nt[1] * k);
        const double bx = fabs(u[BX * b.stride + c]), by = fabs(u[(BX + 1) * b.stride + c]),
                     bz = fabs(u[(BX + 2) * b.stride + c]);
        const double local = bx * dt / b.dx[0] + by * dt / b.dx[1] + bz * dt / b.dx[2];
        m = cfl_max2(m, local);
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = RED_T / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = cfl_max2(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(RED_T) void max_partials_kernel(const double* partial, int n, double* out) {
    __shared__ double sh[RED_T];
    double m = 0.0;
    for (int t = threadIdx.x; t < n; t += RED_T) m = cfl_max2(m, partial[t]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = RED_T / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = cfl_max2(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// setGaugeWaveData (src/spacetime/ccz4.cpp:1190-1221): H = A sin(2 pi x / L) (the product
// 2 pi x formed left to right, then divided by L), chi = (1 + H)^(-1/3), gt_xx = (1 + H) chi,
// gt_yy = gt_zz = chi, alpha = sqrt(1 + H); the fields it does not set are 0, as on a
// freshly constructed GridBlock
__global__ void init_gauge_wave_kernel(BlockDesc b, double* u, double A, double L) {
#pragma clang fp contract(off)
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.stride) return;
    const int i = (int)(t % b.nt[0]);
    const double x = coord(b, 0, i);
    const double H = A * sin(2.0 * PI * x / L);
    const double detg = 1.0 + H;
    const double chi = pow(detg, -1.0 / 3.0);
    for (int v = 0; v < NV; ++v) u[v * b.stride + t] = 0.0;
    u[CHI * b.stride + t] = chi;
    u[(GXX + 0) * b.stride + t] = (1.0 + H) * chi;
    u[(GXX + 3) * b.stride + t] = chi;
    u[(GXX + 5) * b.stride + t] = chi;
    u[ALPHA * b.stride + t] = sqrt(1.0 + H);
}

inline unsigned grid_for(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

} // namespace

Stencil<double> make_stencil(const BlockDesc& b, double ko_sigma) {
    Stencil<double> K;
    static const double wts[4] = {1.0, -6.0, 15.0, -20.0};
    for (int d = 0; d < 3; ++d) {
        K.ih[d] = 1.0 / (12.0 * b.dx[d]);
        K.ih2[d] = 1.0 / (12.0 * b.dx[d] * b.dx[d]);
        K.ko[d] = ko_sigma / (64.0 * b.dx[d]);
        for (int s = 0; s < 4; ++s) K.kw[d][s] = K.ko[d] * wts[s];
    }
    K.ihh[0] = 1.0 / (4.0 * b.dx[0] * b.dx[1]);
    K.ihh[1] = 1.0 / (4.0 * b.dx[0] * b.dx[2]);
    K.ihh[2] = 1.0 / (4.0 * b.dx[1] * b.dx[2]);
    return K;
}

void stage_tile_fp64(int* t3) {
    t3[0] = FP64_TX;
    t3[1] = FP64_TY;
    t3[2] = FP64_TZ;
}

void stage_tile_mixed(int* t3) {
    t3[0] = TILE_X;
    t3[1] = TILE_Y;
    t3[2] = TILE_Z;
}

#ifndef GR_XCD_REMAP
#define GR_XCD_REMAP 1
#endif

static StageArgs stage_args(const Batch& B, const int sel[3], const CellParams& P, double ko_sigma,
                            const StageCoef& cf, const unsigned char* flags, int want) {
    StageArgs A;
    A.batch = B;
    A.sel_in = sel[0];
    A.sel_un = sel[1];
    A.sel_out = sel[2];
    A.ntiles = B.ntiles;
    A.xcd_remap = GR_XCD_REMAP ? 1 : 0;
    BlockDesc g;
    for (int d = 0; d < 3; ++d) g.dx[d] = B.dx[d];
    A.K = make_stencil(g, ko_sigma);
    for (int d = 0; d < 3; ++d) {
        A.K32.ih[d] = (float)A.K.ih[d];
        A.K32.ih2[d] = (float)A.K.ih2[d];
        A.K32.ihh[d] = (float)A.K.ihh[d];
        A.K32.ko[d] = (float)A.K.ko[d];
        for (int t = 0; t < 4; ++t) A.K32.kw[d][t] = (float)A.K.kw[d][t];
    }
    A.P = P;
    A.cf = cf;
    A.ko_on = ko_sigma <= 0.0 ? 0 : 1; // applyDissipation returns early only for sigma <= 0 (NaN dissipates)
    A.flags = flags;
    A.want = want;
    A.scan = 0;
    return A;
}

int stage_tiles(const BlockDesc& b) {
    return ((b.n[0] + TILE_X - 1) / TILE_X) * ((b.n[1] + TILE_Y - 1) / TILE_Y) * ((b.n[2] + TILE_Z - 1) / TILE_Z);
}

static long host_ghost_count(const BlockDesc& b) {
    return 2L * b.ng * b.nt[1] * b.nt[0] + (long)b.n[2] * 2 * b.ng * b.nt[0] + (long)b.n[2] * b.n[1] * 2 * b.ng;
}

int batch_layout(std::vector<BatchBlock>& blocks, std::vector<int>& tile_block, std::vector<int>& ghost_wg_block) {
    int t = 0;
    tile_block.clear();
    ghost_wg_block.clear();
    for (size_t i = 0; i < blocks.size(); ++i) {
        BatchBlock& B = blocks[i];
        B.tiles[0] = (B.b.n[0] + TILE_X - 1) / TILE_X;
        B.tiles[1] = (B.b.n[1] + TILE_Y - 1) / TILE_Y;
        B.tiles[2] = (B.b.n[2] + TILE_Z - 1) / TILE_Z;
        B.tile0 = t;
        const int nt = B.tiles[0] * B.tiles[1] * B.tiles[2];
        tile_block.insert(tile_block.end(), nt, (int)i);
        t += nt;
        B.ghost_wg0 = (int)ghost_wg_block.size();
        ghost_wg_block.insert(ghost_wg_block.end(), grid_for(host_ghost_count(B.b), GHOST_WG), (int)i);
    }
    return t;
}

bool pairs_aligned(const BlockDesc& b, const void* p) {
    return b.nt[0] % 2 == 0 && (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

Batch single_batch(const BlockDesc& b, const double* in, const double* un, double* out, const Sources& src) {
    Batch B;
    B.pairs_ok = pairs_aligned(b, in) && pairs_aligned(b, un) && pairs_aligned(b, out);
    B.variant = 0;
    B.ghost_groups = GR_GHOST_GROUPS;
    B.table = nullptr;
    B.tile_block = nullptr;
    B.ghost_wg_block = nullptr;
    B.nblocks = 1;
    B.one.b = b;
    B.one.buf[0] = const_cast<double*>(in);
    B.one.buf[1] = const_cast<double*>(un);
    B.one.buf[2] = out;
    B.one.src = src;
    B.one.tile0 = 0;
    B.one.tiles[0] = (b.n[0] + TILE_X - 1) / TILE_X;
    B.one.tiles[1] = (b.n[1] + TILE_Y - 1) / TILE_Y;
    B.one.tiles[2] = (b.n[2] + TILE_Z - 1) / TILE_Z;
    B.ntiles = B.one.tiles[0] * B.one.tiles[1] * B.one.tiles[2];
    B.one.ghost_wg0 = 0;
    B.ghost_wgs = (int)grid_for(host_ghost_count(b), GHOST_WG);
    for (int d = 0; d < 3; ++d) B.dx[d] = b.dx[d];
    return B;
}

hipError_t launch_stage_fp64(const Batch& B, const int sel[3], bool has_src, const CellParams& P,
                             double ko_sigma, int mode, const StageCoef& cf, hipStream_t s,
                             const unsigned char* flags) {
    // the direct LDS staging reads 16-byte pairs: other batches run the register staging
    const int variant = (B.variant != 0 && B.pairs_ok) ? B.variant : 0;
    StageArgs A = stage_args(B, sel, P, ko_sigma, cf, flags, 1);
    // with flags (the FP64 tiles of the mixed form: few, near the puncture) a small grid
    // gathers them instead of launching one workgroup per part of every tile of the level
    int nblk = 0;
    if (flags) {
        A.scan = 1;
        nblk = A.ntiles < FP64_SCAN_GRID ? A.ntiles : FP64_SCAN_GRID;
    }
    if (has_src) {
        if (mode == MODE_RHS) return launch_stage_t<MODE_RHS, true>(A, nblk, variant, s);
        if (mode == MODE_STAGE) return launch_stage_t<MODE_STAGE, true>(A, nblk, variant, s);
        return launch_stage_t<MODE_FINAL, true>(A, nblk, variant, s);
    }
    if (mode == MODE_RHS) return launch_stage_t<MODE_RHS, false>(A, nblk, variant, s);
    if (mode == MODE_STAGE) return launch_stage_t<MODE_STAGE, false>(A, nblk, variant, s);
    return launch_stage_t<MODE_FINAL, false>(A, nblk, variant, s);
}

template <int MODE, bool HAS_SRC>
static hipError_t launch_mixed_t(const StageArgs& A, hipStream_t s) {
    constexpr int PX = TILE_X / GR_MIXED_SPLIT;
    // the first stage (u^n is the stage input) has an instance of its own, so that no load
    // of the kernel's register prefetch is conditional
    if (MODE == MODE_STAGE && A.sel_un == A.sel_in)
        stage_kernel_mixed<PX, TILE_Y, TILE_Z, MODE, HAS_SRC, MODE == MODE_STAGE>
            <<<dim3(A.ntiles * GR_MIXED_SPLIT), dim3(PX * TILE_Y * TILE_Z), 0, s>>>(A);
    else
        stage_kernel_mixed<PX, TILE_Y, TILE_Z, MODE, HAS_SRC>
            <<<dim3(A.ntiles * GR_MIXED_SPLIT), dim3(PX * TILE_Y * TILE_Z), 0, s>>>(A);
    return hipGetLastError();
}

hipError_t launch_stage_mixed(const Batch& B, const int sel[3], bool has_src, const CellParams& P,
                              double ko_sigma, int mode, const StageCoef& cf, hipStream_t s, unsigned char* flags,
                              double chi_fp64_below, const AuxStream* aux) {
    const StageArgs A = stage_args(B, sel, P, ko_sigma, cf, flags, 0);
    classify_tiles_kernel<TILE_X, TILE_Y, TILE_Z>
        <<<dim3(A.ntiles), dim3(TILE_X * TILE_Y * TILE_Z), 0, s>>>(A, chi_fp64_below, flags);
    hipError_t st = hipGetLastError();
    if (st != hipSuccess) return st;
    // the FP64 tiles (flag 1) run beside the FP32 kernel on the auxiliary stream: both read
    // the stage input and write disjoint tiles of the output
    if (aux) {
        if ((st = hipEventRecord(aux->fork, s)) != hipSuccess) return st;
        if ((st = hipStreamWaitEvent(aux->stream, aux->fork, 0)) != hipSuccess) return st;
        if ((st = launch_stage_fp64(B, sel, has_src, P, ko_sigma, mode, cf, aux->stream, flags)) != hipSuccess)
            return st;
        if ((st = hipEventRecord(aux->join, aux->stream)) != hipSuccess) return st;
    }
    // FP32 tiles (flag 0)
    if (has_src) {
        if (mode == MODE_RHS) st = launch_mixed_t<MODE_RHS, true>(A, s);
        else if (mode == MODE_STAGE) st = launch_mixed_t<MODE_STAGE, true>(A, s);
        else st = launch_mixed_t<MODE_FINAL, true>(A, s);
    } else {
        if (mode == MODE_RHS) st = launch_mixed_t<MODE_RHS, false>(A, s);
        else if (mode == MODE_STAGE) st = launch_mixed_t<MODE_STAGE, false>(A, s);
        else st = launch_mixed_t<MODE_FINAL, false>(A, s);
    }
    if (st != hipSuccess) return st;
    if (aux) return hipStreamWaitEvent(s, aux->join, 0);
    // FP64 tiles near the puncture (flag 1)
    return launch_stage_fp64(B, sel, has_src, P, ko_sigma, mode, cf, s, flags);
}

bool ghost_groups_valid(int g) { return g == 1 || g == 2 || g == 4 || g == 11; }

template <int G>
static void fill_batch_g(const Batch& B, int sel, int mode, hipStream_t s) {
    fill_ghosts_batch_kernel<G><<<dim3(B.ghost_wgs, G), GHOST_WG, 0, s>>>(B.table, B.ghost_wg_block, sel, mode);
}

hipError_t launch_fill_ghosts_batch(const Batch& B, int sel, int mode, hipStream_t s) {
    if (!B.table) return launch_fill_ghosts(B.one.b, B.one.buf[sel], mode, s, 0x3f, B.ghost_groups);
    switch (B.ghost_groups) {
    case 1: fill_batch_g<1>(B, sel, mode, s); break;
    case 2: fill_batch_g<2>(B, sel, mode, s); break;
    case 11: fill_batch_g<11>(B, sel, mode, s); break;
    default: fill_batch_g<4>(B, sel, mode, s); break;
    }
    return hipGetLastError();
}

template <int G>
static void fill_one_g(const BlockDesc& b, double* u, int mode, hipStream_t s, int faces, long n) {
    fill_ghosts_kernel<G><<<dim3(grid_for(n, GHOST_WG), G), GHOST_WG, 0, s>>>(b, u, mode, faces);
}

hipError_t launch_fill_ghosts(const BlockDesc& b, double* u, int mode, hipStream_t s, int faces, int groups) {
    const long nA = 2L * b.ng * b.nt[1] * b.nt[0];
    const long nB = (long)b.n[2] * 2 * b.ng * b.nt[0];
    const long nC = (long)b.n[2] * b.n[1] * 2 * b.ng;
    const long n = nA + nB + nC;
    switch (groups) {
    case 1: fill_one_g<1>(b, u, mode, s, faces, n); break;
    case 2: fill_one_g<2>(b, u, mode, s, faces, n); break;
    case 11: fill_one_g<11>(b, u, mode, s, faces, n); break;
    default: fill_one_g<4>(b, u, mode, s, faces, n); break;
    }
    return hipGetLastError();
}

hipError_t launch_prolong_axes(const BlockDesc& c, const BlockDesc& f, ProlongAxis* tab, hipStream_t s) {
    prolong_axes_kernel<<<grid_for(f.nt[0] + f.nt[1] + f.nt[2], 256), 256, 0, s>>>(c, f, tab);
    return hipGetLastError();
}

hipError_t launch_prolongate(const BlockDesc& c, const double* cu, const BlockDesc& f, double* fu,
                             const ProlongAxis* tab, hipStream_t s, int groups) {
    const int ntx = (f.nt[0] + PT_X - 1) / PT_X, nty = (f.nt[1] + PT_Y - 1) / PT_Y, ntz = (f.nt[2] + PT_Z - 1) / PT_Z;
    const int g = groups > 0 && groups <= NV ? groups : GR_TRANSFER_GROUPS;
    prolong_kernel<<<dim3(ntx * nty * ntz, g), PT_N, 0, s>>>(c, cu, f, fu, tab, ntx, nty);
    return hipGetLastError();
}

hipError_t launch_restrict(const BlockDesc& f, const double* fu, const BlockDesc& c, double* cu, hipStream_t s,
                           int groups) {
    const long n = (long)c.n[0] * c.n[1] * c.n[2];
    const int g = groups > 0 && groups <= NV ? groups : GR_TRANSFER_GROUPS;
    restrict_kernel<<<dim3(grid_for(n, 256), g), 256, 0, s>>>(f, fu, c, cu);
    return hipGetLastError();
}

hipError_t launch_dissipation(const BlockDesc& b, const double* u, double* rhs, double ko_sigma, hipStream_t s) {
    if (ko_sigma <= 0.0) return hipSuccess;
    const long n = (long)b.n[0] * b.n[1] * b.n[2];
    dissipation_kernel<<<grid_for(n, 256), 256, 0, s>>>(b, u, rhs, make_stencil(b, ko_sigma));
    return hipGetLastError();
}

hipError_t launch_constraints(const BlockDesc& b, const double* u, double* ham, double* mom, hipStream_t s) {
    constraints_kernel<<<grid_for(b.stride, 128), 128, 0, s>>>(b, u, ham, mom,
                                                              make_stencil(b, 0.0));
    return hipGetLastError();
}

hipError_t launch_init_bl(const BlockDesc& b, double* u, int nbh, const double* mass, const double* pos,
                          hipStream_t s) {
    init_bl_kernel<<<grid_for(b.stride, 256), 256, 0, s>>>(b, u, nbh, mass, pos);
    return hipGetLastError();
}

hipError_t launch_add_noise(const BlockDesc& b, double* u, unsigned long long seed, double amp, unsigned mask,
                            hipStream_t s) {
    add_noise_kernel<<<grid_for(b.stride, 256), 256, 0, s>>>(b, u, seed, amp, mask);
    return hipGetLastError();
}

static long reduce_split(const BlockDesc& b, int cap, int* nparts) {
    const long ncell = (long)b.n[0] * b.n[1] * b.n[2];
    long per = (ncell + cap - 1) / cap;
    if (per < RED_T) per = RED_T;
    *nparts = (int)((ncell + per - 1) / per);
    return per;
}

hipError_t launch_sumsq_interior(const BlockDesc& b, const double* ham, int ncomp, double* partial, int cap,
                                 int* nparts, hipStream_t s) {
    const long per = reduce_split(b, cap, nparts);
    sumsq_kernel<<<*nparts, RED_T, 0, s>>>(b, ham, ncomp, partial, per);
    return hipGetLastError();
}

int reduce_parts(const BlockDesc& b, int cap) {
    int n = 0;
    (void)reduce_split(b, cap, &n);
    return n;
}

hipError_t launch_ham_sumsq(const BlockDesc& b, const double* u, double* partial, int cap, int* nparts,
                            hipStream_t s) {
    const long per = reduce_split(b, cap, nparts);
    ham_sumsq_kernel<<<*nparts, RED_T, 0, s>>>(b, u, make_stencil(b, 0.0), partial, per);
    return hipGetLastError();
}

hipError_t launch_count_nonfinite(const BlockDesc& b, const double* u, double* partial, int cap, int* nparts,
                                  hipStream_t s) {
    const long per = reduce_split(b, cap, nparts);
    nonfinite_kernel<<<*nparts, RED_T, 0, s>>>(b, u, partial, per);
    return hipGetLastError();
}

hipError_t launch_first_nonfinite(const BlockDesc& b, const double* u, unsigned long long* partial, int cap,
                                  int* nparts, hipStream_t s) {
    const long total = NV * (long)b.n[0] * b.n[1] * b.n[2];
    long per = (total + cap - 1) / cap;
    if (per < RED_T) per = RED_T;
    *nparts = (int)((total + per - 1) / per);
    first_nonfinite_kernel<<<*nparts, RED_T, 0, s>>>(b, u, partial, per);
    return hipGetLastError();
}

hipError_t launch_min_partials(const unsigned long long* partial, int n, unsigned long long* out, hipStream_t s) {
    min_partials_kernel<<<1, RED_T, 0, s>>>(partial, n, out);
    return hipGetLastError();
}

hipError_t launch_cfl_max(const BlockDesc& b, const double* u, double dt, double* partial, int cap, int* nparts,
                          hipStream_t s) {
    const long per = reduce_split(b, cap, nparts);
    cfl_max_kernel<<<*nparts, RED_T, 0, s>>>(b, u, dt, partial, per);
    return hipGetLastError();
}

hipError_t launch_max_partials(const double* partial, int n, double* out, hipStream_t s) {
    max_partials_kernel<<<1, RED_T, 0, s>>>(partial, n, out);
    return hipGetLastError();
}

hipError_t launch_init_gauge_wave(const BlockDesc& b, double* u, double A, double L, hipStream_t s) {
    init_gauge_wave_kernel<<<grid_for(b.stride, 256), 256, 0, s>>>(b, u, A, L);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double* partial, int n, double* out, hipStream_t s) {
    sum_partials_kernel<<<1, RED_T, 0, s>>>(partial, n, out);
    return hipGetLastError();
}

} // namespace gr
