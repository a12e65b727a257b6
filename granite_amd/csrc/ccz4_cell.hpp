// Per-cell CCZ4 right-hand side, evaluated field by field.
//
// The update of one cell is written against two small interfaces so that the same
// source runs inside the LDS-tiled HIP kernel and in a host build used by the tests:
//   Stager  S : R centre(v)        value of field v at the cell (stage input)
//               const R* field(v)  pointer to field v at the cell inside a staged array;
//                                  neighbours are reached with the strides in Stencil::st.
//                                  Called once per field, in a fixed order, by every thread.
//   Sink    O : put(v, rhs, ko, u_in)  receives the RHS value of field v (without the
//                                  Kreiss-Oliger term `ko`, which the sink adds when enabled)
//               finish()           called once after all 22 fields.
// The arithmetic follows GRANITE's CCZ4Evolution::computeRHS (src/spacetime/ccz4.cpp:251-865)
// and applyDissipation (src/spacetime/ccz4.cpp:871-935), with common subexpressions of the
// Ricci tensor, A~A~ and the lapse Laplacian factored once per cell (same mathematics,
// different rounding order; parity is normwise, not bitwise).
//
// Fields are streamed in the order beta, conformal metric, chi, alpha, K, Gamma-hat, Theta,
// A~ so that each field's stencil results are folded into a small set of accumulators as
// soon as they are produced: the Christoffel symbols exist before chi and alpha are staged,
// so their second derivatives are consumed at once, and every term that needs the
// Christoffel symbols is formed before K is staged; across the later fields only the
// shift's first derivatives, the lapse's Hessian, the Ricci tensor and the Gamma-hat
// accumulators stay live.
#pragma once

#ifndef __HIPCC__
#include <cmath>
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#define GR_INLINE inline
#else
#define GR_INLINE __forceinline__
#endif

namespace gr {

constexpr int NV = 22;
enum : int {
    CHI = 0, GXX = 1, AXX = 7, KTR = 13, GHX = 14, THETA = 17, ALPHA = 18, BX = 19
};
constexpr double PI = 3.14159265358979323846;

// Order in which ccz4_cell stages the fields (Stager::field calls). A stager may rely on it
// to prefetch the next fields while the current one is being evaluated.
constexpr int FIELD_ORDER[NV] = {19, 20, 21, 1, 2, 3, 4, 5, 6, 0, 18, 13, 14, 15, 16, 17, 7, 8, 9, 10, 11, 12};
__host__ __device__ constexpr int field_pos(int v) {
    int p = 0;
    while (FIELD_ORDER[p] != v) ++p;
    return p;
}

// symmetric index pair -> 0..5 (xx, xy, xz, yy, yz, zz), include/granite/core/types.hpp:52-54
__host__ __device__ constexpr int sym(int a, int b) {
    return a <= b ? (a * 3 - a * (a - 1) / 2 + b - a) : (b * 3 - b * (b - 1) / 2 + a - b);
}
// first / second index of a symmetric pair
__host__ __device__ constexpr int sym_a(int p) { return p < 3 ? 0 : (p < 5 ? 1 : 2); }
__host__ __device__ constexpr int sym_b(int p) { return p < 3 ? p : (p < 5 ? p - 2 : 2); }

// Stencil constants of one block, computed on the host with the reference's operations:
// ih = 1/(12 h) (ccz4.cpp:72,111), ih2 = 1/(12 h h) (ccz4.cpp:158), ihh = 1/(4 h_a h_b) for
// (x,y),(x,z),(y,z) (ccz4.cpp:188), ko = sigma/(64 h) (ccz4.cpp:886-887).
// kw[d][s] = ko[d] * (1,-6,15,-20)[s], the per-tap products the reference forms before
// accumulating (ccz4.cpp:921-924); taps 4..6 repeat taps 2..0 (the weights are symmetric, so
// the reference's products are the same numbers), kept once to spare scalar registers.
template <typename R>
struct Stencil {
    R ih[3];
    R ih2[3];
    R ihh[3];
    R ko[3];
    R kw[3][4];
};

// Index strides of a staged array along x, y, z. The LDS tile uses compile-time strides
// so that every stencil tap is an immediate offset from one base address.
template <int SX, int SY, int SZ>
struct FixedStrides {
    __host__ __device__ static constexpr long q(int d) { return d == 0 ? SX : (d == 1 ? SY : SZ); }
};
struct RuntimeStrides {
    long s[3];
    __host__ __device__ long q(int d) const { return s[d]; }
};

struct CellParams {
    double kappa1, kappa2, eta, chi_c, chi_w;
};

template <typename R>
__host__ __device__ GR_INLINE bool finite_(R x) {
#ifdef __HIPCC__
    return __builtin_isfinite(x);
#else
    return std::isfinite(x);
#endif
}

template <typename R>
__host__ __device__ GR_INLINE R tanh_(R x) {
    using std::tanh;
    return tanh(x);
}

// Forces x to be computed at this point of the instruction stream. Used on each field's
// stencil results so that the compiler evaluates them before the next field is staged,
// instead of keeping that field's raw neighbour values alive across later fields.
template <typename R>
__host__ __device__ GR_INLINE void pin(R& x) {
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+v"(x));
#else
    (void)x;
#endif
}

// Kreiss-Oliger taps of one axis accumulated exactly as the reference does: 7 rounded
// products added in tap order onto the running sum over axes (ccz4.cpp:912-925). Kept
// free of FMA contraction, so fields whose RHS is dominated by dissipation (a constant or
// noisy metric) reproduce the reference to the last bit.
template <typename R>
__host__ __device__ GR_INLINE R ko_axis_ordered(R acc, const R* kw, R m3, R m2, R m1, R c, R p1, R p2, R p3) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    acc = acc + kw[0] * m3;
    acc = acc + kw[1] * m2;
    acc = acc + kw[2] * m1;
    acc = acc + kw[3] * c;
    acc = acc + kw[2] * p1;
    acc = acc + kw[1] * p2;
    acc = acc + kw[0] * p3;
    return acc;
}

// All stencils of one field at one cell: 4th-order centred first derivative, one-sided
// 5-point upwind derivative picked by the sign of beta (used only through the advection
// sum), 7-point Kreiss-Oliger sum, and (D2) the 4th-order pure and 2nd-order mixed second
// derivatives. dd is stored by symmetric pair index. KO_EXACT selects the reference's
// accumulation order for the dissipation sum; otherwise it is grouped by symmetric taps.
template <bool D2, bool KO_EXACT, typename R, class ST>
__host__ __device__ GR_INLINE void eval_field(const R* s, const Stencil<R>& K, const ST& st, const R be[3], R w,
                                              R d1[3], R dd[6], R& adv, R& ko) {
    const R c = s[0];
    R upS = R(0), cenS = R(0);
    ko = R(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long q = st.q(d);
        const R m1 = s[-q], p1 = s[q], m2 = s[-2 * q], p2 = s[2 * q];
        const R m3 = s[-3 * q], p3 = s[3 * q], m4 = s[-4 * q], p4 = s[4 * q];
        const R dv = K.ih[d] * (-p2 + R(8) * p1 - R(8) * m1 + m2);
        d1[d] = dv;
        // backward stencil for beta >= 0 equals the forward one mirrored and negated
        const bool bw = be[d] >= R(0);
        const R g1 = bw ? m1 : p1, g2 = bw ? m2 : p2, g3 = bw ? m3 : p3, g4 = bw ? m4 : p4;
        const R up = (bw ? -K.ih[d] : K.ih[d]) * (R(-25) * c + R(48) * g1 - R(36) * g2 + R(16) * g3 - R(3) * g4);
        upS += be[d] * up;
        cenS += be[d] * dv;
        if (KO_EXACT)
            ko = ko_axis_ordered(ko, K.kw[d], m3, m2, m1, c, p1, p2, p3);
        else
            ko += K.ko[d] * ((m3 + p3) - R(6) * (m2 + p2) + R(15) * (m1 + p1) - R(20) * c);
        if (D2) dd[sym(d, d)] = K.ih2[d] * (-p2 + R(16) * p1 - R(30) * c + R(16) * m1 - m2);
    }
    if (D2) {
        const long x = st.q(0), y = st.q(1), z = st.q(2);
        dd[1] = K.ihh[0] * (s[x + y] - s[x - y] - s[-x + y] + s[-x - y]);
        dd[2] = K.ihh[1] * (s[x + z] - s[x - z] - s[-x + z] + s[-x - z]);
        dd[4] = K.ihh[2] * (s[y + z] - s[y - z] - s[-y + z] + s[-y - z]);
    }
    adv = w * upS + (R(1) - w) * cenS;
#ifdef __HIP_DEVICE_COMPILE__
    // scheduling hint for the field's code: its LDS reads go out before its arithmetic, so the
    // reads wait in two batches instead of four and the second batch's latency is covered by
    // the first batch's arithmetic (results unchanged: the same instructions, reordered)
    __builtin_amdgcn_sched_group_barrier(0x100, 40, 0); // LDS reads
    __builtin_amdgcn_sched_group_barrier(0x002, 400, 0); // then VALU
#endif
    pin(adv);
    pin(ko);
#pragma unroll
    for (int d = 0; d < 3; ++d) pin(d1[d]);
    if (D2) {
#pragma unroll
        for (int q = 0; q < 6; ++q) pin(dd[q]);
    }
}

// Full RHS of one cell. `src` holds the matter sources at this cell when HAS_SRC.
struct CellSources {
    double rho, S, Si[3], Sij[6];
};

// R is the arithmetic of the stencils and the RHS algebra (double for the parity form,
// float for the mixed form). The determinant and inverse of the conformal metric are
// always formed in double from the stager's double metric (Stager::metric64).
template <typename R, bool HAS_SRC, bool KO_EXACT, class Stager, class Sink>
__host__ __device__ GR_INLINE void ccz4_cell(Stager& S, Sink& O, const Stencil<R>& K, const CellParams& P,
                                             const CellSources& src) {
    // ---- centre values and the inverse conformal metric (ccz4.cpp:279-330)
    const R chi_raw = S.centre(CHI);
    const R chi = (finite_(chi_raw) && chi_raw >= R(1.0e-4)) ? chi_raw : R(1.0e-4);
    const R alpha = S.centre(ALPHA);
    const R Ktr = S.centre(KTR);
    const R Theta = S.centre(THETA);
    R gt[6], At[6], Gh[3], be[3];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        gt[a] = S.centre(GXX + a);
        At[a] = S.centre(AXX + a);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        Gh[a] = S.centre(GHX + a);
        be[a] = S.centre(BX + a);
    }
    R gu[6];
    {
        double g[6];
        S.metric64(g);
        const double det = g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) +
            g[2] * (g[1] * g[4] - g[3] * g[2]);
        const double idet = 1.0 / (det + 1.0e-30);
        gu[0] = R((g[3] * g[5] - g[4] * g[4]) * idet);
        gu[1] = R((g[2] * g[4] - g[1] * g[5]) * idet);
        gu[2] = R((g[1] * g[4] - g[2] * g[3]) * idet);
        gu[3] = R((g[0] * g[5] - g[2] * g[2]) * idet);
        gu[4] = R((g[1] * g[2] - g[0] * g[4]) * idet);
        gu[5] = R((g[0] * g[3] - g[1] * g[1]) * idet);
    }
    const R w = R(0.5) * (R(1) + tanh_((chi - R(P.chi_c)) / R(P.chi_w)));

    R d1[3], dd[6], adv, ko;

    // Gamma-hat RHS (ccz4.cpp:747-815) is gathered in GA as its ingredients appear:
    // T5 + T6 from the shift's second derivatives, T1 + T4 + T8 once the Christoffels and
    // A^i_j exist, T2 and T3 with K, T7 as each Gamma-hat field is stored.
    R GA[3];

    // ---- shift: keeps d beta; second derivatives give the Laplacian T5 and the grad-div T6
    //      (ccz4.cpp:733-794); RHS (ccz4.cpp:826-831)
    R dbe[3][3], ddiv[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        eval_field<true, KO_EXACT>(S.field(BX + c), K, S.strides(), be, w, d1, dd, adv, ko);
        R t5 = R(0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            dbe[c][a] = d1[a];
#pragma unroll
            for (int b = 0; b < 3; ++b) t5 += gu[sym(a, b)] * dd[sym(a, b)];
            ddiv[a] += dd[sym(c, a)];
        }
        GA[c] = t5;
        pin(GA[c]);
#pragma unroll
        for (int a = 0; a < 3; ++a) pin(ddiv[a]);
        O.put(BX + c, R(0.75) * Gh[c] - R(P.eta) * be[c] + adv, ko, be[c]);
    }
    const R divb = dbe[0][0] + dbe[1][1] + dbe[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        GA[i] += (R(1) / R(3)) * (gu[sym(i, 0)] * ddiv[0] + gu[sym(i, 1)] * ddiv[1] + gu[sym(i, 2)] * ddiv[2]);

    // ---- conformal metric: keeps d gt; second derivatives give Ricci term 1
    //      -1/2 g^lm d_l d_m gt_ij (ccz4.cpp:470-476); RHS (ccz4.cpp:586-607)
    R dgt[6][3], Rt[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const int a = sym_a(p), b = sym_b(p);
        eval_field<true, KO_EXACT>(S.field(GXX + p), K, S.strides(), be, w, dgt[p], dd, adv, ko);
        R r1 = R(0);
#pragma unroll
        for (int q = 0; q < 6; ++q) r1 += ((q == 0 || q == 3 || q == 5) ? R(-0.5) : R(-1)) * gu[q] * dd[q];
        Rt[p] = r1;
        pin(Rt[p]);
        R L = adv;
#pragma unroll
        for (int q = 0; q < 3; ++q) L += gt[sym(a, q)] * dbe[q][b] + gt[sym(b, q)] * dbe[q][a];
        O.put(GXX + p, R(-2) * alpha * At[p] + L - (R(2) / R(3)) * gt[p] * divb, ko, gt[p]);
    }

    // ---- the remaining uses of the shift's first derivatives, formed now so that they die:
    //      the Lie terms of A~ (ccz4.cpp:708-715) and of Gamma-hat's T7 (ccz4.cpp:795-808)
    R LieA[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const int a = sym_a(p), b = sym_b(p);
        R L = R(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) L += At[sym(a, q)] * dbe[q][b] + At[sym(b, q)] * dbe[q][a];
        LieA[p] = L - (R(2) / R(3)) * At[p] * divb;
        pin(LieA[p]);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        GA[q] += -(Gh[0] * dbe[q][0] + Gh[1] * dbe[q][1] + Gh[2] * dbe[q][2]) + (R(2) / R(3)) * Gh[q] * divb;
        pin(GA[q]);
    }

    // ---- Christoffel symbols chr[i][jk] (ccz4.cpp:413-444)
    R chr[3][6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const int j = sym_a(p), k = sym_b(p);
            R s = R(0);
#pragma unroll
            for (int l = 0; l < 3; ++l) s += gu[sym(i, l)] * (dgt[sym(l, k)][j] + dgt[sym(l, j)][k] - dgt[p][l]);
            chr[i][p] = R(0.5) * s;
            pin(chr[i][p]);
        }

    // ---- chi (ccz4.cpp:580); its second derivatives give the conformal-factor Ricci R^chi
    //      (ccz4.cpp:518-541) at once, with the Christoffels already formed
    R dchi[3];
    {
        R ddchi[6];
        eval_field<true, KO_EXACT>(S.field(CHI), K, S.strides(), be, w, dchi, ddchi, adv, ko);
        O.put(CHI, (R(2) / R(3)) * chi * (alpha * Ktr - divb) + adv, ko, chi_raw);
        R lapchi = R(0), gradchi2 = R(0);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int l = sym_a(q), m = sym_b(q);
            const R f = (l == m) ? R(1) : R(2);
            lapchi += f * gu[q] * (ddchi[q] - chr[0][q] * dchi[0] - chr[1][q] * dchi[1] - chr[2][q] * dchi[2]);
            gradchi2 += f * gu[q] * dchi[l] * dchi[m];
        }
        const R c1 = R(0.5) / (chi + R(1e-30));
        const R c2 = R(0.25) / ((chi * chi) + R(1e-30));
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const int a = sym_a(p), b = sym_b(p);
            const R Dab = ddchi[p] - chr[0][p] * dchi[0] - chr[1][p] * dchi[1] - chr[2][p] * dchi[2];
            Rt[p] += c1 * (Dab + gt[p] * lapchi) - c2 * (dchi[a] * dchi[b] - R(3) * gt[p] * gradchi2);
            pin(Rt[p]);
        }
    }

    // ---- lapse (ccz4.cpp:820); its second derivatives give the covariant Hessian
    //      D_i D_j alpha (ccz4.cpp:678-680) at once
    R dal[3], DDa[6];
    {
        R ddal[6];
        eval_field<true, KO_EXACT>(S.field(ALPHA), K, S.strides(), be, w, dal, ddal, adv, ko);
        O.put(ALPHA, R(-2) * alpha * Ktr + adv, ko, alpha);
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            DDa[p] = ddal[p] - chr[0][p] * dal[0] - chr[1][p] * dal[1] - chr[2][p] * dal[2];
            pin(DDa[p]);
        }
    }

    // ---- A~A~ = A^i_j A^j_i with A^i_j = g^ik A_kj (ccz4.cpp:609-619); T1 + T4 + T8 of
    //      Gamma-hat. A^i_j is formed here and again for the A~ RHS rather than kept live.
    R AA = R(0);
    {
        R Au[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Au[i][j] = gu[sym(i, 0)] * At[sym(0, j)] + gu[sym(i, 1)] * At[sym(1, j)] + gu[sym(i, 2)] * At[sym(2, j)];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) AA += Au[i][j] * Au[j][i];
        pin(AA);
        const R c4 = R(3) * alpha / (chi + R(1e-30));
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            R T1 = R(0);
#pragma unroll
            for (int q = 0; q < 6; ++q) T1 += ((q == 0 || q == 3 || q == 5) ? R(1) : R(2)) * chr[i][q] * At[q];
            GA[i] += R(2) * alpha * T1 + c4 * (Au[i][0] * dchi[0] + Au[i][1] * dchi[1] + Au[i][2] * dchi[2]) -
                R(2) * R(P.kappa1) * alpha * Gh[i];
            pin(GA[i]);
        }
    }

    // ---- lapse Laplacian D^2 alpha (ccz4.cpp:621-648), formed while the first derivatives of
    //      chi and alpha are live
    R plap;
    {
        R lap = R(0), cga = R(0);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int l = sym_a(q), m = sym_b(q);
            const R f = (l == m) ? R(1) : R(2);
            lap += f * gu[q] * DDa[q];
            cga += gu[q] * (l == m ? dchi[l] * dal[l] : dchi[l] * dal[m] + dchi[m] * dal[l]);
        }
        plap = chi * lap - R(0.5) * cga;
        pin(plap);
    }

    // ---- remaining Ricci terms (ccz4.cpp:486-515), formed as soon as the Christoffels exist
    //      (so that they die before the next fields are staged), factored per k with
    //      G_k = gt_kn chr^n:
    //   sum_k (Gh^k - g^lm chr^k_lm) G_k,ij + sum_l chr^k_li H_klj + chr^k_lj H_kli,
    //   H_klj = g^lm G_k,jm
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R Gk[6];
#pragma unroll
        for (int p = 0; p < 6; ++p) Gk[p] = gt[sym(k, 0)] * chr[0][p] + gt[sym(k, 1)] * chr[1][p] + gt[sym(k, 2)] * chr[2][p];
        R Gd = R(0);
#pragma unroll
        for (int q = 0; q < 6; ++q) Gd += ((q == 0 || q == 3 || q == 5) ? R(1) : R(2)) * gu[q] * chr[k][q];
        const R cg = Gh[k] - Gd;
#pragma unroll
        for (int p = 0; p < 6; ++p) Rt[p] += cg * Gk[p];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            R H[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) H[j] = gu[sym(l, 0)] * Gk[sym(j, 0)] + gu[sym(l, 1)] * Gk[sym(j, 1)] + gu[sym(l, 2)] * Gk[sym(j, 2)];
#pragma unroll
            for (int p = 0; p < 6; ++p) {
                const int a = sym_a(p), b = sym_b(p);
                Rt[p] += chr[k][sym(l, a)] * H[b] + chr[k][sym(l, b)] * H[a];
            }
        }
    }
#pragma unroll
    for (int p = 0; p < 6; ++p) pin(Rt[p]);

    // ---- K (ccz4.cpp:621-652); its gradient completes T2 of Gamma-hat
    {
        R dK[3];
        eval_field<false, KO_EXACT>(S.field(KTR), K, S.strides(), be, w, dK, dd, adv, ko);
        const R matter = HAS_SRC ? R(4) * R(PI) * alpha * (R(src.rho) + R(src.S)) : R(0);
        O.put(KTR,
              -plap + alpha * (AA + Ktr * Ktr / R(3)) + matter + R(P.kappa1) * (R(1) - R(P.kappa2)) * alpha * Theta + adv,
              ko, Ktr);
        const R c2 = -(R(4) / R(3)) * alpha;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            GA[i] += c2 * (gu[sym(i, 0)] * dK[0] + gu[sym(i, 1)] * dK[1] + gu[sym(i, 2)] * dK[2]);
            if (HAS_SRC)
                GA[i] += -R(16) * R(PI) * alpha *
                    (gu[sym(i, 0)] * R(src.Si[0]) + gu[sym(i, 1)] * R(src.Si[1]) + gu[sym(i, 2)] * R(src.Si[2]));
            pin(GA[i]);
        }
    }

    // ---- Gamma-hat fields: Ricci term 2, 1/2 (gt_ki d_j Gh^k + gt_kj d_i Gh^k) (ccz4.cpp:480-484);
    //      RHS with T7 = -Gh^j d_j beta^i + advection + 2/3 Gh^i div beta, stored at once
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        R advq, koq;
        eval_field<false, KO_EXACT>(S.field(GHX + q), K, S.strides(), be, w, d1, dd, advq, koq);
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const int a = sym_a(p), b = sym_b(p);
            Rt[p] += R(0.5) * (gt[sym(q, a)] * d1[b] + gt[sym(q, b)] * d1[a]);
            pin(Rt[p]);
        }
        O.put(GHX + q, GA[q] + advq, koq, Gh[q]);
    }

    // ---- the Ricci tensor is complete: its trace for Theta, and -D_i D_j alpha + alpha R_ij
    //      for A~ (ccz4.cpp:654-660,674-680), so that Rt and the lapse Hessian die here
    R Rs = R(0), XA[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        Rs += ((q == 0 || q == 3 || q == 5) ? R(1) : R(2)) * gu[q] * Rt[q];
        XA[q] = -DDa[q] + alpha * Rt[q];
        pin(XA[q]);
    }
    Rs *= chi;
    pin(Rs);

    // ---- Theta (ccz4.cpp:654-665)
    eval_field<false, KO_EXACT>(S.field(THETA), K, S.strides(), be, w, d1, dd, adv, ko);
    {
        const R rho = HAS_SRC ? R(src.rho) : R(0);
        O.put(THETA,
              R(0.5) * alpha * (Rs - AA + R(2) * Ktr * Ktr / R(3) - R(16) * R(PI) * rho) -
                  alpha * R(P.kappa1) * (R(2) + R(P.kappa2)) * Theta + adv,
              ko, Theta);
    }

    // ---- A~: raw value, then trace removed with the full raw trace (ccz4.cpp:671-726)
    R raw[6], koA[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const int a = sym_a(p), b = sym_b(p);
        eval_field<false, KO_EXACT>(S.field(AXX + p), K, S.strides(), be, w, d1, dd, adv, koA[p]);
        const R t1 = HAS_SRC ? chi * (XA[p] - alpha * R(8) * R(PI) * R(src.Sij[p])) : chi * XA[p];
        R Aub[3]; // A^q_b = g^qr A_rb
#pragma unroll
        for (int q = 0; q < 3; ++q)
            Aub[q] = gu[sym(q, 0)] * At[sym(0, b)] + gu[sym(q, 1)] * At[sym(1, b)] + gu[sym(q, 2)] * At[sym(2, b)];
        const R Asq = At[sym(a, 0)] * Aub[0] + At[sym(a, 1)] * Aub[1] + At[sym(a, 2)] * Aub[2];
        const R t2 = alpha * (Ktr * At[p] - R(2) * Asq);
        raw[p] = t1 + t2 + (adv + LieA[p]);
        pin(raw[p]);
    }
    {
        R tr = R(0);
#pragma unroll
        for (int q = 0; q < 6; ++q) tr += ((q == 0 || q == 3 || q == 5) ? R(1) : R(2)) * gu[q] * raw[q];
#pragma unroll
        for (int p = 0; p < 6; ++p) O.put(AXX + p, raw[p] - (R(1) / R(3)) * gt[p] * tr, koA[p], At[p]);
    }
    O.finish();
}

} // namespace gr
