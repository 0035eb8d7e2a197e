// Device helpers shared by the generic LDS FFT kernels (gp5_kernels.hip) and the
// register-resident fast kernels (gp5_fast.hip): complex arithmetic, small DFTs,
// the reference k-grid rule, the deterministic block reduction and the fused
// functional / stepper epilogue.
#pragma once

#include "gp5_internal.h"

#include <math.h>

namespace gp5 {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// a * (i * s) for s = +-1
__device__ __forceinline__ double2 cmul_is(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }

__device__ __forceinline__ double2 twiddle(const double2 *tw, int t, int sign)
{
    double2 w = tw[t];
    if (sign > 0) w.y = -w.y;
    return w;
}

// ---------------------------------------------------------------------------
// small DFTs in registers; sign = -1 forward (e^{-2 pi i/R}), +1 inverse
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dft2(double2 *v)
{
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

__device__ __forceinline__ void dft3(double2 *v, double s)
{
    const double c = -0.5, sn = 0.86602540378443864676372317075294;  // sin(2 pi/3)
    double2 t1 = cadd(v[1], v[2]);
    double2 t2 = make_double2(v[0].x + c * t1.x, v[0].y + c * t1.y);
    double2 d = csub(v[1], v[2]);
    double2 t3 = cmul_is(d, s * sn);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

__device__ __forceinline__ void dft4(double2 *v, double s)
{
    double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    double2 b0 = cadd(v[1], v[3]), b1 = cmul_is(csub(v[1], v[3]), s);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
}

// 12 points as 3 x 4 by the prime-factor (Good-Thomas) mapping, no twiddles:
// input n = (4 n1 + 3 n2) mod 12, 3-point DFTs over n1, 4-point DFTs over n2,
// output k = (4 k1 + 9 k2) mod 12 (k = k1 mod 3, k = k2 mod 4).
__device__ __forceinline__ void dft12(double2 *v, double s)
{
    double2 a[3][4];
#pragma unroll
    for (int n1 = 0; n1 < 3; ++n1)
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) a[n1][n2] = v[(4 * n1 + 3 * n2) % 12];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        double2 t[3] = {a[0][n2], a[1][n2], a[2][n2]};
        dft3(t, s);
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) a[k1][n2] = t[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        dft4(a[k1], s);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[(4 * k1 + 9 * k2) % 12] = a[k1][k2];
    }
}

__device__ __forceinline__ void dft5(double2 *v, double s)
{
    const double c1 = 0.30901699437494742410229341718282;   // cos(2 pi/5)
    const double c2 = -0.80901699437494742410229341718282;  // cos(4 pi/5)
    const double s1 = 0.95105651629515357211643933337938;   // sin(2 pi/5)
    const double s2 = 0.58778525229247312916870595463907;   // sin(4 pi/5)
    double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    double2 a0 = v[0];
    double2 b1 = make_double2(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
    double2 b2 = make_double2(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
    double2 d1 = make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    double2 d2 = make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    double2 e1 = cmul_is(d1, s), e2 = cmul_is(d2, s);
    v[0] = cadd(a0, cadd(t1, t2));
    v[1] = cadd(b1, e1);
    v[4] = csub(b1, e1);
    v[2] = cadd(b2, e2);
    v[3] = csub(b2, e2);
}

__device__ __forceinline__ void dft8(double2 *v, double s)
{
    const double h = 0.70710678118654752440084436210485;
    double2 e[4] = {v[0], v[2], v[4], v[6]};
    double2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e, s);
    dft4(o, s);
    // w8^k = cos(pi k/4) + i s sin(pi k/4)
    double2 o1 = cmul(o[1], make_double2(h, s * h));
    double2 o2 = cmul_is(o[2], s);
    double2 o3 = cmul(o[3], make_double2(-h, s * h));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

// 16-point DFT as 4 x 4: r = 4a + b, k = c + 4d;
// X[c + 4d] = sum_b w4^{bd} w16^{bc} sum_a w4^{ac} v[4a + b]
__device__ __forceinline__ void dft16(double2 *v, double s)
{
    const double c1 = 0.92387953251128675612818318939679;  // cos(pi/8)
    const double s1 = 0.38268343236508977172845998403040;  // sin(pi/8)
    const double h = 0.70710678118654752440084436210485;
    double2 y[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double2 t[4] = {v[b], v[b + 4], v[b + 8], v[b + 12]};
        dft4(t, s);
#pragma unroll
        for (int c = 0; c < 4; ++c) y[b][c] = t[c];
    }
    // w16^{bc} = cos(pi bc/8) + i s sin(pi bc/8), bc in {1,2,3,4,6,9}
    y[1][1] = cmul(y[1][1], make_double2(c1, s * s1));
    y[1][2] = cmul(y[1][2], make_double2(h, s * h));
    y[1][3] = cmul(y[1][3], make_double2(s1, s * c1));
    y[2][1] = cmul(y[2][1], make_double2(h, s * h));
    y[2][2] = cmul_is(y[2][2], s);
    y[2][3] = cmul(y[2][3], make_double2(-h, s * h));
    y[3][1] = cmul(y[3][1], make_double2(s1, s * c1));
    y[3][2] = cmul(y[3][2], make_double2(-h, s * h));
    y[3][3] = cmul(y[3][3], make_double2(-c1, -s * s1));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double2 t[4] = {y[0][c], y[1][c], y[2][c], y[3][c]};
        dft4(t, s);
#pragma unroll
        for (int d = 0; d < 4; ++d) v[c + 4 * d] = t[d];
    }
}

// ---------------------------------------------------------------------------
// k-space wavenumber (src/geometry.cpp:127-148)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double momentum(int i, int n, double dk)
{
    if ((n & 1) == 0) return i < n / 2 ? dk * i : (double)(i - n) * dk;
    return i <= (n - 1) / 2 ? dk * i : (double)(i - n) * dk;
}

// ---------------------------------------------------------------------------
// block reduction of kRedQ partial sums
// ---------------------------------------------------------------------------
// Blocks are multiples of 64 threads, at most 512; `sred` is LDS scratch of at
// least kRedScratch doubles.  Each block writes its kRedQ partial sums (wave
// shuffles, then the waves in order); k_finalize, one block launched right
// after on the same stream, sums the partials in a fixed order.  No atomics:
// a single global counter hit by every block serialises tens of thousands of
// blocks at the L2.
constexpr int kRedScratch = kRedQ * 8 + 1;

__device__ __forceinline__ void block_sum(double (&v)[kRedQ], double *sred, double (&out)[kRedQ])
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int NW = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) sred[w * kRedQ + q] = x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) {
        double x = 0.0;
        for (int i = 0; i < NW; ++i) x += sred[i * kRedQ + q];
        out[q] = x;
    }
}

// bid: the block's index in the launch's tile order (blockIdx.x unless remapped)
__device__ __forceinline__ void reduce_block(double (&v)[kRedQ], double *partials, double *sred, long long bid)
{
    double s[kRedQ];
    block_sum(v, sred, s);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < kRedQ; ++q) partials[(size_t)bid * kRedQ + q] = s[q];
    }
}

__device__ __forceinline__ void reduce_block(double (&v)[kRedQ], double *partials, double *sred)
{
    reduce_block(v, partials, sred, (long long)blockIdx.x);
}

// XCD-aware tile order.  The dispatcher deals workgroups to the 8 XCDs of the chip
// round-robin (block b runs on XCD b mod 8); with remap on, the b-th block takes
// tile (b mod 8) * (G / 8) + b / 8, so that each XCD walks one contiguous eighth
// of the tiles (its L2 and TLB see neighbouring lines and columns) instead of
// every eighth tile of the whole field.  The last G mod 8 blocks keep their index,
// so the map is a bijection on [0, G).
__device__ __forceinline__ long long xcd_tile(bool remap)
{
    const long long b = blockIdx.x;
    if (!remap) return b;
    const long long per = (long long)gridDim.x / 8;
    if (b >= per * 8) return b;
    return (b % 8) * per + b / 8;
}

// ---------------------------------------------------------------------------
// epilogue: one element of the last inverse pass
// ---------------------------------------------------------------------------
// Streaming (nontemporal) global accesses.  The epilogue pass touches 5-7
// field streams once each; the nt hint on its loads (bit 0 of GP5_NT_EPI) and
// stores (bit 1) keeps them from displacing each other in L2 and the
// Infinity Cache.  Measured at 256^3: fused x pass 248.7 -> 221.7 us, and the
// following y pass 86.8 -> 83.4 us.  The in-place column passes are faster
// without the hint (GP5_NT_COLS, off: y pass 86.8 -> 98.0 us with it).
#ifndef GP5_NT_EPI
#define GP5_NT_EPI 3
#endif
#ifndef GP5_NT_COLS
#define GP5_NT_COLS 0
#endif
template <bool NT>
__device__ __forceinline__ double2 ldg(const double2 *p)
{
    if constexpr (NT) {
        double2 v;
        v.x = __builtin_nontemporal_load(&p->x);
        v.y = __builtin_nontemporal_load(&p->y);
        return v;
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ void stg(double2 *p, double2 v)
{
    if constexpr (NT) {
        __builtin_nontemporal_store(v.x, &p->x);
        __builtin_nontemporal_store(v.y, &p->y);
    } else {
        *p = v;
    }
}

// v = src[g] * sc[c] - cf * src2[g]: the pre-operation of a first (x) pass
// (terms absent when the pointer is null)
template <bool NT = false>
__device__ __forceinline__ double2 load_pre(const double2 *src, long long g, const double *sc, int c,
                                            const double2 *src2, double2 cf)
{
    double2 v = ldg<NT>(src + g);
    if (sc) {
        const double s = sc[c];
        v.x *= s;
        v.y *= s;
    }
    if (src2) v = csub(v, cmul(cf, ldg<NT>(src2 + g)));
    return v;
}

// Per-thread reduction sums of the epilogue for the thread's component; they
// are placed into the block's component slot once per thread (acc_scatter),
// not per element: a runtime index into the register array would cost a
// compare-and-select chain over every slot at each update.
struct EpiAcc {
    double o2 = 0.0, en = 0.0, rho = 0.0;
};

// slot: the thread's component minus the block's first component (0 or 1)
__device__ __forceinline__ void acc_scatter(const EpiAcc &a, int slot, double (&red)[kRedQ])
{
    const bool s0 = slot == 0;
    red[0] += s0 ? a.o2 : 0.0;
    red[1] += s0 ? a.en : 0.0;
    red[2] += s0 ? a.rho : 0.0;
    red[3] += s0 ? 0.0 : a.o2;
    red[4] += s0 ? 0.0 : a.en;
    red[5] += s0 ? 0.0 : a.rho;
}

// Epilogue of the last inverse x pass over CH elements of one line (one
// component c) at gbase + pos[k].  Every global load of the batch is issued
// before the arithmetic, with the mode / functional tests uniform across the
// wave, so each operand stream has CH independent requests in flight per
// thread.  kin = -1/2 m_c^-1 lap(phi) arrives in v[k]; on return v[k] holds the
// value the pass hands on: the kinetic term (STORE), H (H, ENERGY), or o.
//
// Functionals (src/functional.cpp): H phi = kin + (V + nl) phi with
//   LHYDroplet (:105-122)         nl = a rho + b rho^{3/2}
//   gpFunctional (:13-43,54-89)   nl = g11 rho (1 comp), gcc rho_c + g12 rho_o (2 comps)
//   UnlockedLHYDroplet (:125-150) nl_0 = rho1 + eta rho2 + alpha (rho1 + beta rho2)^{3/2},
//                                 nl_1 = eta beta rho1 + beta rho2 + alpha beta^2 (...)^{3/2}
// Reductions (acc): |o|^2, energy density of phi, |phi|^2 of the thread's component.
// LOCAL = true compiles the form without cross-component reads and without the
// phase constraint (one component, or two uncoupled ones, no phaseConstraint):
// registers are allocated for the longest path of a kernel, so the general
// form costs occupancy even when those paths are never taken.
template <int CH, bool LOCAL = false>
__device__ __forceinline__ void epilogue_batch(const EpiArgs &E, long long gbase, const int (&pos)[CH], int c,
                                               double2 (&v)[CH], EpiAcc &acc)
{
    constexpr bool NTL = (GP5_NT_EPI & 1) != 0, NT = (GP5_NT_EPI & 2) != 0;
    const int mode = E.mode;
    if (mode == EPI_STORE) {
#pragma unroll
        for (int k = 0; k < CH; ++k) stg<NT>(E.out + gbase + pos[k], v[k]);
        return;
    }
    const bool rk = mode >= EPI_RK1 && mode <= EPI_RK4;
    const bool rk_k = mode == EPI_RK2 || mode == EPI_RK3 || mode == EPI_RK4;
    const bool coupled = !LOCAL && (E.fkind == 2 || (E.fkind == 0 && E.has_couplings && E.nc == 2 && !E.uncoupled));
    const bool constrain = !LOCAL && E.phase_exp && (mode == EPI_EULER || mode == EPI_RK4);
    const double2 *src = rk ? E.psi : E.phi;
    const double *ssc = rk ? E.sc_psi : E.sc_phi;
    double2 x[CH], kk[CH], ss[CH], xo[CH], ko[CH], ph[CH];
    double V[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) x[k] = ldg<NTL>(src + gbase + pos[k]);
    if (rk_k) {
#pragma unroll
        for (int k = 0; k < CH; ++k) kk[k] = ldg<NTL>(E.kbuf + gbase + pos[k]);
    }
    if (mode == EPI_RK4) {
#pragma unroll
        for (int k = 0; k < CH; ++k) ss[k] = ldg<NTL>(E.sbuf + gbase + pos[k]);
    }
    if (E.V) {
#pragma unroll
        for (int k = 0; k < CH; ++k) V[k] = E.V[gbase + pos[k]];
    } else {
#pragma unroll
        for (int k = 0; k < CH; ++k) V[k] = 0.0;
    }
    if (coupled) {
        const long long off = c == 0 ? E.M : -E.M;
#pragma unroll
        for (int k = 0; k < CH; ++k) xo[k] = src[gbase + off + pos[k]];
        if (rk_k) {
#pragma unroll
            for (int k = 0; k < CH; ++k) ko[k] = E.kbuf[gbase + off + pos[k]];
        }
    }
    if (constrain) {
#pragma unroll
        for (int k = 0; k < CH; ++k) ph[k] = E.phase_exp[gbase + pos[k]];
    }
    const double sx = ssc ? ssc[c] : 1.0;
    const double sxo = (coupled && ssc) ? ssc[1 - c] : 1.0;
    const double sin_ = (rk_k && E.sc_in) ? E.sc_in[c] : 1.0;
    const double sino = (rk_k && coupled && E.sc_in) ? E.sc_in[1 - c] : 1.0;
    const double2 cin = make_double2(E.cr_in, E.ci_in);
    const double2 cf = make_double2(E.cr, E.ci);
    const double la = E.lhy_ab ? E.lhy_ab[2 * c] : E.lhy_a;
    const double lb = E.lhy_ab ? E.lhy_ab[2 * c + 1] : E.lhy_b;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const long long g = gbase + pos[k];
        const double2 xs = cscale(x[k], sx);  // psi (RK) or phi
        // stage input phi: psi for RK1, norm(u_s) = sc_in (psi - c_in K) for RK2-4
        const double2 p = rk_k ? cscale(csub(xs, cmul(cin, kk[k])), sin_) : xs;
        const double2 kin = v[k];
        const double rho = p.x * p.x + p.y * p.y;
        double nl = 0.0, enl = 0.0;
        if (E.fkind == 1) {
            const double sr = sqrt(rho);
            nl = la * rho + lb * rho * sr;
            enl = 0.5 * la * rho * rho + 0.4 * lb * rho * rho * sr;
        } else if (E.fkind == 0 && !E.has_couplings) {
        } else if (E.fkind == 0 && (E.nc == 1 || E.uncoupled)) {
            // one component, or independent instances sharing g11
            nl = E.g11 * rho;
            enl = 0.5 * E.g11 * rho * rho;
        } else {
            const double2 xso = cscale(xo[k], sxo);
            const double2 q = rk_k ? cscale(csub(xso, cmul(cin, ko[k])), sino) : xso;
            const double ro = q.x * q.x + q.y * q.y;
            if (E.fkind == 0) {
                const double gcc = c == 0 ? E.g11 : E.g22;
                nl = gcc * rho + E.g12 * ro;
                enl = 0.5 * gcc * rho * rho + 0.5 * E.g12 * rho * ro;
            } else {
                const double r1 = c == 0 ? rho : ro, r2 = c == 0 ? ro : rho;
                const double base = r1 + E.beta * r2;
                const double s = base * sqrt(base);
                nl = c == 0 ? r1 + E.eta * r2 + E.alpha * s
                            : E.eta * E.beta * r1 + E.beta * r2 + E.alpha * E.beta * E.beta * s;
            }
        }
        const double a = nl + V[k];
        const double2 h = make_double2(kin.x + a * p.x, kin.y + a * p.y);
        acc.en += p.x * kin.x + p.y * kin.y + V[k] * rho + enl;
        acc.rho += rho;
        if (mode == EPI_H) {
            stg<NT>(E.out + g, h);
            v[k] = h;
            continue;
        }
        if (mode == EPI_ENERGY) {
            v[k] = h;
            continue;
        }
        double2 o;
        if (mode == EPI_RK4) {
            // psi - ((1/6) tau) (((k1 + 2 k2) + 2 k3) + k4), S = k1 + 2 k2, K = k3
            const double2 sum = cadd(make_double2(ss[k].x + 2.0 * kk[k].x, ss[k].y + 2.0 * kk[k].y), h);
            o = csub(xs, cmul(cf, sum));
        } else {
            o = csub(xs, cmul(cf, h));
            if (mode == EPI_RK2) stg<NT>(E.sbuf + g, make_double2(kk[k].x + 2.0 * h.x, kk[k].y + 2.0 * h.y));
            if (rk) stg<NT>(E.kout + g, h);
        }
        if (constrain) {
            const double amp = sqrt(o.x * o.x + o.y * o.y);
            o = make_double2(amp * ph[k].x, amp * ph[k].y);
        }
        if (mode == EPI_EULER || mode == EPI_RK4) stg<NT>(E.out + g, o);
        acc.o2 += o.x * o.x + o.y * o.y;
        v[k] = o;
    }
}

// Single-element form used by the generic LDS kernels (slot: c minus the
// block's first component).
__device__ __forceinline__ double2 epilogue(const EpiArgs &E, long long g, int c, int slot, double2 kin,
                                           double (&red)[kRedQ])
{
    const int pos[1] = {0};
    double2 v[1] = {kin};
    EpiAcc acc;
    epilogue_batch<1>(E, g, pos, c, v, acc);
    acc_scatter(acc, slot, red);
    return v[0];
}

}  // namespace gp5
