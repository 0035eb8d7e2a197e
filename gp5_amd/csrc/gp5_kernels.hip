// HIP kernels of the gfx950 GP core: LDS-tiled Stockham FFT passes along one
// axis of a column-major field, the fused forward/multiply/inverse pass on the
// last axis, and the last inverse pass carrying the nonlinear functional, the
// Euler/RK4 stage update and the wave-shuffle norm/energy reductions.
//
// Reference operations restated here (GP5):
//   fftwFourierTransform::apply     src/fourierTransform.cpp:184-214
//   operators::laplacian::apply     src/operators.cpp:28-36
//   functional::apply (3 kinds)     src/functional.cpp:13-150
//   euleroStepper / RK4Stepper      src/stepper.cpp:30-81
//   gp::norm / gp::normalize        src/geometry.cpp:190-237
//   phaseConstraint::apply          src/constraint.cpp:6-9
#include "gp5_internal.h"

#include <math.h>

#include <algorithm>

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

// LDS index of element `pos` of tile line `b`.  rows: lines contiguous (x axis);
// cols: the B lines interleaved (y/z axes, b fastest as in global memory).
__device__ __forceinline__ int lidx(int b, int pos, int n, int B, bool rows)
{
    return rows ? b * n + pos : pos * B + b;
}

// One Stockham radix-R step over all lines of the tile, a -> b (LDS).
// Butterfly (j, q), j < L, q < m = n/(L R): inputs at q + m (r + R j),
// twiddled by w_{LR}^{j r}, outputs at q + m (j + L k).
template <int R>
__device__ __forceinline__ void stockham_step(const double2 *a, double2 *b, int n, int L, int B, bool rows,
                                              const double2 *tw, int sign, int tid)
{
    const int m = n / (L * R);
    const int nb = n / R;
    const int total = nb * B;
    const double s = (double)sign;
    for (int u = tid; u < total; u += kThreads) {
        int line, bf;
        if (rows) { bf = u % nb; line = u / nb; }
        else      { line = u % B; bf = u / B; }
        const int q = bf % m, j = bf / m;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = a[lidx(line, q + m * (r + R * j), n, B, rows)];
        if (L > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle(tw, j * r * m, sign));
        }
        if constexpr (R == 2) dft2(v);
        else if constexpr (R == 3) dft3(v, s);
        else if constexpr (R == 4) dft4(v, s);
        else if constexpr (R == 5) dft5(v, s);
        else if constexpr (R == 8) dft8(v, s);
#pragma unroll
        for (int k = 0; k < R; ++k) b[lidx(line, q + m * (j + L * k), n, B, rows)] = v[k];
    }
}

// Radix R of any size (used for primes > 5): direct O(R^2) DFT reading LDS.
__device__ void stockham_step_generic(const double2 *a, double2 *b, int n, int L, int R, int B, bool rows,
                                      const double2 *tw, int sign, int tid)
{
    const int m = n / (L * R);
    const int nb = n / R;
    const int total = nb * B;
    const int stride = n / R;
    for (int u = tid; u < total; u += kThreads) {
        int line, bf;
        if (rows) { bf = u % nb; line = u / nb; }
        else      { line = u % B; bf = u / B; }
        const int q = bf % m, j = bf / m;
        for (int k = 0; k < R; ++k) {
            double2 acc = make_double2(0.0, 0.0);
            for (int r = 0; r < R; ++r) {
                double2 x = a[lidx(line, q + m * (r + R * j), n, B, rows)];
                if (r > 0 && L > 1) x = cmul(x, twiddle(tw, j * r * m, sign));
                x = cmul(x, twiddle(tw, ((r * k) % R) * stride, sign));
                acc = cadd(acc, x);
            }
            b[lidx(line, q + m * (j + L * k), n, B, rows)] = acc;
        }
    }
}

// Full Stockham transform of the tile; returns the buffer holding the result.
__device__ double2 *lds_fft(double2 *a, double2 *b, const AxisPlan &P, int B, bool rows, int sign, int tid)
{
    int L = 1;
    for (int s = 0; s < P.nrad; ++s) {
        const int R = P.rad[s];
        switch (R) {
            case 2: stockham_step<2>(a, b, P.n, L, B, rows, P.tw, sign, tid); break;
            case 3: stockham_step<3>(a, b, P.n, L, B, rows, P.tw, sign, tid); break;
            case 4: stockham_step<4>(a, b, P.n, L, B, rows, P.tw, sign, tid); break;
            case 5: stockham_step<5>(a, b, P.n, L, B, rows, P.tw, sign, tid); break;
            case 8: stockham_step<8>(a, b, P.n, L, B, rows, P.tw, sign, tid); break;
            default: stockham_step_generic(a, b, P.n, L, R, B, rows, P.tw, sign, tid); break;
        }
        __syncthreads();
        double2 *t = a; a = b; b = t;
        L *= R;
    }
    return a;
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
// block reduction of kRedQ partial sums; the last block to finish sums the
// per-block partials in a fixed order (deterministic) and publishes results.
// ---------------------------------------------------------------------------
// `sred` is LDS scratch of at least kRedQ*kThreads/64 + 1 doubles.
constexpr int kRedScratch = kRedQ * (kThreads / 64) + 1;

__device__ void reduce_publish(double (&v)[kRedQ], const EpiArgs &E, double *sred)
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    constexpr int NW = kThreads / 64;
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) sred[w * kRedQ + q] = x;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < kRedQ; ++q) {
            double x = 0.0;
            for (int i = 0; i < NW; ++i) x += sred[i * kRedQ + q];
            E.partials[(size_t)blockIdx.x * kRedQ + q] = x;
        }
        __threadfence();
        unsigned int t = atomicAdd(E.counter, 1u);
        sred[NW * kRedQ] = (t == gridDim.x - 1) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (sred[NW * kRedQ] == 0.0) return;
    __threadfence();
    double acc[kRedQ];
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) acc[q] = 0.0;
    for (unsigned int bi = tid; bi < gridDim.x; bi += kThreads) {
#pragma unroll
        for (int q = 0; q < kRedQ; ++q) acc[q] += E.partials[(size_t)bi * kRedQ + q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) {
        double x = acc[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) sred[w * kRedQ + q] = x;
    }
    __syncthreads();
    if (tid == 0) {
        double r[kRedQ];
#pragma unroll
        for (int q = 0; q < kRedQ; ++q) {
            double x = 0.0;
            for (int i = 0; i < NW; ++i) x += sred[i * kRedQ + q];
            r[q] = x * E.dV;
            E.results[q] = r[q];
            if (E.log) E.log[q] = r[q];
        }
        if (E.scale_out) {
            for (int c = 0; c < E.nc; ++c) E.scale_out[c] = sqrt(E.Nnorm[c] * (1.0 / r[c]));
        }
        atomicExch(E.counter, 0u);
    }
}

// ---------------------------------------------------------------------------
// epilogue: one element of the last inverse pass
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 load_scaled(const double2 *p, long long g, const double *sc, int c)
{
    double2 v = p[g];
    if (sc) { double s = sc[c]; v.x *= s; v.y *= s; }
    return v;
}

__device__ __forceinline__ void epilogue(const EpiArgs &E, long long g, int c, double2 kin, double (&red)[kRedQ])
{
    if (E.mode == EPI_STORE) {
        E.out[g] = kin;
        return;
    }
    const long long M = E.M;
    const double2 phi = load_scaled(E.phi, g, E.sc_phi, c);
    const double rho = phi.x * phi.x + phi.y * phi.y;
    double nl = 0.0, enl = 0.0;
    const double V = E.V ? E.V[g] : 0.0;
    if (E.fkind == 1) {  // LHY droplet: a rho + b rho^{3/2}
        const double sr = sqrt(rho);
        nl = E.lhy_a * rho + E.lhy_b * rho * sr;
        enl = 0.5 * E.lhy_a * rho * rho + 0.4 * E.lhy_b * rho * rho * sr;
    } else if (E.fkind == 0) {  // gpFunctional contact couplings
        if (E.has_couplings) {
            if (E.nc == 1) {
                nl = E.g11 * rho;
                enl = 0.5 * E.g11 * rho * rho;
            } else {
                const long long go = c == 0 ? g + M : g - M;
                const double2 po = load_scaled(E.phi, go, E.sc_phi, 1 - c);
                const double ro = po.x * po.x + po.y * po.y;
                const double gcc = c == 0 ? E.g11 : E.g22;
                nl = gcc * rho + E.g12 * ro;
                enl = 0.5 * gcc * rho * rho + 0.5 * E.g12 * rho * ro;
            }
        }
    } else {  // unlocked two-component LHY
        const long long go = c == 0 ? g + M : g - M;
        const double2 po = load_scaled(E.phi, go, E.sc_phi, 1 - c);
        const double ro = po.x * po.x + po.y * po.y;
        const double r1 = c == 0 ? rho : ro, r2 = c == 0 ? ro : rho;
        const double base = r1 + E.beta * r2;
        const double s = base * sqrt(base);
        nl = c == 0 ? r1 + E.eta * r2 + E.alpha * s
                    : E.eta * E.beta * r1 + E.beta * r2 + E.alpha * E.beta * E.beta * s;
    }
    const double a = nl + V;
    const double2 h = make_double2(kin.x + a * phi.x, kin.y + a * phi.y);
    // energy density of the stage input and its norm
    red[2 + c] += phi.x * kin.x + phi.y * kin.y + V * rho + enl;
    red[4 + c] += rho;
    if (E.mode == EPI_H) {
        E.out[g] = h;
        return;
    }
    if (E.mode == EPI_ENERGY) return;
    double2 o;
    if (E.mode == EPI_EULER) {
        o = csub(phi, cmul(make_double2(E.cr, E.ci), h));
    } else if (E.mode == EPI_RK1) {
        E.ksum[g] = h;
        o = csub(phi, cmul(make_double2(E.cr, E.ci), h));
    } else if (E.mode == EPI_RK2 || E.mode == EPI_RK3) {
        const double2 psi = load_scaled(E.psi, g, E.sc_psi, c);
        const double2 ks = E.ksum[g];
        E.ksum[g] = make_double2(ks.x + 2.0 * h.x, ks.y + 2.0 * h.y);
        o = csub(psi, cmul(make_double2(E.cr, E.ci), h));
    } else {  // EPI_RK4: psi - (tau/6) ((k1 + 2k2 + 2k3) + k4)
        const double2 psi = load_scaled(E.psi, g, E.sc_psi, c);
        const double2 ks = E.ksum[g];
        o = csub(psi, cmul(make_double2(E.cr, E.ci), cadd(ks, h)));
    }
    if (E.phase && (E.mode == EPI_EULER || E.mode == EPI_RK4)) {
        const double amp = sqrt(o.x * o.x + o.y * o.y);
        double sn, cs;
        sincos(E.phase[g], &sn, &cs);
        o = make_double2(amp * cs, amp * sn);
    }
    E.out[g] = o;
    red[c] += o.x * o.x + o.y * o.y;
}

// ---------------------------------------------------------------------------
// axis pass kernel
//   MODE 0: FFT along the axis (sign)            src -> dst
//   MODE 1: forward FFT, x k-space multiplier, inverse FFT
//   MODE 2: inverse FFT along x + epilogue (requires I == 1)
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_axis(const double2 *__restrict__ src, double2 *dst, AxisPlan P,
                                                   AxisGeom G, int sign, const double *scale_in,
                                                   long long Opc, MidArgs mid, EpiArgs epi)
{
    extern __shared__ double2 smem[];
    const int n = P.n, B = G.B, tid = threadIdx.x;
    const bool rows = (G.I == 1);
    double2 *bufA = smem, *bufB = smem + (size_t)B * n;
    long long o0 = 0, i0 = 0, ofix = 0;
    int count;
    if (rows) {
        o0 = (long long)blockIdx.x * B;
        count = (int)min((long long)B, G.O - o0);
    } else {
        const long long tpo = (G.I + B - 1) / B;
        ofix = (long long)blockIdx.x / tpo;
        i0 = ((long long)blockIdx.x % tpo) * B;
        count = (int)min((long long)B, G.I - i0);
    }
    const int tileN = B * n;
    // ---- load ----
    for (int e = tid; e < tileN; e += kThreads) {
        int b, j;
        long long gi;
        if (rows) { b = e / n; j = e - b * n; gi = n * o0 + e; }
        else      { b = e % B; j = e / B; gi = (i0 + b) + G.I * (j + (long long)n * ofix); }
        double2 v = make_double2(0.0, 0.0);
        if (b < count) {
            v = src[gi];
            if (scale_in) {
                const int c = (int)((rows ? o0 + b : ofix) / Opc);
                const double s = scale_in[c];
                v.x *= s;
                v.y *= s;
            }
        }
        bufA[lidx(b, j, n, B, rows)] = v;
    }
    __syncthreads();
    double2 *res = lds_fft(bufA, bufB, P, B, rows, sign, tid);
    if constexpr (MODE == 1) {
        for (int e = tid; e < tileN; e += kThreads) {
            int b, j;
            long long gi;
            if (rows) { b = e / n; j = e - b * n; gi = n * o0 + e; }
            else      { b = e % B; j = e / B; gi = (i0 + b) + G.I * (j + (long long)n * ofix); }
            if (b >= count) continue;
            const long long sp = gi % mid.M;
            const int c = (int)(gi / mid.M);
            const int x = (int)(sp % mid.nx);
            const int y = (int)((sp / mid.nx) % mid.ny);
            const int z = (int)(sp / ((long long)mid.nx * mid.ny));
            const double kx = momentum(x, mid.nx, mid.dk[0]);
            const double ky = momentum(y, mid.ny, mid.dk[1]);
            const double kz = momentum(z, mid.nz, mid.dk[2]);
            double f = mid.coef[c] * (kx * kx + ky * ky + kz * kz);
            if (mid.scale) f *= mid.scale[c];
            const int li = lidx(b, j, n, B, rows);
            res[li] = cscale(res[li], f);
        }
        __syncthreads();
        double2 *other = (res == bufA) ? bufB : bufA;
        res = lds_fft(res, other, P, B, rows, -sign, tid);
    }
    if constexpr (MODE == 2) {
        double red[kRedQ];
#pragma unroll
        for (int q = 0; q < kRedQ; ++q) red[q] = 0.0;
        for (int e = tid; e < tileN; e += kThreads) {
            const int b = e / n, j = e - b * n;
            if (b >= count) continue;
            const long long gi = n * o0 + e;
            const int c = (int)((o0 + b) / Opc);
            epilogue(epi, gi, c, res[lidx(b, j, n, B, true)], red);
        }
        if (epi.want_red) {
            __syncthreads();  // the tile in LDS is consumed; reuse it as reduction scratch
            reduce_publish(red, epi, reinterpret_cast<double *>(smem));
        }
        return;
    }
    // ---- store ----
    for (int e = tid; e < tileN; e += kThreads) {
        int b, j;
        long long gi;
        if (rows) { b = e / n; j = e - b * n; gi = n * o0 + e; }
        else      { b = e % B; j = e / B; gi = (i0 + b) + G.I * (j + (long long)n * ofix); }
        if (b < count) dst[gi] = res[lidx(b, j, n, B, rows)];
    }
}

// ---------------------------------------------------------------------------
// pointwise kernels over one component
// ---------------------------------------------------------------------------

// sum |f|^2 over [base, base + len) -> reduce_publish quantity 0
__global__ __launch_bounds__(kThreads) void k_norm(const double2 *__restrict__ f, long long len, EpiArgs E)
{
    __shared__ double sred[kRedScratch];
    double red[kRedQ];
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) red[q] = 0.0;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long long)gridDim.x * kThreads) {
        const double2 v = f[i];
        red[0] += v.x * v.x + v.y * v.y;
    }
    reduce_publish(red, E, sred);
}

// f *= s (host value s when sp is null; else sp[0] for one component, sp[c] for all)
__global__ __launch_bounds__(kThreads) void k_scale(double2 *f, long long M, int nc, int comp, double s,
                                                    const double *sp)
{
    const long long len = comp >= 0 ? M : M * nc;
    const long long base = comp >= 0 ? M * comp : 0;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long long)gridDim.x * kThreads) {
        const long long g = base + i;
        const double f0 = sp ? sp[comp >= 0 ? 0 : g / M] : s;
        double2 v = f[g];
        v.x *= f0;
        v.y *= f0;
        f[g] = v;
    }
}

// max |f|^2 over one component; per-block maxima, reduced on the host
__global__ __launch_bounds__(kThreads) void k_maxdens(const double2 *__restrict__ f, long long len, double *blockmax)
{
    __shared__ double s[kThreads / 64];
    double m = 0.0;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long long)gridDim.x * kThreads) {
        const double2 v = f[i];
        m = fmax(m, v.x * v.x + v.y * v.y);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = s[0];
        for (int i = 1; i < kThreads / 64; ++i) x = fmax(x, s[i]);
        blockmax[blockIdx.x] = x;
    }
}

// psi <- |psi| e^{i phase}
__global__ __launch_bounds__(kThreads) void k_phase(double2 *f, const double *__restrict__ phase, long long len)
{
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long long)gridDim.x * kThreads) {
        const double2 v = f[i];
        const double a = sqrt(v.x * v.x + v.y * v.y);
        double sn, cs;
        sincos(phase[i], &sn, &cs);
        f[i] = make_double2(a * cs, a * sn);
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static size_t lds_bytes(const AxisPlan &P, const AxisGeom &G)
{
    return std::max(2 * sizeof(double2) * (size_t)G.B * P.n, sizeof(double) * (size_t)kRedScratch);
}

static unsigned tiles_of(const AxisGeom &G)
{
    if (G.I == 1) return (unsigned)((G.O + G.B - 1) / G.B);
    return (unsigned)(((G.I + G.B - 1) / G.B) * G.O);
}

hipError_t launch_axis(hipStream_t st, int mode, const double2 *src, double2 *dst, const AxisPlan &P,
                       const AxisGeom &G, int sign, const double *scale_in, long long Opc, const MidArgs &mid,
                       const EpiArgs &epi)
{
    dim3 grid(tiles_of(G)), block(kThreads);
    size_t sh = lds_bytes(P, G);
    if (mode == 0) hipLaunchKernelGGL(k_axis<0>, grid, block, sh, st, src, dst, P, G, sign, scale_in, Opc, mid, epi);
    else if (mode == 1) hipLaunchKernelGGL(k_axis<1>, grid, block, sh, st, src, dst, P, G, sign, scale_in, Opc, mid, epi);
    else hipLaunchKernelGGL(k_axis<2>, grid, block, sh, st, src, dst, P, G, sign, scale_in, Opc, mid, epi);
    return hipGetLastError();
}

unsigned axis_grid(const AxisGeom &G) { return tiles_of(G); }

hipError_t launch_norm(hipStream_t st, const double2 *f, long long len, unsigned blocks, const EpiArgs &E)
{
    hipLaunchKernelGGL(k_norm, dim3(blocks), dim3(kThreads), 0, st, f, len, E);
    return hipGetLastError();
}

hipError_t launch_scale(hipStream_t st, double2 *f, long long M, int nc, int comp, double s, const double *sp)
{
    long long len = comp >= 0 ? M : M * nc;
    unsigned blocks = (unsigned)std::min<long long>((len + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(k_scale, dim3(blocks), dim3(kThreads), 0, st, f, M, nc, comp, s, sp);
    return hipGetLastError();
}

hipError_t launch_maxdens(hipStream_t st, const double2 *f, long long len, unsigned blocks, double *blockmax)
{
    hipLaunchKernelGGL(k_maxdens, dim3(blocks), dim3(kThreads), 0, st, f, len, blockmax);
    return hipGetLastError();
}

hipError_t launch_phase(hipStream_t st, double2 *f, const double *phase, long long len)
{
    unsigned blocks = (unsigned)std::min<long long>((len + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(k_phase, dim3(blocks), dim3(kThreads), 0, st, f, phase, len);
    return hipGetLastError();
}

}  // namespace gp5
