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
#include "gp5_device.h"

#include <math.h>

#include <algorithm>

namespace gp5 {


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
// axis pass kernel
//   MODE 0: FFT along the axis (sign)            src -> dst
//   MODE 1: forward FFT, x k-space multiplier, inverse FFT
//   MODE 2: inverse FFT along x + epilogue (requires I == 1)
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_axis(const double2 *__restrict__ src, double2 *dst, AxisPlan P,
                                                   AxisGeom G, int sign, const double *scale_in, LoadPre pre,
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
    // retired batch instances: the tile's first line decides (block-uniform)
    if (G.active && (rows ? o0 : ofix) >= (long long)*G.active * Opc) return;
    const int tileN = B * n;
    // ---- load ----
    for (int e = tid; e < tileN; e += kThreads) {
        int b, j;
        long long gi;
        if (rows) { b = e / n; j = e - b * n; gi = n * o0 + e; }
        else      { b = e % B; j = e / B; gi = (i0 + b) + G.I * (j + (long long)n * ofix); }
        double2 v = make_double2(0.0, 0.0);
        if (b < count) {
            const int c = (int)((rows ? o0 + b : ofix) / Opc);
            v = load_pre(src, gi, scale_in, c, pre.src2, make_double2(pre.cfr, pre.cfi));
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
            double f = mid.coef[c == 0 ? 0 : 1] * (kx * kx + ky * ky + kz * kz);
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
            epilogue(epi, gi, c, c - (int)(o0 / Opc), res[lidx(b, j, n, B, true)], red);
        }
        if (epi.want_red) {
            __syncthreads();  // the tile in LDS is consumed; reuse it as reduction scratch
            reduce_block(red, epi.partials, reinterpret_cast<double *>(smem));
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

// sum |f|^2 over [0, len) -> per-block partial of quantity 0
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
    reduce_block(red, E.partials, sred);
}

// Sum the per-block partials of the preceding kernel in a fixed order and
// publish results[q][c] = sum * dV, the optional per-step log copy and, for
// lazy renormalisation, scale_out[c] = sqrt(N_c / norm_c)
// (src/geometry.cpp:217-229).  One block per component c: the partial blocks
// whose lines [b*lpb, (b+1)*lpb) meet c's lines [c*lpc, (c+1)*lpc) hold c in
// slot c - b*lpb/lpc (lpb = 0: every block belongs to component 0).
constexpr int kFinalThreads = 512;
__global__ __launch_bounds__(kFinalThreads) void k_finalize(EpiArgs E, unsigned nblocks, long long lpb, long long lpc,
                                                           int ncomp)
{
    __shared__ double sred[kRedScratch];
    const int c = blockIdx.x;
    if (E.active && c >= *E.active) return;  // retired batch instance
    long long b0 = 0, b1 = nblocks;
    if (lpb > 0) {
        b0 = (long long)c * lpc / lpb;
        b1 = min((long long)nblocks, ((long long)(c + 1) * lpc - 1) / lpb + 1);
    }
    double acc[kRedQ];
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) acc[q] = 0.0;
    for (long long b = b0 + threadIdx.x; b < b1; b += kFinalThreads) {
        const int slot = lpb > 0 ? c - (int)(b * lpb / lpc) : 0;
        const double *p = E.partials + (size_t)b * kRedQ + kRedN * slot;
#pragma unroll
        for (int q = 0; q < kRedN; ++q) acc[q] += p[q];
    }
    double s[kRedQ];
    block_sum(acc, sred, s);
    if (threadIdx.x == 0) {
        double r[kRedN];
#pragma unroll
        for (int q = 0; q < kRedN; ++q) {
            r[q] = s[q] * E.dV;
            E.results[q * ncomp + c] = r[q];
            if (E.log) E.log[q * ncomp + c] = r[q];
        }
        if (E.scale_out) {
            const double N = E.Nnorm_dev ? E.Nnorm_dev[c] : E.Nnorm[c < 2 ? c : 1];
            E.scale_out[c] = sqrt(N * (1.0 / r[0]));
        }
    }
}

hipError_t launch_finalize(hipStream_t st, const EpiArgs &E, unsigned nblocks, long long lpb, long long lpc, int ncomp)
{
    hipLaunchKernelGGL(k_finalize, dim3(ncomp), dim3(kFinalThreads), 0, st, E, nblocks, lpb, lpc, ncomp);
    return hipGetLastError();
}

// f *= s (host value s when sp is null; else sp[0] for one component, sp[c] for all)
__global__ __launch_bounds__(kThreads) void k_scale(double2 *f, long long M, int nc, int comp, double s,
                                                    const double *sp, const int *active)
{
    // batch classes: the live components only
    const long long len = comp >= 0 ? M : M * (active ? min(nc, *active) : nc);
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

// out = a - c k, c = cr + i ci (the stage update fieldDataOld - timeStep() * f * k of
// src/stepper.cpp:35,50,55,60 as a separate pass; out may alias a or k)
__global__ __launch_bounds__(kThreads) void k_axpy(double2 *out, const double2 *a, double cr, double ci,
                                                   const double2 *k, long long len)
{
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long long)gridDim.x * kThreads) {
        const double2 kv = k[i], av = a[i];
        out[i] = make_double2(av.x - (cr * kv.x - ci * kv.y), av.y - (cr * kv.y + ci * kv.x));
    }
}

// out = a - c (((k1 + 2 k2) + 2 k3) + k4) (src/stepper.cpp:64, in the reference's sum order)
__global__ __launch_bounds__(kThreads) void k_rk4_combine(double2 *out, const double2 *a, double cr, double ci,
                                                          const double2 *k1, const double2 *k2, const double2 *k3,
                                                          const double2 *k4, long long len)
{
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long long)gridDim.x * kThreads) {
        const double2 v1 = k1[i], v2 = k2[i], v3 = k3[i], v4 = k4[i], av = a[i];
        const double sr = ((v1.x + 2 * v2.x) + 2 * v3.x) + v4.x;
        const double si = ((v1.y + 2 * v2.y) + 2 * v3.y) + v4.y;
        out[i] = make_double2(av.x - (cr * sr - ci * si), av.y - (cr * si + ci * sr));
    }
}

// ex = e^{i phase}: the phaseConstraint factor, precomputed once per advance as
// the reference does in its constructor (src/constraint.cpp:12-17)
__global__ __launch_bounds__(kThreads) void k_phase_exp(double2 *ex, const double *__restrict__ phase, long long len)
{
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long long)gridDim.x * kThreads) {
        double sn, cs;
        sincos(phase[i], &sn, &cs);
        ex[i] = make_double2(cs, sn);
    }
}

hipError_t launch_phase_exp(hipStream_t st, double2 *ex, const double *phase, long long len)
{
    unsigned blocks = (unsigned)std::min<long long>((len + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(k_phase_exp, dim3(blocks), dim3(kThreads), 0, st, ex, phase, len);
    return hipGetLastError();
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
                       const AxisGeom &G, int sign, const double *scale_in, const LoadPre &pre, long long Opc,
                       const MidArgs &mid, const EpiArgs &epi)
{
    dim3 grid(tiles_of(G)), block(kThreads);
    size_t sh = lds_bytes(P, G);
    if (mode == 0)
        hipLaunchKernelGGL(k_axis<0>, grid, block, sh, st, src, dst, P, G, sign, scale_in, pre, Opc, mid, epi);
    else if (mode == 1)
        hipLaunchKernelGGL(k_axis<1>, grid, block, sh, st, src, dst, P, G, sign, scale_in, pre, Opc, mid, epi);
    else
        hipLaunchKernelGGL(k_axis<2>, grid, block, sh, st, src, dst, P, G, sign, scale_in, pre, Opc, mid, epi);
    return hipGetLastError();
}

unsigned axis_grid(const AxisGeom &G) { return tiles_of(G); }

hipError_t launch_norm(hipStream_t st, const double2 *f, long long len, unsigned blocks, const EpiArgs &E)
{
    hipLaunchKernelGGL(k_norm, dim3(blocks), dim3(kThreads), 0, st, f, len, E);
    return hipGetLastError();
}

hipError_t launch_scale(hipStream_t st, double2 *f, long long M, int nc, int comp, double s, const double *sp,
                        const int *active)
{
    long long len = comp >= 0 ? M : M * nc;
    unsigned blocks = (unsigned)std::min<long long>((len + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(k_scale, dim3(blocks), dim3(kThreads), 0, st, f, M, nc, comp, s, sp, active);
    return hipGetLastError();
}

hipError_t launch_maxdens(hipStream_t st, const double2 *f, long long len, unsigned blocks, double *blockmax)
{
    hipLaunchKernelGGL(k_maxdens, dim3(blocks), dim3(kThreads), 0, st, f, len, blockmax);
    return hipGetLastError();
}

hipError_t launch_axpy(hipStream_t st, double2 *out, const double2 *a, double cr, double ci, const double2 *k,
                       long long len)
{
    unsigned blocks = (unsigned)std::min<long long>((len + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(kThreads), 0, st, out, a, cr, ci, k, len);
    return hipGetLastError();
}

hipError_t launch_rk4_combine(hipStream_t st, double2 *out, const double2 *a, double cr, double ci, const double2 *k1,
                              const double2 *k2, const double2 *k3, const double2 *k4, long long len)
{
    unsigned blocks = (unsigned)std::min<long long>((len + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(k_rk4_combine, dim3(blocks), dim3(kThreads), 0, st, out, a, cr, ci, k1, k2, k3, k4, len);
    return hipGetLastError();
}

hipError_t launch_phase(hipStream_t st, double2 *f, const double *phase, long long len)
{
    unsigned blocks = (unsigned)std::min<long long>((len + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(k_phase, dim3(blocks), dim3(kThreads), 0, st, f, phase, len);
    return hipGetLastError();
}

}  // namespace gp5
