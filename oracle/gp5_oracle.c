/*
 * gp5_oracle.c -- CPU restatement of GP5's pseudo-spectral time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in gp5_amd/ links, imports or calls it.
 *
 * It restates, in plain C, the algorithm of the reference files cited at each
 * function.  The reference's third-party arithmetic (FFTW3's unnormalised
 * c2c DFT, Eigen's elementwise tensor expressions) is restated from its
 * published definition: FFTW_FORWARD = -1 computes sum_n x_n e^{-2 pi i k n/N},
 * FFTW_BACKWARD = +1 the same with +i, both unnormalised.
 *
 * Memory layout follows src/traits.h:12-20 (Eigen::Tensor<complex<double>,4>,
 * column-major (x,y,z,component), interleaved re/im doubles).
 *
 * Optional OpenMP parallelism stands in for the reference's MPI ranks; it only
 * splits independent lines/points, so results do not depend on the thread
 * count except through reduction order in or_norm (summed per z-plane, then
 * serially).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXF 64

/* ------------------------------------------------------------------------ */
/* 1-D mixed-radix FFT (Stockham autosort), any length.                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;
    int nf;
    int f[OR_MAXF];
    double *tw; /* tw[2t], tw[2t+1] = cos, sin of -2 pi t / n */
} or_plan1d;

static void or_twiddles(int n, double *tw)
{
    /* exp(-2 pi i t/n), evaluated in long double from the exact integer phase */
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
        long double a = -two_pi * (long double)t / (long double)n;
        tw[2 * t] = (double)cosl(a);
        tw[2 * t + 1] = (double)sinl(a);
    }
}

static int or_plan_init(or_plan1d *p, int n)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    /* radix 4 first, then 2, 3, 5, then any remaining prime */
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    for (int r = 3; m > 1; r += 2) {
        while (m % r == 0) { p->f[p->nf++] = r; m /= r; }
        if (r * r > m && m > 1) { p->f[p->nf++] = m; m = 1; }
    }
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    if (!p->tw) return -1;
    or_twiddles(n, p->tw);
    return 0;
}

static void or_plan_free(or_plan1d *p) { free(p->tw); p->tw = NULL; }

/* One Stockham step: input a (length n), output b.  L = product of radices
 * already applied, R = this radix, m = n/(L*R).  Element (j, k, q) with
 * j in [0,L), q in [0,m): the L-point sub-transforms are already done. */
static void or_stockham(const double *a, double *b, int n, int L, int R, int sign, const double *tw)
{
    int m = n / (L * R);
    double xr[OR_MAXF * 4], xi[OR_MAXF * 4];
    double *vr = xr, *vi = xi;
    double *hr = NULL, *hi = NULL;
    if (R > OR_MAXF * 4) {
        hr = (double *)malloc(sizeof(double) * (size_t)R * 2);
        hi = hr + R;
        vr = hr; vi = hi;
    }
    if (R == 2 || R == 4) {
        /* radix 2 / 4 butterflies written out (same algebra as the direct DFT
         * below, with w_4 = -i forward, +i backward) */
        for (int j = 0; j < L; ++j) {
            double wr[4] = {1, 0, 0, 0}, wi[4] = {0, 0, 0, 0};
            for (int r = 1; r < R; ++r) {
                size_t t = ((size_t)j * (size_t)r * (size_t)m) % (size_t)n;
                wr[r] = tw[2 * t];
                wi[r] = sign > 0 ? -tw[2 * t + 1] : tw[2 * t + 1];
            }
            for (int q = 0; q < m; ++q) {
                double ar[4], ai[4];
                for (int r = 0; r < R; ++r) {
                    size_t idx = (size_t)q + (size_t)m * ((size_t)r + (size_t)R * j);
                    double xr = a[2 * idx], xi = a[2 * idx + 1];
                    ar[r] = xr * wr[r] - xi * wi[r];
                    ai[r] = xr * wi[r] + xi * wr[r];
                }
                if (R == 2) {
                    size_t o0 = (size_t)q + (size_t)m * (size_t)j;
                    size_t o1 = (size_t)q + (size_t)m * ((size_t)j + (size_t)L);
                    b[2 * o0] = ar[0] + ar[1];
                    b[2 * o0 + 1] = ai[0] + ai[1];
                    b[2 * o1] = ar[0] - ar[1];
                    b[2 * o1 + 1] = ai[0] - ai[1];
                } else {
                    double s = sign > 0 ? 1.0 : -1.0; /* w_4 = s*i */
                    double t0r = ar[0] + ar[2], t0i = ai[0] + ai[2];
                    double t1r = ar[0] - ar[2], t1i = ai[0] - ai[2];
                    double t2r = ar[1] + ar[3], t2i = ai[1] + ai[3];
                    double dr = ar[1] - ar[3], di = ai[1] - ai[3];
                    double t3r = -s * di, t3i = s * dr; /* (s i) * d */
                    size_t o[4];
                    for (int k = 0; k < 4; ++k) o[k] = (size_t)q + (size_t)m * ((size_t)j + (size_t)L * k);
                    b[2 * o[0]] = t0r + t2r;
                    b[2 * o[0] + 1] = t0i + t2i;
                    b[2 * o[2]] = t0r - t2r;
                    b[2 * o[2] + 1] = t0i - t2i;
                    b[2 * o[1]] = t1r + t3r;
                    b[2 * o[1] + 1] = t1i + t3i;
                    b[2 * o[3]] = t1r - t3r;
                    b[2 * o[3] + 1] = t1i - t3i;
                }
            }
        }
        free(hr);
        return;
    }
    for (int j = 0; j < L; ++j) {
        for (int q = 0; q < m; ++q) {
            /* gather the R inputs of this butterfly: a[q + m*(j + L*r)]?  We use
             * the "decimation in time" Stockham ordering: input index
             * q + m*(r + R*j), output index q + m*(j + L*k). */
            for (int r = 0; r < R; ++r) {
                size_t idx = (size_t)q + (size_t)m * ((size_t)r + (size_t)R * j);
                vr[r] = a[2 * idx];
                vi[r] = a[2 * idx + 1];
            }
            /* twiddle: input r multiplied by w_{L R}^{sign*j*r}; the table holds
             * e^{-2 pi i t/n} so w_{LR}^{j r} = tw[j*r*m] */
            for (int r = 1; r < R; ++r) {
                size_t t = ((size_t)j * (size_t)r * (size_t)m) % (size_t)n;
                double c = tw[2 * t], s = tw[2 * t + 1];
                if (sign > 0) s = -s;
                double tr = vr[r] * c - vi[r] * s;
                double ti = vr[r] * s + vi[r] * c;
                vr[r] = tr;
                vi[r] = ti;
            }
            /* R-point DFT, direct (the twiddle table of length n contains the
             * R-th roots at stride n/R) */
            for (int k = 0; k < R; ++k) {
                double sr = 0.0, si = 0.0;
                for (int r = 0; r < R; ++r) {
                    size_t t = ((size_t)r * (size_t)k % (size_t)R) * (size_t)(n / R);
                    double c = tw[2 * t], s = tw[2 * t + 1];
                    if (sign > 0) s = -s;
                    sr += vr[r] * c - vi[r] * s;
                    si += vr[r] * s + vi[r] * c;
                }
                size_t o = (size_t)q + (size_t)m * ((size_t)j + (size_t)L * k);
                b[2 * o] = sr;
                b[2 * o + 1] = si;
            }
        }
    }
    free(hr);
}

/* Transform one contiguous line in place; work has 2*n doubles. */
static void or_fft_line(double *x, double *work, const or_plan1d *p, int sign)
{
    int n = p->n;
    if (n <= 1) return;
    double *a = x, *b = work;
    int L = 1;
    for (int s = 0; s < p->nf; ++s) {
        int R = p->f[s];
        or_stockham(a, b, n, L, R, sign, p->tw);
        L *= R;
        double *t = a; a = b; b = t;
    }
    if (a != x) memcpy(x, a, sizeof(double) * 2 * (size_t)n);
}

/* ------------------------------------------------------------------------ */
/* 3-D transform: src/fourierTransform.cpp:146-214 (fftw_plan_many_dft with   */
/* n = {Nz,Ny,Nx} on the column-major buffer = the standard 3-D DFT; loop over */
/* components at offsets c*M; unnormalised; out-of-place with src preserved). */
/* ------------------------------------------------------------------------ */

int or_fft3d(const double *src, double *dst, int nx, int ny, int nz, int nc, int sign)
{
    const int n[3] = {nx, ny, nz};
    const size_t M = (size_t)nx * ny * nz;
    if (dst != src) memcpy(dst, src, sizeof(double) * 2 * M * (size_t)nc);
    for (int d = 0; d < 3; ++d) {
        if (n[d] <= 1) continue;
        or_plan1d p;
        if (or_plan_init(&p, n[d])) return -1;
        size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)nx : (size_t)nx * ny);
        size_t inner = stride;                 /* lines are indexed by (i < inner, o) */
        size_t outer = M / (stride * n[d]) * (size_t)nc;
        long nlines = (long)(inner * outer);
        int len = n[d];
#pragma omp parallel
        {
            double *line = (double *)malloc(sizeof(double) * 4 * (size_t)len);
            double *work = line + 2 * (size_t)len;
#pragma omp for schedule(static)
            for (long l = 0; l < nlines; ++l) {
                size_t i = (size_t)l % inner, o = (size_t)l / inner;
                double *base = dst + 2 * (i + stride * (size_t)len * o);
                for (int j = 0; j < len; ++j) {
                    line[2 * j] = base[2 * stride * j];
                    line[2 * j + 1] = base[2 * stride * j + 1];
                }
                or_fft_line(line, work, &p, sign);
                for (int j = 0; j < len; ++j) {
                    base[2 * stride * j] = line[2 * j];
                    base[2 * stride * j + 1] = line[2 * j + 1];
                }
            }
            free(line);
        }
        or_plan_free(&p);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Geometry: src/geometry.h:88-90, src/geometry.cpp:108-152 (momentums),      */
/* :165-188 (positions), :190-237 (norm / normalize), :241-276 (Gaussian).   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n[3];
    double left[3];
    double right[3];
    int nc;
} or_geom;

static double or_dx(const or_geom *g, int d) { return (g->right[d] - g->left[d]) / g->n[d]; }
static double or_dk(const or_geom *g, int d) { return 2 * M_PI / (g->right[d] - g->left[d]); }

/* k_d for global index i along axis d (geometry.cpp:127-148) */
double or_momentum(int i, int n, double dk)
{
    if (n % 2 == 0) return i < n / 2 ? dk * i : (double)(-n + i) * dk;
    return i <= (n - 1) / 2 ? dk * i : (double)(-n + i) * dk;
}

double or_position(int i, double left, double dx) { return left + dx * (i + 0.5); }

/* gp::norm (geometry.cpp:190-214): sum(re^2 + im^2) * dV.  Summation is per
 * z-plane then serial over planes so the result does not depend on the
 * OpenMP thread count. */
double or_norm(const double *f, const or_geom *g, int c)
{
    const size_t M = (size_t)g->n[0] * g->n[1] * g->n[2];
    const size_t P = (size_t)g->n[0] * g->n[1];
    const double *fc = f + 2 * M * (size_t)c;
    int nz = g->n[2];
    double *plane = (double *)malloc(sizeof(double) * (size_t)nz);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
        double s = 0.0;
        const double *p = fc + 2 * P * (size_t)z;
        for (size_t i = 0; i < P; ++i) {
            s += p[2 * i] * p[2 * i];
            s += p[2 * i + 1] * p[2 * i + 1];
        }
        plane[z] = s;
    }
    double sum = 0.0;
    for (int z = 0; z < nz; ++z) sum += plane[z];
    free(plane);
    double dV = 1.0;
    for (int d = 0; d < 3; ++d) dV *= or_dx(g, d);
    return sum * dV;
}

/* gp::normalize (geometry.cpp:217-229): psi_c *= sqrt(N * (1/norm_c)) */
void or_normalize(double *f, const or_geom *g, int c, double N)
{
    double s = sqrt(N * (1.0 / or_norm(f, g, c)));
    const size_t M = (size_t)g->n[0] * g->n[1] * g->n[2];
    double *fc = f + 2 * M * (size_t)c;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)M; ++i) {
        fc[2 * i] *= s;
        fc[2 * i + 1] *= s;
    }
}

/* ------------------------------------------------------------------------ */
/* Spectral Laplacian: src/operators.cpp:7-36.                               */
/* lap = FFT_bwd( -K2 * FFT_fwd(src) ) * (1/M)                               */
/* ------------------------------------------------------------------------ */

int or_laplacian(const double *src, double *dst, const or_geom *g, double *work)
{
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2], nc = g->nc;
    const size_t M = (size_t)nx * ny * nz;
    if (or_fft3d(src, work, nx, ny, nz, nc, -1)) return -1;
    const double dkx = or_dk(g, 0), dky = or_dk(g, 1), dkz = or_dk(g, 2);
#pragma omp parallel for schedule(static)
    for (long zc = 0; zc < (long)nz * nc; ++zc) {
        int z = (int)(zc % nz);
        double kz = or_momentum(z, nz, dkz);
        for (int y = 0; y < ny; ++y) {
            double ky = or_momentum(y, ny, dky);
            for (int x = 0; x < nx; ++x) {
                double kx = or_momentum(x, nx, dkx);
                double k2 = kx * kx + ky * ky + kz * kz;
                size_t i = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)zc);
                work[2 * i] = -(work[2 * i] * k2);
                work[2 * i + 1] = -(work[2 * i + 1] * k2);
            }
        }
    }
    if (or_fft3d(work, dst, nx, ny, nz, nc, +1)) return -1;
    const double inv = 1.0 / (double)M;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)(M * nc); ++i) {
        dst[2 * i] *= inv;
        dst[2 * i + 1] *= inv;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Functionals: src/functional.cpp:13-150.                                   */
/* ------------------------------------------------------------------------ */

enum { OR_GP = 0, OR_LHY = 1, OR_UNLOCKED = 2 };

typedef struct {
    int kind;
    int has_couplings;     /* gpFunctional::_setCouplings */
    double g[4];           /* couplings(i,j) at g[i + 2*j] */
    double inv_mass[2];    /* gpFunctional::_inverseMasses (init(): 1/m) */
    double lhy_a, lhy_b;   /* LHYDroplet: -3, 5/2 (functional.cpp:117-118) */
    double alpha, beta, eta;
    const double *V;       /* real potential, M*nc doubles, or NULL (addPotential) */
} or_func;

/* dst = H(src); work needs 2*M*nc doubles */
int or_apply(const or_func *F, const double *src, double *dst, const or_geom *g, double *work)
{
    const size_t M = (size_t)g->n[0] * g->n[1] * g->n[2];
    const int nc = g->nc;
    if (or_laplacian(src, dst, g, work)) return -1;
    if (F->kind == OR_GP) {
        /* gpFunctional::apply (functional.cpp:13-43) */
        for (int c = 0; c < nc; ++c) {
            double f = -0.5 * F->inv_mass[c];
#pragma omp parallel for schedule(static)
            for (long i = 0; i < (long)M; ++i) {
                size_t p = (size_t)i + M * c;
                dst[2 * p] *= f;
                dst[2 * p + 1] *= f;
            }
        }
        if (F->V) { /* addPotential (functional.cpp:46-53) */
#pragma omp parallel for schedule(static)
            for (long p = 0; p < (long)(M * nc); ++p) {
                dst[2 * p] += F->V[p] * src[2 * p];
                dst[2 * p + 1] += F->V[p] * src[2 * p + 1];
            }
        }
        if (F->has_couplings) {
            if (nc == 1) { /* functional.cpp:54-68 */
                double gg = F->g[0];
#pragma omp parallel for schedule(static)
                for (long p = 0; p < (long)M; ++p) {
                    double re = src[2 * p], im = src[2 * p + 1];
                    double rho = re * re + im * im;
                    dst[2 * p] += gg * rho * re;
                    dst[2 * p + 1] += gg * rho * im;
                }
            } else if (nc == 2) { /* functional.cpp:71-89 */
                double g11 = F->g[0], g22 = F->g[3], g12 = F->g[2];
#pragma omp parallel for schedule(static)
                for (long p = 0; p < (long)M; ++p) {
                    size_t q = (size_t)p + M;
                    double r0 = src[2 * p] * src[2 * p] + src[2 * p + 1] * src[2 * p + 1];
                    double r1 = src[2 * q] * src[2 * q] + src[2 * q + 1] * src[2 * q + 1];
                    double a0 = g11 * r0 + g12 * r1, a1 = g22 * r1 + g12 * r0;
                    dst[2 * p] += a0 * src[2 * p];
                    dst[2 * p + 1] += a0 * src[2 * p + 1];
                    dst[2 * q] += a1 * src[2 * q];
                    dst[2 * q + 1] += a1 * src[2 * q + 1];
                }
            }
        }
    } else if (F->kind == OR_LHY) {
        /* LHYDropletFunctional::apply (functional.cpp:105-122) */
#pragma omp parallel for schedule(static)
        for (long p = 0; p < (long)(M * nc); ++p) {
            double re = src[2 * p], im = src[2 * p + 1];
            double rho = re * re + im * im;
            double a = F->lhy_a * rho + F->lhy_b * pow(rho, 1.5);
            dst[2 * p] = -0.5 * dst[2 * p] + a * re;
            dst[2 * p + 1] = -0.5 * dst[2 * p + 1] + a * im;
        }
        if (F->V) {
#pragma omp parallel for schedule(static)
            for (long p = 0; p < (long)(M * nc); ++p) {
                dst[2 * p] += F->V[p] * src[2 * p];
                dst[2 * p + 1] += F->V[p] * src[2 * p + 1];
            }
        }
    } else if (F->kind == OR_UNLOCKED) {
        /* LHYDropletUnlockedFunctional::apply (functional.cpp:125-150) */
        if (nc != 2) return -2;
        double al = F->alpha, be = F->beta, et = F->eta;
#pragma omp parallel for schedule(static)
        for (long p = 0; p < (long)M; ++p) {
            size_t q = (size_t)p + M;
            double r1 = src[2 * p] * src[2 * p] + src[2 * p + 1] * src[2 * p + 1];
            double r2 = src[2 * q] * src[2 * q] + src[2 * q + 1] * src[2 * q + 1];
            double s = pow(r1 + be * r2, 1.5);
            double a0 = r1 + et * r2 + al * s;
            double a1 = et * be * r1 + be * r2 + al * be * be * s;
            dst[2 * p] = -0.5 * dst[2 * p] + a0 * src[2 * p];
            dst[2 * p + 1] = -0.5 * dst[2 * p + 1] + a0 * src[2 * p + 1];
            dst[2 * q] = -0.5 * dst[2 * q] + a1 * src[2 * q];
            dst[2 * q + 1] = -0.5 * dst[2 * q + 1] + a1 * src[2 * q + 1];
        }
        if (F->V) {
#pragma omp parallel for schedule(static)
            for (long p = 0; p < (long)(M * nc); ++p) {
                dst[2 * p] += F->V[p] * src[2 * p];
                dst[2 * p + 1] += F->V[p] * src[2 * p + 1];
            }
        }
    } else {
        return -3;
    }
    return 0;
}

/* Energy E = sum_c int [ Re(psi_c^* (-1/2 m_c^-1 lap psi_c)) + V_c rho_c ] + int e(rho),
 * with e the density whose derivative is the functional's nonlinearity
 * (LHY: a/2 rho^2 + 2b/5 rho^{5/2}; gp: 1/2 sum_ij g_ij rho_i rho_j).
 * The reference has no energy; this is the extension used for conservation
 * checks.  Returns NAN for the unlocked functional (not a gradient for
 * general alpha, beta, eta). */
double or_energy(const or_func *F, const double *src, const or_geom *g, double *work, double *lap)
{
    if (F->kind == OR_UNLOCKED) return NAN;
    const size_t M = (size_t)g->n[0] * g->n[1] * g->n[2];
    const int nc = g->nc;
    if (or_laplacian(src, lap, g, work)) return NAN;
    double dV = 1.0;
    for (int d = 0; d < 3; ++d) dV *= or_dx(g, d);
    double tot = 0.0;
    for (int c = 0; c < nc; ++c) {
        double kin = F->kind == OR_GP ? -0.5 * F->inv_mass[c] : -0.5;
        double s = 0.0;
        for (size_t i = 0; i < M; ++i) {
            size_t p = i + M * (size_t)c;
            double re = src[2 * p], im = src[2 * p + 1];
            double rho = re * re + im * im;
            double e = kin * (re * lap[2 * p] + im * lap[2 * p + 1]);
            if (F->V) e += F->V[p] * rho;
            if (F->kind == OR_LHY) {
                e += 0.5 * F->lhy_a * rho * rho + 0.4 * F->lhy_b * rho * rho * sqrt(rho);
            } else if (F->has_couplings) {
                if (nc == 1) e += 0.5 * F->g[0] * rho * rho;
                else {
                    size_t o = i + M * (size_t)(1 - c);
                    double ro = src[2 * o] * src[2 * o] + src[2 * o + 1] * src[2 * o + 1];
                    double gcc = F->g[c + 2 * c], g12 = F->g[2];
                    e += 0.5 * gcc * rho * rho + 0.5 * g12 * rho * ro;
                }
            }
            s += e;
        }
        tot += s;
    }
    return tot * dV;
}

/* ------------------------------------------------------------------------ */
/* Steppers: src/stepper.cpp:7-81, driven as pyInterface::timeStepper::      */
/* advance (src/gpCpp.cpp:255-269).  tau = (dt,0) imaginary, (0,dt) real.     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int kind;              /* 0 eulero, 1 RK4 */
    double tau_re, tau_im;
    int renorm;            /* stepper::_reNormalize */
    double N[2];           /* stepper::_normalizations */
    const double *phase;   /* phaseConstraint phase (M*nc real) or NULL */
} or_stepper;

/* stepper::normalize (stepper.cpp:7-22) */
static void or_renorm(const or_stepper *S, double *f, const or_geom *g)
{
    if (!S->renorm) return;
    for (int c = 0; c < g->nc; ++c) or_normalize(f, g, c, S->N[c]);
}

/* phaseConstraint::apply (constraint.cpp:6-9): psi <- sqrt(psi conj(psi)) e^{i phi} */
static void or_constraint(const or_stepper *S, double *f, const or_geom *g)
{
    if (!S->phase) return;
    const size_t M = (size_t)g->n[0] * g->n[1] * g->n[2] * g->nc;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)M; ++p) {
        double a = sqrt(f[2 * p] * f[2 * p] + f[2 * p + 1] * f[2 * p + 1]);
        f[2 * p] = a * cos(S->phase[p]);
        f[2 * p + 1] = a * sin(S->phase[p]);
    }
}

/* out = a - (cr + i ci) * k  (complex scalar times complex tensor) */
static void or_axpy(double *out, const double *a, double cr, double ci, const double *k, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)n; ++p) {
        double kr = k[2 * p], ki = k[2 * p + 1];
        double tr = cr * kr - ci * ki, ti = cr * ki + ci * kr;
        out[2 * p] = a[2 * p] - tr;
        out[2 * p + 1] = a[2 * p + 1] - ti;
    }
}

/* psi advanced by nsteps in place. */
int or_advance(const or_func *F, const or_stepper *S, double *psi, const or_geom *g, int nsteps)
{
    const size_t n = (size_t)g->n[0] * g->n[1] * g->n[2] * g->nc;
    const size_t bytes = sizeof(double) * 2 * n;
    double *work = (double *)malloc(bytes);
    double *nw = (double *)malloc(bytes);
    double *k1 = NULL, *k2 = NULL, *k3 = NULL, *k4 = NULL;
    if (!work || !nw) { free(work); free(nw); return -1; }
    if (S->kind == 1) {
        k1 = (double *)malloc(bytes); k2 = (double *)malloc(bytes);
        k3 = (double *)malloc(bytes); k4 = (double *)malloc(bytes);
        if (!k1 || !k2 || !k3 || !k4) { free(work); free(nw); free(k1); free(k2); free(k3); free(k4); return -1; }
    }
    const double tr = S->tau_re, ti = S->tau_im;
    int rc = 0;
    for (int st = 0; st < nsteps && rc == 0; ++st) {
        if (S->kind == 0) {
            /* euleroStepper::advance (stepper.cpp:30-43) */
            rc |= or_apply(F, psi, nw, g, work);
            or_axpy(nw, psi, tr, ti, nw, n);
            or_constraint(S, nw, g);
            or_renorm(S, nw, g);
        } else {
            /* RK4Stepper::advance (stepper.cpp:46-70) */
            double hr = tr * 0.5, hi = ti * 0.5;
            rc |= or_apply(F, psi, k1, g, work);
            or_axpy(nw, psi, hr, hi, k1, n);
            or_renorm(S, nw, g);
            rc |= or_apply(F, nw, k2, g, work);
            or_axpy(nw, psi, hr, hi, k2, n);
            or_renorm(S, nw, g);
            rc |= or_apply(F, nw, k3, g, work);
            or_axpy(nw, psi, tr, ti, k3, n);
            or_renorm(S, nw, g);
            rc |= or_apply(F, nw, k4, g, work);
            /* fieldDataOld - (1./6) * timeStep()*(k1 + 2 * k2 + 2*k3 + k4) */
            double sr = (1. / 6) * tr, si = (1. / 6) * ti;
#pragma omp parallel for schedule(static)
            for (long p = 0; p < (long)n; ++p) {
                double sumr = ((k1[2 * p] + 2 * k2[2 * p]) + 2 * k3[2 * p]) + k4[2 * p];
                double sumi = ((k1[2 * p + 1] + 2 * k2[2 * p + 1]) + 2 * k3[2 * p + 1]) + k4[2 * p + 1];
                nw[2 * p] = psi[2 * p] - (sr * sumr - si * sumi);
                nw[2 * p + 1] = psi[2 * p + 1] - (sr * sumi + si * sumr);
            }
            or_constraint(S, nw, g);
            or_renorm(S, nw, g);
        }
        memcpy(psi, nw, bytes);
    }
    free(work); free(nw); free(k1); free(k2); free(k3); free(k4);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Seeded synthetic inputs: splitmix64 over (seed, linear index) -> [-1, 1). */
/* ------------------------------------------------------------------------ */

static uint64_t or_splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* u = 2 * (top 53 bits / 2^53) - 1 for stream value splitmix64(seed*2^40 + i) */
void or_uniform(uint64_t seed, uint64_t start, uint64_t count, double *out)
{
    for (uint64_t i = 0; i < count; ++i) {
        uint64_t v = or_splitmix64((seed << 40) + start + i);
        out[i] = 2.0 * ((double)(v >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Droplet labelling: gp::decompose::decompose / traverseDepthFirst          */
/* (src/cluster.cpp:17-82), driven as pyInterface::decomposition::decompose  */
/* (src/gpCpp.cpp:289-313).  labels has the field's (x,y,z,c) column-major   */
/* shape; only `component` is written, the rest stays 0.  Cells scan with k */
/* fastest and i slowest; an unvisited cell seeds a depth-first traversal   */
/* over its 26 periodic neighbours that gives cells with |psi|^2 > cutoff   */
/* the current colour and marks the others -1; the colour advances when the */
/* seed itself was coloured.  The reference recurses; this restatement uses */
/* an explicit stack (same labels: a traversal colours exactly the seed's    */
/* connected component, whatever the visiting order).  Returns the number of */
/* colours used, or -1 when out of memory.                                   */
/* ------------------------------------------------------------------------ */
static int or_pbc(int i, int L) { return i < 0 ? i + L : (i >= L ? i - L : i); }

int or_decompose(const double *psi, int nx, int ny, int nz, int nc, int component, double cutoff, int *labels)
{
    const size_t M = (size_t)nx * ny * nz;
    memset(labels, 0, sizeof(int) * M * (size_t)nc);
    int *lab = labels + (size_t)component * M;
    const double *f = psi + 2 * (size_t)component * M;
    size_t cap = 1024, top = 0;
    size_t *stack = (size_t *)malloc(cap * sizeof(size_t));
    if (!stack) return -1;
    int color = 1;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t seed = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                if (lab[seed] != 0) continue;
                top = 0;
                stack[top++] = seed;
                while (top) {
                    const size_t p = stack[--top];
                    if (lab[p] != 0) continue;
                    const double re = f[2 * p], im = f[2 * p + 1];
                    if (re * re + im * im > cutoff) {
                        lab[p] = color;
                    } else {
                        lab[p] = -1;
                        continue;
                    }
                    const int pi = (int)(p % nx), pj = (int)((p / nx) % ny), pk = (int)(p / ((size_t)nx * ny));
                    for (int a = -1; a <= 1; ++a)
                        for (int b = -1; b <= 1; ++b)
                            for (int c = -1; c <= 1; ++c) {
                                const size_t q = (size_t)or_pbc(pi + a, nx) +
                                                 (size_t)nx * ((size_t)or_pbc(pj + b, ny) + (size_t)ny * or_pbc(pk + c, nz));
                                if (lab[q] != 0) continue;
                                if (top == cap) {
                                    size_t *s2 = (size_t *)realloc(stack, 2 * cap * sizeof(size_t));
                                    if (!s2) {
                                        free(stack);
                                        return -1;
                                    }
                                    stack = s2;
                                    cap *= 2;
                                }
                                stack[top++] = q;
                            }
                }
                if (lab[seed] != -1) color += 1;
            }
    free(stack);
    return color - 1;
}
