/*
 * gp5.h -- C ABI of the MI355X (gfx950) Gross-Pitaevskii pseudo-spectral core.
 *
 * Every entry point is a plain C function over opaque handles, plain pointers
 * and sizes.  Host buffers use the reference's memory layout
 * (src/traits.h:12-20): complex128 interleaved (re, im), column-major
 * (x fastest, then y, z, component), so an Eigen::Tensor<complex<double>,4>
 * data() pointer or a Fortran-ordered numpy complex128 array passes through
 * unchanged.  Device state stays resident between calls; only scalars come back
 * per step.
 *
 * All functions return GP5_OK (0) or a GP5_ERR_* code; gp5_last_error() holds
 * a thread-local message for the last failure.  Calls are synchronous unless
 * stated otherwise.  One HIP stream per context; contexts are not thread-safe.
 */
#ifndef GP5_H
#define GP5_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP5_OK 0
#define GP5_ERR_ARG 1          /* bad argument / shape mismatch / unknown name */
#define GP5_ERR_HIP 2          /* HIP runtime error (message carries hipGetErrorString) */
#define GP5_ERR_UNSUPPORTED 3  /* e.g. axis length beyond the FFT's LDS tile */
#define GP5_ERR_ALLOC 4        /* device allocation failed */
#define GP5_ERR_NONFINITE 5    /* NaN/Inf detected by a fused reduction */

/* FFT_DIRECTION (reference src/fourierTransform.h:12) */
#define GP5_FORWARD 0
#define GP5_BACKWARD 1

/* functional kinds (reference src/functional.h:51-113) */
#define GP5_FUNC_GP 0            /* gpFunctional */
#define GP5_FUNC_LHY 1           /* LHYDropletFunctional */
#define GP5_FUNC_LHY_UNLOCKED 2  /* LHYDropletUnlockedFunctional */

/* stepper kinds (reference src/stepper.h:70-97) */
#define GP5_STEP_EULERO 0
#define GP5_STEP_RK4 1

typedef struct gp5_ctx gp5_ctx; /* grid + FFT plans + stream + scratch */
typedef struct gp5_buf gp5_buf; /* device array of doubles */

/* Parameters of a functional; plain data, copied at each call.
 * Reference: functional / gpFunctional / LHYDropletFunctional /
 * LHYDropletUnlockedFunctional state (src/functional.h:9-113). */
typedef struct gp5_functional_desc {
    int kind;                 /* GP5_FUNC_* */
    int has_couplings;        /* gpFunctional::_setCouplings */
    double couplings[4];      /* couplings(i,j) at [i + 2*j] */
    double masses[2];         /* gpFunctional::_masses (default 1) */
    double lhy_a, lhy_b;      /* LHY: a*rho + b*rho^{3/2}; reference values -3, 2.5 */
    double alpha, beta, eta;  /* UnlockedLHYDroplet parameters */
    const gp5_buf *potential; /* real FP64 V, M*nC doubles, or NULL (addPotential) */
} gp5_functional_desc;

/* Parameters of a stepper (reference src/stepper.h:8-67). */
typedef struct gp5_stepper_desc {
    int kind;                   /* GP5_STEP_* */
    double tau_re, tau_im;      /* complex time step: (dt,0) imaginary, (0,dt) real */
    int renormalize;            /* stepper::_reNormalize */
    double normalizations[2];   /* stepper::_normalizations */
    const gp5_buf *phase;       /* phaseConstraint phase (real, M*nC) or NULL */
} gp5_stepper_desc;

const char *gp5_last_error(void);
int gp5_version(void);

/* Grid, plans and stream.  Replaces fourierTransformCreator::create +
 * createUniformDiscretization + the fftwFourierTransform constructor
 * (src/fourierTransform.cpp:228-249, :129-143; src/geometry.cpp:56-105).
 * One GPU, so the processor grid is {1,1,1}. */
int gp5_ctx_create(gp5_ctx **out, const int shape[3], const double left[3],
                   const double right[3], int n_components);
int gp5_ctx_destroy(gp5_ctx *ctx);
int gp5_ctx_sync(gp5_ctx *ctx);
/* Points per component (Nx*Ny*Nz). */
size_t gp5_ctx_points(const gp5_ctx *ctx);

/* Register plans of the FFT passes (stage order, values per thread, columns
 * per block), per role: 0 x lines (with the fused epilogue), 1 y columns,
 * 2 the fused z pass.  The reference leaves its plan choice to FFTW_MEASURE
 * (src/fourierTransform.cpp:166-171); here the autotuner (gp5_batch_tune)
 * picks a variant per role, and callers may set one.  count[r] = variants of
 * role r (0: the axis runs on the generic kernels); variant 0 is the default.
 * Every variant computes the same DFT. */
int gp5_ctx_fft_plans(const gp5_ctx *ctx, int variant[3], int count[3]);
int gp5_ctx_set_fft_plans(gp5_ctx *ctx, const int variant[3]);
const char *gp5_ctx_fft_plan_name(const gp5_ctx *ctx, int role); /* e.g. "4x12x4/12" */

/* Device buffers (the Eigen tensors of the reference live here). */
int gp5_buf_alloc(gp5_ctx *ctx, size_t n_doubles, gp5_buf **out);
int gp5_buf_free(gp5_buf *buf);
size_t gp5_buf_size(const gp5_buf *buf);
int gp5_buf_upload(gp5_buf *buf, const double *host, size_t n_doubles);
int gp5_buf_download(const gp5_buf *buf, double *host, size_t n_doubles);
int gp5_buf_copy(gp5_buf *dst, const gp5_buf *src);
int gp5_buf_zero(gp5_buf *buf);

/* Unnormalised 3-D c2c DFT of every component, out-of-place with src
 * preserved (src == dst allowed).  Replaces fftwFourierTransform::apply
 * (src/fourierTransform.cpp:184-214). */
int gp5_fft(gp5_ctx *ctx, const gp5_buf *src, gp5_buf *dst, int direction);

/* dst = IFFT(-K^2 FFT(src)) / M.  Replaces operators::laplacian::apply
 * (src/operators.cpp:28-36). */
int gp5_laplacian(gp5_ctx *ctx, const gp5_buf *src, gp5_buf *dst);

/* dst = H(src).  Replaces functional::apply for gpFunctional
 * (src/functional.cpp:13-43,54-102), LHYDropletFunctional (:105-122) and
 * LHYDropletUnlockedFunctional (:125-150), including addPotential (:46-53).
 * `time` is accepted and unused, as in the reference. */
int gp5_functional_apply(gp5_ctx *ctx, const gp5_functional_desc *f, const gp5_buf *src,
                         gp5_buf *dst, double time);

/* norm_c = sum |psi_c|^2 dV.  Replaces gp::norm (src/geometry.cpp:190-214). */
int gp5_norm(gp5_ctx *ctx, const gp5_buf *f, int component, double *out);

/* psi_c *= sqrt(N / norm_c).  Replaces gp::normalize (src/geometry.cpp:217-229). */
int gp5_normalize(gp5_ctx *ctx, gp5_buf *f, int component, double N);

/* max_x |psi_c(x)|^2.  Replaces the local part of gp.maxDensity
 * (python/gp/gp.py:210-222). */
int gp5_max_density(gp5_ctx *ctx, const gp5_buf *f, int component, double *out);

/* E[psi] = sum_c int psi_c^* (-1/2 m_c^-1 lap) psi_c + V rho + e(rho).  No
 * reference counterpart (the reference computes no energy); GP5_ERR_UNSUPPORTED
 * for the unlocked functional, which is not a gradient for general alpha, beta, eta. */
int gp5_energy(gp5_ctx *ctx, const gp5_functional_desc *f, const gp5_buf *psi, double *out);

/* psi <- |psi| e^{i phase}.  Replaces phaseConstraint::apply (src/constraint.cpp:6-9). */
int gp5_phase_constraint(gp5_ctx *ctx, gp5_buf *psi, const gp5_buf *phase);

/* Stage algebra as separate passes, for steppers that call a functional's own
 * apply per stage (a functional or constraint the fused path does not know):
 * out = a - (cr + i ci) * k replaces the Eigen expression
 * fieldDataOld - timeStep()*f*k (src/stepper.cpp:35,50,55,60), and
 * out = a - (cr + i ci) * (((k1 + 2 k2) + 2 k3) + k4) replaces
 * fieldDataOld - (1./6)*timeStep()*(k1 + 2*k2 + 2*k3 + k4) (src/stepper.cpp:64).
 * out may alias any input. */
int gp5_axpy(gp5_ctx *ctx, gp5_buf *out, const gp5_buf *a, double cr, double ci, const gp5_buf *k);
int gp5_rk4_combine(gp5_ctx *ctx, gp5_buf *out, const gp5_buf *a, double cr, double ci, const gp5_buf *k1,
                    const gp5_buf *k2, const gp5_buf *k3, const gp5_buf *k4);

/* Advance psi in place by nsteps of euleroStepper / RK4Stepper
 * (src/stepper.cpp:30-81) as driven by pyInterface::timeStepper::advance
 * (src/gpCpp.cpp:255-269).  Optional per-step logs (host arrays, may be NULL):
 * norm_log[s*nC + c] = norm_c and energy_log[s] = E of the state at the start
 * of step s (computed inside the fused stage-1 epilogue at no extra pass). */
int gp5_advance(gp5_ctx *ctx, const gp5_functional_desc *f, const gp5_stepper_desc *s,
                gp5_buf *psi, int nsteps, double *norm_log, double *energy_log);

/* gp5_advance (no logs) that allocates nothing: no hipMalloc, hipFree or other
 * allocating call, on the device or the host.  No counterpart in the reference;
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  The stage fields and the
 * reduction partials live in one static device array of 200 MiB per GPU, held
 * until the process ends: RK4 needs 5 fields of 16*M*nC bytes (6 with a phase
 * constraint), Euler 2 (3), plus 48 B per x-pass block, e.g. 160 MiB for one
 * 128^3 component; larger inputs return GP5_ERR_UNSUPPORTED.  The context must
 * exist (gp5_ctx_create allocates, once).  Results equal gp5_advance's.  Calls
 * that overlap on one GPU share the array and race; calls in sequence are safe
 * (nothing in the array needs to start at zero). */
int gp5_advance_noalloc(gp5_ctx *ctx, const gp5_functional_desc *f, const gp5_stepper_desc *s,
                        gp5_buf *psi, int nsteps);

/* Time the hot loop: run nsteps of gp5_advance with no logs and report the
 * elapsed device time between HIP events recorded on the context's stream,
 * in milliseconds. */
int gp5_advance_timed(gp5_ctx *ctx, const gp5_functional_desc *f, const gp5_stepper_desc *s,
                      gp5_buf *psi, int nsteps, double *elapsed_ms);

/* Per-kernel device time of the hot loop (the role of the reference's
 * START_TIMER/STOP_TIMER labels, src/timers.h:69-113, at kernel granularity):
 * runs nsteps exactly as gp5_advance does, with a HIP event pair on the
 * context's stream around every kernel launch, and accumulates per kernel
 * kind.  kind_ms[k] / kind_launches[k] is the average launch duration of kind k
 * (k < GP5_NKINDS, named by gp5_kernel_kind_name). */
#define GP5_NKINDS 8
int gp5_profile_advance(gp5_ctx *ctx, const gp5_functional_desc *f, const gp5_stepper_desc *s,
                        gp5_buf *psi, int nsteps, double *kind_ms, int *kind_launches);
const char *gp5_kernel_kind_name(int kind);

/* Droplet labels: the connected components of {|psi_c|^2 > cutoff} over the 26
 * periodic neighbours of each cell, numbered 1, 2, ... in the order in which
 * the reference's scan (k fastest, i slowest) first meets them; cells at or
 * below the cut-off get -1.  labels: host array of M * nC ints in the field's
 * column-major (x, y, z, c) layout; only component `comp` is labelled, the
 * others are set to 0.  *ndroplets (may be NULL) = number of droplets.
 * Replaces gp::decompose::decompose (src/cluster.cpp:17-82) as driven by
 * pyInterface::decomposition::decompose (src/gpCpp.cpp:289-313). */
int gp5_decompose(gp5_ctx *ctx, const gp5_buf *psi, int comp, double cutoff, int *labels, int *ndroplets);

/* dst = src resampled to dst's grid: the trigonometric interpolant of src
 * (cell-centred samples, Nyquist modes as cosines) evaluated at dst's cell
 * centres; zero-padding of the spectrum when dst is finer, aliases folded when
 * it is coarser.  Both contexts cover the same box and hold the same number of
 * components.  No counterpart in the reference: it carries a state converged
 * on a coarse grid to a finer one (multilevel ground-state starts). */
int gp5_resample(gp5_ctx *src_ctx, const gp5_buf *src, gp5_ctx *dst_ctx, gp5_buf *dst);

/* ---- field files (host side) ----------------------------------------------
 * HDF5 files in the reference's format: dataset "field" of H5T_ARRAY double[2]
 * elements with dims (nC, Nz, Ny, Nx), holding the column-major complex128
 * field (2 * M * nC doubles, x fastest, components last).  Replace gp::save /
 * gp::load (src/io.cpp:56-152, 154-254), serially written (one GPU, one
 * writer).  Files with a plain double dataset of dims (nC, Nz, Ny, Nx, 2) are
 * also read.  libhdf5 is loaded at run time; GP5_ERR_UNSUPPORTED if it cannot
 * be.  No device work. */
int gp5_h5_write(const char *filename, const double *host, const int shape[3], int nc);
int gp5_h5_read(const char *filename, double *host, const int shape[3], int nc);
int gp5_h5_shape(const char *filename, int shape[3], int *nc); /* grid and component count of a file */

/* ---- batches of independent instances ------------------------------------
 * No counterpart in the reference, which evolves one field per process: these
 * run many single-component instances of uneven cubic grids at once, each
 * with the reference's imaginary-time RK4 and per-stage renormalisation
 * (RK4Stepper::advance, src/stepper.cpp:46-70, set up as src/main.cpp:131-145
 * does), for callers that need many ground states (parameter scans, training
 * data).  Instance i has grid n[i]^3 on the box [-L[i]/2, L[i]/2]^3 and
 * normalisation N[i]; instances of one (n, L) share one context as its
 * components.  Host arrays are one instance's column-major complex128 field
 * (2 n^3 doubles) or real potential (n^3 doubles). */
typedef struct gp5_batch gp5_batch;
int gp5_batch_create(gp5_batch **out, int count, const int *n, const double *L, const double *N);
int gp5_batch_destroy(gp5_batch *b);
int gp5_batch_classes(const gp5_batch *b);    /* number of (n, L) classes */
int gp5_batch_class_of(const gp5_batch *b, int i);
int gp5_batch_upload(gp5_batch *b, int i, const double *host);
int gp5_batch_download(const gp5_batch *b, int i, double *host);
/* LHYDropletFunctional coefficients a, b of instance i (default -3, 2.5,
 * src/functional.cpp:117-118) and its external potential (addPotential,
 * src/functional.cpp:46-53); both apply to the active instances only. */
int gp5_batch_set_lhy(gp5_batch *b, int i, double a, double bcoef);
int gp5_batch_set_potential(gp5_batch *b, int i, const double *V);
/* nsteps of imaginary-time RK4 with tau = dt_factor * dx^2 and per-stage
 * renormalisation to N[i], on every active instance. */
int gp5_batch_advance(gp5_batch *b, const gp5_functional_desc *f, double dt_factor, int nsteps);
/* Autotuner: per class, times `steps` steps of each candidate radix plan and
 * tile (lines per block) of every generic-kernel axis and the fused x pass,
 * keeps the fastest; the instance states are restored.  A text summary goes
 * to report (may be NULL). */
int gp5_batch_tune(gp5_batch *b, const gp5_functional_desc *f, double dt_factor, int steps, char *report,
                   size_t report_len);
/* Ground states: hipGraph replays of check_every (even) steps per class until
 * each instance's energy changes by at most tol relative between two
 * consecutive checks; converged instances retire and their class shrinks.
 * steps_out[i]: steps taken (negative: max_steps reached unconverged);
 * energy_out[i]: energy at the last check; elapsed_ms: wall time. */
int gp5_batch_ground_state(gp5_batch *b, const gp5_functional_desc *f, double dt_factor, double tol,
                           int check_every, long long max_steps, long long *steps_out, double *energy_out,
                           double *elapsed_ms);
/* Instance i of dst = instance j of src (its current or final field) resampled
 * to dst's grid, as gp5_resample; the two instances share the box length.
 * Starts a fine-grid batch from a coarse-grid one. */
int gp5_batch_resample(gp5_batch *dst, int i, const gp5_batch *src, int j);

#ifdef __cplusplus
}
#endif

#endif /* GP5_H */
