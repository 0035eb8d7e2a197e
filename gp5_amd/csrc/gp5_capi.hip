// C ABI of the gfx950 GP core (declared in include/gp5.h).
//
// A context owns the grid, the per-axis Stockham plans and twiddle tables,
// one HIP stream and the scratch fields of the fused passes.  Every spectral
// operation is five axis passes over a column-major complex128 field:
//   x forward, y forward, z forward * k-space multiplier * z inverse,
//   y inverse, x inverse + epilogue
// where the epilogue is a store (laplacian), the functional H(phi), or one
// Euler/RK4 stage update with fused norm/energy reductions.
#include "../../include/gp5.h"
#include "gp5_internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <limits.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <thread>
#include <cmath>
#include <string>
#include <vector>

namespace gp5 {
hipError_t launch_axis(hipStream_t st, int mode, const double2 *src, double2 *dst, const AxisPlan &P,
                       const AxisGeom &G, int sign, const double *scale_in, const LoadPre &pre, long long Opc,
                       const MidArgs &mid, const EpiArgs &epi);
unsigned axis_grid(const AxisGeom &G);
hipError_t launch_norm(hipStream_t st, const double2 *f, long long len, unsigned blocks, const EpiArgs &E);
hipError_t launch_scale(hipStream_t st, double2 *f, long long M, int nc, int comp, double s, const double *sp,
                        const int *active);
hipError_t launch_maxdens(hipStream_t st, const double2 *f, long long len, unsigned blocks, double *blockmax);
hipError_t launch_phase(hipStream_t st, double2 *f, const double *phase, long long len);
hipError_t launch_phase_exp(hipStream_t st, double2 *ex, const double *phase, long long len);
hipError_t launch_axpy(hipStream_t st, double2 *out, const double2 *a, double cr, double ci, const double2 *k,
                       long long len);
hipError_t launch_rk4_combine(hipStream_t st, double2 *out, const double2 *a, double cr, double ci, const double2 *k1,
                              const double2 *k2, const double2 *k3, const double2 *k4, long long len);
hipError_t launch_finalize(hipStream_t st, const EpiArgs &E, unsigned nblocks, long long lpb, long long lpc, int ncomp);
}  // namespace gp5

using namespace gp5;

struct gp5_buf {
    gp5_ctx *ctx;
    double *d;
    size_t n;  // doubles
};

struct gp5_ctx {
    int n[3];
    double left[3], right[3];
    int nc;              // components of the field (active components of a batch class)
    int nc_cap;          // components the scratch and the scale slots are sized for
    bool uncoupled;      // components are independent instances (more than two, or a batch class)
    long long M;
    double dV;
    double dk[3];
    int device;
    hipStream_t st;
    bool fast[3];        // axis runs on the register-resident kernels (gp5_fast.hip)
    AxisPlan plan[3];
    AxisGeom geom[3];
    long long Opc[3];
    double2 *tw[3];
    // scratch fields (allocated on first use)
    double2 *T;          // x/y/z transforms of the current stage input
    double2 *kA, *kB;    // RK4: K = k_s, alternating between the two
    double2 *sbuf;       // RK4: S = k1 + 2 k2
    double2 *eout;       // new state, swapped with the caller's buffer after each step
    double2 *phase_ex;   // e^{i phase} of the stepper's phase constraint
    bool fuse_x;         // epilogue pass also forward-transforms its output along x
    int mall_chunks;     // x slabs of the y/z passes (1 = whole field per pass)
    // reduction scratch
    double *partials;
    size_t partials_cap;  // blocks
    double *results;      // [kRedN][nc]
    double *scales;       // 8 slots x nc_cap components (slot k at scales + k * nc_cap)
    double *nnorm;        // nc_cap normalisation targets of a batch (device), or null
    double *lhy_ab;       // nc_cap (a, b) LHY coefficient pairs of a batch (device), or null
    double *Vdev;         // M * nc_cap potential of a batch (device), or null
    double *blockmax;
    double *dlog;
    size_t dlog_cap;      // steps
    hipEvent_t ev0, ev1;
    struct ProfState *prof;  // non-null while gp5_profile_advance runs
    int *ccl;             // droplet-labelling workspace, labels and count (allocated on first use)
    size_t ccl_cap;       // ints
    // batch classes: live components [0, active) on the device (the kernels gate on
    // it; owned by the batch) and its host copy, valid between ground-state calls
    const int *active_dev;
    int active;
    // register-plan variants of the fast passes (0: default): x lines, y columns, z pass
    int pvar[3];
    int xcd_remap;  // XCD-aware tile order of the fast passes (bit 0 columns, bit 1 x lines)
};

// Event pairs around every launch while profiling.
enum KernelKind {
    K_X_FWD = 0,      // x forward pass (first stage of an advance, or unfused stages)
    K_Y = 1,          // y pass, either direction
    K_Z_MID = 2,      // z forward * multiplier * z inverse
    K_X_EPI = 3,      // x inverse + stage epilogue
    K_FFT = 4,        // passes of the standalone transform
    K_NORM = 5,       // norm partials / finalize
    K_SCALE = 6,      // materialised lazy scale
    K_X_EPI_FWD = 7,  // x inverse + stage epilogue + x forward of the new stage vector
};
static const char *kKindNames[GP5_NKINDS] = {"fft_x_forward", "fft_y_pass", "fft_z_fwd_mul_inv", "fft_x_inverse_epilogue",
                                              "fft_plain", "norm_reduce", "scale", "fft_x_inv_epilogue_fwd"};
struct ProfState {
    std::vector<hipEvent_t> ev;  // pairs
    std::vector<int> kind;
    size_t used = 0;
};

static hipError_t prof_begin(gp5_ctx *c, int kind)
{
    ProfState *p = c->prof;
    if (!p) return hipSuccess;
    if (2 * (p->used + 1) > p->ev.size()) {
        for (int i = 0; i < 2; ++i) {
            hipEvent_t e;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            p->ev.push_back(e);
        }
    }
    if (p->kind.size() < p->used + 1) p->kind.push_back(kind);
    else p->kind[p->used] = kind;
    return hipEventRecord(p->ev[2 * p->used], c->st);
}

static hipError_t prof_end(gp5_ctx *c)
{
    ProfState *p = c->prof;
    if (!p) return hipSuccess;
    hipError_t r = hipEventRecord(p->ev[2 * p->used + 1], c->st);
    p->used++;
    return r;
}

#define PLAUNCH(c, kind, call)                   \
    do {                                         \
        HIPCHK(prof_begin((c), (kind)));         \
        HIPCHK(call);                            \
        HIPCHK(prof_end(c));                     \
    } while (0)

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// for the other translation units of the library (gp5_nn.hip)
namespace gp5 {
int set_error(int code, const char *msg)
{
    g_err = msg;
    return code;
}
}  // namespace gp5

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(GP5_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define RCHK(expr)               \
    do {                         \
        int r_ = (expr);         \
        if (r_ != GP5_OK) return r_; \
    } while (0)

// ---------------------------------------------------------------------------
// static scratch of gp5_advance_noalloc
//
// One file-scope device array, held by every GPU the library runs on until the
// process ends.  In host code its name is only a stand-in for the symbol: its
// device address (one per GPU) comes from hipGetSymbolAddress, taken once per
// device at context creation with that device current, and only that address is
// ever handed to kernels and copies.  Calls of gp5_advance_noalloc that overlap
// on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)200 << 20;  // 200 MiB
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];
static std::mutex g_scratch_mtx;
static double2 *g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device, or null

// Called by gp5_ctx_create with `device` current.
static int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(GP5_ERR_UNSUPPORTED, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mtx);
    if (g_scratch_ptr[device]) return GP5_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(GP5_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = (double2 *)ptr;
    return GP5_OK;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
// Radix plans of the generic LDS kernels.  order 0: 8s, 4s, 2, 3s, 5s, other
// primes; order 1: 4s and 2 before 3s and 5s (no radix 8); order 2: 3s and 5s
// first, then 8s, 4s, 2; order 3: 2s only for the power-of-two part.  The
// autotuner picks per axis; the default is order 0.
constexpr int kPlanOrders = 4;
static int make_plan(int n, AxisPlan &P, int order = 0)
{
    P.n = n;
    P.nrad = 0;
    int m = n;
    auto push = [&](int r) -> bool {
        if (P.nrad >= kMaxRadices) return false;
        P.rad[P.nrad++] = r;
        return true;
    };
    auto pow2 = [&]() -> bool {
        if (order == 0 || order == 2) {
            while (m % 8 == 0) { if (!push(8)) return false; m /= 8; }
        }
        if (order != 3) {
            while (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
        }
        while (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
        return true;
    };
    auto small_odd = [&]() -> bool {
        while (m % 3 == 0) { if (!push(3)) return false; m /= 3; }
        while (m % 5 == 0) { if (!push(5)) return false; m /= 5; }
        return true;
    };
    if (order == 2) {
        if (!small_odd() || !pow2()) return -1;
    } else {
        if (!pow2() || !small_odd()) return -1;
    }
    for (int r = 7; m > 1; r += 2) {
        while (m % r == 0) { if (!push(r)) return -1; m /= r; }
        if ((long long)r * r > m && m > 1) { if (!push(m)) return -1; m = 1; }
    }
    return 0;
}

static std::vector<double2> twiddles(int n)
{
    // e^{-2 pi i t / n} from the exact integer phase, evaluated in long double
    const long double two_pi = 6.283185307179586476925286766559005768L;
    std::vector<double2> tw(n);
    for (int t = 0; t < n; ++t) {
        long double a = -two_pi * (long double)t / (long double)n;
        tw[t] = make_double2((double)cosl(a), (double)sinl(a));
    }
    return tw;
}

// Axis geometry of the passes for the context's current component count.
// Generic-kernel tiles (lines per block) default to kTileElems / n; with more
// than two components a row tile never spans more than one component
// boundary (the partial sums carry two component slots).  A nonzero geom[d].B
// (set by the autotuner) is kept, clipped to the same limits.
static void set_geometry(gp5_ctx *c)
{
    long long I = 1;
    for (int d = 0; d < 3; ++d) {
        AxisGeom &G = c->geom[d];
        const int n = c->n[d];
        G.I = I;
        c->Opc[d] = c->M / (I * n);
        G.O = c->Opc[d] * c->nc;
        int B = G.B > 0 ? G.B : std::max(1, kTileElems / n);
        B = std::min(B, std::max(1, kTileElems / n));
        if (I == 1) {
            B = (int)std::min<long long>(B, G.O);
            if (c->nc_cap > 2) B = (int)std::min<long long>(B, c->Opc[d]);
        } else {
            B = (int)std::min<long long>(std::min(B, 64), I);
        }
        G.B = std::max(1, B);
        G.active = c->active_dev;
        I *= n;
    }
}

// Lazy-scale slot k (per component): 0/1 psi, 2/3 RK stage inputs, 7 normalize.
static double *scale_slot(gp5_ctx *c, int k) { return c->scales + (size_t)k * c->nc_cap; }

static int ensure_partials(gp5_ctx *c, size_t blocks)
{
    if (blocks <= c->partials_cap) return GP5_OK;
    if (c->partials) HIPCHK(hipFree(c->partials));
    c->partials = nullptr;
    size_t cap = std::max<size_t>(blocks, 4096);
    HIPCHK(hipMalloc(&c->partials, cap * kRedQ * sizeof(double)));
    c->partials_cap = cap;
    return GP5_OK;
}

// Scratch fields: T always; a stepper adds the output field, RK4 also K (x2) and S.
static int ensure_scratch(gp5_ctx *c, bool rk, bool stepper)
{
    const size_t bytes = (size_t)c->M * c->nc_cap * sizeof(double2);
    auto need = [&](double2 *&p) -> int {
        if (p) return GP5_OK;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return fail(GP5_ERR_ALLOC, "cannot allocate %zu B of scratch", bytes);
        }
        return GP5_OK;
    };
    RCHK(need(c->T));
    if (rk) {
        RCHK(need(c->kA));
        RCHK(need(c->kB));
        RCHK(need(c->sbuf));
    }
    if (stepper) RCHK(need(c->eout));
    return GP5_OK;
}

static int check_field(const gp5_ctx *c, const gp5_buf *b, const char *what)
{
    if (!b) return fail(GP5_ERR_ARG, "%s is NULL", what);
    if (b->ctx != c) return fail(GP5_ERR_ARG, "%s belongs to another context", what);
    if (b->n != (size_t)c->M * c->nc * 2)
        return fail(GP5_ERR_ARG, "%s holds %zu doubles, the field needs %lld", what, b->n, c->M * c->nc * 2);
    return GP5_OK;
}

static int check_real(const gp5_ctx *c, const gp5_buf *b, const char *what)
{
    if (b->ctx != c) return fail(GP5_ERR_ARG, "%s belongs to another context", what);
    if (b->n != (size_t)c->M * c->nc)
        return fail(GP5_ERR_ARG, "%s holds %zu doubles, a real field needs %lld", what, b->n, c->M * c->nc);
    return GP5_OK;
}

static EpiArgs base_epi(gp5_ctx *c)
{
    EpiArgs E;
    memset(&E, 0, sizeof E);
    E.M = c->M;
    E.nc = c->nc;
    E.nx = c->n[0];
    E.dV = c->dV;
    E.partials = c->partials;
    E.results = c->results;
    E.active = c->active_dev;
    return E;
}

static int fill_functional(gp5_ctx *c, const gp5_functional_desc *f, EpiArgs &E, double coef[2])
{
    if (!f) return fail(GP5_ERR_ARG, "functional descriptor is NULL");
    E.fkind = f->kind;
    E.uncoupled = c->uncoupled;
    if (f->kind == GP5_FUNC_GP) {
        // more than two components: uncoupled instances, each with g = couplings[0]
        E.has_couplings = f->has_couplings;
        E.g11 = f->couplings[0];
        E.g12 = f->couplings[2];
        E.g22 = f->couplings[3];
        // more than two (uncoupled) components share masses[0]
        for (int k = 0; k < 2; ++k) {
            double m = c->uncoupled ? f->masses[0] : (k < c->nc ? f->masses[k] : 1.0);
            if (m == 0.0) m = 1.0;
            coef[k] = 0.5 * (1.0 / m) / (double)c->M;
        }
    } else if (f->kind == GP5_FUNC_LHY) {
        E.lhy_a = f->lhy_a;
        E.lhy_b = f->lhy_b;
        E.lhy_ab = c->lhy_ab;  // batch classes: per-instance coefficients
        coef[0] = coef[1] = 0.5 / (double)c->M;
    } else if (f->kind == GP5_FUNC_LHY_UNLOCKED) {
        if (c->nc != 2 || c->uncoupled) return fail(GP5_ERR_ARG, "Incompatible number of components");
        E.alpha = f->alpha;
        E.beta = f->beta;
        E.eta = f->eta;
        coef[0] = coef[1] = 0.5 / (double)c->M;
    } else {
        return fail(GP5_ERR_ARG, " Unkown functional");
    }
    if (f->potential) {
        RCHK(check_real(c, f->potential, "potential"));
        E.V = f->potential->d;
    } else if (c->Vdev) {
        E.V = c->Vdev;  // batch classes: per-instance potentials
    }
    return GP5_OK;
}

static MidArgs make_mid(gp5_ctx *c, const double coef[2], const double *scale)
{
    MidArgs m;
    m.nx = c->n[0];
    m.ny = c->n[1];
    m.nz = c->n[2];
    for (int d = 0; d < 3; ++d) m.dk[d] = c->dk[d];
    m.coef[0] = coef[0];
    m.coef[1] = coef[1];
    m.M = c->M;
    m.scale = scale;
    return m;
}

static FastArgs fast_args(gp5_ctx *c, int axis, const double2 *src, double2 *dst)
{
    FastArgs a;
    memset(&a, 0, sizeof a);
    a.src = src;
    a.dst = dst;
    a.tw = c->tw[axis];
    a.I = c->geom[axis].I;
    a.O = c->geom[axis].O;
    a.Opc = c->Opc[axis];
    a.nx = c->n[0];
    a.active = c->active_dev;
    a.xcd = c->xcd_remap;
    return a;
}

static int pass_plain(gp5_ctx *c, int kind, int axis, const double2 *src, double2 *dst, int sign,
                      const double *scale_in, const LoadPre *pre);

// y forward, z forward * multiplier * z inverse, y inverse on T.  With
// register kernels on y and z the three passes run slab by slab over x
// (mall_chunks slabs): a slab written by one pass is read back by the next
// while it is still resident in the memory-side Infinity Cache, so HBM sees
// each slab read once and written once instead of three times each.
static int middle_passes(gp5_ctx *c, const MidArgs &mid)
{
    const int nx = c->n[0];
    const int nch = (c->fast[1] && c->fast[2]) ? c->mall_chunks : 1;
    if (nch > 1) {
        int step = (nx + nch - 1) / nch;
        step = (step + 15) / 16 * 16;
        for (int x0 = 0; x0 < nx; x0 += step) {
            const int xc = std::min(step, nx - x0);
            FastArgs ay = fast_args(c, 1, c->T, c->T);
            ay.x0 = x0;
            ay.xcount = xc;
            PLAUNCH(c, K_Y, launch_fast(c->st, c->n[1], 0, -1, ay, c->pvar[1]));
            FastArgs az = fast_args(c, 2, c->T, c->T);
            az.x0 = x0;
            az.xcount = xc;
            az.mid = mid;
            PLAUNCH(c, K_Z_MID, launch_fast(c->st, c->n[2], 1, -1, az, c->pvar[2]));
            PLAUNCH(c, K_Y, launch_fast(c->st, c->n[1], 0, +1, ay, c->pvar[1]));
        }
        return GP5_OK;
    }
    RCHK(pass_plain(c, K_Y, 1, c->T, c->T, -1, nullptr, nullptr));
    if (c->fast[2]) {
        FastArgs a = fast_args(c, 2, c->T, c->T);
        a.mid = mid;
        PLAUNCH(c, K_Z_MID, launch_fast(c->st, c->n[2], 1, -1, a, c->pvar[2]));
    } else {
        EpiArgs none;
        memset(&none, 0, sizeof none);
        PLAUNCH(c, K_Z_MID, launch_axis(c->st, 1, c->T, c->T, c->plan[2], c->geom[2], -1, nullptr,
                                        LoadPre{nullptr, 0.0, 0.0}, c->Opc[2], mid, none));
    }
    RCHK(pass_plain(c, K_Y, 1, c->T, c->T, +1, nullptr, nullptr));
    return GP5_OK;
}

// One FFT pass along `axis` (sign -1 forward, +1 inverse), src -> dst, loading
// src * scale_in - cf * src2 when `pre` carries src2.
static int pass_plain(gp5_ctx *c, int kind, int axis, const double2 *src, double2 *dst, int sign,
                      const double *scale_in, const LoadPre *pre)
{
    LoadPre none_pre = {nullptr, 0.0, 0.0};
    if (!pre) pre = &none_pre;
    if (c->fast[axis]) {
        FastArgs a = fast_args(c, axis, src, dst);
        a.scale_in = scale_in;
        a.src2 = pre->src2;
        a.cfr = pre->cfr;
        a.cfi = pre->cfi;
        PLAUNCH(c, kind, launch_fast(c->st, c->n[axis], 0, sign, a, axis == 0 ? c->pvar[0] : c->pvar[1]));
    } else {
        MidArgs mid;
        memset(&mid, 0, sizeof mid);
        EpiArgs none;
        memset(&none, 0, sizeof none);
        PLAUNCH(c, kind, launch_axis(c->st, 0, src, dst, c->plan[axis], c->geom[axis], sign, scale_in, *pre,
                                     c->Opc[axis], mid, none));
    }
    return GP5_OK;
}

// Input of one H application: what an unfused first x pass loads
// (src * src_scale - cf * src2) and the lazy scale the z pass applies.
struct StageIn {
    const double2 *src;
    const double *src_scale;
    LoadPre pre;
    const double *mid_scale;
};

static StageIn plain_input(const double2 *src)
{
    StageIn in;
    in.src = src;
    in.src_scale = nullptr;
    in.pre = {nullptr, 0.0, 0.0};
    in.mid_scale = nullptr;
    return in;
}

// Passes of one H application: x forward, y, z (forward * multiplier *
// inverse), y inverse, x inverse + epilogue.  Fused mode (x axis on the
// register kernels): T already holds the raw x transform of the stage input,
// and the epilogue pass leaves the x transform of its own output in T for the
// next stage, so the x forward pass is skipped.  Either way T carries the raw
// transform and the z pass applies in.mid_scale.
static int stage_passes(gp5_ctx *c, const StageIn &in, const double coef[2], const EpiArgs &E, bool fused)
{
    MidArgs mid = make_mid(c, coef, in.mid_scale);
    if (!fused) RCHK(pass_plain(c, K_X_FWD, 0, in.src, c->T, -1, in.src_scale, &in.pre));
    RCHK(middle_passes(c, mid));
    unsigned nblocks;
    long long lpb;  // lines per block of the epilogue pass (component slots of its partials)
    EpiArgs e = E;
    if (c->fast[0]) {
        nblocks = fast_grid(c->n[0], c->geom[0].I, c->geom[0].O, c->pvar[0]);
        lpb = fast_rows_per_block(c->n[0], c->pvar[0]);
        RCHK(ensure_partials(c, nblocks));
        FastArgs a = fast_args(c, 0, c->T, nullptr);
        e.partials = c->partials;
        a.epi = e;
        a.t_out = fused ? c->T : nullptr;
        // component-local epilogue (fewer registers) unless components couple or a constraint applies
        const bool coupled = e.fkind == GP5_FUNC_LHY_UNLOCKED ||
                             (e.fkind == GP5_FUNC_GP && e.has_couplings && e.nc == 2 && !e.uncoupled);
        const int mode = (coupled || e.phase_exp) ? 2 : 3;
        PLAUNCH(c, fused ? K_X_EPI_FWD : K_X_EPI, launch_fast(c->st, c->n[0], mode, +1, a, c->pvar[0]));
    } else {
        if (fused) return fail(GP5_ERR_UNSUPPORTED, "fused stages need a register-kernel x length");
        nblocks = axis_grid(c->geom[0]);
        lpb = c->geom[0].B;
        RCHK(ensure_partials(c, nblocks));
        e.partials = c->partials;
        PLAUNCH(c, K_X_EPI, launch_axis(c->st, 2, c->T, nullptr, c->plan[0], c->geom[0], +1, nullptr,
                                        LoadPre{nullptr, 0.0, 0.0}, c->Opc[0], mid, e));
    }
    if (e.want_red) PLAUNCH(c, K_NORM, launch_finalize(c->st, e, nblocks, lpb, c->Opc[0], c->nc));
    return GP5_OK;
}

static int spectral(gp5_ctx *c, const double2 *src, const double coef[2], const EpiArgs &E)
{
    return stage_passes(c, plain_input(src), coef, E, false);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *gp5_last_error(void) { return g_err.c_str(); }

int gp5_version(void) { return 1; }

int gp5_ctx_create(gp5_ctx **out, const int shape[3], const double left[3], const double right[3], int nc)
{
    if (!out || !shape || !left || !right) return fail(GP5_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (nc < 1 || nc > kMaxComp) return fail(GP5_ERR_UNSUPPORTED, "nComponents must be 1..%d (got %d)", kMaxComp, nc);
    for (int d = 0; d < 3; ++d) {
        if (shape[d] < 1) return fail(GP5_ERR_ARG, "shape[%d] = %d", d, shape[d]);
        if (shape[d] > kMaxLen)
            return fail(GP5_ERR_UNSUPPORTED, "axis %d has %d points; the LDS FFT supports up to %d", d, shape[d], kMaxLen);
        if (!(right[d] > left[d])) return fail(GP5_ERR_ARG, "domain axis %d is empty", d);
    }
    gp5_ctx *c = new gp5_ctx;
    memset(c, 0, sizeof *c);
    for (int d = 0; d < 3; ++d) {
        c->n[d] = shape[d];
        c->left[d] = left[d];
        c->right[d] = right[d];
        c->dk[d] = 2 * M_PI / (right[d] - left[d]);  // discretization::getFourierStep
    }
    c->nc = nc;
    c->nc_cap = nc;
    c->uncoupled = nc > 2;
    c->M = (long long)shape[0] * shape[1] * shape[2];
    c->dV = 1.0;
    for (int d = 0; d < 3; ++d) c->dV *= (right[d] - left[d]) / shape[d];  // getSpaceStep
    int rc = GP5_OK;
    auto bail = [&](int code) { gp5_ctx_destroy(c); return code; };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(GP5_ERR_HIP, "no HIP device"));
    rc = scratch_setup(c->device);
    if (rc != GP5_OK) return bail(rc);
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(GP5_ERR_HIP, "hipStreamCreate failed"));
    for (int d = 0; d < 3; ++d) {
        if (make_plan(shape[d], c->plan[d])) return bail(fail(GP5_ERR_UNSUPPORTED, "too many radices for %d", shape[d]));
        std::vector<double2> tw = twiddles(shape[d]);
        if (hipMalloc(&c->tw[d], sizeof(double2) * shape[d]) != hipSuccess) return bail(fail(GP5_ERR_ALLOC, "twiddles"));
        if (hipMemcpy(c->tw[d], tw.data(), sizeof(double2) * shape[d], hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(GP5_ERR_HIP, "twiddle upload"));
        c->plan[d].tw = c->tw[d];
        // GP5_GENERIC_FFT=1 forces the generic LDS kernels (used to cross-check paths)
        const char *gen = getenv("GP5_GENERIC_FFT");
        c->fast[d] = fast_supported(shape[d]) && !(gen && gen[0] == '1');
        // GP5_NO_FUSE=1 runs the x forward transform as its own pass (A/B timing)
        const char *nf = getenv("GP5_NO_FUSE");
        c->fuse_x = !(nf && nf[0] == '1');
        // x slabs for the y/z passes: off by default.  Measured at 256^3 on
        // MI355X, 2/4/8/16 slabs of the three passes were slower in sum than
        // whole-field passes (the Infinity Cache did not return the slab
        // bandwidth, and the smaller launches lose tail efficiency);
        // GP5_MALL_CHUNKS=n turns it on for experiments.
        int chunks = 1;
        const char *mc = getenv("GP5_MALL_CHUNKS");
        if (mc && mc[0]) chunks = atoi(mc);
        c->mall_chunks = std::max(1, std::min(chunks, shape[0] / 16));
        // XCD-aware tile order (GP5_XCD_REMAP=n overrides: bit 0 columns, bit 1 x
        // lines).  Columns on by default: measured at 512^3, y pass 823 -> 750 us and
        // z pass 1192 -> 1110 us (52.4 -> 55.0 steps/s), batch of 64 droplets 20.2 ->
        // 19.8 ms per step, 256^3 unchanged; on x lines it slowed the fused epilogue
        // (219 -> 232 us at 256^3, 1940 -> 1989 us at 512^3).
        const char *xr = getenv("GP5_XCD_REMAP");
        c->xcd_remap = (xr && xr[0]) ? (atoi(xr) & 3) : 1;
        c->geom[d].B = 0;  // default tile
    }
    // the epilogue's partial sums carry two component slots: with more than two
    // components an x-line block must not span more than one boundary
    if (nc > 2 && c->fast[0] && fast_rows_per_block(shape[0]) > (long long)shape[1] * shape[2]) c->fast[0] = false;
    set_geometry(c);
    if (hipMalloc(&c->results, kRedN * (size_t)nc * sizeof(double)) != hipSuccess) return bail(fail(GP5_ERR_ALLOC, "results"));
    if (hipMalloc(&c->scales, 8 * (size_t)nc * sizeof(double)) != hipSuccess) return bail(fail(GP5_ERR_ALLOC, "scales"));
    if (hipMalloc(&c->blockmax, 8192 * sizeof(double)) != hipSuccess) return bail(fail(GP5_ERR_ALLOC, "blockmax"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(GP5_ERR_HIP, "events"));
    rc = ensure_partials(c, 4096);
    if (rc != GP5_OK) return bail(rc);
    *out = c;
    return GP5_OK;
}

int gp5_ctx_destroy(gp5_ctx *c)
{
    if (!c) return GP5_OK;
    // teardown continues past individual failures; the first one is reported
    hipError_t first = hipSuccess;
    auto note = [&](hipError_t e) { if (first == hipSuccess) first = e; };
    if (c->st) note(hipStreamSynchronize(c->st));
    for (int d = 0; d < 3; ++d) if (c->tw[d]) note(hipFree(c->tw[d]));
    void *ptrs[] = {c->T,       c->kA,      c->kB,       c->sbuf, c->eout,  c->phase_ex, c->partials,
                    c->results, c->scales, c->blockmax, c->dlog, c->nnorm, c->lhy_ab, c->Vdev, c->ccl};
    for (void *p : ptrs) if (p) note(hipFree(p));
    if (c->ev0) note(hipEventDestroy(c->ev0));
    if (c->ev1) note(hipEventDestroy(c->ev1));
    if (c->st) note(hipStreamDestroy(c->st));
    delete c;
    if (first != hipSuccess) return fail(GP5_ERR_HIP, "context teardown: %s", hipGetErrorString(first));
    return GP5_OK;
}

int gp5_ctx_sync(gp5_ctx *c)
{
    if (!c) return fail(GP5_ERR_ARG, "NULL context");
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

size_t gp5_ctx_points(const gp5_ctx *c) { return c ? (size_t)c->M : 0; }

int gp5_ctx_fft_plans(const gp5_ctx *c, int variant[3], int count[3])
{
    if (!c || !variant || !count) return fail(GP5_ERR_ARG, "NULL argument");
    for (int r = 0; r < 3; ++r) {
        const int axis = r;  // role 0 runs on x, 1 on y, 2 on z
        variant[r] = c->pvar[r];
        count[r] = c->fast[axis] ? fast_variants(c->n[axis]) : 0;
    }
    return GP5_OK;
}

int gp5_ctx_set_fft_plans(gp5_ctx *c, const int variant[3])
{
    if (!c || !variant) return fail(GP5_ERR_ARG, "NULL argument");
    for (int r = 0; r < 3; ++r) {
        const int cnt = c->fast[r] ? fast_variants(c->n[r]) : 1;
        if (variant[r] < 0 || variant[r] >= std::max(1, cnt))
            return fail(GP5_ERR_ARG, "plan variant %d of role %d: %d available", variant[r], r, std::max(1, cnt));
    }
    // x-line blocks of a many-component context must not span two component boundaries
    if (c->fast[0] && c->nc_cap > 2 &&
        fast_rows_per_block(c->n[0], variant[0]) > (long long)c->n[1] * c->n[2])
        return fail(GP5_ERR_UNSUPPORTED, "x-line plan variant %d: blocks would span components", variant[0]);
    HIPCHK(hipStreamSynchronize(c->st));
    for (int r = 0; r < 3; ++r) c->pvar[r] = variant[r];
    return GP5_OK;
}

const char *gp5_ctx_fft_plan_name(const gp5_ctx *c, int role)
{
    static thread_local std::string name;
    if (!c || role < 0 || role > 2) return "";
    name = c->fast[role] ? fast_plan_name(c->n[role], role, c->pvar[role]) : "generic";
    return name.c_str();
}

int gp5_buf_alloc(gp5_ctx *c, size_t n, gp5_buf **out)
{
    if (!c || !out) return fail(GP5_ERR_ARG, "NULL argument");
    *out = nullptr;
    gp5_buf *b = new gp5_buf{c, nullptr, n};
    if (n) {
        if (hipMalloc(&b->d, n * sizeof(double)) != hipSuccess) {
            delete b;
            return fail(GP5_ERR_ALLOC, "cannot allocate %zu B on the device", n * sizeof(double));
        }
        if (hipMemsetAsync(b->d, 0, n * sizeof(double), c->st) != hipSuccess ||
            hipStreamSynchronize(c->st) != hipSuccess) {
            (void)hipFree(b->d);
            delete b;
            return fail(GP5_ERR_HIP, "memset of a new buffer failed");
        }
    }
    *out = b;
    return GP5_OK;
}

int gp5_buf_free(gp5_buf *b)
{
    if (!b) return GP5_OK;
    // hipFree synchronises with the device itself, so the owning context (which
    // may already be destroyed) is not touched here.
    hipError_t e = hipSuccess;
    if (b->d) e = hipFree(b->d);
    delete b;
    if (e != hipSuccess) return fail(GP5_ERR_HIP, "buffer free: %s", hipGetErrorString(e));
    return GP5_OK;
}

size_t gp5_buf_size(const gp5_buf *b) { return b ? b->n : 0; }

int gp5_buf_upload(gp5_buf *b, const double *host, size_t n)
{
    if (!b || (!host && n)) return fail(GP5_ERR_ARG, "NULL argument");
    if (n != b->n) return fail(GP5_ERR_ARG, "upload of %zu doubles into a buffer of %zu", n, b->n);
    HIPCHK(hipMemcpyAsync(b->d, host, n * sizeof(double), hipMemcpyHostToDevice, b->ctx->st));
    HIPCHK(hipStreamSynchronize(b->ctx->st));
    return GP5_OK;
}

int gp5_buf_download(const gp5_buf *b, double *host, size_t n)
{
    if (!b || (!host && n)) return fail(GP5_ERR_ARG, "NULL argument");
    if (n != b->n) return fail(GP5_ERR_ARG, "download of %zu doubles from a buffer of %zu", n, b->n);
    HIPCHK(hipMemcpyAsync(host, b->d, n * sizeof(double), hipMemcpyDeviceToHost, b->ctx->st));
    HIPCHK(hipStreamSynchronize(b->ctx->st));
    return GP5_OK;
}

int gp5_buf_copy(gp5_buf *dst, const gp5_buf *src)
{
    if (!dst || !src) return fail(GP5_ERR_ARG, "NULL argument");
    if (dst->n != src->n) return fail(GP5_ERR_ARG, "copy between buffers of %zu and %zu doubles", src->n, dst->n);
    if (dst == src) return GP5_OK;
    HIPCHK(hipMemcpyAsync(dst->d, src->d, src->n * sizeof(double), hipMemcpyDeviceToDevice, dst->ctx->st));
    HIPCHK(hipStreamSynchronize(dst->ctx->st));
    return GP5_OK;
}

int gp5_buf_zero(gp5_buf *b)
{
    if (!b) return fail(GP5_ERR_ARG, "NULL argument");
    HIPCHK(hipMemsetAsync(b->d, 0, b->n * sizeof(double), b->ctx->st));
    HIPCHK(hipStreamSynchronize(b->ctx->st));
    return GP5_OK;
}

int gp5_fft(gp5_ctx *c, const gp5_buf *src, gp5_buf *dst, int direction)
{
    if (!c) return fail(GP5_ERR_ARG, "NULL context");
    RCHK(check_field(c, src, "src"));
    RCHK(check_field(c, dst, "dst"));
    int sign;
    if (direction == GP5_FORWARD) sign = -1;
    else if (direction == GP5_BACKWARD) sign = +1;
    else return fail(GP5_ERR_ARG, "Unkown direction in fourier transform");
    double2 *d = (double2 *)dst->d;
    RCHK(pass_plain(c, K_FFT, 0, (const double2 *)src->d, d, sign, nullptr, nullptr));
    RCHK(pass_plain(c, K_FFT, 1, d, d, sign, nullptr, nullptr));
    RCHK(pass_plain(c, K_FFT, 2, d, d, sign, nullptr, nullptr));
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

int gp5_laplacian(gp5_ctx *c, const gp5_buf *src, gp5_buf *dst)
{
    if (!c) return fail(GP5_ERR_ARG, "NULL context");
    RCHK(check_field(c, src, "src"));
    RCHK(check_field(c, dst, "dst"));
    RCHK(ensure_scratch(c, false, false));
    EpiArgs E = base_epi(c);
    E.mode = EPI_STORE;
    E.out = (double2 *)dst->d;
    double coef[2] = {-1.0 / (double)c->M, -1.0 / (double)c->M};
    RCHK(spectral(c, (const double2 *)src->d, coef, E));
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

int gp5_functional_apply(gp5_ctx *c, const gp5_functional_desc *f, const gp5_buf *src, gp5_buf *dst, double time)
{
    (void)time;
    if (!c) return fail(GP5_ERR_ARG, "NULL context");
    RCHK(check_field(c, src, "src"));
    RCHK(check_field(c, dst, "dst"));
    if (src == dst || src->d == dst->d) return fail(GP5_ERR_ARG, "functional::apply needs distinct source and destination");
    RCHK(ensure_scratch(c, false, false));
    EpiArgs E = base_epi(c);
    double coef[2];
    RCHK(fill_functional(c, f, E, coef));
    E.mode = EPI_H;
    E.phi = (const double2 *)src->d;
    E.out = (double2 *)dst->d;
    RCHK(spectral(c, (const double2 *)src->d, coef, E));
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

static int norm_impl(gp5_ctx *c, const double2 *f, int comp, double N, double *out, double *scale_out)
{
    const long long M = c->M;
    unsigned blocks = (unsigned)std::min<long long>((M + kThreads - 1) / kThreads, 2048);
    RCHK(ensure_partials(c, blocks));
    EpiArgs E = base_epi(c);
    E.partials = c->partials;
    E.nc = 1;
    E.Nnorm[0] = N;
    E.scale_out = scale_out;
    HIPCHK(launch_norm(c->st, f + M * comp, M, blocks, E));
    HIPCHK(launch_finalize(c->st, E, blocks, 0, 1, 1));
    if (out) {
        double r = 0.0;
        HIPCHK(hipMemcpyAsync(&r, c->results, sizeof r, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        *out = r;
    }
    return GP5_OK;
}

int gp5_norm(gp5_ctx *c, const gp5_buf *f, int comp, double *out)
{
    if (!c || !out) return fail(GP5_ERR_ARG, "NULL argument");
    RCHK(check_field(c, f, "field"));
    if (comp < 0 || comp >= c->nc) return fail(GP5_ERR_ARG, "component %d out of range", comp);
    return norm_impl(c, (const double2 *)f->d, comp, 1.0, out, nullptr);
}

int gp5_normalize(gp5_ctx *c, gp5_buf *f, int comp, double N)
{
    if (!c) return fail(GP5_ERR_ARG, "NULL context");
    RCHK(check_field(c, f, "field"));
    if (comp < 0 || comp >= c->nc) return fail(GP5_ERR_ARG, "component %d out of range", comp);
    double nrm = 0.0;
    RCHK(norm_impl(c, (const double2 *)f->d, comp, N, &nrm, scale_slot(c, 7)));
    if (!std::isfinite(nrm) || nrm <= 0.0) return fail(GP5_ERR_NONFINITE, "norm of component %d is %g", comp, nrm);
    HIPCHK(launch_scale(c->st, (double2 *)f->d, c->M, c->nc, comp, 0.0, scale_slot(c, 7), nullptr));
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

int gp5_max_density(gp5_ctx *c, const gp5_buf *f, int comp, double *out)
{
    if (!c || !out) return fail(GP5_ERR_ARG, "NULL argument");
    RCHK(check_field(c, f, "field"));
    if (comp < 0 || comp >= c->nc) return fail(GP5_ERR_ARG, "component %d out of range", comp);
    unsigned blocks = (unsigned)std::min<long long>((c->M + kThreads - 1) / kThreads, 8192);
    HIPCHK(launch_maxdens(c->st, (const double2 *)f->d + c->M * comp, c->M, blocks, c->blockmax));
    std::vector<double> bm(blocks);
    HIPCHK(hipMemcpyAsync(bm.data(), c->blockmax, blocks * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double m = 0.0;
    for (double v : bm) m = std::max(m, v);
    *out = m;
    return GP5_OK;
}

// Gather table of one axis for resampling ns -> nd points over a box of length
// L with cell-centred samples (the comment of gp5_resample.hip derives it).
struct HostAxisTable {
    std::vector<int> off, src;
    std::vector<double2> w;
};

static HostAxisTable resample_table(int ns, int nd, double L)
{
    std::vector<std::vector<std::pair<int, double2>>> per(nd);
    const double shift = 0.5 * (L / nd - L / ns);  // x0 of the destination minus x0 of the source
    const int kmin = ns % 2 == 0 ? -ns / 2 : -(ns - 1) / 2;
    const int kmax = ns % 2 == 0 ? ns / 2 : (ns - 1) / 2;
    for (int k = kmin; k <= kmax; ++k) {
        const double ck = (ns % 2 == 0 && (k == kmin || k == kmax)) ? 0.5 : 1.0;
        const double ang = 2.0 * M_PI * (double)k * shift / L;
        const double2 w = make_double2(ck * cos(ang), ck * sin(ang));
        const int s = ((k % ns) + ns) % ns;
        const int d = ((k % nd) + nd) % nd;
        bool merged = false;
        for (auto &e : per[d])
            if (e.first == s) {
                e.second.x += w.x;
                e.second.y += w.y;
                merged = true;
            }
        if (!merged) per[d].push_back({s, w});
    }
    HostAxisTable t;
    t.off.push_back(0);
    for (int d = 0; d < nd; ++d) {
        for (auto &e : per[d]) {
            t.src.push_back(e.first);
            t.w.push_back(e.second);
        }
        t.off.push_back((int)t.src.size());
    }
    return t;
}

// dst (cd's grid, cd->nc components) = resampled src (cs's grid); both boxes equal.
static int resample_core(gp5_ctx *cs, const double2 *src, gp5_ctx *cd, double2 *dst)
{
    RCHK(ensure_scratch(cs, false, false));
    RCHK(pass_plain(cs, K_FFT, 0, src, cs->T, -1, nullptr, nullptr));
    RCHK(pass_plain(cs, K_FFT, 1, cs->T, cs->T, -1, nullptr, nullptr));
    RCHK(pass_plain(cs, K_FFT, 2, cs->T, cs->T, -1, nullptr, nullptr));
    HIPCHK(hipStreamSynchronize(cs->st));
    HostAxisTable tab[3];
    size_t bytes = 0;
    for (int d = 0; d < 3; ++d) {
        tab[d] = resample_table(cs->n[d], cd->n[d], cd->right[d] - cd->left[d]);
        bytes += (tab[d].off.size() + tab[d].src.size()) * sizeof(int) + tab[d].w.size() * sizeof(double2);
    }
    // nine pieces, each placed at a 16-byte boundary: up to 15 bytes of padding each
    bytes += 9 * 16;
    char *dev = nullptr;
    HIPCHK(hipMalloc(&dev, bytes));
    ResampleAxis ax[3];
    size_t o = 0;
    int rc = GP5_OK;
    for (int d = 0; d < 3 && rc == GP5_OK; ++d) {
        auto put = [&](const void *h, size_t n) -> void * {
            o = (o + 15) & ~(size_t)15;
            void *p = dev + o;
            if (o + n > bytes) rc = fail(GP5_ERR_HIP, "resample table layout");
            else if (hipMemcpy(p, h, n, hipMemcpyHostToDevice) != hipSuccess) rc = fail(GP5_ERR_HIP, "resample table upload");
            o += n;
            return p;
        };
        ax[d].w = (const double2 *)put(tab[d].w.data(), tab[d].w.size() * sizeof(double2));
        ax[d].off = (const int *)put(tab[d].off.data(), tab[d].off.size() * sizeof(int));
        ax[d].src = (const int *)put(tab[d].src.data(), tab[d].src.size() * sizeof(int));
        ax[d].nd = cd->n[d];
        ax[d].ns = cs->n[d];
    }
    if (rc == GP5_OK) {
        hipError_t e = launch_resample(cd->st, cs->T, dst, ax, cs->M, cd->M, cd->nc);
        if (e != hipSuccess) rc = fail(GP5_ERR_HIP, "resample: %s", hipGetErrorString(e));
    }
    for (int d = 0; d < 3 && rc == GP5_OK; ++d) rc = pass_plain(cd, K_FFT, d, dst, dst, +1, nullptr, nullptr);
    hipError_t es = hipStreamSynchronize(cd->st);
    (void)hipFree(dev);
    if (rc == GP5_OK && es != hipSuccess) rc = fail(GP5_ERR_HIP, "resample: %s", hipGetErrorString(es));
    return rc;
}

static bool same_box(const double *l0, const double *r0, const double *l1, const double *r1)
{
    for (int d = 0; d < 3; ++d) {
        const double s = std::max(1.0, std::max(fabs(r0[d] - l0[d]), fabs(r1[d] - l1[d])));
        if (fabs(l0[d] - l1[d]) > 1e-12 * s || fabs(r0[d] - r1[d]) > 1e-12 * s) return false;
    }
    return true;
}

int gp5_resample(gp5_ctx *cs, const gp5_buf *src, gp5_ctx *cd, gp5_buf *dst)
{
    if (!cs || !cd) return fail(GP5_ERR_ARG, "NULL context");
    RCHK(check_field(cs, src, "src"));
    RCHK(check_field(cd, dst, "dst"));
    if (cs->nc != cd->nc) return fail(GP5_ERR_ARG, "resample: %d components into %d", cs->nc, cd->nc);
    if (!same_box(cs->left, cs->right, cd->left, cd->right))
        return fail(GP5_ERR_ARG, "resample: the two grids must cover the same box");
    if (cs != cd) HIPCHK(hipStreamSynchronize(cd->st));
    return resample_core(cs, (const double2 *)src->d, cd, (double2 *)dst->d);
}

int gp5_decompose(gp5_ctx *c, const gp5_buf *psi, int comp, double cutoff, int *labels, int *ndroplets)
{
    if (!c || !labels) return fail(GP5_ERR_ARG, "NULL argument");
    RCHK(check_field(c, psi, "psi"));
    if (comp < 0 || comp >= c->nc) return fail(GP5_ERR_ARG, "component %d out of range", comp);
    if (c->M >= INT_MAX) return fail(GP5_ERR_UNSUPPORTED, "labelling needs fewer than 2^31 cells");
    const size_t ws = decompose_workspace_ints(c->M);
    const size_t need = ws + (size_t)c->M + 1;
    if (c->ccl_cap < need) {
        if (c->ccl) HIPCHK(hipFree(c->ccl));
        c->ccl = nullptr;
        c->ccl_cap = 0;
        HIPCHK(hipMalloc(&c->ccl, need * sizeof(int)));
        c->ccl_cap = need;
    }
    int *dlab = c->ccl + ws, *dcount = c->ccl + ws + c->M;
    HIPCHK(launch_decompose(c->st, (const double2 *)psi->d + c->M * comp, c->n[0], c->n[1], c->n[2], cutoff, c->ccl,
                            dlab, dcount));
    int count = 0;
    HIPCHK(hipMemcpyAsync(labels + (size_t)c->M * comp, dlab, (size_t)c->M * sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(&count, dcount, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int k = 0; k < c->nc; ++k)
        if (k != comp) memset(labels + (size_t)c->M * k, 0, (size_t)c->M * sizeof(int));
    if (ndroplets) *ndroplets = count;
    return GP5_OK;
}

int gp5_energy(gp5_ctx *c, const gp5_functional_desc *f, const gp5_buf *psi, double *out)
{
    if (!c || !out) return fail(GP5_ERR_ARG, "NULL argument");
    RCHK(check_field(c, psi, "psi"));
    if (f && f->kind == GP5_FUNC_LHY_UNLOCKED)
        return fail(GP5_ERR_UNSUPPORTED, "the unlocked LHY functional has no energy for general alpha, beta, eta");
    RCHK(ensure_scratch(c, false, false));
    EpiArgs E = base_epi(c);
    double coef[2];
    RCHK(fill_functional(c, f, E, coef));
    E.mode = EPI_ENERGY;
    E.phi = (const double2 *)psi->d;
    E.want_red = 1;
    RCHK(spectral(c, (const double2 *)psi->d, coef, E));
    std::vector<double> r((size_t)kRedN * c->nc);
    HIPCHK(hipMemcpyAsync(r.data(), c->results, r.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double e = 0.0;
    for (int k = 0; k < c->nc; ++k) e += r[(size_t)c->nc + k];
    *out = e;
    return GP5_OK;
}

int gp5_axpy(gp5_ctx *c, gp5_buf *out, const gp5_buf *a, double cr, double ci, const gp5_buf *k)
{
    if (!c) return fail(GP5_ERR_ARG, "NULL context");
    RCHK(check_field(c, out, "out"));
    RCHK(check_field(c, a, "a"));
    RCHK(check_field(c, k, "k"));
    HIPCHK(launch_axpy(c->st, (double2 *)out->d, (const double2 *)a->d, cr, ci, (const double2 *)k->d, c->M * c->nc));
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

int gp5_rk4_combine(gp5_ctx *c, gp5_buf *out, const gp5_buf *a, double cr, double ci, const gp5_buf *k1,
                    const gp5_buf *k2, const gp5_buf *k3, const gp5_buf *k4)
{
    if (!c) return fail(GP5_ERR_ARG, "NULL context");
    RCHK(check_field(c, out, "out"));
    RCHK(check_field(c, a, "a"));
    RCHK(check_field(c, k1, "k1"));
    RCHK(check_field(c, k2, "k2"));
    RCHK(check_field(c, k3, "k3"));
    RCHK(check_field(c, k4, "k4"));
    HIPCHK(launch_rk4_combine(c->st, (double2 *)out->d, (const double2 *)a->d, cr, ci, (const double2 *)k1->d,
                              (const double2 *)k2->d, (const double2 *)k3->d, (const double2 *)k4->d, c->M * c->nc));
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

int gp5_phase_constraint(gp5_ctx *c, gp5_buf *psi, const gp5_buf *phase)
{
    if (!c || !phase) return fail(GP5_ERR_ARG, "NULL argument");
    RCHK(check_field(c, psi, "psi"));
    RCHK(check_real(c, phase, "phase"));
    HIPCHK(launch_phase(c->st, (double2 *)psi->d, phase->d, c->M * c->nc));
    HIPCHK(hipStreamSynchronize(c->st));
    return GP5_OK;
}

// One step of Euler or RK4 on c's stream.  psi_slot: scale slot of psi (-1 = none).
// `last`: final step of this advance (its norm feeds the non-finite check).
static int step_once(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi,
                     int &psi_slot, int step, double *dlog, bool fused, bool last)
{
    EpiArgs E = base_epi(c);
    double coef[2];
    RCHK(fill_functional(c, f, E, coef));
    E.Nnorm[0] = s->normalizations[0];
    E.Nnorm[1] = s->normalizations[1];
    E.Nnorm_dev = c->nnorm;  // batch classes: one target per component
    if (s->phase) E.phase_exp = c->phase_ex;  // filled by advance_impl
    const bool rn = s->renormalize != 0;
    double *slot_psi = psi_slot >= 0 ? scale_slot(c, psi_slot) : nullptr;
    const int next_psi_slot = psi_slot == 0 ? 1 : 0;
    double *slotA = scale_slot(c, 2), *slotB = scale_slot(c, 3);
    double2 *P = (double2 *)psi->d;
    const double tr = s->tau_re, ti = s->tau_im;
    // reductions only where a result is consumed: lazy renormalisation scales,
    // the per-step log (first stage), and the non-finite check after the last step
    E.log = dlog ? dlog + (size_t)step * kRedN * c->nc : nullptr;
    // stage 1 input: psi itself; its lazy scale is applied in the z pass
    StageIn in = plain_input(P);
    in.mid_scale = slot_psi;
    if (s->kind == GP5_STEP_EULERO) {
        // euleroStepper::advance (src/stepper.cpp:30-43)
        E.mode = EPI_EULER;
        E.phi = P;
        E.sc_phi = slot_psi;
        E.cr = tr;
        E.ci = ti;
        E.out = c->eout;
        E.scale_out = rn ? scale_slot(c, next_psi_slot) : nullptr;
        E.want_red = rn || dlog || last;
        RCHK(stage_passes(c, in, coef, E, fused));
        // the new state lives in eout: swap device pointers with the caller's buffer
        double2 *fresh = c->eout;
        c->eout = (double2 *)psi->d;
        psi->d = (double *)fresh;
        psi_slot = rn ? next_psi_slot : -1;
        return GP5_OK;
    }
    // RK4Stepper::advance (src/stepper.cpp:46-70); see EpiMode for the K/S form
    E.psi = P;
    E.sc_psi = slot_psi;
    E.sbuf = c->sbuf;
    E.mode = EPI_RK1;
    E.kout = c->kA;  // k1
    E.cr = tr * 0.5;
    E.ci = ti * 0.5;
    E.scale_out = rn ? slotA : nullptr;
    E.want_red = rn || dlog;
    RCHK(stage_passes(c, in, coef, E, fused));
    E.log = nullptr;
    E.want_red = rn;

    // later stages: input u_s = psi - c_in K (psi scaled), renormalised by sc_in
    in.src = P;
    in.src_scale = slot_psi;

    E.mode = EPI_RK2;  // u2 = psi - (tau/2) k1
    E.kbuf = in.pre.src2 = c->kA;
    E.kout = c->kB;  // k2
    E.cr_in = in.pre.cfr = tr * 0.5;
    E.ci_in = in.pre.cfi = ti * 0.5;
    E.sc_in = in.mid_scale = rn ? slotA : nullptr;
    E.scale_out = rn ? slotB : nullptr;
    RCHK(stage_passes(c, in, coef, E, fused));

    E.mode = EPI_RK3;  // u3 = psi - (tau/2) k2
    E.kbuf = in.pre.src2 = c->kB;
    E.kout = c->kA;  // k3
    E.sc_in = in.mid_scale = rn ? slotB : nullptr;
    E.cr = tr;
    E.ci = ti;
    E.scale_out = rn ? slotA : nullptr;
    RCHK(stage_passes(c, in, coef, E, fused));

    E.mode = EPI_RK4;  // u4 = psi - tau k3
    E.kbuf = in.pre.src2 = c->kA;
    E.kout = nullptr;
    E.cr_in = in.pre.cfr = tr;
    E.ci_in = in.pre.cfi = ti;
    E.sc_in = in.mid_scale = rn ? slotA : nullptr;
    E.cr = (1. / 6) * tr;
    E.ci = (1. / 6) * ti;
    E.out = c->eout;
    E.scale_out = rn ? scale_slot(c, next_psi_slot) : nullptr;
    E.want_red = rn || last;
    RCHK(stage_passes(c, in, coef, E, fused));
    // psi' lives in eout: swap device pointers with the caller's buffer
    double2 *fresh = c->eout;
    c->eout = (double2 *)psi->d;
    psi->d = (double *)fresh;
    psi_slot = rn ? next_psi_slot : -1;
    return GP5_OK;
}

// advance = prepare (checks, allocations: nothing of it may run under graph
// capture) + enqueue (kernels only; capturable) + collect (read-back, checks).
static int advance_prepare(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi,
                           int nsteps, bool logging)
{
    if (!c || !f || !s) return fail(GP5_ERR_ARG, "NULL argument");
    RCHK(check_field(c, psi, "psi"));
    if (s->kind != GP5_STEP_EULERO && s->kind != GP5_STEP_RK4) return fail(GP5_ERR_ARG, "unknown stepper kind %d", s->kind);
    if (nsteps < 0) return fail(GP5_ERR_ARG, "nsteps < 0");
    if (s->renormalize) {
        if (c->nc > 2 && !c->nnorm)
            return fail(GP5_ERR_UNSUPPORTED, "renormalisation of more than two components needs per-component targets (batch API)");
        for (int k = 0; k < std::min(c->nc, 2); ++k)
            if (!c->nnorm && !(s->normalizations[k] > 0.0)) return fail(GP5_ERR_ARG, "normalization %d must be > 0", k);
    }
    RCHK(ensure_scratch(c, s->kind == GP5_STEP_RK4, true));
    if (s->phase) {
        RCHK(check_real(c, s->phase, "phase"));
        if (!c->phase_ex) HIPCHK(hipMalloc(&c->phase_ex, (size_t)c->M * c->nc_cap * sizeof(double2)));
        HIPCHK(launch_phase_exp(c->st, c->phase_ex, s->phase->d, c->M * c->nc));
    }
    size_t blocks = axis_grid(c->geom[0]);
    if (c->fast[0]) blocks = std::max<size_t>(blocks, fast_grid(c->n[0], c->geom[0].I, c->geom[0].O, c->pvar[0]));
    RCHK(ensure_partials(c, blocks));
    if (logging && (size_t)nsteps > c->dlog_cap) {
        if (c->dlog) HIPCHK(hipFree(c->dlog));
        c->dlog = nullptr;
        HIPCHK(hipMalloc(&c->dlog, (size_t)nsteps * kRedN * c->nc_cap * sizeof(double)));
        c->dlog_cap = nsteps;
    }
    return GP5_OK;
}

static int advance_enqueue(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi,
                           int nsteps, bool logging)
{
    int psi_slot = -1;
    // fused stages: the epilogue pass of each stage also forward-transforms its
    // output along x, so only the first stage needs a separate x pass
    const bool fused = c->fast[0] && c->fuse_x;
    if (fused) RCHK(pass_plain(c, K_X_FWD, 0, (const double2 *)psi->d, c->T, -1, nullptr, nullptr));
    for (int st = 0; st < nsteps; ++st)
        RCHK(step_once(c, f, s, psi, psi_slot, st, logging ? c->dlog : nullptr, fused, st == nsteps - 1));
    if (psi_slot >= 0)
        PLAUNCH(c, K_SCALE,
                launch_scale(c->st, (double2 *)psi->d, c->M, c->nc, -1, 0.0, scale_slot(c, psi_slot), c->active_dev));
    return GP5_OK;
}

// Reads the last step's output norms (non-finite check) and the per-step logs:
// norm_log[s*nc + c] and energy_log[s] (total), energy_comp_log[s*nc + c].
static int advance_collect(gp5_ctx *c, const gp5_functional_desc *f, int nsteps, double *norm_log, double *energy_log,
                           double *energy_comp_log)
{
    const int nc = c->nc;
    std::vector<double> r((size_t)nc);
    HIPCHK(hipMemcpyAsync(r.data(), c->results, r.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    const bool logging = norm_log || energy_log || energy_comp_log;
    std::vector<double> lg;
    if (logging) {
        lg.resize((size_t)nsteps * kRedN * nc);
        HIPCHK(hipMemcpyAsync(lg.data(), c->dlog, lg.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    const bool no_energy = f->kind == GP5_FUNC_LHY_UNLOCKED;
    for (int st = 0; logging && st < nsteps; ++st) {
        const double *q = &lg[(size_t)st * kRedN * nc];  // [quantity][component]
        double e = 0.0;
        for (int k = 0; k < nc; ++k) {
            if (norm_log) norm_log[(size_t)st * nc + k] = q[2 * nc + k];
            if (energy_comp_log) energy_comp_log[(size_t)st * nc + k] = no_energy ? NAN : q[nc + k];
            e += q[nc + k];
        }
        if (energy_log) energy_log[st] = no_energy ? NAN : e;
    }
    const int live = c->active_dev ? c->active : nc;  // batch classes: slots beyond the live ones are dead
    for (int k = 0; k < live; ++k)
        if (!std::isfinite(r[k])) return fail(GP5_ERR_NONFINITE, "norm of component %d became %g during the step", k, r[k]);
    return GP5_OK;
}

static int advance_impl(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi,
                        int nsteps, double *norm_log, double *energy_log, double *elapsed_ms)
{
    const bool logging = norm_log || energy_log;
    RCHK(advance_prepare(c, f, s, psi, nsteps, logging));
    if (nsteps == 0) return GP5_OK;
    if (elapsed_ms) HIPCHK(hipEventRecord(c->ev0, c->st));
    RCHK(advance_enqueue(c, f, s, psi, nsteps, logging));
    if (elapsed_ms) HIPCHK(hipEventRecord(c->ev1, c->st));
    RCHK(advance_collect(c, f, nsteps, norm_log, energy_log, nullptr));
    if (elapsed_ms) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *elapsed_ms = ms;
    }
    return GP5_OK;
}

int gp5_advance(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi, int nsteps,
                double *norm_log, double *energy_log)
{
    return advance_impl(c, f, s, psi, nsteps, norm_log, energy_log, nullptr);
}

int gp5_advance_timed(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi, int nsteps,
                      double *elapsed_ms)
{
    if (!elapsed_ms) return fail(GP5_ERR_ARG, "NULL elapsed_ms");
    return advance_impl(c, f, s, psi, nsteps, nullptr, nullptr, elapsed_ms);
}

// gp5_advance on the static scratch: the stage fields (T, K twice, S, the output
// field, e^{i phase}) and the epilogue's reduction partials are carved out of
// `scratch` instead of the context's lazily allocated buffers, which are left as
// they are.  Checks and enqueue only: nothing here allocates device or host memory.
int gp5_advance_noalloc(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi, int nsteps)
{
    if (!c || !f || !s) return fail(GP5_ERR_ARG, "NULL argument");
    RCHK(check_field(c, psi, "psi"));
    if (s->kind != GP5_STEP_EULERO && s->kind != GP5_STEP_RK4) return fail(GP5_ERR_ARG, "unknown stepper kind %d", s->kind);
    if (nsteps < 0) return fail(GP5_ERR_ARG, "nsteps < 0");
    if (s->renormalize) {
        if (c->nc > 2 && !c->nnorm)
            return fail(GP5_ERR_UNSUPPORTED, "renormalisation of more than two components needs per-component targets (batch API)");
        for (int k = 0; k < std::min(c->nc, 2); ++k)
            if (!c->nnorm && !(s->normalizations[k] > 0.0)) return fail(GP5_ERR_ARG, "normalization %d must be > 0", k);
    }
    if (s->phase) RCHK(check_real(c, s->phase, "phase"));
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device])
        return fail(GP5_ERR_HIP, "static scratch of device %d is not set up", c->device);
    if (nsteps == 0) return GP5_OK;
    // carve-out: fields of M * nc_cap complex values at 256-byte offsets, then partials
    const bool rk = s->kind == GP5_STEP_RK4;
    const int nfields = (rk ? 5 : 2) + (s->phase ? 1 : 0);
    const size_t field_bytes = ((size_t)c->M * c->nc_cap * sizeof(double2) + 255) & ~(size_t)255;
    size_t blocks = axis_grid(c->geom[0]);
    if (c->fast[0]) blocks = std::max<size_t>(blocks, fast_grid(c->n[0], c->geom[0].I, c->geom[0].O, c->pvar[0]));
    const size_t partial_bytes = blocks * kRedQ * sizeof(double);
    const size_t need = (size_t)nfields * field_bytes + partial_bytes;
    if (need > kScratchBytes)
        return fail(GP5_ERR_UNSUPPORTED,
                    "gp5_advance_noalloc: this field needs %zu B of scratch (%d fields of %zu B + %zu B of partials); "
                    "the static scratch holds %zu B",
                    need, nfields, field_bytes, partial_bytes, kScratchBytes);
    char *base = (char *)g_scratch_ptr[c->device];
    double2 *region[6];
    for (int k = 0; k < nfields; ++k) region[k] = (double2 *)(base + (size_t)k * field_bytes);
    // the context's own scratch is set aside for the call and restored afterwards
    double2 *keep_T = c->T, *keep_kA = c->kA, *keep_kB = c->kB, *keep_s = c->sbuf, *keep_e = c->eout;
    double2 *keep_ph = c->phase_ex;
    double *keep_part = c->partials;
    const size_t keep_cap = c->partials_cap;
    int r = 0;
    c->T = region[r++];
    if (rk) {
        c->kA = region[r++];
        c->kB = region[r++];
        c->sbuf = region[r++];
    }
    c->eout = region[r++];
    if (s->phase) c->phase_ex = region[r++];
    c->partials = (double *)(base + (size_t)nfields * field_bytes);
    c->partials_cap = blocks;
    double *const home = psi->d;
    int rc = GP5_OK;
    if (s->phase) {
        hipError_t e = launch_phase_exp(c->st, c->phase_ex, s->phase->d, c->M * c->nc);
        if (e != hipSuccess) rc = fail(GP5_ERR_HIP, "launch_phase_exp: %s", hipGetErrorString(e));
    }
    if (rc == GP5_OK) rc = advance_enqueue(c, f, s, psi, nsteps, false);
    // after an odd number of steps the state sits in the scratch output field
    if (psi->d != home) {
        hipError_t e = hipMemcpyAsync(home, psi->d, psi->n * sizeof(double), hipMemcpyDeviceToDevice, c->st);
        if (rc == GP5_OK && e != hipSuccess) rc = fail(GP5_ERR_HIP, "copy of the state home: %s", hipGetErrorString(e));
        psi->d = home;
    }
    double last[kMaxComp];
    hipError_t e = hipMemcpyAsync(last, c->results, (size_t)c->nc * sizeof(double), hipMemcpyDeviceToHost, c->st);
    hipError_t es = hipStreamSynchronize(c->st);
    c->T = keep_T;
    c->kA = keep_kA;
    c->kB = keep_kB;
    c->sbuf = keep_s;
    c->eout = keep_e;
    c->phase_ex = keep_ph;
    c->partials = keep_part;
    c->partials_cap = keep_cap;
    RCHK(rc);
    if (e != hipSuccess || es != hipSuccess)
        return fail(GP5_ERR_HIP, "gp5_advance_noalloc: %s", hipGetErrorString(e != hipSuccess ? e : es));
    for (int k = 0; k < c->nc; ++k)
        if (!std::isfinite(last[k])) return fail(GP5_ERR_NONFINITE, "norm of component %d became %g during the step", k, last[k]);
    return GP5_OK;
}

const char *gp5_kernel_kind_name(int kind)
{
    if (kind < 0 || kind >= GP5_NKINDS) return "";
    return kKindNames[kind];
}

int gp5_profile_advance(gp5_ctx *c, const gp5_functional_desc *f, const gp5_stepper_desc *s, gp5_buf *psi,
                        int nsteps, double *kind_ms, int *kind_launches)
{
    if (!c || !kind_ms || !kind_launches) return fail(GP5_ERR_ARG, "NULL argument");
    ProfState p;
    c->prof = &p;
    int rc = advance_impl(c, f, s, psi, nsteps, nullptr, nullptr, nullptr);
    c->prof = nullptr;
    for (int k = 0; k < GP5_NKINDS; ++k) {
        kind_ms[k] = 0.0;
        kind_launches[k] = 0;
    }
    if (rc == GP5_OK) {
        hipError_t e = hipStreamSynchronize(c->st);
        for (size_t i = 0; e == hipSuccess && i < p.used; ++i) {
            float ms = 0.f;
            e = hipEventElapsedTime(&ms, p.ev[2 * i], p.ev[2 * i + 1]);
            kind_ms[p.kind[i]] += ms;
            kind_launches[p.kind[i]] += 1;
        }
        if (e != hipSuccess) rc = fail(GP5_ERR_HIP, "profile events: %s", hipGetErrorString(e));
    }
    for (hipEvent_t ev : p.ev) (void)hipEventDestroy(ev);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches of independent instances (no counterpart in the reference)
//
// The instances that share a grid and box form one class: they are the
// components of one context, so every pass is one launch over all of them.
// Live instances occupy the slots [0, active) of their class; `active` lives
// in device memory and every pass of the class lets the blocks of later slots
// exit, so launches keep their shape while the class shrinks.
//
// Ground states run the reference's imaginary-time RK4 with per-stage
// renormalisation (src/stepper.cpp:46-70 as src/main.cpp:131-145 sets it up).
// One graph holds a replay of check_every (even) steps of every class, each
// class on its own stream (fork/join), followed per class by the on-device
// convergence check and compaction (gp5_batch.hip): an instance whose energy
// moved by at most tol relative since the previous check retires, its field
// goes to its result (preallocated at creation) and the class's last live
// slot takes its place.  The graph is captured once per call and replayed
// until no class has a live slot; the host reads the classes' states into
// pinned memory one replay late, so the device always has the next replay
// queued, and nothing is allocated between replays.
// ---------------------------------------------------------------------------
struct BatchClass {
    int n = 0;
    double L = 0.0;
    gp5_ctx *ctx = nullptr;
    gp5_buf *psi = nullptr;      // 2*M*cap doubles: every slot (slots >= active are dead)
    int cap = 0;
    // host copies of the device bookkeeping, valid between ground-state calls
    std::vector<int> slot_inst;  // slot -> instance; slots < active are live
    int active = 0;
    long long steps_done = 0;
    std::vector<double> prevE;   // per slot: energy at the previous check
    std::vector<int> checks;     // per slot: checks done
    // device bookkeeping (one allocation)
    char *dmem = nullptr;
    BatchClassDev *dev = nullptr;  // entry of gp5_batch::status_dev
    int *d_slot_inst = nullptr, *d_checks = nullptr;
    int *d_copy_slot = nullptr, *d_copy_inst = nullptr, *d_move_src = nullptr, *d_move_dst = nullptr;
    double *d_prevE = nullptr;
    hipEvent_t join = nullptr;
};

struct gp5_batch {
    int count = 0;
    std::vector<int> inst_class, inst_slot;  // instance -> class, slot (-1: retired)
    std::vector<double> N;
    std::vector<BatchClass> cls;
    std::vector<double *> result;            // per instance: its result field (in result_store)
    double *result_store = nullptr;
    double2 **res_ptr_dev = nullptr;         // device copy of the result pointers
    long long *res_steps_dev = nullptr;
    double *res_energy_dev = nullptr;
    std::vector<long long> steps;
    std::vector<double> energy;
    BatchClassDev *status_dev = nullptr;     // [classes]
    BatchClassDev *status_host = nullptr;    // pinned [2][classes]
    hipStream_t st = nullptr;                // the stream the joined graph is launched on
    hipEvent_t fork = nullptr;
    std::vector<gp5_ctx *> one;              // single-component contexts of gp5_batch_resample, by (n, L)
};

static int batch_check_inst(const gp5_batch *b, int i)
{
    if (!b) return fail(GP5_ERR_ARG, "NULL batch");
    if (i < 0 || i >= b->count) return fail(GP5_ERR_ARG, "instance %d out of range (batch of %d)", i, b->count);
    return GP5_OK;
}

static gp5_stepper_desc batch_stepper(const BatchClass &k, double dt_factor)
{
    gp5_stepper_desc s;
    memset(&s, 0, sizeof s);
    s.kind = GP5_STEP_RK4;
    const double dx = k.L / k.n;
    s.tau_re = dt_factor * dx * dx;  // imaginary time: tau = (dt, 0)
    s.tau_im = 0.0;
    s.renormalize = 1;
    s.normalizations[0] = s.normalizations[1] = 1.0;  // the class's device targets apply
    return s;
}

// Median-free timing of `steps` steps of a class on its own stream (ms).
static int batch_time(BatchClass &k, const gp5_functional_desc *f, const gp5_stepper_desc *s, int steps, float *ms)
{
    gp5_ctx *c = k.ctx;
    RCHK(advance_prepare(c, f, s, k.psi, steps + 1, false));
    RCHK(advance_enqueue(c, f, s, k.psi, 1, false));  // warm-up
    HIPCHK(hipEventRecord(c->ev0, c->st));
    RCHK(advance_enqueue(c, f, s, k.psi, steps, false));
    HIPCHK(hipEventRecord(c->ev1, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipEventElapsedTime(ms, c->ev0, c->ev1));
    return GP5_OK;
}

// Host bookkeeping -> device (before a ground-state call; synchronous).
static int batch_push_state(gp5_batch *b, BatchClass &k)
{
    BatchClassDev h;
    memset(&h, 0, sizeof h);
    h.active = k.active;
    h.err_inst = -1;
    h.steps_done = k.steps_done;
    HIPCHK(hipMemcpy(k.dev, &h, sizeof h, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(k.d_slot_inst, k.slot_inst.data(), k.cap * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(k.d_prevE, k.prevE.data(), k.cap * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(k.d_checks, k.checks.data(), k.cap * sizeof(int), hipMemcpyHostToDevice));
    k.ctx->active = k.active;
    (void)b;
    return GP5_OK;
}

// Device bookkeeping -> host (after a ground-state call; the streams are idle).
static int batch_pull_state(gp5_batch *b, BatchClass &k, const BatchClassDev &h, const std::vector<long long> &rs,
                            const std::vector<double> &re)
{
    std::vector<int> prev(k.slot_inst.begin(), k.slot_inst.begin() + k.active);
    HIPCHK(hipMemcpy(k.slot_inst.data(), k.d_slot_inst, k.cap * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(k.prevE.data(), k.d_prevE, k.cap * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(k.checks.data(), k.d_checks, k.cap * sizeof(int), hipMemcpyDeviceToHost));
    k.active = h.active;
    k.steps_done = h.steps_done;
    k.ctx->active = k.active;
    for (int inst : prev) b->inst_slot[inst] = -1;
    for (int sl = 0; sl < k.active; ++sl) b->inst_slot[k.slot_inst[sl]] = sl;
    for (int inst : prev)
        if (b->inst_slot[inst] < 0) {  // retired during this call
            b->steps[inst] = rs[inst];
            b->energy[inst] = re[inst];
        }
    return GP5_OK;
}

extern "C" {

int gp5_batch_create(gp5_batch **out, int count, const int *n, const double *L, const double *N)
{
    if (!out || !n || !L || !N) return fail(GP5_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (count < 1) return fail(GP5_ERR_ARG, "a batch needs at least one instance");
    for (int i = 0; i < count; ++i) {
        if (n[i] < 1 || n[i] > kMaxLen) return fail(GP5_ERR_UNSUPPORTED, "instance %d: grid %d", i, n[i]);
        if (!(L[i] > 0.0)) return fail(GP5_ERR_ARG, "instance %d: box length %g", i, L[i]);
        if (!(N[i] > 0.0)) return fail(GP5_ERR_ARG, "instance %d: normalization %g", i, N[i]);
    }
    gp5_batch *b = new gp5_batch;
    b->count = count;
    b->inst_class.assign(count, -1);
    b->inst_slot.assign(count, -1);
    b->N.assign(N, N + count);
    b->result.assign(count, nullptr);
    b->steps.assign(count, 0);
    b->energy.assign(count, NAN);
    // classes: instances with the same (n, L), at most kMaxComp per class
    for (int i = 0; i < count; ++i) {
        int k = -1;
        for (size_t j = 0; j < b->cls.size(); ++j)
            if (b->cls[j].n == n[i] && b->cls[j].L == L[i] && (int)b->cls[j].slot_inst.size() < kMaxComp) k = (int)j;
        if (k < 0) {
            b->cls.emplace_back();
            k = (int)b->cls.size() - 1;
            b->cls[k].n = n[i];
            b->cls[k].L = L[i];
        }
        b->inst_class[i] = k;
        b->inst_slot[i] = (int)b->cls[k].slot_inst.size();
        b->cls[k].slot_inst.push_back(i);
    }
    auto bail = [&](int rc) {
        gp5_batch_destroy(b);
        return rc;
    };
    const size_t ncls = b->cls.size();
    if (hipMalloc(&b->status_dev, ncls * sizeof(BatchClassDev)) != hipSuccess) return bail(fail(GP5_ERR_ALLOC, "batch state"));
    if (hipHostMalloc((void **)&b->status_host, 2 * ncls * sizeof(BatchClassDev), hipHostMallocDefault) != hipSuccess) {
        b->status_host = nullptr;
        return bail(fail(GP5_ERR_ALLOC, "pinned batch state"));
    }
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        b->st = nullptr;
        return bail(fail(GP5_ERR_HIP, "batch stream"));
    }
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        b->fork = nullptr;
        return bail(fail(GP5_ERR_HIP, "batch events"));
    }
    // the results of every instance, allocated once (retirement copies into them)
    size_t total = 0;
    for (int i = 0; i < count; ++i) total += (size_t)2 * n[i] * n[i] * n[i];
    if (hipMalloc(&b->result_store, total * sizeof(double)) != hipSuccess) {
        b->result_store = nullptr;
        return bail(fail(GP5_ERR_ALLOC, "cannot allocate %zu B for the batch results", total * sizeof(double)));
    }
    {
        size_t o = 0;
        for (int i = 0; i < count; ++i) {
            b->result[i] = b->result_store + o;
            o += (size_t)2 * n[i] * n[i] * n[i];
        }
        if (hipMalloc(&b->res_ptr_dev, count * sizeof(double2 *)) != hipSuccess ||
            hipMalloc(&b->res_steps_dev, count * sizeof(long long)) != hipSuccess ||
            hipMalloc(&b->res_energy_dev, count * sizeof(double)) != hipSuccess)
            return bail(fail(GP5_ERR_ALLOC, "batch result tables"));
        if (hipMemcpy(b->res_ptr_dev, b->result.data(), count * sizeof(double2 *), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(GP5_ERR_HIP, "batch result tables"));
        std::vector<long long> zs(count, 0);
        std::vector<double> ze(count, NAN);
        if (hipMemcpy(b->res_steps_dev, zs.data(), count * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->res_energy_dev, ze.data(), count * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(GP5_ERR_HIP, "batch result tables"));
    }
    for (size_t ki = 0; ki < ncls; ++ki) {
        BatchClass &k = b->cls[ki];
        const int nc = (int)k.slot_inst.size();
        k.cap = nc;
        const int shape[3] = {k.n, k.n, k.n};
        const double left[3] = {-0.5 * k.L, -0.5 * k.L, -0.5 * k.L}, right[3] = {0.5 * k.L, 0.5 * k.L, 0.5 * k.L};
        int rc = gp5_ctx_create(&k.ctx, shape, left, right, nc);
        if (rc != GP5_OK) return bail(rc);
        k.ctx->uncoupled = true;
        k.dev = b->status_dev + ki;
        k.ctx->active_dev = &k.dev->active;
        set_geometry(k.ctx);  // the generic passes carry the gate too
        rc = gp5_buf_alloc(k.ctx, (size_t)2 * k.ctx->M * nc, &k.psi);
        if (rc != GP5_OK) return bail(rc);
        if (hipMalloc(&k.ctx->nnorm, nc * sizeof(double)) != hipSuccess) return bail(fail(GP5_ERR_ALLOC, "batch normalisations"));
        std::vector<double> h(nc);
        for (int sl = 0; sl < nc; ++sl) h[sl] = N[k.slot_inst[sl]];
        if (hipMemcpy(k.ctx->nnorm, h.data(), nc * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(GP5_ERR_HIP, "batch normalisations upload"));
        // slot_inst, checks, copy/move lists (ints), prevE (doubles)
        const size_t ints = 6 * (size_t)nc, bytes = ((ints * sizeof(int) + 15) & ~(size_t)15) + nc * sizeof(double);
        if (hipMalloc(&k.dmem, bytes) != hipSuccess) {
            k.dmem = nullptr;
            return bail(fail(GP5_ERR_ALLOC, "batch bookkeeping"));
        }
        int *ip = (int *)k.dmem;
        k.d_slot_inst = ip;
        k.d_checks = ip + nc;
        k.d_copy_slot = ip + 2 * nc;
        k.d_copy_inst = ip + 3 * nc;
        k.d_move_src = ip + 4 * nc;
        k.d_move_dst = ip + 5 * nc;
        k.d_prevE = (double *)(k.dmem + ((ints * sizeof(int) + 15) & ~(size_t)15));
        if (hipEventCreateWithFlags(&k.join, hipEventDisableTiming) != hipSuccess) {
            k.join = nullptr;
            return bail(fail(GP5_ERR_HIP, "batch events"));
        }
        k.active = nc;
        k.prevE.assign(nc, NAN);
        k.checks.assign(nc, 0);
        rc = batch_push_state(b, k);
        if (rc != GP5_OK) return bail(rc);
    }
    *out = b;
    return GP5_OK;
}

int gp5_batch_destroy(gp5_batch *b)
{
    if (!b) return GP5_OK;
    int rc = GP5_OK;
    auto note = [&](int r) {
        if (rc == GP5_OK) rc = r;
    };
    if (b->st) (void)hipStreamSynchronize(b->st);
    for (BatchClass &k : b->cls) {
        if (k.psi) note(gp5_buf_free(k.psi));
        if (k.ctx) {
            k.ctx->active_dev = nullptr;
            note(gp5_ctx_destroy(k.ctx));
        }
        if (k.dmem) (void)hipFree(k.dmem);
        if (k.join) (void)hipEventDestroy(k.join);
    }
    for (gp5_ctx *c : b->one) note(gp5_ctx_destroy(c));
    void *ptrs[] = {b->result_store, b->res_ptr_dev, b->res_steps_dev, b->res_energy_dev, b->status_dev};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->status_host) (void)hipHostFree(b->status_host);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return rc;
}

static int batch_one_ctx(gp5_batch *b, int n, double L, gp5_ctx **out)
{
    for (gp5_ctx *c : b->one)
        if (c->n[0] == n && c->right[0] - c->left[0] == L) {
            *out = c;
            return GP5_OK;
        }
    const int shape[3] = {n, n, n};
    const double left[3] = {-0.5 * L, -0.5 * L, -0.5 * L}, right[3] = {0.5 * L, 0.5 * L, 0.5 * L};
    gp5_ctx *c = nullptr;
    RCHK(gp5_ctx_create(&c, shape, left, right, 1));
    b->one.push_back(c);
    *out = c;
    return GP5_OK;
}

int gp5_batch_resample(gp5_batch *dst, int i, const gp5_batch *src, int j)
{
    if (!dst || !src) return fail(GP5_ERR_ARG, "NULL batch");
    RCHK(batch_check_inst(dst, i));
    RCHK(batch_check_inst(src, j));
    const int s = dst->inst_slot[i];
    if (s < 0) return fail(GP5_ERR_ARG, "instance %d is retired (its ground state is final)", i);
    BatchClass &kd = dst->cls[dst->inst_class[i]];
    const BatchClass &ks = src->cls[src->inst_class[j]];
    const double Ld = kd.L, Ls = ks.L;
    if (fabs(Ld - Ls) > 1e-12 * std::max(Ld, Ls))
        return fail(GP5_ERR_ARG, "resample: instance boxes differ (%g vs %g)", Ls, Ld);
    gp5_ctx *cs = nullptr, *cd = nullptr;
    RCHK(batch_one_ctx(dst, ks.n, ks.L, &cs));
    RCHK(batch_one_ctx(dst, kd.n, kd.L, &cd));
    HIPCHK(hipStreamSynchronize(ks.ctx->st));
    HIPCHK(hipStreamSynchronize(kd.ctx->st));
    const int ss = src->inst_slot[j];
    const double *from = ss >= 0 ? ks.psi->d + (size_t)2 * ks.ctx->M * ss : src->result[j];
    return resample_core(cs, (const double2 *)from, cd, (double2 *)(kd.psi->d + (size_t)2 * kd.ctx->M * s));
}

int gp5_batch_classes(const gp5_batch *b) { return b ? (int)b->cls.size() : 0; }

int gp5_batch_class_of(const gp5_batch *b, int i)
{
    if (batch_check_inst(b, i) != GP5_OK) return -1;
    return b->inst_class[i];
}

int gp5_batch_upload(gp5_batch *b, int i, const double *host)
{
    RCHK(batch_check_inst(b, i));
    if (!host) return fail(GP5_ERR_ARG, "NULL host array");
    const int s = b->inst_slot[i];
    if (s < 0) return fail(GP5_ERR_ARG, "instance %d is retired (its ground state is final)", i);
    BatchClass &k = b->cls[b->inst_class[i]];
    const long long M = k.ctx->M;
    HIPCHK(hipMemcpy(k.psi->d + (size_t)2 * M * s, host, (size_t)M * sizeof(double2), hipMemcpyHostToDevice));
    return GP5_OK;
}

int gp5_batch_download(const gp5_batch *b, int i, double *host)
{
    RCHK(batch_check_inst(b, i));
    if (!host) return fail(GP5_ERR_ARG, "NULL host array");
    const BatchClass &k = b->cls[b->inst_class[i]];
    const long long M = k.ctx->M;
    const int s = b->inst_slot[i];
    const double *src = s >= 0 ? k.psi->d + (size_t)2 * M * s : b->result[i];
    HIPCHK(hipMemcpy(host, src, (size_t)M * sizeof(double2), hipMemcpyDeviceToHost));
    return GP5_OK;
}

int gp5_batch_set_lhy(gp5_batch *b, int i, double a, double bcoef)
{
    RCHK(batch_check_inst(b, i));
    const int s = b->inst_slot[i];
    if (s < 0) return fail(GP5_ERR_ARG, "instance %d is retired", i);
    BatchClass &k = b->cls[b->inst_class[i]];
    gp5_ctx *c = k.ctx;
    if (!c->lhy_ab) {
        std::vector<double> h(2 * (size_t)c->nc_cap);
        for (int sl = 0; sl < c->nc_cap; ++sl) {
            h[2 * sl] = -3.0;  // LHYDropletFunctional's coefficients (src/functional.cpp:117-118)
            h[2 * sl + 1] = 2.5;
        }
        HIPCHK(hipMalloc(&c->lhy_ab, h.size() * sizeof(double)));
        HIPCHK(hipMemcpy(c->lhy_ab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    const double ab[2] = {a, bcoef};
    HIPCHK(hipMemcpy(c->lhy_ab + 2 * s, ab, sizeof ab, hipMemcpyHostToDevice));
    return GP5_OK;
}

int gp5_batch_set_potential(gp5_batch *b, int i, const double *V)
{
    RCHK(batch_check_inst(b, i));
    if (!V) return fail(GP5_ERR_ARG, "NULL potential");
    const int s = b->inst_slot[i];
    if (s < 0) return fail(GP5_ERR_ARG, "instance %d is retired", i);
    BatchClass &k = b->cls[b->inst_class[i]];
    gp5_ctx *c = k.ctx;
    if (!c->Vdev) {
        HIPCHK(hipMalloc(&c->Vdev, (size_t)c->M * c->nc_cap * sizeof(double)));
        HIPCHK(hipMemset(c->Vdev, 0, (size_t)c->M * c->nc_cap * sizeof(double)));
    }
    HIPCHK(hipMemcpy(c->Vdev + (size_t)c->M * s, V, (size_t)c->M * sizeof(double), hipMemcpyHostToDevice));
    return GP5_OK;
}

int gp5_batch_advance(gp5_batch *b, const gp5_functional_desc *f, double dt_factor, int nsteps)
{
    if (!b || !f) return fail(GP5_ERR_ARG, "NULL argument");
    if (f->potential) return fail(GP5_ERR_ARG, "batch potentials are set per instance (gp5_batch_set_potential)");
    for (BatchClass &k : b->cls) {
        if (k.active == 0) continue;
        gp5_stepper_desc s = batch_stepper(k, dt_factor);
        RCHK(advance_impl(k.ctx, f, &s, k.psi, nsteps, nullptr, nullptr, nullptr));
        k.steps_done += nsteps;
    }
    return GP5_OK;
}

// ---- autotuner --------------------------------------------------------------
// Per class: the register-plan variant of each fast role (x lines, y columns, z
// pass) and the fused x pass are chosen by timing `steps` steps of each
// candidate on the class's own state (coordinate descent, role by role); axes on
// the generic kernels get their radix order and tile.  The register-plan and
// fusion choices are cached in a JSON file keyed by device architecture and
// grid size (GP5_TUNE_CACHE, default $HOME/.cache/gp5_amd/fft_plans.json;
// "off" disables it); a cached class is not timed again.
struct TuneEntry {
    int var[3];
    int fuse;
    double ms;
};

static std::string tune_cache_path()
{
    const char *e = getenv("GP5_TUNE_CACHE");
    if (e) return (strcmp(e, "off") == 0 || !e[0]) ? std::string() : std::string(e);
    const char *home = getenv("HOME");
    if (!home || !home[0]) return std::string();
    return std::string(home) + "/.cache/gp5_amd/fft_plans.json";
}

static std::string device_arch(int device)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

// The cache file: one flat JSON object, "arch/n": [x, y, z, fused, ms] per line.
static std::vector<std::pair<std::string, TuneEntry>> tune_cache_read(const std::string &path)
{
    std::vector<std::pair<std::string, TuneEntry>> out;
    FILE *fp = path.empty() ? nullptr : fopen(path.c_str(), "r");
    if (!fp) return out;
    char line[512];
    while (fgets(line, sizeof line, fp)) {
        char key[128];
        TuneEntry t;
        if (sscanf(line, " \"%127[^\"]\" : [ %d , %d , %d , %d , %lf ]", key, &t.var[0], &t.var[1], &t.var[2], &t.fuse,
                   &t.ms) == 6)
            out.push_back({key, t});
    }
    fclose(fp);
    return out;
}

static void tune_cache_write(const std::string &path, const std::vector<std::pair<std::string, TuneEntry>> &all)
{
    if (path.empty()) return;
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos) {
        // create the directory chain (mkdir -p)
        std::string dir = path.substr(0, slash);
        for (size_t k = 1; k <= dir.size(); ++k)
            if (k == dir.size() || dir[k] == '/') (void)mkdir(dir.substr(0, k).c_str(), 0755);
    }
    const std::string tmp = path + ".tmp";
    FILE *fp = fopen(tmp.c_str(), "w");
    if (!fp) return;
    fprintf(fp, "{\n");
    for (size_t k = 0; k < all.size(); ++k) {
        const TuneEntry &t = all[k].second;
        fprintf(fp, "  \"%s\": [%d, %d, %d, %d, %.6f]%s\n", all[k].first.c_str(), t.var[0], t.var[1], t.var[2], t.fuse,
                t.ms, k + 1 < all.size() ? "," : "");
    }
    fprintf(fp, "}\n");
    fclose(fp);
    (void)rename(tmp.c_str(), path.c_str());
}

int gp5_batch_tune(gp5_batch *b, const gp5_functional_desc *f, double dt_factor, int steps, char *report, size_t report_len)
{
    if (!b || !f) return fail(GP5_ERR_ARG, "NULL argument");
    if (steps < 1) return fail(GP5_ERR_ARG, "steps < 1");
    std::string rep;
    const std::string cache_path = tune_cache_path();
    auto cache = tune_cache_read(cache_path);
    bool cache_dirty = false;
    for (BatchClass &k : b->cls) {
        if (k.active == 0) continue;
        gp5_ctx *c = k.ctx;
        gp5_stepper_desc s = batch_stepper(k, dt_factor);
        char line[512];
        const std::string key = device_arch(c->device) + "/" + std::to_string(k.n);
        const bool tunable = c->fast[0] || c->fast[1] || c->fast[2];
        const TuneEntry *hit = nullptr;
        for (auto &e : cache)
            if (e.first == key) hit = &e.second;
        if (tunable && hit) {
            int v[3] = {hit->var[0], hit->var[1], hit->var[2]};
            if (gp5_ctx_set_fft_plans(c, v) == GP5_OK) {
                c->fuse_x = hit->fuse != 0;
                snprintf(line, sizeof line, "class n=%d L=%g instances=%d: cached (%s) x %s, y %s, z %s, fused x %s\n", k.n,
                         k.L, k.active, key.c_str(), fast_plan_name(k.n, 0, v[0]).c_str(),
                         fast_plan_name(k.n, 1, v[1]).c_str(), fast_plan_name(k.n, 2, v[2]).c_str(),
                         c->fuse_x ? "on" : "off");
                rep += line;
                continue;
            }
        }
        // tune on a copy: the class state is restored afterwards
        const size_t bytes = (size_t)c->M * c->nc * sizeof(double2);
        double *save = nullptr;
        HIPCHK(hipMalloc(&save, bytes));
        HIPCHK(hipMemcpy(save, k.psi->d, bytes, hipMemcpyDeviceToDevice));
        float best = 0.f;
        int rc = batch_time(k, f, &s, steps, &best);
        const float dflt = best;
        snprintf(line, sizeof line, "class n=%d L=%g instances=%d: default %.3f ms/step", k.n, k.L, k.active, best / steps);
        rep += line;
        // register plans, role by role
        for (int role = 0; rc == GP5_OK && role < 3; ++role) {
            if (!c->fast[role]) continue;
            const int cnt = fast_variants(c->n[role]);
            const int keep = c->pvar[role];
            int bestv = keep;
            std::string tried;
            for (int v = 0; rc == GP5_OK && v < cnt; ++v) {
                if (v == keep) continue;
                int vv[3] = {c->pvar[0], c->pvar[1], c->pvar[2]};
                vv[role] = v;
                if (gp5_ctx_set_fft_plans(c, vv) != GP5_OK) continue;  // e.g. blocks spanning components
                float t = 0.f;
                rc = batch_time(k, f, &s, steps, &t);
                snprintf(line, sizeof line, "%s%s %.3f", tried.empty() ? "" : ", ", fast_plan_name(k.n, role, v).c_str(),
                         t / steps);
                tried += line;
                if (rc == GP5_OK && t < best) {
                    best = t;
                    bestv = v;
                }
                vv[role] = bestv;
                (void)gp5_ctx_set_fft_plans(c, vv);
            }
            static const char *role_name[3] = {"x", "y", "z"};
            snprintf(line, sizeof line, "; %s plan %s%s", role_name[role], fast_plan_name(k.n, role, c->pvar[role]).c_str(),
                     bestv != keep ? " (tuned)" : "");
            rep += line;
            if (!tried.empty()) rep += " [tried " + tried + "]";
        }
        // generic-kernel axes: radix order and tile
        for (int d = 0; rc == GP5_OK && d < 3; ++d) {
            if (c->fast[d]) continue;
            AxisPlan keep = c->plan[d];
            int best_order = -1;
            for (int o = 0; rc == GP5_OK && o < kPlanOrders; ++o) {
                AxisPlan P;
                if (make_plan(c->n[d], P, o)) continue;
                P.tw = c->tw[d];
                if (P.nrad == keep.nrad && std::equal(P.rad, P.rad + P.nrad, keep.rad)) continue;
                c->plan[d] = P;
                float t = 0.f;
                rc = batch_time(k, f, &s, steps, &t);
                if (rc == GP5_OK && t < best) {
                    best = t;
                    best_order = o;
                    keep = P;
                }
            }
            c->plan[d] = keep;
            const int b0 = c->geom[d].B;
            int best_B = b0;
            for (int B = 1; rc == GP5_OK && B <= 64; B *= 2) {
                if (B == b0 || B * c->n[d] > kTileElems) continue;
                c->geom[d].B = B;
                set_geometry(c);
                if (c->geom[d].B != B) continue;
                float t = 0.f;
                rc = batch_time(k, f, &s, steps, &t);
                if (rc == GP5_OK && t < best) {
                    best = t;
                    best_B = B;
                }
            }
            c->geom[d].B = best_B;
            set_geometry(c);
            std::string radices;
            for (int r = 0; r < c->plan[d].nrad; ++r) radices += (r ? "x" : "") + std::to_string(c->plan[d].rad[r]);
            snprintf(line, sizeof line, "; axis %d plan %s%s tile %d", d, radices.c_str(), best_order >= 0 ? " (tuned)" : "",
                     c->geom[d].B);
            rep += line;
        }
        if (rc == GP5_OK && c->fast[0]) {
            const bool f0 = c->fuse_x;
            c->fuse_x = !f0;
            float t = 0.f;
            rc = batch_time(k, f, &s, steps, &t);
            if (rc == GP5_OK && t < best) best = t;
            else c->fuse_x = f0;
            snprintf(line, sizeof line, "; fused x %s", c->fuse_x ? "on" : "off");
            rep += line;
        }
        snprintf(line, sizeof line, "; tuned %.3f ms/step (default %.3f)\n", best / steps, dflt / steps);
        rep += line;
        hipError_t e = hipMemcpy(k.psi->d, save, bytes, hipMemcpyDeviceToDevice);
        (void)hipFree(save);
        RCHK(rc);
        HIPCHK(e);
        if (tunable) {
            TuneEntry t = {{c->pvar[0], c->pvar[1], c->pvar[2]}, c->fuse_x ? 1 : 0, (double)best / steps};
            bool found = false;
            for (auto &en : cache)
                if (en.first == key) {
                    en.second = t;
                    found = true;
                }
            if (!found) cache.push_back({key, t});
            cache_dirty = true;
        }
    }
    if (cache_dirty) tune_cache_write(cache_path, cache);
    if (report && report_len) {
        strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return GP5_OK;
}

// One replay of check_every steps of class k on its stream, then the
// convergence check and compaction (capturable: kernels only).
static int batch_enqueue_replay(gp5_batch *b, BatchClass &k, const gp5_functional_desc *f, double dt_factor,
                                double tol, int check_every, long long max_steps)
{
    gp5_ctx *c = k.ctx;
    gp5_stepper_desc s = batch_stepper(k, dt_factor);
    double *const home = k.psi->d;
    RCHK(advance_enqueue(c, f, &s, k.psi, check_every, true));
    if (k.psi->d != home) return fail(GP5_ERR_HIP, "a replay must return the state to its buffer (even step count)");
    BatchCheckArgs a;
    memset(&a, 0, sizeof a);
    a.st = k.dev;
    a.cap = k.cap;
    a.check_every = check_every;
    a.tol = tol;
    a.max_steps = max_steps;
    a.slot_inst = k.d_slot_inst;
    a.prevE = k.d_prevE;
    a.checks = k.d_checks;
    a.log_row = c->dlog + (size_t)(check_every - 1) * kRedN * c->nc;
    a.results = c->results;
    a.nnorm = c->nnorm;
    a.lhy_ab = c->lhy_ab;
    a.res_steps = b->res_steps_dev;
    a.res_energy = b->res_energy_dev;
    a.copy_slot = k.d_copy_slot;
    a.copy_inst = k.d_copy_inst;
    a.move_src = k.d_move_src;
    a.move_dst = k.d_move_dst;
    HIPCHK(launch_batch_check(c->st, a));
    HIPCHK(launch_batch_compact(c->st, (double2 *)k.psi->d, c->Vdev, c->M, a, b->res_ptr_dev));
    return GP5_OK;
}

// Graph of one replay of the classes in `which`: joined (every class on its own
// stream, forked from and joined to `origin`) when origin is given, else the one
// class captured on its own stream.
static int batch_capture(gp5_batch *b, const std::vector<int> &which, hipStream_t origin, const gp5_functional_desc *f,
                         double dt_factor, double tol, int check_every, long long max_steps, hipGraphExec_t *exec)
{
    *exec = nullptr;
    hipStream_t cap_st = origin ? origin : b->cls[which[0]].ctx->st;
    if (hipStreamBeginCapture(cap_st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(GP5_ERR_HIP, "stream capture failed to start");
    int rc = GP5_OK;
    if (origin && hipEventRecord(b->fork, origin) != hipSuccess) rc = fail(GP5_ERR_HIP, "fork event");
    for (size_t q = 0; q < which.size() && rc == GP5_OK; ++q) {
        BatchClass &k = b->cls[which[q]];
        if (origin && hipStreamWaitEvent(k.ctx->st, b->fork, 0) != hipSuccess) {
            rc = fail(GP5_ERR_HIP, "fork wait");
            break;
        }
        rc = batch_enqueue_replay(b, k, f, dt_factor, tol, check_every, max_steps);
        if (rc == GP5_OK && origin &&
            (hipEventRecord(k.join, k.ctx->st) != hipSuccess || hipStreamWaitEvent(origin, k.join, 0) != hipSuccess))
            rc = fail(GP5_ERR_HIP, "join events");
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap_st, &g);
    if (rc == GP5_OK && e != hipSuccess) rc = fail(GP5_ERR_HIP, "stream capture: %s", hipGetErrorString(e));
    if (rc == GP5_OK && hipGraphInstantiate(exec, g, nullptr, nullptr, 0) != hipSuccess) {
        *exec = nullptr;
        rc = fail(GP5_ERR_HIP, "graph instantiation failed");
    }
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

int gp5_batch_ground_state(gp5_batch *b, const gp5_functional_desc *f, double dt_factor, double tol, int check_every,
                           long long max_steps, long long *steps_out, double *energy_out, double *elapsed_ms)
{
    if (!b || !f) return fail(GP5_ERR_ARG, "NULL argument");
    if (f->potential) return fail(GP5_ERR_ARG, "batch potentials are set per instance (gp5_batch_set_potential)");
    if (f->kind == GP5_FUNC_LHY_UNLOCKED) return fail(GP5_ERR_UNSUPPORTED, "batches hold uncoupled instances");
    if (check_every < 2 || check_every % 2) return fail(GP5_ERR_ARG, "check_every must be even and >= 2 (got %d)", check_every);
    if (!(tol > 0.0)) return fail(GP5_ERR_ARG, "tol must be > 0");
    const auto wall0 = std::chrono::steady_clock::now();
    const int ncls = (int)b->cls.size();
    std::vector<int> live;
    for (int ki = 0; ki < ncls; ++ki) {
        BatchClass &k = b->cls[ki];
        HIPCHK(hipStreamSynchronize(k.ctx->st));
        RCHK(batch_push_state(b, k));
        if (k.active == 0) continue;
        gp5_stepper_desc s = batch_stepper(k, dt_factor);
        RCHK(advance_prepare(k.ctx, f, &s, k.psi, check_every, true));  // allocations happen here, not in replays
        live.push_back(ki);
    }
    // units of launch: one joined graph of every live class (default), or one graph
    // per class replayed at its own pace (GP5_BATCH_GRAPH=classes)
    const char *mode = getenv("GP5_BATCH_GRAPH");
    const bool per_class = mode && strcmp(mode, "classes") == 0;
    struct Unit {
        std::vector<int> cls;
        hipStream_t st = nullptr;
        hipGraphExec_t exec = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        int slot[2] = {0, 0};
        int head = 0, inflight = 0, launched = 0;
        bool stop = false;
    };
    std::vector<Unit> units;
    if (per_class) {
        for (int ki : live) {
            units.emplace_back();
            units.back().cls = {ki};
            units.back().st = b->cls[ki].ctx->st;
        }
    } else if (!live.empty()) {
        units.emplace_back();
        units.back().cls = live;
        units.back().st = b->st;
    }
    int rc = GP5_OK;
    for (size_t u = 0; u < units.size() && rc == GP5_OK; ++u) {
        Unit &U = units[u];
        rc = batch_capture(b, U.cls, per_class ? nullptr : b->st, f, dt_factor, tol, check_every, max_steps, &U.exec);
        for (int q = 0; q < 2 && rc == GP5_OK; ++q)
            if (hipEventCreateWithFlags(&U.ev[q], hipEventDisableTiming) != hipSuccess) {
                U.ev[q] = nullptr;
                rc = fail(GP5_ERR_HIP, "batch events");
            }
    }
    // replay loop: every unit keeps up to two replays queued; the state of a
    // replay is read (pinned, one replay late) when its event has completed
    const size_t one = sizeof(BatchClassDev);
    int err_inst = -1;
    while (rc == GP5_OK) {
        bool busy = false, progressed = false;
        for (size_t u = 0; u < units.size() && rc == GP5_OK; ++u) {
            Unit &U = units[u];
            while (!U.stop && U.inflight < 2 && rc == GP5_OK) {
                const int q = (U.head + U.inflight) % 2;
                BatchClassDev *dst = b->status_host + (size_t)q * ncls;
                hipError_t e = hipGraphLaunch(U.exec, U.st);
                if (e == hipSuccess) {
                    if (per_class) e = hipMemcpyAsync(dst + U.cls[0], b->status_dev + U.cls[0], one, hipMemcpyDeviceToHost, U.st);
                    else e = hipMemcpyAsync(dst, b->status_dev, ncls * one, hipMemcpyDeviceToHost, U.st);
                }
                if (e == hipSuccess) e = hipEventRecord(U.ev[q], U.st);
                if (e != hipSuccess) rc = fail(GP5_ERR_HIP, "batch replay launch: %s", hipGetErrorString(e));
                U.inflight++;
                U.launched++;
                progressed = true;
            }
            while (U.inflight > 0 && rc == GP5_OK) {
                const int q = U.head;
                hipError_t e = hipEventQuery(U.ev[q]);
                if (e == hipErrorNotReady) break;
                if (e != hipSuccess) {
                    rc = fail(GP5_ERR_HIP, "batch replay: %s", hipGetErrorString(e));
                    break;
                }
                U.head = (U.head + 1) % 2;
                U.inflight--;
                progressed = true;
                const BatchClassDev *h = b->status_host + (size_t)q * ncls;
                int act = 0;
                for (int ki : U.cls) {
                    act += h[ki].active;
                    if (h[ki].err && rc == GP5_OK) {
                        err_inst = h[ki].err_inst;
                        U.stop = true;
                        rc = GP5_ERR_NONFINITE;
                    }
                }
                if (act == 0) U.stop = true;
            }
            if (!U.stop || U.inflight > 0) busy = true;
        }
        if (!busy || rc != GP5_OK) break;
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    // drain, then bring the bookkeeping back to the host
    for (Unit &U : units) (void)hipStreamSynchronize(U.st);
    for (int ki : live) (void)hipStreamSynchronize(b->cls[ki].ctx->st);
    for (Unit &U : units) {
        if (U.exec) (void)hipGraphExecDestroy(U.exec);
        for (hipEvent_t e : U.ev)
            if (e) (void)hipEventDestroy(e);
    }
    const double wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    // the device may have retired and moved instances before any error: bring its
    // bookkeeping back in every case, so that the host's slot map stays true
    std::string err_msg = rc != GP5_OK ? std::string(gp5_last_error()) : std::string();
    std::vector<BatchClassDev> st(ncls);
    std::vector<long long> rs(b->count);
    std::vector<double> re(b->count);
    HIPCHK(hipMemcpy(st.data(), b->status_dev, ncls * one, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rs.data(), b->res_steps_dev, b->count * sizeof(long long), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(re.data(), b->res_energy_dev, b->count * sizeof(double), hipMemcpyDeviceToHost));
    for (int ki : live) RCHK(batch_pull_state(b, b->cls[ki], st[ki], rs, re));
    if (rc == GP5_ERR_NONFINITE) return fail(GP5_ERR_NONFINITE, "instance %d: non-finite energy or norm", err_inst);
    if (rc != GP5_OK) return fail(rc, "%s", err_msg.c_str());
    for (int i = 0; i < b->count; ++i) {
        if (steps_out) steps_out[i] = b->steps[i];
        if (energy_out) energy_out[i] = b->energy[i];
    }
    if (elapsed_ms) *elapsed_ms = wall_ms;
    return GP5_OK;
}

}  // extern "C"
