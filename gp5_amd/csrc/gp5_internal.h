// Internal structures shared by the HIP kernels (gp5_kernels.hip) and the
// C-ABI host code (gp5_capi.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace gp5 {

constexpr int kMaxRadices = 24;
constexpr int kThreads = 256;          // threads per FFT workgroup (4 wave64)
constexpr int kTileElems = 2048;       // complex elements per generic FFT tile
constexpr int kMaxLen = 2048;          // longest axis the generic LDS FFT accepts
constexpr int kRedQ = 6;               // reduction partials per block: 3 quantities x 2 component slots
constexpr int kRedN = 3;               // quantities: |out|^2, energy density of phi, |phi|^2
constexpr int kMaxComp = 256;          // components of one context (more than 2: uncoupled only)

// Stockham plan of one axis.  rad[] lists radices in execution order;
// tw points to n twiddles e^{-2 pi i t / n} (device memory).
struct AxisPlan {
    int n;
    int nrad;
    int rad[kMaxRadices];
    const double2 *tw;
};

// Geometry of one axis pass over a field of shape (nx, ny, nz, nc) stored
// column-major.  Lines of the axis are indexed by (inner i < I, outer o < O),
// element j of line (i, o) sits at i + I*(j + n*o).
struct AxisGeom {
    long long I;   // product of the dims before the axis (its element stride)
    long long O;   // product of the dims after the axis times nc
    int B;         // lines per tile
    // batch classes: number of live components (device); tiles of components at
    // or beyond it exit at once.  Null: every component is live.
    const int *active;
};

// Spectral multiplier of the fused forward/inverse pass:
// x_hat *= coef[c] * (kx^2 + ky^2 + kz^2), k from the reference rule
// (src/geometry.cpp:127-148), coef already carries the 1/M of
// src/operators.cpp:35 and the kinetic prefactor.
struct MidArgs {
    int nx, ny, nz;
    double dk[3];
    double coef[2];        // component 0, components >= 1 (more than 2 components share one mass)
    long long M;
    const double *scale;   // lazy per-component scale of the transformed field, or null
};

// RK4 of src/stepper.cpp:46-70 with stage vectors that are never stored:
// phi_{s+1} = norm(u_{s+1}), u_{s+1} = psi - c_{s+1} tau k_s, is rebuilt from
// psi and K = k_s wherever it is needed.  Fields between stages: psi, K, S
// (S = k1 + 2 k2, written by stage 2, read by stage 4) and the x transform of
// the raw u_{s+1} (T).  `o` below is the value the pass hands on (transformed
// along x in the fused pass, renormalisation partials taken of it).
enum EpiMode {
    EPI_STORE = 0,   // out = multiplier result (laplacian / kinetic term)
    EPI_H = 1,       // out = H(phi)
    EPI_EULER = 2,   // out = o = psi - tau H(psi)
    EPI_RK1 = 3,     // k1 = H(psi);            K = k1;                  o = u2 = psi - (tau/2) k1
    EPI_RK2 = 4,     // k2 = H(norm(u2(K)));    S = K + 2 k2;  K = k2;   o = u3 = psi - (tau/2) k2
    EPI_RK3 = 5,     // k3 = H(norm(u3(K)));    K = k3;                  o = u4 = psi - tau k3
    EPI_RK4 = 6,     // k4 = H(norm(u4(K)));    out = o = psi - (tau/6) (((S + 2 K) + k4)
    EPI_ENERGY = 7,  // reductions only (energy and norm of phi)
};

struct EpiArgs {
    int mode;
    int fkind;          // GP5_FUNC_*
    int has_couplings;
    int uncoupled;      // components are independent instances sharing g11 (batches, nc > 2)
    double g11, g12, g22;
    double lhy_a, lhy_b;
    const double *lhy_ab;  // per-component (a, b) pairs overriding lhy_a / lhy_b, or null
    double alpha, beta, eta;
    const double *V;    // real potential M*nc or null
    const double2 *phi; // H / ENERGY / EULER: input field (raw), scaled by *sc_phi
    const double *sc_phi;
    const double2 *psi; // RK: step start state (raw), scaled by *sc_psi
    const double *sc_psi;
    // RK: K read (k_{s-1}) and written (k_s) through different buffers, as
    // the coupled functionals read the other component's K and psi at the
    // same point; for the same reason psi' goes to `out`, never onto psi
    const double2 *kbuf;
    double2 *kout;
    double2 *sbuf;      // RK: S = k1 + 2 k2
    const double *sc_in;  // RK2-4: renormalisation scale of u_s (null: 1)
    double cr_in, ci_in;  // RK2-4: u_s = psi - c_in K
    double2 *out;
    double cr, ci;      // output coefficient (tau/2, tau/2, tau or tau/6; tau for Euler)
    const double2 *phase_exp;  // e^{i phase} of phaseConstraint (M*nc) or null
    long long M;
    int nc;
    int nx;
    // fused reductions of quantity q (0: |out|^2, 1: energy density of phi,
    // 2: |phi|^2) per component.  A block's lines span at most two components:
    // slot 0 holds its first line's component cb, slot 1 the component cb + 1.
    int want_red;
    double *partials;       // [gridDim][2 slots][kRedN]
    double *results;        // [kRedN][nc], final sums times dV
    double *log;            // optional copy of results (per-step log) or null
    double *scale_out;      // if non-null: scale_out[c] = sqrt(N[c] / results[0][c])
    double Nnorm[2];        // N[c] for up to two components ...
    const double *Nnorm_dev;  // ... or a device array of nc values when non-null
    double dV;
    const int *active;      // batch classes: live components (device), or null (finalize skips the rest)
};

// Load pre-operation of a first pass: v = src * scale - (cfr + i cfi) * src2.
struct LoadPre {
    const double2 *src2;
    double cfr, cfi;
};

// Arguments of the register-resident passes (gp5_fast.hip).
struct FastArgs {
    const double2 *src;
    double2 *dst;
    const double2 *tw;
    long long I, O, Opc;   // axis geometry as in AxisGeom
    // load: v = src * scale_in[c] - (cfr + i cfi) * src2 (terms absent when null);
    // rebuilds an RK stage vector u_s = psi - c_in K for an unfused x pass
    const double *scale_in;
    const double2 *src2;
    double cfr, cfi;
    MidArgs mid;
    EpiArgs epi;
    double2 *t_out;  // epilogue pass: forward x FFT of its output goes here (null: skip)
    // column kernels restricted to x in [x0, x0 + xcount) (columns i = x + nx*r);
    // xcount == 0: all columns, packed COLS at a time
    int nx, x0, xcount;
    // batch classes: live components (device); blocks whose (first) component is
    // at or beyond it exit at once.  Null: every component is live.
    const int *active;
    int xcd;  // XCD-aware tile order: bit 0 column passes, bit 1 x-line passes
};

// Convergence state of one class of a batch ground state (device memory, one
// entry per class in one array so that the host reads every class at once).
struct BatchClassDev {
    int active;            // live slots: the class's kernels gate on it
    int err;               // nonzero: a non-finite energy or norm was met
    int err_inst;          // the instance that had it
    int ncopy;             // slots retired at the last check (copied to their results)
    int nmove;             // live slots moved into retired ones at the last check
    int pad;
    long long steps_done;  // steps the class has taken
};

// Arguments of the on-device convergence check of one class (gp5_batch.hip):
// after every replay of check_every steps, an instance whose energy moved by at
// most tol relative since the previous check retires; its field is copied to its
// result and the class's last live slot takes its place.
struct BatchCheckArgs {
    BatchClassDev *st;
    int cap;                 // slots (components of the class's context)
    int check_every;
    double tol;
    long long max_steps;     // > 0: every live slot retires once steps_done reaches it
    int *slot_inst;          // [cap] slot -> instance
    double *prevE;           // [cap] energy at the previous check
    int *checks;             // [cap] checks done
    const double *log_row;   // [kRedN][cap] the replay's last per-step log row (energy at [1][slot])
    const double *results;   // [kRedN][cap] the last output's norms at [0][slot]
    double *nnorm;           // [cap] normalisation targets, moved with their slots
    double *lhy_ab;          // [2 cap] LHY coefficient pairs, or null
    long long *res_steps;    // [instances] steps at retirement (negative: max_steps reached unconverged)
    double *res_energy;      // [instances] energy at retirement
    int *copy_slot;          // [cap] retired slots of the last check ...
    int *copy_inst;          // [cap] ... and their instances
    int *move_src;           // [cap] original slots moved ...
    int *move_dst;           // [cap] ... into these
};
hipError_t launch_batch_check(hipStream_t st, const BatchCheckArgs &a);
// copy the retired slots' fields to their results, then move live slots into the holes
hipError_t launch_batch_compact(hipStream_t st, double2 *psi, double *V, long long M, const BatchCheckArgs &a,
                                double2 *const *res_ptr);

int set_error(int code, const char *msg);  // sets gp5_last_error(), returns code
bool fast_supported(int n);
// var: register-plan variant (0 default; fast_variants(n) of them): per context,
// one for the x lines, one for the y columns and one for the fused z pass
unsigned fast_grid(int n, long long I, long long O, int var = 0);
long long fast_rows_per_block(int n, int var = 0);  // x lines per block of the row kernels
hipError_t launch_fast(hipStream_t st, int n, int mode, int dir, const FastArgs &a, int var = 0);
int fast_variants(int n);
std::string fast_plan_name(int n, int role, int var);  // "R0xR1xR2/E" (role 0 x lines, 1 y, 2 z)
// spectral resampling (gp5_resample.hip): per destination index d of one axis,
// the source indices src[off[d] .. off[d+1]) with complex weights w (device)
struct ResampleAxis {
    const int *off;
    const int *src;
    const double2 *w;
    int nd, ns;
};
hipError_t launch_resample(hipStream_t st, const double2 *F, double2 *G, const ResampleAxis ax[3], long long Ms,
                           long long Md, int nc);
// droplet labels of one component (gp5_ccl.hip); ws holds decompose_workspace_ints(M) ints
size_t decompose_workspace_ints(long long M);
hipError_t launch_decompose(hipStream_t st, const double2 *f, int nx, int ny, int nz, double cutoff, int *ws,
                            int *labels, int *ndroplets_dev);

}  // namespace gp5
