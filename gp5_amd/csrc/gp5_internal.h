// Internal structures shared by the HIP kernels (gp5_kernels.hip) and the
// C-ABI host code (gp5_capi.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gp5 {

constexpr int kMaxRadices = 24;
constexpr int kThreads = 256;          // threads per FFT workgroup (4 wave64)
constexpr int kTileElems = 2048;       // complex elements per generic FFT tile
constexpr int kMaxLen = 2048;          // longest axis the generic LDS FFT accepts
constexpr int kRedQ = 6;               // reduction quantities per block (3 x 2 components)

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
};

// Spectral multiplier of the fused forward/inverse pass:
// x_hat *= coef[c] * (kx^2 + ky^2 + kz^2), k from the reference rule
// (src/geometry.cpp:127-148), coef already carries the 1/M of
// src/operators.cpp:35 and the kinetic prefactor.
struct MidArgs {
    int nx, ny, nz;
    double dk[3];
    double coef[2];
    long long M;
    const double *scale;   // lazy per-component scale of the transformed field, or null
};

enum EpiMode {
    EPI_STORE = 0,   // out = multiplier result (laplacian / kinetic term)
    EPI_H = 1,       // out = H(phi)
    EPI_EULER = 2,   // out = psi - tau*H(psi)
    EPI_RK1 = 3,     // ksum = k1;          out = psi - (tau/2) k1
    EPI_RK2 = 4,     // ksum += 2 k2;       out = psi - (tau/2) k2
    EPI_RK3 = 5,     // ksum += 2 k3;       out = psi - tau k3
    EPI_RK4 = 6,     // out = psi - (tau/6) (ksum + k4)
    EPI_ENERGY = 7,  // reductions only (energy and norm of phi)
};

struct EpiArgs {
    int mode;
    int fkind;          // GP5_FUNC_*
    int has_couplings;
    double g11, g12, g22;
    double lhy_a, lhy_b;
    double alpha, beta, eta;
    const double *V;    // real potential M*nc or null
    const double2 *phi; // stage input (raw), scaled by *sc_phi
    const double *sc_phi;
    const double2 *psi; // step start state (raw), scaled by *sc_psi
    const double *sc_psi;
    double2 *ksum;      // RK accumulator
    double2 *out;
    double cr, ci;      // stage coefficient (tau/2, tau or tau/6)
    const double2 *phase_exp;  // e^{i phase} of phaseConstraint (M*nc) or null
    long long M;
    int nc;
    int nx;
    // fused reductions: q = 0,1 norm of out (c=0,1); 2,3 energy of phi; 4,5 norm of phi
    int want_red;
    double *partials;       // [gridDim][kRedQ]
    double *results;        // [kRedQ], final sums times dV
    double *log;            // optional copy of results (per-step log) or null
    double *scale_out;      // if non-null: scale_out[c] = sqrt(N[c] / norm_out[c])
    double Nnorm[2];
    double dV;
};

// Arguments of the register-resident passes (gp5_fast.hip).
struct FastArgs {
    const double2 *src;
    double2 *dst;
    const double2 *tw;
    long long I, O, Opc;   // axis geometry as in AxisGeom
    const double *scale_in;
    MidArgs mid;
    EpiArgs epi;
    double2 *t_out;  // epilogue pass: forward x FFT of its output goes here (null: skip)
};

bool fast_supported(int n);
unsigned fast_grid(int n, long long I, long long O);
hipError_t launch_fast(hipStream_t st, int n, int mode, int dir, const FastArgs &a);

}  // namespace gp5
