// Register-resident Stockham FFT passes for axis lengths 16, 32, 64, 96, 128,
// 192, 256 and 512.
//
// Each thread keeps E complex values of one line in registers; a length-N
// transform runs as 1-3 radix stages (16x16 for 256, 16x8 for 128, 8x8 for 64,
// 8x8x8 for 512, 4x2x12 for 96, 4x4x12 for 192, other orders per layout below;
// the 12-point butterfly is a prime-factor 3x4 in registers) with one LDS
// exchange between stages.  Global loads and stores
// go straight between HBM and registers: element (t + TL*i) + (N/R)*r of a line
// belongs to thread t, so for fixed (i, r) consecutive threads touch consecutive
// elements (x lines) or consecutive columns (y/z axes, columns interleaved).
//
// Forward transforms run the radix list in order and inverse transforms in
// reverse, so the register layout after an inverse transform equals the input
// layout of a forward transform: the fused x pass (inverse x FFT, functional and
// RK4 stage epilogue, forward x FFT of the new stage vector) and the fused z pass
// (forward z FFT, k-space multiplier, inverse z FFT) never leave registers
// between their transforms.
#include "gp5_device.h"

#include <algorithm>
#include <string>
#include <type_traits>

namespace gp5 {

template <int R>
__device__ __forceinline__ void dftR(double2 *v, double s)
{
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v, s);
    else if constexpr (R == 4) dft4(v, s);
    else if constexpr (R == 8) dft8(v, s);
    else if constexpr (R == 16) dft16(v, s);
    else if constexpr (R == 12) dft12(v, s);
}

// Radix plans per layout (ROWS: x lines; else interleaved y/z columns): R0
// first for forward transforms; 0 = no stage; E = complex values per thread.
// The x-line kernels carry the stage epilogue, so they keep E at 8 to leave
// registers for it (16x16 would hold 16 values per thread).
template <int N, bool ROWS> struct FPlan;
#define GP5_PLAN(NN, RW, A, B, C, D, EE)                                                                      \
    template <> struct FPlan<NN, RW> {                                                                     \
        static constexpr int NP = NN, R0 = A, R1 = B, R2 = C, R3 = D, E = EE, COLS = 0; /* 0: default */ \
    }
GP5_PLAN(16, false, 16, 0, 0, 0, 16);
GP5_PLAN(32, false, 8, 4, 0, 0, 8);
GP5_PLAN(64, false, 8, 8, 0, 0, 8);
GP5_PLAN(128, false, 16, 8, 0, 0, 16);
// Plans of the 96 and 192 columns (y) and lines (x), with compile-time
// overrides: R0, R1, R2, R3, E.  On x lines the radix-12 stage early (12.2.4 at
// 96, 4.12.4 at 192) measured 1.4-1.6 % more steps/s than 4.2.12 and 4.4.12,
// which stay the better column plans.
#ifndef GP5_COL96_PLAN
#define GP5_COL96_PLAN 4, 2, 12, 0, 12
#endif
#ifndef GP5_ROW96_PLAN
#define GP5_ROW96_PLAN 12, 2, 4, 0, 12
#endif
#ifndef GP5_COL192_PLAN
#define GP5_COL192_PLAN 4, 4, 12, 0, 12
#endif
#ifndef GP5_ROW192_PLAN
#define GP5_ROW192_PLAN 4, 12, 4, 0, 12
#endif
#define GP5_PLAN_EXPAND(...) GP5_PLAN(__VA_ARGS__)
GP5_PLAN_EXPAND(96, false, GP5_COL96_PLAN);
GP5_PLAN_EXPAND(192, false, GP5_COL192_PLAN);
GP5_PLAN(256, false, 16, 16, 0, 0, 16);
#ifdef GP5_COL512_COLS
template <> struct FPlan<512, false> {
    static constexpr int NP = 512, R0 = 8, R1 = 8, R2 = 8, R3 = 0, E = 8, COLS = GP5_COL512_COLS;
};
#else
GP5_PLAN(512, false, 8, 8, 8, 0, 8);
#endif
GP5_PLAN(16, true, 4, 4, 0, 0, 4);
GP5_PLAN(32, true, 8, 4, 0, 0, 8);
GP5_PLAN(64, true, 8, 8, 0, 0, 8);
GP5_PLAN_EXPAND(96, true, GP5_ROW96_PLAN);
GP5_PLAN_EXPAND(192, true, GP5_ROW192_PLAN);
#ifdef GP5_ROWS_E16
GP5_PLAN(128, true, 16, 8, 0, 0, 16);
GP5_PLAN(256, true, 16, 16, 0, 0, 16);
#else
#ifndef GP5_ROW128_PLAN
#define GP5_ROW128_PLAN 8, 4, 4, 0, 8
#endif
#ifndef GP5_ROW256_PLAN
#define GP5_ROW256_PLAN 8, 8, 4, 0, 8
#endif
GP5_PLAN_EXPAND(128, true, GP5_ROW128_PLAN);
GP5_PLAN_EXPAND(256, true, GP5_ROW256_PLAN);
#endif
GP5_PLAN(512, true, 8, 8, 8, 0, 8);
#undef GP5_PLAN_EXPAND
#undef GP5_PLAN

// Plans of the z pass (MODE 1 columns: forward transform, k-space multiplier,
// inverse): the column plan unless specialised.  At 512 the z columns lie 4 MiB
// apart; 16 values per thread over 16 columns (512-thread blocks, plan 8.8.8,
// 128 VGPRs, 4 waves per SIMD) measured 970 us per pass against 1104 us for
// 16.8.4 over 8 columns and 1020 us for 8.8.8 with 32 values over 16 columns
// (both with the XCD-aware column order); 8.8.8 with 8 values over 8 columns
// stays the better plan for the y pass.  At 192 the radix-12 stage first (12.4.4 over 16 columns)
// measured 3.5 % faster than the 4.4.12 column plan.
template <int N> struct FPlanZ : FPlan<N, false> {};
#ifndef GP5_Z512_DEFAULT
#ifndef GP5_Z512_PLAN
#define GP5_Z512_PLAN 8, 8, 8, 0, 16, 16  // R0, R1, R2, R3, E, COLS
#endif
template <> struct FPlanZ<512> {
    static constexpr int p[6] = {GP5_Z512_PLAN};
    static constexpr int NP = 512, R0 = p[0], R1 = p[1], R2 = p[2], R3 = p[3], E = p[4], COLS = p[5];
};
#endif

// compile-time overrides of other z plans: R0, R1, R2, R3, E, COLS
#define GP5_ZPLAN_OVERRIDE(NN, PLAN)                                                                  \
    template <> struct FPlanZ<NN> {                                                                  \
        static constexpr int p[6] = {PLAN};                                                          \
        static constexpr int NP = NN, R0 = p[0], R1 = p[1], R2 = p[2], R3 = p[3], E = p[4], COLS = p[5]; \
    };
#ifdef GP5_Z256_PLAN
GP5_ZPLAN_OVERRIDE(256, GP5_Z256_PLAN)
#endif
#ifdef GP5_Z96_PLAN
GP5_ZPLAN_OVERRIDE(96, GP5_Z96_PLAN)
#endif
#ifndef GP5_Z192_DEFAULT
#ifndef GP5_Z192_PLAN
#define GP5_Z192_PLAN 12, 4, 4, 0, 12, 16
#endif
GP5_ZPLAN_OVERRIDE(192, GP5_Z192_PLAN)
#endif
#undef GP5_ZPLAN_OVERRIDE

// Alternative plans, chosen per context at run time by the autotuner
// (gp5_batch_tune): variant 0 is the default above, variants 1 and 2 reorder the
// stages (and for 64 and 128 change the values per thread).  Lengths without
// alternatives have one variant.  Per role: x lines (FPlanV<N, true, V>), y
// columns (FPlanV<N, false, V>) and the fused z pass (FPlanZV<N, V>).
template <int N> constexpr int kVariants = (N == 64 || N == 96 || N == 128 || N == 192 || N == 512) ? 3 : 1;
template <int N, bool ROWS, int V> struct FPlanV : FPlan<N, ROWS> {};
template <int N, int V> struct FPlanZV : FPlanZ<N> {};
#define GP5_PLAN_V(NN, RW, V, A, B, C, D, EE)                                                                \
    template <> struct FPlanV<NN, RW, V> {                                                                  \
        static constexpr int NP = NN, R0 = A, R1 = B, R2 = C, R3 = D, E = EE, COLS = 0;                     \
    }
#define GP5_ZPLAN_V(NN, V, A, B, C, D, EE, CC)                                                               \
    template <> struct FPlanZV<NN, V> {                                                                     \
        static constexpr int NP = NN, R0 = A, R1 = B, R2 = C, R3 = D, E = EE, COLS = CC;                    \
    }
GP5_PLAN_V(192, true, 1, 12, 4, 4, 0, 12);
GP5_PLAN_V(192, true, 2, 4, 4, 12, 0, 12);
GP5_PLAN_V(192, false, 1, 12, 4, 4, 0, 12);
GP5_PLAN_V(192, false, 2, 4, 12, 4, 0, 12);
GP5_ZPLAN_V(192, 1, 4, 4, 12, 0, 12, 16);
GP5_ZPLAN_V(192, 2, 12, 4, 4, 0, 12, 8);
GP5_PLAN_V(96, true, 1, 12, 4, 2, 0, 12);
GP5_PLAN_V(96, true, 2, 4, 2, 12, 0, 12);
GP5_PLAN_V(96, false, 1, 12, 2, 4, 0, 12);
GP5_PLAN_V(96, false, 2, 12, 4, 2, 0, 12);
GP5_ZPLAN_V(96, 1, 12, 4, 2, 0, 12, 16);
GP5_ZPLAN_V(96, 2, 12, 2, 4, 0, 12, 16);
GP5_PLAN_V(128, true, 1, 4, 8, 4, 0, 8);
GP5_PLAN_V(128, true, 2, 4, 4, 8, 0, 8);
GP5_PLAN_V(128, false, 1, 8, 16, 0, 0, 16);
GP5_PLAN_V(128, false, 2, 8, 4, 4, 0, 8);
GP5_ZPLAN_V(128, 1, 8, 16, 0, 0, 16, 16);
GP5_ZPLAN_V(128, 2, 8, 4, 4, 0, 8, 16);
GP5_PLAN_V(64, true, 1, 4, 4, 4, 0, 4);
GP5_PLAN_V(64, true, 2, 16, 4, 0, 0, 16);
GP5_PLAN_V(64, false, 1, 16, 4, 0, 0, 16);
GP5_PLAN_V(64, false, 2, 4, 16, 0, 0, 16);
GP5_ZPLAN_V(64, 1, 16, 4, 0, 0, 16, 16);
GP5_ZPLAN_V(64, 2, 4, 16, 0, 0, 16, 16);
GP5_PLAN_V(512, true, 1, 8, 8, 8, 0, 16);
GP5_PLAN_V(512, true, 2, 16, 8, 4, 0, 16);
GP5_PLAN_V(512, false, 1, 16, 8, 4, 0, 16);
GP5_PLAN_V(512, false, 2, 8, 8, 8, 0, 16);
GP5_ZPLAN_V(512, 1, 8, 8, 8, 0, 32, 16);
GP5_ZPLAN_V(512, 2, 16, 8, 4, 0, 16, 8);
#undef GP5_PLAN_V
#undef GP5_ZPLAN_V

template <int N, bool ROWS, int MODE, int V = 0>
using PlanOf = std::conditional_t<ROWS, FPlanV<N, true, V>, std::conditional_t<MODE == 1, FPlanZV<N, V>, FPlanV<N, false, V>>>;

// Block shapes.  rows: x lines, LPB lines per 128-thread block, LDS line stride
// N + N/16 (one pad element every 16 keeps stride-16 accesses conflict free).
// cols: B interleaved columns of TL threads each.
template <class P, bool ROWS> struct FShapeP {
    static constexpr int N = P::NP;
    static constexpr int E = P::E;
    static constexpr int TL = N / E;
#ifndef GP5_ROW_THREADS
#define GP5_ROW_THREADS 128
#endif
    static constexpr int ROW_THREADS = TL > GP5_ROW_THREADS ? TL : GP5_ROW_THREADS;
    static constexpr int LPB = ROW_THREADS / TL;
    static constexpr int ROW_STRIDE = N + N / 16;
#ifndef GP5_COLS
#define GP5_COLS 16
#endif
    static constexpr int COLS = P::COLS > 0 ? P::COLS : (TL >= 64 ? 8 : GP5_COLS);
    static constexpr int COL_THREADS = COLS * TL;
};
template <int N, bool ROWS, int V = 0> using FShape = FShapeP<FPlanV<N, ROWS, V>, ROWS>;

#ifndef GP5_COL_HALVES
#define GP5_COL_HALVES 1
#endif
struct RowMap {
    static constexpr bool kHalves = false;
    int base;
    __device__ __forceinline__ int operator()(int pos) const { return base + pos + (pos >> 4); }
};
// Column blocks exchange through LDS one half of their columns at a time
// (kHalves): half the LDS per block lets 3 blocks share a CU instead of 2
// (measured at 256^3: y pass 88.7 -> 87.0 us, z pass 92.7 -> 89.4 us).
struct ColMap {
    static constexpr bool kHalves = GP5_COL_HALVES != 0;
    int b, B;
    __device__ __forceinline__ int operator()(int pos) const { return b + B * pos; }
    __device__ __forceinline__ int half(int pos) const { return (b % (B / 2)) + (B / 2) * pos; }
};

// One radix-R stage with L already-combined points: E/R butterflies per thread.
template <int N, int E, int L, int R>
__device__ __forceinline__ void stage_compute(double2 (&v)[E], int t, const double2 *__restrict__ tw, double s)
{
    constexpr int TL = N / E;
    constexpr int m = N / (L * R);
#pragma unroll
    for (int i = 0; i < E / R; ++i) {
        if constexpr (L > 1) {
            // w_{LR}^{j r}, r = 1..R-1, by binary powering of one table entry
            // (at most 4 products deep for R = 16): one load instead of R-1
            // keeps the twiddles out of the register file until they are used
            const int bf = t + TL * i;
            const int j = bf / m;
            double2 w1 = tw[j * m];
            w1.y *= -s;  // table holds e^{-2 pi i t/N}; s = -1 forward keeps it
            double2 wp[R];
            wp[1] = w1;
#pragma unroll
            for (int r = 2; r < R; ++r) wp[r] = (r & 1) ? cmul(wp[r - 1], w1) : cmul(wp[r / 2], wp[r / 2]);
#pragma unroll
            for (int r = 1; r < R; ++r) v[i * R + r] = cmul(v[i * R + r], wp[r]);
        }
        dftR<R>(&v[i * R], s);
    }
}

// Stage outputs -> LDS, then next stage inputs <- LDS (in place, two barriers).
template <int N, int E, int L, int R, int RN, class Map>
__device__ __forceinline__ void exchange(double2 (&v)[E], int t, double2 *lds, const Map &map)
{
    constexpr int TL = N / E;
    constexpr int m = N / (L * R);
    constexpr int L2 = L * R;
    constexpr int m2 = N / (L2 * RN);
    if constexpr (Map::kHalves) {
        const int mine = map.b / (map.B / 2);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (mine == h) {
#pragma unroll
                for (int i = 0; i < E / R; ++i) {
                    const int bf = t + TL * i;
                    const int q = bf % m, j = bf / m;
#pragma unroll
                    for (int k = 0; k < R; ++k) lds[map.half(q + m * (j + L * k))] = v[i * R + k];
                }
            }
            __syncthreads();
            if (mine == h) {
#pragma unroll
                for (int i = 0; i < E / RN; ++i) {
                    const int bf = t + TL * i;
                    const int q = bf % m2, j = bf / m2;
#pragma unroll
                    for (int r = 0; r < RN; ++r) v[i * RN + r] = lds[map.half(q + m2 * (r + RN * j))];
                }
            }
            __syncthreads();
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < E / R; ++i) {
        const int bf = t + TL * i;
        const int q = bf % m, j = bf / m;
#pragma unroll
        for (int k = 0; k < R; ++k) lds[map(q + m * (j + L * k))] = v[i * R + k];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E / RN; ++i) {
        const int bf = t + TL * i;
        const int q = bf % m2, j = bf / m2;
#pragma unroll
        for (int r = 0; r < RN; ++r) v[i * RN + r] = lds[map(q + m2 * (r + RN * j))];
    }
    __syncthreads();
}

// Full transform over the stages Ra, Rb, Rc, Rd (0 = absent, trailing), sign s.
// In:  v[i*Ra + r] = x[(t + TL i) + (N/Ra) r];  out: v[i*Rl + k] = X[(t + TL i) + (N/Rl) k], Rl last radix.
template <int N, int E, int Ra, int Rb, int Rc, int Rd, class Map>
__device__ __forceinline__ void fft_run(double2 (&v)[E], int t, const double2 *__restrict__ tw, double s,
                                        double2 *lds, const Map &map)
{
    stage_compute<N, E, 1, Ra>(v, t, tw, s);
    if constexpr (Rb > 0) {
        exchange<N, E, 1, Ra, Rb>(v, t, lds, map);
        stage_compute<N, E, Ra, Rb>(v, t, tw, s);
        if constexpr (Rc > 0) {
            exchange<N, E, Ra, Rb, Rc>(v, t, lds, map);
            stage_compute<N, E, Ra * Rb, Rc>(v, t, tw, s);
            if constexpr (Rd > 0) {
                exchange<N, E, Ra * Rb, Rc, Rd>(v, t, lds, map);
                stage_compute<N, E, Ra * Rb * Rc, Rd>(v, t, tw, s);
            }
        }
    }
}

template <class P, class Map>
__device__ __forceinline__ void fft_forward(double2 (&v)[P::E], int t, const double2 *tw, double2 *lds, const Map &map)
{
    fft_run<P::NP, P::E, P::R0, P::R1, P::R2, P::R3>(v, t, tw, -1.0, lds, map);
}

template <class P, class Map>
__device__ __forceinline__ void fft_inverse(double2 (&v)[P::E], int t, const double2 *tw, double2 *lds, const Map &map)
{
    constexpr int N = P::NP;
    if constexpr (P::R3 > 0) fft_run<N, P::E, P::R3, P::R2, P::R1, P::R0>(v, t, tw, 1.0, lds, map);
    else if constexpr (P::R2 > 0) fft_run<N, P::E, P::R2, P::R1, P::R0, 0>(v, t, tw, 1.0, lds, map);
    else if constexpr (P::R1 > 0) fft_run<N, P::E, P::R1, P::R0, 0, 0>(v, t, tw, 1.0, lds, map);
    else fft_run<N, P::E, P::R0, 0, 0, 0>(v, t, tw, 1.0, lds, map);
}

template <class P> struct FirstLast {
    static constexpr int FWD_FIRST = P::R0;
    static constexpr int FWD_LAST = P::R3 > 0 ? P::R3 : (P::R2 > 0 ? P::R2 : (P::R1 > 0 ? P::R1 : P::R0));
    static constexpr int INV_FIRST = FWD_LAST;
    static constexpr int INV_LAST = P::R0;
};

// MODE 0: one transform (sign from DIR: -1 forward, +1 inverse) src -> dst.
// MODE 1: forward, k-space multiplier, inverse (the last axis).
// MODE 2: rows only: inverse, epilogue, optional forward of the epilogue output.
// MODE 3: MODE 2 with the component-local epilogue (no coupling, no constraint).
template <int N, bool ROWS, int MODE, int DIR, int V>
__device__ __forceinline__ void fast_body(const FastArgs &a)
{
    using PL = PlanOf<N, ROWS, MODE, V>;
    using S = FShapeP<PL, ROWS>;
    using FL = FirstLast<PL>;
    constexpr int E = S::E, TL = S::TL;
    extern __shared__ double2 lds[];
    const int tid = threadIdx.x;
    int t;
    long long gbase, gstride;
    bool valid;
    int c;
    long long colidx = 0;
    RowMap rmap{0};
    ColMap cmap{0, S::COLS};
    // tile of this block: XCD-aware order when a.xcd asks for it (bit 0 columns, bit 1 lines)
    const long long bid = xcd_tile(ROWS ? (a.xcd & 2) != 0 : (a.xcd & 1) != 0);
    if constexpr (ROWS) {
        const int l = tid / TL;
        t = tid - l * TL;
        const long long line = bid * S::LPB + l;
        valid = line < a.O;
        gbase = line * N;
        gstride = 1;
        c = valid ? (int)(line / a.Opc) : 0;
        rmap.base = l * S::ROW_STRIDE;
        // retired batch instances: the block's first line decides (block-uniform)
        if (a.active && bid * S::LPB >= (long long)*a.active * a.Opc) return;
    } else {
        const int b = tid % S::COLS;
        t = tid / S::COLS;
        long long o;
        if (a.xcount == 0) {
            const long long tpo = (a.I + S::COLS - 1) / S::COLS;
            o = bid / tpo;
            colidx = (bid - o * tpo) * S::COLS + b;
            valid = colidx < a.I;
        } else {
            // x slab [x0, x0 + xcount) of every row r of columns i = x + nx*r
            const long long tpr = (a.xcount + S::COLS - 1) / S::COLS;
            const long long rows_per_o = a.I / a.nx;
            const long long per_o = tpr * rows_per_o;
            o = bid / per_o;
            const long long rem = bid - o * per_o;
            const long long r = rem / tpr;
            const int xo = (int)(rem - r * tpr) * S::COLS + b;
            valid = xo < a.xcount;
            colidx = a.x0 + xo + a.nx * r;
        }
        gbase = colidx + a.I * (long long)N * o;
        gstride = a.I;
        c = (int)(o / a.Opc);
        cmap.b = b;
        if (a.active && c >= *a.active) return;  // retired batch instance (block-uniform)
    }
    double2 v[E];
    // ---- load: inputs of the first stage of the first transform ----
    constexpr bool first_fwd = (MODE == 1) || (MODE == 0 && DIR < 0);
    constexpr int R_IN = first_fwd ? FL::FWD_FIRST : FL::INV_FIRST;
    // streaming hints: epilogue passes follow GP5_NT_EPI (bit 0 loads, bit 1
    // stores), column passes GP5_NT_COLS
    constexpr bool NTL = (MODE >= 2) ? (GP5_NT_EPI & 1) != 0 : (!ROWS && GP5_NT_COLS != 0);
    constexpr bool NTS = (MODE >= 2) ? (GP5_NT_EPI & 2) != 0 : (!ROWS && GP5_NT_COLS != 0);
    // the epilogue pass's forward x transform of the next stage input (T, read next by
    // the y pass): GP5_NT_TOUT, else the epilogue's store hint
#ifdef GP5_NT_TOUT
    constexpr bool NTT = GP5_NT_TOUT != 0;
#else
    constexpr bool NTT = NTS;
#endif
    {
        const double2 cf = make_double2(a.cfr, a.cfi);
#pragma unroll
        for (int i = 0; i < E / R_IN; ++i) {
#pragma unroll
            for (int r = 0; r < R_IN; ++r) {
                const int pos = (t + TL * i) + (N / R_IN) * r;
                v[i * R_IN + r] = valid ? load_pre<NTL>(a.src, gbase + gstride * pos, a.scale_in, c, a.src2, cf)
                                        : make_double2(0.0, 0.0);
            }
        }
    }
    if constexpr (MODE == 0) {
        if constexpr (DIR < 0) {
            if constexpr (ROWS) fft_forward<PL>(v, t, a.tw, lds, rmap);
            else fft_forward<PL>(v, t, a.tw, lds, cmap);
        } else {
            if constexpr (ROWS) fft_inverse<PL>(v, t, a.tw, lds, rmap);
            else fft_inverse<PL>(v, t, a.tw, lds, cmap);
        }
        constexpr int RL = DIR < 0 ? FL::FWD_LAST : FL::INV_LAST;
        if (valid) {
#pragma unroll
            for (int i = 0; i < E / RL; ++i)
#pragma unroll
                for (int k = 0; k < RL; ++k) stg<NTS>(a.dst + gbase + gstride * ((t + TL * i) + (N / RL) * k), v[i * RL + k]);
        }
        return;
    }
    if constexpr (MODE == 1) {
        if constexpr (ROWS) fft_forward<PL>(v, t, a.tw, lds, rmap);
        else fft_forward<PL>(v, t, a.tw, lds, cmap);
        // multiplier coef[c] * |k|^2 (* lazy scale); this pass runs on z, so the
        // column index carries (x, y)
        const MidArgs &md = a.mid;
        double kxy2 = 0.0;
        if constexpr (!ROWS) {
            const int x = (int)(colidx % md.nx);
            const int y = (int)(colidx / md.nx);
            const double kx = momentum(x, md.nx, md.dk[0]);
            const double ky = momentum(y, md.ny, md.dk[1]);
            kxy2 = kx * kx + ky * ky;
        }
        double f0 = valid ? md.coef[c == 0 ? 0 : 1] : 0.0;
        if (md.scale && valid) f0 *= md.scale[c];
        constexpr int RL = FL::FWD_LAST;
#pragma unroll
        for (int i = 0; i < E / RL; ++i)
#pragma unroll
            for (int k = 0; k < RL; ++k) {
                const int pos = (t + TL * i) + (N / RL) * k;
                const double kz = momentum(pos, N, md.dk[2]);
                v[i * RL + k] = cscale(v[i * RL + k], f0 * (kxy2 + kz * kz));
            }
        if constexpr (ROWS) fft_inverse<PL>(v, t, a.tw, lds, rmap);
        else fft_inverse<PL>(v, t, a.tw, lds, cmap);
        constexpr int RO = FL::INV_LAST;
        if (valid) {
#pragma unroll
            for (int i = 0; i < E / RO; ++i)
#pragma unroll
                for (int k = 0; k < RO; ++k) stg<NTS>(a.dst + gbase + gstride * ((t + TL * i) + (N / RO) * k), v[i * RO + k]);
        }
        return;
    }
    if constexpr (MODE == 2 || MODE == 3) {
        static_assert(ROWS, "the epilogue pass runs on x lines");
        fft_inverse<PL>(v, t, a.tw, lds, rmap);
        EpiAcc acc;
        constexpr int RL = FL::INV_LAST;  // == FWD_FIRST
#ifndef GP5_EPI_CHUNK
#define GP5_EPI_CHUNK 4
#endif
        constexpr int CH = GP5_EPI_CHUNK < E ? GP5_EPI_CHUNK : E;
        static_assert(E % CH == 0, "epilogue chunking");
        if (valid) {
#pragma unroll
            for (int e0 = 0; e0 < E; e0 += CH) {
                // keep each chunk's loads inside its chunk: hoisting all chunks
                // at once exhausts the VGPR file and drops occupancy to 1 wave
                __builtin_amdgcn_sched_barrier(0);
                int pos[CH];
                double2 w[CH];
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    const int e = e0 + u, i = e / RL, k = e % RL;
                    pos[u] = (t + TL * i) + (N / RL) * k;
                    w[u] = v[e];
                }
                epilogue_batch<CH, MODE == 3>(a.epi, gbase, pos, c, w, acc);
#pragma unroll
                for (int u = 0; u < CH; ++u) v[e0 + u] = w[u];
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = make_double2(0.0, 0.0);
        }
        if (a.t_out) {
            fft_forward<PL>(v, t, a.tw, lds, rmap);
            constexpr int RF = FL::FWD_LAST;
            if (valid) {
#pragma unroll
                for (int i = 0; i < E / RF; ++i)
#pragma unroll
                    for (int k = 0; k < RF; ++k) stg<NTT>(a.t_out + gbase + (t + TL * i) + (N / RF) * k, v[i * RF + k]);
            }
        }
        if (a.epi.want_red) {
            double red[kRedQ];
#pragma unroll
            for (int q = 0; q < kRedQ; ++q) red[q] = 0.0;
            const int cb = (int)(bid * S::LPB / a.Opc);  // block's first component
            acc_scatter(acc, c - cb, red);
            __syncthreads();
            reduce_block(red, a.epi.partials, reinterpret_cast<double *>(lds), bid);
        }
    }
}

#ifdef GP5_WAVES_PER_EU
#define GP5_KATTR __attribute__((amdgpu_waves_per_eu(GP5_WAVES_PER_EU)))
#else
#define GP5_KATTR
#endif
template <int N, bool ROWS, int MODE, int DIR, int V>
__global__ void __launch_bounds__((ROWS ? FShapeP<PlanOf<N, ROWS, MODE, V>, ROWS>::ROW_THREADS
                                          : FShapeP<PlanOf<N, ROWS, MODE, V>, ROWS>::COL_THREADS))
    GP5_KATTR k_fast(FastArgs a)
{
    fast_body<N, ROWS, MODE, DIR, V>(a);
}

// The same pass held to 3 waves per SIMD (at most 168 VGPRs).  Used for the
// 256-point z pass, which needs 170 registers otherwise and then runs 2 waves
// per SIMD: without spills at 3 it measured 96.3 -> 88.0 us at 256^3.  The
// epilogue kernels would spill under this limit and stay on k_fast (at 192
// the x epilogue needs 175 registers; held to 168 it spills 32 bytes per lane
// and measured 106.2 -> 109.6 us).  More waves do not help the 96 and 192 z
// passes either: held to 5 waves per SIMD (no spills at 96) they measured
// 17.7 -> 18.1 us and 47.3 -> 50.0 us.
template <int N, bool ROWS, int MODE, int DIR, int V>
__global__ void __launch_bounds__((ROWS ? FShapeP<PlanOf<N, ROWS, MODE, V>, ROWS>::ROW_THREADS
                                          : FShapeP<PlanOf<N, ROWS, MODE, V>, ROWS>::COL_THREADS))
    __attribute__((amdgpu_waves_per_eu(3))) k_fast_w3(FastArgs a)
{
    fast_body<N, ROWS, MODE, DIR, V>(a);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool fast_supported(int n)
{
    return n == 16 || n == 32 || n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 512;
}

template <class S>
static void col_shape(long long I, long long O, int nx, int xcount, unsigned &grid, unsigned &threads, size_t &lds)
{
    if (xcount == 0) grid = (unsigned)(((I + S::COLS - 1) / S::COLS) * O);
    else grid = (unsigned)(((xcount + S::COLS - 1) / S::COLS) * (I / nx) * O);
    threads = S::COL_THREADS;
    lds = sizeof(double2) * (size_t)(ColMap::kHalves ? S::COLS / 2 : S::COLS) * S::N;
}

// launch shape of a pass; mode selects the z-pass plan for columns (MODE 1)
template <int N, int V>
static void fast_shape(bool rows, int mode, long long I, long long O, int nx, int xcount, unsigned &grid,
                       unsigned &threads, size_t &lds)
{
    if (rows) {
        using S = FShape<N, true, V>;
        grid = (unsigned)((O + S::LPB - 1) / S::LPB);
        threads = S::ROW_THREADS;
        lds = sizeof(double2) * (size_t)S::LPB * S::ROW_STRIDE;
    } else if (mode == 1) {
        col_shape<FShapeP<FPlanZV<N, V>, false>>(I, O, nx, xcount, grid, threads, lds);
    } else {
        col_shape<FShape<N, false, V>>(I, O, nx, xcount, grid, threads, lds);
    }
    lds = std::max(lds, sizeof(double) * (size_t)kRedScratch);
}

template <int N, int V>
static hipError_t launch_fast_nv(hipStream_t st, int mode, int dir, const FastArgs &a)
{
    const bool rows = (a.I == 1);
    unsigned grid, threads;
    size_t lds;
    fast_shape<N, V>(rows, mode, a.I, a.O, a.nx, a.xcount, grid, threads, lds);
    dim3 g(grid), b(threads);
#define GP5_FAST_LAUNCH(R, M, D) hipLaunchKernelGGL((k_fast<N, R, M, D, V>), g, b, lds, st, a)
    if (rows) {
        if (mode == 0 && dir < 0) GP5_FAST_LAUNCH(true, 0, -1);
        else if (mode == 0) GP5_FAST_LAUNCH(true, 0, 1);
        else if (mode == 1) GP5_FAST_LAUNCH(true, 1, -1);
        else if (mode == 2) GP5_FAST_LAUNCH(true, 2, 1);
        else GP5_FAST_LAUNCH(true, 3, 1);
    } else {
        if (mode == 0 && dir < 0) GP5_FAST_LAUNCH(false, 0, -1);
        else if (mode == 0) GP5_FAST_LAUNCH(false, 0, 1);
        else if (mode == 1) {
            if constexpr (N == 256) hipLaunchKernelGGL((k_fast_w3<N, false, 1, -1, V>), g, b, lds, st, a);
            else GP5_FAST_LAUNCH(false, 1, -1);
        } else return hipErrorInvalidValue;  // the epilogue pass needs x lines
    }
#undef GP5_FAST_LAUNCH
    return hipGetLastError();
}

template <int N>
static hipError_t launch_fast_n(hipStream_t st, int mode, int dir, const FastArgs &a, int var)
{
    if constexpr (kVariants<N> > 1) {
        if (var == 1) return launch_fast_nv<N, 1>(st, mode, dir, a);
        if (var == 2) return launch_fast_nv<N, 2>(st, mode, dir, a);
    }
    return launch_fast_nv<N, 0>(st, mode, dir, a);
}

template <int N>
static long long rows_per_block_n(int var)
{
    if constexpr (kVariants<N> > 1) {
        if (var == 1) return FShape<N, true, 1>::LPB;
        if (var == 2) return FShape<N, true, 2>::LPB;
    }
    return FShape<N, true, 0>::LPB;
}

template <int N>
static unsigned grid_n(long long I, long long O, int var)
{
    unsigned grid = 0, threads;
    size_t lds;
    const bool rows = (I == 1);
    if constexpr (kVariants<N> > 1) {
        if (var == 1) {
            fast_shape<N, 1>(rows, 2, I, O, 1, 0, grid, threads, lds);
            return grid;
        }
        if (var == 2) {
            fast_shape<N, 2>(rows, 2, I, O, 1, 0, grid, threads, lds);
            return grid;
        }
    }
    fast_shape<N, 0>(rows, 2, I, O, 1, 0, grid, threads, lds);
    return grid;
}

#define GP5_FAST_SWITCH(n, EXPR_N) \
    switch (n) {                   \
        case 16: return EXPR_N(16); \
        case 32: return EXPR_N(32); \
        case 64: return EXPR_N(64); \
        case 96: return EXPR_N(96); \
        case 128: return EXPR_N(128); \
        case 192: return EXPR_N(192); \
        case 256: return EXPR_N(256); \
        case 512: return EXPR_N(512); \
        default: break;             \
    }

long long fast_rows_per_block(int n, int var)
{
#define GP5_E(NN) rows_per_block_n<NN>(var)
    GP5_FAST_SWITCH(n, GP5_E)
#undef GP5_E
    return 1;
}

unsigned fast_grid(int n, long long I, long long O, int var)
{
#define GP5_E(NN) grid_n<NN>(I, O, var)
    GP5_FAST_SWITCH(n, GP5_E)
#undef GP5_E
    return 0;
}

hipError_t launch_fast(hipStream_t st, int n, int mode, int dir, const FastArgs &a, int var)
{
#define GP5_E(NN) launch_fast_n<NN>(st, mode, dir, a, var)
    GP5_FAST_SWITCH(n, GP5_E)
#undef GP5_E
    return hipErrorInvalidValue;
}

int fast_variants(int n)
{
#define GP5_E(NN) kVariants<NN>
    GP5_FAST_SWITCH(n, GP5_E)
#undef GP5_E
    return 0;
}

// "R0xR1xR2/E" of a plan, for tuning reports (role 0: x lines, 1: y columns, 2: z pass)
template <class P> static std::string plan_text()
{
    std::string t = std::to_string(P::R0);
    for (int r : {P::R1, P::R2, P::R3})
        if (r > 0) t += "x" + std::to_string(r);
    t += "/" + std::to_string(P::E);
    return t;
}

template <int N, int V> static std::string plan_name_nv(int role)
{
    if (role == 0) return plan_text<FPlanV<N, true, V>>();
    if (role == 1) return plan_text<FPlanV<N, false, V>>() + " cols " + std::to_string(FShape<N, false, V>::COLS);
    return plan_text<FPlanZV<N, V>>() + " cols " + std::to_string(FShapeP<FPlanZV<N, V>, false>::COLS);
}

template <int N> static std::string plan_name_n(int role, int var)
{
    if constexpr (kVariants<N> > 1) {
        if (var == 1) return plan_name_nv<N, 1>(role);
        if (var == 2) return plan_name_nv<N, 2>(role);
    }
    return plan_name_nv<N, 0>(role);
}

std::string fast_plan_name(int n, int role, int var)
{
#define GP5_E(NN) plan_name_n<NN>(role, var)
    GP5_FAST_SWITCH(n, GP5_E)
#undef GP5_E
    return "generic";
}
#undef GP5_FAST_SWITCH

}  // namespace gp5
