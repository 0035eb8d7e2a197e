// Kernels of the learned warm start (include/gp5_nn.h): 3x3x3 convolutions as
// implicit GEMMs on bf16 matrix-core tiles (v_mfma_f32_16x16x16_bf16, float
// accumulation), their data and weight gradients, and the elementwise pieces
// of the training step.
//
// MFMA 16x16x16 operand layout (wave64): lane l holds, for row/column
// r = l % 16 and k-group g = l / 16,
//   A[r][4g .. 4g+3]  (M x K, row r)      B[4g .. 4g+3][r]  (K x N, column r)
//   D[4g .. 4g+3][r]  (M x N, column r; 4 floats)
#include "../../include/gp5.h"
#include "../../include/gp5_nn.h"
#include "gp5_internal.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <mutex>
#include <vector>

namespace gp5 {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr float kLeak = 0.1f;
constexpr int kActYf0 = 2;  // act flag of the convolutions: the float output keeps channel 0 only
// block partials of a bias gradient: rows of C floats, summed in a fixed order
constexpr int kBiasBlocks = 2048;

// The stream that the calling thread's launches and copies go to
// (gp5_nn_set_stream; the null stream until set)
thread_local hipStream_t g_stream = nullptr;
inline hipStream_t nn_stream() { return g_stream; }

// Positions are linear indices p = ((b*D + z)*D + y)*D + x.

// 8-byte load of 4 consecutive bf16 (the caller guarantees 8-byte alignment:
// channel offsets are multiples of 4)
__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) { return *reinterpret_cast<const bf16x4 *>(p); }

// Raw buffer resource over [base, base + bytes).  A load whose lane offset is
// at or beyond `bytes` returns zero (the buffer range check), which is how the
// convolution reads its zero padding without branches; kOffNone is such an
// offset for every resource here (all are below 2 GiB).
constexpr int kOffNone = (int)0x80000000u;

#ifndef GP5_CONV_PD
#define GP5_CONV_PD 2
#endif
constexpr int kConvPD = GP5_CONV_PD;  // taps of loads in flight ahead of the MFMAs

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}

// LV consecutive bf16 (LV = 4, 8 or 16) at byte offset voff + soff, as LV/4
// MFMA fragments; voff + soff is a multiple of 2 LV bytes
template <int LV>
struct Frag {
    bf16x4 v[LV / 4];
};

template <int LV>
__device__ __forceinline__ Frag<LV> bld(__amdgpu_buffer_rsrc_t rs, int voff, int soff)
{
    Frag<LV> f;
    if constexpr (LV == 4) {
        f.v[0] = __builtin_bit_cast(bf16x4, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
    } else {
#pragma unroll
        for (int h = 0; h < LV / 8; ++h) {
            const auto q = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff + 16 * h, 0);
            f.v[2 * h] = __builtin_bit_cast(bf16x4, __builtin_shufflevector(q, q, 0, 1));
            f.v[2 * h + 1] = __builtin_bit_cast(bf16x4, __builtin_shufflevector(q, q, 2, 3));
        }
    }
    return f;
}

// Convolution epilogue: lane holds D[4g + i][r] of each block, i.e. position
// base + 16 t + 4g + i and output channel 16 (nb0 + j) + r: bias, leaky ReLU
// (act bit 0), then the float output Yf, the bf16 output A and the bf16 output Y
// (plus the skip input when given), each only when its pointer is set.  With
// act bit 1 (kActYf0) Yf holds channel 0 only, one float per position.
template <int NB, int MT>
__device__ __forceinline__ void conv_epilogue(const f32x4 (&acc)[MT][NB], int base, int nb0, int P, int Cout,
                                              const float *__restrict__ bias, int act, uint16_t *A,
                                              const uint16_t *skip, uint16_t *Y, float *Yf)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const bool yf0 = (act & kActYf0) != 0;
    act &= 1;
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int po = base + 16 * t + 4 * g + i;
            if (po >= P) continue;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int co = 16 * (nb0 + j) + r;
                float v = acc[t][j][i] + (bias ? bias[co] : 0.f);
                if (act == 1) v = v > 0.f ? v : kLeak * v;
                const size_t o = (size_t)po * Cout + co;
                if (Yf && !yf0) Yf[o] = v;
                if (Yf && yf0 && co == 0) Yf[po] = v;
                const uint16_t h = f2bf(v);
                if (A) A[o] = h;
                if (Y) Y[o] = skip ? f2bf(v + bf2f(skip[o])) : h;
            }
        }
}

// The same epilogue for the LDS kernels, through a per-wave LDS scratch of
// kEpiScratch bytes: two tiles (32 positions x 16 channels of one output block)
// at a time go to the scratch in float, in the MFMA layout, and come back one
// row piece per lane, so that the outputs are written, and the skip input read,
// with 16-byte accesses that cover 1 KB of consecutive memory per instruction.
// Same values as conv_epilogue.  Every wave of the workgroup must call it (it
// synchronises the workgroup).
constexpr int kEpiStride = 20;                       // floats per scratch row (16 channels, conflict-free)
constexpr int kEpiScratch = 32 * kEpiStride * 4;     // bytes per wave

// The outputs of 32 positions (16 channels cb .. cb + 15 of each) from the float
// scratch T (position-major rows of kEpiStride floats): rows 0..15 go to positions
// pb .. pb + 15, rows 16..31 to pb + 16 .. pb + 31, or with SPLIT to pb2 .. pb2 + 15.  Yf (all channels, or channel 0 only with yf0), A = bf16(v) and
// Y = bf16(v + skip) or bf16(v), each when set, with 16-byte accesses (called by one
// whole wave).
template <bool SPLIT = false>
__device__ __forceinline__ void epi_write32(const float *T, int pb, int P, int Cout, int cb, int lane, uint16_t *A,
                                            const uint16_t *skip, uint16_t *Y, float *Yf, bool yf0, int pb2 = 0)
{
    auto at = [&](int pos) { return SPLIT ? (pos < 16 ? pb + pos : pb2 + pos - 16) : pb + pos; };
    if (Yf && !yf0) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int pos = lane / 4 + 16 * k, c4 = 4 * (lane & 3);
            if (at(pos) < P)
                *reinterpret_cast<float4 *>(Yf + (size_t)at(pos) * Cout + cb + c4) =
                    *reinterpret_cast<const float4 *>(T + pos * kEpiStride + c4);
        }
    } else if (Yf && cb == 0 && lane < 32) {  // channel 0 of the 32 positions
        if (at(lane) < P) Yf[at(lane)] = T[lane * kEpiStride];
    }
    if (A || Y) {
        const int pos = lane / 2, c8 = 8 * (lane & 1);
        if (at(pos) < P) {
            const float4 lo = *reinterpret_cast<const float4 *>(T + pos * kEpiStride + c8);
            const float4 hi = *reinterpret_cast<const float4 *>(T + pos * kEpiStride + c8 + 4);
            const float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            const size_t o = (size_t)at(pos) * Cout + cb + c8;
            uint16_t hb[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) hb[e] = f2bf(v[e]);
            if (A) *reinterpret_cast<uint4 *>(A + o) = __builtin_bit_cast(uint4, hb);
            if (Y) {
                if (skip) {
                    const uint4 sk4 = *reinterpret_cast<const uint4 *>(skip + o);
                    const uint16_t *sk = reinterpret_cast<const uint16_t *>(&sk4);
                    uint16_t yb[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) yb[e] = f2bf(v[e] + bf2f(sk[e]));
                    *reinterpret_cast<uint4 *>(Y + o) = __builtin_bit_cast(uint4, yb);
                } else {
                    *reinterpret_cast<uint4 *>(Y + o) = __builtin_bit_cast(uint4, hb);
                }
            }
        }
    }
}

// With WS the scratch is wave-private and the exchange needs no workgroup
// barrier: a wave's LDS writes and reads execute in issue order, and the wave
// barrier keeps the compiler from moving them across each other.  The caller
// then owns the workgroup barrier before the scratch is reused by other waves.
//
// ABWD (the data gradient of a layer followed by the backward of the leaky ReLU
// below it): A is that layer's activation (read), Y receives H = bf16(v act'(A))
// in the arithmetic of k_act_bwd_cast_bias, and bs[j][e] (this lane's channels
// 16 (nb0 + j) + 8 (lane & 1) + e) adds up v act'(A) over its positions.
template <int NB, int MT, bool WS = false, bool ABWD = false>
__device__ __forceinline__ void conv_epilogue_lds_impl(const f32x4 (&acc)[MT][NB], int base, int nb0, int P, int Cout,
                                                       const float *__restrict__ bias, int act, uint16_t *A,
                                                       const uint16_t *skip, uint16_t *Y, float *Yf, unsigned char *scr,
                                                       float (&bs)[NB][8])
{
    static_assert(MT % 2 == 0, "tiles go through the scratch in pairs");
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const bool yf0 = (act & kActYf0) != 0;
    act &= 1;
    float *T = reinterpret_cast<float *>(scr);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int cb = 16 * (nb0 + j);
        const float bco = bias ? bias[cb + r] : 0.f;
#pragma unroll
        for (int h = 0; h < MT / 2; ++h) {
#pragma unroll
            for (int tt = 0; tt < 2; ++tt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float v = acc[2 * h + tt][j][i] + bco;
                    if (act == 1) v = v > 0.f ? v : kLeak * v;
                    T[(16 * tt + 4 * g + i) * kEpiStride + r] = v;
                }
            if constexpr (WS) __builtin_amdgcn_wave_barrier();
            else __syncthreads();
            const int pb = base + 32 * h;  // first position of the pair
            if constexpr (ABWD) {
                epi_write32(T, pb, P, Cout, cb, lane, nullptr, nullptr, nullptr, Yf, yf0);
                const int pos = lane / 2, c8 = 8 * (lane & 1);
                if (pb + pos < P) {
                    const float4 lo = *reinterpret_cast<const float4 *>(T + pos * kEpiStride + c8);
                    const float4 hi = *reinterpret_cast<const float4 *>(T + pos * kEpiStride + c8 + 4);
                    const float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
                    const size_t o = (size_t)(pb + pos) * Cout + cb + c8;
                    const uint4 a4 = *reinterpret_cast<const uint4 *>(A + o);
                    const uint16_t *ah = reinterpret_cast<const uint16_t *>(&a4);
                    uint16_t hb[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float ov = v[e] * (bf2f(ah[e]) > 0.f ? 1.f : kLeak);
                        hb[e] = f2bf(ov);
                        bs[j][e] += ov;
                    }
                    *reinterpret_cast<uint4 *>(Y + o) = __builtin_bit_cast(uint4, hb);
                }
            } else {
                epi_write32(T, pb, P, Cout, cb, lane, A, skip, Y, Yf, yf0);
            }
            if constexpr (WS) __builtin_amdgcn_wave_barrier();  // the scratch is rewritten by the next pair
            else __syncthreads();
        }
    }
}

template <int NB, int MT, bool WS = false>
__device__ __forceinline__ void conv_epilogue_lds(const f32x4 (&acc)[MT][NB], int base, int nb0, int P, int Cout,
                                                  const float *__restrict__ bias, int act, uint16_t *A,
                                                  const uint16_t *skip, uint16_t *Y, float *Yf, unsigned char *scr)
{
    float bs[NB][8];  // unused without ABWD
    conv_epilogue_lds_impl<NB, MT, WS, false>(acc, base, nb0, P, Cout, bias, act, A, skip, Y, Yf, scr, bs);
}

// Forward / data-gradient convolution.  A workgroup of 4 waves covers 64 * MT
// consecutive positions: each wave holds MT tiles of 16 positions that share
// every weight fragment it loads, and NB 16-wide blocks of output channels.
//
// The reduction over K = 27 taps x Cin channels runs in chunks of KC channels
// (KC = 16, 32 or 64 divides Cin).  Within a chunk, lane (r, g) loads the KC/4
// consecutive channels [g KC/4, (g+1) KC/4) of its position (A) and of its output
// channel (B) in one contiguous piece, and MFMA step s takes channels
// g KC/4 + 4s .. +3 from both: a permutation of K that A and B share, so that
// the 4 lanes of a position read its KC channels as one run and a load
// instruction covers whole rows of 16 positions.
//
// Positions are decoded once; a tap is a scalar byte offset on a resource that
// starts (D^2 + D + 1) Cin elements before X (the most negative tap), so every
// tap offset is non-negative, and a neighbour outside the cube gets the lane
// offset kOffNone (reads zero).  Offsets are 32-bit: the host keeps every tensor
// below 2 GiB.  The loads run PD taps ahead of the MFMAs.  Each accumulator sees
// its MFMAs in the same (chunk, tap, step) order whatever NB, MT and PD are.
template <int KC, int NB, int MT, int PD>
__global__ __launch_bounds__(256) void k_conv(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wk,
                                              const float *__restrict__ bias, uint16_t *A, const uint16_t *skip,
                                              uint16_t *Y, float *Yf, int P, int D, int Cin, int Cout, int mirror,
                                              int act)
{
    constexpr int LV = KC / 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int nb0 = blockIdx.y * NB;  // first 16-wide output block of this group
    const int base = ((int)blockIdx.x * 4 + wave) * (16 * MT);
    const int halo = (D * D + D + 1) * Cin * 2;
    const __amdgpu_buffer_rsrc_t xr = make_rsrc((const char *)X - halo, halo + P * Cin * 2);
    const __amdgpu_buffer_rsrc_t wr = make_rsrc(Wk, 27 * Cout * Cin * 2);
    int vo[MT];
    bool xm[MT], xp[MT], ym[MT], yp[MT], zm[MT], zp[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int p = base + 16 * t + r;  // A row of this lane in tile t
        const unsigned q1 = (unsigned)p / (unsigned)D, q2 = q1 / (unsigned)D;
        const int x = p - (int)q1 * D, y = (int)(q1 - q2 * D), z = (int)(q2 % (unsigned)D);
        vo[t] = p < P ? (p * Cin + g * LV) * 2 : kOffNone;
        xm[t] = x > 0;
        xp[t] = x < D - 1;
        ym[t] = y > 0;
        yp[t] = y < D - 1;
        zm[t] = z > 0;
        zp[t] = z < D - 1;
    }
    const int wvo = ((nb0 * 16 + r) * Cin + g * LV) * 2;
    f32x4 acc[MT][NB];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // fragments of tap `tap` into ring slot `slot`
    Frag<LV> a[PD + 1][MT], b[PD + 1][NB];
    auto fetch = [&](int c0, int tap, int slot) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const int soff = halo + ((dz * D + dy) * D + dx) * Cin * 2 + 2 * c0;
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const bool ok = (dx < 0 ? xm[t] : dx > 0 ? xp[t] : true) && (dy < 0 ? ym[t] : dy > 0 ? yp[t] : true) &&
                            (dz < 0 ? zm[t] : dz > 0 ? zp[t] : true);
            a[slot][t] = bld<LV>(xr, ok ? vo[t] : kOffNone, soff);
        }
        const int wso = ((mirror ? 26 - tap : tap) * Cout * Cin + c0) * 2;
#pragma unroll
        for (int j = 0; j < NB; ++j) b[slot][j] = bld<LV>(wr, wvo, wso + j * 16 * Cin * 2);
    };
    for (int c0 = 0; c0 < Cin; c0 += KC) {
        // keeps the 27 MT per-tap offsets from being hoisted out of this loop
        // into registers (they cost one select each where they are used)
#pragma unroll
        for (int t = 0; t < MT; ++t) asm volatile("" : "+v"(vo[t]));
        // software pipeline: the loads of tap + PD are in flight during the MFMAs of tap
#pragma unroll
        for (int tap = 0; tap < PD; ++tap) fetch(c0, tap, tap);
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            if (tap + PD < 27) fetch(c0, tap + PD, (tap + PD) % (PD + 1));
            const int sl = tap % (PD + 1);
#pragma unroll
            for (int st = 0; st < LV / 4; ++st)
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int t = 0; t < MT; ++t)
                        acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[sl][t].v[st], b[sl][j].v[st], acc[t][j],
                                                                              0, 0, 0);
        }
    }
    conv_epilogue<NB, MT>(acc, base, nb0, P, Cout, bias, act, A, skip, Y, Yf);
}

// n / d for 0 <= n < 2^22 and 1 <= d < 2^11, with inv = 1 / d: the float
// quotient is within one of the true one, and the two corrections make it exact
__device__ __forceinline__ int udiv_small(int n, int d, float inv)
{
    int q = (int)((float)n * inv);
    q -= q * d > n;
    q += (q + 1) * d <= n;
    return q;
}

constexpr int kStageBatch = 8;  // staging loads in flight per thread

// Bank swizzle of staged positions of KC = 16 or 32 channels (2 or 4 units of 16
// bytes): unit c of the position in column ix sits in slot c ^ swz_unit(ix).  A
// wave reads 16 consecutive columns; with the plain layout the columns 8 apart
// (KC = 16, 8-byte reads) or 4 apart (KC = 32, 16-byte reads) hit the same banks,
// and the swizzle gives each of them a different slot.
template <int KC>
__device__ __forceinline__ int swz_unit(int ix)
{
    return KC == 16 ? (ix >> 3) & 1 : (ix >> 2) & 3;
}

// byte offset, within its position, of lane group g's fragment (LV = KC / 4 channels)
template <int KC>
__device__ __forceinline__ int swz_frag(int g, int ix)
{
    constexpr int FB = KC / 2;  // fragment bytes
    return 16 * (((g * FB) >> 4) ^ swz_unit<KC>(ix)) + ((g * FB) & 15);
}

// A fragment from LDS at byte offset off.  The offset goes through an empty asm
// so that the compiler cannot see the common base of neighbouring taps: it would
// pair the reads of x-neighbours (32 or 64 bytes apart) into ds_read2 pairs, in
// which each lane's second read hits the banks of the next lane's first (2-way
// conflicts, measured 40 % of the kernel time); single reads of 16 consecutive
// positions cover 512 or 1024 consecutive bytes without conflicts.
// VOL = false uses a non-volatile asm: the offset stays opaque (no pairing) but
// the read may be scheduled freely, e.g. one tap ahead of its MFMAs.
template <int LV, bool VOL = true>
__device__ __forceinline__ Frag<LV> lds_frag(const unsigned char *sx, int off)
{
    if constexpr (VOL) asm volatile("" : "+v"(off));
    else asm("" : "+v"(off));
    if constexpr (LV == 4) {
        return *reinterpret_cast<const Frag<LV> *>(sx + off);
    } else {  // 16-byte pieces, 16-byte aligned (off is a multiple of 2 LV)
        const uint4 *q = reinterpret_cast<const uint4 *>(__builtin_assume_aligned(sx + off, 16));
        Frag<LV> f;
#pragma unroll
        for (int h = 0; h < LV / 8; ++h) {
            const uint4 v = q[h];
            f.v[2 * h] = __builtin_bit_cast(bf16x4, make_uint2(v.x, v.y));
            f.v[2 * h + 1] = __builtin_bit_cast(bf16x4, make_uint2(v.z, v.w));
        }
        return f;
    }
}

// LDS-staged convolution for cubes of side D = 8 .. 128, a power of two.  A
// workgroup of 4 waves owns 256 consecutive positions: Rb = min(D, 256 / D)
// whole rows of PZ = 256 / (D Rb) planes of one sample.  Per chunk of KC input
// channels it stages the input region those positions read, (PZ + 2) planes x
// (Rb + 2) rows x (D + 2) columns x KC channels with zeros outside the cube,
// into LDS with 16-byte loads (each input element is fetched about 4.5 times per
// output instead of 27), then every wave runs 4 tiles of 16 positions x NB
// 16-wide output blocks over the 27 taps, A fragments from LDS and weight
// fragments from global memory PD taps ahead.  Same K permutation and
// accumulation order as k_conv (chunk, tap, step), so for the same KC the two
// kernels give bit-identical results.  With up = 1, X has side D/2 and each
// staged cell is read from its parent cell (the nearest upsampling of X).
//
// Split over channel chunks (csplit > 0, grid z = the split): workgroup z takes
// the chunks [z csplit, (z + 1) csplit) and writes its raw float sums (no bias,
// no activation) to Yf + z P Cout; k_conv_split_finish adds the splits in order.
template <int KC, int NB, int PD>
__global__ __launch_bounds__(256) void k_conv_lds(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wk,
                                                  const float *__restrict__ bias, uint16_t *A, const uint16_t *skip,
                                                  uint16_t *Y, float *Yf, int P, int D, int Cin, int Cout, int mirror,
                                                  int act, int Rb, int PZ, int up, int csplit)
{
    constexpr int MT = 4, LV = KC / 4, U8 = KC / 8;  // U8: 16-byte units per staged position
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int nb0 = blockIdx.y * NB;
    const int RX = D + 2, RY = Rb + 2, RZ = PZ + 2;
    const int UR = RX * U8, U = RZ * RY * UR;  // 16-byte units per region row, in the region
    const float invUR = 1.f / (float)UR, invRY = 1.f / (float)RY;
    const int p0 = (int)blockIdx.x * 256;
    const int y0 = (p0 / D) % D, z0 = (p0 / (D * D)) % D;
    // up = 1: X is stored at side Ds = D/2 and read at the parent cell (nearest upsampling)
    const int Ds = D >> up;
    const int sb = p0 / (D * D * D) * Ds * Ds * Ds;  // stored position of this sample's origin
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(X, (P >> (3 * up)) * Cin * 2);
    const __amdgpu_buffer_rsrc_t wr = make_rsrc(Wk, 27 * Cout * Cin * 2);
    int lo[MT][3];  // LDS byte offset of this lane's fragment in tile t at taps (0, 0, dx = d - 1)
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int q = 64 * wave + 16 * t + r;
        const int lx = q & (D - 1), ly = (q / D) % Rb, lz = q / (D * Rb);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int ix = lx + d;  // column of the neighbour dx = d - 1 (column 0 is x = -1)
            lo[t][d] = (((lz + 1) * RY + ly + 1) * RX + ix) * KC * 2 + swz_frag<KC>(g, ix);
        }
    }
    const int wvo = ((nb0 * 16 + r) * Cin + g * LV) * 2;
    f32x4 acc[MT][NB];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    Frag<LV> b[PD + 1][NB];
    auto fetch_w = [&](int c0, int tap, int slot) {
        const int wso = ((mirror ? 26 - tap : tap) * Cout * Cin + c0) * 2;
#pragma unroll
        for (int j = 0; j < NB; ++j) b[slot][j] = bld<LV>(wr, wvo, wso + j * 16 * Cin * 2);
    };
    const int cb = csplit ? (int)blockIdx.z * csplit * KC : 0, ce = csplit ? cb + csplit * KC : Cin;
    for (int c0 = cb; c0 < ce; c0 += KC) {
        if (c0 != cb) __syncthreads();  // every wave is done with the previous chunk
        // keeps the 27 MT tap offsets from being hoisted out of this loop into registers
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int d = 0; d < 3; ++d) asm volatile("" : "+v"(lo[t][d]));
        // stage the region, unit u = 16 bytes at LDS byte 16 u, row-major over
        // (plane, row, column, 8-channel group); kStageBatch loads in flight per
        // thread before their LDS writes
        for (int u0 = tid; u0 < U; u0 += 256 * kStageBatch) {
            uint4 v[kStageBatch];
            int us[kStageBatch];  // swizzled unit of the LDS write
#pragma unroll
            for (int k = 0; k < kStageBatch; ++k) {
                const int u = u0 + 256 * k;
                const int row = udiv_small(u, UR, invUR), e = u - row * UR;
                const int iz = udiv_small(row, RY, invRY), iy = row - iz * RY;
                const int ix = e / U8, c8 = e - ix * U8;
                us[k] = u - c8 + (c8 ^ swz_unit<KC>(ix));
                const int y = y0 - 1 + iy, z = z0 - 1 + iz;
                const bool in = u < U && y >= 0 && y < D && z >= 0 && z < D && ix >= 1 && ix <= D;
                const int voff =
                    in ? ((sb + ((z >> up) * Ds + (y >> up)) * Ds + ((ix - 1) >> up)) * Cin + c0 + 8 * c8) * 2 : kOffNone;
                v[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, voff, 0, 0));
            }
#pragma unroll
            for (int k = 0; k < kStageBatch; ++k)
                if (u0 + 256 * k < U) *reinterpret_cast<uint4 *>(sx + 16 * us[k]) = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < PD; ++tap) fetch_w(c0, tap, tap);
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            if (tap + PD < 27) fetch_w(c0, tap + PD, (tap + PD) % (PD + 1));
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            const int toff = (dz * RY + dy) * RX * KC * 2;
            Frag<LV> a[MT];
#pragma unroll
            for (int t = 0; t < MT; ++t) a[t] = lds_frag<LV>(sx, lo[t][dx + 1] + toff);
            const int sl = tap % (PD + 1);
#pragma unroll
            for (int st = 0; st < LV / 4; ++st)
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int t = 0; t < MT; ++t)
                        acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t].v[st], b[sl][j].v[st], acc[t][j],
                                                                              0, 0, 0);
        }
    }
    __syncthreads();  // the staged region becomes the epilogue scratch
    if (csplit)
        conv_epilogue_lds<NB, MT>(acc, p0 + 64 * wave, nb0, P, Cout, nullptr, 0, nullptr, nullptr, nullptr,
                                  Yf + (size_t)blockIdx.z * P * Cout, sx + wave * kEpiScratch);
    else
        conv_epilogue_lds<NB, MT>(acc, p0 + 64 * wave, nb0, P, Cout, bias, act, A, skip, Y, Yf,
                                  sx + wave * kEpiScratch);
}

// The splits of k_conv_lds added in order (z = 0, 1, ...), then the epilogue of
// conv_epilogue: bias, activation (act bit 0; bit 1 keeps channel 0 of Yf), the
// float output Yf, the bf16 output A and Y (plus skip); 4 channels per thread.
__global__ __launch_bounds__(256) void k_conv_split_finish(const float *__restrict__ part, int splits, int P, int Cout,
                                                           const float *__restrict__ bias, int act, uint16_t *A,
                                                           const uint16_t *skip, uint16_t *Y, float *Yf)
{
    const bool yf0 = (act & kActYf0) != 0;
    act &= 1;
    const int C4 = Cout / 4;
    const long long n = (long long)P * C4;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        float4 v = reinterpret_cast<const float4 *>(part)[i];
        for (int z = 1; z < splits; ++z) {
            const float4 w = reinterpret_cast<const float4 *>(part + (size_t)z * P * Cout)[i];
            v = make_float4(v.x + w.x, v.y + w.y, v.z + w.z, v.w + w.w);
        }
        const int c = (int)(i % C4) * 4;
        float o[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            o[k] += bias ? bias[c + k] : 0.f;
            if (act == 1) o[k] = o[k] > 0.f ? o[k] : kLeak * o[k];
        }
        if (Yf && !yf0) reinterpret_cast<float4 *>(Yf)[i] = make_float4(o[0], o[1], o[2], o[3]);
        if (Yf && yf0 && c == 0) Yf[i / C4] = o[0];
        uint16_t h[4] = {f2bf(o[0]), f2bf(o[1]), f2bf(o[2]), f2bf(o[3])};
        if (A) reinterpret_cast<uint2 *>(A)[i] = __builtin_bit_cast(uint2, h);
        if (Y) {
            if (skip) {
                const uint2 s2 = reinterpret_cast<const uint2 *>(skip)[i];
                const uint16_t *sk = reinterpret_cast<const uint16_t *>(&s2);
#pragma unroll
                for (int k = 0; k < 4; ++k) h[k] = f2bf(o[k] + bf2f(sk[k]));
            }
            reinterpret_cast<uint2 *>(Y)[i] = __builtin_bit_cast(uint2, h);
        }
    }
}

// slab units (16 bytes) per thread of k_conv_zs: a slab has at most
// 4 x 130 x CIN/8 units (D = 128)
__host__ __device__ constexpr int zs_units_per_thread(int cin) { return cin == 16 ? 5 : 9; }

// z-sliding convolution for inputs of CIN = 16 or 32 channels on cubes of side
// D = 16 .. 128 (a power of two).  A workgroup of 4 waves owns Rb = 256 / D
// rows of ZL consecutive planes of one sample and walks them in z.  LDS holds a
// ring of three slabs, one per plane z-1, z, z+1, each (Rb + 2) rows x (D + 2)
// columns x CIN channels with zeros outside the cube.  While the waves run the
// 27 taps of plane z, the global loads of the slab of plane z+2 are in flight
// in registers; they go to the ring slot of plane z-1 once every wave is done
// with it.  Each input slab is thus fetched once per plane instead of three
// times, and its latency overlaps the MFMAs.  All 27 weight fragments of the
// wave's NB output blocks stay in registers for the whole walk.  Same K
// permutation and accumulation order (tap, step) as k_conv with KC = CIN, so
// the two agree bit for bit.  With up = 1, X has side D/2 and each slab cell is
// read from its parent cell (the nearest upsampling of X).  PF reads the
// fragments of tap + 1 while the MFMAs of tap run (scheduling barriers keep
// the compiler from sinking the reads to their uses); WS gives the epilogue
// wave-private barriers.  Neither changes any accumulation order.
//
// POOL (D = 32, 64 or 128, no bias or activation, float output only): Yf receives,
// at side D/2, the sum over the 2x2x2 children of each parent cell (the
// adjoint of nearest upsampling applied to the convolution), in the order
// k = 0..7 of k_pool2 (z, y, x bits), so that it equals the convolution
// followed by gp5_nn_pool2_f32 with scale 8 bit for bit.  A parent's x pair
// lies in one lane's registers, its y pair in one wave at D = 32 (tiles t and
// t + 2) and in two waves at D >= 64 (the even row's wave owns the parent and
// reads the odd row's values through LDS), its z pair in consecutive planes of
// the walk (the even plane's sum waits in registers).
//
// ABWD (no bias, activation or skip): the epilogue also forms the backward of the
// leaky ReLU below (conv_epilogue_lds_impl: A read, H = Y written), and at the end
// workgroup x writes its bias-gradient partial, the sum over its positions, to row
// x of bpart (kBiasBlocks rows of Cout floats, gridDim.x <= kBiasBlocks) and zeros
// to rows x + gridDim.x, x + 2 gridDim.x, ..., so that the rows add up to the bias
// gradient in k_bias_final's fixed order.
template <int CIN, int NB, bool PF = false, bool WS = false, bool POOL = false, bool ABWD = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ABWD && CIN == 16 ? 3 : 2))) void k_conv_zs(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wk,
                                                 const float *__restrict__ bias, uint16_t *A, const uint16_t *skip,
                                                 uint16_t *Y, float *Yf, int P, int D, int Cout, int mirror, int act,
                                                 int nzs, int up, float *bpart)
{
    static_assert(!(ABWD && POOL), "one epilogue form at a time");
    constexpr int MT = 4, LV = CIN / 4, U8 = CIN / 8;
    constexpr int NU = zs_units_per_thread(CIN);
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int nb0 = blockIdx.y * NB;
    const int Rb = 256 / D, RY = Rb + 2, RX = D + 2;
    const int UR = RX * U8, SU = RY * UR;  // 16-byte units per slab row, per slab
    const float invUR = 1.f / (float)UR;
    const int nyb = D / Rb;
    const int yb = (int)blockIdx.x % nyb, zsg = ((int)blockIdx.x / nyb) % nzs, smp = (int)blockIdx.x / (nyb * nzs);
    // nzs runs of planes per cube side, ZL = D / nzs or one more plane per run
    const int y0 = yb * Rb, zb = zsg * D / nzs, ZL = (zsg + 1) * D / nzs - zb;
    const int s0 = smp * D * D * D;
    // up = 1: X is stored at side Ds = D/2 and read at the parent cell (nearest upsampling)
    const int Ds = D >> up;
    const int sb = smp * Ds * Ds * Ds;
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(X, (P >> (3 * up)) * CIN * 2);
    const __amdgpu_buffer_rsrc_t wr = make_rsrc(Wk, 27 * Cout * CIN * 2);
    int lo[MT][3];  // LDS byte offset of this lane's fragment in tile t at taps (0, 0, dx = d - 1), slot 0
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int q = 64 * wave + 16 * t + r;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int ix = (q & (D - 1)) + d;
            lo[t][d] = (((q / D) + 1) * RX + ix) * CIN * 2 + swz_frag<CIN>(g, ix);
        }
    }
    // weights of all taps
    Frag<LV> w[27][NB];
    {
        const int wvo = ((nb0 * 16 + r) * CIN + g * LV) * 2;
#pragma unroll
        for (int tap = 0; tap < 27; ++tap)
#pragma unroll
            for (int j = 0; j < NB; ++j)
                w[tap][j] = bld<LV>(wr, wvo, ((mirror ? 26 - tap : tap) * Cout * CIN + j * 16 * CIN) * 2);
    }
    uint4 v[NU];
    int us[NU];  // swizzled unit of the LDS write
    auto load_slab = [&](int z) {
#pragma unroll
        for (int k = 0; k < NU; ++k) {
            const int u = tid + 256 * k;
            const int row = udiv_small(u, UR, invUR), e = u - row * UR;
            const int ix = e / U8, c8 = e - ix * U8;
            us[k] = u - c8 + (c8 ^ swz_unit<CIN>(ix));
            const int y = y0 - 1 + row;
            const bool in = u < SU && z >= 0 && z < D && y >= 0 && y < D && ix >= 1 && ix <= D;
            const int voff =
                in ? ((sb + ((z >> up) * Ds + (y >> up)) * Ds + ((ix - 1) >> up)) * CIN + 8 * c8) * 2 : kOffNone;
            v[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, voff, 0, 0));
        }
    };
    auto store_slab = [&](int slot) {
        unsigned char *dst = sx + (size_t)slot * SU * 16;
#pragma unroll
        for (int k = 0; k < NU; ++k)
            if (tid + 256 * k < SU) *reinterpret_cast<uint4 *>(dst + 16 * us[k]) = v[k];
    };
    load_slab(zb - 1);
    store_slab(0);
    load_slab(zb);
    store_slab(1);
    load_slab(zb + 1);
    store_slab(2);
    __syncthreads();
    float S[POOL ? MT : 1][POOL ? NB : 1][2];  // POOL: the even plane's partial parent sums
    float bs[NB][8];                           // ABWD: this lane's bias-gradient sums
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int e = 0; e < 8; ++e) bs[j][e] = 0.f;
    for (int zi = 0; zi < ZL; ++zi) {
        const int z = zb + zi;
        // keeps the 27 MT tap offsets from being hoisted out of this loop into registers
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int d = 0; d < 3; ++d) asm volatile("" : "+v"(lo[t][d]));
        if (zi + 1 < ZL) load_slab(z + 2);  // in flight during this plane's MFMAs
        f32x4 acc[MT][NB];
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        // slots of planes z-1, z, z+1
        const int sl[3] = {zi % 3, (zi + 1) % 3, (zi + 2) % 3};
        auto load_a = [&](int tap, Frag<LV> (&dst)[MT]) {
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            const int toff = sl[dz + 1] * SU * 16 + dy * RX * CIN * 2;
#pragma unroll
            for (int t = 0; t < MT; ++t) dst[t] = lds_frag<LV, !PF>(sx, lo[t][dx + 1] + toff);
        };
        auto mfmas = [&](int tap, const Frag<LV> (&a)[MT]) {
#pragma unroll
            for (int st = 0; st < LV / 4; ++st)
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int t = 0; t < MT; ++t)
                        acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t].v[st], w[tap][j].v[st], acc[t][j],
                                                                              0, 0, 0);
        };
        if constexpr (PF) {
            // the fragments of tap + 1 are read while the MFMAs of tap run
            Frag<LV> a[2][MT];
            load_a(0, a[0]);
#pragma unroll
            for (int tap = 0; tap < 27; ++tap) {
                if (tap + 1 < 27) load_a(tap + 1, a[(tap + 1) & 1]);
                // keeps the scheduler from pulling the reads of tap + 1 down to their MFMAs
                __builtin_amdgcn_sched_barrier(0);
                mfmas(tap, a[tap & 1]);
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int tap = 0; tap < 27; ++tap) {
                Frag<LV> a[MT];
                load_a(tap, a);
                mfmas(tap, a);
            }
        }
        __syncthreads();  // every wave is done with the slab of plane z-1: it becomes the epilogue scratch
        if constexpr (POOL) {
          if (D == 32) {  // a wave holds rows 2w (tiles 0, 1) and 2w + 1 (tiles 2, 3): its own y pairs
            const int Dh = 16, yp = (y0 + 2 * wave) / 2;
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const float a0 = acc[t][j][2 * h], a1 = acc[t][j][2 * h + 1];
                        const float b0 = acc[t + 2][j][2 * h], b1 = acc[t + 2][j][2 * h + 1];
                        if ((zi & 1) == 0) {
                            S[t][j][h] = ((a0 + a1) + b0) + b1;
                        } else {
                            const float v = (((S[t][j][h] + a0) + a1) + b0) + b1;
                            const int pp = ((smp * Dh + z / 2) * Dh + yp) * Dh + (16 * t + 4 * g) / 2 + h;
                            Yf[(size_t)pp * Cout + 16 * (nb0 + j) + r] = v;
                        }
                    }
          } else {
            float *xs = reinterpret_cast<float *>(sx + (size_t)sl[0] * SU * 16);  // 8 KB of the free slot
            const int wpr = D / 64;                                              // waves per row
            const bool owner = ((wave / wpr) & 1) == 0;                          // even row
            const int pair = wave / (2 * wpr) * wpr + wave % wpr;
            const int Dh = D / 2;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (!owner)
#pragma unroll
                    for (int t = 0; t < MT; ++t)
#pragma unroll
                        for (int i = 0; i < 4; ++i) xs[((pair * 16) + t * 4 + i) * 64 + lane] = acc[t][j][i];
                __syncthreads();
                if (owner) {
#pragma unroll
                    for (int t = 0; t < MT; ++t) {
                        const int q0 = 64 * wave + 16 * t;  // first plane-local position of the tile
                        const int yp = (y0 + q0 / D) / 2, xb = (q0 % D + 4 * g) / 2;
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const float a0 = acc[t][j][2 * h], a1 = acc[t][j][2 * h + 1];
                            const float b0 = xs[((pair * 16) + t * 4 + 2 * h) * 64 + lane];
                            const float b1 = xs[((pair * 16) + t * 4 + 2 * h + 1) * 64 + lane];
                            if ((zi & 1) == 0) {
                                S[t][j][h] = ((a0 + a1) + b0) + b1;
                            } else {
                                const float v = (((S[t][j][h] + a0) + a1) + b0) + b1;
                                const int pp = ((smp * Dh + z / 2) * Dh + yp) * Dh + xb + h;
                                Yf[(size_t)pp * Cout + 16 * (nb0 + j) + r] = v;
                            }
                        }
                    }
                }
                __syncthreads();
            }
          }
        } else {
            conv_epilogue_lds_impl<NB, MT, WS, ABWD>(acc, s0 + (z * D + y0) * D + 64 * wave, nb0, P, Cout, bias, act,
                                                     A, skip, Y, Yf, sx + (size_t)sl[0] * SU * 16 + wave * kEpiScratch,
                                                     bs);
        }
        if (zi + 1 < ZL) {
            if constexpr (WS) __syncthreads();  // every wave is done with its scratch in the slot
            store_slab(sl[0]);
            __syncthreads();
        }
    }
    if constexpr (ABWD) {
        // channel c of block j: the lanes with (lane & 1) == c / 8 of the four waves,
        // in wave and lane order
        __syncthreads();  // every wave is done with the LDS
        float *red = reinterpret_cast<float *>(sx);  // NB x 8 x 256 floats
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int e = 0; e < 8; ++e) red[(j * 8 + e) * 256 + tid] = bs[j][e];
        __syncthreads();
        if (tid < 16 * NB) {
            const int j = tid / 16, c = tid % 16;
            const float *rr = red + (j * 8 + c % 8) * 256;
            float t = 0.f;
            for (int w = 0; w < 4; ++w)
                for (int l = c / 8; l < 64; l += 2) t += rr[64 * w + l];
            const int co = 16 * (nb0 + j) + c;
            bpart[(size_t)blockIdx.x * Cout + co] = t;
            for (int row = (int)blockIdx.x + (int)gridDim.x; row < kBiasBlocks; row += (int)gridDim.x)
                bpart[(size_t)row * Cout + co] = 0.f;
        }
    }
}

// Convolution of the nearest upsampling of X (side Dc = D/2, CIN = 32 channels)
// by output parity.  Fine position 2q + p (p in {0,1}^3) sees, along each axis,
// the coarse cells q + o with o in {p - 1, p}: the 3x3x3 fine taps t fold into
// 2x2x2 coarse taps with weights W_p[o] = sum of W[t] over floor((p + t) / 2) = o.
// A workgroup owns one (py, pz) and both x parities of 16 output channels: its
// 12 coarse taps (oz, oy in the window, ox = -1, 0, 1) make 16 (tap, px) pairs
// against 27 taps for each fine tile, and every A fragment (16 coarse positions
// x 32 channels, one 16x16x32 MFMA operand) serves one or both x parities.  The
// folded weights are summed in float from the bf16 ones and kept as bf16 hi + lo
// (hi = bf16(s), lo = bf16(s - hi)), two MFMAs per pair, so that the result
// agrees with the direct convolution to float rounding.  The walk over coarse
// planes, the three-slab ring and the epilogue follow k_conv_zs (side Dc, no
// upsampling in the staging); the epilogue interleaves the two x parities into
// 32 consecutive fine positions.
// f(std::integral_constant<int, K>) for K = 0 .. N - 1: indices known at compile time
template <typename F, int... Ks>
__device__ __forceinline__ void static_for(F &&f, std::integer_sequence<int, Ks...>)
{
    (f(std::integral_constant<int, Ks>{}), ...);
}

constexpr int up2p_pair(int k, int px)  // index of (tap k, x parity px) among the 16 pairs, or -1
{
    int n = 0;
    for (int kk = 0; kk < k; ++kk) n += (kk % 3 == 1) ? 2 : 1;
    const int ox = k % 3 - 1;
    if (ox == -1) return px == 0 ? n : -1;
    if (ox == 1) return px == 1 ? n : -1;
    return n + px;
}

template <int CIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv_up2p(
    const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wk, const float *__restrict__ bias, uint16_t *A,
    const uint16_t *skip, uint16_t *Y, float *Yf, int P, int D, int Cout, int act, int nzs)
{
    static_assert(CIN == 32 || CIN == 64 || CIN == 128, "32-channel 16x16x32 operands");
    typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
    // MT = 2 tiles of 16 coarse positions per wave: the 32 folded weight fragments
    // of one 32-channel part take 128 VGPRs.  CIN = 64 and 128 split the channels over
    // KS = CIN / 32 wave groups (waves 0, 1 channels 0..31, waves 2, 3 channels 32..63
    // with 64; one wave per 32 channels with 128), which cover the same positions and
    // whose sums are added in channel order through LDS before the epilogue of the
    // first group.  A workgroup covers NPOS coarse positions per plane (Rb = NPOS / Dc
    // rows); with Dc = 8 a tile of 16 positions covers two coarse rows.
    constexpr int KS = CIN / 32, WG = 4 / KS, MT = 2, NPOS = 16 * MT * WG, U8 = CIN / 8;
    constexpr int NU = CIN == 32 ? 5 : 7;  // slab units per thread: (Rb + 2) (Dc + 2) U8 <= 256 NU
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int wv = wave % WG, kh = wave / WG;  // position group, channel half
    const int Dc = D >> 1;
    const int nb0 = (int)blockIdx.y >> 2, py = ((int)blockIdx.y >> 1) & 1, pz = (int)blockIdx.y & 1;
    const int Rb = NPOS / Dc, RY = Rb + 2, RX = Dc + 2;
    const int UR = RX * U8, SU = RY * UR;
    const float invUR = 1.f / (float)UR;
    const int nyb = Dc / Rb;
    const int yb = (int)blockIdx.x % nyb, zsg = ((int)blockIdx.x / nyb) % nzs, smp = (int)blockIdx.x / (nyb * nzs);
    const int y0 = yb * Rb, zb = zsg * Dc / nzs, ZL = (zsg + 1) * Dc / nzs - zb;  // coarse rows and planes
    const int sb = smp * Dc * Dc * Dc;
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(X, (P >> 3) * CIN * 2);
    const __amdgpu_buffer_rsrc_t wr = make_rsrc(Wk, 27 * Cout * CIN * 2);
    int lo[MT][3];  // LDS byte offset of this lane's fragment in tile t at coarse taps (0, 0, ox = d - 1), slot 0
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int q = 16 * MT * wv + 16 * t + r;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int ix = (q & (Dc - 1)) + d;
            // 16-byte unit 4 kh + g of the position, swizzled within its group of four
            lo[t][d] = (((q / Dc) + 1) * RX + ix) * CIN * 2 + 16 * ((4 * kh + g) ^ swz_unit<32>(ix));
        }
    }
    // folded weights of the 16 (tap, px) pairs: lane (r, g) holds output channel
    // 16 nb0 + r, input channels 32 kh + 8g .. + 7
    uint4 whi[16], wlo[16];
    {
        const int wvo = ((nb0 * 16 + r) * CIN + 32 * kh + g * 8) * 2;
        auto fold = [&](auto kc, auto pc) {
            constexpr int k = decltype(kc)::value, px = decltype(pc)::value, pi = up2p_pair(k, px);
            if constexpr (pi >= 0) {
                constexpr int ox = k % 3 - 1;
                const int oz = pz - 1 + k / 6, oy = py - 1 + (k / 3) % 2;
                float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int tz = -1; tz <= 1; ++tz)
#pragma unroll
                    for (int ty = -1; ty <= 1; ++ty)
#pragma unroll
                        for (int tx = -1; tx <= 1; ++tx) {
                            if ((px + tx + 2) / 2 - 1 != ox) continue;  // compile time
                            if ((pz + tz + 2) / 2 - 1 != oz || (py + ty + 2) / 2 - 1 != oy) continue;
                            const int tap = 9 * (tz + 1) + 3 * (ty + 1) + (tx + 1);
                            const uint4 wv = __builtin_bit_cast(
                                uint4, __builtin_amdgcn_raw_buffer_load_b128(wr, wvo, tap * Cout * CIN * 2, 0));
                            const uint32_t wd[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
                            for (int e = 0; e < 8; ++e)
                                sum[e] += __uint_as_float(e & 1 ? wd[e / 2] & 0xffff0000u : wd[e / 2] << 16);
                        }
                uint32_t hd[4], ld[4];
#pragma unroll
                for (int e2 = 0; e2 < 4; ++e2) {
                    const uint32_t h0 = f2bf(sum[2 * e2]), h1 = f2bf(sum[2 * e2 + 1]);
                    const uint32_t l0 = f2bf(sum[2 * e2] - bf2f((uint16_t)h0)),
                                   l1 = f2bf(sum[2 * e2 + 1] - bf2f((uint16_t)h1));
                    hd[e2] = h0 | (h1 << 16);
                    ld[e2] = l0 | (l1 << 16);
                }
                whi[pi] = make_uint4(hd[0], hd[1], hd[2], hd[3]);
                wlo[pi] = make_uint4(ld[0], ld[1], ld[2], ld[3]);
            }
        };
        static_for([&](auto kc) {
            fold(kc, std::integral_constant<int, 0>{});
            fold(kc, std::integral_constant<int, 1>{});
        }, std::make_integer_sequence<int, 12>{});
    }
    uint4 v[NU];
    int us[NU];  // swizzled unit of the LDS write
    auto load_slab = [&](int z) {
#pragma unroll
        for (int k = 0; k < NU; ++k) {
            const int u = tid + 256 * k;
            const int row = udiv_small(u, UR, invUR), e = u - row * UR;
            const int ix = e / U8, c8 = e - ix * U8;
            us[k] = u - c8 + (c8 ^ swz_unit<32>(ix));
            const int y = y0 - 1 + row;
            const bool in = u < SU && z >= 0 && z < Dc && y >= 0 && y < Dc && ix >= 1 && ix <= Dc;
            const int voff = in ? ((sb + (z * Dc + y) * Dc + (ix - 1)) * CIN + 8 * c8) * 2 : kOffNone;
            v[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, voff, 0, 0));
        }
    };
    auto store_slab = [&](int slot) {
        unsigned char *dst = sx + (size_t)slot * SU * 16;
#pragma unroll
        for (int k = 0; k < NU; ++k)
            if (tid + 256 * k < SU) *reinterpret_cast<uint4 *>(dst + 16 * us[k]) = v[k];
    };
    load_slab(zb - 1);
    store_slab(0);
    load_slab(zb);
    store_slab(1);
    load_slab(zb + 1);
    store_slab(2);
    __syncthreads();
    const int yf0 = (act & kActYf0) != 0;
    const int a1 = act & 1;
    const int cb = 16 * nb0;
    const float bco = bias ? bias[cb + r] : 0.f;
    for (int zi = 0; zi < ZL; ++zi) {
        const int z = zb + zi;  // coarse plane; the outputs are fine plane 2 z + pz
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int d = 0; d < 3; ++d) asm volatile("" : "+v"(lo[t][d]));
        if (zi + 1 < ZL) load_slab(z + 2);  // in flight during this plane's MFMAs
        f32x4 acc[MT][2];
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int px = 0; px < 2; ++px) acc[t][px] = f32x4{0.f, 0.f, 0.f, 0.f};
        // the slab of plane z + o is in slot (zi + 1 + o) % 3
        static_for([&](auto kc) {
            constexpr int k = decltype(kc)::value, ox = k % 3 - 1;
            const int oz = pz - 1 + k / 6, oy = py - 1 + (k / 3) % 2;
            const int toff = ((zi + 1 + oz) % 3) * SU * 16 + oy * RX * CIN * 2;
            uint4 a[MT];
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                int off = lo[t][ox + 1] + toff;
                asm volatile("" : "+v"(off));
                a[t] = *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(sx + off, 16));
            }
            auto pass = [&](auto pc) {
                constexpr int px = decltype(pc)::value, pi = up2p_pair(k, px);
                if constexpr (pi >= 0) {
                    const bf8 bh = __builtin_bit_cast(bf8, whi[pi]), bl = __builtin_bit_cast(bf8, wlo[pi]);
#pragma unroll
                    for (int t = 0; t < MT; ++t) {
                        const bf8 av = __builtin_bit_cast(bf8, a[t]);
                        acc[t][px] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bh, acc[t][px], 0, 0, 0);
                        acc[t][px] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bl, acc[t][px], 0, 0, 0);
                    }
                }
            };
            pass(std::integral_constant<int, 0>{});
            pass(std::integral_constant<int, 1>{});
        }, std::make_integer_sequence<int, 12>{});
        const int free_slot = zi % 3;  // plane z - 1
        __syncthreads();  // every wave is done with the slab of plane z-1: it becomes the epilogue scratch
        unsigned char *freeb = sx + (size_t)free_slot * SU * 16;
        constexpr int RED = (KS - 1) * WG * MT * 2 * 4 * 64;  // floats of the other groups' sums
        if constexpr (KS > 1) {  // the other channel parts' sums, added in order by the first group
            float *red = reinterpret_cast<float *>(freeb);
            if (kh > 0)
#pragma unroll
                for (int t = 0; t < MT; ++t)
#pragma unroll
                    for (int px = 0; px < 2; ++px)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            red[((((kh - 1) * WG + wv) * MT + t) * 2 + px) * 4 * 64 + i * 64 + lane] = acc[t][px][i];
            __syncthreads();
            if (kh == 0)
#pragma unroll
                for (int k2 = 1; k2 < KS; ++k2)
#pragma unroll
                    for (int t = 0; t < MT; ++t)
#pragma unroll
                        for (int px = 0; px < 2; ++px)
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                                acc[t][px][i] += red[((((k2 - 1) * WG + wv) * MT + t) * 2 + px) * 4 * 64 + i * 64 + lane];
        }
        if (kh == 0) {
            float *T = reinterpret_cast<float *>(freeb + RED * 4 + wv * kEpiScratch);
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                // coarse tile: 16 positions from q0 (row q0 / Dc, columns q0 % Dc ..) -> fine
                // positions 2 (q0 % Dc) .. + 31 of fine row 2 (y0 + q0 / Dc) + py, plane 2 z + pz
#pragma unroll
                for (int px = 0; px < 2; ++px)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        float vv = acc[t][px][i] + bco;
                        if (a1) vv = vv > 0.f ? vv : kLeak * vv;
                        T[(2 * (4 * g + i) + px) * kEpiStride + r] = vv;
                    }
                __builtin_amdgcn_wave_barrier();
                const int q0 = 16 * MT * wv + 16 * t;
                const int fy = 2 * (y0 + q0 / Dc) + py, fz = 2 * z + pz, fx = 2 * (q0 & (Dc - 1));
                const int pb = ((smp * D + fz) * D + fy) * D + fx;
                // Dc = 8: the tile's second coarse row is the fine row two below
                epi_write32<CIN == 128>(T, pb, P, Cout, cb, lane, A, skip, Y, Yf, yf0 != 0,
                                        Dc == 8 ? pb + 2 * D : pb + 16);
                __builtin_amdgcn_wave_barrier();  // the scratch is rewritten by the next tile
            }
        }
        if (zi + 1 < ZL) {
            __syncthreads();  // every wave is done with its scratch in the slot
            store_slab(free_slot);
            __syncthreads();
        }
    }
}

// Data gradient of the convolution of an upsampled input, summed over each cell's
// children (what k_conv_zs's POOL form gives), as a stride-2 convolution of the
// fine gradient: dX[q] = sum over d in {-1, 0, 1, 2}^3 of V[d]^T dY[2q + d], where
// V[d] sums the fine weights W[t] with 0 <= d + t <= 1 along each axis (1, 2, 4 or
// 8 of them), which is the adjoint of the fold that k_conv_up2p uses.  dY has CO =
// 16 or 32 channels (the layer's outputs), dX CI (the layer's inputs, 16 per
// blockIdx.y).  The four waves of a workgroup cover the same MT tiles of 16 coarse
// positions, wave w the taps with dz = w - 1; the folded weights of its 16 taps are
// summed in float from the bf16 ones and held as bf16 hi + lo, one 16x16x32 MFMA
// operand per tap pair (CO = 16: x-neighbours) or tap (CO = 32).  The four partial
// sums go through LDS and are added in dz order.  A ring of four fine planes (two
// fetched during each coarse plane's MFMAs) feeds the walk over coarse planes.
template <int CO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv_pool2p(
    const uint16_t *__restrict__ dY, const uint16_t *__restrict__ Wb, float *dX, int P, int D, int CI, int nzs)
{
    static_assert(CO == 16 || CO == 32, "16 or 32 gradient channels");
    typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
    constexpr int MT = CO == 16 ? 4 : 2, NPOS = 16 * MT, U = CO / 8;
    constexpr int NS = CO == 16 ? 8 : 16;  // MFMA steps per wave: (dy, x pair) or (dy, dx)
    constexpr int NUF = 9;                 // staging units per thread: two fine planes, at most 2304
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int dz = wave - 1;
    const int Dc = D >> 1, Rb = NPOS / Dc;
    const int FR = 2 * Rb + 2, FX = D + 2, SUF = FR * FX * U;  // fine rows, columns, 16-byte units of a plane
    const int nb0 = (int)blockIdx.y;
    const int nyb = Dc / Rb;
    const int yb = (int)blockIdx.x % nyb, zsg = ((int)blockIdx.x / nyb) % nzs, smp = (int)blockIdx.x / (nyb * nzs);
    const int y0 = yb * Rb, zb = zsg * Dc / nzs, ZL = (zsg + 1) * Dc / nzs - zb;  // coarse
    const __amdgpu_buffer_rsrc_t yr = make_rsrc(dY, P * CO * 2);
    const __amdgpu_buffer_rsrc_t wr = make_rsrc(Wb, 27 * CI * CO * 2);
    float *red = reinterpret_cast<float *>(sx + (size_t)4 * SUF * 16);  // 4 x MT x 4 x 64 partial sums
    auto swz = [&](int ix) { return (ix >> (U == 2 ? 3 : 2)) & (U - 1); };
    // lane's unit and x offset: CO = 16, step (dy, xp): dx = 2 xp - 1 + g / 2, unit g % 2;
    // CO = 32, step (dy, dx): unit g
    const int cu = CO == 16 ? (g & 1) : g;
    int lo[MT][CO == 16 ? 2 : 4];  // byte offset in a plane slab of tile t's fragment at dy = 0 and each x choice
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int q = 16 * t + r;
        const int row = 2 * (q / Dc) + 1, ixb = 2 * (q & (Dc - 1)) + 1;
#pragma unroll
        for (int xi = 0; xi < (CO == 16 ? 2 : 4); ++xi) {
            const int ddx = CO == 16 ? 2 * xi - 1 + (g >> 1) : xi - 1;
            const int ix = ixb + ddx;
            lo[t][xi] = ((row * FX + ix) * U + (cu ^ swz(ix))) * 16;
        }
    }
    // folded weights: lane (r, g) holds dX channel 16 nb0 + r and 8 dY channels
    uint4 whi[NS], wlo[NS];
    {
        auto taps = [](int d, int k) { return k - d; };  // the fine taps t of offset d: t = k - d for k = 0, 1 in range
        static_for([&](auto sc) {
            constexpr int st = decltype(sc)::value;
            const int ddy = CO == 16 ? st / 2 - 1 : st / 4 - 1;
            const int ddx = CO == 16 ? 2 * (st % 2) - 1 + (g >> 1) : st % 4 - 1;
            const int co0 = 8 * cu;
            float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            for (int kz = 0; kz < 2; ++kz) {
                const int tz = taps(dz, kz);
                if (tz < -1 || tz > 1) continue;
                for (int ky = 0; ky < 2; ++ky) {
                    const int ty = taps(ddy, ky);
                    if (ty < -1 || ty > 1) continue;
                    for (int kx = 0; kx < 2; ++kx) {
                        const int tx = taps(ddx, kx);
                        if (tx < -1 || tx > 1) continue;
                        const int tap = 9 * (tz + 1) + 3 * (ty + 1) + (tx + 1);
                        const uint4 wv = __builtin_bit_cast(
                            uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                       wr, ((tap * CI + 16 * nb0 + r) * CO + co0) * 2, 0, 0));
                        const uint32_t wd[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            sum[e] += __uint_as_float(e & 1 ? wd[e / 2] & 0xffff0000u : wd[e / 2] << 16);
                    }
                }
            }
            uint32_t hd[4], ld[4];
#pragma unroll
            for (int e2 = 0; e2 < 4; ++e2) {
                const uint32_t h0 = f2bf(sum[2 * e2]), h1 = f2bf(sum[2 * e2 + 1]);
                const uint32_t l0 = f2bf(sum[2 * e2] - bf2f((uint16_t)h0)),
                               l1 = f2bf(sum[2 * e2 + 1] - bf2f((uint16_t)h1));
                hd[e2] = h0 | (h1 << 16);
                ld[e2] = l0 | (l1 << 16);
            }
            whi[st] = make_uint4(hd[0], hd[1], hd[2], hd[3]);
            wlo[st] = make_uint4(ld[0], ld[1], ld[2], ld[3]);
        }, std::make_integer_sequence<int, NS>{});
    }
    // two fine planes f0, f0 + 1 through registers into ring slots (f + 4) % 4
    uint4 v[NUF];
    int us[NUF];
    auto load_pair = [&](int f0) {
#pragma unroll
        for (int k = 0; k < NUF; ++k) {
            const int u = tid + 256 * k;
            const int pl = u >= SUF ? 1 : 0, uu = u - pl * SUF;
            const int row = uu / (FX * U), rem = uu - row * FX * U;
            const int ix = rem / U, c = rem - ix * U;
            us[k] = pl * SUF + (row * FX + ix) * U + (c ^ swz(ix));
            const int fz = f0 + pl, fy = 2 * y0 - 1 + row, fx = ix - 1;
            const bool in = u < 2 * SUF && fz >= 0 && fz < D && fy >= 0 && fy < D && fx >= 0 && fx < D;
            const int voff = in ? ((((smp * D + fz) * D + fy) * D + fx) * CO + 8 * c) * 2 : kOffNone;
            v[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(yr, voff, 0, 0));
        }
    };
    auto store_pair = [&](int f0) {
        const int s0 = (f0 + 4) & 3, s1 = (f0 + 5) & 3;
#pragma unroll
        for (int k = 0; k < NUF; ++k) {
            const int u = tid + 256 * k;
            if (u < 2 * SUF) {
                const int pl = us[k] >= SUF ? 1 : 0;
                *reinterpret_cast<uint4 *>(sx + ((size_t)(pl ? s1 : s0) * SUF + us[k] - pl * SUF) * 16) = v[k];
            }
        }
    };
    load_pair(2 * zb - 1);
    store_pair(2 * zb - 1);
    load_pair(2 * zb + 1);
    store_pair(2 * zb + 1);
    __syncthreads();
    for (int zi = 0; zi < ZL; ++zi) {
        const int z = zb + zi;  // coarse plane: fine planes 2z - 1 .. 2z + 2 are in the ring
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int xi = 0; xi < (CO == 16 ? 2 : 4); ++xi) asm volatile("" : "+v"(lo[t][xi]));
        if (zi + 1 < ZL) load_pair(2 * z + 3);  // in flight during this plane's MFMAs
        f32x4 acc[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int pbase = ((2 * z + dz + 4) & 3) * SUF * 16;
        static_for([&](auto sc) {
            constexpr int st = decltype(sc)::value;
            constexpr int ddy = CO == 16 ? st / 2 - 1 : st / 4 - 1, xi = CO == 16 ? st % 2 : st % 4;
            const bf8 bh = __builtin_bit_cast(bf8, whi[st]), bl = __builtin_bit_cast(bf8, wlo[st]);
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                int off = pbase + ddy * FX * U * 16 + lo[t][xi];
                asm volatile("" : "+v"(off));
                const bf8 a = __builtin_bit_cast(bf8, *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(sx + off, 16)));
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc[t], 0, 0, 0);
            }
        }, std::make_integer_sequence<int, NS>{});
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[((wave * MT + t) * 4 + i) * 64 + lane] = acc[t][i];
        __syncthreads();  // every wave's MFMAs and partial sums are done
        if (wave < MT) {  // wave w finishes tile w: the dz partials in order
            const int t = wave, q0 = 16 * t;
            const int y = y0 + q0 / Dc, xb = (q0 & (Dc - 1)) + 4 * g;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float o = red[((0 * MT + t) * 4 + i) * 64 + lane];
#pragma unroll
                for (int w = 1; w < 4; ++w) o += red[((w * MT + t) * 4 + i) * 64 + lane];
                dX[((size_t)((smp * Dc + z) * Dc + y) * Dc + xb + i) * CI + 16 * nb0 + r] = o;
            }
        }
        if (zi + 1 < ZL) store_pair(2 * z + 3);  // into the slots of planes 2z - 1, 2z
        __syncthreads();
    }
}

// Weight gradient of a layer whose output gradient dY is nonzero in channel 0 only
// (the output layer, whose loss reads channel 0; dY rows of ldy channels, X of 16):
// dW[t][ci][0] = sum_q X[q][ci] dY[q - t][0], one 16x16x32 MFMA per 16 taps and 32
// positions (M = ci, N = the 27 taps, K = positions).  A workgroup (4 waves) owns
// Rb = 256 / D rows (D >= 32) of a run of planes: X of plane z channel-major in LDS,
// channel 0 of dY in a ring of three planes (rows y0 - 1 .. y0 + Rb, 8 zero
// elements on either side of each row); lane (n, g) builds the tap n's shifted dY
// fragment from an aligned read and its two funnel shifts.  Each workgroup's sums
// (27 x 16 floats, the waves added in order) go to part[blk].
constexpr int kWo0Pad = 8;
__global__ __launch_bounds__(256) void k_wgrad_o0(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                                  float *part, int P, int D, int ldy, int nzs)
{
    constexpr int CIN = 16;
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int Rb = 256 / D, HS = D + 2 * kWo0Pad, HR = Rb + 2;
    uint16_t *XC = reinterpret_cast<uint16_t *>(sx);    // [16][Rb * D]
    uint16_t *HY = XC + CIN * 256;                       // [3][HR][HS]
    const int nyb = D / Rb;
    const int yb = (int)blockIdx.x % nyb, zsg = ((int)blockIdx.x / nyb) % nzs, smp = (int)blockIdx.x / (nyb * nzs);
    const int y0 = yb * Rb, zb = zsg * D / nzs, ZL = (zsg + 1) * D / nzs - zb;
    const int s0 = smp * D * D * D;
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(X, P * CIN * 2);
    const __amdgpu_buffer_rsrc_t yr = make_rsrc(dY, P * ldy * 2);
    for (int i = tid; i < 3 * HR * HS / 8; i += 256) reinterpret_cast<uint4 *>(HY)[i] = make_uint4(0, 0, 0, 0);
    // X: thread = (position pair, channel half); dY: (Rb + 2) x D elements of channel 0
    const int pr = tid % (128), hx = tid / 128;  // pair 0..127 of the 256 positions, channel half
    uint4 xa, xb;
    uint16_t hv[2];  // at most (Rb + 2) D = 2 x 256 + ... elements: two per thread for D >= 32
    auto load_x = [&](int z) {
        const int q = 2 * pr, row = q / D, x = q - row * D;
        const int o = ((s0 + (z * D + y0 + row) * D + x) * CIN + 8 * hx) * 2;
        xa = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, o, 0, 0));
        xb = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, o + CIN * 2, 0, 0));
    };
    auto store_x = [&]() {
        const uint32_t wa[4] = {xa.x, xa.y, xa.z, xa.w}, wb[4] = {xb.x, xb.y, xb.z, xb.w};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            // channel 8 hx + c of positions 2 pr (low half) and 2 pr + 1 (high half)
            const uint32_t d = (c & 1) ? __builtin_amdgcn_perm(wb[c / 2], wa[c / 2], 0x07060302u)
                                       : __builtin_amdgcn_perm(wb[c / 2], wa[c / 2], 0x05040100u);
            *reinterpret_cast<uint32_t *>(XC + (8 * hx + c) * 256 + 2 * pr) = d;
        }
    };
    auto load_h = [&](int z) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int e = tid + 256 * k, row = e / D, x = e - row * D;
            const int y = y0 - 1 + row;
            const bool in = row < HR && z >= 0 && z < D && y >= 0 && y < D;
            hv[k] = in ? dY[(size_t)(s0 + (z * D + y) * D + x) * ldy] : (uint16_t)0;
        }
    };
    auto store_h = [&](int slot) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int e = tid + 256 * k, row = e / D, x = e - row * D;
            if (row < HR) HY[(slot * HR + row) * HS + kWo0Pad + x] = hv[k];
        }
    };
    __syncthreads();  // the zero fill is complete
    load_h(zb - 1);
    store_h(0);
    load_h(zb);
    store_h(1);
    load_h(zb + 1);
    store_h(2);
    load_x(zb);
    store_x();
    __syncthreads();
    // this lane's taps t = 16 nt + r: (dz, dy, dx)
    int tdz[2], tdy[2], tdx[2];
    bool tok[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int t = 16 * nt + r;
        tok[nt] = t < 27;
        tdz[nt] = tok[nt] ? t / 9 - 1 : 0;
        tdy[nt] = tok[nt] ? (t / 3) % 3 - 1 : 0;
        tdx[nt] = tok[nt] ? t % 3 - 1 : 0;
    }
    f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
    typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
    for (int zi = 0; zi < ZL; ++zi) {
        const int z = zb + zi;
        if (zi + 1 < ZL) {  // in flight during this plane's MFMAs
            load_x(z + 1);
            load_h(z + 2);
        }
        for (int ck = wave; ck < 8; ck += 4) {  // 32-position chunks of the plane block
            const int q0 = 32 * ck, row = q0 / D, x0 = q0 - row * D;
            const bf8 a = __builtin_bit_cast(
                bf8, *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(XC + r * 256 + q0 + 8 * g, 16)));
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                // dY[q - t] for the lane's 8 positions q: plane z - dz (slot), row - dy, x - dx
                const int slot = (zi + 1 - tdz[nt]) % 3;
                const uint16_t *hp = HY + (slot * HR + row + 1 - tdy[nt]) * HS + kWo0Pad + x0 + 8 * g;
                const uint4 c = *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(hp, 16));
                const uint32_t pv = *reinterpret_cast<const uint32_t *>(__builtin_assume_aligned(hp - 2, 4));
                const uint32_t nx = *reinterpret_cast<const uint32_t *>(__builtin_assume_aligned(hp + 8, 4));
                const uint32_t a1 = __builtin_amdgcn_alignbit(c.x, pv, 16), a2 = __builtin_amdgcn_alignbit(c.y, c.x, 16),
                               a3 = __builtin_amdgcn_alignbit(c.z, c.y, 16), a4 = __builtin_amdgcn_alignbit(c.w, c.z, 16),
                               a5 = __builtin_amdgcn_alignbit(nx, c.w, 16);
                // dx = 1: positions q - 1 .. q + 6; dx = -1: q + 1 .. q + 8
                uint4 bv = tdx[nt] == 0 ? c : (tdx[nt] == 1 ? make_uint4(a1, a2, a3, a4) : make_uint4(a2, a3, a4, a5));
                if (!tok[nt]) bv = make_uint4(0, 0, 0, 0);
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf8, bv), acc[nt], 0, 0, 0);
            }
        }
        if (zi + 1 < ZL) {
            __syncthreads();  // every wave is done with X of plane z and dY of plane z - 1
            store_x();
            store_h((zi + 3) % 3);  // plane z + 2 takes the slot of plane z - 1
            __syncthreads();
        }
    }
    // D[m = ci 4g + i][n = tap 16 nt + r]; the four waves' sums added in order
    __syncthreads();
    float *red = reinterpret_cast<float *>(sx);  // 4 x 2 x 4 x 64
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) red[((wave * 2 + nt) * 4 + i) * 64 + lane] = acc[nt][i];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int t = 16 * nt + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = red[((0 * 2 + nt) * 4 + i) * 64 + lane];
#pragma unroll
                for (int w = 1; w < 4; ++w) v += red[((w * 2 + nt) * 4 + i) * 64 + lane];
                if (t < 27) part[((size_t)blockIdx.x * 27 + t) * CIN + 4 * g + i] = v;
            }
        }
    }
}

// dW[t][ci][co] = sum over the blocks of part[blk][t][ci] for co = 0, 0 for the other
// output channels.  Workgroup t: thread (ci = tid % 16, k = tid / 16) adds blocks k,
// k + 16, ... in order (four loads in flight), then a fixed tree over k.
__global__ __launch_bounds__(256) void k_wgrad_o0_reduce(const float *__restrict__ part, float *dW, int blocks, int Cout)
{
    __shared__ float s[256];
    const int t = blockIdx.x, ci = threadIdx.x & 15, k = threadIdx.x >> 4;
    float v = 0.f;
    int b = k;
    for (; b + 48 < blocks; b += 64) {  // four loads in flight, added in the same order
        const float p0 = part[((size_t)b * 27 + t) * 16 + ci], p1 = part[((size_t)(b + 16) * 27 + t) * 16 + ci],
                    p2 = part[((size_t)(b + 32) * 27 + t) * 16 + ci], p3 = part[((size_t)(b + 48) * 27 + t) * 16 + ci];
        v += p0;
        v += p1;
        v += p2;
        v += p3;
    }
    for (; b < blocks; b += 16) v += part[((size_t)b * 27 + t) * 16 + ci];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 8; w > 0; w >>= 1) {
        if (k < w) s[threadIdx.x] += s[threadIdx.x + 16 * w];
        __syncthreads();
    }
    for (int e = threadIdx.x; e < 16 * Cout; e += 256) {  // (ci, co) of tap t
        const int c = e / Cout, co = e - c * Cout;
        dW[(size_t)t * 16 * Cout + e] = co == 0 ? s[c] : 0.f;
    }
}

// Data gradient of the output layer from channel 0 of its output gradient (one
// bf16 per position, G0; the loss reads channel 0, so the other channels are
// zero), with the backward of the leaky ReLU below in the epilogue (mirror = 1):
//   dX[p][n] = sum_tap Wk[26 - tap][n][0] G0[p + off(tap)],
// one 16x16x32 MFMA per 16 positions and 16 channels with the 27 taps as K (5 of
// them zero) against 27 16x16x16 MFMAs for the full 16-channel gradient.  A
// workgroup of 4 waves owns Rb = 256 / D rows of a run of ZL planes and 16
// channels (16 blockIdx.y ..); LDS holds G0 of planes z-1, z, z+1 in a ring
// (rows y0 - 1 .. y0 + Rb, 8 zero elements on either side of each row) and a
// wave-private epilogue scratch.  Lane (r, g) gathers taps 8g .. 8g + 7 of
// position r of its tile.  The epilogue writes Gd = bf16(dX) and H =
// bf16(dX act'(A)) (A: the activation of the layer below, read one plane ahead)
// and adds dX act'(A) to this lane's bias sums in the order of
// conv_epilogue_lds_impl with ABWD; the bias partials (row x of bpart for
// workgroup x, zeros to rows x + gridDim.x, ...) are those of k_conv_zs with ABWD.
constexpr int kDo0Pad = 8;
__host__ inline size_t dgrad_o0_lds(int D)
{
    const size_t ring = (size_t)3 * (256 / D + 2) * (D + 2 * kDo0Pad) * 2;
    return std::max((ring + 15) / 16 * 16 + 4 * (size_t)kEpiScratch, (size_t)8 * 256 * sizeof(float));
}

__global__ __launch_bounds__(256) void k_dgrad_o0(const uint16_t *__restrict__ G0, const uint16_t *__restrict__ Wk,
                                                  const uint16_t *__restrict__ A, uint16_t *Y, uint16_t *Gd, int P,
                                                  int D, int Cin, int Cout, int nzs, float *bpart)
{
    constexpr int MT = 4;
    typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int nb0 = (int)blockIdx.y;
    const int Rb = 256 / D, HR = Rb + 2, HS = D + 2 * kDo0Pad;
    uint16_t *HY = reinterpret_cast<uint16_t *>(sx);  // [3][HR][HS]
    unsigned char *scr = sx + ((size_t)3 * HR * HS * 2 + 15) / 16 * 16 + wave * kEpiScratch;
    const int nyb = D / Rb;
    const int yb = (int)blockIdx.x % nyb, zsg = ((int)blockIdx.x / nyb) % nzs, smp = (int)blockIdx.x / (nyb * nzs);
    const int y0 = yb * Rb, zb = zsg * D / nzs, ZL = (zsg + 1) * D / nzs - zb;
    const int s0 = smp * D * D * D;
    for (int i = tid; i < 3 * HR * HS / 8; i += 256) reinterpret_cast<uint4 *>(HY)[i] = make_uint4(0, 0, 0, 0);
    // B[k = tap 8g + j][n = r]: the weights of output channel 16 nb0 + r, input channel 0, mirrored taps
    bf8 b;
    {
        uint32_t wv[4];
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            const int t0 = 8 * g + j, t1 = t0 + 1;
            const uint32_t lo = t0 < 27 ? Wk[((size_t)(26 - t0) * Cout + 16 * nb0 + r) * Cin] : 0u;
            const uint32_t hi = t1 < 27 ? Wk[((size_t)(26 - t1) * Cout + 16 * nb0 + r) * Cin] : 0u;
            wv[j / 2] = lo | (hi << 16);
        }
        b = __builtin_bit_cast(bf8, make_uint4(wv[0], wv[1], wv[2], wv[3]));
    }
    // this lane's taps 8g + j: ring plane dz + 1 and the element offset within a slab
    // (row + 1 + dy, x + dx), or none past tap 26
    int tpl[8], toff[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int t = 8 * g + j;
        tpl[j] = t < 27 ? t / 9 : -1;
        toff[j] = t < 27 ? ((t / 3) % 3) * HS + kDo0Pad + t % 3 - 1 : 0;
    }
    uint16_t hv[2];  // (Rb + 2) D <= 512 elements of a plane: two per thread
    auto load_h = [&](int z) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int e = tid + 256 * k, row = e / D, x = e - row * D;
            const int y = y0 - 1 + row;
            const bool in = row < HR && z >= 0 && z < D && y >= 0 && y < D;
            hv[k] = in ? G0[(size_t)s0 + ((size_t)z * D + y) * D + x] : (uint16_t)0;
        }
    };
    auto store_h = [&](int slot) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int e = tid + 256 * k, row = e / D, x = e - row * D;
            if (row < HR) HY[(slot * HR + row) * HS + kDo0Pad + x] = hv[k];
        }
    };
    // epilogue rows: lane's position lane / 2 of the 32 of a tile pair, channels cb + c8 .. + 7
    const int epos = lane / 2, c8 = 8 * (lane & 1), cb = 16 * nb0;
    auto eoff = [&](int z, int h) {  // element offset of the lane's row piece, tile pair h of plane z
        return (size_t)(s0 + (z * D + y0) * D + 64 * wave + 32 * h + epos) * Cout + cb + c8;
    };
    uint4 acur[2], anext[2];  // A of this plane (and of the next, in flight)
    __syncthreads();  // the zero fill is complete
    load_h(zb - 1);
    store_h(0);
    load_h(zb);
    store_h(1);
    load_h(zb + 1);
    store_h(2);
#pragma unroll
    for (int h = 0; h < 2; ++h) acur[h] = *reinterpret_cast<const uint4 *>(A + eoff(zb, h));
    __syncthreads();
    float bs[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) bs[e] = 0.f;
    float *T = reinterpret_cast<float *>(scr);
    for (int zi = 0; zi < ZL; ++zi) {
        const int z = zb + zi;
        if (zi + 1 < ZL) {  // in flight during this plane's MFMAs and epilogue
            load_h(z + 2);
#pragma unroll
            for (int h = 0; h < 2; ++h) anext[h] = *reinterpret_cast<const uint4 *>(A + eoff(z + 1, h));
        }
        const int sb[3] = {(zi % 3) * HR * HS, ((zi + 1) % 3) * HR * HS, ((zi + 2) % 3) * HR * HS};  // planes z-1, z, z+1
        f32x4 acc[MT][1];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const int q = 64 * wave + 16 * t + r, row = q / D, x = q - row * D;
            const int po = row * HS + x;
            uint32_t av[4];
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                const uint32_t lo = tpl[j] >= 0 ? HY[sb[tpl[j]] + toff[j] + po] : 0u;
                const uint32_t hi = tpl[j + 1] >= 0 ? HY[sb[tpl[j + 1]] + toff[j + 1] + po] : 0u;
                av[j / 2] = lo | (hi << 16);
            }
            acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, make_uint4(av[0], av[1], av[2], av[3])),
                                                               b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        }
        // epilogue, a tile pair at a time through the wave's scratch (D[m = 4g + i][n = r])
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int tt = 0; tt < 2; ++tt)
#pragma unroll
                for (int i = 0; i < 4; ++i) T[(16 * tt + 4 * g + i) * kEpiStride + r] = acc[2 * h + tt][0][i];
            __builtin_amdgcn_wave_barrier();
            const float4 lo = *reinterpret_cast<const float4 *>(T + epos * kEpiStride + c8);
            const float4 hi = *reinterpret_cast<const float4 *>(T + epos * kEpiStride + c8 + 4);
            const float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            const uint16_t *ah = reinterpret_cast<const uint16_t *>(&acur[h]);
            uint16_t hb[8], gb[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float ov = v[e] * (bf2f(ah[e]) > 0.f ? 1.f : kLeak);
                hb[e] = f2bf(ov);
                gb[e] = f2bf(v[e]);
                bs[e] += ov;
            }
            const size_t o = eoff(z, h);
            *reinterpret_cast<uint4 *>(Y + o) = __builtin_bit_cast(uint4, hb);
            *reinterpret_cast<uint4 *>(Gd + o) = __builtin_bit_cast(uint4, gb);
            __builtin_amdgcn_wave_barrier();  // the scratch is rewritten by the next pair
        }
        if (zi + 1 < ZL) {
            acur[0] = anext[0];
            acur[1] = anext[1];
            __syncthreads();  // every wave is done with the slab of plane z-1
            store_h(zi % 3);  // plane z + 2 takes the slot of plane z - 1
            __syncthreads();
        }
    }
    // channel c: the lanes with (lane & 1) == c / 8 of the four waves, in wave and lane order
    __syncthreads();  // every wave is done with the LDS
    float *red = reinterpret_cast<float *>(sx);  // 8 x 256 floats
#pragma unroll
    for (int e = 0; e < 8; ++e) red[e * 256 + tid] = bs[e];
    __syncthreads();
    if (tid < 16) {
        const int c = tid;
        const float *rr = red + (c % 8) * 256;
        float t = 0.f;
        for (int w = 0; w < 4; ++w)
            for (int l = c / 8; l < 64; l += 2) t += rr[64 * w + l];
        const int co = 16 * nb0 + c;
        bpart[(size_t)blockIdx.x * Cout + co] = t;
        for (int row = (int)blockIdx.x + (int)gridDim.x; row < kBiasBlocks; row += (int)gridDim.x)
            bpart[(size_t)row * Cout + co] = 0.f;
    }
}

// Weight gradient: one wave per (16 input channels, 16 output channels, chunk of
// positions) accumulates all 27 taps at once, so each X and dY element of the
// chunk is fetched by one wave instead of 27; partial sums go to
// ws[chunk][tap][ci][co] and are summed over chunks in a fixed order.  X and dY
// are first copied channel-major (XT[c][p], dYT[c][p]), so a lane's four
// consecutive positions of one channel are one aligned 8-byte load; the x-shifted
// taps come from the previous / current / next 4-position chunks of the same row.
#ifndef GP5_WGRAD_TILES
#define GP5_WGRAD_TILES 1024
#endif
constexpr long long kWgradTiles = GP5_WGRAD_TILES;
__host__ __device__ inline long long wgrad_chunk(long long P, int Cin, int Cout)
{
    // about kWgradTiles (chunk, channel tile) pairs over the layer, at least 256
    // positions each, a multiple of 32
    const long long tiles = (long long)(Cin / 16) * (Cout / 16);
    long long c = (P * tiles + kWgradTiles - 1) / kWgradTiles;
    c = c < 256 ? 256 : c;
    return (c + 31) / 32 * 32;
}

// out[c * P + p] = in[p * C + c] (bf16), 64 positions per block through LDS, 128
// channels at a time
__global__ __launch_bounds__(256) void k_transpose_pc(const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                                                      long long P, int C)
{
    __shared__ uint16_t t[64][129];
    const long long p0 = (long long)blockIdx.x * 64;
    for (int c0 = 0; c0 < C; c0 += 128) {
        const int cc = min(128, C - c0);
        for (int e = threadIdx.x; e < 64 * cc; e += 256) {
            const int pl = e / cc, c = e % cc;
            t[pl][c] = p0 + pl < P ? in[(size_t)(p0 + pl) * C + c0 + c] : (uint16_t)0;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * cc; e += 256) {
            const int c = e / 64, pl = e % 64;
            if (p0 + pl < P) out[(size_t)(c0 + c) * P + p0 + pl] = t[pl][c];
        }
        __syncthreads();
    }
}

// Blocked channel-major copy for P a multiple of 128 and C a multiple of 16:
// element (p, c) goes to out[((p / 8) C + c) 8 + p % 8], i.e. runs of 8
// positions of one channel, the channels of one 8-position block side by side.
// A workgroup moves 128 positions x 16 channels: each thread loads 8 channels of
// one position (16 bytes), and after the LDS transpose stores 8 positions of one
// channel (16 bytes; 16 threads cover 256 consecutive bytes).
// With up = 1 the input has side D/2 and position p (side D) reads its parent
// cell: the copy of the nearest upsampling.
__global__ __launch_bounds__(256) void k_transpose_b8(const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                                                      int C, int D, int up)
{
    constexpr int kRow = 128 + 8;  // LDS row of one channel, padded (16-byte aligned rows)
    __shared__ __align__(16) uint16_t t[16 * kRow];
    const int tid = threadIdx.x;
    const long long p0 = (long long)blockIdx.x * 128;
    const int c0 = 16 * (int)blockIdx.y;
    {
        const int pos = tid >> 1, h = tid & 1;
        long long q = p0 + pos;
        if (up) {
            const int x = (int)(q % D), y = (int)((q / D) % D), z = (int)((q / ((long long)D * D)) % D);
            const long long smp = q / ((long long)D * D * D), Ds = D / 2;
            q = ((smp * Ds + z / 2) * Ds + y / 2) * Ds + x / 2;
        }
        const uint4 v = *reinterpret_cast<const uint4 *>(in + (size_t)q * C + c0 + 8 * h);
        const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
        for (int k = 0; k < 8; ++k) t[(8 * h + k) * kRow + pos] = e[k];
    }
    __syncthreads();
    {
        const int blk = tid >> 4, c = tid & 15;  // 8-position block of the 16, channel
        *reinterpret_cast<uint4 *>(out + ((size_t)(p0 / 8 + blk) * C + c0 + c) * 8) =
            *reinterpret_cast<const uint4 *>(t + c * kRow + 8 * blk);
    }
}

// The same for cube sides that are not a multiple of 4 (small test cubes):
// scalar gathers from the position-major X and dY.
__global__ __launch_bounds__(64) void k_wgrad_gather(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                                     float *ws, long long P, int D, int Cin, int Cout, long long chunk)
{
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int nco = Cout / 16;
    const int cob = blockIdx.x % nco, cib = blockIdx.x / nco;
    const long long ck = blockIdx.y;
    const long long p0 = ck * chunk, p1 = min(P, p0 + chunk);
    const int ci = cib * 16 + r, co = cob * 16 + r;
    f32x4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long pb = p0; pb < p1; pb += 16) {
        long long pq[4];
        int px[4], py[4], pz[4];
        bool in[4];
        bf16x4 bv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long p = pb + 4 * g + i;
            in[i] = p < p1;
            pq[i] = p;
            px[i] = (int)(p % D);
            py[i] = (int)((p / D) % D);
            pz[i] = (int)((p / ((long long)D * D)) % D);
            bv[i] = in[i] ? (short)dY[(size_t)p * Cout + co] : (short)0;
        }
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            const long long off = ((long long)dz * D + dy) * D + dx;
            bf16x4 a;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xx = px[i] + dx, yy = py[i] + dy, zz = pz[i] + dz;
                const bool ok = in[i] && xx >= 0 && xx < D && yy >= 0 && yy < D && zz >= 0 && zz < D;
                a[i] = ok ? (short)X[(size_t)(pq[i] + off) * Cin + ci] : (short)0;
            }
            acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[tap], 0, 0, 0);
        }
    }
#pragma unroll
    for (int tap = 0; tap < 27; ++tap) {
        float *o = ws + (((size_t)ck * 27 + tap) * Cin + cib * 16) * Cout + cob * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) o[(size_t)(4 * g + i) * Cout] = acc[tap][i];
    }
}

__global__ __launch_bounds__(64) void k_wgrad(const uint16_t *__restrict__ XT, const uint16_t *__restrict__ dYT,
                                              float *ws, long long P, int D, int Cin, int Cout, long long chunk)
{
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int nco = Cout / 16;
    const int cob = blockIdx.x % nco, cib = blockIdx.x / nco;
    const long long ck = blockIdx.y;
    const long long p0 = ck * chunk, p1 = min(P, p0 + chunk);
    const int ci = cib * 16 + r, co = cob * 16 + r;
    const uint16_t *xrow = XT + (size_t)ci * P;
    const uint16_t *yrow = dYT + (size_t)co * P;
    const bf16x4 zero4 = bf16x4{0, 0, 0, 0};
    f32x4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long pb = p0; pb < p1; pb += 16) {
        // A[m = ci][k = position]: positions pg .. pg+3 (one x-row, x a multiple of 4);
        // B[k = position][n = co]: the same positions of channel co
        const long long pg = pb + 4 * g;
        if (pg >= p1) continue;  // p1 is a multiple of 16: groups are wholly in or out
        const int x = (int)(pg % D), y = (int)((pg / D) % D), z = (int)((pg / ((long long)D * D)) % D);
        const bf16x4 bv = ld4(yrow + pg);
#pragma unroll
        for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
                const int t0 = (dz + 1) * 9 + (dy + 1) * 3;  // tap of dx = -1
                const bool row_ok = z + dz >= 0 && z + dz < D && y + dy >= 0 && y + dy < D;
                bf16x4 cur = zero4, prv = zero4, nxt = zero4;
                if (row_ok) {
                    const long long q = pg + ((long long)dz * D + dy) * D;
                    cur = ld4(xrow + q);
                    if (x > 0) prv = ld4(xrow + q - 4);
                    if (x + 4 < D) nxt = ld4(xrow + q + 4);
                }
                const bf16x4 am = bf16x4{prv[3], cur[0], cur[1], cur[2]};
                const bf16x4 ap = bf16x4{cur[1], cur[2], cur[3], nxt[0]};
                acc[t0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bv, acc[t0], 0, 0, 0);
                acc[t0 + 1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(cur, bv, acc[t0 + 1], 0, 0, 0);
                acc[t0 + 2] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap, bv, acc[t0 + 2], 0, 0, 0);
            }
    }
    // D[m = ci row 4g+i][n = co column r]
#pragma unroll
    for (int tap = 0; tap < 27; ++tap) {
        float *o = ws + (((size_t)ck * 27 + tap) * Cin + cib * 16) * Cout + cob * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) o[(size_t)(4 * g + i) * Cout] = acc[tap][i];
    }
}

// The same for cube sides that are a multiple of 8, on the blocked copies of X
// and dY (k_transpose_b8), 32 positions per step: a lane loads 8 consecutive
// positions of its channel (16 bytes; the 16 lanes of a k-group read 256
// consecutive bytes) and feeds two
// MFMA k-steps with them, positions 8g .. 8g+3 and 8g+4 .. 8g+7 (the same split
// for dY), so one aligned load serves 6 MFMAs of a row (three x-shifted taps x
// two steps) with the 4 positions on either side from two 8-byte loads.  The
// taps are split by dz over blockIdx.z: each wave holds the 9 taps of one dz
// and reads only the 3 input rows of that plane, which triples the waves per
// chunk and cuts the accumulators per wave to 9.  The loads are buffer loads
// without branches (a piece outside the cube reads zero through the range
// check), and the 10 loads of the next 32-position step are in flight while
// the 18 MFMAs of the current one run.
__global__ __launch_bounds__(64) void k_wgrad8(const uint16_t *__restrict__ XT, const uint16_t *__restrict__ dYT,
                                               float *ws, int P, int D, int Cin, int Cout, int chunk)
{
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int nco = Cout / 16;
    const int cob = blockIdx.x % nco, cib = blockIdx.x / nco;
    const int ck = blockIdx.y;
    const int dz = (int)blockIdx.z - 1;
    const int p0 = ck * chunk, p1 = min(P, p0 + chunk);
    const int ci = cib * 16 + r, co = cob * 16 + r;
    // blocked layout: positions q .. q+7 (q a multiple of 8) of channel c at bytes ((q / 8) C + c) 16;
    // loads out of the cube get kOffNone and read zero
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(XT, P * Cin * 2);
    const __amdgpu_buffer_rsrc_t yr = make_rsrc(dYT, P * Cout * 2);
    typedef short bf16x8 __attribute__((ext_vector_type(8)));
    struct Step {
        bf16x8 c[3];
        bf16x4 pv[3], nx[3];
        bf16x8 b;
    };
    auto fetch = [&](int pb, Step &st) {
        const int pg = pb + 8 * g;
        const bool in = pg < p1;  // p1 is a multiple of 32 or P: groups of 8 are wholly in or out
        const unsigned q1 = (unsigned)pg / (unsigned)D, q2 = q1 / (unsigned)D;
        const int x = pg - (int)q1 * D, y = (int)(q1 - q2 * D), z = (int)(q2 % (unsigned)D);
        st.b = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(
                                              yr, in ? ((pg / 8) * Cout + co) * 16 : kOffNone, 0, 0));
        const bool plane_ok = in && z + dz >= 0 && z + dz < D;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int dy = k - 1;
            const bool row_ok = plane_ok && y + dy >= 0 && y + dy < D;
            const int qb = (pg + (dz * D + dy) * D) / 8;  // 8-position block of the row piece
            const int co0 = (qb * Cin + ci) * 16;
            st.c[k] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(xr, row_ok ? co0 : kOffNone, 0, 0));
            st.pv[k] = __builtin_bit_cast(
                bf16x4, __builtin_amdgcn_raw_buffer_load_b64(xr, row_ok && x > 0 ? co0 - Cin * 16 + 8 : kOffNone, 0, 0));
            st.nx[k] = __builtin_bit_cast(
                bf16x4, __builtin_amdgcn_raw_buffer_load_b64(xr, row_ok && x + 8 < D ? co0 + Cin * 16 : kOffNone, 0, 0));
        }
    };
    f32x4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    Step cur, nxt;
    fetch(p0, cur);
    for (int pb = p0; pb < p1; pb += 32) {
        fetch(pb + 32, nxt);  // in flight during this step's MFMAs (past p1: out of range, zeros)
        const bf16x4 b1 = bf16x4{cur.b[0], cur.b[1], cur.b[2], cur.b[3]};
        const bf16x4 b2 = bf16x4{cur.b[4], cur.b[5], cur.b[6], cur.b[7]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int t0 = 3 * k;  // tap of dx = -1 within this dz, dy = k - 1
            const bf16x8 c = cur.c[k];
            const bf16x4 pv = cur.pv[k], nx = cur.nx[k];
            acc[t0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4{pv[3], c[0], c[1], c[2]}, b1, acc[t0], 0, 0, 0);
            acc[t0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4{c[3], c[4], c[5], c[6]}, b2, acc[t0], 0, 0, 0);
            acc[t0 + 1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4{c[0], c[1], c[2], c[3]}, b1, acc[t0 + 1], 0, 0, 0);
            acc[t0 + 1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4{c[4], c[5], c[6], c[7]}, b2, acc[t0 + 1], 0, 0, 0);
            acc[t0 + 2] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4{c[1], c[2], c[3], c[4]}, b1, acc[t0 + 2], 0, 0, 0);
            acc[t0 + 2] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4{c[5], c[6], c[7], nx[0]}, b2, acc[t0 + 2], 0, 0, 0);
        }
        cur = nxt;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int tap = (dz + 1) * 9 + t;
        float *o = ws + (((size_t)ck * 27 + tap) * Cin + cib * 16) * Cout + cob * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) o[(size_t)(4 * g + i) * Cout] = acc[t][i];
    }
}

// Weight gradient straight from the channel-last X and dY, for cube sides D =
// 16 .. 128 (a power of two), on v_mfma_f32_16x16x32_bf16 (M = 16 input
// channels, N = 16 output channels, K = 32 positions).  A workgroup of 3 waves
// owns one (16 ci, 16 co) tile pair and Rb = 256 / D rows of ZL consecutive
// planes of one sample; wave w takes the 9 taps of dz = w - 1 and keeps their
// accumulators in registers for the whole walk.  LDS holds a ring of three X
// slabs (planes z-1, z, z+1; Rb + 2 rows with zeros outside the cube) and the
// dY plane z, both channel-major, so that a lane's 8 consecutive positions of
// one channel are one aligned 16-byte read.  The slabs are transposed on their
// way in: a thread loads 8 channels of two x-neighbours (two 16-byte loads) and
// writes 8 dwords, one per channel.  Row layout: x at element x + 8 of a row of
// D + 8 elements; elements 0 .. 7 stay zero, so x = -1 reads element 7 and
// x = D reads the next row's element 0 (or the zero padding after the last
// row).  The x shift of a tap moves to dY: sum_p X[p + dx] dY[p] =
// sum_q X[q] dY[q - dx] over the cube (dY is zero outside it), so each 32-
// position chunk builds its three x-shifted dY fragments once, from the aligned
// read and the dword on either side by funnel shifts, and every X fragment is
// one aligned read.  The global loads of plane z+2 (X) and z+1 (dY) are in
// flight while the waves run plane z.  Per-workgroup sums go to
// ws[blk][tap][ci][co] for the chunk reduction.  With up = 1, X has side D/2 and
// position (z, y, x) reads its parent cell (the weight gradient of the
// convolution of the nearest upsampling).
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8m __attribute__((ext_vector_type(8)));

// channel stride of a staged slab (elements): at least `elems`, a multiple of 8,
// and (in dwords) 4 mod 32, so that the 16 channels of one read spread over the banks
__host__ __device__ constexpr int wz_chstride(int elems)
{
    int dw = elems / 2 + 1;  // at least two elements of padding
    dw += ((4 - dw) % 32 + 32) % 32;
    return 2 * dw;
}
constexpr int kWzThreads = 192;

__global__ __launch_bounds__(kWzThreads) void k_wgrad_zs(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                                         float *ws, int P, int D, int Cin, int Cout, int nzs, int up)
{
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int Rb = 256 / D, XS = D + 8;
    const int CSX = wz_chstride((Rb + 2) * XS), CSY = wz_chstride(Rb * XS);
    const int SLOT = 16 * CSX;  // elements per X slab
    uint16_t *LX = reinterpret_cast<uint16_t *>(sx);
    uint16_t *LY = LX + 3 * SLOT;
    const int nco = Cout / 16;
    const int cib = (int)blockIdx.y / nco, cob = (int)blockIdx.y % nco;
    const int nyb = D / Rb;
    const int yb = (int)blockIdx.x % nyb, zsg = ((int)blockIdx.x / nyb) % nzs, smp = (int)blockIdx.x / (nyb * nzs);
    const int y0 = yb * Rb, zb = zsg * D / nzs, ZL = (zsg + 1) * D / nzs - zb;  // nzs runs of planes
    const int Ds = D >> up;
    const int sb = smp * Ds * Ds * Ds;  // stored position of this sample's origin (X)
    const int s0 = smp * D * D * D;     // position of this sample's origin (dY)
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(X, (P >> (3 * up)) * Cin * 2);
    const __amdgpu_buffer_rsrc_t yr = make_rsrc(dY, P * Cout * 2);
    // zero the whole staging area once: the row pads are never written again
    {
        const int n16 = (3 * SLOT + 16 * CSY) / 8;
        for (int i = tid; i < n16; i += kWzThreads) reinterpret_cast<uint4 *>(sx)[i] = make_uint4(0, 0, 0, 0);
    }
    // staging items: X slab (Rb + 2) rows x D/2 x-pairs x 2 channel halves; dY plane Rb rows x D/2 x 2
    const int NXI = (Rb + 2) * D, NYI = Rb * D;  // items (both are at most 2.67 per thread)
    constexpr int MXI = 3, MYI = 2;
    uint4 xv[MXI][2], yv[MYI][2];
    auto load_x = [&](int zz) {
#pragma unroll
        for (int k = 0; k < MXI; ++k) {
            const int it = tid + kWzThreads * k;
            const int pr = it % (D / 2), h = (it / (D / 2)) & 1, row = it / D;
            const int x = 2 * pr, y = y0 - 1 + row;
            const bool in = it < NXI && zz >= 0 && zz < D && y >= 0 && y < D;
            const int c8 = cib * 16 + 8 * h;
            const int o0 = in ? ((sb + ((zz >> up) * Ds + (y >> up)) * Ds + (x >> up)) * Cin + c8) * 2 : kOffNone;
            const int o1 = in ? ((sb + ((zz >> up) * Ds + (y >> up)) * Ds + ((x + 1) >> up)) * Cin + c8) * 2 : kOffNone;
            xv[k][0] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, o0, 0, 0));
            xv[k][1] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, o1, 0, 0));
        }
    };
    auto load_y = [&](int z) {
#pragma unroll
        for (int k = 0; k < MYI; ++k) {
            const int it = tid + kWzThreads * k;
            const int pr = it % (D / 2), h = (it / (D / 2)) & 1, row = it / D;
            const int x = 2 * pr;
            const bool in = it < NYI && z < D;
            const int c8 = cob * 16 + 8 * h;
            const int o0 = in ? ((s0 + (z * D + y0 + row) * D + x) * Cout + c8) * 2 : kOffNone;
            yv[k][0] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(yr, o0, 0, 0));
            yv[k][1] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(yr, in ? o0 + Cout * 2 : kOffNone, 0, 0));
        }
    };
    // dword of channel c (0..7 of the half) for the x-pair: low half from position x, high from x + 1
    auto pair_dw = [](const uint4 &a, const uint4 &b, int c) -> uint32_t {
        const uint32_t wa = c < 2 ? a.x : (c < 4 ? a.y : (c < 6 ? a.z : a.w));
        const uint32_t wb = c < 2 ? b.x : (c < 4 ? b.y : (c < 6 ? b.z : b.w));
        // byte selector of v_perm_b32(wb, wa, sel): bytes 0-3 from wa, 4-7 from wb
        return (c & 1) ? __builtin_amdgcn_perm(wb, wa, 0x07060302u) : __builtin_amdgcn_perm(wb, wa, 0x05040100u);
    };
    auto store_x = [&](int slot) {
        uint16_t *base = LX + slot * SLOT;
#pragma unroll
        for (int k = 0; k < MXI; ++k) {
            const int it = tid + kWzThreads * k;
            if (it >= NXI) continue;
            const int pr = it % (D / 2), h = (it / (D / 2)) & 1, row = it / D;
            uint16_t *p = base + (8 * h) * CSX + row * XS + 2 * pr + 8;
#pragma unroll
            for (int c = 0; c < 8; ++c) *reinterpret_cast<uint32_t *>(p + c * CSX) = pair_dw(xv[k][0], xv[k][1], c);
        }
    };
    auto store_y = [&]() {
#pragma unroll
        for (int k = 0; k < MYI; ++k) {
            const int it = tid + kWzThreads * k;
            if (it >= NYI) continue;
            const int pr = it % (D / 2), h = (it / (D / 2)) & 1, row = it / D;
            uint16_t *p = LY + (8 * h) * CSY + row * XS + 2 * pr + 8;
#pragma unroll
            for (int c = 0; c < 8; ++c) *reinterpret_cast<uint32_t *>(p + c * CSY) = pair_dw(yv[k][0], yv[k][1], c);
        }
    };
    __syncthreads();  // the zero fill is complete
    load_x(zb - 1);
    store_x(0);
    load_x(zb);
    store_x(1);
    load_x(zb + 1);
    store_x(2);
    load_y(zb);
    store_y();
    __syncthreads();
    const int dz = wave - 1;
    f32x4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int zi = 0; zi < ZL; ++zi) {
        const int z = zb + zi;
        if (zi + 1 < ZL) {  // in flight during this plane's MFMAs
            load_x(z + 2);
            load_y(z + 1);
        }
        const int slot = (zi + 1 + dz) % 3;  // ring slot of plane z + dz
        const uint16_t *xs = LX + slot * SLOT + r * CSX;
        const uint16_t *ys = LY + r * CSY;
#pragma unroll 2
        for (int ck = 0; ck < 8; ++ck) {
            const int q = 32 * ck + 8 * g;  // plane-local position of this lane group's 8 positions
            const int row = q / D, x = q - row * D;
            // dY[q - dx] for dx = -1, 0, +1: positions q+1 .. q+8, q .. q+7, q-1 .. q+6
            const uint16_t *yp = ys + row * XS + x + 8;
            const uint4 c = *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(yp, 16));
            const uint32_t pv = *reinterpret_cast<const uint32_t *>(__builtin_assume_aligned(yp - 2, 4));
            const uint32_t nx = *reinterpret_cast<const uint32_t *>(__builtin_assume_aligned(yp + 8, 4));
            const uint32_t a1 = __builtin_amdgcn_alignbit(c.x, pv, 16), a2 = __builtin_amdgcn_alignbit(c.y, c.x, 16),
                           a3 = __builtin_amdgcn_alignbit(c.z, c.y, 16), a4 = __builtin_amdgcn_alignbit(c.w, c.z, 16),
                           a5 = __builtin_amdgcn_alignbit(nx, c.w, 16);
            const bf16x8m bm = __builtin_bit_cast(bf16x8m, make_uint4(a2, a3, a4, a5));
            const bf16x8m b0 = __builtin_bit_cast(bf16x8m, c);
            const bf16x8m bp = __builtin_bit_cast(bf16x8m, make_uint4(a1, a2, a3, a4));
#pragma unroll
            for (int k = 0; k < 3; ++k) {  // dy = k - 1
                const bf16x8m a =
                    *reinterpret_cast<const bf16x8m *>(__builtin_assume_aligned(xs + (row + k) * XS + x + 8, 16));
                acc[3 * k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bm, acc[3 * k], 0, 0, 0);
                acc[3 * k + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[3 * k + 1], 0, 0, 0);
                acc[3 * k + 2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bp, acc[3 * k + 2], 0, 0, 0);
            }
        }
        if (zi + 1 < ZL) {
            __syncthreads();  // every wave is done with the slab of plane z-1 and the dY plane z
            store_x(zi % 3);  // plane z + 2 takes the slot of plane z - 1
            store_y();
            __syncthreads();
        }
    }
    // D[m = ci 4g+i][n = co r]
    const int n = 27 * Cin * Cout;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int tap = (dz + 1) * 9 + t;
        float *o = ws + (size_t)blockIdx.x * n + ((size_t)tap * Cin + cib * 16) * Cout + cob * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) o[(size_t)(4 * g + i) * Cout] = acc[t][i];
    }
}

// LDS bytes of k_wgrad_zs for cube side D
__host__ inline size_t wgrad_zs_lds(int D)
{
    const int Rb = 256 / D;
    return (size_t)(3 * 16 * wz_chstride((Rb + 2) * (D + 8)) + 16 * wz_chstride(Rb * (D + 8))) * 2;
}

// Weight gradient of the convolution of a nearest upsampling, by output parity
// (v_mfma_f32_16x16x32_bf16; X has side Dc = D/2, dY side D, D in {16, 32, 64,
// 128}).  A fine position p = 2q + a (a in {0,1}^3) reads the parent cell
// q + o under tap t, o = floor((a + t) / 2) in {a - 1, a}^3, so
//   dW[t] = sum_a G_a[o(a, t)],   G_a[o] = sum_q X(q + o) dY(2q + a):
// 64 products per coarse cell against 216 for the 27 taps of its 8 children.
// A workgroup of 4 waves owns one (16 ci, 16 co) tile pair and Rb = min(Dc,
// 128/Dc) coarse rows of a run of ZL coarse planes; wave w takes the parities
// a_z = w >> 1, a_y = w & 1 and keeps the 16 accumulators (s_z, s_y, a_x, s_x),
// o = a - 1 + s, for the whole walk.  LDS holds a ring of three coarse X slabs
// (laid out as in k_wgrad_zs) and the two fine dY planes of the current coarse
// plane split by x parity: row ((a_z * 2Rb + fine row) * 2 + a_x) holds dY at
// fine x = 2 x_c + a_x in element x_c + 8, elements 0 .. 7 zero.  The x offset
// moves to dY as in k_wgrad_zs: o_x = -1 reads dY_a at x_c + 1 and o_x = +1 at
// x_c - 1 (funnel shifts of one aligned read and a neighbouring dword), so every
// X fragment is one aligned read.  At the end each tap's sum of 8 G's is formed
// in a fixed order (x parities in registers, then the four waves through LDS)
// and written to ws[blk][tap][ci][co] for k_wgrad_reduce1.
constexpr int kWuThreads = 256;

__host__ __device__ constexpr int wu_rows(int Dc) { return Dc < 128 / Dc ? Dc : 128 / Dc; }

// LDS bytes of k_wgrad_up2 for fine cube side D: the staging area, or the 48 KiB
// of the final combination when that is larger
__host__ inline size_t wgrad_up2_lds(int D)
{
    const int Dc = D / 2, Rb = wu_rows(Dc), XS = Dc + 8;
    const size_t stage = (size_t)(3 * 16 * wz_chstride((Rb + 2) * XS) + 16 * wz_chstride(8 * Rb * XS)) * 2;
    return std::max(stage, (size_t)4 * 4 * 3 * 256 * sizeof(float));
}

__global__ __launch_bounds__(kWuThreads) void k_wgrad_up2(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                                          float *ws, int P, int D, int Cin, int Cout, int nzs)
{
    extern __shared__ __align__(16) unsigned char sx[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int az = wave >> 1, ay = wave & 1;
    const int Dc = D >> 1, Rb = wu_rows(Dc), XS = Dc + 8;
    const int CSX = wz_chstride((Rb + 2) * XS), CSY = wz_chstride(8 * Rb * XS);
    const int SLOT = 16 * CSX;  // elements per X slab
    uint16_t *LX = reinterpret_cast<uint16_t *>(sx);
    uint16_t *LY = LX + 3 * SLOT;
    const int nco = Cout / 16;
    const int cib = (int)blockIdx.y / nco, cob = (int)blockIdx.y % nco;
    const int nyb = Dc / Rb;
    const int yb = (int)blockIdx.x % nyb, zsg = ((int)blockIdx.x / nyb) % nzs, smp = (int)blockIdx.x / (nyb * nzs);
    const int y0 = yb * Rb, zb = zsg * Dc / nzs, ZL = (zsg + 1) * Dc / nzs - zb;  // coarse rows and planes
    const int sb = smp * Dc * Dc * Dc;  // coarse position of this sample's origin (X)
    const int s0 = smp * D * D * D;     // fine position of this sample's origin (dY)
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(X, (P >> 3) * Cin * 2);
    const __amdgpu_buffer_rsrc_t yr = make_rsrc(dY, P * Cout * 2);
    {  // zero the staging area once: the row pads are never written again
        const int n16 = (3 * SLOT + 16 * CSY) / 8;
        for (int i = tid; i < n16; i += kWuThreads) reinterpret_cast<uint4 *>(sx)[i] = make_uint4(0, 0, 0, 0);
    }
    // X items: (Rb + 2) rows x Dc/2 x-pairs x 2 channel halves (at most 256);
    // dY items: 2 fine planes x 2Rb fine rows x Dc/2 groups of 4 fine x x 2 halves (at most 512)
    const int NXI = (Rb + 2) * Dc, NYI = 4 * Rb * Dc;
    constexpr int MYI = 2;
    uint4 xv[2], yv[MYI][4];
    auto load_x = [&](int zz) {
        const int it = tid;
        const int pr = it % (Dc / 2), h = (it / (Dc / 2)) & 1, row = it / Dc;
        const int x = 2 * pr, y = y0 - 1 + row;
        const bool in = it < NXI && zz >= 0 && zz < Dc && y >= 0 && y < Dc;
        const int o0 = in ? ((sb + (zz * Dc + y) * Dc + x) * Cin + cib * 16 + 8 * h) * 2 : kOffNone;
        xv[0] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, o0, 0, 0));
        xv[1] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, in ? o0 + Cin * 2 : kOffNone, 0, 0));
    };
    auto load_y = [&](int zc) {
#pragma unroll
        for (int k = 0; k < MYI; ++k) {
            const int it = tid + kWuThreads * k;
            const int x4 = it % (Dc / 2), h = (it / (Dc / 2)) & 1, fr = it / Dc;  // fr = a_z * 2Rb + fine row
            const int fz = 2 * zc + fr / (2 * Rb), fy = 2 * y0 + fr % (2 * Rb);
            const bool in = it < NYI && zc < Dc;
            const int o0 = in ? ((s0 + (fz * D + fy) * D + 4 * x4) * Cout + cob * 16 + 8 * h) * 2 : kOffNone;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                yv[k][j] = __builtin_bit_cast(
                    uint4, __builtin_amdgcn_raw_buffer_load_b128(yr, in ? o0 + j * Cout * 2 : kOffNone, 0, 0));
        }
    };
    // dword of channel c (0..7 of the half) from two positions: low half from a, high from b
    auto pair_dw = [](const uint4 &a, const uint4 &b, int c) -> uint32_t {
        const uint32_t wa = c < 2 ? a.x : (c < 4 ? a.y : (c < 6 ? a.z : a.w));
        const uint32_t wb = c < 2 ? b.x : (c < 4 ? b.y : (c < 6 ? b.z : b.w));
        return (c & 1) ? __builtin_amdgcn_perm(wb, wa, 0x07060302u) : __builtin_amdgcn_perm(wb, wa, 0x05040100u);
    };
    auto store_x = [&](int slot) {
        const int it = tid;
        if (it >= NXI) return;
        const int pr = it % (Dc / 2), h = (it / (Dc / 2)) & 1, row = it / Dc;
        uint16_t *p = LX + slot * SLOT + (8 * h) * CSX + row * XS + 2 * pr + 8;
#pragma unroll
        for (int c = 0; c < 8; ++c) *reinterpret_cast<uint32_t *>(p + c * CSX) = pair_dw(xv[0], xv[1], c);
    };
    auto store_y = [&]() {
#pragma unroll
        for (int k = 0; k < MYI; ++k) {
            const int it = tid + kWuThreads * k;
            if (it >= NYI) continue;
            const int x4 = it % (Dc / 2), h = (it / (Dc / 2)) & 1, fr = it / Dc;
            // fine x 4 x4 + {0, 2} -> parity 0 at coarse x 2 x4, 2 x4 + 1; 4 x4 + {1, 3} -> parity 1
            uint16_t *p = LY + (8 * h) * CSY + (2 * fr) * XS + 2 * x4 + 8;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                *reinterpret_cast<uint32_t *>(p + c * CSY) = pair_dw(yv[k][0], yv[k][2], c);
                *reinterpret_cast<uint32_t *>(p + XS + c * CSY) = pair_dw(yv[k][1], yv[k][3], c);
            }
        }
    };
    __syncthreads();  // the zero fill is complete
    load_x(zb - 1);
    store_x(0);
    load_x(zb);
    store_x(1);
    load_x(zb + 1);
    store_x(2);
    load_y(zb);
    store_y();
    __syncthreads();
    f32x4 acc[2][2][2][2];  // [s_z][s_y][a_x][s_x]
#pragma unroll
    for (int i = 0; i < 16; ++i) (&acc[0][0][0][0])[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nck = Rb * Dc / 32;  // 32-cell chunks of a coarse plane slab (2 or 4)
    for (int zi = 0; zi < ZL; ++zi) {
        if (zi + 1 < ZL) {  // in flight during this plane's MFMAs
            load_x(zb + zi + 2);
            load_y(zb + zi + 1);
        }
        const uint16_t *ys = LY + r * CSY;
        const uint16_t *xs = LX + r * CSX;
#pragma unroll 2
        for (int ck = 0; ck < nck; ++ck) {
            const int q = 32 * ck + 8 * g;  // slab-local coarse cell of this lane group's 8 cells
            const int row = q / Dc, x = q - row * Dc;
            const uint16_t *y0p = ys + ((az * 2 * Rb + 2 * row + ay) * 2) * XS + x + 8;  // a_x = 0
            const uint16_t *y1p = y0p + XS;                                                // a_x = 1
            const uint4 c0 = *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(y0p, 16));
            const uint32_t nx0 = *reinterpret_cast<const uint32_t *>(__builtin_assume_aligned(y0p + 8, 4));
            const uint4 c1 = *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(y1p, 16));
            const uint32_t pv1 = *reinterpret_cast<const uint32_t *>(__builtin_assume_aligned(y1p - 2, 4));
            // a_x = 0: s_x = 0 is o_x = -1 (cells x+1 .. x+8), s_x = 1 is o_x = 0;
            // a_x = 1: s_x = 0 is o_x = 0, s_x = 1 is o_x = +1 (cells x-1 .. x+6)
            const bf16x8m b00 = __builtin_bit_cast(
                bf16x8m, make_uint4(__builtin_amdgcn_alignbit(c0.y, c0.x, 16), __builtin_amdgcn_alignbit(c0.z, c0.y, 16),
                                    __builtin_amdgcn_alignbit(c0.w, c0.z, 16), __builtin_amdgcn_alignbit(nx0, c0.w, 16)));
            const bf16x8m b01 = __builtin_bit_cast(bf16x8m, c0);
            const bf16x8m b10 = __builtin_bit_cast(bf16x8m, c1);
            const bf16x8m b11 = __builtin_bit_cast(
                bf16x8m, make_uint4(__builtin_amdgcn_alignbit(c1.x, pv1, 16), __builtin_amdgcn_alignbit(c1.y, c1.x, 16),
                                    __builtin_amdgcn_alignbit(c1.z, c1.y, 16), __builtin_amdgcn_alignbit(c1.w, c1.z, 16)));
#pragma unroll
            for (int sz = 0; sz < 2; ++sz) {
                const int slot = (zi + az + sz) % 3;  // ring slot of coarse plane z + a_z - 1 + s_z
#pragma unroll
                for (int sy = 0; sy < 2; ++sy) {  // slab row of coarse row + a_y - 1 + s_y
                    const bf16x8m a = *reinterpret_cast<const bf16x8m *>(
                        __builtin_assume_aligned(xs + slot * SLOT + (row + ay + sy) * XS + x + 8, 16));
                    acc[sz][sy][0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b00, acc[sz][sy][0][0], 0, 0, 0);
                    acc[sz][sy][0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b01, acc[sz][sy][0][1], 0, 0, 0);
                    acc[sz][sy][1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b10, acc[sz][sy][1][0], 0, 0, 0);
                    acc[sz][sy][1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b11, acc[sz][sy][1][1], 0, 0, 0);
                }
            }
        }
        if (zi + 1 < ZL) {
            __syncthreads();  // every wave is done with the slab of plane z-1 and the dY planes of z
            store_x(zi % 3);  // plane z + 2 takes the slot of plane z - 1
            store_y();
            __syncthreads();
        }
    }
    // V[wave][s_z s_y][t_x + 1][ci 16][co 16]: the two x parities combined, in a fixed order
    // (t_x = -1: s_x = 0 of both; t_x = 0: s_x = 1 of a_x = 0 and s_x = 0 of a_x = 1; t_x = 1: s_x = 1 of both)
    __syncthreads();  // the staging area is free
    float *V = reinterpret_cast<float *>(sx);
#pragma unroll
    for (int sz = 0; sz < 2; ++sz)
#pragma unroll
        for (int sy = 0; sy < 2; ++sy) {
            const f32x4 v[3] = {acc[sz][sy][0][0] + acc[sz][sy][1][0], acc[sz][sy][0][1] + acc[sz][sy][1][0],
                                acc[sz][sy][0][1] + acc[sz][sy][1][1]};
#pragma unroll
            for (int tx = 0; tx < 3; ++tx)
#pragma unroll
                for (int i = 0; i < 4; ++i) V[((wave * 4 + sz * 2 + sy) * 3 + tx) * 256 + (4 * g + i) * 16 + r] = v[tx][i];
        }
    __syncthreads();
    // thread tid owns (ci, co) = (tid / 16, tid % 16) of the tile pair; tap (t_z, t_y, t_x) sums,
    // over the waves (a_z, a_y) in order, the G with s = (a == 0 ? t >= 0 : t == 1) per axis
    const int n = 27 * Cin * Cout;
    float *o = ws + (size_t)blockIdx.x * n + (size_t)(cib * 16 + (tid >> 4)) * Cout + cob * 16 + (tid & 15);
#pragma unroll
    for (int tz = -1; tz <= 1; ++tz)
#pragma unroll
        for (int ty = -1; ty <= 1; ++ty)
#pragma unroll
            for (int tx = 0; tx < 3; ++tx) {
                float s = 0.f;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const int sz = (w >> 1) == 0 ? (tz >= 0) : (tz == 1);
                    const int sy = (w & 1) == 0 ? (ty >= 0) : (ty == 1);
                    s += V[((w * 4 + sz * 2 + sy) * 3 + tx) * 256 + tid];
                }
                o[(size_t)(((tz + 1) * 3 + (ty + 1)) * 3 + tx) * Cin * Cout] = s;
            }
}

// Sum of the chunk partials in one fixed-order pass: a workgroup owns OPB
// consecutive outputs, its Q =
// 256 / OPB thread groups take the chunks c = q, q + Q, q + 2Q, ... (q the
// group) in increasing order, and the Q group sums are added in the order
// q = 0 .. Q-1.  OPB = 16 keeps small layers spread over enough workgroups.
template <int OPB>
__global__ __launch_bounds__(256) void k_wgrad_reduce1(const float *__restrict__ ws, float *dW, long long n, int chunks)
{
    constexpr int Q = 256 / OPB;
    __shared__ float s[Q][OPB];
    const int o = threadIdx.x % OPB, q = threadIdx.x / OPB;
    const long long i = (long long)blockIdx.x * OPB + o;
    float a = 0.f;
    if (i < n) {
        int c = q;
        for (; c + 3 * Q < chunks; c += 4 * Q) {  // four loads in flight
            const float v0 = ws[(size_t)c * n + i], v1 = ws[(size_t)(c + Q) * n + i],
                        v2 = ws[(size_t)(c + 2 * Q) * n + i], v3 = ws[(size_t)(c + 3 * Q) * n + i];
            a += v0;
            a += v1;
            a += v2;
            a += v3;
        }
        for (; c < chunks; c += Q) a += ws[(size_t)c * n + i];
    }
    s[q][o] = a;
    __syncthreads();
    if (q == 0 && i < n) {
        float t = s[0][o];
#pragma unroll
        for (int k = 1; k < Q; ++k) t += s[k][o];
        dW[i] = t;
    }
}

// db[c] = sum_p dY[p][c] in two fixed-order passes: kBiasBlocks blocks each sum a
// contiguous range of positions for every channel (consecutive threads read
// consecutive channels), then one thread per channel adds the block sums in order.
__global__ __launch_bounds__(256) void k_bias_partial(const float *__restrict__ dY, float *part, long long P, int C)
{
    __shared__ float s[256];
    const int per_pos = 256 / C;  // C divides 256 (a multiple of 16, at most 256)
    const int c = threadIdx.x % C, sub = threadIdx.x / C;
    const long long per = (P + gridDim.x - 1) / gridDim.x;
    const long long p0 = (long long)blockIdx.x * per, p1 = min(P, p0 + per);
    float a = 0.f;
    if (sub < per_pos)
        for (long long p = p0 + sub; p < p1; p += per_pos) a += dY[(size_t)p * C + c];
    s[threadIdx.x] = a;
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float t = 0.f;
        for (int k = 0; k < per_pos; ++k) t += s[threadIdx.x + k * C];
        part[(size_t)blockIdx.x * C + threadIdx.x] = t;
    }
}

// one thread's share of k_act_bwd_cast_bias without G: positions p, p + step, ...
// below p1, four per pass with their loads issued together, dP summed into a in
// position order (64^3, 16 channels: 43.4 -> 41.3 us against one per pass)
__device__ __forceinline__ void act_bwd_bias_pass4(const uint16_t *__restrict__ A, const float *__restrict__ dA,
                                                   uint16_t *__restrict__ H, long long p, long long p1, int per_pos,
                                                   int C4, int c4, float4 &a)
{
    constexpr int U = 4;
    for (; p < p1; p += U * per_pos) {
        uint2 av[U];
        float4 d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long q = p + (long long)u * per_pos;
            if (q < p1) {
                av[u] = reinterpret_cast<const uint2 *>(A)[q * C4 + c4];
                d[u] = reinterpret_cast<const float4 *>(dA)[q * C4 + c4];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long q = p + (long long)u * per_pos;
            if (q < p1) {
                const uint16_t *ah = reinterpret_cast<const uint16_t *>(&av[u]);
                const float o[4] = {d[u].x * (bf2f(ah[0]) > 0.f ? 1.f : kLeak), d[u].y * (bf2f(ah[1]) > 0.f ? 1.f : kLeak),
                                    d[u].z * (bf2f(ah[2]) > 0.f ? 1.f : kLeak), d[u].w * (bf2f(ah[3]) > 0.f ? 1.f : kLeak)};
                uint16_t h[4] = {f2bf(o[0]), f2bf(o[1]), f2bf(o[2]), f2bf(o[3])};
                reinterpret_cast<uint2 *>(H)[q * C4 + c4] = __builtin_bit_cast(uint2, h);
                a.x += o[0];
                a.y += o[1];
                a.z += o[2];
                a.w += o[3];
            }
        }
    }
}

// Backward of the leaky ReLU with the bias gradient folded in: dP = dA act'(A)
// goes out as its bf16 copy H, and each of the kBiasBlocks blocks sums dP over a
// contiguous range of positions per channel (threads take 4 channels of one
// position; C / 4 divides 256), for k_bias_final.  With G set, the incoming
// gradient is dA + scale * G at the parent cell (G has side D/2: the adjoint of
// pooling added on the fly, the same fused multiply-add as k_up2_add4).
// (one instance per form, so that each gets its own register allocation: in one
// kernel the unrolled form slowed the G form 47.2 -> 50.5 us).  DBF (with G): dA
// is bf16.
template <bool HAS_G, bool DBF = false>
__global__ __launch_bounds__(256) void k_act_bwd_cast_bias(const uint16_t *__restrict__ A, const void *__restrict__ dAv,
                                                           uint16_t *__restrict__ H, float *part, long long P, int C,
                                                           const float *__restrict__ G, float scale, int D)
{
    static_assert(HAS_G || !DBF, "bf16 dA with the up-add form only");
    const float *__restrict__ dA = reinterpret_cast<const float *>(dAv);
    __shared__ float4 s[256];
    const int C4 = C / 4, per_pos = 256 / C4;
    const int c4 = threadIdx.x % C4, sub = threadIdx.x / C4;
    const long long per = (P + gridDim.x - 1) / gridDim.x;
    const long long p0 = (long long)blockIdx.x * per, p1 = min(P, p0 + per);
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (!HAS_G) {
        act_bwd_bias_pass4(A, dA, H, p0 + sub, p1, per_pos, C4, c4, a);
    } else {
        for (long long p = p0 + sub; p < p1; p += per_pos) {
            const long long i = p * C4 + c4;
            const uint2 av = reinterpret_cast<const uint2 *>(A)[i];
            float4 d;
            if constexpr (DBF) {
                const uint2 dv = reinterpret_cast<const uint2 *>(dAv)[i];
                d = make_float4(__uint_as_float(dv.x << 16), __uint_as_float(dv.x & 0xffff0000u),
                                __uint_as_float(dv.y << 16), __uint_as_float(dv.y & 0xffff0000u));
            } else {
                d = reinterpret_cast<const float4 *>(dA)[i];
            }
            const int pp = (int)p, Dh = D / 2;  // P = B D^3 < 2^31
            const int x = pp % D, y = (pp / D) % D, z = (pp / (D * D)) % D, b = pp / (D * D * D);
            const float4 v = reinterpret_cast<const float4 *>(G)[(((b * Dh + z / 2) * Dh + y / 2) * Dh + x / 2) * C4 + c4];
            d = make_float4(__builtin_fmaf(scale, v.x, d.x), __builtin_fmaf(scale, v.y, d.y),
                            __builtin_fmaf(scale, v.z, d.z), __builtin_fmaf(scale, v.w, d.w));
            const uint16_t *ah = reinterpret_cast<const uint16_t *>(&av);
            const float o[4] = {d.x * (bf2f(ah[0]) > 0.f ? 1.f : kLeak), d.y * (bf2f(ah[1]) > 0.f ? 1.f : kLeak),
                                d.z * (bf2f(ah[2]) > 0.f ? 1.f : kLeak), d.w * (bf2f(ah[3]) > 0.f ? 1.f : kLeak)};
            uint16_t h[4] = {f2bf(o[0]), f2bf(o[1]), f2bf(o[2]), f2bf(o[3])};
            reinterpret_cast<uint2 *>(H)[i] = __builtin_bit_cast(uint2, h);
            a.x += o[0];
            a.y += o[1];
            a.z += o[2];
            a.w += o[3];
        }
    }
    s[threadIdx.x] = a;
    __syncthreads();
    if ((int)threadIdx.x < C4) {
        float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int k = 0; k < per_pos; ++k) {
            const float4 v = s[threadIdx.x + k * C4];
            t.x += v.x;
            t.y += v.y;
            t.z += v.z;
            t.w += v.w;
        }
        *reinterpret_cast<float4 *>(part + (size_t)blockIdx.x * C + 4 * threadIdx.x) = t;
    }
}

// one block per channel: each thread sums a strided subset of the block partials,
// then a fixed tree over the threads
// the block partials of one channel in a fixed order: each thread a strided
// subset, then a fixed tree over the threads
__device__ __forceinline__ float bias_channel_sum(const float *part, int nblk, int C, int c)
{
    __shared__ float s[256];
    float t = 0.f;
    for (int b = threadIdx.x; b < nblk; b += 256) t += part[(size_t)b * C + c];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

// gp5_nn_bias_final_multi: workgroup b finishes channel b of the segment whose
// channel range holds b (the same sum as k_bias_final)
struct BiasMulti {
    gp5_nn_bias_seg s[GP5_NN_BIAS_MAX_SEGS];
    int n;
};
__global__ __launch_bounds__(256) void k_bias_final_multi(BiasMulti a)
{
    int b = blockIdx.x, k = 0;
    while (k + 1 < a.n && b >= a.s[k].C) b -= a.s[k++].C;  // block-uniform
    const float v = bias_channel_sum(a.s[k].part, kBiasBlocks, a.s[k].C, b);
    if (threadIdx.x == 0) a.s[k].db[b] = v;
}

__global__ __launch_bounds__(256) void k_bias_final(const float *part, float *db, int nblk, int C)
{
    __shared__ float s[256];
    const int c = blockIdx.x;
    float t = 0.f;
    for (int b = threadIdx.x; b < nblk; b += 256) t += part[(size_t)b * C + c];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = s[0];
}

// bf16 pooling and upsampling, 4 channels (8 bytes) per thread; same arithmetic
// as the scalar templates below (the 8 children summed in order, then scaled)
// (I: the index type; 32 bits when B D^3 C < 2^31)
template <typename I>
__global__ void k_pool2_bf16x4(const uint16_t *__restrict__ X, uint16_t *__restrict__ Y, int B, int D, int C,
                               float scale)
{
    const int Dh = D / 2, C4 = C / 4;
    const I n = (I)B * Dh * Dh * Dh * C4;
    for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
        const int c4 = (int)(i % C4);
        const I p = i / C4;
        const int x = (int)(p % Dh), y = (int)((p / Dh) % Dh), z = (int)((p / ((I)Dh * Dh)) % Dh);
        const I b = p / ((I)Dh * Dh * Dh);
        float s[4] = {0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; ++k) {
            const I q = ((b * D + 2 * z + (k >> 2)) * D + 2 * y + ((k >> 1) & 1)) * D + 2 * x + (k & 1);
            const bf16x4 v = ld4(X + (size_t)q * C + 4 * c4);
#pragma unroll
            for (int e = 0; e < 4; ++e) s[e] += bf2f((uint16_t)v[e]);
        }
        bf16x4 o;
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = (short)f2bf(s[e] * (0.125f * scale));
        *reinterpret_cast<bf16x4 *>(Y + (size_t)p * C + 4 * c4) = o;
    }
}

__global__ void k_up2_bf16x4(const uint16_t *__restrict__ X, uint16_t *__restrict__ Y, int B, int D, int C, float scale)
{
    const int D2 = 2 * D, C4 = C / 4;
    const long long n = (long long)B * D2 * D2 * D2 * C4;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int c4 = (int)(i % C4);
        const long long p = i / C4;
        const int x = (int)(p % D2), y = (int)((p / D2) % D2), z = (int)((p / ((long long)D2 * D2)) % D2);
        const long long b = p / ((long long)D2 * D2 * D2);
        const long long q = ((b * D + z / 2) * D + y / 2) * D + x / 2;
        bf16x4 v = ld4(X + (size_t)q * C + 4 * c4);
        if (scale != 1.0f) {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = (short)f2bf(scale * bf2f((uint16_t)v[e]));
        }
        *reinterpret_cast<bf16x4 *>(Y + (size_t)p * C + 4 * c4) = v;
    }
}

// element access of the float pool / upsample kernels (the bf16 ones move 4
// channels per thread)
template <typename T> __device__ __forceinline__ float ldv(const T *p);
template <> __device__ __forceinline__ float ldv<float>(const float *p) { return *p; }
template <typename T> __device__ __forceinline__ void stv(T *p, float v);
template <> __device__ __forceinline__ void stv<float>(float *p, float v) { *p = v; }

// Y (side D/2) = scale * mean of the 8 children
template <typename T>
__global__ void k_pool2(const T *X, T *Y, int B, int D, int C, float scale)
{
    const int Dh = D / 2;
    const long long n = (long long)B * Dh * Dh * Dh * C;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long long p = i / C;
        const int x = (int)(p % Dh), y = (int)((p / Dh) % Dh), z = (int)((p / ((long long)Dh * Dh)) % Dh);
        const long long b = p / ((long long)Dh * Dh * Dh);
        float s = 0.f;
        for (int k = 0; k < 8; ++k) {
            const long long q = ((b * D + 2 * z + (k >> 2)) * D + 2 * y + ((k >> 1) & 1)) * D + 2 * x + (k & 1);
            s += ldv(X + (size_t)q * C + c);
        }
        stv(Y + i, s * (0.125f * scale));
    }
}

// k_pool2 for float with C a multiple of 4 and B D^3 C < 2^31: four channels per thread
// and 32-bit index arithmetic, the same sums in the same order (children k = 0..7)
__global__ void k_pool2_f32x4(const float *__restrict__ X, float *__restrict__ Y, int B, int D, int C, float scale)
{
    const int Dh = D / 2, C4 = C / 4;
    const int n = B * Dh * Dh * Dh * C4;
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) {
        const int c4 = i % C4, p = i / C4;
        const int x = p % Dh, y = (p / Dh) % Dh, z = (p / (Dh * Dh)) % Dh, b = p / (Dh * Dh * Dh);
        float4 v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int q = ((b * D + 2 * z + (k >> 2)) * D + 2 * y + ((k >> 1) & 1)) * D + 2 * x + (k & 1);
            v[k] = reinterpret_cast<const float4 *>(X)[(size_t)q * C4 + c4];
        }
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            s.x += v[k].x;
            s.y += v[k].y;
            s.z += v[k].z;
            s.w += v[k].w;
        }
        const float f = 0.125f * scale;
        reinterpret_cast<float4 *>(Y)[i] = make_float4(s.x * f, s.y * f, s.z * f, s.w * f);
    }
}

// Y (side 2D) = scale * X at the parent
template <typename T>
__global__ void k_up2(const T *X, T *Y, int B, int D, int C, float scale)
{
    const int D2 = 2 * D;
    const long long n = (long long)B * D2 * D2 * D2 * C;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long long p = i / C;
        const int x = (int)(p % D2), y = (int)((p / D2) % D2), z = (int)((p / ((long long)D2 * D2)) % D2);
        const long long b = p / ((long long)D2 * D2 * D2);
        const long long q = ((b * D + z / 2) * D + y / 2) * D + x / 2;
        stv(Y + i, scale * ldv(X + (size_t)q * C + c));
    }
}

__global__ void k_act_bwd(const uint16_t *A, const float *dA, float *dP, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dP[i] = dA[i] * (bf2f(A[i]) > 0.f ? 1.f : kLeak);
}

// dP = dA act'(A) and its bf16 copy H in one pass, 4 elements per thread (n4 = n / 4);
// dP may be dA
__global__ void k_act_bwd_cast4(const uint16_t *__restrict__ A, const float *dA, float *dP,
                                uint16_t *__restrict__ H, long long n4)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
        const uint2 a = reinterpret_cast<const uint2 *>(A)[i];
        const float4 d = reinterpret_cast<const float4 *>(dA)[i];
        const uint16_t *ah = reinterpret_cast<const uint16_t *>(&a);
        const float o[4] = {d.x * (bf2f(ah[0]) > 0.f ? 1.f : kLeak), d.y * (bf2f(ah[1]) > 0.f ? 1.f : kLeak),
                            d.z * (bf2f(ah[2]) > 0.f ? 1.f : kLeak), d.w * (bf2f(ah[3]) > 0.f ? 1.f : kLeak)};
        reinterpret_cast<float4 *>(dP)[i] = make_float4(o[0], o[1], o[2], o[3]);
        uint16_t h[4] = {f2bf(o[0]), f2bf(o[1]), f2bf(o[2]), f2bf(o[3])};
        reinterpret_cast<uint2 *>(H)[i] = __builtin_bit_cast(uint2, h);
    }
}

// Y (side 2D) += scale * X at the parent, 4 channels per thread (C % 4 == 0)
__global__ void k_up2_add4(const float *__restrict__ X, float *Y, int B, int D, int C, float scale)
{
    const int D2 = 2 * D, C4 = C / 4;
    const long long n = (long long)B * D2 * D2 * D2 * C4;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int c4 = (int)(i % C4);
        const long long p = i / C4;
        const int x = (int)(p % D2), y = (int)((p / D2) % D2), z = (int)((p / ((long long)D2 * D2)) % D2);
        const long long b = p / ((long long)D2 * D2 * D2);
        const long long q = ((b * D + z / 2) * D + y / 2) * D + x / 2;
        const float4 v = reinterpret_cast<const float4 *>(X)[q * C4 + c4];
        float4 &o = reinterpret_cast<float4 *>(Y)[i];
        o = make_float4(__builtin_fmaf(scale, v.x, o.x), __builtin_fmaf(scale, v.y, o.y), __builtin_fmaf(scale, v.z, o.z),
                        __builtin_fmaf(scale, v.w, o.w));
    }
}

__global__ void k_add(float *d, const float *s, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] += s[i];
}

__global__ void k_cast(const float *s, uint16_t *d, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] = f2bf(s[i]);
}

// Input channels of the droplet warm start on the scaled grid u in [-ub, ub]^3,
// element (i, j, k) of a C-ordered [D][D][D][16] array: 0 the flat-top baseline
// 0.5 (1 - tanh((r - 1) s)), 1 the trap shape a0 u_i^2 + a1 u_j^2 + a2 u_k^2,
// 2..6 constants, 7 r / ub, 8..15 zero, in float32 with the operation order of
// the host construction, then rounded to bf16; base = channel 0 in float.
struct FeatArgs {
    int D;
    double ub;
    float s, a0, a1, a2, ub_f;
    float c[5];
};

__device__ __forceinline__ float u_axis_f(int i, int D, double ub) { return (float)(-ub + (i + 0.5) * (2.0 * ub / D)); }

__global__ __launch_bounds__(256) void k_droplet_features(uint16_t *X, float *base, FeatArgs a)
{
    const long long P = (long long)a.D * a.D * a.D;
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const int k = (int)(p % a.D), j = (int)((p / a.D) % a.D), i = (int)(p / ((long long)a.D * a.D));
    const float ui = u_axis_f(i, a.D, a.ub), uj = u_axis_f(j, a.D, a.ub), uk = u_axis_f(k, a.D, a.ub);
    const float u2i = __fmul_rn(ui, ui), u2j = __fmul_rn(uj, uj), u2k = __fmul_rn(uk, uk);
    const float r = sqrtf(__fadd_rn(__fadd_rn(u2i, u2j), u2k));
    const float t = tanhf(__fmul_rn(__fsub_rn(r, 1.0f), a.s));
    const float c0 = __fmul_rn(0.5f, __fsub_rn(1.0f, t));
    const float c1 = __fadd_rn(__fadd_rn(__fmul_rn(a.a0, u2i), __fmul_rn(a.a1, u2j)), __fmul_rn(a.a2, u2k));
    uint16_t h[16];
    h[0] = f2bf(c0);
    h[1] = f2bf(c1);
#pragma unroll
    for (int q = 0; q < 5; ++q) h[2 + q] = f2bf(a.c[q]);
    h[7] = f2bf(__fdiv_rn(r, a.ub_f));
#pragma unroll
    for (int q = 8; q < 16; ++q) h[q] = 0;
    uint4 *dst = reinterpret_cast<uint4 *>(X + p * 16);
    uint4 w0, w1;
    w0.x = h[0] | (uint32_t)h[1] << 16;
    w0.y = h[2] | (uint32_t)h[3] << 16;
    w0.z = h[4] | (uint32_t)h[5] << 16;
    w0.w = h[6] | (uint32_t)h[7] << 16;
    w1.x = w1.y = w1.z = w1.w = 0u;
    dst[0] = w0;
    dst[1] = w1;
    base[p] = c0;
}

// y[p] = base[p] + Yf[p * C] (the network's prediction: baseline plus channel 0)
__global__ __launch_bounds__(256) void k_predict_out(const float *Yf, int C, const float *base, float *y, long long P)
{
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p < P) y[p] = base[p] + Yf[p * C];
}

__global__ void k_fill0(uint32_t *p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// The prediction is y = base + Yf[p * ldy] (channel 0 of rows of ldy floats).
// dYf and H (each optional) receive the gradient rows of C channels in float and
// bf16; partial[b] and partial[gridDim.x + b] the block sums of e^2 and of the
// channel-0 gradient 2 e / P (in double)
__global__ __launch_bounds__(256) void k_mse(const float *Yf, int ldy, int C, const float *base, const float *target,
                                             float *dYf, uint16_t *H, int hc, long long P, double *partial)
{
    __shared__ double s[256], sg[256];
    double a = 0.0, ag = 0.0;
    const float inv = 2.0f / (float)P;
    const long long st = (long long)gridDim.x * 256;
    // position p's error e: the sums, then the gradient row (float and/or bf16)
    auto put = [&](long long p, float e) {
        a += (double)e * e;
        ag += (double)(inv * e);
        if (dYf && C % 4 == 0) {  // the row as 16-byte stores (rows are 16-byte aligned)
            float4 *row = reinterpret_cast<float4 *>(dYf + (size_t)p * C);
            row[0] = make_float4(inv * e, 0.f, 0.f, 0.f);
            for (int c = 1; c < C / 4; ++c) row[c] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (dYf) {
            dYf[(size_t)p * C] = inv * e;
            for (int c = 1; c < C; ++c) dYf[(size_t)p * C + c] = 0.f;
        }
        if (H && hc == 1) {  // channel 0 alone, one bf16 per position
            H[p] = f2bf(inv * e);
        } else if (H) {  // C % 8 == 0 (checked on the host): 16-byte stores
            uint4 *hrow = reinterpret_cast<uint4 *>(H + (size_t)p * C);
            hrow[0] = make_uint4((uint32_t)f2bf(inv * e), 0u, 0u, 0u);
            for (int c = 1; c < C / 8; ++c) hrow[c] = make_uint4(0u, 0u, 0u, 0u);
        }
    };
    long long p0 = (long long)blockIdx.x * 256 + threadIdx.x;
    // four positions' loads in flight, then their sums in position order
    for (; p0 + 3 * st < P; p0 += 4 * st) {
        float y[4], t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long p = p0 + k * st;
            y[k] = base[p] + Yf[(size_t)p * ldy];
            t[k] = target[p];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) put(p0 + k * st, y[k] - t[k]);
    }
    for (long long p = p0; p < P; p += st) put(p, base[p] + Yf[(size_t)p * ldy] - target[p]);
    s[threadIdx.x] = a;
    sg[threadIdx.x] = ag;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[threadIdx.x] += s[threadIdx.x + w];
            sg[threadIdx.x] += sg[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s[0];
        partial[gridDim.x + blockIdx.x] = sg[0];
    }
}

// one block: the block sums of k_mse added in a fixed order (double); the loss
// (their mean) to loss[0], also added to acc[0] when set, and the output layer's
// bias gradient (channel 0 the gradient sum, the other C - 1 channels 0) to db
// when set
__global__ __launch_bounds__(256) void k_mse_final(const double *partial, int blocks, long long P, double *loss, float *db,
                                                   int C, double *acc)
{
    __shared__ double s[256], sg[256];
    double a = 0.0, ag = 0.0;
    for (int b = threadIdx.x; b < blocks; b += 256) {
        a += partial[b];
        ag += partial[blocks + b];
    }
    s[threadIdx.x] = a;
    sg[threadIdx.x] = ag;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[threadIdx.x] += s[threadIdx.x + w];
            sg[threadIdx.x] += sg[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        loss[0] = s[0] / (double)P;
        if (acc) acc[0] += s[0] / (double)P;
    }
    if (db)
        for (int c = threadIdx.x; c < C; c += 256) db[c] = c == 0 ? (float)sg[0] : 0.f;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// gp5_nn_adam_multi: blockIdx.y selects the tensor (a block-uniform index into
// the argument table), the blocks along x stride over its elements.
struct AdamMulti {
    gp5_nn_adam_seg s[GP5_NN_ADAM_MAX_SEGS];
    float lr, b1, b2, eps, c1, c2;
};

__global__ __launch_bounds__(256) void k_adam_multi(AdamMulti a)
{
    const gp5_nn_adam_seg &sg = a.s[blockIdx.y];
    float *p = sg.p;
    const float *g = sg.g;
    float *m = sg.m, *v = sg.v;
    uint16_t *wf = (uint16_t *)sg.wf, *wb = (uint16_t *)sg.wb;
    const long long n = sg.n;
    const int Cin = sg.cin, Cout = sg.cout;
    const float lr = a.lr, b1 = a.b1, b2 = a.b2, eps = a.eps, c1 = a.c1, c2 = a.c2;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        float pi = p[i];
        pi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        if (wf || wb) {  // a weight tensor: 27 Cin Cout < 2^31 elements, 32-bit index arithmetic
            const int ii = (int)i, co = ii % Cout, ci = (ii / Cout) % Cin, tap = ii / (Cin * Cout);
            const uint16_t h = f2bf(pi);
            if (wb) wb[i] = h;
            if (wf) wf[((size_t)tap * Cout + co) * Cin + ci] = h;
        }
    }
}

__global__ void k_relayout(const float *W, uint16_t *Wf, uint16_t *Wb, int Cin, int Cout)
{
    const long long n = 27LL * Cin * Cout;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int co = (int)(i % Cout);
        const int ci = (int)((i / Cout) % Cin);
        const int tap = (int)(i / ((long long)Cin * Cout));
        const uint16_t h = f2bf(W[i]);
        if (Wb) Wb[i] = h;
        if (Wf) Wf[((size_t)tap * Cout + co) * Cin + ci] = h;
    }
}

unsigned grid_for(long long n, int threads = 256)
{
    return (unsigned)std::max<long long>(1, std::min<long long>((n + threads - 1) / threads, 16384));
}

int hip_fail(hipError_t e, const char *what)
{
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return set_error(GP5_ERR_HIP, buf);
}

#define NNCHK(expr, what)                                  \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(e_, what);   \
    } while (0)

int check_channels(int c, const char *what)
{
    if (c < 16 || c % 16) {
        char buf[128];
        snprintf(buf, sizeof buf, "%s = %d: channel counts must be positive multiples of 16", what, c);
        return set_error(GP5_ERR_ARG, buf);
    }
    return GP5_OK;
}

}  // namespace
}  // namespace gp5

using namespace gp5;

// Launch of k_conv_zs with `per_run` x `gy` workgroups per plane run.  The number of
// plane runs per cube side is the one that fills the device's workgroup slots once
// (CUs x resident workgroups of this instance), between 1 and D/2, and for the
// pooled form one that pairs planes; `runs` > 0 overrides it.  Measured at 64^3,
// batch 8 (one full round of workgroups against the 1024-workgroup rule of before):
// 16 -> 16 62.3 -> 57.9 us at 6 runs, 32 -> 16 upsampled 91.1 -> 85.4 us at 4 runs.
template <int C, int K, bool W, bool S, bool PL = false, bool AB = false>
static int zs_launch(int per_run, int gy, size_t shm, const uint16_t *x, const uint16_t *w, const float *bias,
                     uint16_t *a, const uint16_t *sk, uint16_t *y, float *Yf, int P, int D, int Cout, int mirror,
                     int act, int runs, int up, float *bpart = nullptr)
{
    static size_t shm_set = 65536;  // the default limit of dynamic LDS
    const void *fn = (const void *)k_conv_zs<C, K, W, S, PL, AB>;
    if (shm > shm_set) {
        NNCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm), "conv3d LDS size");
        shm_set = shm;
    }
    int nzs = runs;
    if (nzs <= 0) {
        static size_t occ_shm = 0;
        static int slots = 0;
        if (occ_shm != shm) {
            int dev = 0, cus = 0, per_cu = 0;
            NNCHK(hipGetDevice(&dev), "conv3d device");
            NNCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "conv3d CU count");
            NNCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, shm), "conv3d occupancy");
            slots = std::max(1, cus * per_cu);
            occ_shm = shm;
        }
        nzs = std::max(1, std::min(D / 2, slots / (per_run * gy)));
        if (PL)
            while (nzs > 1 && D % (2 * nzs)) --nzs;
    }
    if (AB) nzs = std::max(1, std::min(nzs, kBiasBlocks / per_run));  // one bias partial row per workgroup
    hipLaunchKernelGGL((k_conv_zs<C, K, W, S, PL, AB>), dim3((unsigned)(per_run * nzs), (unsigned)gy), dim3(256), shm,
                       nn_stream(), x, w, bias, a, sk, y, Yf, P, D, Cout, mirror, act, nzs, up, bpart);
    return GP5_OK;
}

// Launch of k_conv_up2p (X of side D/2, 32 or 64 channels, D = 32, 64 or 128; slabs
// of at most 5 or 7 x 256 units of 16 bytes): coarse
// plane runs that fill the device's workgroup slots once, as for k_conv_zs, or
// `runs` > 0 of them.
template <int CIN>
static int up2p_launch(const uint16_t *x, const uint16_t *w, const float *bias, uint16_t *a, const uint16_t *sk,
                       uint16_t *y, float *Yf, int B, int P, int D, int Cout, int act, int runs)
{
    const int Dc = D / 2, Rb = (CIN == 32 ? 128 : (CIN == 64 ? 64 : 32)) / Dc;
    const size_t shm = (size_t)3 * (Rb + 2) * (Dc + 2) * (CIN / 8) * 16;
    const void *fn = (const void *)k_conv_up2p<CIN>;
    static size_t shm_set = 65536;  // the default limit of dynamic LDS
    if (shm > shm_set) {
        NNCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm), "conv3d LDS size");
        shm_set = shm;
    }
    const int per_run = B * (Dc / Rb), gy = (Cout / 16) * 4;
    int nzs = runs;
    if (nzs <= 0) {
        static size_t occ_shm = 0;
        static int slots = 0;
        if (occ_shm != shm) {
            int dev = 0, cus = 0, per_cu = 0;
            NNCHK(hipGetDevice(&dev), "conv3d device");
            NNCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "conv3d CU count");
            NNCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, shm), "conv3d occupancy");
            slots = std::max(1, cus * per_cu);
            occ_shm = shm;
        }
        nzs = slots / (per_run * gy);
    }
    nzs = std::max(1, std::min(Dc / 2, nzs));
    hipLaunchKernelGGL(k_conv_up2p<CIN>, dim3((unsigned)(per_run * nzs), (unsigned)gy), dim3(256), shm, nn_stream(), x,
                       w, bias, a, sk, y, Yf, P, D, Cout, act, nzs);
    NNCHK(hipGetLastError(), "conv3d launch");
    return GP5_OK;
}

// Launch of k_conv_pool2p (dY of side D with CO = 16 or 32 channels, D = 32, 64 or
// 128, and D <= 64 for CO = 32): coarse plane runs that fill the device's workgroup
// slots once, or `runs` > 0 of them.
template <int CO>
static int pool2p_launch(const uint16_t *dy, const uint16_t *w, float *dX, int B, int P, int D, int CI, int runs)
{
    constexpr int NPOS = CO == 16 ? 64 : 32, U = CO / 8, MT = CO == 16 ? 4 : 2;
    const int Dc = D / 2, Rb = NPOS / Dc;
    const int SUF = (2 * Rb + 2) * (D + 2) * U;
    const size_t shm = (size_t)4 * SUF * 16 + (size_t)4 * MT * 4 * 64 * 4;
    const void *fn = (const void *)k_conv_pool2p<CO>;
    static size_t shm_set = 65536;  // the default limit of dynamic LDS
    if (shm > shm_set) {
        NNCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm), "conv3d LDS size");
        shm_set = shm;
    }
    const int per_run = B * (Dc / Rb), gy = CI / 16;
    int nzs = runs;
    if (nzs <= 0) {
        static size_t occ_shm = 0;
        static int slots = 0;
        if (occ_shm != shm) {
            int dev = 0, cus = 0, per_cu = 0;
            NNCHK(hipGetDevice(&dev), "conv3d device");
            NNCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "conv3d CU count");
            NNCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, shm), "conv3d occupancy");
            slots = std::max(1, cus * per_cu);
            occ_shm = shm;
        }
        nzs = slots / (per_run * gy);
    }
    nzs = std::max(1, std::min(Dc / 2, nzs));
    hipLaunchKernelGGL(k_conv_pool2p<CO>, dim3((unsigned)(per_run * nzs), (unsigned)gy), dim3(256), shm, nn_stream(),
                       dy, w, dX, P, D, CI, nzs);
    NNCHK(hipGetLastError(), "conv3d launch");
    return GP5_OK;
}

extern "C" {

int gp5_nn_malloc(void **p, size_t bytes)
{
    if (!p) return set_error(GP5_ERR_ARG, "NULL argument");
    NNCHK(hipMalloc(p, std::max<size_t>(bytes, 16)), "gp5_nn_malloc");
    return GP5_OK;
}

int gp5_nn_free(void *p)
{
    if (p) NNCHK(hipFree(p), "gp5_nn_free");
    return GP5_OK;
}

// Copies run in order on the thread's stream; upload and download return when
// theirs is done.
int gp5_nn_upload(void *dst, const void *host, size_t bytes)
{
    NNCHK(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, nn_stream()), "gp5_nn_upload");
    NNCHK(hipStreamSynchronize(nn_stream()), "gp5_nn_upload");
    return GP5_OK;
}

int gp5_nn_download(void *host, const void *src, size_t bytes)
{
    NNCHK(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, nn_stream()), "gp5_nn_download");
    NNCHK(hipStreamSynchronize(nn_stream()), "gp5_nn_download");
    return GP5_OK;
}

int gp5_nn_copy(void *dst, const void *src, size_t bytes)
{
    NNCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, nn_stream()), "gp5_nn_copy");
    return GP5_OK;
}

int gp5_nn_stream_create(void **stream)
{
    if (!stream) return set_error(GP5_ERR_ARG, "NULL argument");
    hipStream_t s = nullptr;
    NNCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "gp5_nn_stream_create");
    *stream = (void *)s;
    return GP5_OK;
}

int gp5_nn_stream_destroy(void *stream)
{
    if (stream) NNCHK(hipStreamDestroy((hipStream_t)stream), "gp5_nn_stream_destroy");
    return GP5_OK;
}

int gp5_nn_set_stream(void *stream)
{
    g_stream = (hipStream_t)stream;
    return GP5_OK;
}

int gp5_nn_event_create(void **event)
{
    if (!event) return set_error(GP5_ERR_ARG, "NULL argument");
    hipEvent_t e = nullptr;
    NNCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming), "gp5_nn_event_create");
    *event = (void *)e;
    return GP5_OK;
}

int gp5_nn_event_destroy(void *event)
{
    if (event) NNCHK(hipEventDestroy((hipEvent_t)event), "gp5_nn_event_destroy");
    return GP5_OK;
}

int gp5_nn_event_record(void *event, void *stream)
{
    if (!event) return set_error(GP5_ERR_ARG, "NULL argument");
    NNCHK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "gp5_nn_event_record");
    return GP5_OK;
}

int gp5_nn_stream_wait(void *stream, void *event)
{
    if (!event) return set_error(GP5_ERR_ARG, "NULL argument");
    NNCHK(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0), "gp5_nn_stream_wait");
    return GP5_OK;
}

int gp5_nn_zero(void *p, size_t bytes)
{
    if (bytes % 4) return set_error(GP5_ERR_ARG, "gp5_nn_zero: size must be a multiple of 4 bytes");
    const long long n = (long long)(bytes / 4);
    hipLaunchKernelGGL(k_fill0, dim3(grid_for(n)), dim3(256), 0, nn_stream(), (uint32_t *)p, n);
    NNCHK(hipGetLastError(), "fill");
    return GP5_OK;
}

int gp5_nn_sync(void)
{
    NNCHK(hipDeviceSynchronize(), "gp5_nn_sync");
    return GP5_OK;
}

// up = 1: X has side D/2 and the convolution reads its nearest upsampling (only
// the LDS-staged and z-sliding kernels do that)
// The float partial sums of split convolutions, in a process-wide buffer that
// only grows (the kernels of successive calls run in order on one stream).
static int conv_split_scratch(size_t floats, float **out)
{
    static std::mutex mu;
    static float *buf = nullptr;
    static size_t cap = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (cap < floats) {
        if (buf) NNCHK(hipFree(buf), "conv split scratch");
        buf = nullptr;
        cap = 0;
        NNCHK(hipMalloc(&buf, floats * sizeof(float)), "conv split scratch");
        cap = floats;
    }
    *out = buf;
    return GP5_OK;
}

// pool = 1: Yf (side D/2) receives the 2x2x2 child sums of the convolution (the
// POOL form of k_conv_zs; only where that kernel applies, else an error)
// bpart set (gp5_nn_conv3d_act_bwd): where the z-sliding kernel takes the shape
// with one output block per wave and at most kBiasBlocks workgroups per plane run,
// its ABWD form runs (A read, Y = H written, the bias partials to bpart) and
// *fused is set; otherwise only Yf is written.
static int conv3d_run(const void *X, const void *Wk, const float *bias, void *A, const void *skip, void *Y,
                      float *Yf, int B, int D, int Cin, int Cout, int mirror, int act, int up, int pool = 0,
                      float *bpart = nullptr, int *fused = nullptr)
{
    if (!X || !Wk) return set_error(GP5_ERR_ARG, "NULL input or weights");
    if (skip && !Y) return set_error(GP5_ERR_ARG, "a skip input needs the Y output");
    if (B < 1 || D < 1) return set_error(GP5_ERR_ARG, "empty tensor");
    int rc = check_channels(Cin, "Cin");
    if (rc == GP5_OK) rc = check_channels(Cout, "Cout");
    if (rc != GP5_OK) return rc;
    const long long P = (long long)B * D * D * D;
    // 32-bit byte offsets in the kernel (halo included) and element offsets of the outputs
    const long long halo = ((long long)D * D + D + 1) * Cin * 2;
    if (halo + P * Cin * 2 > 0x7fffffffLL || P * Cout * 4 > 0x7fffffffLL || 27LL * Cin * Cout * 2 > 0x7fffffffLL)
        return set_error(GP5_ERR_ARG, "conv3d: tensors must stay below 2 GiB (B*D^3*C elements)");
    // Output channels in groups of NB 16-wide blocks (a divisor of Cout / 16, at
    // most 8; blockIdx.y is the group) and MT position tiles per wave: the widest
    // group and the most tiles that still give about 2048 waves, so that small
    // cubes fill the machine.
    const int nblk = Cout / 16;
    int NB = 1, MT = 1;
    bool found = false;
    for (int nb = 8; nb >= 1 && !found; --nb) {
        if (nblk % nb) continue;
        const int mts[2] = {nb == 1 ? 4 : (nb <= 4 ? 2 : 1), 1};
        for (int k = 0; k < 2 && !found; ++k) {
            const long long waves = (P + 16LL * mts[k] - 1) / (16LL * mts[k]) * (nblk / nb);
            if (waves >= 2048) {
                NB = nb;
                MT = mts[k];
                found = true;
            }
        }
    }
    const uint16_t *x = (const uint16_t *)X, *w = (const uint16_t *)Wk, *sk = (const uint16_t *)skip;
    uint16_t *a = (uint16_t *)A, *y = (uint16_t *)Y;
    const dim3 block(256);
    const bool pow2 = D >= 8 && D <= 128 && (D & (D - 1)) == 0;
    // z-sliding kernel: 16 or 32 input channels, D a power of two in [16, 128], the
    // longest plane run ZL (a power of two, at least 2) that keeps 1024 workgroups,
    // as long as there are at least 512.  GP5_CONV_ZS=0 turns it off.
    // the upsampled input by output parity (k_conv_up2p): 32 or 64 input channels at
    // D = 32, 64 or 128, or 128 at D = 16, 32 or 64 (a workgroup covers whole coarse
    // rows), forward taps; GP5_CONV_UP2P=0 turns it off, GP5_UP2P_RUNS sets the coarse
    // plane runs (measurement)
    const char *pp_env = getenv("GP5_CONV_UP2P");
    // (Dc = 8, a tile over two coarse rows, only with 128 channels)
    if (up && !mirror && !pool && !bpart && !(pp_env && pp_env[0] == '0') &&
        (((Cin == 32 || Cin == 64) && (D == 32 || D == 64 || D == 128)) || (Cin == 128 && (D == 16 || D == 32 || D == 64)))) {
        const char *pr = getenv("GP5_UP2P_RUNS");
        const int runs = pr ? atoi(pr) : 0;
        if (Cin == 32) return up2p_launch<32>(x, w, bias, a, sk, y, Yf, B, (int)P, D, Cout, act, runs);
        if (Cin == 64) return up2p_launch<64>(x, w, bias, a, sk, y, Yf, B, (int)P, D, Cout, act, runs);
        return up2p_launch<128>(x, w, bias, a, sk, y, Yf, B, (int)P, D, Cout, act, runs);
    }
    // the pooled data gradient as a stride-2 convolution of dY (k_conv_pool2p): 16 or
    // 32 dY channels, D = 32, 64 or 128 (at most 64 with 32 channels), mirrored taps;
    // GP5_CONV_POOL2P=0 turns it off, GP5_POOL2P_RUNS sets the coarse plane runs
    const char *p2_env = getenv("GP5_CONV_POOL2P");
    if (pool && mirror == 1 && !bias && !act && !A && !Y && Yf && (Cin == 16 || Cin == 32) &&
        (D == 32 || D == 64 || (D == 128 && Cin == 16)) && !(p2_env && p2_env[0] == '0')) {
        const char *pr = getenv("GP5_POOL2P_RUNS");
        const int runs = pr ? atoi(pr) : 0;
        if (Cin == 16) return pool2p_launch<16>(x, w, Yf, B, (int)P, D, Cout, runs);
        return pool2p_launch<32>(x, w, Yf, B, (int)P, D, Cout, runs);
    }
    const char *zs_env = getenv("GP5_CONV_ZS");
    if (pow2 && D >= 16 && (Cin == 16 || Cin == 32) && !(zs_env && zs_env[0] == '0')) {
        const int NBz = (Cin == 16 && nblk % 2 == 0) ? 2 : 1;
        const int Rb = 256 / D;
        const long long per_z = (long long)B * (D / Rb) * (nblk / NBz);  // workgroups per plane run
        int ZL = 0;
        for (int zl = D; zl >= 2; zl /= 2)
            if (!ZL && per_z * (D / zl) >= 1024) ZL = zl;
        if (!ZL && per_z * (D / 2) >= 512) ZL = 2;
        const int SU = (Rb + 2) * (D + 2) * (Cin / 8);
        if (ZL && SU <= 256 * zs_units_per_thread(Cin) && SU * 16 >= 4 * kEpiScratch) {
            // plane runs per cube side: from the occupancy in zs_launch, or GP5_ZS_RUNS
            // (measurement; for the pooled form only counts whose runs pair planes)
            int nzs = 0;
            const char *runs = getenv("GP5_ZS_RUNS");
            if (runs && atoi(runs) >= 1 && atoi(runs) <= D / 2 && (!pool || (D % (2 * atoi(runs)) == 0)))
                nzs = atoi(runs);
            const int per_run = (int)(per_z / (nblk / NBz)), gy = nblk / NBz;
            const size_t shm = (size_t)3 * SU * 16;
            // GP5_CONV_ZSV overrides the variant: bit 0 wave-private epilogue barriers, bit 1
            // fragment reads one tap ahead.  Defaults (measured at 64^3, batch 8): both for
            // 16 -> 16 (70.9 -> 62.5 us), the barriers only for 16 -> 32 and 32 input channels
            // (reads ahead: 87.9 -> 98.6 us on 32 -> 16)
            const char *v_env = getenv("GP5_CONV_ZSV");
            const int Vz = v_env ? atoi(v_env) & 3 : (Cin == 16 && NBz == 1 ? 3 : 1);
#define GP5_CONV_ZS_CASE(C, K, V)                                                                                \
    if (Cin == C && NBz == K && Vz == V)                                                                         \
        zrc = zs_launch<C, K, (V & 2) != 0, (V & 1) != 0>(per_run, gy, shm, x, w, bias, a, sk, y, Yf, (int)P, D, Cout, \
                                                          mirror, act, nzs, up);
#define GP5_CONV_ZS_V(C, K) GP5_CONV_ZS_CASE(C, K, 0) GP5_CONV_ZS_CASE(C, K, 1) GP5_CONV_ZS_CASE(C, K, 2) \
    GP5_CONV_ZS_CASE(C, K, 3)
            int zrc = GP5_OK;
            if (bpart && NBz == 1 && per_run <= kBiasBlocks) {  // one bias partial row per workgroup
                if (bias || act || skip || !A || !Y)
                    return set_error(GP5_ERR_ARG, "conv3d_act_bwd: no bias, activation or skip");
                if (Cin == 16)
                    zrc = zs_launch<16, 1, true, true, false, true>(per_run, gy, shm, x, w, nullptr, a, nullptr, y, Yf,
                                                                     (int)P, D, Cout, mirror, 0, nzs, up, bpart);
                else
                    zrc = zs_launch<32, 1, false, true, false, true>(per_run, gy, shm, x, w, nullptr, a, nullptr, y, Yf,
                                                                      (int)P, D, Cout, mirror, 0, nzs, up, bpart);
                if (zrc != GP5_OK) return zrc;
                NNCHK(hipGetLastError(), "conv3d launch");
                *fused = 1;
                return GP5_OK;
            }
            if (bpart) {  // the plain data gradient below: only Yf
                a = nullptr;
                y = nullptr;
            }
            if (pool) {
                if (D < 32 || bias || act || A || Y || !Yf)
                    return set_error(GP5_ERR_ARG, "conv3d_pool: needs 16 or 32 input channels, D = 32, 64 or 128, "
                                                  "and only the float output");
                if (Cin == 16 && NBz == 2)
                    zrc = zs_launch<16, 2, false, true, true>(per_run, gy, shm, x, w, nullptr, nullptr, nullptr, nullptr, Yf,
                                                               (int)P, D, Cout, mirror, 0, nzs, up);
                else if (Cin == 32)
                    zrc = zs_launch<32, 1, false, true, true>(per_run, gy, shm, x, w, nullptr, nullptr, nullptr, nullptr, Yf,
                                                               (int)P, D, Cout, mirror, 0, nzs, up);
                else
                    return set_error(GP5_ERR_ARG, "conv3d_pool: 16 input channels need a multiple of 32 output "
                                                  "channels");
                if (zrc != GP5_OK) return zrc;
                NNCHK(hipGetLastError(), "conv3d launch");
                return GP5_OK;
            }
            GP5_CONV_ZS_V(16, 1) GP5_CONV_ZS_V(16, 2) GP5_CONV_ZS_V(32, 1)
#undef GP5_CONV_ZS_V
            if (zrc != GP5_OK) return zrc;
            NNCHK(hipGetLastError(), "conv3d launch");
            return GP5_OK;
        }
    }
    if (pool) return set_error(GP5_ERR_ARG, "conv3d_pool: the shape does not take the z-sliding kernel");
    if (bpart) {  // not fused: only Yf
        a = nullptr;
        y = nullptr;
    }
    // LDS-staged kernel: D a power of two in [8, 128] (workgroups of 256 positions
    // are whole rows of one sample), channel chunk KC = 32 or 16 with a staged
    // region of at most 64 KB, output blocks per wave NB = 4, 2 or 1, the widest
    // that still gives 512 workgroups.  GP5_CONV_LDS=0 selects the direct kernel.
    const char *lds_env = getenv("GP5_CONV_LDS");
    if (pow2 && !(lds_env && lds_env[0] == '0')) {
        const int Rb = std::min(D, 256 / D), PZ = 256 / (D * Rb);
        const long long region = (long long)(PZ + 2) * (Rb + 2) * (D + 2);
        int KCl = 0;
        for (int kc : {32, 16})
            if (!KCl && Cin % kc == 0 && region * kc * 2 <= 65536 && region * kc * 2 >= 4 * kEpiScratch) KCl = kc;
        if (KCl) {
            const long long wgs = P / 256;
            int NBl = 1;
            // (512 workgroups: c5's data gradient at 16^3, batch 8, 41.8 -> 37.4 us with 2
            // blocks; the other layers' choices are unchanged, and near their best in a
            // sweep of 1, 2, 4 blocks and 1, 2, 4 splits)
            for (int nb : {4, 2})
                if (NBl == 1 && nblk % nb == 0 && wgs * (nblk / nb) >= 512) NBl = nb;
            const char *nb_env = getenv("GP5_CONV_LDS_NB");  // measurement: 1, 2 or 4
            if (nb_env && (atoi(nb_env) == 1 || atoi(nb_env) == 2 || atoi(nb_env) == 4) && nblk % atoi(nb_env) == 0)
                NBl = atoi(nb_env);
            // fewer workgroups than CUs: split the channel chunks S ways (S a power of two
            // dividing the chunk count, up to 256 workgroups), the splits' float sums added
            // in order by k_conv_split_finish; GP5_CONV_SPLIT=0 turns it off.  Measured at
            // 64^3 batch 8: c4 (8^3) data gradient 28.1 -> 15.3 us with 4 splits, forward
            // 17.0 -> 15.2 us with 2; with 256-512 workgroups the split lost (c3 data
            // gradient 19.0 -> 22.2 us, c5 forward 41.2 -> 51.6 us)
            const char *sp_env = getenv("GP5_CONV_SPLIT");
            const int chunks = Cin / KCl;
            int S = 1;
            if (!(sp_env && sp_env[0] == '0'))
                while (chunks % (2 * S) == 0 && wgs * (nblk / NBl) * S < 256) S *= 2;
            if (sp_env && atoi(sp_env) > 1 && chunks % atoi(sp_env) == 0) S = atoi(sp_env);  // measurement
            float *ws = nullptr;
            if (S > 1) {
                const int rc2 = conv_split_scratch((size_t)S * P * Cout, &ws);
                if (rc2 != GP5_OK) return rc2;
            }
            const dim3 grid((unsigned)wgs, (unsigned)(nblk / NBl), (unsigned)S);
            const size_t shm = (size_t)region * KCl * 2;
            const int csplit = S > 1 ? chunks / S : 0;
            float *yf_k = S > 1 ? ws : Yf;
#define GP5_CONV_LDS_CASE(C, K)                                                                                   \
    if (KCl == C && NBl == K)                                                                                     \
        hipLaunchKernelGGL((k_conv_lds<C, K, kConvPD>), grid, block, shm, nn_stream(), x, w, bias, a, sk, y, yf_k, (int)P, D,    \
                           Cin, Cout, mirror, act, Rb, PZ, up, csplit);
            GP5_CONV_LDS_CASE(16, 1) GP5_CONV_LDS_CASE(16, 2) GP5_CONV_LDS_CASE(16, 4)
            GP5_CONV_LDS_CASE(32, 1) GP5_CONV_LDS_CASE(32, 2) GP5_CONV_LDS_CASE(32, 4)
#undef GP5_CONV_LDS_CASE
            NNCHK(hipGetLastError(), "conv3d launch");
            if (S > 1) {
                hipLaunchKernelGGL(k_conv_split_finish, dim3(grid_for(P * (Cout / 4))), dim3(256), 0, nn_stream(), ws, S, (int)P,
                                   Cout, bias, act, a, sk, y, Yf);
                NNCHK(hipGetLastError(), "conv3d split finish");
            }
            return GP5_OK;
        }
    }
    if (up) return set_error(GP5_ERR_ARG, "conv3d_up: the cube side must be a power of two in [8, 128]");
    {
        const unsigned gy = (unsigned)(nblk / NB);
        // channels per chunk: the widest of 64, 32, 16 that divides Cin
        const int KC = Cin % 64 == 0 ? 64 : (Cin % 32 == 0 ? 32 : 16);
#define GP5_CONV_CASE(C, K, T)                                                                                 \
    if (KC == C && NB == K && MT == T)                                                                         \
        hipLaunchKernelGGL((k_conv<C, K, T, kConvPD>), dim3((unsigned)((P + 64 * T - 1) / (64 * T)), gy), block, 0, nn_stream(), x, w, \
                           bias, a, sk, y, Yf, (int)P, D, Cin, Cout, mirror, act);
#define GP5_CONV_KC(C)                                                                                         \
    GP5_CONV_CASE(C, 1, 4) GP5_CONV_CASE(C, 1, 1) GP5_CONV_CASE(C, 2, 2) GP5_CONV_CASE(C, 2, 1)                \
    GP5_CONV_CASE(C, 3, 2) GP5_CONV_CASE(C, 3, 1) GP5_CONV_CASE(C, 4, 2) GP5_CONV_CASE(C, 4, 1)                \
    GP5_CONV_CASE(C, 5, 1) GP5_CONV_CASE(C, 6, 1) GP5_CONV_CASE(C, 7, 1) GP5_CONV_CASE(C, 8, 1)
        GP5_CONV_KC(16) GP5_CONV_KC(32) GP5_CONV_KC(64)
#undef GP5_CONV_KC
#undef GP5_CONV_CASE
        NNCHK(hipGetLastError(), "conv3d launch");
    }
    return GP5_OK;
}

int gp5_nn_conv3d(const void *X, const void *Wk, const float *bias, void *A, const void *skip, void *Y, float *Yf,
                  int B, int D, int Cin, int Cout, int mirror, int act)
{
    if (act != 0 && act != 1) return set_error(GP5_ERR_ARG, "conv3d: act is 0 (none) or 1 (leaky ReLU)");
    return conv3d_run(X, Wk, bias, A, skip, Y, Yf, B, D, Cin, Cout, mirror, act, 0);
}

int gp5_nn_conv3d_pool(const void *X, const void *Wk, float *Yh, int B, int D, int Cin, int Cout, int mirror)
{
    if (!Yh) return set_error(GP5_ERR_ARG, "NULL argument");
    return conv3d_run(X, Wk, nullptr, nullptr, nullptr, nullptr, Yh, B, D, Cin, Cout, mirror, 0, 0, 1);
}

int gp5_nn_conv3d_out0(const void *X, const void *Wk, const float *bias, float *Yf0, int B, int D, int Cin, int Cout)
{
    if (!Yf0) return set_error(GP5_ERR_ARG, "NULL argument");
    return conv3d_run(X, Wk, bias, nullptr, nullptr, nullptr, Yf0, B, D, Cin, Cout, 0, kActYf0, 0);
}

int gp5_nn_conv3d_up(const void *Xh, const void *Wk, const float *bias, void *A, const void *skip, void *Y, float *Yf,
                     int B, int D, int Cin, int Cout, int mirror, int act)
{
    if (act != 0 && act != 1) return set_error(GP5_ERR_ARG, "conv3d: act is 0 (none) or 1 (leaky ReLU)");
    return conv3d_run(Xh, Wk, bias, A, skip, Y, Yf, B, D, Cin, Cout, mirror, act, 1);
}

static int act_bwd_bias_run(const void *A, const void *dA, const float *G, float scale, int D, void *H, float *db,
                            long long P, int C, float *out_part = nullptr, bool dbf = false);

int gp5_nn_conv3d_act_bwd(const void *dY, const void *Wk, const void *Ab, float *Yf, void *H, float *part, int B,
                          int D, int Cin, int Cout)
{
    if (!dY || !Wk || !Ab || !Yf || !H || !part) return set_error(GP5_ERR_ARG, "NULL argument");
    if (Cout % 4 || 256 % (Cout / 4)) return set_error(GP5_ERR_ARG, "conv3d_act_bwd: Cout / 4 must divide 256");
    int fused = 0;
    const int rc = conv3d_run(dY, Wk, nullptr, const_cast<void *>(Ab), nullptr, H, Yf, B, D, Cin, Cout, 1, 0, 0, 0,
                              part, &fused);
    if (rc != GP5_OK || fused) return rc;
    return act_bwd_bias_run(Ab, Yf, nullptr, 0.f, D, H, nullptr, (long long)B * D * D * D, Cout, part);
}

int gp5_nn_conv3d_act_bwd_out0(const void *dY0, const void *Wk, const void *Ab, void *Gd, void *H, float *part, int B,
                               int D, int Cin, int Cout)
{
    if (!dY0 || !Wk || !Ab || !Gd || !H || !part) return set_error(GP5_ERR_ARG, "NULL argument");
    if (D < 16 || D > 128 || (D & (D - 1)) || Cin < 1 || Cout < 16 || Cout % 16 || B < 1)
        return set_error(GP5_ERR_ARG, "conv3d_act_bwd_out0: D a power of two in [16, 128], Cout a multiple of 16");
    if ((long long)B * D * D * D * Cout * 4 > 0x7fffffffLL)
        return set_error(GP5_ERR_ARG, "conv3d_act_bwd_out0: tensors must stay below 2 GiB");
    const int per_run = B * (D / (256 / D)), gy = Cout / 16;
    if (per_run > kBiasBlocks) return set_error(GP5_ERR_ARG, "conv3d_act_bwd_out0: batch too large for the bias rows");
    const size_t shm = dgrad_o0_lds(D);
    int nzs = 0;
    const char *rn = getenv("GP5_DO0_RUNS");  // plane runs per cube side (measurement)
    if (rn && atoi(rn) >= 1 && atoi(rn) <= D / 2) {
        nzs = atoi(rn);
    } else {  // fill the device's workgroup slots once
        static size_t occ_shm = 0;
        static int slots = 0;
        if (occ_shm != shm) {
            int dev = 0, cus = 0, per_cu = 0;
            NNCHK(hipGetDevice(&dev), "act_bwd_out0 device");
            NNCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "act_bwd_out0 CU count");
            NNCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_dgrad_o0, 256, shm),
                  "act_bwd_out0 occupancy");
            slots = std::max(1, cus * per_cu);
            occ_shm = shm;
        }
        nzs = std::max(1, std::min(D / 2, slots / (per_run * gy)));
    }
    nzs = std::max(1, std::min(nzs, kBiasBlocks / per_run));  // one bias partial row per workgroup
    hipLaunchKernelGGL(k_dgrad_o0, dim3((unsigned)(per_run * nzs), (unsigned)gy), dim3(256), shm, nn_stream(),
                       (const uint16_t *)dY0, (const uint16_t *)Wk, (const uint16_t *)Ab, (uint16_t *)H,
                       (uint16_t *)Gd, B * D * D * D, D, Cin, Cout, nzs, part);
    NNCHK(hipGetLastError(), "act_bwd_out0 launch");
    return GP5_OK;
}

// Plan of k_wgrad_zs for a cube side D (a power of two in [16, 128]): the longest
// plane run ZL (a power of two, at least 2) that keeps 512 workgroups; returns
// the number of position blocks (the chunks of the reduction), 0 when the
// kernel does not apply (or GP5_WGRAD_ZS=0 turns it off).
static long long wgrad_zs_plan(int B, int D, int Cin, int Cout, int *nzs_out)
{
    const char *env = getenv("GP5_WGRAD_ZS");
    if (env && env[0] == '0') return 0;
    if (D < 16 || D > 128 || (D & (D - 1)) != 0) return 0;
    const long long per_z = (long long)B * (D / (256 / D)), tiles = (long long)(Cin / 16) * (Cout / 16);
    // 512 workgroups (two per CU) against 1024 before, measured at batch 8: 64^3 16 x 16
    // 63.8 -> 57.7 us, 64^3 32 x 16 upsampled 103.6 -> 96.9, 32^3 16 x 32 36.1 -> 26.8,
    // 16^3 128 x 64 upsampled 44.4 -> 38.2 (fewer, longer runs halve the overlap planes)
    const char *tg = getenv("GP5_WZ_TARGET");  // measurement: workgroup target
    const long long target = tg && atoll(tg) >= 16 ? atoll(tg) : 512;
    int zl = 2;
    for (int c = D; c >= 2; c /= 2)
        if (per_z * (D / c) * tiles >= target) {
            zl = c;
            break;
        }
    int nzs = D / zl;
    const char *runs = getenv("GP5_WZ_RUNS");  // plane runs per cube side (measurement)
    if (runs && atoi(runs) >= 1 && atoi(runs) <= D / 2) nzs = atoi(runs);
    *nzs_out = nzs;
    return per_z * nzs;
}

// Plan of k_wgrad_up2 for a fine cube side D in {16, 32, 64, 128}: the longest
// coarse plane run (a power of two, at least 2) that keeps 512 workgroups;
// returns the number of position blocks (the chunks of the reduction), 0 when
// the kernel does not apply (or GP5_WGRAD_UP2=0 turns it off).
static long long wgrad_up2_plan(int B, int D, int Cin, int Cout, int *nzs_out)
{
    const char *env = getenv("GP5_WGRAD_UP2");
    if (env && env[0] == '0') return 0;
    if (D != 16 && D != 32 && D != 64 && D != 128) return 0;
    const int Dc = D / 2;
    const long long per_z = (long long)B * (Dc / wu_rows(Dc)), tiles = (long long)(Cin / 16) * (Cout / 16);
    int zl = 2;
    for (int c = Dc; c >= 2; c /= 2)
        if (per_z * (Dc / c) * tiles >= 512) {
            zl = c;
            break;
        }
    int nzs = Dc / zl;
    const char *runs = getenv("GP5_WU2_RUNS");  // coarse plane runs per cube side (measurement)
    if (runs && atoi(runs) >= 1 && atoi(runs) <= Dc / 2) nzs = atoi(runs);
    *nzs_out = nzs;
    return per_z * nzs;
}

size_t gp5_nn_wgrad_ws(int B, int D, int Cin, int Cout)
{
    const long long P = (long long)B * D * D * D;
    const long long chunk = wgrad_chunk(P, Cin, Cout);
    const long long chunks = (P + chunk - 1) / chunk;
    int nzs = 0;
    const long long zchunks = wgrad_zs_plan(B, D, Cin, Cout, &nzs);
    const long long uchunks = wgrad_up2_plan(B, D, Cin, Cout, &nzs);
    // channel-major copies of X and dY (blocked by 8 positions when D % 8 == 0), then the partial sums;
    // or the partial sums of the position blocks of k_wgrad_zs or k_wgrad_up2
    const size_t n = (size_t)27 * Cin * Cout * sizeof(float);
    return std::max((size_t)P * (Cin + Cout) * sizeof(uint16_t) + (size_t)chunks * n,
                    (size_t)std::max(zchunks, uchunks) * n);
}

// dW = the sum of `chunks` partials of n values each (fixed order)
static int wgrad_reduce(const float *ws, float *dW, long long n, int chunks)
{
    if (n >= 64 * 256)
        hipLaunchKernelGGL(k_wgrad_reduce1<64>, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, nn_stream(), ws, dW, n, chunks);
    else
        hipLaunchKernelGGL(k_wgrad_reduce1<16>, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, nn_stream(), ws, dW, n, chunks);
    NNCHK(hipGetLastError(), "wgrad reduce launch");
    return GP5_OK;
}

// up = 1: X has side D/2 and the gradient is that of a convolution of its
// nearest upsampling (cube sides that are a multiple of 8)
static int wgrad_run(const void *X, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout, int up)
{
    if (!X || !dY || !dW || !ws) return set_error(GP5_ERR_ARG, "NULL argument");
    if (up && D % 8) return set_error(GP5_ERR_ARG, "conv3d_wgrad_up: the cube side must be a multiple of 8");
    if ((long long)B * D * D * D * std::max(Cin, Cout) * 2 > 0x7fffffffLL)
        return set_error(GP5_ERR_ARG, "conv3d_wgrad: tensors must stay below 2 GiB (B*D^3*C elements)");
    int rc = check_channels(Cin, "Cin");
    if (rc == GP5_OK) rc = check_channels(Cout, "Cout");
    if (rc != GP5_OK) return rc;
    const long long P = (long long)B * D * D * D;
    const long long n = 27LL * Cin * Cout;
    int nzs = 0;
    const long long ublocks = up ? wgrad_up2_plan(B, D, Cin, Cout, &nzs) : 0;
    if (ublocks > 0) {  // by output parity
        const size_t shm = wgrad_up2_lds(D);
        static size_t shm_up2 = 65536;  // the default limit of dynamic LDS
        if (shm > shm_up2) {
            NNCHK(hipFuncSetAttribute((const void *)k_wgrad_up2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm),
                  "wgrad up2 LDS size");
            shm_up2 = shm;
        }
        hipLaunchKernelGGL(k_wgrad_up2, dim3((unsigned)ublocks, (unsigned)((Cin / 16) * (Cout / 16))), dim3(kWuThreads),
                           shm, nn_stream(), (const uint16_t *)X, (const uint16_t *)dY, ws, (int)P, D, Cin, Cout, nzs);
        NNCHK(hipGetLastError(), "wgrad up2 launch");
        return wgrad_reduce(ws, dW, n, (int)ublocks);
    }
    const long long zblocks = wgrad_zs_plan(B, D, Cin, Cout, &nzs);
    if (zblocks > 0) {
        const size_t shm = wgrad_zs_lds(D);
        static size_t shm_set = 65536;  // the default limit of dynamic LDS
        if (shm > shm_set) {
            NNCHK(hipFuncSetAttribute((const void *)k_wgrad_zs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm),
                  "wgrad LDS size");
            shm_set = shm;
        }
        hipLaunchKernelGGL(k_wgrad_zs, dim3((unsigned)zblocks, (unsigned)((Cin / 16) * (Cout / 16))), dim3(kWzThreads),
                           shm, nn_stream(), (const uint16_t *)X, (const uint16_t *)dY, ws, (int)P, D, Cin, Cout, nzs, up);
        NNCHK(hipGetLastError(), "wgrad launch");
        return wgrad_reduce(ws, dW, n, (int)zblocks);
    }
    const long long chunk = wgrad_chunk(P, Cin, Cout);
    const long long chunks = (P + chunk - 1) / chunk;
    uint16_t *XT = reinterpret_cast<uint16_t *>(ws);
    uint16_t *dYT = XT + (size_t)P * Cin;
    float *part0 = reinterpret_cast<float *>(dYT + (size_t)P * Cout);
    const dim3 grid((unsigned)((Cin / 16) * (Cout / 16)), (unsigned)chunks);
    if (D % 4 == 0) {
        if (D % 8 == 0) {  // P = B D^3 is a multiple of 512
            hipLaunchKernelGGL(k_transpose_b8, dim3((unsigned)(P / 128), (unsigned)(Cin / 16)), dim3(256), 0, nn_stream(),
                               (const uint16_t *)X, XT, Cin, D, up);
            NNCHK(hipGetLastError(), "wgrad transpose");
            hipLaunchKernelGGL(k_transpose_b8, dim3((unsigned)(P / 128), (unsigned)(Cout / 16)), dim3(256), 0, nn_stream(),
                               (const uint16_t *)dY, dYT, Cout, D, 0);
        } else {
            const unsigned tb = (unsigned)((P + 63) / 64);
            hipLaunchKernelGGL(k_transpose_pc, dim3(tb), dim3(256), 0, nn_stream(), (const uint16_t *)X, XT, P, Cin);
            NNCHK(hipGetLastError(), "wgrad transpose");
            hipLaunchKernelGGL(k_transpose_pc, dim3(tb), dim3(256), 0, nn_stream(), (const uint16_t *)dY, dYT, P, Cout);
        }
        NNCHK(hipGetLastError(), "wgrad transpose");
        if (D % 8 == 0)  // 32-bit offsets: P * C * 2 < 2^31 (checked above)
            hipLaunchKernelGGL(k_wgrad8, dim3(grid.x, grid.y, 3), dim3(64), 0, nn_stream(), XT, dYT, part0, (int)P, D, Cin, Cout,
                               (int)chunk);
        else
            hipLaunchKernelGGL(k_wgrad, grid, dim3(64), 0, nn_stream(), XT, dYT, part0, P, D, Cin, Cout, chunk);
    } else {
        hipLaunchKernelGGL(k_wgrad_gather, grid, dim3(64), 0, nn_stream(), (const uint16_t *)X, (const uint16_t *)dY, part0, P, D,
                           Cin, Cout, chunk);
    }
    NNCHK(hipGetLastError(), "wgrad launch");
    return wgrad_reduce(part0, dW, n, (int)chunks);
}

int gp5_nn_conv3d_wgrad(const void *X, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout)
{
    return wgrad_run(X, dY, dW, ws, B, D, Cin, Cout, 0);
}

// k_wgrad_o0 and its reduction; dY rows of ldy channels (Cout, or 1: channel 0 alone).
// Without the kernel's shape the full weight gradient runs (ldy = Cout) or the call
// fails (ldy = 1).
static int wgrad_out0_run(const void *X, const void *dY, int ldy, float *dW, float *ws, int B, int D, int Cin, int Cout)
{
    if (!X || !dY || !dW || !ws) return set_error(GP5_ERR_ARG, "NULL argument");
    int nzs = 0;
    const long long zblocks = (Cin == 16 && Cout % 16 == 0 && Cout >= 16 && (D == 32 || D == 64 || D == 128))
                                  ? wgrad_zs_plan(B, D, Cin, 16, &nzs) : 0;
    const char *env = getenv("GP5_WGRAD_OUT0");
    if (zblocks <= 0 || (long long)B * D * D * D * Cout * 2 > 0x7fffffffLL || (ldy > 1 && env && env[0] == '0')) {
        if (ldy == 1)
            return set_error(GP5_ERR_ARG, "conv3d_wgrad_out0c: Cin = 16, Cout a multiple of 16, D in {32, 64, 128}");
        return wgrad_run(X, dY, dW, ws, B, D, Cin, Cout, 0);  // every output channel (the others give 0)
    }
    const int Rb = 256 / D;
    const size_t shm = (size_t)16 * 256 * 2 + (size_t)3 * (Rb + 2) * (D + 2 * kWo0Pad) * 2;
    // D / 8 plane runs (measured at 64^3, batch 8, with the reduction's four loads in
    // flight: 4 runs 35.1 us, 8 33.4; with one load at a time 4 runs 41.3 us, 8 53.4, 16
    // 73.1; the full weight gradient 58.2); the partials, 27 x 16 floats per block, fit
    // the plan's workspace of 27 x 16 x 16 per block
    const char *rn = getenv("GP5_WO0_RUNS");  // measurement
    const int runs = std::max(1, std::min(D / 2, rn && atoi(rn) > 0 ? atoi(rn) : D / 8));
    const long long blocks = (long long)B * (D / Rb) * runs;
    if (blocks * 27 * 16 > zblocks * 27 * 16 * 16) return set_error(GP5_ERR_ARG, "wgrad out0: workspace");
    hipLaunchKernelGGL(k_wgrad_o0, dim3((unsigned)blocks), dim3(256), std::max<size_t>(shm, 8192), nn_stream(),
                       (const uint16_t *)X, (const uint16_t *)dY, ws, B * D * D * D, D, ldy, runs);
    NNCHK(hipGetLastError(), "wgrad out0 launch");
    hipLaunchKernelGGL(k_wgrad_o0_reduce, dim3(27), dim3(256), 0, nn_stream(), ws, dW, (int)blocks, Cout);
    NNCHK(hipGetLastError(), "wgrad out0 reduce");
    return GP5_OK;
}

int gp5_nn_conv3d_wgrad_out0(const void *X, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout)
{
    return wgrad_out0_run(X, dY, Cout, dW, ws, B, D, Cin, Cout);
}

int gp5_nn_conv3d_wgrad_out0c(const void *X, const void *dY0, float *dW, float *ws, int B, int D, int Cin, int Cout)
{
    return wgrad_out0_run(X, dY0, 1, dW, ws, B, D, Cin, Cout);
}

int gp5_nn_conv3d_wgrad_up(const void *Xh, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout)
{
    return wgrad_run(Xh, dY, dW, ws, B, D, Cin, Cout, 1);
}

// Block partials of the bias gradients in a process-wide scratch that only grows
// (the kernels of successive calls run in order on one stream, so they can share
// it).
static int bias_scratch(size_t need, float **out)
{
    static std::mutex mu;
    static float *part = nullptr;
    static size_t cap = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (cap < need) {
        if (part) NNCHK(hipFree(part), "bias grad scratch free");
        part = nullptr;
        cap = 0;
        NNCHK(hipMalloc(&part, need * sizeof(float)), "bias grad scratch");
        cap = need;
    }
    *out = part;
    return GP5_OK;
}

int gp5_nn_bias_grad(const float *dY, float *db, long long P, int C)
{
    if (!dY || !db || C < 1 || C > 256 || 256 % C) return set_error(GP5_ERR_ARG, "bias grad: C must divide 256");
    float *part = nullptr;
    const int rc = bias_scratch((size_t)kBiasBlocks * C, &part);
    if (rc != GP5_OK) return rc;
    hipLaunchKernelGGL(k_bias_partial, dim3(kBiasBlocks), dim3(256), 0, nn_stream(), dY, part, P, C);
    NNCHK(hipGetLastError(), "bias grad launch");
    hipLaunchKernelGGL(k_bias_final, dim3(C), dim3(256), 0, nn_stream(), part, db, kBiasBlocks, C);
    NNCHK(hipGetLastError(), "bias grad launch");
    return GP5_OK;
}

int gp5_nn_pool2_bf16(const void *X, void *Y, int B, int D, int C, float scale)
{
    if (D % 2) return set_error(GP5_ERR_ARG, "pool of an odd side");
    if (C % 4) return set_error(GP5_ERR_ARG, "pool2: C must be a multiple of 4");
    const long long n = (long long)B * (D / 2) * (D / 2) * (D / 2) * (C / 4);
    if ((long long)B * D * D * D * C < 0x7fffffffLL)
        hipLaunchKernelGGL(k_pool2_bf16x4<int>, dim3(grid_for(n)), dim3(256), 0, nn_stream(), (const uint16_t *)X,
                           (uint16_t *)Y, B, D, C, scale);
    else
        hipLaunchKernelGGL(k_pool2_bf16x4<long long>, dim3(grid_for(n)), dim3(256), 0, nn_stream(), (const uint16_t *)X,
                           (uint16_t *)Y, B, D, C, scale);
    NNCHK(hipGetLastError(), "pool2");
    return GP5_OK;
}

int gp5_nn_up2_bf16(const void *X, void *Y, int B, int D, int C, float scale)
{
    if (C % 4) return set_error(GP5_ERR_ARG, "up2: C must be a multiple of 4");
    const long long n = (long long)B * 8 * D * D * D * (C / 4);
    hipLaunchKernelGGL(k_up2_bf16x4, dim3(grid_for(n)), dim3(256), 0, nn_stream(), (const uint16_t *)X, (uint16_t *)Y, B, D,
                       C, scale);
    NNCHK(hipGetLastError(), "up2");
    return GP5_OK;
}

int gp5_nn_pool2_f32(const float *X, float *Y, int B, int D, int C, float scale)
{
    if (D % 2) return set_error(GP5_ERR_ARG, "pool of an odd side");
    const long long n = (long long)B * (D / 2) * (D / 2) * (D / 2) * C;
    const bool a16 = ((uintptr_t)X | (uintptr_t)Y) % 16 == 0;  // the four-channel form's 16-byte accesses
    if (C % 4 == 0 && a16 && (long long)B * D * D * D * C < 0x7fffffffLL)
        hipLaunchKernelGGL(k_pool2_f32x4, dim3(grid_for(n / 4)), dim3(256), 0, nn_stream(), X, Y, B, D, C, scale);
    else
        hipLaunchKernelGGL(k_pool2<float>, dim3(grid_for(n)), dim3(256), 0, nn_stream(), X, Y, B, D, C, scale);
    NNCHK(hipGetLastError(), "pool2");
    return GP5_OK;
}

int gp5_nn_up2_f32(const float *X, float *Y, int B, int D, int C, float scale)
{
    const long long n = (long long)B * 8 * D * D * D * C;
    hipLaunchKernelGGL(k_up2<float>, dim3(grid_for(n)), dim3(256), 0, nn_stream(), X, Y, B, D, C, scale);
    NNCHK(hipGetLastError(), "up2");
    return GP5_OK;
}

int gp5_nn_act_bwd(const void *A, const float *dA, float *dPre, long long n)
{
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(n)), dim3(256), 0, nn_stream(), (const uint16_t *)A, dA, dPre, n);
    NNCHK(hipGetLastError(), "act_bwd");
    return GP5_OK;
}

int gp5_nn_act_bwd_cast(const void *A, const float *dA, float *dPre, void *H, long long n)
{
    if (!A || !dA || !dPre || !H) return set_error(GP5_ERR_ARG, "NULL argument");
    if (n % 4) return set_error(GP5_ERR_ARG, "act_bwd_cast: n must be a multiple of 4");
    hipLaunchKernelGGL(k_act_bwd_cast4, dim3(grid_for(n / 4)), dim3(256), 0, nn_stream(), (const uint16_t *)A, dA, dPre,
                       (uint16_t *)H, n / 4);
    NNCHK(hipGetLastError(), "act_bwd_cast");
    return GP5_OK;
}

// db set: the bias gradient is finished here (partials in the shared scratch);
// else the partials go to the caller's `part` (kBiasBlocks * C floats)
static int act_bwd_bias_run(const void *A, const void *dA, const float *G, float scale, int D, void *H, float *db,
                            long long P, int C, float *out_part, bool dbf)
{
    if (!A || !dA || !H || (!db && !out_part)) return set_error(GP5_ERR_ARG, "NULL argument");
    if (dbf && !G) return set_error(GP5_ERR_ARG, "act_bwd: a bf16 dA needs G");
    if (C < 4 || C % 4 || 256 % (C / 4)) return set_error(GP5_ERR_ARG, "act_bwd_cast_bias: C / 4 must divide 256");
    float *part = out_part;
    if (!part) {
        const int rc = bias_scratch((size_t)kBiasBlocks * C, &part);
        if (rc != GP5_OK) return rc;
    }
    if (G && dbf)
        hipLaunchKernelGGL((k_act_bwd_cast_bias<true, true>), dim3(kBiasBlocks), dim3(256), 0, nn_stream(),
                           (const uint16_t *)A, dA, (uint16_t *)H, part, P, C, G, scale, D);
    else if (G)
        hipLaunchKernelGGL(k_act_bwd_cast_bias<true>, dim3(kBiasBlocks), dim3(256), 0, nn_stream(), (const uint16_t *)A,
                           dA, (uint16_t *)H, part, P, C, G, scale, D);
    else
        hipLaunchKernelGGL(k_act_bwd_cast_bias<false>, dim3(kBiasBlocks), dim3(256), 0, nn_stream(), (const uint16_t *)A,
                           dA, (uint16_t *)H, part, P, C, nullptr, scale, D);
    NNCHK(hipGetLastError(), "act_bwd_cast_bias");
    if (!db) return GP5_OK;
    hipLaunchKernelGGL(k_bias_final, dim3(C), dim3(256), 0, nn_stream(), part, db, kBiasBlocks, C);
    NNCHK(hipGetLastError(), "act_bwd_cast_bias");
    return GP5_OK;
}

int gp5_nn_act_bwd_cast_bias(const void *A, const float *dA, void *H, float *db, long long P, int C)
{
    return act_bwd_bias_run(A, dA, nullptr, 0.f, 0, H, db, P, C);
}

int gp5_nn_up2_add_act_bwd_cast_bias(const void *A, const float *dA, const float *G, float scale, void *H, float *db,
                                     int B, int D, int C)
{
    if (!G) return set_error(GP5_ERR_ARG, "NULL argument");
    if (B < 1 || D < 2 || D % 2) return set_error(GP5_ERR_ARG, "up2_add_act_bwd_cast_bias: D must be even");
    if ((long long)B * D * D * D * C >= 0x7fffffffLL)
        return set_error(GP5_ERR_ARG, "up2_add_act_bwd_cast_bias: tensors must stay below 2^31 elements");
    return act_bwd_bias_run(A, dA, G, scale, D, H, db, (long long)B * D * D * D, C);
}

int gp5_nn_act_bwd_part_bf16(const void *A, const void *dA, const float *G, float scale, void *H, float *part, int B,
                             int D, int C)
{
    if (!part || !G) return set_error(GP5_ERR_ARG, "NULL argument");
    if (B < 1 || D < 2 || D % 2) return set_error(GP5_ERR_ARG, "act_bwd_part_bf16: D must be even");
    if ((long long)B * D * D * D * C >= 0x7fffffffLL)
        return set_error(GP5_ERR_ARG, "act_bwd_part_bf16: tensors must stay below 2^31 elements");
    return act_bwd_bias_run(A, dA, G, scale, D, H, nullptr, (long long)B * D * D * D, C, part, true);
}

size_t gp5_nn_bias_part_floats(int C)
{
    return (size_t)kBiasBlocks * (size_t)(C > 0 ? C : 0);
}

int gp5_nn_act_bwd_part(const void *A, const float *dA, const float *G, float scale, void *H, float *part, int B, int D,
                        int C)
{
    if (!part) return set_error(GP5_ERR_ARG, "NULL argument");
    if (B < 1 || D < 1 || (G && D % 2)) return set_error(GP5_ERR_ARG, "act_bwd_part: D must be even when G is set");
    if ((long long)B * D * D * D * C >= 0x7fffffffLL)
        return set_error(GP5_ERR_ARG, "act_bwd_part: tensors must stay below 2^31 elements");
    return act_bwd_bias_run(A, dA, G, scale, D, H, nullptr, (long long)B * D * D * D, C, part);
}

int gp5_nn_bias_final_multi(const gp5_nn_bias_seg *segs, int nseg)
{
    if (!segs || nseg < 1 || nseg > GP5_NN_BIAS_MAX_SEGS)
        return set_error(GP5_ERR_ARG, "bias_final_multi: 1 .. GP5_NN_BIAS_MAX_SEGS segments");
    BiasMulti a{};
    a.n = nseg;
    int blocks = 0;
    for (int k = 0; k < nseg; ++k) {
        if (!segs[k].part || !segs[k].db || segs[k].C < 1)
            return set_error(GP5_ERR_ARG, "bias_final_multi: incomplete segment");
        a.s[k] = segs[k];
        blocks += segs[k].C;
    }
    hipLaunchKernelGGL(k_bias_final_multi, dim3(blocks), dim3(256), 0, nn_stream(), a);
    NNCHK(hipGetLastError(), "bias_final_multi");
    return GP5_OK;
}

int gp5_nn_up2_add_f32(const float *X, float *Y, int B, int D, int C, float scale)
{
    if (!X || !Y) return set_error(GP5_ERR_ARG, "NULL argument");
    if (C % 4) return set_error(GP5_ERR_ARG, "up2_add: C must be a multiple of 4");
    const long long n = (long long)B * 8 * D * D * D * (C / 4);
    hipLaunchKernelGGL(k_up2_add4, dim3(grid_for(n)), dim3(256), 0, nn_stream(), X, Y, B, D, C, scale);
    NNCHK(hipGetLastError(), "up2_add");
    return GP5_OK;
}

int gp5_nn_add_f32(float *dst, const float *src, long long n)
{
    hipLaunchKernelGGL(k_add, dim3(grid_for(n)), dim3(256), 0, nn_stream(), dst, src, n);
    NNCHK(hipGetLastError(), "add");
    return GP5_OK;
}

int gp5_nn_droplet_features(void *X, float *base, int D, double ub, float s, const float a[3], const float c[5])
{
    if (!X || !base || !a || !c || D < 1) return set_error(GP5_ERR_ARG, "gp5_nn_droplet_features: bad argument");
    FeatArgs f;
    f.D = D;
    f.ub = ub;
    f.ub_f = (float)ub;
    f.s = s;
    f.a0 = a[0];
    f.a1 = a[1];
    f.a2 = a[2];
    for (int q = 0; q < 5; ++q) f.c[q] = c[q];
    const long long P = (long long)D * D * D;
    hipLaunchKernelGGL(k_droplet_features, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, nn_stream(), (uint16_t *)X, base, f);
    NNCHK(hipGetLastError(), "droplet features");
    return GP5_OK;
}

int gp5_nn_predict_out(const float *Yf, int C, const float *base, float *y, long long P)
{
    if (!Yf || !base || !y || C < 1 || P < 0) return set_error(GP5_ERR_ARG, "gp5_nn_predict_out: bad argument");
    hipLaunchKernelGGL(k_predict_out, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, nn_stream(), Yf, C, base, y, P);
    NNCHK(hipGetLastError(), "predict out");
    return GP5_OK;
}

int gp5_nn_cast_bf16(const float *src, void *dst, long long n)
{
    hipLaunchKernelGGL(k_cast, dim3(grid_for(n)), dim3(256), 0, nn_stream(), src, (uint16_t *)dst, n);
    NNCHK(hipGetLastError(), "cast");
    return GP5_OK;
}

// Scratch of the loss: block sums and the loss, in a process-wide buffer that
// only grows (the kernels of successive calls run in order on one stream).
static int mse_scratch(size_t blocks, double **out)
{
    static std::mutex mu;
    static double *buf = nullptr;
    static size_t cap = 0;
    std::lock_guard<std::mutex> lock(mu);
    const size_t need = 2 * blocks + 1;
    if (cap < need) {
        if (buf) NNCHK(hipFree(buf), "mse scratch");
        buf = nullptr;
        cap = 0;
        NNCHK(hipMalloc(&buf, need * sizeof(double)), "mse scratch");
        cap = need;
    }
    *out = buf;
    return GP5_OK;
}

// db (optional, device, C floats): the bias gradient of the output layer, sum
// over positions of the gradient rows: channel 0 from the double block sums, 0
// elsewhere.  The sums finish on the device; the one host copy is the loss
// (none when acc, a device double the loss is added to, is given instead).
// hc: channels of the rows of H (C, or 1 for channel 0 alone)
static int mse_run(const float *Yf, int ldy, int C, const float *base, const float *target, float *dYf, uint16_t *H,
                   float *db, long long P, double *loss, double *acc = nullptr, int hc = 0)
{
    if (!Yf || !base || !target || (!loss && !acc)) return set_error(GP5_ERR_ARG, "NULL argument");
    if (hc != 1) hc = C;
    if (H && hc == C && C % 8) return set_error(GP5_ERR_ARG, "mse_bf16: C must be a multiple of 8");
    if (C < 1 || P < 1) return set_error(GP5_ERR_ARG, "mse: bad size");
    const unsigned blocks = (unsigned)std::min<long long>((P + 255) / 256, 1024);
    double *part = nullptr;
    const int rc = mse_scratch(blocks, &part);
    if (rc != GP5_OK) return rc;
    hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, nn_stream(), Yf, ldy, C, base, target, dYf, H, hc, P, part);
    NNCHK(hipGetLastError(), "mse");
    hipLaunchKernelGGL(k_mse_final, dim3(1), dim3(256), 0, nn_stream(), part, (int)blocks, P, part + 2 * blocks, db, C, acc);
    NNCHK(hipGetLastError(), "mse");
    if (loss) {
        NNCHK(hipMemcpyAsync(loss, part + 2 * blocks, sizeof(double), hipMemcpyDeviceToHost, nn_stream()), "mse loss");
        NNCHK(hipStreamSynchronize(nn_stream()), "mse loss");
    }
    return GP5_OK;
}

int gp5_nn_mse(const float *Yf, int C, const float *base, const float *target, float *dYf, long long P, double *loss)
{
    if (!dYf) return set_error(GP5_ERR_ARG, "NULL argument");
    return mse_run(Yf, C, C, base, target, dYf, nullptr, nullptr, P, loss);
}

int gp5_nn_mse_bf16(const float *Yf, int C, const float *base, const float *target, float *dYf, void *H, float *db,
                    long long P, double *loss)
{
    if (!H) return set_error(GP5_ERR_ARG, "NULL argument");
    return mse_run(Yf, C, C, base, target, dYf, (uint16_t *)H, db, P, loss);
}

int gp5_nn_mse_out0(const float *y0, const float *base, const float *target, void *H, int C, float *db, long long P,
                    double *loss)
{
    if (!H) return set_error(GP5_ERR_ARG, "NULL argument");
    return mse_run(y0, 1, C, base, target, nullptr, (uint16_t *)H, db, P, loss);
}

int gp5_nn_mse_out0_acc(const float *y0, const float *base, const float *target, void *H, int C, float *db,
                        long long P, double *loss_acc)
{
    if (!H || !loss_acc) return set_error(GP5_ERR_ARG, "NULL argument");
    return mse_run(y0, 1, C, base, target, nullptr, (uint16_t *)H, db, P, nullptr, loss_acc);
}

int gp5_nn_mse_out0c(const float *y0, const float *base, const float *target, void *H0, int C, float *db, long long P,
                     double *loss, double *loss_acc)
{
    if (!H0) return set_error(GP5_ERR_ARG, "NULL argument");
    return mse_run(y0, 1, C, base, target, nullptr, (uint16_t *)H0, db, P, loss, loss_acc, 1);
}

int gp5_nn_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2, float eps, int t)
{
    if (t < 1) return set_error(GP5_ERR_ARG, "adam step count starts at 1");
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, nn_stream(), p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    NNCHK(hipGetLastError(), "adam");
    return GP5_OK;
}

int gp5_nn_adam_multi(const gp5_nn_adam_seg *segs, int nseg, float lr, float b1, float b2, float eps, int t)
{
    if (t < 1) return set_error(GP5_ERR_ARG, "adam step count starts at 1");
    char buf[160];
    if (!segs || nseg < 1 || nseg > GP5_NN_ADAM_MAX_SEGS) {
        snprintf(buf, sizeof buf, "adam_multi takes 1 to %d tensors, got %d", GP5_NN_ADAM_MAX_SEGS, nseg);
        return set_error(GP5_ERR_ARG, buf);
    }
    AdamMulti a;
    memset(&a, 0, sizeof a);
    long long most = 0;
    for (int k = 0; k < nseg; ++k) {
        const gp5_nn_adam_seg &sg = segs[k];
        if (!sg.p || !sg.g || !sg.m || !sg.v || sg.n < 0) {
            snprintf(buf, sizeof buf, "adam_multi: tensor %d is incomplete", k);
            return set_error(GP5_ERR_ARG, buf);
        }
        if ((sg.wf || sg.wb) && (sg.cin < 1 || sg.cout < 1 || 27LL * sg.cin * sg.cout != sg.n || sg.n > 0x7fffffffLL)) {
            snprintf(buf, sizeof buf, "adam_multi: tensor %d has %lld values, not 27 x %d x %d", k, sg.n, sg.cin, sg.cout);
            return set_error(GP5_ERR_ARG, buf);
        }
        a.s[k] = sg;
        most = std::max(most, sg.n);
    }
    a.lr = lr;
    a.b1 = b1;
    a.b2 = b2;
    a.eps = eps;
    a.c1 = 1.f - powf(b1, (float)t);
    a.c2 = 1.f - powf(b2, (float)t);
    if (most == 0) return GP5_OK;
    hipLaunchKernelGGL(k_adam_multi, dim3(std::min(grid_for(most), 1024u), (unsigned)nseg), dim3(256), 0, nn_stream(), a);
    NNCHK(hipGetLastError(), "adam_multi");
    return GP5_OK;
}

int gp5_nn_relayout(const float *W, void *Wf, void *Wb, int Cin, int Cout)
{
    const long long n = 27LL * Cin * Cout;
    hipLaunchKernelGGL(k_relayout, dim3(grid_for(n)), dim3(256), 0, nn_stream(), W, (uint16_t *)Wf, (uint16_t *)Wb, Cin, Cout);
    NNCHK(hipGetLastError(), "relayout");
    return GP5_OK;
}

}  // extern "C"
