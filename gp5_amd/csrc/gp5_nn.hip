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

#include <algorithm>
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

// Positions are linear indices p = ((b*D + z)*D + y)*D + x.  Neighbour of p at
// tap (dz, dy, dx) in {-1, 0, 1}^3, or -1 outside the cube (zero padding).
__device__ __forceinline__ long long neighbour(long long p, int D, int tap)
{
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int x = (int)(p % D), y = (int)((p / D) % D), z = (int)((p / ((long long)D * D)) % D);
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    if (xx < 0 || xx >= D || yy < 0 || yy >= D || zz < 0 || zz >= D) return -1;
    return p + ((long long)dz * D + dy) * D + dx;
}

// 8-byte load of 4 consecutive bf16 (the caller guarantees 8-byte alignment:
// channel offsets are multiples of 4)
__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) { return *reinterpret_cast<const bf16x4 *>(p); }

// Forward / data-gradient convolution.  A workgroup of 4 waves covers 64
// consecutive positions (16 per wave) and NB 16-wide blocks of output
// channels; the K loop runs over 27 taps x Cin/16 channel chunks.
template <int NB>
__global__ __launch_bounds__(256) void k_conv(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wk,
                                              const float *__restrict__ bias, uint16_t *A, const uint16_t *skip,
                                              uint16_t *Y, float *Yf, long long P, int D, int Cin, int Cout,
                                              int nb0, int mirror, int act)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const long long p = (long long)blockIdx.x * 64 + wave * 16 + r;  // A row of this lane
    f32x4 acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x4 zero4 = bf16x4{0, 0, 0, 0};
    for (int tap = 0; tap < 27; ++tap) {
        const long long q = p < P ? neighbour(p, D, tap) : -1;
        const int wt = mirror ? 26 - tap : tap;
        const uint16_t *wrow = Wk + ((size_t)wt * Cout + (size_t)nb0 * 16 + r) * Cin + 4 * g;
        for (int c0 = 0; c0 < Cin; c0 += 16) {
            const bf16x4 a = q >= 0 ? ld4(X + (size_t)q * Cin + c0 + 4 * g) : zero4;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const bf16x4 bv = ld4(wrow + (size_t)j * 16 * Cin + c0);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[j], 0, 0, 0);
            }
        }
    }
    // epilogue: lane holds D[4g + i][r] of each block: position row 4g + i, channel 16 (nb0 + j) + r
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long long po = (long long)blockIdx.x * 64 + wave * 16 + 4 * g + i;
        if (po >= P) continue;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int co = 16 * (nb0 + j) + r;
            float v = acc[j][i] + (bias ? bias[co] : 0.f);
            if (act == 1) v = v > 0.f ? v : kLeak * v;
            const size_t o = (size_t)po * Cout + co;
            if (Yf) Yf[o] = v;
            const uint16_t h = f2bf(v);
            if (A) A[o] = h;
            if (Y) Y[o] = skip ? f2bf(v + bf2f(skip[o])) : h;
        }
    }
}

// Weight gradient: one wave per (tap, 16 input channels, 16 output channels,
// position chunk); partial sums go to ws[chunk][tap][ci][co].
constexpr int kWChunk = 4096;  // positions per wave
__global__ __launch_bounds__(64) void k_wgrad(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                              float *ws, long long P, int D, int Cin, int Cout)
{
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int nci = Cin / 16, nco = Cout / 16;
    int blk = blockIdx.x;
    const int cob = blk % nco;
    blk /= nco;
    const int cib = blk % nci;
    const int tap = blk / nci;
    const long long chunk = blockIdx.y;
    const long long p0 = chunk * kWChunk, p1 = min(P, p0 + kWChunk);
    const int ci = cib * 16 + r, co = cob * 16 + r;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long pb = p0; pb < p1; pb += 16) {
        // A[m = ci][k = position]: this lane's row ci, positions pb + 4g .. 4g+3
        // B[k = position][n = co]: this lane's column co, same positions
        bf16x4 a, bv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long p = pb + 4 * g + i;
            short xa = 0, yb = 0;
            if (p < p1) {
                const long long q = neighbour(p, D, tap);
                if (q >= 0) xa = (short)X[(size_t)q * Cin + ci];
                yb = (short)dY[(size_t)p * Cout + co];
            }
            a[i] = xa;
            bv[i] = yb;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc, 0, 0, 0);
    }
    // D[m = ci row 4g+i][n = co column r]
    float *o = ws + (((size_t)chunk * 27 + tap) * Cin + cib * 16) * Cout + cob * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i) o[(size_t)(4 * g + i) * Cout] = acc[i];
}

__global__ void k_wgrad_reduce(const float *ws, float *dW, long long n, int chunks)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < chunks; ++c) s += ws[(size_t)c * n + i];
        dW[i] = s;
    }
}

// db[c] = sum_p dY[p][c]: one block per channel, fixed-order reduction
__global__ __launch_bounds__(256) void k_bias_grad(const float *dY, float *db, long long P, int C)
{
    __shared__ float s[256];
    const int c = blockIdx.x;
    float a = 0.f;
    for (long long p = threadIdx.x; p < P; p += 256) a += dY[(size_t)p * C + c];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = s[0];
}

template <typename T> __device__ __forceinline__ float ldv(const T *p);
template <> __device__ __forceinline__ float ldv<uint16_t>(const uint16_t *p) { return bf2f(*p); }
template <> __device__ __forceinline__ float ldv<float>(const float *p) { return *p; }
template <typename T> __device__ __forceinline__ void stv(T *p, float v);
template <> __device__ __forceinline__ void stv<uint16_t>(uint16_t *p, float v) { *p = f2bf(v); }
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

__global__ __launch_bounds__(256) void k_mse(const float *Yf, int C, const float *base, const float *target, float *dYf,
                                             long long P, double *partial)
{
    __shared__ double s[256];
    double a = 0.0;
    const float inv = 2.0f / (float)P;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < P; p += (long long)gridDim.x * 256) {
        const float y = base[p] + Yf[(size_t)p * C];
        const float e = y - target[p];
        a += (double)e * e;
        dYf[(size_t)p * C] = inv * e;
        for (int c = 1; c < C; ++c) dYf[(size_t)p * C + c] = 0.f;
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
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

int gp5_nn_upload(void *dst, const void *host, size_t bytes)
{
    NNCHK(hipMemcpy(dst, host, bytes, hipMemcpyHostToDevice), "gp5_nn_upload");
    return GP5_OK;
}

int gp5_nn_download(void *host, const void *src, size_t bytes)
{
    NNCHK(hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost), "gp5_nn_download");
    return GP5_OK;
}

int gp5_nn_copy(void *dst, const void *src, size_t bytes)
{
    NNCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "gp5_nn_copy");
    return GP5_OK;
}

int gp5_nn_zero(void *p, size_t bytes)
{
    if (bytes % 4) return set_error(GP5_ERR_ARG, "gp5_nn_zero: size must be a multiple of 4 bytes");
    const long long n = (long long)(bytes / 4);
    hipLaunchKernelGGL(k_fill0, dim3(grid_for(n)), dim3(256), 0, 0, (uint32_t *)p, n);
    NNCHK(hipGetLastError(), "fill");
    return GP5_OK;
}

int gp5_nn_sync(void)
{
    NNCHK(hipDeviceSynchronize(), "gp5_nn_sync");
    return GP5_OK;
}

int gp5_nn_conv3d(const void *X, const void *Wk, const float *bias, void *A, const void *skip, void *Y, float *Yf,
                  int B, int D, int Cin, int Cout, int mirror, int act)
{
    if (!X || !Wk) return set_error(GP5_ERR_ARG, "NULL input or weights");
    if (skip && !Y) return set_error(GP5_ERR_ARG, "a skip input needs the Y output");
    if (B < 1 || D < 1) return set_error(GP5_ERR_ARG, "empty tensor");
    int rc = check_channels(Cin, "Cin");
    if (rc == GP5_OK) rc = check_channels(Cout, "Cout");
    if (rc != GP5_OK) return rc;
    const long long P = (long long)B * D * D * D;
    const dim3 grid((unsigned)((P + 63) / 64)), block(256);
    // output channels in groups of up to 8 16-wide blocks per launch
    for (int nb0 = 0; nb0 < Cout / 16; nb0 += 8) {
        const int nb = std::min(8, Cout / 16 - nb0);
        const uint16_t *x = (const uint16_t *)X, *w = (const uint16_t *)Wk, *sk = (const uint16_t *)skip;
        uint16_t *a = (uint16_t *)A, *y = (uint16_t *)Y;
        switch (nb) {
#define GP5_CONV_CASE(K)                                                                                   \
    case K:                                                                                                \
        hipLaunchKernelGGL(k_conv<K>, grid, block, 0, 0, x, w, bias, a, sk, y, Yf, P, D, Cin, Cout, nb0, mirror, act); \
        break;
            GP5_CONV_CASE(1) GP5_CONV_CASE(2) GP5_CONV_CASE(3) GP5_CONV_CASE(4)
            GP5_CONV_CASE(5) GP5_CONV_CASE(6) GP5_CONV_CASE(7) GP5_CONV_CASE(8)
#undef GP5_CONV_CASE
        }
        NNCHK(hipGetLastError(), "conv3d launch");
    }
    return GP5_OK;
}

size_t gp5_nn_wgrad_ws(int B, int D, int Cin, int Cout)
{
    const long long P = (long long)B * D * D * D;
    const long long chunks = (P + kWChunk - 1) / kWChunk;
    return (size_t)chunks * 27 * Cin * Cout * sizeof(float);
}

int gp5_nn_conv3d_wgrad(const void *X, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout)
{
    if (!X || !dY || !dW || !ws) return set_error(GP5_ERR_ARG, "NULL argument");
    int rc = check_channels(Cin, "Cin");
    if (rc == GP5_OK) rc = check_channels(Cout, "Cout");
    if (rc != GP5_OK) return rc;
    const long long P = (long long)B * D * D * D;
    const long long chunks = (P + kWChunk - 1) / kWChunk;
    const dim3 grid((unsigned)(27 * (Cin / 16) * (Cout / 16)), (unsigned)chunks);
    hipLaunchKernelGGL(k_wgrad, grid, dim3(64), 0, 0, (const uint16_t *)X, (const uint16_t *)dY, ws, P, D, Cin, Cout);
    NNCHK(hipGetLastError(), "wgrad launch");
    const long long n = 27LL * Cin * Cout;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid_for(n)), dim3(256), 0, 0, ws, dW, n, (int)chunks);
    NNCHK(hipGetLastError(), "wgrad reduce launch");
    return GP5_OK;
}

int gp5_nn_bias_grad(const float *dY, float *db, long long P, int C)
{
    if (!dY || !db || C < 1) return set_error(GP5_ERR_ARG, "bad argument");
    hipLaunchKernelGGL(k_bias_grad, dim3(C), dim3(256), 0, 0, dY, db, P, C);
    NNCHK(hipGetLastError(), "bias grad launch");
    return GP5_OK;
}

int gp5_nn_pool2_bf16(const void *X, void *Y, int B, int D, int C, float scale)
{
    if (D % 2) return set_error(GP5_ERR_ARG, "pool of an odd side");
    const long long n = (long long)B * (D / 2) * (D / 2) * (D / 2) * C;
    hipLaunchKernelGGL(k_pool2<uint16_t>, dim3(grid_for(n)), dim3(256), 0, 0, (const uint16_t *)X, (uint16_t *)Y, B, D,
                       C, scale);
    NNCHK(hipGetLastError(), "pool2");
    return GP5_OK;
}

int gp5_nn_up2_bf16(const void *X, void *Y, int B, int D, int C, float scale)
{
    const long long n = (long long)B * 8 * D * D * D * C;
    hipLaunchKernelGGL(k_up2<uint16_t>, dim3(grid_for(n)), dim3(256), 0, 0, (const uint16_t *)X, (uint16_t *)Y, B, D,
                       C, scale);
    NNCHK(hipGetLastError(), "up2");
    return GP5_OK;
}

int gp5_nn_pool2_f32(const float *X, float *Y, int B, int D, int C, float scale)
{
    if (D % 2) return set_error(GP5_ERR_ARG, "pool of an odd side");
    const long long n = (long long)B * (D / 2) * (D / 2) * (D / 2) * C;
    hipLaunchKernelGGL(k_pool2<float>, dim3(grid_for(n)), dim3(256), 0, 0, X, Y, B, D, C, scale);
    NNCHK(hipGetLastError(), "pool2");
    return GP5_OK;
}

int gp5_nn_up2_f32(const float *X, float *Y, int B, int D, int C, float scale)
{
    const long long n = (long long)B * 8 * D * D * D * C;
    hipLaunchKernelGGL(k_up2<float>, dim3(grid_for(n)), dim3(256), 0, 0, X, Y, B, D, C, scale);
    NNCHK(hipGetLastError(), "up2");
    return GP5_OK;
}

int gp5_nn_act_bwd(const void *A, const float *dA, float *dPre, long long n)
{
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(n)), dim3(256), 0, 0, (const uint16_t *)A, dA, dPre, n);
    NNCHK(hipGetLastError(), "act_bwd");
    return GP5_OK;
}

int gp5_nn_add_f32(float *dst, const float *src, long long n)
{
    hipLaunchKernelGGL(k_add, dim3(grid_for(n)), dim3(256), 0, 0, dst, src, n);
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
    hipLaunchKernelGGL(k_droplet_features, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, 0, (uint16_t *)X, base, f);
    NNCHK(hipGetLastError(), "droplet features");
    return GP5_OK;
}

int gp5_nn_predict_out(const float *Yf, int C, const float *base, float *y, long long P)
{
    if (!Yf || !base || !y || C < 1 || P < 0) return set_error(GP5_ERR_ARG, "gp5_nn_predict_out: bad argument");
    hipLaunchKernelGGL(k_predict_out, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, 0, Yf, C, base, y, P);
    NNCHK(hipGetLastError(), "predict out");
    return GP5_OK;
}

int gp5_nn_cast_bf16(const float *src, void *dst, long long n)
{
    hipLaunchKernelGGL(k_cast, dim3(grid_for(n)), dim3(256), 0, 0, src, (uint16_t *)dst, n);
    NNCHK(hipGetLastError(), "cast");
    return GP5_OK;
}

int gp5_nn_mse(const float *Yf, int C, const float *base, const float *target, float *dYf, long long P, double *loss)
{
    if (!Yf || !base || !target || !dYf || !loss) return set_error(GP5_ERR_ARG, "NULL argument");
    const unsigned blocks = (unsigned)std::min<long long>((P + 255) / 256, 1024);
    double *part = nullptr;
    NNCHK(hipMalloc(&part, blocks * sizeof(double)), "mse scratch");
    hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, 0, Yf, C, base, target, dYf, P, part);
    hipError_t e = hipGetLastError();
    std::vector<double> h(blocks);
    if (e == hipSuccess) e = hipMemcpy(h.data(), part, blocks * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(part);
    if (e != hipSuccess) return hip_fail(e, "mse");
    double s = 0.0;
    for (double v : h) s += v;
    *loss = s / (double)P;
    return GP5_OK;
}

int gp5_nn_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2, float eps, int t)
{
    if (t < 1) return set_error(GP5_ERR_ARG, "adam step count starts at 1");
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, 0, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    NNCHK(hipGetLastError(), "adam");
    return GP5_OK;
}

int gp5_nn_relayout(const float *W, void *Wf, void *Wb, int Cin, int Cout)
{
    const long long n = 27LL * Cin * Cout;
    hipLaunchKernelGGL(k_relayout, dim3(grid_for(n)), dim3(256), 0, 0, W, (uint16_t *)Wf, (uint16_t *)Wb, Cin, Cout);
    NNCHK(hipGetLastError(), "relayout");
    return GP5_OK;
}

}  // extern "C"
