// Droplet labelling on the GPU: the connected components of {|psi_c|^2 > cutoff}
// over the 26 periodic neighbours of each cell, numbered as the reference's
// depth-first scan numbers them (gp::decompose::decompose, src/cluster.cpp:17-82).
//
// The reference scans cells with k fastest and i slowest and gives the next
// colour to every unvisited cell above the cut-off together with all it
// reaches, so colour n is the droplet whose first cell in that scan order is
// the n-th such first cell.  Here every cell has a key in that scan order,
// key = (i ny + j) nz + k, and the components come from a concurrent
// union-find over keys (the atomic-min hooking of Playne and Hawick, IEEE TPDS
// 29(6), 2018): a root only ever hooks under a smaller key, so each component
// ends rooted at its smallest key.  The rank of that root among all roots, an
// exclusive prefix count over keys, is the colour minus one.  Cells at or
// below the cut-off get -1.
//
// Passes (one thread per cell each; integer work, HBM-bound):
//   init     parent[key] = key if |psi|^2 > cutoff else -1   (reads the field once, coalesced)
//   union    each foreground key unions with its 13 forward neighbours (the other 13
//            are covered from the neighbour's side)
//   flatten  parent[key] = root(key)
//   count    roots per chunk of kChunk keys
//   scan     exclusive prefix of the chunk counts (one block)
//   rank     rank[root] = prefix of roots before it
//   label    labels[p] = rank[parent[key(p)]] + 1, or -1            (coalesced writes)
#include <hip/hip_runtime.h>

#include <climits>

#include "gp5_internal.h"

namespace gp5 {
namespace {

constexpr int kCclThreads = 256;
constexpr int kPerThread = 16;
constexpr int kChunk = kCclThreads * kPerThread;  // keys per block of the count / rank passes
constexpr int kScanThreads = 1024;

__device__ __forceinline__ int ld(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ __forceinline__ int find_root(int *parent, int x)
{
    int p = ld(parent + x);
    while (p != x) {
        x = p;
        p = ld(parent + x);
    }
    return x;
}

// Every pointer goes to a smaller key of the same component, so the walks end
// and the smallest key of a component is its only root once all unions are done.
__device__ __forceinline__ void unite(int *parent, int a, int b)
{
    while (true) {
        a = find_root(parent, a);
        b = find_root(parent, b);
        if (a == b) return;
        if (a > b) {
            int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(parent + b, a);
        if (old == b) return;  // b was a root and now hangs under a
        b = old;               // b had been hooked meanwhile: join a with its new parent
    }
}

__global__ void __launch_bounds__(kCclThreads) k_init(const double2 *__restrict__ f, long long M, int nx, int ny,
                                                      int nz, double cutoff, int *__restrict__ parent)
{
    const long long p = (long long)blockIdx.x * kCclThreads + threadIdx.x;
    if (p >= M) return;
    const int i = (int)(p % nx);
    const long long r = p / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const int key = (i * ny + j) * nz + k;
    const double2 v = f[p];
    // the reference's density re^2 + im^2 (src/cluster.cpp:53), rounded after
    // each operation (no fused multiply-add) so cells at the cut-off fall on
    // the same side as in a plain IEEE evaluation
    const double d = __dadd_rn(__dmul_rn(v.x, v.x), __dmul_rn(v.y, v.y));
    parent[key] = d > cutoff ? key : -1;
}

__device__ __forceinline__ int wrap(int i, int L) { return i < 0 ? i + L : (i >= L ? i - L : i); }

__global__ void __launch_bounds__(kCclThreads) k_union(int *parent, int M, int nx, int ny, int nz)
{
    const long long q = (long long)blockIdx.x * kCclThreads + threadIdx.x;
    if (q >= M) return;
    if (ld(parent + q) < 0) return;
    const int k = (int)(q % nz);
    const int j = (int)((q / nz) % ny);
    const int i = (int)(q / ((long long)nz * ny));
    // the 13 offsets (a, b, c) that come after (0, 0, 0) in scan order
    for (int a = 0; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int c = -1; c <= 1; ++c) {
                if (a == 0 && (b < 0 || (b == 0 && c <= 0))) continue;
                const int n = (wrap(i + a, nx) * ny + wrap(j + b, ny)) * nz + wrap(k + c, nz);
                if (n != (int)q && ld(parent + n) >= 0) unite(parent, (int)q, n);
            }
}

__global__ void __launch_bounds__(kCclThreads) k_flatten(int *parent, int M)
{
    const long long q = (long long)blockIdx.x * kCclThreads + threadIdx.x;
    if (q >= M) return;
    if (ld(parent + q) < 0) return;
    const int r = find_root(parent, (int)q);
    __hip_atomic_store(parent + q, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// roots among the kPerThread keys of this thread
__device__ __forceinline__ int thread_roots(const int *parent, int M, long long base)
{
    int cnt = 0;
    for (int e = 0; e < kPerThread; ++e) {
        const long long q = base + e;
        if (q < M && parent[q] == (int)q) ++cnt;
    }
    return cnt;
}

// exclusive prefix over the block's threads (Hillis-Steele in LDS); returns
// this thread's offset, *total the block sum
__device__ __forceinline__ int block_exclusive(int v, int *total)
{
    __shared__ int s[kCclThreads];
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int d = 1; d < kCclThreads; d <<= 1) {
        const int add = t >= d ? s[t - d] : 0;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    *total = s[kCclThreads - 1];
    return s[t] - v;
}

__global__ void __launch_bounds__(kCclThreads) k_count(const int *parent, int M, int *blocksum)
{
    const long long base = (long long)blockIdx.x * kChunk + (long long)threadIdx.x * kPerThread;
    int total;
    block_exclusive(thread_roots(parent, M, base), &total);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
}

// blocksum[0..nb) -> exclusive prefix; blocksum[nb] = total
__global__ void __launch_bounds__(kScanThreads) k_scan(int *blocksum, int nb)
{
    __shared__ int s[kScanThreads];
    const int t = threadIdx.x;
    const int per = (nb + kScanThreads - 1) / kScanThreads;
    const int lo = t * per, hi = min(nb, lo + per);
    int sum = 0;
    for (int b = lo; b < hi; ++b) sum += blocksum[b];
    s[t] = sum;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const int add = t >= d ? s[t - d] : 0;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    int run = s[t] - sum;
    for (int b = lo; b < hi; ++b) {
        const int v = blocksum[b];
        blocksum[b] = run;
        run += v;
    }
    if (t == kScanThreads - 1) blocksum[nb] = s[t];
}

__global__ void __launch_bounds__(kCclThreads) k_rank(const int *parent, int M, const int *blockoff, int *rank)
{
    const long long base = (long long)blockIdx.x * kChunk + (long long)threadIdx.x * kPerThread;
    int total;
    int r = blockoff[blockIdx.x] + block_exclusive(thread_roots(parent, M, base), &total);
    for (int e = 0; e < kPerThread; ++e) {
        const long long q = base + e;
        if (q < M && parent[q] == (int)q) rank[q] = r++;
    }
}

__global__ void __launch_bounds__(kCclThreads) k_label(const int *__restrict__ parent, const int *__restrict__ rank,
                                                       long long M, int nx, int ny, int nz, int *__restrict__ labels)
{
    const long long p = (long long)blockIdx.x * kCclThreads + threadIdx.x;
    if (p >= M) return;
    const int i = (int)(p % nx);
    const long long r = p / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const int root = parent[(i * ny + j) * nz + k];
    labels[p] = root < 0 ? -1 : rank[root] + 1;
}

}  // namespace

size_t decompose_workspace_ints(long long M)
{
    const long long nb = (M + kChunk - 1) / kChunk;
    return (size_t)(2 * M + nb + 1);
}

hipError_t launch_decompose(hipStream_t st, const double2 *f, int nx, int ny, int nz, double cutoff, int *ws,
                            int *labels, int *ndroplets_dev)
{
    const long long M = (long long)nx * ny * nz;
    if (M <= 0 || M >= INT_MAX) return hipErrorInvalidValue;
    const int Mi = (int)M;
    const int nb = (int)((M + kChunk - 1) / kChunk);
    int *parent = ws, *rank = ws + M, *blocksum = ws + 2 * M;
    const unsigned g = (unsigned)((M + kCclThreads - 1) / kCclThreads);
    k_init<<<g, kCclThreads, 0, st>>>(f, M, nx, ny, nz, cutoff, parent);
    k_union<<<g, kCclThreads, 0, st>>>(parent, Mi, nx, ny, nz);
    k_flatten<<<g, kCclThreads, 0, st>>>(parent, Mi);
    k_count<<<nb, kCclThreads, 0, st>>>(parent, Mi, blocksum);
    k_scan<<<1, kScanThreads, 0, st>>>(blocksum, nb);
    k_rank<<<nb, kCclThreads, 0, st>>>(parent, Mi, blocksum, rank);
    k_label<<<g, kCclThreads, 0, st>>>(parent, rank, M, nx, ny, nz, labels);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ndroplets_dev) return hipMemcpyAsync(ndroplets_dev, blocksum + nb, sizeof(int), hipMemcpyDeviceToDevice, st);
    return hipSuccess;
}

}  // namespace gp5
