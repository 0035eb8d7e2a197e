// On-device convergence and compaction of batched ground states.
//
// A class of a batch holds its live instances in slots [0, active) of one
// multi-component context; every pass of the class reads `active` from device
// memory and lets blocks of later slots exit.  After each replay of
// check_every imaginary-time steps (one graph launch), three kernels run on the
// class's stream:
//   k_batch_check   one block: the energy-change test per live slot, the
//                   retirement bookkeeping and the plan of moves;
//   k_batch_copyout the fields of the slots retired at this check -> their
//                   instances' result fields;
//   k_batch_move    live slots beyond the new count -> the holes below it.
// The host never waits on a check: it reads the classes' states a replay late.
#include "gp5_internal.h"

#include <math.h>

namespace gp5 {

constexpr int kCheckThreads = 256;  // >= kMaxComp: one thread per slot
static_assert(kCheckThreads >= kMaxComp, "one thread per slot");

// Retirement rule (the host loop it replaces is described in gp5_capi.hip):
// slot s converged when it has been checked before and
// |E - E_previous| <= tol |E|.  Slots retire from the highest down; each
// retiree's place goes to the class's last live slot.  orig[p] tracks which
// original slot position p holds, so that every field move reads the layout
// of before the check: sources are live slots, destinations retired ones.
__global__ __launch_bounds__(kCheckThreads) void k_batch_check(BatchCheckArgs a)
{
    __shared__ int conv[kMaxComp];
    __shared__ int orig[kMaxComp];
    __shared__ double E[kMaxComp];
    __shared__ int bad;
    const int t = threadIdx.x;
    const int active = a.st->active;
    if (t == 0) bad = 0;
    __syncthreads();
    for (int s = t; s < active; s += kCheckThreads) {
        const double e = a.log_row[a.cap + s];
        const double nrm = a.results[s];
        if (!isfinite(e) || !isfinite(nrm)) {
            bad = 1;
            a.st->err_inst = a.slot_inst[s];
        }
        E[s] = e;
        conv[s] = a.checks[s] > 0 && fabs(e - a.prevE[s]) <= a.tol * fabs(e);
        a.prevE[s] = e;
        a.checks[s] += 1;
        orig[s] = s;
    }
    __syncthreads();
    if (t != 0) return;
    BatchClassDev *st = a.st;
    st->ncopy = 0;
    st->nmove = 0;
    if (active == 0) return;
    if (bad) {
        st->err = 1;
        return;
    }
    const long long steps = st->steps_done + a.check_every;
    st->steps_done = steps;
    const bool out_of_steps = a.max_steps > 0 && steps >= a.max_steps;
    int n = active, ncopy = 0;
    for (int j = active - 1; j >= 0; --j) {
        if (!(conv[j] || out_of_steps)) continue;
        const int inst = a.slot_inst[j];
        a.res_steps[inst] = conv[j] ? steps : -steps;
        a.res_energy[inst] = E[j];
        a.copy_slot[ncopy] = j;  // orig[j] == j: position j is rewritten only now
        a.copy_inst[ncopy] = inst;
        ++ncopy;
        const int last = n - 1;
        if (j != last) {
            orig[j] = orig[last];
            a.slot_inst[j] = a.slot_inst[last];
            a.prevE[j] = a.prevE[last];
            a.checks[j] = a.checks[last];
            E[j] = E[last];
            a.nnorm[j] = a.nnorm[last];
            if (a.lhy_ab) {
                a.lhy_ab[2 * j] = a.lhy_ab[2 * last];
                a.lhy_ab[2 * j + 1] = a.lhy_ab[2 * last + 1];
            }
        }
        n = last;
    }
    int nmove = 0;
    for (int p = 0; p < n; ++p)
        if (orig[p] != p) {
            a.move_src[nmove] = orig[p];
            a.move_dst[nmove] = p;
            ++nmove;
        }
    st->active = n;
    st->ncopy = ncopy;
    st->nmove = nmove;
}

constexpr int kCopyThreads = 256;
constexpr unsigned kCopyBlocks = 2048;

__global__ __launch_bounds__(kCopyThreads) void k_batch_copyout(const double2 *__restrict__ psi, long long M,
                                                                const BatchClassDev *st, const int *slot,
                                                                const int *inst, double2 *const *res_ptr)
{
    const long long total = (long long)st->ncopy * M;
    for (long long g = (long long)blockIdx.x * kCopyThreads + threadIdx.x; g < total;
         g += (long long)gridDim.x * kCopyThreads) {
        const long long q = g / M, i = g - q * M;
        res_ptr[inst[q]][i] = psi[(long long)slot[q] * M + i];
    }
}

__global__ __launch_bounds__(kCopyThreads) void k_batch_move(double2 *psi, double *V, long long M,
                                                             const BatchClassDev *st, const int *src, const int *dst)
{
    const long long total = (long long)st->nmove * M;
    for (long long g = (long long)blockIdx.x * kCopyThreads + threadIdx.x; g < total;
         g += (long long)gridDim.x * kCopyThreads) {
        const long long q = g / M, i = g - q * M;
        const long long from = (long long)src[q] * M + i, to = (long long)dst[q] * M + i;
        psi[to] = psi[from];
        if (V) V[to] = V[from];
    }
}

hipError_t launch_batch_check(hipStream_t st, const BatchCheckArgs &a)
{
    hipLaunchKernelGGL(k_batch_check, dim3(1), dim3(kCheckThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_batch_compact(hipStream_t st, double2 *psi, double *V, long long M, const BatchCheckArgs &a,
                                double2 *const *res_ptr)
{
    hipLaunchKernelGGL(k_batch_copyout, dim3(kCopyBlocks), dim3(kCopyThreads), 0, st, (const double2 *)psi, M,
                       (const BatchClassDev *)a.st, (const int *)a.copy_slot, (const int *)a.copy_inst, res_ptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_move, dim3(kCopyBlocks), dim3(kCopyThreads), 0, st, psi, V, M,
                       (const BatchClassDev *)a.st, (const int *)a.move_src, (const int *)a.move_dst);
    return hipGetLastError();
}

}  // namespace gp5
