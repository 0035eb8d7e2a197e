// Spectral resampling between two grids of one periodic box: the trigonometric
// interpolant of the source samples evaluated at the destination's cell
// centres.  Used to carry a ground state converged on a coarse grid to a finer
// one (and back).
//
// With cell centres x_m = left + (m + 1/2) L / n on both grids, the source
// interpolant is f(x) = (1/ns) sum_k c_k F[k mod ns] e^{2 pi i k (x - x0s) / L}
// over k in [-ns/2, ns/2] (c = 1/2 at +-ns/2 for even ns: the Nyquist mode is a
// cosine) or |k| <= (ns-1)/2 for odd ns, F the unnormalised forward DFT.  At
// the destination points the exponent depends on k mod nd only, so the
// destination spectrum is
//     G[d] = (1/Ms) sum_{k = d mod nd} c_k e^{2 pi i k (x0d - x0s) / L} F[k mod ns]
// per axis (a gather table), followed by an unnormalised inverse DFT.
// Upsampling puts at most one source mode in each destination mode; equal
// sizes give the identity; downsampling folds the aliases.
#include <hip/hip_runtime.h>

#include "gp5_internal.h"

namespace gp5 {
namespace {

__global__ void __launch_bounds__(256) k_resample(const double2 *__restrict__ F, double2 *__restrict__ G,
                                                  ResampleAxis ax, ResampleAxis ay, ResampleAxis az, long long Ms,
                                                  long long Md, int nc, double scale)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= Md * nc) return;
    const int comp = (int)(e / Md);
    const long long p = e - (long long)comp * Md;
    const int a = (int)(p % ax.nd);
    const long long r = p / ax.nd;
    const int b = (int)(r % ay.nd);
    const int c = (int)(r / ay.nd);
    const double2 *Fc = F + (long long)comp * Ms;
    double sr = 0.0, si = 0.0;
    for (int iz = az.off[c]; iz < az.off[c + 1]; ++iz) {
        const double2 wz = az.w[iz];
        const long long oz = (long long)az.src[iz] * ay.ns;
        for (int iy = ay.off[b]; iy < ay.off[b + 1]; ++iy) {
            const double2 wy = ay.w[iy];
            const double2 wyz = make_double2(wy.x * wz.x - wy.y * wz.y, wy.x * wz.y + wy.y * wz.x);
            const long long oyz = (oz + ay.src[iy]) * ax.ns;
            for (int ix = ax.off[a]; ix < ax.off[a + 1]; ++ix) {
                const double2 wx = ax.w[ix];
                const double2 w = make_double2(wx.x * wyz.x - wx.y * wyz.y, wx.x * wyz.y + wx.y * wyz.x);
                const double2 v = Fc[oyz + ax.src[ix]];
                sr += w.x * v.x - w.y * v.y;
                si += w.x * v.y + w.y * v.x;
            }
        }
    }
    G[e] = make_double2(sr * scale, si * scale);
}

}  // namespace

hipError_t launch_resample(hipStream_t st, const double2 *F, double2 *G, const ResampleAxis ax[3], long long Ms,
                           long long Md, int nc)
{
    const long long total = Md * nc;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    k_resample<<<blocks, 256, 0, st>>>(F, G, ax[0], ax[1], ax[2], Ms, Md, nc, 1.0 / (double)Ms);
    return hipGetLastError();
}

}  // namespace gp5
