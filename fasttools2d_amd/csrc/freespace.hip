// Free-space (non-periodic) Stokes layer potentials by direct summation.
//
// The reference evaluates these with a Fortran fast multipole method behind
// the mex functions stokesSLPfmm and stokesDLPfmm (FMM/src/StokesSLP/
// stokesSLPfmm.F:10-149, stokesSLP.f; FMM/src/StokesDLP/stokesDLPfmm.F,
// stokesDLP.f), checked against the direct sums of FMM/tests/
// stokes_SLP_FMM_test.m:29-51 and stokes_DLP_FMM_test.m:17-35.  Here the sums
// themselves run on the GPU in FP64 -- exact to rounding rather than to the
// FMM's tolerance, with O(Ns Nt) work:
//   SLP: u(x) = sum_s (-1/2 log(rho^2) f_s + (r.f_s) r/rho^2)/(4 pi),
//   DLP: u(x) = sum_s (r.n_s)(r.f_s) r/rho^4 /pi,   r = x - x_s, rho = |r|,
// skipping coincident pairs (and every source's own index when the targets
// are the sources).
//
// Mapping: a workgroup of 256 lanes takes 256 targets and one slice of the
// sources; the slice is staged through LDS 256 records at a time and every
// lane sums it for its target.  With few targets the sources are split into
// slices whose partial sums a second pass adds in slice order, so the result
// does not depend on the split's timing.
#include "se2p_internal.h"
#include "e1_eval.h"

#include <algorithm>

namespace se2p {

namespace {

constexpr int kFreeThreads = 256;

template <bool DLP>
__global__ __launch_bounds__(kFreeThreads) void k_free(const double* __restrict__ src,
                                                       int Ns, const double* __restrict__ tar,
                                                       int Nt, int self, int slice,
                                                       double* __restrict__ out)
{
    // src: Ns records {x, y, f1, f2} (SLP) or {x, y, f1, f2, n1, n2} (DLP)
    constexpr int REC = DLP ? 6 : 4;
    __shared__ double rec[kFreeThreads * REC];
    const int t = blockIdx.x * kFreeThreads + threadIdx.x;
    const bool act = t < Nt;
    const double xt = act ? tar[2 * t] : 0.0, yt = act ? tar[2 * t + 1] : 0.0;
    const int s0 = blockIdx.y * slice, s1 = min(Ns, s0 + slice);
    double a1 = 0, a2 = 0;
    for (int b = s0; b < s1; b += kFreeThreads) {
        const int nb = min(kFreeThreads, s1 - b);
        __syncthreads();
        for (int i = threadIdx.x; i < nb * REC; i += kFreeThreads) rec[i] = src[(size_t)b * REC + i];
        __syncthreads();
        if (!act) continue;
        for (int k = 0; k < nb; ++k) {
            const double* q = rec + k * REC;
            const double dx = xt - q[0], dy = yt - q[1];
            const double r2 = dx * dx + dy * dy;
            if (r2 == 0.0 || (self && b + k == t)) continue;
            const double rf = dx * q[2] + dy * q[3];
            const double ir2 = rcp64(r2);
            if (DLP) {
                const double rn = dx * q[4] + dy * q[5];
                const double c = rn * rf * (ir2 * ir2);
                a1 += c * dx;
                a2 += c * dy;
            } else {
                const double L = -0.5 * log_pos(r2);
                const double c = rf * ir2;
                a1 += L * q[2] + c * dx;
                a2 += L * q[3] + c * dy;
            }
        }
    }
    if (act) {
        const double sc = DLP ? 1.0 / kPi : 1.0 / (4 * kPi);
        double* o = out + (size_t)blockIdx.y * 2 * Nt;
        o[2 * t] = a1 * sc;
        o[2 * t + 1] = a2 * sc;
    }
}

// out[i] = sum over slices of part[slice][i], in slice order.
__global__ void k_free_reduce(const double* __restrict__ part, int n, int nslice,
                              double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int k = 0; k < nslice; ++k) s += part[(size_t)k * n + i];
    out[i] = s;
}

}  // namespace

int free_slices(int Ns, int Nt)
{
    // about 2048 workgroups in flight, slices of at least 1024 sources
    const int tb = (Nt + kFreeThreads - 1) / kFreeThreads;
    const int want = std::max(1, 2048 / std::max(1, tb));
    const int maxs = std::max(1, Ns / 1024);
    return std::min(want, maxs);
}

int launch_free(hipStream_t s, bool dlp, const double* src, int Ns, const double* tar, int Nt,
                bool self, int nslice, double* part, double* out)
{
    if (Nt <= 0) return SE2P_OK;
    const int slice = (Ns + nslice - 1) / nslice;
    dim3 grid((Nt + kFreeThreads - 1) / kFreeThreads, nslice);
    double* dst = nslice > 1 ? part : out;
    if (dlp)
        hipLaunchKernelGGL(k_free<true>, grid, dim3(kFreeThreads), 0, s, src, Ns, tar, Nt,
                           self ? 1 : 0, slice, dst);
    else
        hipLaunchKernelGGL(k_free<false>, grid, dim3(kFreeThreads), 0, s, src, Ns, tar, Nt,
                           self ? 1 : 0, slice, dst);
    SE2P_HIP_CHECK(hipGetLastError());
    if (nslice > 1) {
        const int n = 2 * Nt;
        hipLaunchKernelGGL(k_free_reduce, dim3((n + 255) / 256), dim3(256), 0, s, part, n, nslice,
                           out);
        SE2P_HIP_CHECK(hipGetLastError());
    }
    return SE2P_OK;
}

}  // namespace se2p
