// Real-space cell-list sums for the Stokeslet (SLP) and stresslet (DLP).
//
// Follows mex_stokes_slp_real.cpp:58-357 and mex_stokes_dlp_real.cpp:15-211
// (PeriodicEwald/src/mex/...): cells of size Lx/nside_x x Ly/nside_y; for a
// target, every source of its own cell with no cutoff (SLP: coincident points
// take the self term -(1 + gamma/2 + ln xi) f; DLP: coincident points are
// skipped), then the 8 neighbour cells in the fixed stencil order with the
// cutoff r^2 < Lx Ly/(nside_x nside_y), each neighbour's partial sum added
// after it is complete.  Output scaled by 1/(4 pi) and written in input order.
//
// Mapping: one workgroup per target cell, one lane per target, the sources of
// each stencil cell staged through LDS in chunks and read by broadcast.  Each
// target accumulates in the reference's order.
#include "se2p_internal.h"
#include "e1_eval.h"

namespace se2p {

namespace {

struct alignas(64) DlpRec {
    double x, y, f1, f2, n1, n2, fn, pad;
};

__constant__ int c_nbdx[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
__constant__ int c_nbdy[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

// 4 pi times the SLP pair term for r = target - source:
//   E1(xi^2 r^2)/2 f + e^{-xi^2 r^2} ((r.f) r/r^2 - f)
// in the cancellation-free form u1 += E/2 f1 + (e/r^2) dy (dx f2 - dy f1),
// u2 += E/2 f2 - (e/r^2) dx (dx f2 - dy f1).
template <int TIER>
__device__ __forceinline__ void slp_pair(double dx, double dy, double r2, double xi2, double f1,
                                         double f2, double& a1, double& a2)
{
    double x = xi2 * r2;
    double e = exp(-x);
    double t = 1.0 / x;
    double E = e1_eval<TIER>(x, e, t);
    double c = e / r2;
    double cr = dx * f2 - dy * f1;
    a1 += 0.5 * E * f1 + c * dy * cr;
    a2 += 0.5 * E * f2 - c * dx * cr;
}

// 4 pi times the DLP pair term for r = target - source:
//   e^{-xi^2 r^2} [ -4 (1 + xi^2 r^2)/r^4 r (r.f)(r.n) + 2 xi^2 (f (r.n) + n (r.f) + r (f.n)) ]
__device__ __forceinline__ void dlp_pair(double r1, double r2, double rr, double xi2,
                                         const DlpRec& s, double& a1, double& a2)
{
    double e = exp(-xi2 * rr);
    double rf = r1 * s.f1 + r2 * s.f2;
    double rn = r1 * s.n1 + r2 * s.n2;
    double b = -4.0 * (1.0 + xi2 * rr) / (rr * rr) * rf * rn;
    double c = 2.0 * xi2;
    a1 += e * (b * r1 + c * (s.f1 * rn + s.n1 * rf + s.fn * r1));
    a2 += e * (b * r2 + c * (s.f2 * rn + s.n2 * rf + s.fn * r2));
}

template <int BLOCK, int TIER>
__global__ __launch_bounds__(BLOCK) void k_slp_real(
    int nsx, int nsy, double Lx, double Ly, double xi2, double cut2, double self,
    const int* __restrict__ tstart, const int* __restrict__ tperm,
    const double2* __restrict__ tsorted, const int* __restrict__ sstart,
    const double4* __restrict__ ssorted, double* __restrict__ u)
{
    constexpr int CH = 256;
    __shared__ double4 buf[CH];
    const int b = blockIdx.x;
    const int bx = b % nsx, by = b / nsx;
    const int t0 = tstart[b], t1 = tstart[b + 1];
    for (int tb = t0; tb < t1; tb += BLOCK) {
        const int ti = tb + (int)threadIdx.x;
        const bool act = ti < t1;
        const double2 pt = act ? tsorted[ti] : make_double2(0.0, 0.0);
        double a1 = 0, a2 = 0;
        for (int j = 0; j < 9; ++j) {
            int qx = bx, qy = by;
            if (j > 0) {
                qx += c_nbdx[j - 1];
                qy += c_nbdy[j - 1];
            }
            const int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
            const int sb = wy * nsx + wx;
            const double ox = (Lx * (qx - wx)) / nsx, oy = (Ly * (qy - wy)) / nsy;
            const int s0 = sstart[sb], s1 = sstart[sb + 1];
            double p1 = 0, p2 = 0;
            for (int sc = s0; sc < s1; sc += CH) {
                const int cnt = min(CH, s1 - sc);
                __syncthreads();
                for (int i = threadIdx.x; i < cnt; i += BLOCK) {
                    double4 r = ssorted[sc + i];
                    r.x += ox;
                    r.y += oy;
                    buf[i] = r;
                }
                __syncthreads();
                if (!act) continue;
                if (j == 0) {
                    for (int i = 0; i < cnt; ++i) {
                        const double4 s = buf[i];
                        double dx = pt.x - s.x, dy = pt.y - s.y;
                        double r2 = dx * dx + dy * dy;
                        if (fabs(r2) < 1e-15) {
                            p1 += self * s.z;
                            p2 += self * s.w;
                        } else {
                            slp_pair<TIER>(dx, dy, r2, xi2, s.z, s.w, p1, p2);
                        }
                    }
                } else {
                    for (int i = 0; i < cnt; ++i) {
                        const double4 s = buf[i];
                        double dx = pt.x - s.x, dy = pt.y - s.y;
                        double r2 = dx * dx + dy * dy;
                        if (r2 < cut2) slp_pair<TIER>(dx, dy, r2, xi2, s.z, s.w, p1, p2);
                    }
                }
            }
            a1 += p1;
            a2 += p2;
        }
        if (act) {
            const int o = tperm[ti];
            u[2 * o] = a1 / (4 * kPi);
            u[2 * o + 1] = a2 / (4 * kPi);
        }
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_dlp_real(
    int nsx, int nsy, double Lx, double Ly, double xi2, double cut2,
    const int* __restrict__ tstart, const int* __restrict__ tperm,
    const double2* __restrict__ tsorted, const int* __restrict__ sstart,
    const DlpRec* __restrict__ ssorted, double* __restrict__ u)
{
    constexpr int CH = 256;
    __shared__ DlpRec buf[CH];
    const int b = blockIdx.x;
    const int bx = b % nsx, by = b / nsx;
    const int t0 = tstart[b], t1 = tstart[b + 1];
    for (int tb = t0; tb < t1; tb += BLOCK) {
        const int ti = tb + (int)threadIdx.x;
        const bool act = ti < t1;
        const double2 pt = act ? tsorted[ti] : make_double2(0.0, 0.0);
        double a1 = 0, a2 = 0;
        for (int j = 0; j < 9; ++j) {
            int qx = bx, qy = by;
            if (j > 0) {
                qx += c_nbdx[j - 1];
                qy += c_nbdy[j - 1];
            }
            const int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
            const int sb = wy * nsx + wx;
            const double ox = (Lx * (qx - wx)) / nsx, oy = (Ly * (qy - wy)) / nsy;
            const int s0 = sstart[sb], s1 = sstart[sb + 1];
            for (int sc = s0; sc < s1; sc += CH) {
                const int cnt = min(CH, s1 - sc);
                __syncthreads();
                for (int i = threadIdx.x; i < cnt; i += BLOCK) {
                    DlpRec r = ssorted[sc + i];
                    r.x += ox;
                    r.y += oy;
                    buf[i] = r;
                }
                __syncthreads();
                if (!act) continue;
                for (int i = 0; i < cnt; ++i) {
                    const DlpRec& s = buf[i];
                    double r1 = pt.x - s.x, r2 = pt.y - s.y;
                    double rr = r1 * r1 + r2 * r2;
                    bool take = (j == 0) ? (rr != 0) : (rr < cut2);
                    if (take) dlp_pair(r1, r2, rr, xi2, s, a1, a2);
                }
            }
        }
        if (act) {
            const int o = tperm[ti];
            u[2 * o] = a1 / (4 * kPi);
            u[2 * o + 1] = a2 / (4 * kPi);
        }
    }
}

__global__ void k_pack_slp(const int* __restrict__ perm, int n, const double* __restrict__ p,
                           const double* __restrict__ f, double4* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    out[i] = make_double4(p[2 * o], p[2 * o + 1], f[2 * o], f[2 * o + 1]);
}

__global__ void k_pack_dlp(const int* __restrict__ perm, int n, const double* __restrict__ p,
                           const double* __restrict__ f, const double* __restrict__ nrm,
                           DlpRec* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    DlpRec r;
    r.x = p[2 * o];
    r.y = p[2 * o + 1];
    r.f1 = f[2 * o];
    r.f2 = f[2 * o + 1];
    r.n1 = nrm[2 * o];
    r.n2 = nrm[2 * o + 1];
    r.fn = r.f1 * r.n1 + r.f2 * r.n2;
    r.pad = 0;
    out[i] = r;
}

__global__ void k_pack_tar(const int* __restrict__ perm, int n, const double* __restrict__ p,
                           double2* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    out[i] = make_double2(p[2 * o], p[2 * o + 1]);
}

template <int TIER>
__global__ void k_e1(const double* __restrict__ x, int n, double* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    out[i] = e1_eval<TIER>(v, exp(-v), 1.0 / v);
}

}  // namespace

int launch_pack_sources(hipStream_t s, int kind, const int* perm, int n, const double* p,
                        const double* f, const double* nrm, double* out)
{
    if (n <= 0) return SE2P_OK;
    dim3 g((n + 255) / 256), b(256);
    if (kind == SE2P_SLP)
        hipLaunchKernelGGL(k_pack_slp, g, b, 0, s, perm, n, p, f, (double4*)out);
    else
        hipLaunchKernelGGL(k_pack_dlp, g, b, 0, s, perm, n, p, f, nrm, (DlpRec*)out);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_pack_targets(hipStream_t s, const int* perm, int n, const double* p, double2* out)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_pack_tar, dim3((n + 255) / 256), dim3(256), 0, s, perm, n, p, out);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_real(hipStream_t s, int kind, int nsx, int nsy, double Lx, double Ly, double xi,
                const int* tstart, const int* tperm, const double2* tsorted,
                const int* sstart, const double* ssorted, int mean_targets, double* u)
{
    const int nbox = nsx * nsy;
    const double xi2 = xi * xi;
    const double cut2 = Lx * Ly / nsx / nsy;
    // -(1 + gamma/2 + ln xi) (mex_stokes_slp_real.cpp:124)
    const double self = -(1.0 + 0.5772156649015328606 / 2 + log(xi));
    const bool big = mean_targets > 96;
    if (kind == SE2P_SLP) {
        if (big)
            hipLaunchKernelGGL((k_slp_real<256, kE1Full>), dim3(nbox), dim3(256), 0, s, nsx, nsy,
                               Lx, Ly, xi2, cut2, self, tstart, tperm, tsorted, sstart,
                               (const double4*)ssorted, u);
        else
            hipLaunchKernelGGL((k_slp_real<64, kE1Full>), dim3(nbox), dim3(64), 0, s, nsx, nsy,
                               Lx, Ly, xi2, cut2, self, tstart, tperm, tsorted, sstart,
                               (const double4*)ssorted, u);
    } else {
        if (big)
            hipLaunchKernelGGL((k_dlp_real<256>), dim3(nbox), dim3(256), 0, s, nsx, nsy, Lx, Ly,
                               xi2, cut2, tstart, tperm, tsorted, sstart,
                               (const DlpRec*)ssorted, u);
        else
            hipLaunchKernelGGL((k_dlp_real<64>), dim3(nbox), dim3(64), 0, s, nsx, nsy, Lx, Ly,
                               xi2, cut2, tstart, tperm, tsorted, sstart,
                               (const DlpRec*)ssorted, u);
    }
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_e1(hipStream_t s, const double* x, int n, int tier, double* out)
{
    if (n <= 0) return SE2P_OK;
    if (tier == kE1Full)
        hipLaunchKernelGGL(k_e1<kE1Full>, dim3((n + 255) / 256), dim3(256), 0, s, x, n, out);
    else
        hipLaunchKernelGGL(k_e1<kE1Fast>, dim3((n + 255) / 256), dim3(256), 0, s, x, n, out);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

}  // namespace se2p
