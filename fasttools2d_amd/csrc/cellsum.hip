// Real-space cell-list sums for the Stokeslet (SLP) and stresslet (DLP).
//
// Follows mex_stokes_slp_real.cpp:58-357 and mex_stokes_dlp_real.cpp:15-211
// (PeriodicEwald/src/mex/...): cells of size Lx/nside_x x Ly/nside_y; a target
// takes every source of its own cell with no cutoff (SLP: coincident points
// r^2 < 1e-15 take the self term -(1 + gamma/2 + ln xi) f; DLP: coincident
// points are skipped) and every source of the 8 neighbour cells (periodic
// images) with r^2 < Lx Ly/(nside_x nside_y).  Output scaled by 1/(4 pi) and
// written in input order.  Sources further than one cell away are never within
// the cutoff, so "own cell, or r^2 < cutoff^2" over any window that contains
// the 3 x 3 stencil selects exactly the reference's pairs.
//
// Mapping: one workgroup of 256 lanes per block of bs x bs cells, one lane per
// target.  The sources of the block's window ((bs+2)^2 cells, image shifts
// applied) are staged in LDS with a tag naming their cell position in the
// window.  Candidates are tested 64 at a time in strided chunks (chunk c holds
// staged sources c, c + nch, c + 2 nch, ...), so every chunk samples the whole
// window and each lane's hit count per chunk is close to the mean.  Each lane
// collects its hits as two 64-bit masks -- pairs with xi^2 r^2 >= 1 (E1 by
// its rational fit) and below 1 (E1 by log and series) -- and then walks its
// own masks, so no lane evaluates a pair it does not sum and the two E1
// regimes do not diverge inside a wave.
#include "se2p_internal.h"
#include "e1_eval.h"

#include <algorithm>
#include <cmath>

namespace se2p {

namespace {

struct alignas(64) DlpRec {
    double x, y, f1, f2, n1, n2, fn, pad;
};

template <int KIND>
struct SrcOf;
template <>
struct SrcOf<SE2P_SLP> {
    using T = double4;
    static constexpr int cap = 1024;  // 32 KB of staged records: 4 workgroups per CU
};
template <>
struct SrcOf<SE2P_DLP> {
    using T = DlpRec;
    static constexpr int cap = 512;
};

constexpr int kBlkThreads = 256;
constexpr int kMaxBlockSide = 4;
constexpr int kMaxEntries = (kMaxBlockSide + 2) * (kMaxBlockSide + 2);

// 4 pi times the SLP pair term for r = target - source:
//   E1(xi^2 r^2)/2 f + e^{-xi^2 r^2} ((r.f) r/r^2 - f)
// in the cancellation-free form u1 += E/2 f1 + (e/r^2) dy (dx f2 - dy f1),
// u2 += E/2 f2 - (e/r^2) dx (dx f2 - dy f1).  NEAR selects the E1 regime
// (xi^2 r^2 < 1), decided by the caller.
template <int TIER, bool NEAR>
__device__ __forceinline__ void slp_pair(double dx, double dy, double r2, double xi2,
                                         double ixi2, double f1, double f2, double& a1,
                                         double& a2)
{
    const double ir2 = rcp64(r2);
    const double x = xi2 * r2;
    const double e = exp(-x);
    const double E = NEAR ? e1_near<TIER>(x) : e1_far<TIER>(e, ir2 * ixi2);
    const double c = e * ir2;
    const double cr = dx * f2 - dy * f1;
    a1 += 0.5 * E * f1 + c * dy * cr;
    a2 += 0.5 * E * f2 - c * dx * cr;
}

// 4 pi times the DLP pair term for r = target - source:
//   e^{-xi^2 r^2} [ -4 (1 + xi^2 r^2)/r^4 r (r.f)(r.n) + 2 xi^2 (f (r.n) + n (r.f) + r (f.n)) ]
__device__ __forceinline__ void dlp_pair(double r1, double r2, double rr, double xi2,
                                         const DlpRec& s, double& a1, double& a2)
{
    const double e = exp(-xi2 * rr);
    const double ir = rcp64(rr);
    const double rf = r1 * s.f1 + r2 * s.f2;
    const double rn = r1 * s.n1 + r2 * s.n2;
    const double b = -4.0 * (1.0 + xi2 * rr) * (ir * ir) * rf * rn;
    const double c = 2.0 * xi2;
    a1 += e * (b * r1 + c * (s.f1 * rn + s.n1 * rf + s.fn * r1));
    a2 += e * (b * r2 + c * (s.f2 * rn + s.n2 * rf + s.fn * r2));
}

__device__ __forceinline__ int wrapi(int q, int n) { return ((q % n) + n) % n; }

template <int KIND, int TIER>
__global__ __launch_bounds__(kBlkThreads) void k_real_blk(
    int nsx, int nsy, int bs, int nbx, double Lx, double Ly, double xi2, double cut2,
    double self, const int* __restrict__ tstart, const int* __restrict__ tperm,
    const double2* __restrict__ tsorted, const int* __restrict__ sstart,
    const typename SrcOf<KIND>::T* __restrict__ ssorted, double* __restrict__ u)
{
    using Rec = typename SrcOf<KIND>::T;
    constexpr int CAP = SrcOf<KIND>::cap;
    __shared__ Rec buf[CAP];
    __shared__ int tagb[CAP];
    __shared__ int e_src[kMaxEntries], e_off[kMaxEntries + 1], e_tag[kMaxEntries];
    __shared__ double e_ox[kMaxEntries], e_oy[kMaxEntries];
    __shared__ int c_t0[kMaxBlockSide * kMaxBlockSide], c_off[kMaxBlockSide * kMaxBlockSide + 1];
    __shared__ int c_tag[kMaxBlockSide * kMaxBlockSide];

    const int tid = threadIdx.x;
    const int bx0 = (blockIdx.x % nbx) * bs, by0 = (blockIdx.x / nbx) * bs;
    const int bw = min(bs, nsx - bx0), bh = min(bs, nsy - by0);
    const int W = bs + 2;
    const int nent = (bw + 2) * (bh + 2), ncell = bw * bh;
    if (tid < nent) {
        const int ux = tid / (bh + 2) - 1, uy = tid % (bh + 2) - 1;
        const int qx = bx0 + ux, qy = by0 + uy;
        const int wx = wrapi(qx, nsx), wy = wrapi(qy, nsy);
        const int cell = wy * nsx + wx;
        e_src[tid] = sstart[cell];
        e_off[tid + 1] = sstart[cell + 1] - sstart[cell];  // count, prefixed below
        e_ox[tid] = (Lx * (qx - wx)) / nsx;
        e_oy[tid] = (Ly * (qy - wy)) / nsy;
        e_tag[tid] = (ux + 1) * W + (uy + 1);
    }
    if (tid < ncell) {
        const int cx = tid / bh, cy = tid % bh;
        const int cell = (by0 + cy) * nsx + (bx0 + cx);
        c_t0[tid] = tstart[cell];
        c_off[tid + 1] = tstart[cell + 1] - tstart[cell];
        c_tag[tid] = (cx + 1) * W + (cy + 1);
    }
    __syncthreads();
    if (tid == 0) {
        e_off[0] = 0;
        for (int e = 0; e < nent; ++e) e_off[e + 1] += e_off[e];
        c_off[0] = 0;
        for (int c = 0; c < ncell; ++c) c_off[c + 1] += c_off[c];
    }
    __syncthreads();
    const int total = e_off[nent], ntar = c_off[ncell];
    const double ixi2 = 1.0 / xi2;

    for (int tb = 0; tb < ntar; tb += kBlkThreads) {
        const int i = tb + tid;
        const bool act = i < ntar;
        int ti = 0, ttag = -1;
        double2 pt = make_double2(0.0, 0.0);
        if (act) {
            int c = 0;
            while (c + 1 < ncell && c_off[c + 1] <= i) ++c;
            ti = c_t0[c] + (i - c_off[c]);
            ttag = c_tag[c];
            pt = tsorted[ti];
        }
        double a1 = 0, a2 = 0;
        // Staging rounds take every nround-th source of the window, so each round
        // (and each strided chunk inside it) samples the whole window.
        const int nround = (total + CAP - 1) / CAP;
        for (int r = 0; r < nround; ++r) {
            const int rn = (total - r + nround - 1) / nround;
            __syncthreads();
            for (int k = tid; k < rn; k += kBlkThreads) {
                const int g = r + k * nround;
                int e = 0;
                while (e_off[e + 1] <= g) ++e;
                Rec rec = ssorted[e_src[e] + (g - e_off[e])];
                rec.x += e_ox[e];
                rec.y += e_oy[e];
                buf[k] = rec;
                tagb[k] = e_tag[e];
            }
            __syncthreads();
            const int nch = (rn + 63) >> 6;
            for (int c = 0; c < nch; ++c) {
                const int kmax = (rn - c + nch - 1) / nch;  // <= 64
                unsigned long long far = 0, near = 0;
#pragma unroll 8
                for (int k = 0; k < kmax; ++k) {
                    const int idx = c + k * nch;
                    const double2 sp = *reinterpret_cast<const double2*>(&buf[idx]);
                    const double dx = pt.x - sp.x, dy = pt.y - sp.y;
                    const double r2 = dx * dx + dy * dy;
                    const bool own = tagb[idx] == ttag;
                    const unsigned long long bit = 1ull << k;
                    if (KIND == SE2P_SLP) {
                        const bool hit = act & (own | (r2 < cut2));
                        const bool nr = xi2 * r2 < 1.0;
                        far |= (hit & !nr) ? bit : 0ull;
                        near |= (hit & nr) ? bit : 0ull;
                    } else {
                        const bool hit = act & (own ? (r2 != 0) : (r2 < cut2));
                        far |= hit ? bit : 0ull;
                    }
                }
                while (far) {
                    const int k = __builtin_ctzll(far);
                    far &= far - 1;
                    const Rec& s = buf[c + k * nch];
                    const double dx = pt.x - s.x, dy = pt.y - s.y;
                    const double r2 = dx * dx + dy * dy;
                    if (KIND == SE2P_SLP) {
                        const double4& q = reinterpret_cast<const double4&>(s);
                        slp_pair<TIER, false>(dx, dy, r2, xi2, ixi2, q.z, q.w, a1, a2);
                    } else {
                        dlp_pair(dx, dy, r2, xi2, reinterpret_cast<const DlpRec&>(s), a1, a2);
                    }
                }
                if (KIND == SE2P_SLP) {
                    while (near) {
                        const int k = __builtin_ctzll(near);
                        near &= near - 1;
                        const int idx = c + k * nch;
                        const double4& q = reinterpret_cast<const double4&>(buf[idx]);
                        const double dx = pt.x - q.x, dy = pt.y - q.y;
                        const double r2 = dx * dx + dy * dy;
                        if (tagb[idx] == ttag && fabs(r2) < 1e-15) {
                            a1 += self * q.z;
                            a2 += self * q.w;
                        } else {
                            slp_pair<TIER, true>(dx, dy, r2, xi2, ixi2, q.z, q.w, a1, a2);
                        }
                    }
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

int launch_real(hipStream_t s, int kind, int tier, int nsx, int nsy, double Lx, double Ly,
                double xi, const int* tstart, const int* tperm, const double2* tsorted,
                const int* sstart, const double* ssorted, int mean_targets, double* u)
{
    const double xi2 = xi * xi;
    const double cut2 = Lx * Ly / nsx / nsy;
    // -(1 + gamma/2 + ln xi) (mex_stokes_slp_real.cpp:124)
    const double self = -(1.0 + 0.5772156649015328606 / 2 + log(xi));
    // blocks of bs x bs cells holding about 256 targets
    int bs = (int)std::floor(std::sqrt(256.0 / (mean_targets > 1 ? mean_targets : 1)));
    bs = std::max(1, std::min(bs, kMaxBlockSide));
    bs = std::min(bs, std::min(nsx, nsy));
    const int nbx = (nsx + bs - 1) / bs, nby = (nsy + bs - 1) / bs;
    dim3 gr(nbx * nby), bl(kBlkThreads);
#define SE2P_LAUNCH_REAL(K, T, REC)                                                              \
    hipLaunchKernelGGL((k_real_blk<K, T>), gr, bl, 0, s, nsx, nsy, bs, nbx, Lx, Ly, xi2, cut2,  \
                       self, tstart, tperm, tsorted, sstart, (const REC*)ssorted, u)
    if (kind == SE2P_SLP && tier == kE1Full)
        SE2P_LAUNCH_REAL(SE2P_SLP, kE1Full, double4);
    else if (kind == SE2P_SLP)
        SE2P_LAUNCH_REAL(SE2P_SLP, kE1Fast, double4);
    else
        SE2P_LAUNCH_REAL(SE2P_DLP, kE1Full, DlpRec);
#undef SE2P_LAUNCH_REAL
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
