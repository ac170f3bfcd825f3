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
// u2 += E/2 f2 - (e/r^2) dx (dx f2 - dy f1).  ixi2 = 1/xi^2.
template <int TIER>
__device__ __forceinline__ void slp_pair(double dx, double dy, double r2, double xi2,
                                         double ixi2, double f1, double f2, double& a1,
                                         double& a2)
{
    const double ir2 = rcp64(r2);
    const double x = xi2 * r2;
    const double e = exp(-x);
    const double E = e1_eval<TIER>(x, e, ir2 * ixi2);
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

template <int KIND>
struct SrcOf;
template <>
struct SrcOf<SE2P_SLP> {
    using T = double4;
};
template <>
struct SrcOf<SE2P_DLP> {
    using T = DlpRec;
};

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// Cluster-pair cell-list kernel.  One workgroup of 16 waves per target cell.
// The sources of the 3 x 3 stencil cells, with their periodic image shift,
// are staged in LDS as one list in which every cell is padded to a multiple of
// 8, so each run of 8 (a source cluster, Morton-ordered inside its cell)
// belongs to a single cell.  The cell's targets, also Morton-ordered, form
// clusters of 8; a wave takes one target cluster at a time and its lane
// (tl, sl) pairs target tl with slot sl of every source cluster, evaluating an
// 8 x 8 block of pairs per step.  Blocks from neighbour cells whose bounding
// boxes are at least the cutoff apart are skipped by a wave-uniform test; the
// remaining lanes apply the reference's rule (own cell: every pair; neighbour
// cells: r^2 < cutoff^2).  Per-target sums go through LDS when the
// neighbourhood needs more than one staging round or the cell more than one
// pass of targets.
constexpr int kRealTargets = 256;  // targets per pass

// Staged sources per round: 80 KB (16-wave variant, cells of ~100+ targets) or
// 32 KB (4-wave variant, small cells: several workgroups share a CU).
template <int KIND, int NW>
struct RealCap {
    static constexpr int rec = KIND == SE2P_SLP ? 32 : 64;
    static constexpr int value = (NW >= 16 ? 81920 : 32768) / rec;
};

template <int KIND, int TIER, int NW>
__global__ __launch_bounds__(NW * 64) void k_real_cp(
    int nsx, int nsy, double Lx, double Ly, double xi2, double cut2, double self,
    const int* __restrict__ tstart, const int* __restrict__ tperm,
    const double2* __restrict__ tsorted, const int* __restrict__ sstart,
    const typename SrcOf<KIND>::T* __restrict__ ssorted, double* __restrict__ u)
{
    using Rec = typename SrcOf<KIND>::T;
    constexpr int CAP = RealCap<KIND, NW>::value;
    constexpr int kRealWaves = NW;
    constexpr int kRealThreads = NW * 64;
    __shared__ Rec buf[CAP];
    __shared__ double4 sbox[CAP / 8];
    __shared__ int sinfo[CAP / 8];  // (own cell << 4) | valid count
    __shared__ double2 acc[kRealTargets];
    __shared__ int cell_src[9], cell_n[9], cell_pad[10];
    __shared__ double cell_ox[9], cell_oy[9];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tl = lane >> 3, sl = lane & 7;
    const int b = blockIdx.x;
    const int bx = b % nsx, by = b / nsx;
    const int t0 = tstart[b], t1 = tstart[b + 1];
    if (t0 == t1) return;
    const double ixi2 = 1.0 / xi2;
    const double skip2 = cut2 * (1.0 + 1e-12);

    if (threadIdx.x < 9) {
        const int j = threadIdx.x;
        int qx = bx, qy = by;
        if (j > 0) {
            qx += c_nbdx[j - 1];
            qy += c_nbdy[j - 1];
        }
        const int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
        const int sb = wy * nsx + wx;
        cell_src[j] = sstart[sb];
        cell_n[j] = sstart[sb + 1] - sstart[sb];
        cell_ox[j] = (Lx * (qx - wx)) / nsx;
        cell_oy[j] = (Ly * (qy - wy)) / nsy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int o = 0;
        for (int j = 0; j < 9; ++j) {
            cell_pad[j] = o;
            o += (cell_n[j] + 7) & ~7;
        }
        cell_pad[9] = o;
    }
    __syncthreads();
    const int total = cell_pad[9];
    const bool multi = total > CAP || total == 0 || (t1 - t0) > kRealTargets;

    for (int tb = t0; tb < t1; tb += kRealTargets) {
        const int nt = min(kRealTargets, t1 - tb);
        const int ntc = (nt + 7) >> 3;
        for (int i = threadIdx.x; i < nt; i += kRealThreads) acc[i] = make_double2(0.0, 0.0);
        for (int r0 = 0; r0 < total; r0 += CAP) {
            const int rn = min(CAP, total - r0);
            __syncthreads();
            for (int i = threadIdx.x; i < rn; i += kRealThreads) {
                const int g = r0 + i;
                int j = 0;
#pragma unroll
                for (int q = 1; q < 9; ++q) j += (g >= cell_pad[q]);
                const int k = g - cell_pad[j];
                Rec rec;
                if (k < cell_n[j]) {
                    rec = ssorted[cell_src[j] + k];
                    rec.x += cell_ox[j];
                    rec.y += cell_oy[j];
                } else {
                    rec = Rec{};
                }
                buf[i] = rec;
            }
            __syncthreads();
            for (int c = threadIdx.x; c < (rn >> 3); c += kRealThreads) {
                const int g = r0 + c * 8;
                int j = 0;
#pragma unroll
                for (int q = 1; q < 9; ++q) j += (g >= cell_pad[q]);
                const int cnt = min(8, cell_n[j] - (g - cell_pad[j]));
                double lx = INFINITY, hx = -INFINITY, ly = INFINITY, hy = -INFINITY;
                for (int i = 0; i < cnt; ++i) {
                    const Rec& q = buf[c * 8 + i];
                    lx = fmin(lx, q.x);
                    hx = fmax(hx, q.x);
                    ly = fmin(ly, q.y);
                    hy = fmax(hy, q.y);
                }
                sbox[c] = make_double4(lx, hx, ly, hy);
                sinfo[c] = ((j == 0) << 4) | cnt;
            }
            __syncthreads();
            const int ncl = rn >> 3;
            for (int kc = wave; kc < ntc; kc += kRealWaves) {
                const int ti = tb + kc * 8 + tl;
                const bool tact = ti < tb + nt;
                const double2 pt = tact ? tsorted[ti] : make_double2(0.0, 0.0);
                double lx = tact ? pt.x : INFINITY, hx = tact ? pt.x : -INFINITY;
                double ly = tact ? pt.y : INFINITY, hy = tact ? pt.y : -INFINITY;
#pragma unroll
                for (int m = 8; m < 64; m <<= 1) {
                    lx = fmin(lx, shfl_xor_d(lx, m));
                    hx = fmax(hx, shfl_xor_d(hx, m));
                    ly = fmin(ly, shfl_xor_d(ly, m));
                    hy = fmax(hy, shfl_xor_d(hy, m));
                }
                double a1 = 0, a2 = 0;
                for (int c = 0; c < ncl; ++c) {
                    const int info = sinfo[c];
                    const bool own = info >> 4;
                    if (!own) {
                        const double4 bb = sbox[c];
                        const double gx = fmax(0.0, fmax(bb.x - hx, lx - bb.y));
                        const double gy = fmax(0.0, fmax(bb.z - hy, ly - bb.w));
                        if (gx * gx + gy * gy > skip2) continue;
                    }
                    const bool valid = tact && sl < (info & 15);
                    const Rec s = buf[c * 8 + sl];
                    const double dx = pt.x - s.x, dy = pt.y - s.y;
                    const double r2 = dx * dx + dy * dy;
                    if (KIND == SE2P_SLP) {
                        const double4& q = reinterpret_cast<const double4&>(s);
                        if (valid && own && fabs(r2) < 1e-15) {
                            a1 += self * q.z;
                            a2 += self * q.w;
                        } else if (valid && (own || r2 < cut2)) {
                            slp_pair<TIER>(dx, dy, r2, xi2, ixi2, q.z, q.w, a1, a2);
                        }
                    } else {
                        const DlpRec& q = reinterpret_cast<const DlpRec&>(s);
                        if (valid && (own ? (r2 != 0) : (r2 < cut2)))
                            dlp_pair(dx, dy, r2, xi2, q, a1, a2);
                    }
                }
#pragma unroll
                for (int m = 1; m < 8; m <<= 1) {
                    a1 += shfl_xor_d(a1, m);
                    a2 += shfl_xor_d(a2, m);
                }
                if (sl == 0 && tact) {
                    if (multi) {
                        double2 v = acc[kc * 8 + tl];
                        acc[kc * 8 + tl] = make_double2(v.x + a1, v.y + a2);
                    } else {
                        const int o = tperm[ti];
                        u[2 * o] = a1 / (4 * kPi);
                        u[2 * o + 1] = a2 / (4 * kPi);
                    }
                }
            }
        }
        if (multi) {
            __syncthreads();
            for (int i = threadIdx.x; i < nt; i += kRealThreads) {
                const int o = tperm[tb + i];
                u[2 * o] = acc[i].x / (4 * kPi);
                u[2 * o + 1] = acc[i].y / (4 * kPi);
            }
            __syncthreads();
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
    const int nbox = nsx * nsy;
    const double xi2 = xi * xi;
    const double cut2 = Lx * Ly / nsx / nsy;
    // -(1 + gamma/2 + ln xi) (mex_stokes_slp_real.cpp:124)
    const double self = -(1.0 + 0.5772156649015328606 / 2 + log(xi));
    dim3 gr(nbox);
    // small cells: 4 waves per cell; the 3 x 3 neighbourhood fits the small staging area
    const bool small = 9L * mean_targets <= 900;
#define SE2P_LAUNCH_REAL(K, T, NWV, REC)                                                        \
    hipLaunchKernelGGL((k_real_cp<K, T, NWV>), gr, dim3(NWV * 64), 0, s, nsx, nsy, Lx, Ly, xi2, \
                       cut2, self, tstart, tperm, tsorted, sstart, (const REC*)ssorted, u)
    if (kind == SE2P_SLP && tier == kE1Full) {
        if (small)
            SE2P_LAUNCH_REAL(SE2P_SLP, kE1Full, 4, double4);
        else
            SE2P_LAUNCH_REAL(SE2P_SLP, kE1Full, 16, double4);
    } else if (kind == SE2P_SLP) {
        if (small)
            SE2P_LAUNCH_REAL(SE2P_SLP, kE1Fast, 4, double4);
        else
            SE2P_LAUNCH_REAL(SE2P_SLP, kE1Fast, 16, double4);
    } else {
        if (small)
            SE2P_LAUNCH_REAL(SE2P_DLP, kE1Full, 4, DlpRec);
        else
            SE2P_LAUNCH_REAL(SE2P_DLP, kE1Full, 16, DlpRec);
    }
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
