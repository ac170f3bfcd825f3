// Real-space cell-list sums for the Stokeslet (SLP) and stresslet (DLP).
//
// Follows mex_stokes_slp_real.cpp:58-357 and mex_stokes_dlp_real.cpp:15-211
// (PeriodicEwald/src/mex/...): cells of size Lx/nside_x x Ly/nside_y; a target
// takes every source of its own cell with no cutoff (SLP: coincident points
// r^2 < 1e-15 take the self term -(1 + gamma/2 + ln xi) f -- free-space sums:
// exactly coincident points only, SysDesc::self_r2; DLP: coincident
// points are skipped) and every source of the 8 neighbour cells (periodic
// images) with r^2 < Lx Ly/(nside_x nside_y).  Output scaled by 1/(4 pi) and
// written in input order.  The Stokeslet pressure
// (mex_stokes_slp_pressure_real.cpp:60-160) uses the same pair set as the DLP
// (coincident points of the own cell skipped) with the pair term
// (r.f) e^{-xi^2 r^2}/r^2 and one output scaled by 1/(2 pi); the stresslet
// pressure (mex_stokes_dlp_pressure_real.cpp:100-192) likewise with its own pair
// term.  The other derived quantities (velocity gradient, pressure gradient,
// vorticity and stress of both layers, mex_stokes_{slp,dlp}_{gradient,
// pressure_grad,vorticity,stress}_real.cpp) share that pair set too (the
// stresslet vorticity skips own-cell pairs with r^2 < 1e-13,
// mex_stokes_dlp_vorticity_real.cpp:128) and write 1, 2 or 4 values per
// target.  Sources further than one cell away are never within
// the cutoff, so "own cell, or r^2 < cutoff^2" over any window that contains
// the 3 x 3 stencil selects exactly the reference's pairs.
//
// Mapping: one workgroup of 256 lanes per block of bs x bs cells, one lane per
// target.  The sources of the block's window ((bs+2)^2 cells, image shifts
// applied) are staged in LDS, window column by window column, so the 3 x 3
// stencil of any target cell is three contiguous runs of the staging buffer
// (one per stencil column) and its own cell a sub-run of the middle one.  Each
// lane walks only its own stencil, 64 candidates at a time: it tests them
// (own cell, or r^2 < cutoff^2) into two 64-bit masks -- pairs with
// xi^2 r^2 >= 1 (E1 by its rational fit) and below 1 (E1 by log and series) --
// and then evaluates the pairs of each mask, so no lane evaluates a pair it
// does not sum and the two E1 regimes do not diverge inside a wave.  Windows
// larger than the staging buffer are staged in rounds; round r holds every
// nround-th source, which keeps each stencil run contiguous in every round.
#include "se2p_internal.h"
#include "e1_eval.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace se2p {

namespace {

struct alignas(64) DlpRec {
    double x, y, f1, f2, n1, n2, fn, pad;
};

template <int KIND>
struct SrcOf {
    using T = typename std::conditional<dlp_family(KIND), DlpRec, double4>::type;
    // 32 KB of staged records: 4 workgroups per CU
    static constexpr int cap = dlp_family(KIND) ? 512 : 1024;
};

// Divisor of the accumulated sums at output.
constexpr double out_div(int kind)
{
    return (kind == SE2P_SLP_PRESSURE || kind == SE2P_DLP_PRESSURE || kind == SE2P_SLP_PRESSURE_GRAD ||
            kind == SE2P_DLP_PRESSURE_GRAD || kind == SE2P_SLP_VORTICITY)
               ? 2 * kPi
           : kind == SE2P_DLP_VORTICITY ? kPi
                                        : 4 * kPi;
}

constexpr int kBlkThreads = kRealLanes;
constexpr int kMaxBlockSide = 4;
constexpr int kMaxEntries = (kMaxBlockSide + 2) * (kMaxBlockSide + 2);
constexpr int kWords = 2, kChunk = 64 * kWords;  // candidates per test chunk

// E = E1(xi^2 r^2) and c = e^{-xi^2 r^2}/r^2.  In the far regime (x = xi^2 r^2 >=
// 1) the fit E1 = e^{-x} t P(t)/Q(t), t = 1/x, is evaluated as e^{-x} P~(x)/(x
// Q~(x)) with the reversed polynomials P~(x) = x^d P(1/x), Q~(x) = x^d Q(1/x)
// (positive coefficients, so Horner in x has no cancellation), and 1/r^2 =
// xi^2 Q~/(x Q~): one reciprocal serves both.
template <int TIER, bool NEAR>
__device__ __forceinline__ void slp_radial(double r2, double xi2, double& E, double& c)
{
    const double x = xi2 * r2;
    const double e = TIER == kE1Fast ? exp_neg(x) : exp(-x);
    if (NEAR) {
        E = e1_near<TIER>(x);
        c = e * (TIER == kE1Fast ? rcp64_1(r2) : rcp64(r2));
    } else {
        using C = E1Coeffs<TIER>;
        const double pr = horner_rev(C::farp, x), qr = horner_rev(C::farq, x);
        const double eD = e * (TIER == kE1Fast ? rcp64_1(x * qr) : rcp64(x * qr));
        E = eD * pr;
        c = eD * qr * xi2;
    }
}

// 4 pi times the SLP pair term for r = target - source:
//   E1(xi^2 r^2)/2 f + e^{-xi^2 r^2} ((r.f) r/r^2 - f)
// in the cancellation-free form u1 += E/2 f1 + (e/r^2) dy (dx f2 - dy f1),
// u2 += E/2 f2 - (e/r^2) dx (dx f2 - dy f1).  NEAR selects the E1 regime
// (xi^2 r^2 < 1), decided by the caller.
template <int TIER, bool NEAR>
__device__ __forceinline__ void slp_pair(double dx, double dy, double r2, double xi2,
                                         double f1, double f2, double& a1,
                                         double& a2)
{
    double E, c;
    slp_radial<TIER, NEAR>(r2, xi2, E, c);
    const double cr = dx * f2 - dy * f1;
    a1 += 0.5 * E * f1 + c * dy * cr;
    a2 += 0.5 * E * f2 - c * dx * cr;
}

// slp_pair for a pair evaluated once for both points: (a1, a2) += the term for
// density f at the target, (b1, b2) = the term for density g at the source (the
// pair term is even in r).
template <int TIER, bool NEAR>
__device__ __forceinline__ void slp_pair2(double dx, double dy, double r2, double xi2,
                                          double f1, double f2, double g1,
                                          double g2, double& a1, double& a2, double& b1,
                                          double& b2)
{
    double E, c;
    slp_radial<TIER, NEAR>(r2, xi2, E, c);
    const double hE = 0.5 * E;
    const double cf = dx * f2 - dy * f1, cg = dx * g2 - dy * g1;
    a1 += hE * f1 + c * dy * cf;
    a2 += hE * f2 - c * dx * cf;
    b1 = hE * g1 + c * dy * cg;
    b2 = hE * g2 - c * dx * cg;
}

// The FP32 tier's far pairs (x = xi^2 r^2 >= 1): e^{-x} from the hardware
// exp2, E1 from the 'f32' fit, one hardware reciprocal for both E1 and 1/r^2;
// the term (in the same cancellation-free form as slp_pair) is added to FP32
// partial sums that the caller folds into its FP64 sums after every chunk.
__device__ __forceinline__ void slp_radial_f32(float r2, float xi2, float& E, float& c)
{
    const float x = xi2 * r2;
    const float e = __expf(-x);
    const float pr = horner_rev_f(E1F32::farp, x), qr = horner_rev_f(E1F32::farq, x);
    const float eD = e * __builtin_amdgcn_rcpf(x * qr);
    E = eD * pr;
    c = eD * qr * xi2;
}

__device__ __forceinline__ void slp_pair_f32(float dx, float dy, float r2, float xi2, float f1,
                                             float f2, float& a1, float& a2)
{
    float E, c;
    slp_radial_f32(r2, xi2, E, c);
    const float cr = dx * f2 - dy * f1;
    a1 += 0.5f * E * f1 + c * dy * cr;
    a2 += 0.5f * E * f2 - c * dx * cr;
}

__device__ __forceinline__ void slp_pair2_f32(float dx, float dy, float r2, float xi2, float f1,
                                              float f2, float g1, float g2, float& a1, float& a2,
                                              float& b1, float& b2)
{
    float E, c;
    slp_radial_f32(r2, xi2, E, c);
    const float hE = 0.5f * E;
    const float cf = dx * f2 - dy * f1, cg = dx * g2 - dy * g1;
    a1 += hE * f1 + c * dy * cf;
    a2 += hE * f2 - c * dx * cf;
    b1 = hE * g1 + c * dy * cg;
    b2 = hE * g2 - c * dx * cg;
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

// dlp_pair in FP32 (the FP32 tier): every pair of the stresslet is a far pair
// in this sense (no E1), so the whole term is FP32, added to FP32 partial sums.
__device__ __forceinline__ void dlp_pair_f32(float r1, float r2, float rr, float xi2,
                                             const DlpRec& s, float& a1, float& a2)
{
    const float e = __expf(-xi2 * rr);
    const float ir = __builtin_amdgcn_rcpf(rr);
    const float f1 = (float)s.f1, f2 = (float)s.f2, n1 = (float)s.n1, n2 = (float)s.n2;
    const float rf = r1 * f1 + r2 * f2;
    const float rn = r1 * n1 + r2 * n2;
    const float b = -4.0f * (1.0f + xi2 * rr) * (ir * ir) * rf * rn;
    const float c = 2.0f * xi2;
    const float fn = (float)s.fn;
    a1 += e * (b * r1 + c * (f1 * rn + n1 * rf + fn * r1));
    a2 += e * (b * r2 + c * (f2 * rn + n2 * rf + fn * r2));
}

// dlp_pair for a pair evaluated once for both points: (a1, a2) += the term of
// the source (f, n) at the target; (b1, b2) = the term of the target's (g, m)
// at the source, which is the same expression negated (the kernel is odd in r).
__device__ __forceinline__ void dlp_pair2(double r1, double r2, double rr, double xi2, double f1,
                                          double f2, double n1, double n2, double g1, double g2,
                                          double m1, double m2, double& a1, double& a2,
                                          double& b1, double& b2)
{
    const double e = exp(-xi2 * rr);
    const double ir = rcp64(rr);
    const double k4 = -4.0 * (1.0 + xi2 * rr) * (ir * ir);
    const double c = 2.0 * xi2;
    const double rf = r1 * f1 + r2 * f2, rn = r1 * n1 + r2 * n2, fn = f1 * n1 + f2 * n2;
    const double bs = k4 * rf * rn;
    a1 += e * (bs * r1 + c * (f1 * rn + n1 * rf + fn * r1));
    a2 += e * (bs * r2 + c * (f2 * rn + n2 * rf + fn * r2));
    const double rg = r1 * g1 + r2 * g2, rm = r1 * m1 + r2 * m2, gm = g1 * m1 + g2 * m2;
    const double bt = k4 * rg * rm;
    b1 = -e * (bt * r1 + c * (g1 * rm + m1 * rg + gm * r1));
    b2 = -e * (bt * r2 + c * (g2 * rm + m2 * rg + gm * r2));
}

__device__ __forceinline__ void dlp_pair2_f32(float r1, float r2, float rr, float xi2, float f1,
                                              float f2, float n1, float n2, float g1, float g2,
                                              float m1, float m2, float& a1, float& a2,
                                              float& b1, float& b2)
{
    const float e = __expf(-xi2 * rr);
    const float ir = __builtin_amdgcn_rcpf(rr);
    const float k4 = -4.0f * (1.0f + xi2 * rr) * (ir * ir);
    const float c = 2.0f * xi2;
    const float rf = r1 * f1 + r2 * f2, rn = r1 * n1 + r2 * n2, fn = f1 * n1 + f2 * n2;
    const float bs = k4 * rf * rn;
    a1 += e * (bs * r1 + c * (f1 * rn + n1 * rf + fn * r1));
    a2 += e * (bs * r2 + c * (f2 * rn + n2 * rf + fn * r2));
    const float rg = r1 * g1 + r2 * g2, rm = r1 * m1 + r2 * m2, gm = g1 * m1 + g2 * m2;
    const float bt = k4 * rg * rm;
    b1 = -e * (bt * r1 + c * (g1 * rm + m1 * rg + gm * r1));
    b2 = -e * (bt * r2 + c * (g2 * rm + m2 * rg + gm * r2));
}

// 2 pi times the stresslet pressure pair term for r = target - source
// (mex_stokes_dlp_pressure_real.cpp:113-116):
//   -e^{-xi^2 r^2} ((f.n - 2 xi^2 (r.n)(r.f))/r^2 - 2 (r.n)(r.f)/r^4)
__device__ __forceinline__ double dlp_pressure_pair(double r1, double r2, double rr, double xi2,
                                                    const DlpRec& s)
{
    const double ir = rcp64(rr);
    const double rf = r1 * s.f1 + r2 * s.f2, rn = r1 * s.n1 + r2 * s.n2;
    return -exp(-xi2 * rr) * ((s.fn - 2 * xi2 * rn * rf) * ir - 2 * rn * rf * ir * ir);
}

// Derived quantities, r = target - source, before the output divisor.  Tensor
// slot 2p + j holds the derivative along p of component j.
//   SLP gradient (mex_stokes_slp_gradient_real.cpp:118-132):
//     e [2 xi^2 r_p f_j + (d_pj (r.f) + r_j f_p - r_p f_j)/r^2 - 2 r_p r_j (r.f)(xi^2 + 1/r^2)/r^2]
//   SLP pressure gradient (mex_stokes_slp_pressure_grad_real.cpp:120-121):
//     -e [(f_p - 2 xi^2 (r.f) r_p)/r^2 - 2 (r.f) r_p/r^4]
//   SLP vorticity (mex_stokes_slp_vorticity_real.cpp:118-121):
//     e (1/r^2 - xi^2)(f1 r2 - f2 r1)
//   SLP stress (mex_stokes_slp_stress_real.cpp:118-129):
//     e [2 xi^2 (d_pj (r.f) + r_p f_j + r_j f_p) - 4 r_p r_j (r.f)(1 + xi^2 r^2)/r^4]
//   DLP gradient (mex_stokes_dlp_gradient_real.cpp:130-145), C = 8 xi^4/r^2 +
//   16 xi^2/r^4 + 16/r^6:
//     e [r_p r_j (r.f)(r.n) C - 4 (1 + xi^2 r^2)(d_pj (r.f)(r.n) + r_j f_p (r.n) + r_j n_p (r.f))/r^4
//        + 2 xi^2 (f_p n_j + n_p f_j + d_pj (f.n) - 2 xi^2 r_p (f_j (r.n) + n_j (r.f) + r_j (f.n)))]
//   DLP pressure gradient (mex_stokes_dlp_pressure_grad_real.cpp:122-128), with
//   a_p = r_p (f.n) + f_p (r.n) + n_p (r.f):
//     e [2 xi^2 (a_p - 2 xi^2 r_p (r.f)(r.n))/r^2 + 2 (a_p - 4 xi^2 r_p (r.n)(r.f))/r^4
//        - 8 r_p (r.n)(r.f)/r^6]
//   DLP vorticity (mex_stokes_dlp_vorticity_real.cpp:136-145), with
//   v = (r.f)(n1 r2 - n2 r1) + (r.n)(f1 r2 - f2 r1):
//     e [-(1 + xi^2 r^2) v/r^4 + xi^4 v]
//   DLP stress (mex_stokes_dlp_stress_real.cpp:132-151, mu = 1), twice
//     e [d_pj ((f.n) - 2 xi^2 (r.f)(r.n))/r^2 - 2 d_pj (r.f)(r.n)/r^4 + r_p r_j (r.f)(r.n) C
//        - 2 (1 + xi^2 r^2)(f_p r_j (r.n) + n_p r_j (r.f) + r_p f_j (r.n) + r_p n_j (r.f)
//                            + 2 d_pj (r.f)(r.n))/r^4
//        + 2 xi^2 (d_pj (f.n) + n_p f_j + f_p n_j
//                  - xi^2 (2 r_p r_j (f.n) + r_p f_j (r.n) + r_p n_j (r.f) + f_p r_j (r.n) + n_p r_j (r.f)))]
template <int KIND>
__device__ __forceinline__ void derived_pair(double r1, double r2, double rr, double xi2,
                                             const double* __restrict__ s, double* acc)
{
    const double e = exp(-xi2 * rr);
    const double ir = rcp64(rr);
    const double f[2] = {s[2], s[3]};
    const double r[2] = {r1, r2};
    const double rf = r1 * f[0] + r2 * f[1];
    if (KIND == SE2P_SLP_GRADIENT) {
        const double c = 2 * (xi2 + ir) * ir * rf;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[2 * p + j] += e * (2 * xi2 * r[p] * f[j] +
                                       ((p == j ? rf : 0.0) + r[j] * f[p] - r[p] * f[j]) * ir -
                                       r[p] * r[j] * c);
    } else if (KIND == SE2P_SLP_PRESSURE_GRAD) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
            acc[p] -= e * ((f[p] - 2 * xi2 * rf * r[p]) * ir - 2 * rf * r[p] * ir * ir);
    } else if (KIND == SE2P_SLP_VORTICITY) {
        acc[0] += e * (ir - xi2) * (f[0] * r2 - f[1] * r1);
    } else if (KIND == SE2P_SLP_STRESS) {
        const double c = 4 * rf * (1 + xi2 * rr) * ir * ir;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[2 * p + j] += e * (2 * xi2 * ((p == j ? rf : 0.0) + r[p] * f[j] + r[j] * f[p]) -
                                       r[p] * r[j] * c);
    } else {
        const double n[2] = {s[4], s[5]};
        const double fn = s[6];
        const double rn = r1 * n[0] + r2 * n[1];
        const double ir2 = ir * ir;
        if (KIND == SE2P_DLP_GRADIENT) {
            const double C = (8 * xi2 * xi2 + (16 * xi2 + 16 * ir) * ir) * ir;
            const double B = 4 * (1 + xi2 * rr) * ir2;
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[2 * p + j] +=
                        e * (r[p] * r[j] * rf * rn * C -
                             B * ((p == j ? rf * rn : 0.0) + r[j] * f[p] * rn + r[j] * n[p] * rf) +
                             2 * xi2 *
                                 (f[p] * n[j] + n[p] * f[j] + (p == j ? fn : 0.0) -
                                  2 * xi2 * r[p] * (f[j] * rn + n[j] * rf + r[j] * fn)));
        } else if (KIND == SE2P_DLP_PRESSURE_GRAD) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const double a = r[p] * fn + f[p] * rn + n[p] * rf;
                acc[p] += e * (2 * xi2 * (a - 2 * xi2 * r[p] * rf * rn) * ir +
                               2 * (a - 4 * xi2 * r[p] * rn * rf) * ir2 -
                               8 * r[p] * rn * rf * ir2 * ir);
            }
        } else if (KIND == SE2P_DLP_VORTICITY) {
            const double v = rf * (n[0] * r2 - n[1] * r1) + rn * (f[0] * r2 - f[1] * r1);
            acc[0] += e * (-(1 + xi2 * rr) * v * ir2 + xi2 * xi2 * v);
        } else {  // SE2P_DLP_STRESS
            const double C = (8 * xi2 * xi2 + (16 * xi2 + 16 * ir) * ir) * ir;
            const double B = 2 * (1 + xi2 * rr) * ir2;
            const double P = (fn - 2 * xi2 * rf * rn) * ir - 2 * rf * rn * ir2;
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const double d = p == j ? 1.0 : 0.0;
                    const double G =
                        r[p] * r[j] * rf * rn * C -
                        B * (f[p] * r[j] * rn + n[p] * r[j] * rf + r[p] * f[j] * rn +
                             r[p] * n[j] * rf + 2 * d * rf * rn) +
                        2 * xi2 *
                            (d * fn + n[p] * f[j] + f[p] * n[j] -
                             xi2 * (2 * r[p] * r[j] * fn + r[p] * f[j] * rn + r[p] * n[j] * rf +
                                    f[p] * r[j] * rn + n[p] * r[j] * rf));
                    acc[2 * p + j] += 2 * e * (d * P + G);
                }
        }
    }
}

__device__ __forceinline__ int wrapi(int q, int n) { return ((q % n) + n) % n; }

// Bits [lo, hi) of a 64-bit word (0 <= lo <= hi <= 64).
__device__ __forceinline__ unsigned long long bit_range(int lo, int hi)
{
    const unsigned long long m = hi - lo >= 64 ? ~0ull : (1ull << (hi - lo)) - 1;
    return lo < hi ? m << lo : 0ull;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// First staged index at or after window position g in round r of nround.
__device__ __forceinline__ int round_index(int g, int r, int nround)
{
    return nround == 1 ? g : (g - r + nround - 1) / nround;
}

// UNI: blocks of one cell (bs = 1), where every lane has the same stencil and
// the candidate bookkeeping is wave-uniform (scalar registers, LDS broadcast).
//
// SYM (velocity sums, sources = targets in one y order): each pair is evaluated
// once, by the point the other lies "after": the source's image in the stencil
// row above, or in the same row with a larger in-cell y key (ties: the image in
// the column to the right; in the own cell: later in the sorted order).  This
// order is antisymmetric for every pair of images, so each pair is taken once.
// A target's candidates are then, per stencil column, the part of the row-y
// cell after its key (own cell: after itself) and the row above up to y +
// trim_w.  The window holds rows 0 .. bh only; the term for the source is
// accumulated in LDS (racc) and added after each round, the target's own sum
// after the last round, with atomics into sacc, which follows the sorted order
// (so a cell's atomics share cache lines); the gather adds it to u.  Both
// kernels are even (Stokeslet) or odd (stresslet) in r, so a pair's term for
// the source is the target's expression with the target's density.
// For the stresslet the staged sources carry their normals, in a smaller window.
template <int KIND, int TIER, bool UNI, bool SYM = false>
__global__ __launch_bounds__(kBlkThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_real_stencil(
    const SysDesc* __restrict__ sys, const Job* __restrict__ jobs, double* __restrict__ u_base)
{
    static_assert(!SYM || ((KIND == SE2P_SLP || KIND == SE2P_DLP) && !UNI),
                  "pair-symmetric sums: velocity blocks");
    using Rec = typename SrcOf<KIND>::T;
    // with the reflected sums and keys, 4 workgroups per CU still fit in LDS
    constexpr int CAP = SYM ? (KIND == SE2P_SLP ? 768 : 576) : SrcOf<KIND>::cap;
    // Stokeslet records, and the stresslet's in pair-symmetric sums, are staged
    // as arrays of 16-byte entries -- positions, densities (and normals) -- so a
    // candidate test reads 16-byte entries at a 16-byte stride
    constexpr bool SOA = KIND == SE2P_SLP || SYM;
    constexpr int NW = KIND == SE2P_SLP ? 2 : 3;  // staged 16-byte entries per source
    constexpr int BUF = SOA ? NW * CAP : CAP * (int)(sizeof(Rec) / 16);
    __shared__ __align__(64) double2 raw[BUF];
    Rec* const buf = reinterpret_cast<Rec*>(raw);
    double2* const bxy = raw;
    double2* const bfd = bxy + CAP;
    double2* const bnr = bxy + 2 * CAP;  // stresslet normals (pair-symmetric sums)
    auto rec_xy = [&](int i) -> double2 {
        if constexpr (SOA)
            return bxy[i];
        else
            return *reinterpret_cast<const double2*>(&buf[i]);
    };
    __shared__ double2 racc[SYM ? CAP : 1];
    __shared__ int rkey[SYM ? CAP : 1];
    __shared__ int e_src[kMaxEntries], e_off[kMaxEntries + 1];
    __shared__ double e_ox[kMaxEntries], e_oy[kMaxEntries];
    __shared__ int c_t0[kMaxBlockSide * kMaxBlockSide], c_off[kMaxBlockSide * kMaxBlockSide + 1];

    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int nsx = d.nsx, nsy = d.nsy, bs = d.bs, nbx = d.nbx;
    const double Lx = d.Lx, Ly = d.Ly, xi2 = d.xi2, cut2 = d.cut2, self = d.self;
    const glob<const int>* __restrict__ tstart = gp(d.t_start);
    const glob<const int>* __restrict__ tperm = gp(d.t_perm);
    const glob<const double2>* __restrict__ tsorted = gp(d.t_sorted);
    const glob<const int>* __restrict__ sstart = gp(d.s_start);
    const glob<const Rec>* __restrict__ ssorted = gp(reinterpret_cast<const Rec*>(d.s_rec));
    const glob<const int>* __restrict__ ykey = gp(d.ykey);
    glob<double>* __restrict__ sacc = gp(reinterpret_cast<double*>(d.sacc));
    double* __restrict__ u = u_base + d.u_off;
    const int tid = threadIdx.x;
    // job: block blk, and the block's targets [part, part + 1) x kBlkThreads
    const int nblk = nbx * ((nsy + bs - 1) / bs);
    const int blk = jb.y % nblk, part = jb.y / nblk;
    const int bx0 = (blk % nbx) * bs, by0 = (blk / nbx) * bs;
    const int bw = min(bs, nsx - bx0), bh = min(bs, nsy - by0);
    // window column height; entry (ux, uy) is (ux+1) H + (uy+1), or (ux+1) H + uy
    // for SYM, whose window starts at row 0
    const int H = SYM ? bh + 1 : bh + 2;
    const int nent = (bw + 2) * H, ncell = bw * bh;
    static_assert(kMaxEntries <= 64, "the window entries are set up by one wave");
    if (tid < 64) {
        // wave 0: one lane per window entry and per target cell, then the
        // offsets as inclusive scans of the counts across the wave
        int ecnt = 0, ccnt = 0;
        if (tid < nent) {
            const int ux = tid / H - 1, uy = tid % H - (SYM ? 0 : 1);
            const int qx = bx0 + ux, qy = by0 + uy;
            const int wx = wrapi(qx, nsx), wy = wrapi(qy, nsy);
            const int cell = wy * nsx + wx;
            e_src[tid] = sstart[cell];
            ecnt = sstart[cell + 1] - sstart[cell];
            e_ox[tid] = (Lx * (qx - wx)) / nsx;
            e_oy[tid] = (Ly * (qy - wy)) / nsy;
        }
        if (tid < ncell) {
            const int cx = tid / bh, cy = tid % bh;
            const int cell = (by0 + cy) * nsx + (bx0 + cx);
            c_t0[tid] = tstart[cell];
            ccnt = tstart[cell + 1] - tstart[cell];
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int a = __shfl_up(ecnt, o), b = __shfl_up(ccnt, o);
            if (tid >= o) {
                ecnt += a;
                ccnt += b;
            }
        }
        if (tid < nent) e_off[tid + 1] = ecnt;
        if (tid < ncell) c_off[tid + 1] = ccnt;
        if (tid == 0) {
            e_off[0] = 0;
            c_off[0] = 0;
        }
    }
    __syncthreads();
    // window entry of window position g (the last entry starting at or before g)
    auto entry_of = [&](int g) {
        int lo = 0, hi = nent;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (e_off[mid] <= g)
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    };
    // block-uniform: kept in scalar registers
    const int total = __builtin_amdgcn_readfirstlane(e_off[nent]);
    const int ntar = __builtin_amdgcn_readfirstlane(c_off[ncell]);
    const double ixi2 = 1.0 / xi2;
    const int nround = (total + CAP - 1) / CAP;
    const double trim_w = d.trim_w;
    const double far32_r2 = d.far32_r2;
    const double self_r2 = d.self_r2;
    const bool trim = !UNI && trim_w > 0 && d.unsorted && *d.unsorted == 0;

    // the job list holds one job per 256 targets of the block (at least one)
    const int nparts = max(1, (ntar + kBlkThreads - 1) / kBlkThreads);
    for (int tb = part * kBlkThreads; tb < ntar; tb += nparts * kBlkThreads) {
        const int i = tb + tid;
        const bool act = i < ntar;
        int ti = 0, cx = 0, cy = 0, kin = 0;
        double2 pt = make_double2(0.0, 0.0);
        double fi1 = 0, fi2 = 0;  // SYM: the target's density (and normal) and y key
        double ni1 = 0, ni2 = 0;
        int qyi = 0;
        if (act) {
            int c = 0;
            while (c + 1 < ncell && c_off[c + 1] <= i) ++c;
            kin = i - c_off[c];
            ti = c_t0[c] + kin;
            cx = c / bh;
            cy = c % bh;
            pt = tsorted[ti];
            if constexpr (SYM && KIND == SE2P_SLP) {
                const double4 q = ssorted[ti];
                fi1 = q.z;
                fi2 = q.w;
            } else if constexpr (SYM) {
                const DlpRec q = ssorted[ti];
                fi1 = q.f1;
                fi2 = q.f2;
                ni1 = q.n1;
                ni2 = q.n2;
            }
            if constexpr (SYM) qyi = ykey[ti];
        }
        constexpr int NC = out_comps(KIND);
        double a1 = 0, a2 = 0;                // velocity and pressure kinds
        double acc[4] = {0.0, 0.0, 0.0, 0.0};  // derived kinds
        for (int r = 0; r < nround; ++r) {
            const int rn = (total - r + nround - 1) / nround;
            __syncthreads();
            for (int k = tid; k < rn; k += kBlkThreads) {
                const int g = r + k * nround;
                const int e = entry_of(g);
                const int sk = e_src[e] + (g - e_off[e]);
                Rec rec = ssorted[sk];
                rec.x += e_ox[e];
                rec.y += e_oy[e];
                if constexpr (SOA && KIND == SE2P_SLP) {
                    bxy[k] = make_double2(rec.x, rec.y);
                    bfd[k] = make_double2(rec.z, rec.w);
                } else if constexpr (SOA) {
                    const DlpRec& q = reinterpret_cast<const DlpRec&>(rec);
                    bxy[k] = make_double2(q.x, q.y);
                    bfd[k] = make_double2(q.f1, q.f2);
                    bnr[k] = make_double2(q.n1, q.n2);
                } else {
                    buf[k] = rec;
                }
                if constexpr (SYM) {
                    rkey[k] = ykey[sk];
                    racc[k] = make_double2(0.0, 0.0);
                }
            }
            __syncthreads();
            // this lane's candidates: staged runs [s_j, s_j + L_j), j = 0, 1, 2 (the
            // stencil columns: window entries (cx + j) H + cy .. + 2), numbered
            // 0 .. n-1 in that order; the own cell (entry (cx + 1) H + cy + 1) is [o0, o1)
            const int e0 = cx * H + cy;
            int s0, s1, s2, L0, L01, n, o0, o1;
            if constexpr (SYM) {
                // column j: its row-y cell is entry e0 + j H, the row above e0 + j H + 1
                const double yhi = pt.y + trim_w;
                auto first_at = [&](int a, int b, double v) {  // first of [a, b) with y >= v
                    while (a < b) {
                        const int mid = (a + b) >> 1;
                        if (rec_xy(mid).y < v)
                            a = mid + 1;
                        else
                            b = mid;
                    }
                    return a;
                };
                auto first_key = [&](int a, int b, int v) {  // first of [a, b) with key >= v
                    while (a < b) {
                        const int mid = (a + b) >> 1;
                        if (rkey[mid] < v)
                            a = mid + 1;
                        else
                            b = mid;
                    }
                    return a;
                };
                const int eo = e0 + H;
                const int w1 = round_index(e_off[e0 + 1], r, nround);
                const int m1 = round_index(e_off[eo + 1], r, nround);
                const int q1 = round_index(e_off[e0 + 2 * H + 1], r, nround);
                s0 = first_key(round_index(e_off[e0], r, nround), w1, qyi + 1);
                s1 = round_index(e_off[eo] + kin + 1, r, nround);
                s2 = first_key(round_index(e_off[e0 + 2 * H], r, nround), q1, qyi);
                const int b0 = first_at(w1, round_index(e_off[e0 + 2], r, nround), yhi);
                const int b1 = first_at(m1, round_index(e_off[eo + 2], r, nround), yhi);
                const int b2 = first_at(q1, round_index(e_off[e0 + 2 * H + 2], r, nround), yhi);
                L0 = b0 - s0;
                L01 = L0 + (b1 - s1);
                n = act ? L01 + (b2 - s2) : 0;
                o0 = s1;  // the own cell after the target
                o1 = m1;
            } else {
                s0 = round_index(e_off[e0], r, nround);
                s1 = round_index(e_off[e0 + H], r, nround);
                s2 = round_index(e_off[e0 + 2 * H], r, nround);
                L0 = round_index(e_off[e0 + 3], r, nround) - s0;
                L01 = L0 + round_index(e_off[e0 + H + 3], r, nround) - s1;
                // a lane without a target has no candidates, so a wave without
                // targets (a block's second pass past 256 targets, or its last wave)
                // skips the candidate tests
                n = act ? L01 + round_index(e_off[e0 + 2 * H + 3], r, nround) - s2 : 0;
                o0 = round_index(e_off[e0 + H + 1], r, nround);
                o1 = round_index(e_off[e0 + H + 2], r, nround);
                if (trim && act) {
                    // Each run is ordered by y (sources by y inside their cells, the
                    // cells of a column stacked, the periodic shift staged with y):
                    // keep the part within trim_w of the target's y, found by binary
                    // search.  Every pair the sums take lies in it.
                    const double ylo = pt.y - trim_w, yhi = pt.y + trim_w;
                    auto first_at = [&](int a, int b, double v) {  // first of [a, b) with y >= v
                        while (a < b) {
                            const int mid = (a + b) >> 1;
                            if (rec_xy(mid).y < v)
                                a = mid + 1;
                            else
                                b = mid;
                        }
                        return a;
                    };
                    const int e0b = s0 + L0, e1b = s1 + (L01 - L0), e2b = s2 + (n - L01);
                    const int a0 = first_at(s0, e0b, ylo), b0 = first_at(a0, e0b, yhi);
                    const int a1 = first_at(s1, e1b, ylo), b1 = first_at(a1, e1b, yhi);
                    const int a2 = first_at(s2, e2b, ylo), b2 = first_at(a2, e2b, yhi);
                    s0 = a0;
                    s1 = a1;
                    s2 = a2;
                    L0 = b0 - a0;
                    L01 = L0 + (b1 - a1);
                    n = L01 + (b2 - a2);
                    o0 = min(max(o0, a1), b1);
                    o1 = min(max(o1, o0), b1);
                }
            }
            // candidate pos -> staged index: pos + s0, shifted at the two run ends
            int j1 = s1 - L0 - s0, j2 = s2 - L01 - (s1 - L0);
            if (UNI) {
                // one cell: the window is the stencil, its three runs follow each
                // other from 0, and candidate pos is staged source pos
                s0 = j1 = j2 = 0;
                L0 = __builtin_amdgcn_readfirstlane(L0);
                L01 = __builtin_amdgcn_readfirstlane(L01);
                n = __builtin_amdgcn_readfirstlane(n);
                o0 = __builtin_amdgcn_readfirstlane(o0);
                o1 = __builtin_amdgcn_readfirstlane(o1);
            }
            auto staged = [&](int pos) {
                return pos + s0 + (pos >= L0 ? j1 : 0) + (pos >= L01 ? j2 : 0);
            };
            // Chunk c takes candidates c, c + nch, c + 2 nch, ... (nch the same
            // for the whole wave), so each chunk samples all three runs and the
            // hit counts of the lanes in a chunk stay close to their mean.
            int nmax = n;
            if (!UNI) {
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
            }
            const int nch = (nmax + kChunk - 1) / kChunk;
            const int kmax = nch ? (nmax + nch - 1) / nch : 0;  // <= kChunk
            // own cell and valid candidates as candidate-position ranges: the own
            // cell is a sub-run of run 1, which holds positions [L0, L01)
            const int po0 = o0 - (s0 + j1), po1 = o1 - (s0 + j1);
            for (int c = 0; c < nch; ++c) {
                // chunk candidates kg = 0 .. kc-1 at positions c + kg nch
                const int kc = UNI ? (n - c + nch - 1) / nch : kmax;
                const int kvalid = act ? (n - c + nch - 1) / nch : 0;
                const int kown0 = nch == 1 ? po0 - c : (po0 - c + nch - 1) / nch;
                const int kown1 = nch == 1 ? po1 - c : (po1 - c + nch - 1) / nch;
                // per word: comparison bits shifted in (candidate k of the word, k <
                // kw[w], lands on bit kw[w] - 1 - k), then combined with the range masks;
                // kw[w] is rounded up to a multiple of 8 and the candidates past the
                // chunk's end are masked off as not valid
                unsigned long long fm[kWords], nm[kWords], g32[kWords];
                int kw[kWords];
#pragma unroll
                for (int w = 0; w < kWords; ++w) {
                    kw[w] = (min(64, max(0, kc - 64 * w)) + 7) & ~7;
                    unsigned long long cm = 0, zm = 0;  // r^2 < cut^2; SLP r^2 < 1/xi^2, DLP r^2 != 0
                    unsigned long long gm = 0;  // fast tier: r^2 >= far32_r2 (FP32 far pairs)
                    for (int k0 = 0; k0 < kw[w]; k0 += 8) {
                        double2 sp[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const int pos = c + (64 * w + k0 + j) * nch;
                            sp[j] = rec_xy(min(staged(pos), CAP - 1));
                        }
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const double dx = pt.x - sp[j].x, dy = pt.y - sp[j].y;
                            const double r2 = dx * dx + dy * dy;
                            cm = (cm << 1) | (r2 < cut2 ? 1ull : 0ull);
                            if (KIND == SE2P_SLP && TIER == kE1Fast)
                                gm = (gm << 1) | (r2 >= far32_r2 ? 1ull : 0ull);
                            if (KIND == SE2P_SLP)
                                zm = (zm << 1) | (r2 < ixi2 ? 1ull : 0ull);
                            else if (KIND == SE2P_DLP_VORTICITY)  // own-cell pairs r^2 < 1e-13 skipped
                                zm = (zm << 1) | (r2 >= 1e-13 ? 1ull : 0ull);
                            else  // coincident points of the own cell are skipped
                                zm = (zm << 1) | (r2 != 0 ? 1ull : 0ull);
                        }
                    }
                    const int K = kw[w];
                    const unsigned long long valid = bit_range(K - clampi(kvalid - 64 * w, 0, K), K);
                    const unsigned long long own = bit_range(K - clampi(kown1 - 64 * w, 0, K),
                                                             K - clampi(kown0 - 64 * w, 0, K));
                    if (KIND == SE2P_SLP) {
                        // every pair with xi^2 r^2 < 1 is inside the cutoff
                        const unsigned long long hit = (cm | own) & valid;
                        fm[w] = hit & ~zm;
                        nm[w] = hit & zm;
                    } else {
                        fm[w] = ((cm & ~own) | (own & zm)) & valid;
                        nm[w] = 0;
                    }
                    g32[w] = fm[w] & gm;
                }
                // candidate position of the lowest set bit of (lo, hi), which is cleared
                // candidate position of bit 0 of each word
                const int pos_w0 = c + (kw[0] - 1) * nch, pos_w1 = c + (64 + kw[1] - 1) * nch;
                auto next = [&](unsigned long long& lo, unsigned long long& hi) {
                    const bool z = lo == 0;
                    const unsigned long long m = z ? hi : lo;
                    const int bit = __builtin_ctzll(m);
                    const unsigned long long m2 = m & (m - 1);
                    lo = z ? lo : m2;
                    hi = z ? m2 : hi;
                    return (z ? pos_w1 : pos_w0) - bit * nch;
                };
                // Far pairs of the chunk: FP64 terms into (a1, a2), FP32 terms (the FP32
                // tier; in the fast tier the pairs with r^2 >= far32_r2) into (c1, c2).
                float c1 = 0.f, c2 = 0.f;
                auto far_loop = [&](unsigned long long lo, unsigned long long hi, auto f32tag) {
                constexpr bool F32 = decltype(f32tag)::value;
                if constexpr (SYM) {
                    // the next hit's record is read before this hit's LDS atomics
                    // are issued, so waiting for it does not wait for them
                    bool more = (lo | hi) != 0;
                    int idx = more ? staged(next(lo, hi)) : 0;
                    double2 q = bxy[idx], g = bfd[idx];
                    double2 m = KIND == SE2P_DLP ? bnr[idx] : make_double2(0.0, 0.0);
                    while (more) {
                        more = (lo | hi) != 0;
                        const int nidx = more ? staged(next(lo, hi)) : idx;
                        const double2 qn = bxy[nidx], gn = bfd[nidx];
                        const double2 mn = KIND == SE2P_DLP ? bnr[nidx] : m;
                        const double dx = pt.x - q.x, dy = pt.y - q.y;
                        const double r2 = dx * dx + dy * dy;
                        if constexpr (KIND == SE2P_DLP && F32) {
                            float b1, b2;
                            dlp_pair2_f32((float)dx, (float)dy, (float)r2, (float)xi2,
                                          (float)g.x, (float)g.y, (float)m.x, (float)m.y,
                                          (float)fi1, (float)fi2, (float)ni1, (float)ni2, c1, c2,
                                          b1, b2);
                            atomicAdd(&racc[idx].x, (double)b1);
                            atomicAdd(&racc[idx].y, (double)b2);
                        } else if constexpr (KIND == SE2P_DLP) {
                            double b1, b2;
                            dlp_pair2(dx, dy, r2, xi2, g.x, g.y, m.x, m.y, fi1, fi2, ni1, ni2, a1,
                                      a2, b1, b2);
                            atomicAdd(&racc[idx].x, b1);
                            atomicAdd(&racc[idx].y, b2);
                        } else if constexpr (F32) {
                            float b1, b2;
                            slp_pair2_f32((float)dx, (float)dy, (float)r2, (float)xi2,
                                          (float)g.x, (float)g.y, (float)fi1, (float)fi2, c1, c2,
                                          b1, b2);
                            atomicAdd(&racc[idx].x, (double)b1);
                            atomicAdd(&racc[idx].y, (double)b2);
                        } else {
                            double b1, b2;
                            slp_pair2<fp64_tier(TIER), false>(dx, dy, r2, xi2, g.x, g.y, fi1,
                                                              fi2, a1, a2, b1, b2);
                            atomicAdd(&racc[idx].x, b1);
                            atomicAdd(&racc[idx].y, b2);
                        }
                        idx = nidx;
                        q = qn;
                        g = gn;
                        m = mn;
                    }
                    lo = hi = 0;
                }
                while (lo | hi) {
                    const int idx = staged(next(lo, hi));
                    const Rec& s = buf[idx];
                    const double2 sxy = rec_xy(idx);
                    const double dx = pt.x - sxy.x, dy = pt.y - sxy.y;
                    const double r2 = dx * dx + dy * dy;
                    if (KIND == SE2P_SLP) {
                        const double2 g = bfd[idx];
                        if constexpr (F32)
                            slp_pair_f32((float)dx, (float)dy, (float)r2, (float)xi2, (float)g.x,
                                         (float)g.y, c1, c2);
                        else
                            slp_pair<fp64_tier(TIER), false>(dx, dy, r2, xi2, g.x, g.y, a1, a2);
                    } else if (KIND == SE2P_DLP) {
                        if constexpr (F32)
                            dlp_pair_f32((float)dx, (float)dy, (float)r2, (float)xi2,
                                         reinterpret_cast<const DlpRec&>(s), c1, c2);
                        else
                            dlp_pair(dx, dy, r2, xi2, reinterpret_cast<const DlpRec&>(s), a1, a2);
                    } else if (KIND == SE2P_SLP_PRESSURE) {
                        const double4& q = reinterpret_cast<const double4&>(s);
                        a1 += (dx * q.z + dy * q.w) * exp(-xi2 * r2) * rcp64(r2);
                    } else if (KIND == SE2P_DLP_PRESSURE) {
                        a1 += dlp_pressure_pair(dx, dy, r2, xi2, reinterpret_cast<const DlpRec&>(s));
                    } else {
                        derived_pair<KIND>(dx, dy, r2, xi2, reinterpret_cast<const double*>(&s), acc);
                    }
                }
                };
                if constexpr (KIND == SE2P_SLP && TIER == kE1Fast) {
                    far_loop(fm[0] & ~g32[0], fm[1] & ~g32[1], std::false_type{});
                    far_loop(g32[0], g32[1], std::true_type{});
                } else {
                    far_loop(fm[0], fm[1], std::integral_constant<bool, TIER == kF32>{});
                }
                if constexpr (TIER == kF32 || (TIER == kE1Fast && KIND == SE2P_SLP)) {
                    // the chunk's FP32 partial sums join the FP64 sums exactly
                    a1 += (double)c1;
                    a2 += (double)c2;
                }
                if (KIND == SE2P_SLP) {
                    unsigned long long lo = nm[0], hi = nm[1];
                    while (lo | hi) {
                        const int idx = staged(next(lo, hi));
                        const double2 q = rec_xy(idx), g = bfd[idx];
                        const double dx = pt.x - q.x, dy = pt.y - q.y;
                        const double r2 = dx * dx + dy * dy;
                        if (idx >= o0 && idx < o1 && fabs(r2) < self_r2) {
                            a1 += self * g.x;
                            a2 += self * g.y;
                            if (SYM) {
                                atomicAdd(&racc[idx].x, self * fi1);
                                atomicAdd(&racc[idx].y, self * fi2);
                            }
                        } else if (SYM) {
                            double b1, b2;
                            slp_pair2<fp64_tier(TIER), true>(dx, dy, r2, xi2, g.x, g.y, fi1,
                                                             fi2, a1, a2, b1, b2);
                            atomicAdd(&racc[idx].x, b1);
                            atomicAdd(&racc[idx].y, b2);
                        } else {
                            slp_pair<fp64_tier(TIER), true>(dx, dy, r2, xi2, g.x, g.y, a1, a2);
                        }
                    }
                }
            }
            if constexpr (SYM) {
                // the sources' sums of this round, before the next round restages
                __syncthreads();
                for (int k = tid; k < rn; k += kBlkThreads) {
                    const double2 v = racc[k];
                    if (v.x != 0.0 || v.y != 0.0) {
                        const int g = r + k * nround;
                        const int e = entry_of(g);
                        glob<double>* a = sacc + 2 * (e_src[e] + (g - e_off[e]));
                        __hip_atomic_fetch_add(a, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_fetch_add(a + 1, v.y, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        }
        if (act) {
            const int o = tperm[ti];
            constexpr double dv = out_div(KIND);
            if (SYM) {
                // the Stokeslet's pair of the target with itself: the self term
                const double s1 = KIND == SE2P_SLP ? a1 + self * fi1 : a1;
                const double s2 = KIND == SE2P_SLP ? a2 + self * fi2 : a2;
                glob<double>* a = sacc + 2 * ti;
                __hip_atomic_fetch_add(a, s1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(a + 1, s2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if (KIND >= SE2P_SLP_GRADIENT) {
#pragma unroll
                for (int c = 0; c < NC; ++c) u[NC * o + c] = acc[c] / dv;
            } else if (NC == 1) {
                u[o] = a1 / dv;
            } else {
                u[2 * o] = a1 / dv;
                u[2 * o + 1] = a2 / dv;
            }
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

__global__ void k_inverse_perm(const int* __restrict__ perm, int n, int* __restrict__ pos)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) pos[perm[k]] = k;
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
    if (rec_layout(kind) == 0)
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

int real_block_side(int nsx, int nsy, int mean_targets)
{
    // blocks of bs x bs cells holding about 256 targets
    int bs = (int)std::floor(std::sqrt(256.0 / (mean_targets > 1 ? mean_targets : 1)));
    bs = std::max(1, std::min(bs, kMaxBlockSide));
    return std::min(bs, std::min(nsx, nsy));
}

int launch_real(hipStream_t s, int kind, int tier, int mode, const SysDesc* sys, const Job* jobs,
                int njobs, double* u)
{
    if (njobs <= 0) return SE2P_OK;
    dim3 gr(njobs), bl(kBlkThreads);
    const bool uni = mode == RM_ONE_CELL;
    if (mode == RM_SYM) {
        if (kind == SE2P_SLP) {
            if (tier == kE1Full)
                hipLaunchKernelGGL((k_real_stencil<SE2P_SLP, kE1Full, false, true>), gr, bl, 0, s,
                                   sys, jobs, u);
            else if (tier == kF32)
                hipLaunchKernelGGL((k_real_stencil<SE2P_SLP, kF32, false, true>), gr, bl, 0, s,
                                   sys, jobs, u);
            else
                hipLaunchKernelGGL((k_real_stencil<SE2P_SLP, kE1Fast, false, true>), gr, bl, 0, s,
                                   sys, jobs, u);
        } else if (kind == SE2P_DLP) {
            // the stresslet has no E1: its FP64 tiers are the same sums
            if (tier == kF32)
                hipLaunchKernelGGL((k_real_stencil<SE2P_DLP, kF32, false, true>), gr, bl, 0, s,
                                   sys, jobs, u);
            else
                hipLaunchKernelGGL((k_real_stencil<SE2P_DLP, kE1Full, false, true>), gr, bl, 0, s,
                                   sys, jobs, u);
        } else {
            return set_error(SE2P_ERR_UNSUPPORTED, "pair-symmetric sums are velocity sums only");
        }
        SE2P_HIP_CHECK(hipGetLastError());
        return SE2P_OK;
    }
#define SE2P_LAUNCH_REAL(K, T)                                                                  \
    do {                                                                                        \
        if (uni)                                                                                \
            hipLaunchKernelGGL((k_real_stencil<K, T, true>), gr, bl, 0, s, sys, jobs, u);        \
        else                                                                                    \
            hipLaunchKernelGGL((k_real_stencil<K, T, false>), gr, bl, 0, s, sys, jobs, u);       \
    } while (0)
    switch (kind) {
    case SE2P_SLP:
        if (tier == kE1Full)
            SE2P_LAUNCH_REAL(SE2P_SLP, kE1Full);
        else if (tier == kF32)
            SE2P_LAUNCH_REAL(SE2P_SLP, kF32);
        else
            SE2P_LAUNCH_REAL(SE2P_SLP, kE1Fast);
        break;
    case SE2P_DLP:
        if (tier == kF32)
            SE2P_LAUNCH_REAL(SE2P_DLP, kF32);
        else
            SE2P_LAUNCH_REAL(SE2P_DLP, kE1Full);
        break;
    case SE2P_SLP_PRESSURE: SE2P_LAUNCH_REAL(SE2P_SLP_PRESSURE, kE1Full); break;
    case SE2P_DLP_PRESSURE: SE2P_LAUNCH_REAL(SE2P_DLP_PRESSURE, kE1Full); break;
    case SE2P_SLP_GRADIENT: SE2P_LAUNCH_REAL(SE2P_SLP_GRADIENT, kE1Full); break;
    case SE2P_SLP_PRESSURE_GRAD: SE2P_LAUNCH_REAL(SE2P_SLP_PRESSURE_GRAD, kE1Full); break;
    case SE2P_SLP_VORTICITY: SE2P_LAUNCH_REAL(SE2P_SLP_VORTICITY, kE1Full); break;
    case SE2P_SLP_STRESS: SE2P_LAUNCH_REAL(SE2P_SLP_STRESS, kE1Full); break;
    case SE2P_DLP_GRADIENT: SE2P_LAUNCH_REAL(SE2P_DLP_GRADIENT, kE1Full); break;
    case SE2P_DLP_PRESSURE_GRAD: SE2P_LAUNCH_REAL(SE2P_DLP_PRESSURE_GRAD, kE1Full); break;
    case SE2P_DLP_VORTICITY: SE2P_LAUNCH_REAL(SE2P_DLP_VORTICITY, kE1Full); break;
    default: SE2P_LAUNCH_REAL(SE2P_DLP_STRESS, kE1Full); break;
    }
#undef SE2P_LAUNCH_REAL
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_inverse_perm(hipStream_t s, const int* perm, int n, int* pos)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_inverse_perm, dim3((n + 255) / 256), dim3(256), 0, s, perm, n, pos);
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
