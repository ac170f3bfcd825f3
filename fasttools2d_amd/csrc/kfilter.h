// Spectral filters of the spectral Ewald k-space stage, shared by the filter
// kernel and the fused x-line transform.  Device-only inline functions.
#pragma once

#include "se2p_internal.h"

namespace se2p {
namespace kf {

// Wavenumber of grid index j: indices up to M/2 non-negative
// (mex_stokes_slp_kspace.cpp:115-128,141-150).
__device__ __forceinline__ double wavenum(int j, int M, double L)
{
    return 2.0 * kPi / L * (double)(j <= M / 2 ? j : j - M);
}

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }

// Real-field spectra packed as Z = A + i B: A(k) = (Z(k) + conj Z(-k))/2,
// B(k) = (Z(k) - conj Z(-k))/(2i).
__device__ __forceinline__ void unpack(double2 zk, double2 zm, cplx& a, cplx& b)
{
    a = {0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y)};
    b = {0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x)};
}

// SLP filter at (k1, k2) applied to q = (q1, q2) (mex_stokes_slp_kspace.cpp:130-139):
//   (K^2 q - k (k.q)) (1/K^4 + 1/(4 xi^2 K^2)) exp(-(1 - eta) K^2/(4 xi^2))
// The scalar factor e is passed in: the periodic one above, or the free-space
// one of free_factor.
__device__ __forceinline__ void slp_apply_e(double k1, double k2, double e, cplx q1, cplx q2,
                                            cplx& g1, cplx& g2)
{
    const double K2 = k1 * k1 + k2 * k2;
    const cplx kd = cadd(cscale(q1, k1), cscale(q2, k2));
    g1 = cscale(csub(cscale(q1, K2), cscale(kd, k1)), e);
    g2 = cscale(csub(cscale(q2, K2), cscale(kd, k2)), e);
}

__device__ __forceinline__ void slp_apply(double k1, double k2, double a, double eta, cplx q1,
                                          cplx q2, cplx& g1, cplx& g2)
{
    const double K2 = k1 * k1 + k2 * k2;
    const double e = (1.0 / (K2 * K2) + a / K2) * exp(-(1 - eta) * a * K2);
    slp_apply_e(k1, k2, e, q1, q2, g1, g2);
}

// Fourier transform of the biharmonic kernel B(r) = r^2 (ln r - 1)/(8 pi)
// (Delta^2 B = delta, so its transform is 1/K^4) less the quadratic that
// matches its value and slope at R, truncated to r < R:
//   B~(r) = B(r) - B(R) - beta (r^2 - R^2),  beta = B'(R)/(2 R) = (2 ln R - 1)/(16 pi),
//   B~_R(K) = (1 - J0(K R))/K^4 - R J1(K R)/(2 K^3)
// (the transform of B's truncation has R^3 J1/K and R^2 J2/K^2 terms that the
// quadratic's cancels).  B~ and its slope vanish at R, so the tail of B~_R is
// ~R^{1/2} K^{-7/2}, (R K)^{1/2} times the periodic 1/K^4 instead of the
// (R K)^{5/2} a plain truncation leaves, and rounding errors are not
// amplified at large K.  (d_i d_j - delta_ij Lap) of the quadratic is the
// constant -2 beta delta_ij, which the sum's constant term takes back.  For
// x = K R < 2 the two terms cancel, and the power series is used instead:
//   B~_R(K) = R^4/16 sum_{m>=2} (-1)^m (m - 1) (x/2)^{2m-4}/(m!)^2
// (16 terms: the last is below 1e-30 of the first at x = 2).
__device__ __forceinline__ double bhat_trunc(double K, double R)
{
    const double x = K * R;
    if (x < 2.0) {
        const double t = 0.25 * x * x;
        double s = 0.0, c = 1.0, fac = 2.0;  // fac = m!
#pragma unroll
        for (int m = 2; m < 18; ++m) {
            s += ((m & 1) ? -1.0 : 1.0) * (m - 1) * c / (fac * fac);
            c *= t;
            fac *= m + 1;
        }
        const double R2 = R * R;
        return R2 * R2 / 16.0 * s;
    }
    const double iK = 1.0 / K, iK2 = iK * iK;
    return (1.0 - j0(x)) * iK2 * iK2 - 0.5 * R * j1(x) * iK2 * iK;
}

// Free-space Stokeslet factor: the periodic 1/K^4 (1 + K^2/(4 xi^2)) becomes
// B~_R(K) (1 + K^2/(4 xi^2)), so that the k-space kernel is the truncated
// biharmonic's, screened: it equals the free-space k-space kernel at distances
// below R - delta and vanishes beyond R + delta, and sampled on the periodic
// grid of a box wider than the points' extent plus R + delta it gives the
// free-space sum without images.
__device__ __forceinline__ double free_factor(double K2, double a, double eta, double R)
{
    return bhat_trunc(sqrt(K2), R) * (1.0 + a * K2) * exp(-(1 - eta) * a * K2);
}

// DLP bracket at (k1, k2) (mex_stokes_dlp_kspace.cpp:164-191): with Qs = Q12 + Q21,
//   v_i = 2 k_i Q_ii + k_j Qs + k_i (Q11 + Q22) - 2 k_i (k.Q.k)/K^2,
// scaled by e = exp(-(1-eta) K^2/(4 xi^2)) (1 + K^2/(4 xi^2))/K^2.  The filter
// output is i e v.
__device__ __forceinline__ void dlp_bracket(double k1, double k2, double a, double eta,
                                            cplx q11, cplx qs, cplx q22, cplx& v1, cplx& v2)
{
    const double K2 = k1 * k1 + k2 * k2;
    const double e = exp(-(1 - eta) * a * K2) * (1 + a * K2) / K2;
    const cplx tr = cadd(q11, q22);
    const cplx m = cscale(cadd(cadd(cscale(q11, k1 * k1), cscale(qs, k1 * k2)), cscale(q22, k2 * k2)),
                          1.0 / K2);
    v1 = cscale(csub(cadd(cadd(cscale(q11, 2 * k1), cscale(qs, k2)), cscale(tr, k1)), cscale(m, 2 * k1)), e);
    v2 = cscale(csub(cadd(cadd(cscale(q22, 2 * k2), cscale(qs, k1)), cscale(tr, k2)), cscale(m, 2 * k2)), e);
}

// Derived-quantity filters at (k1, k2), with e0 = exp(-(1 - eta) K^2/(4 xi^2))
// and a = 1/(4 xi^2).  Each output spectrum is i^s R with R real-coefficient in
// the spread spectra: s = 1 for the Stokeslet gradient, vorticity and stress
// and the stresslet pressure gradient, s = 0 for the rest.  Slot 2p + j: the
// derivative along p of component j.  Stokeslet, q = spectra of f:
//   gradient (mex_stokes_slp_gradient_kspace.cpp:143-172):
//     R = k_p (K^2 q_j - k_j (k.q)) (1/K^4 + a/K^2) e0
//   pressure gradient (mex_stokes_slp_pressure_grad_kspace.cpp:126-135):
//     R = -k_p (k.q) e0/K^2
//   vorticity (mex_stokes_slp_vorticity_kspace.cpp:142-156):
//     R = -(q1 k2 - q2 k1)(1/K^2 + a) e0
//   stress (mex_stokes_slp_stress_kspace.cpp:170-185):
//     R = (d_pj (k.q) + q_j k_p + q_p k_j - 2 k_p k_j (k.q)/K^2) e0 (1 + a K^2)/K^2
// Stresslet, with Q the spectra of f_i n_j entering only through A = Q11,
// S = Q12 + Q21, D = Q22, T = A + D and kQk = k1^2 A + k1 k2 S + k2^2 D:
//   gradient (mex_stokes_dlp_gradient_kspace.cpp:167-206): R = -k_p v_j with v
//     the velocity bracket of dlp_bracket
//   pressure gradient (mex_stokes_dlp_pressure_grad_kspace.cpp:170-173):
//     R = -k_p kQk e0/K^2
//   vorticity (mex_stokes_dlp_vorticity_kspace.cpp:160-169):
//     R = (2 k1 k2 (A - D) + (k2^2 - k1^2) S)(1/K^2 + a) e0
//   stress (mex_stokes_dlp_stress_kspace.cpp:166-201, mu = 1), c = 1/K^2 + a:
//     R_0 = e0 (kQk/K^2 - c (2 k1^2 T + 4 k1^2 A + 2 k1 k2 S - 4 k1^2 kQk/K^2))
//     R_1 = R_2 = -e0 c (2 k1 k2 (T + A + D) + K^2 S - 4 k1 k2 kQk/K^2)
//     R_3 = e0 (kQk/K^2 - c (2 k2^2 T + 4 k2^2 D + 2 k1 k2 S - 4 k2^2 kQk/K^2))
template <int KIND>
__device__ __forceinline__ void derived_op(double k1, double k2, double a, double eta,
                                           const cplx* q, cplx* R)
{
    const double K2 = k1 * k1 + k2 * k2, iK2 = 1.0 / K2;
    const double e0 = exp(-(1 - eta) * a * K2);
    const double k[2] = {k1, k2};
    if (!dlp_family(KIND)) {
        const cplx kq = cadd(cscale(q[0], k1), cscale(q[1], k2));
        if (KIND == SE2P_SLP_GRADIENT) {
            const double c = (iK2 * iK2 + a * iK2) * e0;
            for (int p = 0; p < 2; ++p)
                for (int j = 0; j < 2; ++j)
                    R[2 * p + j] = cscale(csub(cscale(q[j], K2), cscale(kq, k[j])), k[p] * c);
        } else if (KIND == SE2P_SLP_PRESSURE_GRAD) {
            for (int p = 0; p < 2; ++p) R[p] = cscale(kq, -k[p] * e0 * iK2);
        } else if (KIND == SE2P_SLP_VORTICITY) {
            R[0] = cscale(csub(cscale(q[0], k2), cscale(q[1], k1)), -(iK2 + a) * e0);
        } else {  // SE2P_SLP_STRESS
            const double e = e0 * (1 + a * K2) * iK2;
            for (int p = 0; p < 2; ++p)
                for (int j = 0; j < 2; ++j) {
                    cplx v = cadd(cscale(q[j], k[p]), cscale(q[p], k[j]));
                    v = cadd(v, cscale(kq, (p == j ? 1.0 : 0.0) - 2 * k[p] * k[j] * iK2));
                    R[2 * p + j] = cscale(v, e);
                }
        }
        return;
    }
    const cplx A = q[0], S = cadd(q[1], q[2]), D = q[3];
    const cplx kQk = cadd(cadd(cscale(A, k1 * k1), cscale(S, k1 * k2)), cscale(D, k2 * k2));
    if (KIND == SE2P_DLP_GRADIENT) {
        cplx v[2];
        dlp_bracket(k1, k2, a, eta, A, S, D, v[0], v[1]);
        for (int p = 0; p < 2; ++p)
            for (int j = 0; j < 2; ++j) R[2 * p + j] = cscale(v[j], -k[p]);
    } else if (KIND == SE2P_DLP_PRESSURE_GRAD) {
        for (int p = 0; p < 2; ++p) R[p] = cscale(kQk, -k[p] * e0 * iK2);
    } else if (KIND == SE2P_DLP_VORTICITY) {
        R[0] = cscale(cadd(cscale(csub(A, D), 2 * k1 * k2), cscale(S, k2 * k2 - k1 * k1)),
                      (iK2 + a) * e0);
    } else {  // SE2P_DLP_STRESS
        const double c = iK2 + a;
        const cplx T = cadd(A, D);
        const cplx m = cscale(kQk, iK2);
        cplx b0 = cadd(cadd(cscale(T, 2 * k1 * k1), cscale(A, 4 * k1 * k1)), cscale(S, 2 * k1 * k2));
        b0 = csub(b0, cscale(m, 4 * k1 * k1));
        cplx b3 = cadd(cadd(cscale(T, 2 * k2 * k2), cscale(D, 4 * k2 * k2)), cscale(S, 2 * k1 * k2));
        b3 = csub(b3, cscale(m, 4 * k2 * k2));
        cplx b1 = cadd(cscale(T, 4 * k1 * k2), cscale(S, K2));  // 2 k1 k2 (T + A + D)
        b1 = csub(b1, cscale(m, 4 * k1 * k2));
        R[0] = cscale(csub(m, cscale(b0, c)), e0);
        R[1] = R[2] = cscale(b1, -e0 * c);
        R[3] = cscale(csub(m, cscale(b3, c)), e0);
    }
}

// Filter, k = 0 removal and the 1/(Mx My) of the inverse transform.  The
// reference keeps only the real part of each inverse transform
// (mex_stokes_slp_kspace.cpp:180-181), i.e. the Hermitian part of each filtered
// spectrum; the output W = G1h + i G2h packs both Hermitian parts so a single
// inverse FFT returns g1 + i g2.  G(-k) uses the reference's wavenumbers at
// index -k, which differ from -k only on Nyquist lines.
// The filter of one mode (ix, iy) of the transposed layout, k = 0 removal and
// the 1/(Mx My) of the inverse transform.  wk, dk: the forward spectra W (and D
// for stresslet kinds) at the mode; wm, dm: at index (-ix, -iy).  o0 (and o1 for
// kinds with 4 values per target): the packed output spectra.
template <int KIND>
__device__ __forceinline__ void filter_mode(const SysDesc& d, int ix, int iy, double2 wk,
                                            double2 wm, double2 dk, double2 dm, double2& o0,
                                            double2& o1)
{
    const int Mx = d.Mx, My = d.My;
    o1 = make_double2(0.0, 0.0);
    if (ix == 0 && iy == 0) {
        o0 = make_double2(0.0, 0.0);
        return;
    }
    const int jx = ix ? Mx - ix : 0, jy = iy ? My - iy : 0;
    const double k1 = wavenum(ix, Mx, d.Lx), k2 = wavenum(iy, My, d.Ly);
    const double m1 = wavenum(jx, Mx, d.Lx), m2 = wavenum(jy, My, d.Ly);
    const double a = 0.25 / (d.xi * d.xi), eta = d.eta;
    const double scale = 1.0 / ((double)Mx * (double)My);
    if (KIND >= SE2P_SLP_GRADIENT) {
        constexpr int NC = out_comps(KIND);
        constexpr bool imag = KIND == SE2P_SLP_GRADIENT || KIND == SE2P_SLP_VORTICITY ||
                              KIND == SE2P_SLP_STRESS || KIND == SE2P_DLP_PRESSURE_GRAD;
        cplx q[4];
        unpack(wk, wm, q[0], q[1]);
        if (dlp_family(KIND)) unpack(dk, dm, q[2], q[3]);  // (f1 n2, f2 n2)
        cplx Rk[4], Rm[4], G[4];
        derived_op<KIND>(k1, k2, a, eta, q, Rk);
        derived_op<KIND>(m1, m2, a, eta, q, Rm);
        // Hermitian part of i^s R: i^s (R(k) + (-1)^s R(k'))/2
        for (int c = 0; c < NC; ++c) {
            const cplx h = cscale(imag ? csub(Rk[c], Rm[c]) : cadd(Rk[c], Rm[c]), 0.5);
            G[c] = imag ? cplx{-h.im, h.re} : h;
        }
        if (NC == 1) {
            o0 = make_double2(G[0].re * scale, G[0].im * scale);
        } else {
            o0 = make_double2((G[0].re - G[1].im) * scale, (G[0].im + G[1].re) * scale);
            if (NC == 4) o1 = make_double2((G[2].re - G[3].im) * scale, (G[2].im + G[3].re) * scale);
        }
        return;
    }
    cplx g1, g2;
    if (KIND == SE2P_SLP) {
        cplx q1, q2;
        unpack(wk, wm, q1, q2);
        cplx p1, p2, r1, r2;
        if (d.free_R > 0) {
            // the partner index has the same |K| (its wavenumbers are -k, or k's
            // mirror on the Nyquist lines)
            const double e = free_factor(k1 * k1 + k2 * k2, a, eta, d.free_R);
            slp_apply_e(k1, k2, e, q1, q2, p1, p2);
            slp_apply_e(m1, m2, e, q1, q2, r1, r2);
        } else {
            slp_apply(k1, k2, a, eta, q1, q2, p1, p2);
            slp_apply(m1, m2, a, eta, q1, q2, r1, r2);
        }
        g1 = cscale(cadd(p1, r1), 0.5);
        g2 = cscale(cadd(p2, r2), 0.5);
    } else if (KIND == SE2P_SLP_PRESSURE) {
        // P(k) = -i (k.q) e^{-(1-eta) K^2/(4 xi^2)}/K^2
        // (mex_stokes_slp_pressure_kspace.cpp:120-150); its Hermitian part with
        // k' the index -k's wavenumbers is -i ((k.q) g(k) - (k'.q) g(k'))/2.
        cplx q1, q2;
        unpack(wk, wm, q1, q2);
        const double K2 = k1 * k1 + k2 * k2, M2 = m1 * m1 + m2 * m2;
        const double gk = exp(-(1 - eta) * a * K2) / K2, gm_ = exp(-(1 - eta) * a * M2) / M2;
        const cplx dk = cadd(cscale(q1, k1 * gk), cscale(q2, k2 * gk));
        const cplx dm = cadd(cscale(q1, m1 * gm_), cscale(q2, m2 * gm_));
        const cplx dd = cscale(csub(dk, dm), 0.5);
        g1 = {dd.im, -dd.re};
        g2 = {0.0, 0.0};
    } else if (KIND == SE2P_DLP_PRESSURE) {
        // P(k) = -(k.Q.k) e^{-(1-eta) K^2/(4 xi^2)}/K^2 with Q_ij the spread f_i n_j
        // (mex_stokes_dlp_pressure_kspace.cpp:150-190); a real operator, so its
        // Hermitian part is the mean of the operators at k and k'.
        cplx q11, q21, q12, q22;
        unpack(wk, wm, q11, q21);  // spread (f1 n1, f2 n1)
        unpack(dk, dm, q12, q22);  // spread (f1 n2, f2 n2)
        const cplx qs = cadd(q12, q21);
        const double K2 = k1 * k1 + k2 * k2, M2 = m1 * m1 + m2 * m2;
        const double gk = -exp(-(1 - eta) * a * K2) / K2, gm_ = -exp(-(1 - eta) * a * M2) / M2;
        const cplx pk = cscale(cadd(cadd(cscale(q11, k1 * k1), cscale(qs, k1 * k2)),
                                    cscale(q22, k2 * k2)), gk);
        const cplx pm = cscale(cadd(cadd(cscale(q11, m1 * m1), cscale(qs, m1 * m2)),
                                    cscale(q22, m2 * m2)), gm_);
        g1 = cscale(cadd(pk, pm), 0.5);
        g2 = {0.0, 0.0};
    } else {
        cplx q11, q21, q12, q22;
        unpack(wk, wm, q11, q21);  // spread (f1 n1, f2 n1)
        unpack(dk, dm, q12, q22);  // spread (f1 n2, f2 n2)
        const cplx qs = cadd(q12, q21);
        cplx p1, p2, r1, r2;
        dlp_bracket(k1, k2, a, eta, q11, qs, q22, p1, p2);
        dlp_bracket(m1, m2, a, eta, q11, qs, q22, r1, r2);
        // Hermitian part of i e v(k): i (e v(k) - e' v(-k index)) / 2
        const cplx d1 = cscale(csub(p1, r1), 0.5), d2 = cscale(csub(p2, r2), 0.5);
        g1 = {-d1.im, d1.re};
        g2 = {-d2.im, d2.re};
    }
    // W = g1 + i g2
    o0 = make_double2((g1.re - g2.im) * scale, (g1.im + g2.re) * scale);
}

// filter_mode<SE2P_SLP> of a periodic system (free_R = 0), the same arithmetic
// in the same order without the free-space branch, for the fused x-line
// transform, whose registers it has to share with the transform stages.
__device__ __forceinline__ double2 slp_filter_periodic(const SysDesc& d, int ix, int iy,
                                                       double2 wk, double2 wm)
{
    const int Mx = d.Mx, My = d.My;
    if (ix == 0 && iy == 0) return make_double2(0.0, 0.0);
    const int jx = ix ? Mx - ix : 0, jy = iy ? My - iy : 0;
    const double k1 = wavenum(ix, Mx, d.Lx), k2 = wavenum(iy, My, d.Ly);
    const double m1 = wavenum(jx, Mx, d.Lx), m2 = wavenum(jy, My, d.Ly);
    const double a = 0.25 / (d.xi * d.xi), eta = d.eta;
    const double scale = 1.0 / ((double)Mx * (double)My);
    cplx q1, q2;
    unpack(wk, wm, q1, q2);
    cplx p1, p2, r1, r2;
    slp_apply(k1, k2, a, eta, q1, q2, p1, p2);
    slp_apply(m1, m2, a, eta, q1, q2, r1, r2);
    const cplx g1 = cscale(cadd(p1, r1), 0.5), g2 = cscale(cadd(p2, r2), 0.5);
    return make_double2((g1.re - g2.im) * scale, (g1.im + g2.re) * scale);
}

}  // namespace kf
}  // namespace se2p
