// k-space part of the spectral Ewald sum: FGG spreading, halo merge, Stokes
// filters, FGG gather.
//
// Follows PeriodicEwald/src/mex/common/ewald_tools.cpp:120-271 (Spread,
// Gather) and mex_stokes_{slp,dlp}_kspace.cpp (filters).  The Gaussian
// exp(-2 xi^2 |x - x_node|^2 / eta) over the (P+1)^2 nodes that start at a
// point's start node factors into wx[i] wy[j]; both 1D factors come from the
// fast-Gaussian-gridding recurrence wx[i] = exp(t(px^2 + 2 w px)) e3x^i e1[i],
// e3x = exp(-2 t h px), e1[i] = exp(t h^2 (i - P/2)^2), t = -2 xi^2/eta.
//
// Spreading.  Points are binned by the tile (TX x TY start nodes) their start
// node falls in, in increasing index order.  A tile's points touch an R x R
// node region (R = TX + P).  Its contribution to the two real grids is the
// product  H_c = A^T diag(d_c) B  with A[k][gx] = wx_k(gx), B[k][gy] = wy_k(gy)
// (zero outside the support), evaluated as a register-blocked FP64 outer-
// product accumulation: 64 lanes, each owning a (R/8) x (R/8) block of nodes
// of both grids.  Every node sums its points in index order, so the result is
// bit-reproducible.  The R x R partial regions are then merged into the grid
// by a node-parallel pass that adds the overlapping tiles in a fixed order.
//
// Two real grids travel as one complex grid Z = H_0 + i H_1, so the SLP needs
// one forward and one inverse complex FFT; the DLP spreads (f1 n1, f2 n1) and
// (f1 n2, f2 n2) into two complex grids.
#include "se2p_internal.h"

namespace se2p {

namespace {

constexpr int kMaxP = 40;

__device__ __forceinline__ void point_density(int mode, int p, const double* __restrict__ f,
                                              const double* __restrict__ nrm, double& d0,
                                              double& d1)
{
    const double f1 = f[2 * p], f2 = f[2 * p + 1];
    if (mode == 0) {  // SLP: (f1, f2)
        d0 = f1;
        d1 = f2;
    } else {  // DLP: (f1 n_c, f2 n_c), c = mode - 1 (mex_stokes_dlp_kspace.cpp:79-87)
        const double nc = nrm[2 * p + (mode - 1)];
        d0 = f1 * nc;
        d1 = f2 * nc;
    }
}

template <int BLK>
__global__ __launch_bounds__(64) void k_spread(const SysDesc* __restrict__ sys,
                                               const Job* __restrict__ jobs, int c,
                                               const double* __restrict__ f_base)
{
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int mode = dlp_family(d.kind) ? 1 + c : 0;
    const int* __restrict__ tstart = d.sp_start;
    const int* __restrict__ perm = d.sp_perm;
    const int2* __restrict__ node = d.sp_node;
    const double2* __restrict__ off = d.sp_off;
    const double* __restrict__ f = f_base + d.f_off;
    const double* __restrict__ nrm = d.nrm;
    const int P = d.P, TX = d.TX, TY = d.TY, ntx = d.ntx;
    const double t = d.t, w = d.w, h = d.h;
    const double* __restrict__ e1tab = d.e1tab;
    constexpr int R = 8 * BLK;
    constexpr int KC = 16;
    __shared__ double A[KC][R];
    __shared__ double B0[KC][R];
    __shared__ double B1[KC][R];
    __shared__ double e1s[kMaxP + 1];

    const int tile = jb.y;
    const int tx = tile % ntx, ty = tile / ntx;
    const int lane = threadIdx.x;
    const int bi = (lane >> 3) * BLK, bj = (lane & 7) * BLK;
    for (int i = lane; i <= P; i += 64) e1s[i] = e1tab[i];

    double acc0[BLK][BLK], acc1[BLK][BLK];
#pragma unroll
    for (int i = 0; i < BLK; ++i)
#pragma unroll
        for (int j = 0; j < BLK; ++j) acc0[i][j] = acc1[i][j] = 0;

    const int k0 = tstart[tile], k1 = tstart[tile + 1];
    for (int kb = k0; kb < k1; kb += KC) {
        const int cnt = min(KC, k1 - kb);
        __syncthreads();
        for (int i = lane; i < KC * R; i += 64) {
            (&A[0][0])[i] = 0;
            (&B0[0][0])[i] = 0;
            (&B1[0][0])[i] = 0;
        }
        __syncthreads();
        if (lane < cnt) {
            const int p = perm[kb + lane];
            const int2 m = node[p];
            const double2 o = off[p];
            const int ox = m.x - tx * TX, oy = m.y - ty * TY;
            double d0, d1;
            point_density(mode, p, f, nrm, d0, d1);
            double ex = exp(t * (o.x * o.x + 2 * w * o.x));
            double ey = exp(t * (o.y * o.y + 2 * w * o.y));
            const double e3x = exp(-2 * t * h * o.x);
            const double e3y = exp(-2 * t * h * o.y);
            for (int i = 0; i <= P; ++i) {
                A[lane][ox + i] = ex * e1s[i];
                const double wy = ey * e1s[i];
                B0[lane][oy + i] = d0 * wy;
                B1[lane][oy + i] = d1 * wy;
                ex *= e3x;
                ey *= e3y;
            }
        }
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
            double a[BLK], b0[BLK], b1[BLK];
#pragma unroll
            for (int i = 0; i < BLK; ++i) {
                a[i] = A[k][bi + i];
                b0[i] = B0[k][bj + i];
                b1[i] = B1[k][bj + i];
            }
#pragma unroll
            for (int i = 0; i < BLK; ++i)
#pragma unroll
                for (int j = 0; j < BLK; ++j) {
                    acc0[i][j] = fma(a[i], b0[j], acc0[i][j]);
                    acc1[i][j] = fma(a[i], b1[j], acc1[i][j]);
                }
        }
    }
    double* base = d.partial + (size_t)tile * 2 * R * R;
#pragma unroll
    for (int i = 0; i < BLK; ++i)
#pragma unroll
        for (int j = 0; j < BLK; ++j) {
            base[(bi + i) * R + bj + j] = acc0[i][j];
            base[R * R + (bi + i) * R + bj + j] = acc1[i][j];
        }
}

// Node-parallel halo merge: node (ix, iy) adds, in a fixed order, every tile
// region that covers it (regions wrap periodically, possibly more than once
// on grids smaller than R).  Job b covers node column ix = b / ceil(My/256)
// and 256 nodes of it; indices wrap by a single conditional add (tile and node
// differences stay within one period).
__global__ void k_merge(const SysDesc* __restrict__ sys, const Job* __restrict__ jobs, int out)
{
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int Mx = d.Mx, My = d.My, TX = d.TX, TY = d.TY, ntx = d.ntx, nty = d.nty, R = d.R;
    const int nyb = (My + 255) / 256;
    const int ix = jb.y / nyb;
    const int iy = (jb.y - ix * nyb) * 256 + threadIdx.x;
    if (iy >= My) return;
    const double* __restrict__ partial = d.partial;
    const int ndx = min(ntx, R / TX + 2), ndy = min(nty, R / TY + 2);
    const int bx = ix / TX, by = iy / TY;
    double s0 = 0, s1 = 0;
    for (int dx = 0; dx < ndx; ++dx) {
        int tx = bx - dx;
        tx += tx < 0 ? ntx : 0;
        int lx = ix - tx * TX;
        lx += lx < 0 ? Mx : 0;
        for (; lx < R; lx += Mx) {
            for (int dy = 0; dy < ndy; ++dy) {
                int ty = by - dy;
                ty += ty < 0 ? nty : 0;
                const double* q = partial + (size_t)(ty * ntx + tx) * 2 * R * R + (size_t)lx * R;
                int ly = iy - ty * TY;
                ly += ly < 0 ? My : 0;
                for (; ly < R; ly += My) {
                    s0 += q[ly];
                    s1 += q[R * R + ly];
                }
            }
        }
    }
    d.grid[out][(size_t)ix * My + iy] = make_double2(s0, s1);
}

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
__device__ __forceinline__ void slp_apply(double k1, double k2, double a, double eta, cplx q1,
                                          cplx q2, cplx& g1, cplx& g2)
{
    const double K2 = k1 * k1 + k2 * k2;
    const double e = (1.0 / (K2 * K2) + a / K2) * exp(-(1 - eta) * a * K2);
    const cplx kd = cadd(cscale(q1, k1), cscale(q2, k2));
    g1 = cscale(csub(cscale(q1, K2), cscale(kd, k1)), e);
    g2 = cscale(csub(cscale(q2, K2), cscale(kd, k2)), e);
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
template <int KIND>
__global__ void k_filter(const SysDesc* __restrict__ sys, const Job* __restrict__ jobs)
{
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int Mx = d.Mx, My = d.My;
    const long gid = (long)jb.y * blockDim.x + threadIdx.x;
    if (gid >= (long)Mx * My) return;
    // transposed layout Z[iy*Mx + ix]
    const int iy = (int)(gid / Mx);
    const int ix = (int)(gid - (long)iy * Mx);
    const double2* __restrict__ Z0 = d.grid[G_W];
    const double2* __restrict__ Z1 = d.grid[G_D];
    double2* __restrict__ W = d.grid[G_Z0];
    if (ix == 0 && iy == 0) {
        W[gid] = make_double2(0.0, 0.0);
        if (out_grids(KIND) == 2) d.grid[G_Z1][gid] = make_double2(0.0, 0.0);
        return;
    }
    const int jx = ix ? Mx - ix : 0, jy = iy ? My - iy : 0;
    const long gm = (long)jy * Mx + jx;
    const double k1 = wavenum(ix, Mx, d.Lx), k2 = wavenum(iy, My, d.Ly);
    const double m1 = wavenum(jx, Mx, d.Lx), m2 = wavenum(jy, My, d.Ly);
    const double a = 0.25 / (d.xi * d.xi), eta = d.eta;
    const double scale = 1.0 / ((double)Mx * (double)My);
    if (KIND >= SE2P_SLP_GRADIENT) {
        constexpr int NC = out_comps(KIND);
        constexpr bool imag = KIND == SE2P_SLP_GRADIENT || KIND == SE2P_SLP_VORTICITY ||
                              KIND == SE2P_SLP_STRESS || KIND == SE2P_DLP_PRESSURE_GRAD;
        cplx q[4];
        unpack(Z0[gid], Z0[gm], q[0], q[1]);
        if (dlp_family(KIND)) unpack(Z1[gid], Z1[gm], q[2], q[3]);  // (f1 n2, f2 n2)
        cplx Rk[4], Rm[4], G[4];
        derived_op<KIND>(k1, k2, a, eta, q, Rk);
        derived_op<KIND>(m1, m2, a, eta, q, Rm);
        // Hermitian part of i^s R: i^s (R(k) + (-1)^s R(k'))/2
        for (int c = 0; c < NC; ++c) {
            const cplx h = cscale(imag ? csub(Rk[c], Rm[c]) : cadd(Rk[c], Rm[c]), 0.5);
            G[c] = imag ? cplx{-h.im, h.re} : h;
        }
        if (NC == 1) {
            W[gid] = make_double2(G[0].re * scale, G[0].im * scale);
        } else {
            W[gid] = make_double2((G[0].re - G[1].im) * scale, (G[0].im + G[1].re) * scale);
            if (NC == 4)
                d.grid[G_Z1][gid] =
                    make_double2((G[2].re - G[3].im) * scale, (G[2].im + G[3].re) * scale);
        }
        return;
    }
    cplx g1, g2;
    if (KIND == SE2P_SLP) {
        cplx q1, q2;
        unpack(Z0[gid], Z0[gm], q1, q2);
        cplx p1, p2, r1, r2;
        slp_apply(k1, k2, a, eta, q1, q2, p1, p2);
        slp_apply(m1, m2, a, eta, q1, q2, r1, r2);
        g1 = cscale(cadd(p1, r1), 0.5);
        g2 = cscale(cadd(p2, r2), 0.5);
    } else if (KIND == SE2P_SLP_PRESSURE) {
        // P(k) = -i (k.q) e^{-(1-eta) K^2/(4 xi^2)}/K^2
        // (mex_stokes_slp_pressure_kspace.cpp:120-150); its Hermitian part with
        // k' the index -k's wavenumbers is -i ((k.q) g(k) - (k'.q) g(k'))/2.
        cplx q1, q2;
        unpack(Z0[gid], Z0[gm], q1, q2);
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
        unpack(Z0[gid], Z0[gm], q11, q21);  // spread (f1 n1, f2 n1)
        unpack(Z1[gid], Z1[gm], q12, q22);  // spread (f1 n2, f2 n2)
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
        unpack(Z0[gid], Z0[gm], q11, q21);  // spread (f1 n1, f2 n1)
        unpack(Z1[gid], Z1[gm], q12, q22);  // spread (f1 n2, f2 n2)
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
    W[gid] = make_double2((g1.re - g2.im) * scale, (g1.im + g2.re) * scale);
}

__global__ __launch_bounds__(256) void k_gather(const SysDesc* __restrict__ sys,
                                               const Job* __restrict__ jobs,
                                               double* __restrict__ u_base, int g,
                                               int accumulate, int use_zm)
{
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int* __restrict__ tstart = d.gt_start;
    const int* __restrict__ perm = d.gt_perm;
    const int2* __restrict__ node = d.gt_node;
    const double2* __restrict__ off = d.gt_off;
    const double2* __restrict__ W = d.grid[g ? G_D : G_W];
    const int Mx = d.Mx, My = d.My, TX = d.gTX, TY = d.gTY, ntx = d.gntx, P = d.P, R = d.gR;
    const double t = d.t, w = d.w, h = d.h, scale = d.gscale;
    const double* __restrict__ e1tab = d.e1tab;
    const double* zm = use_zm ? d.zm : nullptr;
    double* u = u_base + d.u_off;
    extern __shared__ double2 reg[];
    __shared__ double e1s[kMaxP + 1];
    const int tile = jb.y;
    const int tx = tile % ntx, ty = tile / ntx;
    const int X0 = tx * TX, Y0 = ty * TY;
    const int k0 = tstart[tile], k1 = tstart[tile + 1];
    if (k0 == k1) return;
    for (int i = threadIdx.x; i <= P; i += blockDim.x) e1s[i] = e1tab[i];
    for (int idx = threadIdx.x; idx < R * R; idx += blockDim.x) {
        const int lx = idx / R, ly = idx % R;
        reg[idx] = W[(size_t)((X0 + lx) % Mx) * My + (Y0 + ly) % My];
    }
    __syncthreads();
    const double z0 = zm ? zm[0] : 0.0, z1 = zm ? zm[1] : 0.0;
    for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const int p = perm[k];
        const int2 m = node[p];
        const double2 o = off[p];
        const int ox = m.x - X0, oy = m.y - Y0;
        double ex = exp(t * (o.x * o.x + 2 * w * o.x));
        const double ey0 = exp(t * (o.y * o.y + 2 * w * o.y));
        const double e3x = exp(-2 * t * h * o.x);
        const double e3y = exp(-2 * t * h * o.y);
        double s0 = 0, s1 = 0;
        for (int i = 0; i <= P; ++i) {
            double ey = ex * ey0 * e1s[i];
            const double2* row = reg + (ox + i) * R + oy;
            for (int j = 0; j <= P; ++j) {
                const double g = ey * e1s[j];
                const double2 v = row[j];
                s0 += g * v.x;
                s1 += g * v.y;
                ey *= e3y;
            }
            ex *= e3x;
        }
        double v0 = s0 * scale + z0, v1 = s1 * scale + z1;
        if (d.nc == 1) {  // one real field: the real part of W
            u[p] = accumulate ? u[p] + v0 : v0;
            continue;
        }
        double* up = u + (size_t)d.nc * p + 2 * g;
        if (accumulate) {
            v0 = up[0] + v0;
            v1 = up[1] + v1;
        }
        up[0] = v0;
        up[1] = v1;
    }
}

__global__ void k_e1_table(double xi, double eta, double h, int P, double* tab)
{
    int j = threadIdx.x;
    if (j > P) return;
    // ewald_tools.cpp:125-127
    const double c = -2 * xi * xi / eta * h * h;
    const int q = j - P / 2;
    tab[j] = exp(c * q * q);
}

// Deterministic two-level reduction of sum_s (f.n)_s (x_s, y_s) over the
// sorted DLP records {x, y, f1, f2, n1, n2, f.n, 0}: job (s, b) sums records
// b*256 + tid + k zm_nblk*256 of system s; then one job per system adds the
// zm_nblk partials in order and divides by the box area.
__global__ __launch_bounds__(256) void k_zero_mode_part(const SysDesc* __restrict__ sys,
                                                        const Job* __restrict__ jobs)
{
    __shared__ double sx[256], sy[256];
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const double* __restrict__ rec = d.s_rec;
    double ax = 0, ay = 0;
    for (int i = jb.y * 256 + threadIdx.x; i < d.Ns; i += d.zm_nblk * 256) {
        if (d.kind == SE2P_DLP) {  // {x, y, f1, f2, n1, n2, f.n, 0}: (f.n) x
            const double* r = rec + 8 * (size_t)i;
            ax += r[6] * r[0];
            ay += r[6] * r[1];
        } else if (d.kind == SE2P_DLP_PRESSURE) {  // f.n
            ax += rec[8 * (size_t)i + 6];
        } else if (d.kind == SE2P_SLP_VORTICITY) {  // {x, y, f1, f2}: f1 y - f2 x
            const double* r = rec + 4 * (size_t)i;
            ax += r[2] * r[1] - r[3] * r[0];
        } else {  // {x, y, f1, f2}: f.x
            const double* r = rec + 4 * (size_t)i;
            ax += r[2] * r[0] + r[3] * r[1];
        }
    }
    sx[threadIdx.x] = ax;
    sy[threadIdx.x] = ay;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            sx[threadIdx.x] += sx[threadIdx.x + k];
            sy[threadIdx.x] += sy[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        d.zm_part[2 * jb.y] = sx[0];
        d.zm_part[2 * jb.y + 1] = sy[0];
    }
}

__global__ void k_zero_mode_final(const SysDesc* __restrict__ sys, const Job* __restrict__ jobs)
{
    if (threadIdx.x != 0) return;
    const SysDesc& d = sys[jobs[blockIdx.x].x];
    double sx = 0, sy = 0;
    for (int b = 0; b < d.zm_nblk; ++b) {
        sx += d.zm_part[2 * b];
        sy += d.zm_part[2 * b + 1];
    }
    const double inv_area = 1.0 / (d.Lx * d.Ly);
    if (d.kind == SE2P_DLP) {  // StokesDLP_ewald_2p.m:149-151
        d.zm[0] = sx * inv_area;
        d.zm[1] = sy * inv_area;
    } else {  // StokesSLP_pressure_ewald_2p.m:127-128, StokesDLP_pressure_ewald_2p.m:156-157,
              // StokesSLP_vorticity_ewald_2p.m:145
        d.zm[0] = -sx * inv_area / 2;
        d.zm[1] = 0;
    }
}

// Density fill of the sorted source records (positions and normals are set at
// build): layout 0 {x, y, f1, f2}; layout 1 (DLP) {x, y, f1, f2, n1, n2, f.n, 0}.
template <int LAYOUT>
__global__ void k_fill_dens(const SysDesc* __restrict__ sys, const Job* __restrict__ jobs,
                            const double* __restrict__ f_base)
{
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int i = jb.y * 256 + threadIdx.x;
    if (i >= d.Ns) return;
    const double* __restrict__ f = f_base + d.f_off;
    const int o = d.s_perm[i];
    const double f1 = f[2 * o], f2 = f[2 * o + 1];
    if (LAYOUT == 0) {
        double4* rec = reinterpret_cast<double4*>(d.s_rec);
        rec[i].z = f1;
        rec[i].w = f2;
    } else {
        double* r = d.s_rec + 8 * (size_t)i;
        r[2] = f1;
        r[3] = f2;
        r[6] = f1 * r[4] + f2 * r[5];
    }
}

}  // namespace

int kgeom_init(KGeom* g, int Mx, int My, int P, double Lx, double Ly, double xi, double eta,
               double w)
{
    if (P < 2 || (P & 1) || P > kMaxP)
        return set_error(SE2P_ERR_UNSUPPORTED, "P = %d: need an even P in [2, %d]", P, kMaxP);
    g->Mx = Mx;
    g->My = My;
    g->P = P;
    g->Lx = Lx;
    g->Ly = Ly;
    g->h = Lx / Mx;
    g->xi = xi;
    g->eta = eta;
    g->w = w;
    g->R = P <= 24 ? 32 : 48;
    g->TX = g->TY = g->R - P;
    g->ntx = (Mx + g->TX - 1) / g->TX;
    g->nty = (My + g->TY - 1) / g->TY;
    // gathering tiles: 32 x 32 start nodes (region up to 56 x 56 for P <= 24) for
    // 256 lanes per tile
    g->gTX = g->gTY = P <= 24 ? 32 : g->TX;
    g->gR = g->gTX + P;
    g->gntx = (Mx + g->gTX - 1) / g->gTX;
    g->gnty = (My + g->gTY - 1) / g->gTY;
    return SE2P_OK;
}

int launch_e1_table(hipStream_t s, const KGeom& g, double* tab)
{
    hipLaunchKernelGGL(k_e1_table, dim3(1), dim3(64), 0, s, g.xi, g.eta, g.h, g.P, tab);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_spread(hipStream_t s, int R, int c, const SysDesc* sys, const Job* jobs, int njobs,
                  const double* f)
{
    if (njobs <= 0) return SE2P_OK;
    if (R == 32)
        hipLaunchKernelGGL(k_spread<4>, dim3(njobs), dim3(64), 0, s, sys, jobs, c, f);
    else
        hipLaunchKernelGGL(k_spread<6>, dim3(njobs), dim3(64), 0, s, sys, jobs, c, f);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_merge(hipStream_t s, int out, const SysDesc* sys, const Job* jobs, int njobs)
{
    if (njobs <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_merge, dim3(njobs), dim3(256), 0, s, sys, jobs, out);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_filter(hipStream_t s, int kind, const SysDesc* sys, const Job* jobs, int njobs)
{
    if (njobs <= 0) return SE2P_OK;
#define SE2P_LAUNCH_FILTER(K) \
    hipLaunchKernelGGL(k_filter<K>, dim3(njobs), dim3(256), 0, s, sys, jobs)
    switch (kind) {
    case SE2P_SLP: SE2P_LAUNCH_FILTER(SE2P_SLP); break;
    case SE2P_DLP: SE2P_LAUNCH_FILTER(SE2P_DLP); break;
    case SE2P_SLP_PRESSURE: SE2P_LAUNCH_FILTER(SE2P_SLP_PRESSURE); break;
    case SE2P_DLP_PRESSURE: SE2P_LAUNCH_FILTER(SE2P_DLP_PRESSURE); break;
    case SE2P_SLP_GRADIENT: SE2P_LAUNCH_FILTER(SE2P_SLP_GRADIENT); break;
    case SE2P_SLP_PRESSURE_GRAD: SE2P_LAUNCH_FILTER(SE2P_SLP_PRESSURE_GRAD); break;
    case SE2P_SLP_VORTICITY: SE2P_LAUNCH_FILTER(SE2P_SLP_VORTICITY); break;
    case SE2P_SLP_STRESS: SE2P_LAUNCH_FILTER(SE2P_SLP_STRESS); break;
    case SE2P_DLP_GRADIENT: SE2P_LAUNCH_FILTER(SE2P_DLP_GRADIENT); break;
    case SE2P_DLP_PRESSURE_GRAD: SE2P_LAUNCH_FILTER(SE2P_DLP_PRESSURE_GRAD); break;
    case SE2P_DLP_VORTICITY: SE2P_LAUNCH_FILTER(SE2P_DLP_VORTICITY); break;
    default: SE2P_LAUNCH_FILTER(SE2P_DLP_STRESS); break;
    }
#undef SE2P_LAUNCH_FILTER
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_gather(hipStream_t s, int maxR, int g, const SysDesc* sys, const Job* jobs, int njobs,
                  double* u, int accumulate, int use_zm)
{
    if (njobs <= 0) return SE2P_OK;
    const size_t lds = sizeof(double2) * (size_t)maxR * maxR;
    hipLaunchKernelGGL(k_gather, dim3(njobs), dim3(256), lds, s, sys, jobs, u, g, accumulate,
                       use_zm);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_zero_mode(hipStream_t s, const SysDesc* sys, const Job* part_jobs, int npart,
                     const Job* sys_jobs, int nsys)
{
    if (npart <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_zero_mode_part, dim3(npart), dim3(256), 0, s, sys, part_jobs);
    SE2P_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_zero_mode_final, dim3(nsys), dim3(64), 0, s, sys, sys_jobs);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_fill_dens(hipStream_t s, int layout, const SysDesc* sys, const Job* jobs, int njobs,
                     const double* f)
{
    if (njobs <= 0) return SE2P_OK;
    if (layout == 0)
        hipLaunchKernelGGL(k_fill_dens<0>, dim3(njobs), dim3(256), 0, s, sys, jobs, f);
    else
        hipLaunchKernelGGL(k_fill_dens<1>, dim3(njobs), dim3(256), 0, s, sys, jobs, f);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

}  // namespace se2p
