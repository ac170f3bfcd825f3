/*
 * CPU oracle for the doubly-periodic 2D Stokes spectral Ewald path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker (and the timed CPU
 * baseline of the benchmark).  The product library never links or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Results do
 * not depend on the OpenMP thread count: every parallel loop writes disjoint
 * outputs and every floating-point reduction runs in a fixed order.
 *
 * The same source built with -DSE2P_REFCPU (-O3, AVX2/FMA, contraction allowed)
 * is the benchmark's CPU baseline, "refcpu": its real-space E1 then comes from
 * rational fits in double precision (accuracy 1e-18 before rounding, the
 * coefficient header e1_coeffs.h) as the reference's mex does with its own fits
 * (mex_stokes_slp_real.cpp:19-48,166-205), instead of the checker's long-double
 * series and continued fraction.
 *
 * Reference paths below are relative to the reference repository root.
 */
#include "se2p_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846
#define ORC_PIL 3.141592653589793238462643383279502884L
#define ORC_EULER 0.577215664901532860606512090082402431L

/* ------------------------------------------------------------------------ */
/* a1: Assign, PeriodicEwald/src/mex/common/ewald_tools.cpp:7-80            */
/* ------------------------------------------------------------------------ */
static int orc_cell_of(double x, double y, double Lx, double Ly, int nsx, int nsy)
{
    /* ewald_tools.cpp:28-34: floor(nside*(x/L+0.5)), clamped to [0, nside-1]. */
    int bx = (int)floor(nsx * (x / Lx + 0.5));
    int by = (int)floor(nsy * (y / Ly + 0.5));
    if (bx < 0) bx = 0;
    if (bx >= nsx) bx = nsx - 1;
    if (by < 0) by = 0;
    if (by >= nsy) by = nsy - 1;
    return by * nsx + bx;
}

void orc_assign(const double* p, int n, double Lx, double Ly, int nside_x, int nside_y,
                int* particle_offsets, int* box_offsets, int* n_in_box)
{
    const int nbox = nside_x * nside_y;
    int* key = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    int* cursor = (int*)malloc(sizeof(int) * (nbox + 1));
    for (int b = 0; b < nbox; ++b) n_in_box[b] = 0;
    for (int i = 0; i < n; ++i) {
        key[i] = orc_cell_of(p[2 * i], p[2 * i + 1], Lx, Ly, nside_x, nside_y);
        n_in_box[key[i]]++;
    }
    /* Exclusive prefix (ewald_tools.cpp:52-68), then a stable scatter in input
     * order (ewald_tools.cpp:71-74). */
    box_offsets[0] = 0;
    for (int b = 0; b < nbox; ++b) box_offsets[b + 1] = box_offsets[b] + n_in_box[b];
    for (int b = 0; b <= nbox; ++b) cursor[b] = box_offsets[b];
    for (int i = 0; i < n; ++i) particle_offsets[cursor[key[i]]++] = i;
    free(key);
    free(cursor);
}

/* ------------------------------------------------------------------------ */
/* a7: FindClosestNode, ewald_tools.cpp:86-114                              */
/* ------------------------------------------------------------------------ */
void orc_find_closest_node(double x, double y, double Lx, double Ly, double h, int P,
                           int* mx, int* my, double* px, double* py)
{
    const double snap = 1e-13;
    double ox = x - h * floor(x / h);
    double oy = y - h * floor(y / h);
    double hx = Lx / 2.0, hy = Ly / 2.0;
    int sx = (int)(floor((x + hx) / h) - P / 2);
    int sy = (int)(floor((y + hy) / h) - P / 2);
    /* A point within 1e-13 of a node is moved onto it; when that node is the
     * upper one the start index advances (ewald_tools.cpp:103-113). */
    if (fabs(ox) < snap || fabs(ox - h) < snap) ox = 0;
    if (fabs(oy) < snap || fabs(oy - h) < snap) oy = 0;
    if (ox == 0 && remainder((x + hx) / h - P / 2, 1) < 0) sx++;
    if (oy == 0 && remainder((y + hy) / h - P / 2, 1) < 0) sy++;
    *mx = sx;
    *my = sy;
    *px = ox;
    *py = oy;
}

/* ------------------------------------------------------------------------ */
/* a4: E1(x).  The reference evaluates E1 with rational fits inline          */
/* (mex_stokes_slp_real.cpp:19-48,166-205).  The oracle uses the textbook    */
/* power series (x <= 1.5) and the even continued fraction (x > 1.5) in long */
/* double, which is independent of any fitted coefficients.                  */
/* ------------------------------------------------------------------------ */
double orc_e1(double xd)
{
    long double x = xd;
    if (!(x > 0)) return INFINITY;
    if (x <= 1.5L) {
        /* E1(x) = -gamma - ln x - sum_{k>=1} (-x)^k / (k k!) */
        long double t = 1.0L, s = 0.0L;
        for (int k = 1; k < 400; ++k) {
            t *= -x / k;
            long double a = t / k;
            s += a;
            if (fabsl(a) <= 1e-22L * fabsl(s)) break;
        }
        return (double)(-ORC_EULER - logl(x) - s);
    }
    /* E1(x) = e^{-x} / (x+1 - 1^2/(x+3 - 2^2/(x+5 - ...))), modified Lentz. */
    const long double tiny = 1e-4000L;
    long double b = x + 1.0L;
    long double c = 1.0L / tiny;
    long double d = 1.0L / b;
    long double hcf = d;
    for (int i = 1; i < 100000; ++i) {
        long double a = -(long double)i * (long double)i;
        b += 2.0L;
        d = 1.0L / (a * d + b);
        c = b + a / c;
        long double del = c * d;
        hcf *= del;
        if (fabsl(del - 1.0L) < 1e-21L) break;
    }
    return (double)(hcf * expl(-x));
}

#ifdef SE2P_REFCPU
#include "e1_coeffs.h"
/* E1 from rational fits: x < 1: -gamma - ln x + x A(x); x >= 1: e^{-x} t P(t)/Q(t),
 * t = 1/x (e = e^{-x} given by the caller). */
static inline double refcpu_e1(double x, double e)
{
    static const double near[] = SE2P_E1_FULL_NEAR;
    static const double fp[] = SE2P_E1_FULL_FARP;
    static const double fq[] = SE2P_E1_FULL_FARQ;
    if (x < 1.0) {
        double a = near[SE2P_E1_FULL_NEAR_DEG];
        for (int i = SE2P_E1_FULL_NEAR_DEG - 1; i >= 0; --i) a = a * x + near[i];
        return -(double)ORC_EULER - log(x) + x * a;
    }
    const double t = 1.0 / x;
    double p = fp[SE2P_E1_FULL_FAR_DEG], q = fq[SE2P_E1_FULL_FAR_DEG];
    for (int i = SE2P_E1_FULL_FAR_DEG - 1; i >= 0; --i) {
        p = p * t + fp[i];
        q = q * t + fq[i];
    }
    return e * t * p / q;
}
#endif

/* ------------------------------------------------------------------------ */
/* Real-space pair kernels                                                   */
/* ------------------------------------------------------------------------ */
/* SLP pair (times 4 pi), r = target - source: (1/2)E1(xi^2 r^2) f + e^{-xi^2 r^2}
 * ((r.f) r / r^2 - f), written as the cancellation-free
 *   u1 += E/2 f1 + (e/r^2) dy (dx f2 - dy f1),  u2 += E/2 f2 - (e/r^2) dx (dx f2 - dy f1)
 * (mex_stokes_slp_real.cpp:213-225; stokes_slp_real_ds.m:61-67). */
static inline void slp_pair(double dx, double dy, double r2, double xi2, double f1, double f2,
                            double* a1, double* a2)
{
    double x = xi2 * r2;
    double e = exp(-x);
#ifdef SE2P_REFCPU
    double E = refcpu_e1(x, e);
#else
    double E = orc_e1(x);
#endif
    double c = e / r2;
    double cr = dx * f2 - dy * f1;
    *a1 += 0.5 * E * f1 + c * dy * cr;
    *a2 += 0.5 * E * f2 - c * dx * cr;
}

/* DLP pair (times 4 pi), r = target - source (mex_stokes_dlp_real.cpp:117-138;
 * stokes_dlp_real_ds.m:208-220):
 *   e^{-xi^2 r^2} [ -4 (1 + xi^2 r^2)/r^4 r_i (r.f)(r.n) + 2 xi^2 (f_i (r.n) + n_i (r.f) + r_i (f.n)) ] */
static inline void dlp_pair(double r1, double r2v, double xi2, double f1, double f2,
                            double n1, double n2, double* a1, double* a2)
{
    double rr = r1 * r1 + r2v * r2v;
    double e = exp(-xi2 * rr);
    double rf = r1 * f1 + r2v * f2;
    double rn = r1 * n1 + r2v * n2;
    double fn = f1 * n1 + f2 * n2;
    double b = -4.0 * (1.0 + xi2 * rr) / (rr * rr) * rf * rn;
    double c = 2.0 * xi2;
    *a1 += e * (b * r1 + c * (f1 * rn + n1 * rf + fn * r1));
    *a2 += e * (b * r2v + c * (f2 * rn + n2 * rf + fn * r2v));
}

/* Neighbour order of the 3x3 stencil (mex_stokes_slp_real.cpp:61-62). */
static const int nb_dx[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
static const int nb_dy[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

typedef struct {
    int* po_s; int* bo_s; int* nb_s;
    int* po_t; int* bo_t; int* nb_t;
} cells_t;

static void cells_build(cells_t* c, const double* psrc, int Ns, const double* ptar, int Nt,
                        double Lx, double Ly, int nsx, int nsy)
{
    int nbox = nsx * nsy;
    c->po_s = (int*)malloc(sizeof(int) * (Ns > 0 ? Ns : 1));
    c->bo_s = (int*)malloc(sizeof(int) * (nbox + 1));
    c->nb_s = (int*)malloc(sizeof(int) * nbox);
    c->po_t = (int*)malloc(sizeof(int) * (Nt > 0 ? Nt : 1));
    c->bo_t = (int*)malloc(sizeof(int) * (nbox + 1));
    c->nb_t = (int*)malloc(sizeof(int) * nbox);
    orc_assign(psrc, Ns, Lx, Ly, nsx, nsy, c->po_s, c->bo_s, c->nb_s);
    orc_assign(ptar, Nt, Lx, Ly, nsx, nsy, c->po_t, c->bo_t, c->nb_t);
}

static void cells_free(cells_t* c)
{
    free(c->po_s); free(c->bo_s); free(c->nb_s);
    free(c->po_t); free(c->bo_t); free(c->nb_t);
}

/* a3/a5: mex_stokes_slp_real.cpp:58-357.  For each target cell: every source of
 * the cell itself with no cutoff, coincident points (r^2 < 1e-15) taking the self
 * term -(1 + gamma/2 + ln xi) f (:124,161-163); then the 8 neighbour cells in the
 * fixed stencil order with cutoff r^2 < Lx Ly/(nside_x nside_y) (:127,233-332),
 * each neighbour's sum kept apart and added at the end (:255,323,327).  Output
 * scaled by 1/(4 pi) in input order (:340-346). */
void orc_slp_real(const double* psrc, const double* ptar, const double* f, int Ns, int Nt,
                  double xi, int nside_x, int nside_y, double Lx, double Ly, double* u)
{
    cells_t c;
    cells_build(&c, psrc, Ns, ptar, Nt, Lx, Ly, nside_x, nside_y);
    const int nbox = nside_x * nside_y;
    const double xi2 = xi * xi;
    const double self = -(1.0 + (double)(ORC_EULER / 2) + log(xi));
    const double cut2 = Lx * Ly / nside_x / nside_y;

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        int bx = b % nside_x, by = b / nside_x;
        for (int it = c.bo_t[b]; it < c.bo_t[b + 1]; ++it) {
            int t = c.po_t[it];
            double xt = ptar[2 * t], yt = ptar[2 * t + 1];
            double a1 = 0, a2 = 0;
            for (int is = c.bo_s[b]; is < c.bo_s[b + 1]; ++is) {
                int s = c.po_s[is];
                double dx = xt - psrc[2 * s], dy = yt - psrc[2 * s + 1];
                double r2 = dx * dx + dy * dy;
                if (fabs(r2) < 1e-15) {
                    a1 += self * f[2 * s];
                    a2 += self * f[2 * s + 1];
                } else {
                    slp_pair(dx, dy, r2, xi2, f[2 * s], f[2 * s + 1], &a1, &a2);
                }
            }
            for (int j = 0; j < 8; ++j) {
                int qx = bx + nb_dx[j], qy = by + nb_dy[j];
                int wx = (qx + nside_x) % nside_x, wy = (qy + nside_y) % nside_y;
                int sb = wy * nside_x + wx;
                double ox = (Lx * (qx - wx)) / nside_x;
                double oy = (Ly * (qy - wy)) / nside_y;
                double t1 = 0, t2 = 0;
                for (int is = c.bo_s[sb]; is < c.bo_s[sb + 1]; ++is) {
                    int s = c.po_s[is];
                    double dx = xt - (psrc[2 * s] + ox), dy = yt - (psrc[2 * s + 1] + oy);
                    double r2 = dx * dx + dy * dy;
                    if (r2 < cut2) slp_pair(dx, dy, r2, xi2, f[2 * s], f[2 * s + 1], &t1, &t2);
                }
                a1 += t1;
                a2 += t2;
            }
            u[2 * t] = a1 / (4 * ORC_PI);
            u[2 * t + 1] = a2 / (4 * ORC_PI);
        }
    }
    cells_free(&c);
}

/* a6: mex_stokes_dlp_real.cpp:15-211.  Same cells and stencil; coincident
 * points (r^2 == 0) are skipped (:122); output scaled by 1/(4 pi) (:197-200). */
void orc_dlp_real(const double* psrc, const double* ptar, const double* f, const double* n,
                  int Ns, int Nt, double xi, int nside_x, int nside_y, double Lx, double Ly,
                  double* u)
{
    cells_t c;
    cells_build(&c, psrc, Ns, ptar, Nt, Lx, Ly, nside_x, nside_y);
    const int nbox = nside_x * nside_y;
    const double xi2 = xi * xi;
    const double cut2 = Lx * Ly / nside_x / nside_y;

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        int bx = b % nside_x, by = b / nside_x;
        for (int it = c.bo_t[b]; it < c.bo_t[b + 1]; ++it) {
            int t = c.po_t[it];
            double xt = ptar[2 * t], yt = ptar[2 * t + 1];
            double a1 = 0, a2 = 0;
            for (int is = c.bo_s[b]; is < c.bo_s[b + 1]; ++is) {
                int s = c.po_s[is];
                double r1 = xt - psrc[2 * s], r2 = yt - psrc[2 * s + 1];
                if (r1 * r1 + r2 * r2 == 0) continue;
                dlp_pair(r1, r2, xi2, f[2 * s], f[2 * s + 1], n[2 * s], n[2 * s + 1], &a1, &a2);
            }
            for (int j = 0; j < 8; ++j) {
                int qx = bx + nb_dx[j], qy = by + nb_dy[j];
                int wx = (qx + nside_x) % nside_x, wy = (qy + nside_y) % nside_y;
                int sb = wy * nside_x + wx;
                double ox = (Lx * (qx - wx)) / nside_x;
                double oy = (Ly * (qy - wy)) / nside_y;
                for (int is = c.bo_s[sb]; is < c.bo_s[sb + 1]; ++is) {
                    int s = c.po_s[is];
                    double r1 = xt - (psrc[2 * s] + ox), r2 = yt - (psrc[2 * s + 1] + oy);
                    if (r1 * r1 + r2 * r2 < cut2)
                        dlp_pair(r1, r2, xi2, f[2 * s], f[2 * s + 1], n[2 * s], n[2 * s + 1],
                                 &a1, &a2);
                }
            }
            u[2 * t] = a1 / (4 * ORC_PI);
            u[2 * t + 1] = a2 / (4 * ORC_PI);
        }
    }
    cells_free(&c);
}

/* ------------------------------------------------------------------------ */
/* a8: Spread, ewald_tools.cpp:120-199                                       */
/* ------------------------------------------------------------------------ */
/* The Gaussian exp(-2 xi^2 |x - x_node|^2 / eta) on the (P+1)^2 nodes starting
 * at the node returned by find_closest_node is a product wx[i] wy[j].  Like the
 * reference, the 1D factors come from the fast-Gaussian-gridding recurrence
 *   wx[i] = exp(t (px^2 + P h px)) e3x^i e1[i],  e3x = exp(-2 t h px),
 *   e1[i] = exp(t h^2 (i - P/2)^2),  t = -2 xi^2 / eta
 * (ewald_tools.cpp:125-127,155-190). */
static void fgg_weights(double px, double py, double h, double t, double w, int P,
                        const double* e1, double* wx, double* wy)
{
    double ex = exp(t * (px * px + 2 * w * px));
    double ey = exp(t * (py * py + 2 * w * py));
    double e3x = exp(-2 * t * h * px);
    double e3y = exp(-2 * t * h * py);
    for (int i = 0; i <= P; ++i) {
        wx[i] = ex * e1[i];
        wy[i] = ey * e1[i];
        ex *= e3x;
        ey *= e3y;
    }
}

static void fgg_table(double xi, double eta, double h, int P, double* e1)
{
    double t = -2 * xi * xi / eta * h * h;
    for (int j = -P / 2; j <= P / 2; ++j) e1[j + P / 2] = exp(t * j * j);
}

void orc_spread(const double* psrc, const double* f, int Ns, double Lx, double Ly,
                double xi, double w, double eta, int P, int Mx, int My,
                double* H1, double* H2)
{
    const double h = Lx / Mx;
    const double t = -2 * xi * xi / eta;
    double* e1 = (double*)malloc(sizeof(double) * (P + 1));
    fgg_table(xi, eta, h, P, e1);
    memset(H1, 0, sizeof(double) * (size_t)Mx * My);
    memset(H2, 0, sizeof(double) * (size_t)Mx * My);

    /* Every thread owns a contiguous range of grid columns and visits all
     * sources in input order, so each node accumulates in the same order as a
     * serial spread whatever the thread count (the reference serialises
     * columns with locks instead, ewald_tools.cpp:133-138,167,188). */
#pragma omp parallel
    {
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        int c0 = (int)((long)Mx * tid / nth), c1 = (int)((long)Mx * (tid + 1) / nth);
        double wx[64], wy[64];
        for (int k = 0; k < Ns; ++k) {
            int mx, my;
            double px, py;
            orc_find_closest_node(psrc[2 * k], psrc[2 * k + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
            /* Does the support [mx, mx+P] (mod Mx) meet [c0, c1)? */
            int hit = 0;
            for (int i = 0; i <= P && !hit; ++i) {
                int cx = ((i + mx) % Mx + Mx) % Mx;
                hit = (cx >= c0 && cx < c1);
            }
            if (!hit) continue;
            fgg_weights(px, py, h, t, w, P, e1, wx, wy);
            for (int i = 0; i <= P; ++i) {
                int cx = ((i + mx) % Mx + Mx) % Mx;
                if (cx < c0 || cx >= c1) continue;
                double a = wx[i];
                for (int j = 0; j <= P; ++j) {
                    int cy = ((j + my) % My + My) % My;
                    double g = a * wy[j];
                    H1[(size_t)cx * My + cy] += g * f[2 * k];
                    H2[(size_t)cx * My + cy] += g * f[2 * k + 1];
                }
            }
        }
    }
    free(e1);
}

/* ------------------------------------------------------------------------ */
/* a12: Gather, ewald_tools.cpp:206-271                                      */
/* ------------------------------------------------------------------------ */
void orc_gather(const double* H, int ncomp, int comp, const double* ptar, int Nt,
                double Lx, double Ly, double xi, double w, double eta, int P, int Mx,
                int My, double* out)
{
    const double h = Lx / Mx;
    const double t = -2 * xi * xi / eta;
    double* e1 = (double*)malloc(sizeof(double) * (P + 1));
    fgg_table(xi, eta, h, P, e1);
    /* (4 xi^2/eta)^2 h^2 / pi / (4 pi) (ewald_tools.cpp:267-269). */
    double s = 4 * xi * xi / eta;
    s = s * s * h * h / ORC_PI / (4 * ORC_PI);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nt; ++k) {
        int mx, my;
        double px, py, wx[64], wy[64];
        orc_find_closest_node(ptar[2 * k], ptar[2 * k + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
        fgg_weights(px, py, h, t, w, P, e1, wx, wy);
        double acc = 0;
        for (int i = 0; i <= P; ++i) {
            int cx = ((i + mx) % Mx + Mx) % Mx;
            double a = wx[i];
            for (int j = 0; j <= P; ++j) {
                int cy = ((j + my) % My + My) % My;
                acc += a * wy[j] * H[(size_t)cx * My + cy];
            }
        }
        out[ncomp * k + comp] = acc * s;
    }
    free(e1);
}

/* ------------------------------------------------------------------------ */
/* a9: 2D DFT.  The reference calls MATLAB's fft2/ifft2 through mexCallMATLAB */
/* (mex_stokes_slp_kspace.cpp:80-81,176-177).  The oracle has its own        */
/* recursive mixed-radix DFT for any length: split off the smallest prime    */
/* factor p, transform the p decimated subsequences, combine with O(p) work  */
/* per output.  Twiddles are rounded from long double.                       */
/* ------------------------------------------------------------------------ */
static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* x: n complex values at stride s (interleaved re,im); y: output, contiguous;
 * tw: table of exp(sign 2 pi i k / N) for k < N, with N / n = step. */
static void dft_rec(const double* x, int n, int s, double* y, const double* tw, int step,
                    double* scratch)
{
    if (n == 1) {
        y[0] = x[0];
        y[1] = x[1];
        return;
    }
    int p = smallest_factor(n);
    int m = n / p;
    /* Sub-transforms of x[r + p j], j < m, land in y[r*m .. r*m+m). */
    for (int r = 0; r < p; ++r)
        dft_rec(x + 2 * (size_t)r * s, m, s * p, y + 2 * (size_t)r * m, tw, step * p,
                scratch + 2 * (size_t)n);
    /* X[k + m q] = sum_r W_n^{r (k + m q)} Y_r[k] */
    const int N = n * step;
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            int kk = k + m * q;
            double re = 0, im = 0;
            for (int r = 0; r < p; ++r) {
                long e = ((long)r * kk) % n;
                const double* w = tw + 2 * (size_t)((e * step) % N);
                double yr = y[2 * ((size_t)r * m + k)], yi = y[2 * ((size_t)r * m + k) + 1];
                re += yr * w[0] - yi * w[1];
                im += yr * w[1] + yi * w[0];
            }
            scratch[2 * kk] = re;
            scratch[2 * kk + 1] = im;
        }
    }
    memcpy(y, scratch, sizeof(double) * 2 * (size_t)n);
}

static double* twiddles(int N, int sign)
{
    double* tw = (double*)malloc(sizeof(double) * 2 * (size_t)N);
    for (int k = 0; k < N; ++k) {
        long double a = 2.0L * ORC_PIL * (long double)k / (long double)N;
        tw[2 * k] = (double)cosl(a);
        tw[2 * k + 1] = (double)(sign * sinl(a));
    }
    return tw;
}

static void dft_lines(double* re, double* im, int len, int count, long stride_elem,
                      long line_step, int inverse)
{
    double* tw = twiddles(len, inverse ? 1 : -1);
    const double sc = inverse ? 1.0 / len : 1.0;
#pragma omp parallel
    {
        double* in = (double*)malloc(sizeof(double) * 2 * (size_t)len);
        double* out = (double*)malloc(sizeof(double) * 2 * (size_t)len);
        /* recursion scratch: 2n + 2n/p + ... < 4n complex */
        double* scr = (double*)malloc(sizeof(double) * 8 * (size_t)len + 16);
#pragma omp for schedule(static)
        for (int c = 0; c < count; ++c) {
            for (int i = 0; i < len; ++i) {
                size_t o = (size_t)(c * line_step + i * stride_elem);
                in[2 * i] = re[o];
                in[2 * i + 1] = im[o];
            }
            dft_rec(in, len, 1, out, tw, 1, scr);
            for (int i = 0; i < len; ++i) {
                size_t o = (size_t)(c * line_step + i * stride_elem);
                re[o] = out[2 * i] * sc;
                im[o] = out[2 * i + 1] * sc;
            }
        }
        free(in);
        free(out);
        free(scr);
    }
    free(tw);
}

void orc_fft2(double* re, double* im, int My, int Mx, int inverse)
{
    /* columns: contiguous, length My; rows: stride My, length Mx */
    dft_lines(re, im, My, Mx, 1, My, inverse);
    dft_lines(re, im, Mx, My, My, 1, inverse);
}

/* ------------------------------------------------------------------------ */
/* a10/a11: k-space filters                                                  */
/* ------------------------------------------------------------------------ */
/* Wavenumber of grid index j on an M-point axis of length L: indices up to
 * M/2 are non-negative (mex_stokes_slp_kspace.cpp:115-128,141-150). */
static inline double wavenumber(int j, int M, double L)
{
    return 2.0 * ORC_PI / L * (j <= M / 2 ? j : j - M);
}

/* SLP: H <- (K^2 q - k (k.q)) (1/K^4 + 1/(4 xi^2 K^2)) exp(-(1-eta) K^2/(4 xi^2)),
 * k = 0 set to zero (mex_stokes_slp_kspace.cpp:110-169). */
void orc_slp_filter(double* h1re, double* h1im, double* h2re, double* h2im, int Mx, int My,
                    double Lx, double Ly, double xi, double eta)
{
    const double a = 0.25 / (xi * xi);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Mx; ++ix) {
        double k1 = wavenumber(ix, Mx, Lx);
        for (int iy = 0; iy < My; ++iy) {
            size_t o = (size_t)ix * My + iy;
            if (ix == 0 && iy == 0) {
                h1re[o] = h1im[o] = h2re[o] = h2im[o] = 0;
                continue;
            }
            double k2 = wavenumber(iy, My, Ly);
            double K2 = k1 * k1 + k2 * k2;
            double g = (1.0 / (K2 * K2) + a / K2) * exp(-(1 - eta) * a * K2);
            double dr = k1 * h1re[o] + k2 * h2re[o];
            double di = k1 * h1im[o] + k2 * h2im[o];
            double q1r = h1re[o], q1i = h1im[o], q2r = h2re[o], q2i = h2im[o];
            h1re[o] = (K2 * q1r - k1 * dr) * g;
            h1im[o] = (K2 * q1i - k1 * di) * g;
            h2re[o] = (K2 * q2r - k2 * dr) * g;
            h2im[o] = (K2 * q2i - k2 * di) * g;
        }
    }
}

/* DLP: with Q the 2x2 tensor of spread f_i n_j and e = exp(-(1-eta)K^2/(4xi^2))
 * (1 + K^2/(4 xi^2))/K^2, the two outputs are
 *   G_i = i [ 2 k_i Q_ii + k_j (Q_12 + Q_21) + k_i (Q_11 + Q_22) - 2 k_i (k.Q.k)/K^2 ] e,
 * j != i, k = 0 set to zero (mex_stokes_dlp_kspace.cpp:151-232).  Inputs follow
 * the reference's spreading: h1 = f1 n1, h2 = f2 n1, h3 = f1 n2, h4 = f2 n2
 * (:79-92); only Q_12 + Q_21 = h2 + h3 enters. */
void orc_dlp_filter(double* h1re, double* h1im, double* h2re, double* h2im,
                    const double* h3re, const double* h3im, const double* h4re,
                    const double* h4im, int Mx, int My, double Lx, double Ly, double xi,
                    double eta)
{
    const double a = 0.25 / (xi * xi);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Mx; ++ix) {
        double k1 = wavenumber(ix, Mx, Lx);
        for (int iy = 0; iy < My; ++iy) {
            size_t o = (size_t)ix * My + iy;
            if (ix == 0 && iy == 0) {
                h1re[o] = h1im[o] = h2re[o] = h2im[o] = 0;
                continue;
            }
            double k2 = wavenumber(iy, My, Ly);
            double K2 = k1 * k1 + k2 * k2;
            double e = exp(-(1 - eta) * a * K2) * (1 + a * K2) / K2;
            double q11[2] = {h1re[o], h1im[o]}, q22[2] = {h4re[o], h4im[o]};
            double qs[2] = {h2re[o] + h3re[o], h2im[o] + h3im[o]};
            double v[2][2];
            for (int c = 0; c < 2; ++c) {
                double tr = q11[c] + q22[c];
                double kqk = k1 * k1 * q11[c] + k1 * k2 * qs[c] + k2 * k2 * q22[c];
                double m = kqk / K2;
                v[0][c] = 2 * k1 * q11[c] + k2 * qs[c] + k1 * tr - 2 * k1 * m;
                v[1][c] = 2 * k2 * q22[c] + k1 * qs[c] + k2 * tr - 2 * k2 * m;
            }
            /* multiply by i e */
            h1re[o] = -v[0][1] * e;
            h1im[o] = v[0][0] * e;
            h2re[o] = -v[1][1] * e;
            h2im[o] = v[1][0] * e;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Whole k-space parts                                                       */
/* ------------------------------------------------------------------------ */
/* mex_stokes_slp_kspace.cpp:7-215: spread (f1, f2), forward DFT of each real
 * grid, filter, inverse DFT, real parts gathered into components 1 and 2. */
void orc_slp_kspace(const double* psrc, const double* ptar, double xi, double eta,
                    const double* f, int Ns, int Nt, int Mx, int My, double Lx, double Ly,
                    double w, int P, double* uk)
{
    size_t G = (size_t)Mx * My;
    double* b = (double*)calloc(4 * G, sizeof(double));
    double *h1r = b, *h1i = b + G, *h2r = b + 2 * G, *h2i = b + 3 * G;
    orc_spread(psrc, f, Ns, Lx, Ly, xi, w, eta, P, Mx, My, h1r, h2r);
    orc_fft2(h1r, h1i, My, Mx, 0);
    orc_fft2(h2r, h2i, My, Mx, 0);
    orc_slp_filter(h1r, h1i, h2r, h2i, Mx, My, Lx, Ly, xi, eta);
    orc_fft2(h1r, h1i, My, Mx, 1);
    orc_fft2(h2r, h2i, My, Mx, 1);
    orc_gather(h1r, 2, 0, ptar, Nt, Lx, Ly, xi, w, eta, P, Mx, My, uk);
    orc_gather(h2r, 2, 1, ptar, Nt, Lx, Ly, xi, w, eta, P, Mx, My, uk);
    free(b);
}

/* mex_stokes_dlp_kspace.cpp:10-282: spread v1 = (f1 n1, f2 n1) and
 * v2 = (f1 n2, f2 n2), four forward DFTs, filter, two inverse DFTs, gather.
 * The zero mode is not added here (the driver adds it, StokesDLP_ewald_2p.m:149-151). */
void orc_dlp_kspace(const double* psrc, const double* ptar, double xi, double eta,
                    const double* f, const double* n, int Ns, int Nt, int Mx, int My,
                    double Lx, double Ly, double w, int P, double* uk)
{
    size_t G = (size_t)Mx * My;
    double* b = (double*)calloc(8 * G, sizeof(double));
    double* v1 = (double*)malloc(sizeof(double) * 2 * (size_t)(Ns > 0 ? Ns : 1));
    double* v2 = (double*)malloc(sizeof(double) * 2 * (size_t)(Ns > 0 ? Ns : 1));
    for (int i = 0; i < Ns; ++i) {
        v1[2 * i] = f[2 * i] * n[2 * i];
        v1[2 * i + 1] = f[2 * i + 1] * n[2 * i];
        v2[2 * i] = f[2 * i] * n[2 * i + 1];
        v2[2 * i + 1] = f[2 * i + 1] * n[2 * i + 1];
    }
    double* h[8];
    for (int j = 0; j < 8; ++j) h[j] = b + j * G;
    orc_spread(psrc, v1, Ns, Lx, Ly, xi, w, eta, P, Mx, My, h[0], h[2]);
    orc_spread(psrc, v2, Ns, Lx, Ly, xi, w, eta, P, Mx, My, h[4], h[6]);
    for (int j = 0; j < 4; ++j) orc_fft2(h[2 * j], h[2 * j + 1], My, Mx, 0);
    orc_dlp_filter(h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], Mx, My, Lx, Ly, xi, eta);
    orc_fft2(h[0], h[1], My, Mx, 1);
    orc_fft2(h[2], h[3], My, Mx, 1);
    orc_gather(h[0], 2, 0, ptar, Nt, Lx, Ly, xi, w, eta, P, Mx, My, uk);
    orc_gather(h[2], 2, 1, ptar, Nt, Lx, Ly, xi, w, eta, P, Mx, My, uk);
    free(v1);
    free(v2);
    free(b);
}

/* ------------------------------------------------------------------------ */
/* Direct sums (tests/direct_sums/matlab_sums/stokes_*_ds.m files)           */
/* ------------------------------------------------------------------------ */
/* stokes_slp_real_ds.m:21-69: all 3x3 images, no cutoff, self term when r < 1e-13. */
void orc_slp_real_ds(const double* psrc, const double* ptar, const double* f, int Ns,
                     int Nt, double Lx, double Ly, double xi, double* u)
{
    const double xi2 = xi * xi;
    const double self = -((double)(ORC_EULER / 2) + log(xi) + 1.0);
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double a1 = 0, a2 = 0;
        for (int s = 0; s < Ns; ++s) {
            for (int jx = -1; jx <= 1; ++jx) {
                for (int jy = -1; jy <= 1; ++jy) {
                    double dx = ptar[2 * m] - (psrc[2 * s] + jx * Lx);
                    double dy = ptar[2 * m + 1] - (psrc[2 * s + 1] + jy * Ly);
                    double r2 = dx * dx + dy * dy;
                    if (sqrt(r2) < 1e-13) {
                        a1 += self * f[2 * s];
                        a2 += self * f[2 * s + 1];
                    } else {
                        slp_pair(dx, dy, r2, xi2, f[2 * s], f[2 * s + 1], &a1, &a2);
                    }
                }
            }
        }
        u[2 * m] = a1 / (4 * ORC_PI);
        u[2 * m + 1] = a2 / (4 * ORC_PI);
    }
}

/* stokes_dlp_real_ds.m:168-221: all 3x3 images, coincident images skipped. */
void orc_dlp_real_ds(const double* psrc, const double* ptar, const double* f,
                     const double* n, int Ns, int Nt, double Lx, double Ly, double xi,
                     double* u)
{
    const double xi2 = xi * xi;
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double a1 = 0, a2 = 0;
        for (int s = 0; s < Ns; ++s) {
            for (int jx = -1; jx <= 1; ++jx) {
                for (int jy = -1; jy <= 1; ++jy) {
                    double r1 = ptar[2 * m] - (psrc[2 * s] + jx * Lx);
                    double r2 = ptar[2 * m + 1] - (psrc[2 * s + 1] + jy * Ly);
                    if (sqrt(r1 * r1 + r2 * r2) < 1e-13) continue;
                    dlp_pair(r1, r2, xi2, f[2 * s], f[2 * s + 1], n[2 * s], n[2 * s + 1], &a1, &a2);
                }
            }
        }
        u[2 * m] = a1 / (4 * ORC_PI);
        u[2 * m + 1] = a2 / (4 * ORC_PI);
    }
}

/* Fourier sums over modes (-kinf, kinf]^2 \ {0}, written through structure
 * factors S(k) = sum_s v_s e^{-i k.x_s} so the cost is O(modes (Ns + Nt))
 * instead of O(modes Ns Nt); the sum is the same as stokes_slp_kspace_ds.m:90-134
 * (the SLP filter is even in k, so e^{-ik.r} and e^{ik.r} agree) and
 * stokes_dlp_kspace_ds.m:244-306 (e^{+ik.(x_t - x_s)}). */
static void kspace_ds(int kind, const double* psrc, const double* ptar, const double* f,
                      const double* n, int Ns, int Nt, double Lx, double Ly, double xi,
                      int kinf, double* u)
{
    const int K = 2 * kinf;
    const size_t nm = (size_t)K * K;
    double* S = (double*)calloc(nm * 8, sizeof(double)); /* up to 4 complex per mode */
    const double a = 0.25 / (xi * xi);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)nm; ++q) {
        int j1 = (int)(q / K) - kinf + 1, j2 = (int)(q % K) - kinf + 1;
        double k1 = 2 * ORC_PI / Lx * j1, k2 = 2 * ORC_PI / Ly * j2;
        double* s = S + 8 * q;
        for (int i = 0; i < Ns; ++i) {
            double ph = -(k1 * psrc[2 * i] + k2 * psrc[2 * i + 1]);
            double c = cos(ph), sn = sin(ph);
            if (kind == 0) {
                s[0] += f[2 * i] * c; s[1] += f[2 * i] * sn;
                s[2] += f[2 * i + 1] * c; s[3] += f[2 * i + 1] * sn;
            } else {
                /* DLP per-source vector: f (k.n) + n (k.f) + k (f.n) - 2 k (k.n)(k.f)/K^2 */
                double K2 = k1 * k1 + k2 * k2;
                if (K2 == 0) continue;
                double kn = k1 * n[2 * i] + k2 * n[2 * i + 1];
                double kf = k1 * f[2 * i] + k2 * f[2 * i + 1];
                double fn = f[2 * i] * n[2 * i] + f[2 * i + 1] * n[2 * i + 1];
                double v1 = f[2 * i] * kn + n[2 * i] * kf + k1 * fn - 2 * k1 * kn * kf / K2;
                double v2 = f[2 * i + 1] * kn + n[2 * i + 1] * kf + k2 * fn - 2 * k2 * kn * kf / K2;
                s[0] += v1 * c; s[1] += v1 * sn;
                s[2] += v2 * c; s[3] += v2 * sn;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double a1 = 0, a2 = 0;
        for (size_t q = 0; q < nm; ++q) {
            int j1 = (int)(q / K) - kinf + 1, j2 = (int)(q % K) - kinf + 1;
            if (j1 == 0 && j2 == 0) continue;
            double k1 = 2 * ORC_PI / Lx * j1, k2 = 2 * ORC_PI / Ly * j2;
            double K2 = k1 * k1 + k2 * k2;
            double g = (1 + a * K2) * exp(-a * K2) / K2;
            double ph = k1 * ptar[2 * m] + k2 * ptar[2 * m + 1];
            double c = cos(ph), sn = sin(ph);
            const double* s = S + 8 * q;
            if (kind == 0) {
                /* (f - k (k.f)/K^2) g, real part of times e^{ik.x_t} */
                double kr = (k1 * s[0] + k2 * s[2]) / K2, ki = (k1 * s[1] + k2 * s[3]) / K2;
                double v1r = s[0] - k1 * kr, v1i = s[1] - k1 * ki;
                double v2r = s[2] - k2 * kr, v2i = s[3] - k2 * ki;
                a1 += g * (v1r * c - v1i * sn);
                a2 += g * (v2r * c - v2i * sn);
            } else {
                /* i v g e^{ik.x_t}: real part is -(v_r sin + v_i cos) g */
                a1 += -g * (s[0] * sn + s[1] * c);
                a2 += -g * (s[2] * sn + s[3] * c);
            }
        }
        u[2 * m] = a1 / (Lx * Ly);
        u[2 * m + 1] = a2 / (Lx * Ly);
    }
    free(S);
}

void orc_slp_kspace_ds(const double* psrc, const double* ptar, const double* f, int Ns,
                       int Nt, double Lx, double Ly, double xi, int kinf, double* u)
{
    kspace_ds(0, psrc, ptar, f, NULL, Ns, Nt, Lx, Ly, xi, kinf, u);
}

void orc_dlp_kspace_ds(const double* psrc, const double* ptar, const double* f,
                       const double* n, int Ns, int Nt, double Lx, double Ly, double xi,
                       int kinf, double* u)
{
    kspace_ds(1, psrc, ptar, f, n, Ns, Nt, Lx, Ly, xi, kinf, u);
    double z[2];
    orc_dlp_zero_mode(psrc, f, n, Ns, Lx, Ly, z);
    for (int m = 0; m < Nt; ++m) {
        u[2 * m] += z[0];
        u[2 * m + 1] += z[1];
    }
}

void orc_dlp_zero_mode(const double* psrc, const double* f, const double* n, int Ns,
                       double Lx, double Ly, double out[2])
{
    double sx = 0, sy = 0;
    for (int i = 0; i < Ns; ++i) {
        double fn = f[2 * i] * n[2 * i] + f[2 * i + 1] * n[2 * i + 1];
        sx += fn * psrc[2 * i];
        sy += fn * psrc[2 * i + 1];
    }
    out[0] = sx / (Lx * Ly);
    out[1] = sy / (Lx * Ly);
}

/* ------------------------------------------------------------------------ */
/* a14: parameter recipe                                                     */
/* ------------------------------------------------------------------------ */
/* MATLAB rat(): continued-fraction expansion of X until |X - N/D| < tol with
 * tol = 1e-6 |X| (MATLAB's documented default). */
void orc_rat(double X, int* N, int* D)
{
    double tol = 1e-6 * fabs(X);
    double y = X;
    double lastn = 1, lastd = 0;
    double nn = round(y), dd = 1;
    double frac = y - nn;
    while (fabs(y - nn / dd) >= tol) {
        double flip = 1 / frac;
        double step = round(flip);
        frac = flip - step;
        double sn = nn, sd = dd;
        nn = nn * step + lastn;
        dd = dd * step + lastd;
        lastn = sn;
        lastd = sd;
    }
    *N = (int)(nn * (dd < 0 ? -1 : 1));
    *D = (int)fabs(dd);
}

/* Newton iterations of StokesSLP_ewald_2p.m:147-202 and StokesDLP_ewald_2p.m:171-225:
 * step limited to 0.9 x, at most 101 steps, stop when |f| <= tol. */
static double newton(int kind, int which, double x, double Q, double Lx, double Ly, double rc,
                     double xi, double tol, int* ok)
{
    *ok = 1;
    for (int it = 0;; ++it) {
        double fv, fp;
        if (which == 0) { /* xi from rc */
            if (kind == 0) {
                fv = sqrt(Q * ORC_PI / (4 * Lx * Ly * x)) * exp(-x * x * rc * rc) - tol;
                fp = -sqrt(Q * ORC_PI / (4 * Lx * Ly)) * exp(-x * x * rc * rc) *
                     (0.5 * pow(x, -1.5) + 2 * rc * rc * x * sqrt(1 / x));
            } else {
                fv = exp(-x * x * rc * rc) * x * rc * sqrt(2 * ORC_PI * Q / (Lx * Ly)) - tol;
                fp = (1 - 2 * x * x * rc * rc) * exp(-x * x * rc * rc) *
                     sqrt(2 * ORC_PI * Q / (Lx * Ly)) * rc;
            }
        } else { /* k_inf from xi */
            double Lm = Lx > Ly ? Lx : Ly;
            double g = exp(-x * x / (4 * xi * xi));
            if (kind == 0) {
                double c = 4 * Q * ORC_PI * Lm / (pow(Lx, 3) * pow(Ly, 3));
                fv = sqrt(c / x) * g - tol;
                fp = -sqrt(c) * g * (0.5 * pow(x, -1.5) + sqrt(1 / x) * 2 * x / (4 * xi * xi));
            } else {
                /* StokesDLP_ewald_2p.m:177-179, with MATLAB's left-to-right precedence
                 * (.../Lx^3*Ly^3 multiplies by Ly^3). */
                double c = 8 * ORC_PI * Q * Lm / pow(Lx, 3) * pow(Ly, 3);
                fv = sqrt(x * c) * g - tol;
                fp = sqrt(c) * g * (0.5 * pow(x, -0.5) - 2 * pow(x, 1.5) / (4 * xi * xi));
            }
        }
        if (!(fabs(fv) > tol)) break;
        if (it > 100) {
            *ok = 0;
            break;
        }
        double d = fv / fp;
        if (!(d < 0.9 * x)) d = 0.9 * x;
        x = x - d;
    }
    return x;
}

int orc_params(int kind, int npts, double Q, double Lx, double Ly, int P, double Nb,
               double tol, double* out, int* iout)
{
    int A, B, ok1, ok2;
    orc_rat(Lx / Ly, &A, &B);
    if (A <= 0 || B <= 0) return -1;
    double a = ceil(sqrt(npts / (Nb * A * B)));
    double m = 0.95 * sqrt(ORC_PI * P);
    int nsx = (int)(a * A), nsy = (int)(a * B);
    double rc = Lx / nsx;
    double xi = newton(kind, 0, kind == 0 ? 4 / rc : 1 / rc, Q, Lx, Ly, rc, 0, tol, &ok1);
    /* Both drivers pass Lx for Ly here (StokesSLP_ewald_2p.m:96). */
    double k0 = round(5 * xi);
    double kinf = newton(kind, 1, k0, Q, Lx, Lx, rc, xi, tol, &ok2);
    int kx = (int)round(Lx * kinf / (2 * ORC_PI));
    int M0 = 2 * kx < 10000 ? 2 * kx : 10000;
    int Mx = A * M0, My = B * M0;
    double w = P * Lx / Mx / 2;
    double eta = pow(2 * xi * w / m, 2);
    out[0] = xi;
    out[1] = rc;
    out[2] = w;
    out[3] = eta;
    iout[0] = nsx;
    iout[1] = nsy;
    iout[2] = Mx;
    iout[3] = My;
    iout[4] = kx;
    iout[5] = A;
    iout[6] = B;
    return (ok1 && ok2) ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* Derived quantity: SLP pressure                                            */
/* ------------------------------------------------------------------------ */
/* mex_stokes_slp_pressure_real.cpp:58-160: the velocity's cells and stencil;
 * coincident points of the own cell skipped (r^2 == 0, :104), neighbours with
 * r^2 < Lx Ly/(nside_x nside_y) (:139); pair term (r.f) e^{-xi^2 r^2}/r^2
 * (:107-108), output / (2 pi) (:150-152). */
void orc_slp_pressure_real(const double* psrc, const double* ptar, const double* f, int Ns,
                           int Nt, double xi, int nside_x, int nside_y, double Lx, double Ly,
                           double* p)
{
    cells_t c;
    cells_build(&c, psrc, Ns, ptar, Nt, Lx, Ly, nside_x, nside_y);
    const int nbox = nside_x * nside_y;
    const double xi2 = xi * xi;
    const double cut2 = Lx * Ly / nside_x / nside_y;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        int bx = b % nside_x, by = b / nside_x;
        for (int it = c.bo_t[b]; it < c.bo_t[b + 1]; ++it) {
            int t = c.po_t[it];
            double xt = ptar[2 * t], yt = ptar[2 * t + 1];
            double a = 0;
            for (int is = c.bo_s[b]; is < c.bo_s[b + 1]; ++is) {
                int s = c.po_s[is];
                double r1 = xt - psrc[2 * s], r2 = yt - psrc[2 * s + 1];
                double rr = r1 * r1 + r2 * r2;
                if (rr == 0) continue;
                a += (r1 * f[2 * s] + r2 * f[2 * s + 1]) * exp(-xi2 * rr) / rr;
            }
            for (int j = 0; j < 8; ++j) {
                int qx = bx + nb_dx[j], qy = by + nb_dy[j];
                int wx = (qx + nside_x) % nside_x, wy = (qy + nside_y) % nside_y;
                int sb = wy * nside_x + wx;
                double ox = (Lx * (qx - wx)) / nside_x;
                double oy = (Ly * (qy - wy)) / nside_y;
                for (int is = c.bo_s[sb]; is < c.bo_s[sb + 1]; ++is) {
                    int s = c.po_s[is];
                    double r1 = xt - (psrc[2 * s] + ox), r2 = yt - (psrc[2 * s + 1] + oy);
                    double rr = r1 * r1 + r2 * r2;
                    if (rr < cut2) a += (r1 * f[2 * s] + r2 * f[2 * s + 1]) * exp(-xi2 * rr) / rr;
                }
            }
            p[t] = a / (2 * ORC_PI);
        }
    }
    cells_free(&c);
}

/* mex_stokes_slp_pressure_kspace.cpp:100-150: with Q the spread (f1, f2)
 * spectra, P(k) = -i (k.Q) exp(-(1-eta) K^2/(4 xi^2)) / K^2 into h1; k = 0 set
 * to zero (:152-154). */
void orc_slp_pressure_filter(double* h1re, double* h1im, const double* h2re,
                             const double* h2im, int Mx, int My, double Lx, double Ly,
                             double xi, double eta)
{
    const double a = 0.25 / (xi * xi);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Mx; ++ix) {
        double k1 = wavenumber(ix, Mx, Lx);
        for (int iy = 0; iy < My; ++iy) {
            size_t o = (size_t)ix * My + iy;
            if (ix == 0 && iy == 0) {
                h1re[o] = h1im[o] = 0;
                continue;
            }
            double k2 = wavenumber(iy, My, Ly);
            double K2 = k1 * k1 + k2 * k2;
            double e = exp(-(1 - eta) * a * K2);
            double dr = k1 * h1re[o] + k2 * h2re[o];
            double di = k1 * h1im[o] + k2 * h2im[o];
            h1re[o] = di * e / K2;
            h1im[o] = -dr * e / K2;
        }
    }
}

/* mex_stokes_slp_pressure_kspace.cpp:7-190: spread (f1, f2), two forward DFTs,
 * filter, inverse DFT, gather of the real part.  The zero mode is not added
 * (the driver adds it, StokesSLP_pressure_ewald_2p.m:127-128). */
void orc_slp_pressure_kspace(const double* psrc, const double* ptar, const double* f,
                             double xi, double eta, int Ns, int Nt, int Mx, int My, double Lx,
                             double Ly, double w, int P, double* pk)
{
    size_t G = (size_t)Mx * My;
    double* b = (double*)calloc(4 * G, sizeof(double));
    double *h1r = b, *h1i = b + G, *h2r = b + 2 * G, *h2i = b + 3 * G;
    orc_spread(psrc, f, Ns, Lx, Ly, xi, w, eta, P, Mx, My, h1r, h2r);
    orc_fft2(h1r, h1i, My, Mx, 0);
    orc_fft2(h2r, h2i, My, Mx, 0);
    orc_slp_pressure_filter(h1r, h1i, h2r, h2i, Mx, My, Lx, Ly, xi, eta);
    orc_fft2(h1r, h1i, My, Mx, 1);
    for (int m = 0; m < Nt; ++m) pk[m] = 0;
    orc_gather(h1r, 1, 0, ptar, Nt, Lx, Ly, xi, w, eta, P, Mx, My, pk);
    free(b);
}

/* StokesSLP_pressure_ewald_2p.m:127-128: -sum_s (f1 x + f2 y)_s / (2 Lx Ly). */
double orc_slp_pressure_zero_mode(const double* psrc, const double* f, int Ns, double Lx,
                                  double Ly)
{
    double s = 0;
    for (int i = 0; i < Ns; ++i) s += f[2 * i] * psrc[2 * i] + f[2 * i + 1] * psrc[2 * i + 1];
    return -s / (2 * Lx * Ly);
}

/* stokes_slp_pressure_real_ds.m: all 3 x 3 images, |r| < 1e-13 skipped, / (2 pi). */
void orc_slp_pressure_real_ds(const double* psrc, const double* ptar, const double* f, int Ns,
                              int Nt, double Lx, double Ly, double xi, double* p)
{
    const double xi2 = xi * xi;
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double a = 0;
        for (int s = 0; s < Ns; ++s)
            for (int jx = -1; jx <= 1; ++jx)
                for (int jy = -1; jy <= 1; ++jy) {
                    double r1 = ptar[2 * m] - (psrc[2 * s] + jx * Lx);
                    double r2 = ptar[2 * m + 1] - (psrc[2 * s + 1] + jy * Ly);
                    double rr = r1 * r1 + r2 * r2;
                    if (sqrt(rr) < 1e-13) continue;
                    a += (r1 * f[2 * s] + r2 * f[2 * s + 1]) * exp(-xi2 * rr) / rr;
                }
        p[m] = a / (2 * ORC_PI);
    }
}

/* stokes_slp_pressure_kspace_ds.m: real part of
 *   sum_{k != 0, |j_i| <= kinf} -i (k.f_s) e^{-K^2/(4 xi^2)}/K^2 e^{i k.(x_t - x_s)} / (Lx Ly)
 * plus the zero mode, through the structure factors S_c(k) = sum_s f_c e^{-i k.x_s}. */
void orc_slp_pressure_kspace_ds(const double* psrc, const double* ptar, const double* f,
                                int Ns, int Nt, double Lx, double Ly, double xi, int kinf,
                                double* p)
{
    const int K = 2 * kinf;
    const size_t nm = (size_t)K * K;
    double* S = (double*)calloc(nm * 4, sizeof(double));
    const double a = 0.25 / (xi * xi);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)nm; ++q) {
        int j1 = (int)(q / K) - kinf + 1, j2 = (int)(q % K) - kinf + 1;
        double k1 = 2 * ORC_PI / Lx * j1, k2 = 2 * ORC_PI / Ly * j2;
        double* s = S + 4 * q;
        for (int i = 0; i < Ns; ++i) {
            double ph = -(k1 * psrc[2 * i] + k2 * psrc[2 * i + 1]);
            double c = cos(ph), sn = sin(ph);
            s[0] += f[2 * i] * c;
            s[1] += f[2 * i] * sn;
            s[2] += f[2 * i + 1] * c;
            s[3] += f[2 * i + 1] * sn;
        }
    }
    const double zm = orc_slp_pressure_zero_mode(psrc, f, Ns, Lx, Ly);
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double acc = 0;
        for (size_t q = 0; q < nm; ++q) {
            int j1 = (int)(q / K) - kinf + 1, j2 = (int)(q % K) - kinf + 1;
            if (j1 == 0 && j2 == 0) continue;
            double k1 = 2 * ORC_PI / Lx * j1, k2 = 2 * ORC_PI / Ly * j2;
            double K2 = k1 * k1 + k2 * k2;
            double g = exp(-a * K2) / K2;
            const double* s = S + 4 * q;
            double Ar = k1 * s[0] + k2 * s[2], Ai = k1 * s[1] + k2 * s[3];
            double ph = k1 * ptar[2 * m] + k2 * ptar[2 * m + 1];
            /* Re(-i (Ar + i Ai) e^{i ph}) = Ai cos ph + Ar sin ph */
            acc += g * (Ai * cos(ph) + Ar * sin(ph));
        }
        p[m] = acc / (Lx * Ly) + zm;
    }
    free(S);
}

/* ------------------------------------------------------------------------ */
/* Derived quantity: DLP (stresslet) pressure                                */
/* ------------------------------------------------------------------------ */
static inline double dlp_pressure_pair(double r1, double r2, double rr, double xi2, double f1,
                                       double f2, double n1, double n2)
{
    double rf = r1 * f1 + r2 * f2, rn = r1 * n1 + r2 * n2, fn = f1 * n1 + f2 * n2;
    return -exp(-xi2 * rr) * ((fn - 2 * xi2 * rn * rf) / rr - 2 * rn * rf / rr / rr);
}

/* mex_stokes_dlp_pressure_real.cpp:100-192: the DLP's cells and pair set
 * (coincident points of the own cell skipped), pair term
 * -e^{-xi^2 r^2} ((f.n - 2 xi^2 (r.n)(r.f))/r^2 - 2 (r.n)(r.f)/r^4), / (2 pi). */
void orc_dlp_pressure_real(const double* psrc, const double* ptar, const double* f,
                           const double* n, int Ns, int Nt, double xi, int nside_x,
                           int nside_y, double Lx, double Ly, double* p)
{
    cells_t c;
    cells_build(&c, psrc, Ns, ptar, Nt, Lx, Ly, nside_x, nside_y);
    const int nbox = nside_x * nside_y;
    const double xi2 = xi * xi;
    const double cut2 = Lx * Ly / nside_x / nside_y;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        int bx = b % nside_x, by = b / nside_x;
        for (int it = c.bo_t[b]; it < c.bo_t[b + 1]; ++it) {
            int t = c.po_t[it];
            double xt = ptar[2 * t], yt = ptar[2 * t + 1];
            double a = 0;
            for (int is = c.bo_s[b]; is < c.bo_s[b + 1]; ++is) {
                int s = c.po_s[is];
                double r1 = xt - psrc[2 * s], r2 = yt - psrc[2 * s + 1];
                double rr = r1 * r1 + r2 * r2;
                if (rr == 0) continue;
                a += dlp_pressure_pair(r1, r2, rr, xi2, f[2 * s], f[2 * s + 1], n[2 * s],
                                       n[2 * s + 1]);
            }
            for (int j = 0; j < 8; ++j) {
                int qx = bx + nb_dx[j], qy = by + nb_dy[j];
                int wx = (qx + nside_x) % nside_x, wy = (qy + nside_y) % nside_y;
                int sb = wy * nside_x + wx;
                double ox = (Lx * (qx - wx)) / nside_x;
                double oy = (Ly * (qy - wy)) / nside_y;
                for (int is = c.bo_s[sb]; is < c.bo_s[sb + 1]; ++is) {
                    int s = c.po_s[is];
                    double r1 = xt - (psrc[2 * s] + ox), r2 = yt - (psrc[2 * s + 1] + oy);
                    double rr = r1 * r1 + r2 * r2;
                    if (rr < cut2)
                        a += dlp_pressure_pair(r1, r2, rr, xi2, f[2 * s], f[2 * s + 1],
                                               n[2 * s], n[2 * s + 1]);
                }
            }
            p[t] = a / (2 * ORC_PI);
        }
    }
    cells_free(&c);
}

/* mex_stokes_dlp_pressure_kspace.cpp:140-200: with h1 = f1 n1, h2 = f2 n1,
 * h3 = f1 n2, h4 = f2 n2 (:86-89), P(k) = -(k1^2 h1 + k1 k2 (h2 + h3) + k2^2 h4)
 * exp(-(1-eta) K^2/(4 xi^2)) / K^2 into h1; k = 0 set to zero. */
void orc_dlp_pressure_filter(double* h1re, double* h1im, const double* h2re,
                             const double* h2im, const double* h3re, const double* h3im,
                             const double* h4re, const double* h4im, int Mx, int My, double Lx,
                             double Ly, double xi, double eta)
{
    const double a = 0.25 / (xi * xi);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Mx; ++ix) {
        double k1 = wavenumber(ix, Mx, Lx);
        for (int iy = 0; iy < My; ++iy) {
            size_t o = (size_t)ix * My + iy;
            if (ix == 0 && iy == 0) {
                h1re[o] = h1im[o] = 0;
                continue;
            }
            double k2 = wavenumber(iy, My, Ly);
            double K2 = k1 * k1 + k2 * k2;
            double e = exp(-(1 - eta) * a * K2);
            double qr = k1 * k1 * h1re[o] + k1 * k2 * (h2re[o] + h3re[o]) + k2 * k2 * h4re[o];
            double qi = k1 * k1 * h1im[o] + k1 * k2 * (h2im[o] + h3im[o]) + k2 * k2 * h4im[o];
            h1re[o] = -qr * e / K2;
            h1im[o] = -qi * e / K2;
        }
    }
}

/* mex_stokes_dlp_pressure_kspace.cpp:10-230: spread v1 = (f1 n1, f2 n1) and
 * v2 = (f1 n2, f2 n2), four forward DFTs, filter, inverse DFT, gather.  The
 * zero mode is not added (StokesDLP_pressure_ewald_2p.m:156-157 adds it). */
void orc_dlp_pressure_kspace(const double* psrc, const double* ptar, const double* f,
                             const double* n, double xi, double eta, int Ns, int Nt, int Mx,
                             int My, double Lx, double Ly, double w, int P, double* pk)
{
    size_t G = (size_t)Mx * My;
    double* b = (double*)calloc(8 * G, sizeof(double));
    double* v1 = (double*)malloc(sizeof(double) * 2 * (size_t)(Ns > 0 ? Ns : 1));
    double* v2 = (double*)malloc(sizeof(double) * 2 * (size_t)(Ns > 0 ? Ns : 1));
    for (int i = 0; i < Ns; ++i) {
        v1[2 * i] = f[2 * i] * n[2 * i];
        v1[2 * i + 1] = f[2 * i + 1] * n[2 * i];
        v2[2 * i] = f[2 * i] * n[2 * i + 1];
        v2[2 * i + 1] = f[2 * i + 1] * n[2 * i + 1];
    }
    double *h1r = b, *h1i = b + G, *h2r = b + 2 * G, *h2i = b + 3 * G;
    double *h3r = b + 4 * G, *h3i = b + 5 * G, *h4r = b + 6 * G, *h4i = b + 7 * G;
    orc_spread(psrc, v1, Ns, Lx, Ly, xi, w, eta, P, Mx, My, h1r, h2r);
    orc_spread(psrc, v2, Ns, Lx, Ly, xi, w, eta, P, Mx, My, h3r, h4r);
    orc_fft2(h1r, h1i, My, Mx, 0);
    orc_fft2(h2r, h2i, My, Mx, 0);
    orc_fft2(h3r, h3i, My, Mx, 0);
    orc_fft2(h4r, h4i, My, Mx, 0);
    orc_dlp_pressure_filter(h1r, h1i, h2r, h2i, h3r, h3i, h4r, h4i, Mx, My, Lx, Ly, xi, eta);
    orc_fft2(h1r, h1i, My, Mx, 1);
    for (int m = 0; m < Nt; ++m) pk[m] = 0;
    orc_gather(h1r, 1, 0, ptar, Nt, Lx, Ly, xi, w, eta, P, Mx, My, pk);
    free(v1);
    free(v2);
    free(b);
}

/* StokesDLP_pressure_ewald_2p.m:156-157: -sum_s (n.f)_s / (2 Lx Ly). */
double orc_dlp_pressure_zero_mode(const double* f, const double* n, int Ns, double Lx,
                                  double Ly)
{
    double s = 0;
    for (int i = 0; i < Ns; ++i) s += n[2 * i] * f[2 * i] + n[2 * i + 1] * f[2 * i + 1];
    return -s / (2 * Lx * Ly);
}

/* stokes_dlp_pressure_real_ds.m: all 3 x 3 images, |r| < 1e-13 skipped, / (2 pi). */
void orc_dlp_pressure_real_ds(const double* psrc, const double* ptar, const double* f,
                              const double* n, int Ns, int Nt, double Lx, double Ly, double xi,
                              double* p)
{
    const double xi2 = xi * xi;
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double a = 0;
        for (int s = 0; s < Ns; ++s)
            for (int jx = -1; jx <= 1; ++jx)
                for (int jy = -1; jy <= 1; ++jy) {
                    double r1 = ptar[2 * m] - (psrc[2 * s] + jx * Lx);
                    double r2 = ptar[2 * m + 1] - (psrc[2 * s + 1] + jy * Ly);
                    double rr = r1 * r1 + r2 * r2;
                    if (sqrt(rr) < 1e-13) continue;
                    a += dlp_pressure_pair(r1, r2, rr, xi2, f[2 * s], f[2 * s + 1], n[2 * s],
                                           n[2 * s + 1]);
                }
        p[m] = a / (2 * ORC_PI);
    }
}

/* stokes_dlp_pressure_kspace_ds.m: real part of
 *   sum_{k != 0} -(k.f_s)(k.n_s) e^{-K^2/(4 xi^2)}/K^2 e^{i k.(x_t - x_s)} / (Lx Ly)
 * plus the zero mode, through S(k) = sum_s (k.f_s)(k.n_s) e^{-i k.x_s}. */
void orc_dlp_pressure_kspace_ds(const double* psrc, const double* ptar, const double* f,
                                const double* n, int Ns, int Nt, double Lx, double Ly,
                                double xi, int kinf, double* p)
{
    const int K = 2 * kinf;
    const size_t nm = (size_t)K * K;
    double* S = (double*)calloc(nm * 2, sizeof(double));
    const double a = 0.25 / (xi * xi);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)nm; ++q) {
        int j1 = (int)(q / K) - kinf + 1, j2 = (int)(q % K) - kinf + 1;
        double k1 = 2 * ORC_PI / Lx * j1, k2 = 2 * ORC_PI / Ly * j2;
        double* s = S + 2 * q;
        for (int i = 0; i < Ns; ++i) {
            double ph = -(k1 * psrc[2 * i] + k2 * psrc[2 * i + 1]);
            double w = (k1 * f[2 * i] + k2 * f[2 * i + 1]) * (k1 * n[2 * i] + k2 * n[2 * i + 1]);
            s[0] += w * cos(ph);
            s[1] += w * sin(ph);
        }
    }
    const double zm = orc_dlp_pressure_zero_mode(f, n, Ns, Lx, Ly);
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double acc = 0;
        for (size_t q = 0; q < nm; ++q) {
            int j1 = (int)(q / K) - kinf + 1, j2 = (int)(q % K) - kinf + 1;
            if (j1 == 0 && j2 == 0) continue;
            double k1 = 2 * ORC_PI / Lx * j1, k2 = 2 * ORC_PI / Ly * j2;
            double K2 = k1 * k1 + k2 * k2;
            double ph = k1 * ptar[2 * m] + k2 * ptar[2 * m + 1];
            const double* s = S + 2 * q;
            /* Re(-S e^{i ph}) g */
            acc -= exp(-a * K2) / K2 * (s[0] * cos(ph) - s[1] * sin(ph));
        }
        p[m] = acc / (Lx * Ly) + zm;
    }
    free(S);
}

/* ------------------------------------------------------------------------ */
/* Derived quantities: velocity gradient, pressure gradient, vorticity and   */
/* stress of both layers.  Kind codes 4..11 (gradient, pressure gradient,    */
/* vorticity, stress; Stokeslet then stresslet).  Tensors hold 4 values per  */
/* target in the mex order Ts[4j..4j+3].                                     */
/* ------------------------------------------------------------------------ */
int orc_kind_ncomp(int kind)
{
    switch (kind) {
    case 4: case 7: case 8: case 11: return 4;
    case 5: case 9: return 2;
    case 6: case 10: return 1;
    default: return 0;
    }
}

static int kind_is_dlp(int kind) { return kind >= 8; }

/* Pair terms of the real-space mex files, r = target - source, before the
 * output scaling; each line restates the file's own expression. */
static void derived_pair_orc(int kind, double r1, double r2, double rr, double xi, double f1,
                             double f2, double n1, double n2, double* T)
{
    const double xi2 = xi * xi;
    const double e2 = exp(-xi2 * rr);
    const double rdotf = r1 * f1 + r2 * f2;
    switch (kind) {
    case 4: /* mex_stokes_slp_gradient_real.cpp:118-132 */
        T[0] += e2 * (2 * xi * xi * r1 * f1 + rdotf / rr - 2 * r1 * r1 * rdotf * (xi * xi + 1 / rr) / rr);
        T[1] += e2 * (2 * xi * xi * r1 * f2 + (-r1 * f2 + r2 * f1) / rr - 2 * r1 * r2 * rdotf * (xi * xi + 1 / rr) / rr);
        T[2] += e2 * (2 * xi * xi * r2 * f1 + (r1 * f2 - r2 * f1) / rr - 2 * r1 * r2 * rdotf * (xi * xi + 1 / rr) / rr);
        T[3] += e2 * (2 * xi * xi * r2 * f2 + rdotf / rr - 2 * r2 * r2 * rdotf * (xi * xi + 1 / rr) / rr);
        break;
    case 5: /* mex_stokes_slp_pressure_grad_real.cpp:120-121 */
        T[0] -= e2 * ((f1 - 2 * xi * xi * rdotf * r1) / rr - 2 * rdotf * r1 / rr / rr);
        T[1] -= e2 * ((f2 - 2 * xi * xi * rdotf * r2) / rr - 2 * rdotf * r2 / rr / rr);
        break;
    case 6: /* mex_stokes_slp_vorticity_real.cpp:118-121 */
        T[0] += e2 * (1 / rr - xi2) * (f1 * r2 - f2 * r1);
        break;
    case 7: /* mex_stokes_slp_stress_real.cpp:118-129 */
        T[0] += e2 * (2 * xi2 * (rdotf + r1 * f1 + r1 * f1) - 4 * r1 * r1 * rdotf / (rr * rr) * (1 + xi2 * rr));
        T[1] += e2 * (2 * xi2 * (r2 * f1 + r1 * f2) - 4 * r1 * r2 * rdotf / (rr * rr) * (1 + xi2 * rr));
        T[2] += e2 * (2 * xi2 * (r1 * f2 + r2 * f1) - 4 * r2 * r1 * rdotf / (rr * rr) * (1 + xi2 * rr));
        T[3] += e2 * (2 * xi2 * (rdotf + r2 * f2 + r2 * f2) - 4 * r2 * r2 * rdotf / (rr * rr) * (1 + xi2 * rr));
        break;
    default: {
        const double rdotn = r1 * n1 + r2 * n2, fdotn = f1 * n1 + f2 * n2;
        const double c6 = 8 * xi2 * xi2 / rr + 16 * xi2 / rr / rr + 16 / rr / rr / rr;
        if (kind == 8) { /* mex_stokes_dlp_gradient_real.cpp:130-145 */
            T[0] += e2 * (r1 * r1 * rdotf * rdotn * c6 - 4 * (1 + xi2 * rr) * (rdotf * rdotn + r1 * f1 * rdotn + r1 * n1 * rdotf) / rr / rr
                          + 2 * xi2 * (2 * f1 * n1 + fdotn - 2 * xi2 * r1 * (f1 * rdotn + n1 * rdotf + r1 * fdotn)));
            T[1] += e2 * (r1 * r2 * rdotf * rdotn * c6 - 4 * (1 + xi2 * rr) * (r2 * f1 * rdotn + r2 * n1 * rdotf) / rr / rr
                          + 2 * xi2 * (f1 * n2 + n1 * f2 - 2 * xi2 * r1 * (f2 * rdotn + n2 * rdotf + r2 * fdotn)));
            T[2] += e2 * (r1 * r2 * rdotf * rdotn * c6 - 4 * (1 + xi2 * rr) * (r1 * f2 * rdotn + r1 * n2 * rdotf) / rr / rr
                          + 2 * xi2 * (f2 * n1 + f1 * n2 - 2 * xi2 * r2 * (f1 * rdotn + n1 * rdotf + r1 * fdotn)));
            T[3] += e2 * (r2 * r2 * rdotf * rdotn * c6 - 4 * (1 + xi2 * rr) * (rdotf * rdotn + r2 * f2 * rdotn + r2 * n2 * rdotf) / rr / rr
                          + 2 * xi2 * (2 * f2 * n2 + fdotn - 2 * xi2 * r2 * (f2 * rdotn + n2 * rdotf + r2 * fdotn)));
        } else if (kind == 9) { /* mex_stokes_dlp_pressure_grad_real.cpp:122-128 */
            T[0] += e2 * (2 * xi2 * (r1 * fdotn + f1 * rdotn + n1 * rdotf - 2 * xi2 * r1 * rdotf * rdotn) / rr
                          + 2 * (r1 * fdotn + f1 * rdotn + n1 * rdotf - 4 * xi2 * r1 * rdotn * rdotf) / rr / rr
                          - 8 * r1 * rdotn * rdotf / rr / rr / rr);
            T[1] += e2 * (2 * xi2 * (r2 * fdotn + f2 * rdotn + n2 * rdotf - 2 * xi2 * r2 * rdotf * rdotn) / rr
                          + 2 * (r2 * fdotn + f2 * rdotn + n2 * rdotf - 4 * xi2 * r2 * rdotn * rdotf) / rr / rr
                          - 8 * r2 * rdotn * rdotf / rr / rr / rr);
        } else if (kind == 10) { /* mex_stokes_dlp_vorticity_real.cpp:136-145 */
            double rdot = r1 * (n2 * rdotf + f2 * rdotn) - r2 * (n1 * rdotf + f1 * rdotn);
            double ndot = n1 * (r2 * rdotf) - n2 * (r1 * rdotf);
            double fdot = f1 * (r2 * rdotn) - f2 * (r1 * rdotn);
            T[0] += e2 * ((1 + xi2 * rr) * rdot / (rr * rr) + xi2 * xi2 * (ndot + fdot));
        } else { /* mex_stokes_dlp_stress_real.cpp:132-151, mu = 1 */
            const double mu = 1, pi = ORC_PI;
            T[0] += e2 * (1 / (2 * pi) * ((fdotn - 2 * xi2 * rdotf * rdotn) / rr - 2 * rdotf * rdotn / (rr * rr))
                          + 1 / (4 * pi) * 2 * mu * (r1 * r1 * rdotf * rdotn * c6
                          - 2 * (1 + xi2 * rr) * (f1 * r1 * rdotn + n1 * r1 * rdotf + r1 * f1 * rdotn + r1 * n1 * rdotf + 2 * rdotf * rdotn) / (rr * rr)
                          + 2 * xi2 * (fdotn + n1 * f1 + f1 * n1 - xi2 * (2 * r1 * r1 * fdotn + r1 * f1 * rdotn + r1 * n1 * rdotf + f1 * r1 * rdotn + n1 * r1 * rdotf))));
            T[1] += 1 / (4 * pi) * 2 * mu * e2 * (r1 * r2 * rdotf * rdotn * c6
                          - 2 * (1 + xi2 * rr) * (f1 * r2 * rdotn + n1 * r2 * rdotf + r1 * f2 * rdotn + r1 * n2 * rdotf) / (rr * rr)
                          + 2 * xi2 * (n1 * f2 + f1 * n2 - xi2 * (2 * r1 * r2 * fdotn + r1 * f2 * rdotn + r1 * n2 * rdotf + f1 * r2 * rdotn + n1 * r2 * rdotf)));
            T[2] += 1 / (4 * pi) * 2 * mu * e2 * (r2 * r1 * rdotf * rdotn * c6
                          - 2 * (1 + xi2 * rr) * (f2 * r1 * rdotn + n2 * r1 * rdotf + r2 * f1 * rdotn + r2 * n1 * rdotf) / (rr * rr)
                          + 2 * xi2 * (n2 * f1 + f2 * n1 - xi2 * (2 * r2 * r1 * fdotn + r2 * f1 * rdotn + r2 * n1 * rdotf + f2 * r1 * rdotn + n2 * r1 * rdotf)));
            T[3] += e2 * (1 / (2 * pi) * ((fdotn - 2 * xi2 * rdotf * rdotn) / rr - 2 * rdotf * rdotn / (rr * rr))
                          + 1 / (4 * pi) * 2 * mu * (r2 * r2 * rdotf * rdotn * c6
                          - 2 * (1 + xi2 * rr) * (f2 * r2 * rdotn + n2 * r2 * rdotf + r2 * f2 * rdotn + r2 * n2 * rdotf + 2 * rdotf * rdotn) / (rr * rr)
                          + 2 * xi2 * (fdotn + n2 * f2 + f2 * n2 - xi2 * (2 * r2 * r2 * fdotn + r2 * f2 * rdotn + r2 * n2 * rdotf + f2 * r2 * rdotn + n2 * r2 * rdotf))));
        }
    }
    }
}

/* Output scaling at the end of each real-space mex (e.g. /4/pi at
 * mex_stokes_slp_gradient_real.cpp:202-205, /(2*pi) at
 * mex_stokes_slp_pressure_grad_real.cpp:173-174, /pi at
 * mex_stokes_dlp_vorticity_real.cpp:209; the stresslet stress keeps the
 * factors inside its pair term). */
static double derived_out_scale(int kind)
{
    switch (kind) {
    case 4: case 7: case 8: return 1 / (4 * ORC_PI);
    case 5: case 6: case 9: return 1 / (2 * ORC_PI);
    case 10: return 1 / ORC_PI;
    default: return 1;
    }
}

/* The real-space mex files of the derived quantities: the cells, stencil and
 * cutoff of mex_stokes_dlp_real.cpp (own cell uncut, coincident points skipped;
 * the stresslet vorticity skips own-cell pairs with r^2 < 1e-13,
 * mex_stokes_dlp_vorticity_real.cpp:128).  out: nc x Nt. */
void orc_derived_real(int kind, const double* psrc, const double* ptar, const double* f,
                      const double* n, int Ns, int Nt, double xi, int nside_x, int nside_y,
                      double Lx, double Ly, double* out)
{
    const int nc = orc_kind_ncomp(kind);
    const int dlp = kind_is_dlp(kind);
    cells_t c;
    cells_build(&c, psrc, Ns, ptar, Nt, Lx, Ly, nside_x, nside_y);
    const int nbox = nside_x * nside_y;
    const double cut2 = Lx * Ly / nside_x / nside_y;
    const double own_skip = kind == 10 ? 1e-13 : 0.0;
    const double sc = derived_out_scale(kind);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        int bx = b % nside_x, by = b / nside_x;
        for (int it = c.bo_t[b]; it < c.bo_t[b + 1]; ++it) {
            int t = c.po_t[it];
            double xt = ptar[2 * t], yt = ptar[2 * t + 1];
            double T[4] = {0, 0, 0, 0};
            for (int is = c.bo_s[b]; is < c.bo_s[b + 1]; ++is) {
                int s = c.po_s[is];
                double r1 = xt - psrc[2 * s], r2 = yt - psrc[2 * s + 1];
                double rr = r1 * r1 + r2 * r2;
                if (kind == 10 ? rr < own_skip : rr == 0) continue;
                derived_pair_orc(kind, r1, r2, rr, xi, f[2 * s], f[2 * s + 1],
                                 dlp ? n[2 * s] : 0, dlp ? n[2 * s + 1] : 0, T);
            }
            for (int j = 0; j < 8; ++j) {
                int qx = bx + nb_dx[j], qy = by + nb_dy[j];
                int wx = (qx + nside_x) % nside_x, wy = (qy + nside_y) % nside_y;
                int sb = wy * nside_x + wx;
                double ox = (Lx * (qx - wx)) / nside_x;
                double oy = (Ly * (qy - wy)) / nside_y;
                for (int is = c.bo_s[sb]; is < c.bo_s[sb + 1]; ++is) {
                    int s = c.po_s[is];
                    double r1 = xt - (psrc[2 * s] + ox), r2 = yt - (psrc[2 * s + 1] + oy);
                    double rr = r1 * r1 + r2 * r2;
                    if (rr < cut2)
                        derived_pair_orc(kind, r1, r2, rr, xi, f[2 * s], f[2 * s + 1],
                                         dlp ? n[2 * s] : 0, dlp ? n[2 * s + 1] : 0, T);
                }
            }
            for (int q = 0; q < nc; ++q) out[(size_t)nc * t + q] = T[q] * sc;
        }
    }
    cells_free(&c);
}

/* The k-space filters of the derived-quantity mex files, on one mode.  in:
 * the forward spectra as the mex names them (Stokeslet: Hhat1, Hhat2 = spread
 * f1, f2; stresslet: Hhat1..4 = spread f1 n1, f2 n1, f1 n2, f2 n2, which the
 * files call f1n1, f1n2, f2n1, f2n2 -- the middle two enter only through their
 * sum).  out: the nc filtered spectra. */
static void derived_filter_mode(int kind, double k1, double k2, double xi, double eta,
                                const double complex* H, double complex* o)
{
    const double Ksq = k1 * k1 + k2 * k2;
    const double xi2 = xi * xi;
    const double complex I1 = I;
    if (!kind_is_dlp(kind)) {
        const double complex q1 = H[0], q2 = H[1];
        const double complex kdotq = k1 * q1 + k2 * q2;
        if (kind == 4) { /* mex_stokes_slp_gradient_kspace.cpp:143-172 */
            double e = (1.0 / (Ksq * Ksq) + 0.25 / (Ksq * xi * xi)) * exp(-0.25 * (1 - eta) / (xi * xi) * Ksq);
            o[0] = I1 * k1 * (Ksq * q1 - k1 * kdotq) * e;
            o[1] = I1 * k1 * (Ksq * q2 - k2 * kdotq) * e;
            o[2] = I1 * k2 * (Ksq * q1 - k1 * kdotq) * e;
            o[3] = I1 * k2 * (Ksq * q2 - k2 * kdotq) * e;
        } else if (kind == 5) { /* mex_stokes_slp_pressure_grad_kspace.cpp:126-135 */
            double e = exp(-Ksq * (1 - eta) / (4 * xi * xi));
            o[0] = -k1 * kdotq * e / Ksq;
            o[1] = -k2 * kdotq * e / Ksq;
        } else if (kind == 6) { /* mex_stokes_slp_vorticity_kspace.cpp:142-156 */
            double e = (1.0 / Ksq + 0.25 / (xi * xi)) * exp(-0.25 * (1 - eta) / (xi * xi) * Ksq);
            double complex fdotkperp = q1 * k2 - q2 * k1;
            o[0] = -I1 * fdotkperp * e;
        } else { /* mex_stokes_slp_stress_kspace.cpp:170-185 */
            double e = exp(-0.25 * (1 - eta) / xi2 * Ksq) * (1.0 + Ksq / (4 * xi2)) / Ksq;
            o[0] = I1 * (kdotq + q1 * k1 + q1 * k1 - 2 * (k1 * k1 * kdotq) / Ksq) * e;
            o[1] = I1 * (q2 * k1 + q1 * k2 - 2 * (k2 * k1 * kdotq) / Ksq) * e;
            o[2] = I1 * (q1 * k2 + q2 * k1 - 2 * (k1 * k2 * kdotq) / Ksq) * e;
            o[3] = I1 * (kdotq + q2 * k2 + q2 * k2 - 2 * (k2 * k2 * kdotq) / Ksq) * e;
        }
        return;
    }
    const double complex f1n1 = H[0], f1n2 = H[1], f2n1 = H[2], f2n2 = H[3];
    const double complex kqk = k1 * k1 * f1n1 + k1 * k2 * (f1n2 + f2n1) + k2 * k2 * f2n2;
    if (kind == 8) { /* mex_stokes_dlp_gradient_kspace.cpp:167-206 */
        double e = exp(-Ksq * (1 - eta) / (4 * xi * xi)) * (1 + Ksq / (4 * xi * xi)) / Ksq;
        double complex v1 = 2 * f1n1 * k1 + k2 * (f1n2 + f2n1) + k1 * (f1n1 + f2n2) - 2 * k1 * kqk / Ksq;
        double complex v2 = 2 * f2n2 * k2 + k1 * (f1n2 + f2n1) + k2 * (f1n1 + f2n2) - 2 * k2 * kqk / Ksq;
        o[0] = -k1 * v1 * e;
        o[1] = -k1 * v2 * e;
        o[2] = -k2 * v1 * e;
        o[3] = -k2 * v2 * e;
    } else if (kind == 9) { /* mex_stokes_dlp_pressure_grad_kspace.cpp:170-173 */
        double e = exp(-Ksq * (1 - eta) / (4 * xi * xi));
        o[0] = -I1 * k1 * kqk * e / Ksq;
        o[1] = -I1 * k2 * kqk * e / Ksq;
    } else if (kind == 10) { /* mex_stokes_dlp_vorticity_kspace.cpp:160-169 */
        double e = (1.0 / Ksq + 1 / (4 * xi2)) * exp(-Ksq * (1 - eta) / (4 * xi2));
        double complex t1 = k1 * k2 * f1n1 + k2 * k2 * f1n2 - (k1 * k1 * f2n1 + k1 * k2 * f2n2);
        double complex t2 = k1 * k2 * f1n1 + k2 * k2 * f2n1 - (k1 * k1 * f1n2 + k1 * k2 * f2n2);
        o[0] = (t1 + t2) * e;
    } else { /* mex_stokes_dlp_stress_kspace.cpp:166-201, mu = 1 */
        const double mu = 1;
        double e = exp(-Ksq * (1 - eta) / (4 * xi * xi));
        double c = 1 / Ksq + 0.25 / (xi * xi);
        double complex a1 = k1 * f1n1 + k2 * f1n2, a2 = k1 * f1n1 + k2 * f2n1;
        double complex b1 = k1 * f2n1 + k2 * f2n2, b2 = k1 * f1n2 + k2 * f2n2;
        o[0] = e * (kqk / Ksq - mu * c * (2 * k1 * k1 * (f1n1 + f2n2) + k1 * a1 + k1 * a2 + k1 * a1 + k1 * a2
                                          - 4 * k1 * k1 * kqk / Ksq));
        o[1] = -e * mu * c * (2 * k2 * k1 * (f1n1 + f2n2) + k2 * a1 + k2 * a2 + k1 * b1 + k1 * b2
                              - 4 * k2 * k1 * kqk / Ksq);
        o[2] = -e * mu * c * (2 * k1 * k2 * (f1n1 + f2n2) + k1 * b1 + k1 * b2 + k2 * a1 + k2 * a2
                              - 4 * k1 * k2 * kqk / Ksq);
        o[3] = e * (kqk / Ksq - mu * c * (2 * k2 * k2 * (f1n1 + f2n2) + k2 * b1 + k2 * b2 + k2 * b1 + k2 * b2
                                          - 4 * k2 * k2 * kqk / Ksq));
    }
}

/* The k-space mex files of the derived quantities: spread (f, or f n1 and
 * f n2), forward DFT of each real grid, the filter on every mode with k = 0
 * set to zero, inverse DFT of each output, gather of its real part into
 * component c (Gather(Ht_c, nc, c, ...)).  No zero mode.  out: nc x Nt. */
void orc_derived_kspace(int kind, const double* psrc, const double* ptar, const double* f,
                        const double* n, double xi, double eta, int Ns, int Nt, int Mx, int My,
                        double Lx, double Ly, double w, int P, double* out)
{
    const int nc = orc_kind_ncomp(kind);
    const int nin = kind_is_dlp(kind) ? 4 : 2;
    const size_t G = (size_t)Mx * My;
    double* hin = (double*)calloc(2 * nin * G, sizeof(double));
    double* hout = (double*)calloc(2 * nc * G, sizeof(double));
    if (kind_is_dlp(kind)) {
        double* v1 = (double*)malloc(sizeof(double) * 2 * (size_t)(Ns > 0 ? Ns : 1));
        double* v2 = (double*)malloc(sizeof(double) * 2 * (size_t)(Ns > 0 ? Ns : 1));
        for (int i = 0; i < Ns; ++i) {
            v1[2 * i] = f[2 * i] * n[2 * i];
            v1[2 * i + 1] = f[2 * i + 1] * n[2 * i];
            v2[2 * i] = f[2 * i] * n[2 * i + 1];
            v2[2 * i + 1] = f[2 * i + 1] * n[2 * i + 1];
        }
        orc_spread(psrc, v1, Ns, Lx, Ly, xi, w, eta, P, Mx, My, hin, hin + 2 * G);
        orc_spread(psrc, v2, Ns, Lx, Ly, xi, w, eta, P, Mx, My, hin + 4 * G, hin + 6 * G);
        free(v1);
        free(v2);
    } else {
        orc_spread(psrc, f, Ns, Lx, Ly, xi, w, eta, P, Mx, My, hin, hin + 2 * G);
    }
    for (int q = 0; q < nin; ++q) orc_fft2(hin + 2 * q * G, hin + (2 * q + 1) * G, My, Mx, 0);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Mx; ++ix) {
        double k1 = wavenumber(ix, Mx, Lx);
        for (int iy = 0; iy < My; ++iy) {
            size_t o = (size_t)ix * My + iy;
            if (ix == 0 && iy == 0) continue; /* zero frequency removed */
            double k2 = wavenumber(iy, My, Ly);
            double complex H[4], out4[4];
            for (int q = 0; q < nin; ++q) H[q] = hin[2 * q * G + o] + I * hin[(2 * q + 1) * G + o];
            derived_filter_mode(kind, k1, k2, xi, eta, H, out4);
            for (int q = 0; q < nc; ++q) {
                hout[2 * q * G + o] = creal(out4[q]);
                hout[(2 * q + 1) * G + o] = cimag(out4[q]);
            }
        }
    }
    for (int m = 0; m < nc * Nt; ++m) out[m] = 0;
    for (int q = 0; q < nc; ++q) {
        orc_fft2(hout + 2 * q * G, hout + (2 * q + 1) * G, My, Mx, 1);
        orc_gather(hout + 2 * q * G, nc, q, ptar, Nt, Lx, Ly, xi, w, eta, P, Mx, My, out);
    }
    free(hin);
    free(hout);
}

/* Real-space direct sums, tests/direct_sums/matlab_sums/stokes_*_real_ds.m:
 * all 3 x 3 images, |r| < 1e-13 skipped; gradient and stress projected on the
 * per-target vector b (2 x Nt), the others as they are.  out: 2 x Nt (1 x Nt
 * for the vorticity). */
void orc_derived_real_ds(int kind, const double* psrc, const double* ptar, const double* f,
                         const double* n, const double* b, int Ns, int Nt, double Lx, double Ly,
                         double xi, double* out)
{
    const double xi2 = xi * xi, pi = ORC_PI;
    const int no = kind == 6 || kind == 10 ? 1 : 2;
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double acc[2] = {0, 0};
        const double b1 = b ? b[2 * m] : 0, b2 = b ? b[2 * m + 1] : 0;
        for (int s = 0; s < Ns; ++s)
            for (int jx = -1; jx <= 1; ++jx)
                for (int jy = -1; jy <= 1; ++jy) {
                    double r1 = ptar[2 * m] - (psrc[2 * s] + jx * Lx);
                    double r2 = ptar[2 * m + 1] - (psrc[2 * s + 1] + jy * Ly);
                    double rr = r1 * r1 + r2 * r2, r = sqrt(rr);
                    if (r < 1e-13) continue;
                    double f1 = f[2 * s], f2 = f[2 * s + 1];
                    double n1 = n ? n[2 * s] : 0, n2 = n ? n[2 * s + 1] : 0;
                    double rdotf = r1 * f1 + r2 * f2, rdotb = r1 * b1 + r2 * b2;
                    double bdotf = f1 * b1 + f2 * b2, e = exp(-xi2 * rr);
                    double rdotn = r1 * n1 + r2 * n2, fdotn = n1 * f1 + n2 * f2;
                    double bdotn = n1 * b1 + n2 * b2;
                    switch (kind) {
                    case 4: /* stokes_slp_gradient_real_ds.m:59-75 */
                        acc[0] += e * (2 * xi2 * rdotb * f1 + (b1 * rdotf + r1 * bdotf - f1 * rdotb) / rr
                                       - 2 * r1 * rdotb * rdotf * (xi2 + 1 / rr) / rr);
                        acc[1] += e * (2 * xi2 * rdotb * f2 + (b2 * rdotf + r2 * bdotf - f2 * rdotb) / rr
                                       - 2 * r2 * rdotb * rdotf * (xi2 + 1 / rr) / rr);
                        break;
                    case 5: /* stokes_slp_pressure_grad_real_ds.m:57-64 */
                        acc[0] += e * ((f1 - r1 * 2 * xi2 * rdotf) / rr - 2 * r1 * rdotf / (rr * rr));
                        acc[1] += e * ((f2 - r2 * 2 * xi2 * rdotf) / rr - 2 * r2 * rdotf / (rr * rr));
                        break;
                    case 6: /* stokes_slp_vorticity_real_ds.m:55-63 */
                        acc[0] += e * (1 / rr - xi2) * (f1 * r2 - f2 * r1) / (2 * pi);
                        break;
                    case 7: /* stokes_slp_stress_real_ds.m:61-78 */
                        acc[0] += e * (2 * xi2 * (b1 * rdotf + f1 * rdotb + r1 * bdotf)
                                       - (4 * r1 * rdotb * rdotf / (rr * rr)) * (1 + xi2 * rr));
                        acc[1] += e * (2 * xi2 * (b2 * rdotf + f2 * rdotb + r2 * bdotf)
                                       - (4 * r2 * rdotb * rdotf / (rr * rr)) * (1 + xi2 * rr));
                        break;
                    case 8: { /* stokes_dlp_gradient_real_ds.m:61-82 */
                        double c6 = 8 * xi2 * xi2 / rr + 16 * xi2 / (rr * rr) + 16 / (rr * rr * rr);
                        acc[0] += e * (r1 * rdotf * rdotn * rdotb * c6
                                       - 4 * (1 + xi2 * rr) * (b1 * rdotf * rdotn + r1 * bdotf * rdotn + r1 * bdotn * rdotf) / (rr * rr)
                                       + 2 * xi2 * (f1 * bdotn + n1 * bdotf + b1 * fdotn - 2 * xi2 * rdotb * (f1 * rdotn + n1 * rdotf + r1 * fdotn)));
                        acc[1] += e * (r2 * rdotf * rdotn * rdotb * c6
                                       - 4 * (1 + xi2 * rr) * (b2 * rdotf * rdotn + r2 * bdotf * rdotn + r2 * bdotn * rdotf) / (rr * rr)
                                       + 2 * xi2 * (f2 * bdotn + n2 * bdotf + b2 * fdotn - 2 * xi2 * rdotb * (f2 * rdotn + n2 * rdotf + r2 * fdotn)));
                        break;
                    }
                    case 9: /* stokes_dlp_pressure_grad_real_ds.m:63-76 */
                        acc[0] += e * (2 * xi2 * (r1 * fdotn + f1 * rdotn + n1 * rdotf - 2 * xi2 * r1 * rdotn * rdotf) / rr
                                       + 2 * (-4 * xi2 * r1 * rdotn * rdotf + r1 * fdotn + f1 * rdotn + n1 * rdotf) / (rr * rr)
                                       - 8 * r1 * rdotn * rdotf / (rr * rr * rr));
                        acc[1] += e * (2 * xi2 * (r2 * fdotn + f2 * rdotn + n2 * rdotf - 2 * xi2 * r2 * rdotn * rdotf) / rr
                                       + 2 * (-4 * xi2 * r2 * rdotn * rdotf + r2 * fdotn + f2 * rdotn + n2 * rdotf) / (rr * rr)
                                       - 8 * r2 * rdotn * rdotf / (rr * rr * rr));
                        break;
                    case 10: { /* stokes_dlp_vorticity_real_ds.m:57-86 */
                        double rdot = r1 * (n2 * rdotf + f2 * rdotn) - r2 * (n1 * rdotf + f1 * rdotn);
                        double ndot = n1 * (r2 * rdotf) - n2 * (r1 * rdotf);
                        double fdot = f1 * (r2 * rdotn) - f2 * (r1 * rdotn);
                        acc[0] += e * ((1 + xi2 * rr) * rdot / (rr * rr) + xi2 * xi2 * (ndot + fdot)) / pi;
                        break;
                    }
                    default: { /* stokes_dlp_stress_real_ds.m:58-84, mu = 1 */
                        double fdotb = bdotf, mu = 1;
                        double c6 = (8 * xi2 * xi2) / rr + (16 * xi2) / (rr * rr) + 16 / (rr * rr * rr);
                        double ptmp = (fdotn - 2 * xi2 * rdotf * rdotn) / rr - 2 * rdotf * rdotn / (rr * rr);
                        acc[0] += e * (b1 / (2 * pi) * ptmp + 1 / (4 * pi) * 2 * mu * (r1 * rdotb * rdotf * rdotn * c6
                                  - 2 * (1 + xi2 * rr) * (f1 * rdotb * rdotn + n1 * rdotb * rdotf + r1 * fdotb * rdotn + r1 * bdotn * rdotf + 2 * b1 * rdotf * rdotn) / (rr * rr)
                                  + 2 * xi2 * (b1 * fdotn + n1 * fdotb + f1 * bdotn - xi2 * (2 * r1 * rdotb * fdotn + r1 * rdotn * fdotb + r1 * rdotf * bdotn + f1 * rdotb * rdotn + n1 * rdotb * rdotf))));
                        acc[1] += e * (b2 / (2 * pi) * ptmp + 1 / (4 * pi) * 2 * mu * (r2 * rdotb * rdotf * rdotn * c6
                                  - 2 * (1 + xi2 * rr) * (f2 * rdotb * rdotn + n2 * rdotb * rdotf + r2 * fdotb * rdotn + r2 * bdotn * rdotf + 2 * b2 * rdotf * rdotn) / (rr * rr)
                                  + 2 * xi2 * (b2 * fdotn + n2 * fdotb + f2 * bdotn - xi2 * (2 * r2 * rdotb * fdotn + r2 * rdotn * fdotb + r2 * rdotf * bdotn + f2 * rdotb * rdotn + n2 * rdotb * rdotf))));
                    }
                    }
                }
        double sc = kind == 4 || kind == 7 || kind == 8 ? 1 / (4 * pi)
                    : kind == 5 ? -1 / (2 * pi) : kind == 9 ? 1 / (2 * pi) : 1.0;
        for (int q = 0; q < no; ++q) out[no * m + q] = acc[q] * sc;
    }
}

/* k-space direct sums, stokes_*_kspace_ds.m: modes |j_i| <= kinf (j from
 * -kinf+1 to kinf), k = 0 skipped, real part / (Lx Ly); the stresslet stress
 * adds sum(n.f)/(2 Lx Ly) to both values (stokes_dlp_stress_kspace_ds.m:67).
 * Evaluated pair by pair (Ns Nt (2 kinf)^2 terms): small cases only. */
void orc_derived_kspace_ds(int kind, const double* psrc, const double* ptar, const double* f,
                           const double* n, const double* b, int Ns, int Nt, double Lx,
                           double Ly, double xi, int kinf, double* out)
{
    const double xi2 = xi * xi;
    const int no = kind == 6 || kind == 10 ? 1 : 2;
    double zm = 0;
    if (kind == 11)
        for (int s = 0; s < Ns; ++s) zm += n[2 * s] * f[2 * s] + n[2 * s + 1] * f[2 * s + 1];
    zm /= 2 * Lx * Ly;
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nt; ++m) {
        double complex acc[2] = {0, 0};
        const double b1 = b ? b[2 * m] : 0, b2 = b ? b[2 * m + 1] : 0;
        for (int s = 0; s < Ns; ++s) {
            double r1 = ptar[2 * m] - psrc[2 * s], r2 = ptar[2 * m + 1] - psrc[2 * s + 1];
            double f1 = f[2 * s], f2 = f[2 * s + 1];
            double n1 = n ? n[2 * s] : 0, n2 = n ? n[2 * s + 1] : 0;
            for (int j1 = -kinf + 1; j1 <= kinf; ++j1)
                for (int j2 = -kinf + 1; j2 <= kinf; ++j2) {
                    if (j1 == 0 && j2 == 0) continue;
                    double k1 = 2 * ORC_PI / Lx * j1, k2 = 2 * ORC_PI / Ly * j2;
                    double kSq = k1 * k1 + k2 * k2;
                    double kdotf = k1 * f1 + k2 * f2, kdotb = k1 * b1 + k2 * b2;
                    double kdotn = k1 * n1 + k2 * n2, fdotn = f1 * n1 + f2 * n2;
                    double bdotf = f1 * b1 + f2 * b2, ndotb = n1 * b1 + n2 * b2;
                    double g = exp(-kSq / (4 * xi2));
                    double complex t[2] = {0, 0};
                    switch (kind) {
                    case 4: /* stokes_slp_gradient_kspace_ds.m:66-81 */
                        t[0] = I * (f1 * kdotb / kSq - k1 * kdotb * kdotf / (kSq * kSq)) * (1 + kSq / (4 * xi2)) * g;
                        t[1] = I * (f2 * kdotb / kSq - k2 * kdotb * kdotf / (kSq * kSq)) * (1 + kSq / (4 * xi2)) * g;
                        break;
                    case 5: /* stokes_slp_pressure_grad_kspace_ds.m:64-72 */
                        t[0] = -k1 * kdotf * g / kSq;
                        t[1] = -k2 * kdotf * g / kSq;
                        break;
                    case 6: /* stokes_slp_vorticity_kspace_ds.m:64-70 */
                        t[0] = -I * (f1 * k2 - f2 * k1) * (1 / kSq + 1 / (4 * xi2)) * g;
                        break;
                    case 7: { /* stokes_slp_stress_kspace_ds.m:66-86 */
                        double e = g * (1 + kSq / (4 * xi2)) / kSq;
                        t[0] = I * (b1 * kdotf + f1 * kdotb + k1 * bdotf - 2 * (k1 * kdotb * kdotf) / kSq) * e;
                        t[1] = I * (b2 * kdotf + f2 * kdotb + k2 * bdotf - 2 * (k2 * kdotb * kdotf) / kSq) * e;
                        break;
                    }
                    case 8: /* stokes_dlp_gradient_kspace_ds.m:76-92 */
                        t[0] = -kdotb * (f1 * kdotn + n1 * kdotf + k1 * fdotn - 2 * k1 * kdotn * kdotf / kSq) * (1 + kSq / (4 * xi2)) * g / kSq;
                        t[1] = -kdotb * (f2 * kdotn + n2 * kdotf + k2 * fdotn - 2 * k2 * kdotn * kdotf / kSq) * (1 + kSq / (4 * xi2)) * g / kSq;
                        break;
                    case 9: /* stokes_dlp_pressure_grad_kspace_ds.m:69-81 */
                        t[0] = -I * k1 * kdotn * kdotf * g / kSq;
                        t[1] = -I * k2 * kdotn * kdotf * g / kSq;
                        break;
                    case 10: { /* stokes_dlp_vorticity_kspace_ds.m:67-95 */
                        double fdot = f1 * k2 * kdotn - f2 * k1 * kdotn;
                        double ndot = n1 * k2 * kdotf - n2 * k1 * kdotf;
                        t[0] = (fdot + ndot) * (1 / kSq + 1 / (4 * xi2)) * g;
                        break;
                    }
                    default: { /* stokes_dlp_stress_kspace_ds.m:71-97, mu = 1 */
                        double mu = 1, fdotb = bdotf;
                        t[0] = g * (b1 * kdotf * kdotn / kSq - mu * (1 / kSq + 1 / (4 * xi2)) * (2 * k1 * kdotb * fdotn
                                    + k1 * kdotn * fdotb + k1 * kdotf * ndotb + f1 * kdotb * kdotn + n1 * kdotb * kdotf
                                    - 4 * k1 * kdotb * kdotf * kdotn / kSq));
                        t[1] = g * (b2 * kdotf * kdotn / kSq - mu * (1 / kSq + 1 / (4 * xi2)) * (2 * k2 * kdotb * fdotn
                                    + k2 * kdotn * fdotb + k2 * kdotf * ndotb + f2 * kdotb * kdotn + n2 * kdotb * kdotf
                                    - 4 * k2 * kdotb * kdotf * kdotn / kSq));
                    }
                    }
                    double complex ph = cexp(I * (k1 * r1 + k2 * r2));
                    acc[0] += t[0] * ph;
                    acc[1] += t[1] * ph;
                }
        }
        for (int q = 0; q < no; ++q) out[no * m + q] = creal(acc[q]) / (Lx * Ly) + zm;
    }
}

/* ------------------------------------------------------------------------ */
/* Free-space layer potentials: the direct sums the reference's FMM tests    */
/* check against (FMM/tests/stokes_SLP_FMM_test.m:29-51,                     */
/* stokes_DLP_FMM_test.m:17-35).                                             */
/* ------------------------------------------------------------------------ */
/* u_t = sum_s (-1/2 log rho^2 f_s + (r.f_s) r/rho^2)/(4 pi), r = x_t - x_s;
 * self: leave out s = t (the test's srcEqualsTar); pairs with rho = 0 skipped. */
void orc_free_slp(const double* f, const double* psrc, int Ns, const double* ptar, int Nt,
                  int self, double* u)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nt; ++k) {
        double ud = 0, vd = 0;
        for (int s = 0; s < Ns; ++s) {
            if (self && s == k) continue;
            double rx = ptar[2 * k] - psrc[2 * s], ry = ptar[2 * k + 1] - psrc[2 * s + 1];
            double rho2 = rx * rx + ry * ry;
            if (rho2 == 0) continue;
            double rdotf = rx * f[2 * s] + ry * f[2 * s + 1];
            ud += -0.5 * log(rho2) * f[2 * s] + rdotf / rho2 * rx;
            vd += -0.5 * log(rho2) * f[2 * s + 1] + rdotf / rho2 * ry;
        }
        u[2 * k] = ud / 4 / ORC_PI;
        u[2 * k + 1] = vd / 4 / ORC_PI;
    }
}

/* u_t = 4 sum_{s != t} (r.n_s)(r.f_s) r/rho^4 / (4 pi) at the sources. */
void orc_free_dlp(const double* f, const double* n, const double* p, int N, double* u)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) {
        double ud = 0, vd = 0;
        for (int s = 0; s < N; ++s) {
            if (s == k) continue;
            double rx = p[2 * k] - p[2 * s], ry = p[2 * k + 1] - p[2 * s + 1];
            double rho2 = rx * rx + ry * ry;
            if (rho2 == 0) continue;
            double rho4 = rho2 * rho2;
            double rdotf = rx * f[2 * s] + ry * f[2 * s + 1];
            double rdotn = rx * n[2 * s] + ry * n[2 * s + 1];
            ud += rdotn * rdotf / rho4 * rx;
            vd += rdotn * rdotf / rho4 * ry;
        }
        u[2 * k] = 4 * ud / 4 / ORC_PI;
        u[2 * k + 1] = 4 * vd / 4 / ORC_PI;
    }
}
