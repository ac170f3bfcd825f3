/*
 * se2p -- spectral Ewald for doubly-periodic 2D Stokes sums on AMD Instinct MI355X.
 *
 * C ABI of libse2p.so.  Plain pointers and sizes only.  All functions return
 * SE2P_OK (0) or a negative status; se2p_last_error() gives the text of the
 * last failure on the calling thread.  No function aborts the process.
 *
 * Layouts follow the reference mex interface:
 *   points, densities f and normals n are 2 x N column-major double arrays,
 *   i.e. interleaved [x0, y0, x1, y1, ...];
 *   velocities are returned in the same 2 x N layout.
 * Host-pointer entry points are synchronous (like a mex call): inputs are only
 * read, outputs are fully overwritten, and nothing is retained after return.
 *
 * Reference paths are relative to the reference repository root
 * (PeriodicEwald/src/...).
 */
#ifndef SE2P_H
#define SE2P_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SE2P_OK 0
#define SE2P_ERR_ARG (-1)         /* bad argument, shape or violated precondition */
#define SE2P_ERR_HIP (-2)         /* HIP runtime failure */
#define SE2P_ERR_NOMEM (-3)       /* device or host allocation failed */
#define SE2P_ERR_UNSUPPORTED (-4) /* valid input outside what this build handles */
#define SE2P_ERR_NODEV (-5)       /* no usable GPU */

#define SE2P_SLP 0 /* Stokeslet, single-layer potential */
#define SE2P_DLP 1 /* stresslet, double-layer potential */
#define SE2P_SLP_PRESSURE 2 /* pressure of the Stokeslet (one value per target) */
#define SE2P_DLP_PRESSURE 3 /* pressure of the stresslet (one value per target) */
/* Further derived quantities.  Tensors are 4 values per target in the order
 * [d/dx of component 1, d/dx of component 2, d/dy of component 1, d/dy of
 * component 2] (e.g. mex_stokes_slp_gradient_real.cpp:118-132). */
#define SE2P_SLP_GRADIENT 4      /* Stokeslet velocity gradient (4 per target) */
#define SE2P_SLP_PRESSURE_GRAD 5 /* Stokeslet pressure gradient (2 per target) */
#define SE2P_SLP_VORTICITY 6     /* Stokeslet vorticity (1 per target) */
#define SE2P_SLP_STRESS 7        /* Stokeslet stress tensor (4 per target) */
#define SE2P_DLP_GRADIENT 8      /* stresslet velocity gradient (4 per target) */
#define SE2P_DLP_PRESSURE_GRAD 9 /* stresslet pressure gradient (2 per target) */
#define SE2P_DLP_VORTICITY 10    /* stresslet vorticity (1 per target) */
#define SE2P_DLP_STRESS 11       /* stresslet stress tensor (4 per target) */
#define SE2P_NUM_KINDS 12

/* Arithmetic of an evaluation (se2p_params.precision). */
#define SE2P_FP64 0 /* every stage in FP64 */
/* FP32 tier, for tolerances down to 1e-6 (Stokeslet and stresslet velocity; the
 * derived kinds run FP64 whatever is asked): spreading on the FP32 matrix
 * cores with two-sum accumulation per chunk, float2 grids through the FFT and
 * filter (each mode filtered in FP64), gathering in FP32 rows added in FP64,
 * real-space far pairs (xi^2 r^2 >= 1) in FP32 added per chunk to FP64 sums;
 * near pairs (the E1 series near r = 0), self terms, zero modes and every
 * output sum stay FP64. */
#define SE2P_FP32C 1

typedef struct se2p_ctx se2p_ctx;
typedef struct se2p_plan se2p_plan;
typedef struct se2p_batch se2p_batch;

/* ------------------------------------------------------------------------ */
/* Context: one GPU, one HIP stream, a grow-only device workspace.           */
/* One context per host thread.                                              */
/* ------------------------------------------------------------------------ */
int se2p_ctx_create(int device, se2p_ctx** out);
int se2p_ctx_destroy(se2p_ctx* ctx);
const char* se2p_last_error(void);
int se2p_version(void);
/* Number of visible GPUs (0 when none). */
int se2p_device_count(int* count);
/* Precision (SE2P_FP64, the default, or SE2P_FP32C) of this context's one-shot
 * entry points: the mex equivalents and the drivers.  Plans, batches and
 * se2p_eval_device take theirs from se2p_params.precision.  No counterpart in
 * the reference, which computes in FP64 only. */
int se2p_ctx_set_precision(se2p_ctx* ctx, int precision);
/* Method of this context's free-space Stokeslet sums (se2p_free_slp):
 * SE2P_FREE_AUTO (the default) takes the spectral Ewald sum when Nsrc x Ntar >=
 * 1e9 and iprec <= 4, direct summation otherwise; SE2P_FREE_DIRECT and
 * SE2P_FREE_EWALD force one.  No counterpart in the reference, whose FMM has no
 * such choice. */
#define SE2P_FREE_AUTO 0
#define SE2P_FREE_DIRECT 1
#define SE2P_FREE_EWALD 2
int se2p_ctx_set_free_method(se2p_ctx* ctx, int method);

/* ------------------------------------------------------------------------ */
/* The four mex kernels, same argument order, sizes made explicit.          */
/* ------------------------------------------------------------------------ */

/* Replaces mex_stokes_slp_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly)
 * (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_real.cpp:58-83).
 * ur: 2 x Ntar.  Points must lie in [-Lx/2, Lx/2) x [-Ly/2, Ly/2). */
int se2p_slp_real(se2p_ctx* ctx, const double* psrc, const double* ptar, const double* f,
                  int Nsrc, int Ntar, double xi, int nside_x, int nside_y, double Lx,
                  double Ly, double* ur);

/* Replaces mex_stokes_slp_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P)
 * (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_kspace.cpp:7-45).
 * Requires Lx/Mx == Ly/My, Mx and My even, P even, P + 2 <= min(Mx, My). */
int se2p_slp_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                    double eta, const double* f, int Nsrc, int Ntar, int Mx, int My,
                    double Lx, double Ly, double w, int P, double* uk);

/* Replaces mex_stokes_dlp_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly)
 * (PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_real.cpp:15-52). */
int se2p_dlp_real(se2p_ctx* ctx, const double* psrc, const double* ptar, const double* f,
                  const double* n, int Nsrc, int Ntar, double xi, int nside_x, int nside_y,
                  double Lx, double Ly, double* ur);

/* Replaces mex_stokes_dlp_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P)
 * (PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_kspace.cpp:10-56).  Like the
 * mex, the zero mode is not included. */
int se2p_dlp_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                    double eta, const double* f, const double* n, int Nsrc, int Ntar, int Mx,
                    int My, double Lx, double Ly, double w, int P, double* uk);

/* ------------------------------------------------------------------------ */
/* Drivers: StokesSLP_ewald_2p / StokesDLP_ewald_2p restated.                */
/* ------------------------------------------------------------------------ */
typedef struct {
    int P;      /* Gaussian support nodes per direction minus one, even (default 24) */
    double Nb;  /* mean points per cell; <= 0 selects min(P round(log2 npts), npts/4) */
    double tol; /* error tolerance of the analytic estimates (default 1e-16) */
    int verbose;
} se2p_options;

typedef struct {
    double xi, rc, w, eta;
    int nside_x, nside_y, Mx, My, P, kinf;
    int converged; /* 0 when a Newton solve hit its iteration cap */
    int e1_tier;   /* real-space E1: 0 full FP64 (reference), 1 rel. error < 1e-12 */
    /* Real space of the velocity sums (Stokeslet and stresslet) with the
     * sources as the targets (plans and batch systems whose d_psrc and d_ptar
     * are the same pointer, with a k-space part): 1 evaluates each pair once and
     * adds it to both points with atomics, so the sums' order, and their last
     * bits, vary from run to run; 0 (the reference's order) sums each target's
     * pairs in a fixed order.  Ignored for other kinds and inputs, and where
     * a point lies outside the box (se2p_plan_params reports what is in use). */
    int symmetric;
    int precision; /* SE2P_FP64 or SE2P_FP32C */
    /* Fast E1 tier of the Stokeslet (e1_tier 1, FP64): real-space pairs with
     * xi^2 r^2 >= far32 are evaluated in FP32 (their terms are below e^{-far32}
     * of the largest, so FP32 rounding stays that far below the tolerance) and
     * added per chunk to the FP64 sums; 0 evaluates every pair in FP64. */
    double far32;
    /* Stokeslet only: > 0 evaluates the free-space (non-periodic) sum instead of
     * the periodic one.  The k-space filter then uses the Fourier transform of
     * the biharmonic kernel, less the quadratic matching its value and slope at
     * free_R, truncated at radius free_R, and (1 - ln free_R) sum(f)/(4 pi) is
     * added to every target.  Every source and target must lie in a sub-box of
     * sides wx, wy with free_R >= sqrt(wx^2 + wy^2) + delta, Lx >= wx + free_R +
     * delta and Ly >= wy + free_R + delta, delta several 1/xi (se2p_free_params
     * sets all of it); both the real-space and the k-space parts are needed.
     * 0: the periodic sum. */
    double free_R;
} se2p_params;

/* ------------------------------------------------------------------------ */
/* Derived quantity: pressure of the Stokeslet.                              */
/* ------------------------------------------------------------------------ */
/* Replaces mex_stokes_slp_pressure_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly)
 * (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_pressure_real.cpp:17-160).
 * pr: 1 x Ntar. */
int se2p_slp_pressure_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                           const double* f, int Nsrc, int Ntar, double xi, int nside_x,
                           int nside_y, double Lx, double Ly, double* pr);
/* Replaces mex_stokes_slp_pressure_kspace(psrc, ptar, f, xi, eta, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_slp_pressure_kspace.cpp:7-190; note f before xi).  pk: 1 x Ntar,
 * without the zero mode (the driver adds it). */
int se2p_slp_pressure_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                             const double* f, double xi, double eta, int Nsrc, int Ntar, int Mx,
                             int My, double Lx, double Ly, double w, int P, double* pk);
/* StokesSLP_pressure_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, Lx, Ly, ...)
 * (StokesSLP_pressure_ewald_2p.m:1-138): wrapping, the SLP parameter recipe, p =
 * pr + pk with the zero mode -sum(f1 x + f2 y)/(2 Lx Ly) in pk.  p, pr, pk: Ntar
 * values each (pr, pk and xi may be NULL). */
int se2p_slp_pressure_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc,
                            const double* xtar, const double* ytar, const double* f1,
                            const double* f2, int Nsrc, int Ntar, double Lx, double Ly,
                            const se2p_options* opt, double* p, double* pr, double* pk,
                            double* xi);

/* ------------------------------------------------------------------------ */
/* Derived quantity: pressure of the stresslet.                              */
/* ------------------------------------------------------------------------ */
/* Replaces mex_stokes_dlp_pressure_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly)
 * (PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_pressure_real.cpp:17-195).
 * pr: 1 x Ntar. */
int se2p_dlp_pressure_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                           const double* f, const double* n, int Nsrc, int Ntar, double xi,
                           int nside_x, int nside_y, double Lx, double Ly, double* pr);
/* Replaces mex_stokes_dlp_pressure_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_dlp_pressure_kspace.cpp:10-230).  pk: 1 x Ntar, without the zero mode. */
int se2p_dlp_pressure_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                             const double* f, const double* n, double xi, double eta, int Nsrc,
                             int Ntar, int Mx, int My, double Lx, double Ly, double w, int P,
                             double* pk);
/* StokesDLP_pressure_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Lx, Ly, ...)
 * (StokesDLP_pressure_ewald_2p.m): wrapping, the DLP parameter recipe, p = pr + pk
 * with the zero mode -sum(n1 f1 + n2 f2)/(2 Lx Ly) in pk, plus xi^2 (f.n)/(2 pi)
 * of each source at the targets with its exact (unwrapped) coordinates
 * (:64-70,165-172). */
int se2p_dlp_pressure_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc,
                            const double* xtar, const double* ytar, const double* n1,
                            const double* n2, const double* f1, const double* f2, int Nsrc,
                            int Ntar, double Lx, double Ly, const se2p_options* opt, double* p,
                            double* pr, double* pk, double* xi);

/* ------------------------------------------------------------------------ */
/* Derived quantities: velocity gradient, pressure gradient, vorticity and   */
/* stress of both layers.  Each pair replaces the reference's real-space and */
/* k-space mex functions with their argument order; outputs are written in  */
/* full (nc x Ntar, nc = 4, 2, 1, 4).  The k-space parts hold no zero mode   */
/* and no self term: the MATLAB drivers add those.  SLP arguments: psrc,     */
/* ptar, f; DLP arguments add the normals n after f.                         */
/* ------------------------------------------------------------------------ */
/* mex_stokes_slp_gradient_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly)
 * (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_gradient_real.cpp:9-217) and
 * mex_stokes_slp_gradient_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_slp_gradient_kspace.cpp:3-288); out 4 x Ntar. */
int se2p_slp_gradient_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                           const double* f, int Nsrc, int Ntar, double xi, int nside_x,
                           int nside_y, double Lx, double Ly, double* out);
int se2p_slp_gradient_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                             double eta, const double* f, int Nsrc, int Ntar, int Mx, int My,
                             double Lx, double Ly, double w, int P, double* out);
/* mex_stokes_slp_pressure_grad_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly)
 * (mex_stokes_slp_pressure_grad_real.cpp:17-186) and
 * mex_stokes_slp_pressure_grad_kspace(psrc, ptar, f, xi, eta, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_slp_pressure_grad_kspace.cpp:7-195, f before xi); out 2 x Ntar. */
int se2p_slp_pressure_grad_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                                const double* f, int Nsrc, int Ntar, double xi, int nside_x,
                                int nside_y, double Lx, double Ly, double* out);
int se2p_slp_pressure_grad_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                                  const double* f, double xi, double eta, int Nsrc, int Ntar,
                                  int Mx, int My, double Lx, double Ly, double w, int P,
                                  double* out);
/* mex_stokes_slp_vorticity_real (mex_stokes_slp_vorticity_real.cpp:9-187) and
 * mex_stokes_slp_vorticity_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_slp_vorticity_kspace.cpp:3-203); out 1 x Ntar. */
int se2p_slp_vorticity_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                            const double* f, int Nsrc, int Ntar, double xi, int nside_x,
                            int nside_y, double Lx, double Ly, double* out);
int se2p_slp_vorticity_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                              double eta, const double* f, int Nsrc, int Ntar, int Mx, int My,
                              double Lx, double Ly, double w, int P, double* out);
/* mex_stokes_slp_stress_real (mex_stokes_slp_stress_real.cpp:9-217) and
 * mex_stokes_slp_stress_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_slp_stress_kspace.cpp:3-293); out 4 x Ntar. */
int se2p_slp_stress_real(se2p_ctx* ctx, const double* psrc, const double* ptar, const double* f,
                         int Nsrc, int Ntar, double xi, int nside_x, int nside_y, double Lx,
                         double Ly, double* out);
int se2p_slp_stress_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                           double eta, const double* f, int Nsrc, int Ntar, int Mx, int My,
                           double Lx, double Ly, double w, int P, double* out);
/* mex_stokes_dlp_gradient_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly)
 * (PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_gradient_real.cpp:9-240) and
 * mex_stokes_dlp_gradient_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_dlp_gradient_kspace.cpp:3-345); out 4 x Ntar. */
int se2p_dlp_gradient_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                           const double* f, const double* n, int Nsrc, int Ntar, double xi,
                           int nside_x, int nside_y, double Lx, double Ly, double* out);
int se2p_dlp_gradient_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                             double eta, const double* f, const double* n, int Nsrc, int Ntar,
                             int Mx, int My, double Lx, double Ly, double w, int P, double* out);
/* mex_stokes_dlp_pressure_grad_real (mex_stokes_dlp_pressure_grad_real.cpp:17-207) and
 * mex_stokes_dlp_pressure_grad_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_dlp_pressure_grad_kspace.cpp:10-269); out 2 x Ntar. */
int se2p_dlp_pressure_grad_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                                const double* f, const double* n, int Nsrc, int Ntar, double xi,
                                int nside_x, int nside_y, double Lx, double Ly, double* out);
int se2p_dlp_pressure_grad_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                                  const double* f, const double* n, double xi, double eta,
                                  int Nsrc, int Ntar, int Mx, int My, double Lx, double Ly,
                                  double w, int P, double* out);
/* mex_stokes_dlp_vorticity_real (mex_stokes_dlp_vorticity_real.cpp:19-221) and
 * mex_stokes_dlp_vorticity_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_dlp_vorticity_kspace.cpp:3-252); out 1 x Ntar. */
int se2p_dlp_vorticity_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                            const double* f, const double* n, int Nsrc, int Ntar, double xi,
                            int nside_x, int nside_y, double Lx, double Ly, double* out);
int se2p_dlp_vorticity_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                              const double* f, const double* n, double xi, double eta, int Nsrc,
                              int Ntar, int Mx, int My, double Lx, double Ly, double w, int P,
                              double* out);
/* mex_stokes_dlp_stress_real (mex_stokes_dlp_stress_real.cpp:9-248) and
 * mex_stokes_dlp_stress_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P)
 * (mex_stokes_dlp_stress_kspace.cpp:3-340); out 4 x Ntar. */
int se2p_dlp_stress_real(se2p_ctx* ctx, const double* psrc, const double* ptar, const double* f,
                         const double* n, int Nsrc, int Ntar, double xi, int nside_x,
                         int nside_y, double Lx, double Ly, double* out);
int se2p_dlp_stress_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                           double eta, const double* f, const double* n, int Nsrc, int Ntar,
                           int Mx, int My, double Lx, double Ly, double w, int P, double* out);

/* ------------------------------------------------------------------------ */
/* Free-space (non-periodic) layer potentials.  The reference evaluates them */
/* with a Fortran FMM.  Here: direct sums on the GPU in FP64 (exact to       */
/* rounding, O(Nsrc Ntar) work), and for the Stokeslet a free-space spectral */
/* Ewald sum (O(N log N), accurate to the FMM's tolerance for iprec).        */
/* ------------------------------------------------------------------------ */
/* [u1, u2] = stokesSLPfmm(f1, f2, xsrc, ysrc, xtar, ytar, iequal, iprec)
 * (FMM/src/StokesSLP/stokesSLPfmm.F:10-149, stokesSLP.f): u = sum_s (-1/2
 * log(rho^2) f + (r.f) r/rho^2)/(4 pi), r = x - x_s (the direct sum of
 * FMM/tests/stokes_SLP_FMM_test.m:29-51).  iequal = 1: the targets are the
 * sources (Ntar == Nsrc; xtar, ytar may be NULL) and each source's own term is
 * left out.  iprec (-2..5) is the FMM's accuracy level (cfmm2dpart.f:351-361:
 * relative tolerance 0.5 10^-k, k = 0, 1, 2, 3, 6, 9, 12, 15).  Coincident
 * source-target pairs are skipped.  The context's free-space method
 * (se2p_ctx_set_free_method) picks direct summation or the free-space spectral
 * Ewald sum at the tolerance of iprec (at most 5e-3, at least 5e-13; iprec 5
 * always sums directly). */
int se2p_free_slp(se2p_ctx* ctx, const double* f1, const double* f2, const double* xsrc,
                  const double* ysrc, int Nsrc, const double* xtar, const double* ytar, int Ntar,
                  int iequal, int iprec, double* u1, double* u2);
/* [u1, u2] = stokesDLPfmm(f1, f2, xsrc, ysrc, n1, n2) (FMM/src/StokesDLP/
 * stokesDLPfmm.F, stokesDLP.f): at the N sources, u = sum_{s != t} (r.n)(r.f)
 * r/rho^4 / pi (FMM/tests/stokes_DLP_FMM_test.m:17-35); zero when sum |f1| +
 * |f2| <= 1e-14 (stokesDLP.f:31). */
int se2p_free_dlp(se2p_ctx* ctx, const double* f1, const double* f2, const double* xsrc,
                  const double* ysrc, const double* n1, const double* n2, int N, double* u1,
                  double* u2);
/* Parameters of a free-space spectral Ewald Stokeslet sum (se2p_params.free_R)
 * for Nsrc sources and Ntar targets inside a wx x wy box at relative tolerance
 * tol (>= 1e-13): xi, cells, grid and support from the tolerance rules of the
 * periodic tuner with L = ln(100/tol) (xi rc = sqrt(L - 3.2), about 22 points
 * per cell of the point box, grids at most 10240 per side) and k_inf/xi =
 * 2 sqrt(L - 2.4 + 0.5 ln(free_R k_inf)) for the truncated kernel's transform,
 * delta = sqrt(L + 2 ln(xi D))/xi with D = sqrt(wx^2 + wy^2), free_R = D + delta and the
 * periodic box *Lx x *Ly = (wx + free_R + delta, wy + free_R + delta) rounded
 * up to whole grid steps.  The points go in centred: subtract the centre of
 * their wx x wy box.  No counterpart in the reference, which uses an FMM.  */
int se2p_free_params(int Nsrc, int Ntar, double wx, double wy, double tol, se2p_params* prm,
                     double* Lx, double* Ly);

void se2p_options_default(se2p_options* opt);

/* Parameter recipe of StokesSLP_ewald_2p.m:27-37,84-104,147-202 (kind SE2P_SLP)
 * or StokesDLP_ewald_2p.m:28-38,104-125,171-225 (kind SE2P_DLP).  Q is
 * sum(f.^2) + 1 over both components of the density.  A derived kind takes the
 * recipe its MATLAB driver uses: the SLP recipe for every Stokeslet quantity
 * and for the stresslet vorticity (StokesDLP_vorticity_ewald_2p.m:113-131,
 * whose find_xi/find_kinfb are the SLP ones), the DLP recipe for the other
 * stresslet quantities. */
int se2p_params_ref(int kind, int Nsrc, int Ntar, double Q, double Lx, double Ly,
                    const se2p_options* opt, se2p_params* out);

/* StokesSLP_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, Lx, Ly, ...) ->
 * [u1, u2, ur, uk, xi] (StokesSLP_ewald_2p.m:1-138).  Points are wrapped into
 * the box first.  ur, uk (2 x Ntar) and xi may be NULL. */
int se2p_slp_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc,
                   const double* xtar, const double* ytar, const double* f1,
                   const double* f2, int Nsrc, int Ntar, double Lx, double Ly,
                   const se2p_options* opt, double* u1, double* u2, double* ur, double* uk,
                   double* xi);

/* StokesDLP_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Lx, Ly, ...)
 * (StokesDLP_ewald_2p.m:1-163), zero mode Σ(f.n)x_s/(Lx Ly) added to uk. */
int se2p_dlp_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc,
                   const double* xtar, const double* ytar, const double* n1,
                   const double* n2, const double* f1, const double* f2, int Nsrc, int Ntar,
                   double Lx, double Ly, const se2p_options* opt, double* u1, double* u2,
                   double* ur, double* uk, double* xi);

/* ------------------------------------------------------------------------ */
/* Device-resident evaluation.                                               */
/* ------------------------------------------------------------------------ */
/* Device buffers owned by the caller, allocated through the context. */
int se2p_dev_alloc(se2p_ctx* ctx, size_t bytes, void** dptr);
int se2p_dev_free(se2p_ctx* ctx, void* dptr);
int se2p_memcpy_h2d(se2p_ctx* ctx, void* dst, const void* src, size_t bytes);
int se2p_memcpy_d2h(se2p_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy, stream-ordered on the context's stream (no host sync). */
int se2p_memcpy_d2d(se2p_ctx* ctx, void* dst, const void* src, size_t bytes);
int se2p_ctx_sync(se2p_ctx* ctx);

/* One complete evaluation on device pointers with explicit parameters: binning,
 * real space and k space recomputed on every call, as the reference does.
 * d_n must be given for the stresslet kinds (NULL otherwise).  d_u (nc x Ntar,
 * nc the kind's values per target) receives ur + uk, plus the zero mode for the
 * kinds whose driver adds one that is the same for every target: DLP velocity
 * (StokesDLP_ewald_2p.m:149-151), both pressures and the SLP vorticity
 * (StokesSLP_vorticity_ewald_2p.m:145).  Stream-ordered on the context's
 * stream; call se2p_ctx_sync() before reading d_u. */
int se2p_eval_device(se2p_ctx* ctx, int kind, const double* d_psrc, const double* d_ptar,
                     const double* d_f, const double* d_n, int Nsrc, int Ntar, double Lx,
                     double Ly, const se2p_params* prm, double* d_u);

/* As se2p_eval_device, allocating nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * workspace buffer of the evaluation is carved from a fixed scratch array of
 * 256 MiB per GPU (a static device array the library holds until the process
 * ends); a problem that needs more (about 200 bytes per source or target plus
 * 32 bytes per grid node for a Stokeslet, so roughly a million points in all)
 * is rejected with SE2P_ERR_UNSUPPORTED before any work is queued.
 * All calls on one GPU share its array and it keeps the last call's contents
 * (the kernels clear what must start at zero): calls that overlap on one GPU,
 * from two contexts or threads, race.  Results are those of se2p_eval_device,
 * bit for bit. */
int se2p_eval_device_noalloc(se2p_ctx* ctx, int kind, const double* d_psrc,
                             const double* d_ptar, const double* d_f, const double* d_n,
                             int Nsrc, int Ntar, double Lx, double Ly, const se2p_params* prm,
                             double* d_u);

/* ------------------------------------------------------------------------ */
/* Plans: fixed geometry, repeated matvecs.  No counterpart in the reference, */
/* which recomputes binning and Gaussian factors on every mex call           */
/* (mex_stokes_slp_real.cpp:98-100, ewald_tools.cpp:125-127,152-159).        */
/* ------------------------------------------------------------------------ */

/* Bins sources and targets, orders them, computes the start nodes and keeps
 * every buffer an evaluation needs.  Device pointers; d_n for stresslet kinds.
 * The point arrays are read during creation and not retained. */
int se2p_plan_create(se2p_ctx* ctx, int kind, const double* d_psrc, const double* d_ptar,
                     const double* d_n, int Nsrc, int Ntar, double Lx, double Ly,
                     const se2p_params* prm, se2p_plan** out);
/* d_u (nc x Ntar) = ur + uk (+ the zero mode, as for se2p_eval_device) for
 * density d_f (2 x Nsrc).  Stream-ordered on the context's stream; allocates
 * nothing. */
int se2p_plan_apply(se2p_plan* plan, const double* d_f, double* d_u);
/* As se2p_plan_apply, replayed from a HIP graph captured on the first call for
 * this (d_f, d_u) pair; a new pair re-captures. */
int se2p_plan_apply_graph(se2p_plan* plan, const double* d_f, double* d_u);
int se2p_plan_destroy(se2p_plan* plan);
/* The parameters a plan's evaluations run with: a copy of those it was created
 * with, except that `symmetric` is 1 only when pair-symmetric sums are in use
 * (they need the targets to be the sources and every cell ordered by y; a cell
 * of more than 4096 points or a point outside the box selects the fixed order). */
int se2p_plan_params(const se2p_plan* plan, se2p_params* used);


/* ------------------------------------------------------------------------ */
/* Batches: many independent periodic systems (own points, box, kind and     */
/* parameters each), fixed geometry, evaluated together: every stage runs as */
/* one launch over the blocks of all systems.  No counterpart in the         */
/* reference (one system per mex call, StokesSLP_ewald_2p.m:119,126).        */
/* ------------------------------------------------------------------------ */
typedef struct se2p_system {
    int kind;             /* SE2P_SLP .. SE2P_DLP_STRESS */
    int Nsrc, Ntar;
    double Lx, Ly;
    const double* d_psrc; /* 2 x Nsrc device array, read during creation only */
    const double* d_ptar; /* 2 x Ntar */
    const double* d_n;    /* 2 x Nsrc normals (stresslet kinds), else NULL */
    se2p_params prm;      /* e.g. from se2p_params_ref at the system's tolerance */
} se2p_system;

int se2p_batch_create(se2p_ctx* ctx, int nsys, const se2p_system* systems, se2p_batch** out);
/* Doubles of the concatenated densities (sum of 2 Nsrc) and outputs (sum of nc Ntar). */
int se2p_batch_sizes(const se2p_batch* batch, long long* nf, long long* nu);
/* d_u = u of every system, systems back to back (nc x Ntar each), for the
 * densities d_f (2 x Nsrc each, same order).  Stream-ordered; allocates nothing. */
int se2p_batch_apply(se2p_batch* batch, const double* d_f, double* d_u);
/* As se2p_batch_apply, replayed from a HIP graph captured on the first call for
 * this (d_f, d_u) pair; a new pair re-captures. */
int se2p_batch_apply_graph(se2p_batch* batch, const double* d_f, double* d_u);
/* As se2p_plan_params for system `sys` (0-based, creation order) of a batch. */
int se2p_batch_params(const se2p_batch* batch, int sys, se2p_params* used);
int se2p_batch_destroy(se2p_batch* batch);

/* ------------------------------------------------------------------------ */
/* Parameter selector network.  No counterpart in the reference, whose       */
/* drivers pick (xi, M, P, rc) from analytic error estimates                 */
/* (StokesSLP_ewald_2p.m:84-104,147-202).  An MLP of SE2P_MLP_IN inputs, two */
/* hidden layers of SE2P_MLP_HIDDEN GELU units (gelu(z) = z Phi(z)) and      */
/* SE2P_MLP_OUT outputs, in FP32, its SE2P_MLP_PARAMS parameters packed as   */
/* W1[64][8], b1[64], W2[64][64], b2[64], W3[4][64], b3[4].  Device arrays,  */
/* stream-ordered on the context's stream.                                   */
/* ------------------------------------------------------------------------ */
#define SE2P_MLP_IN 8
#define SE2P_MLP_HIDDEN 64
#define SE2P_MLP_OUT 4
#define SE2P_MLP_PARAMS 4996
/* y (n x 4) = the network applied to x (n x 8). */
int se2p_mlp_forward(se2p_ctx* ctx, const float* d_w, const float* d_x, int n, float* d_y);
/* d_loss (1 float) = mean over the n x 4 outputs of (y - t)^2 and d_grad
 * (SE2P_MLP_PARAMS floats) its gradient, summed in a fixed order. */
int se2p_mlp_loss_grad(se2p_ctx* ctx, const float* d_w, const float* d_x, const float* d_t,
                       int n, float* d_grad, float* d_loss);
/* One Adam step (bias-corrected, step >= 1) on the packed parameters with
 * gradient d_g and moment arrays d_m, d_v (updated in place). */
int se2p_mlp_adam(se2p_ctx* ctx, float* d_w, const float* d_g, float* d_m, float* d_v, int step,
                  float lr, float beta1, float beta2, float eps);

/* ------------------------------------------------------------------------ */
/* Stage-level entry points (host buffers, synchronous): the helpers of      */
/* PeriodicEwald/src/mex/common/ewald_tools.h, each replaced by the GPU path  */
/* the mex kernels above use internally.                                     */
/* ------------------------------------------------------------------------ */

/* Replaces Assign (ewald_tools.cpp:7-80) for one point set: stable counting
 * sort into nside_x x nside_y cells.  particle_offsets: n; box_offsets:
 * nside_x*nside_y + 1; n_in_box: nside_x*nside_y.  Bit-identical. */
int se2p_assign(se2p_ctx* ctx, const double* p, int n, double Lx, double Ly, int nside_x,
                int nside_y, int* particle_offsets, int* box_offsets, int* n_in_box);

/* Replaces FindClosestNode (ewald_tools.cpp:86-114) for n points; mx, my are
 * the unwrapped start nodes as in the reference.  Bit-identical. */
int se2p_find_closest_node(se2p_ctx* ctx, const double* p, int n, double Lx, double Ly,
                           double h, int P, int* mx, int* my, double* px, double* py);

/* Replaces Spread (ewald_tools.cpp:120-199): H1, H2 are My x Mx column-major. */
int se2p_spread(se2p_ctx* ctx, const double* psrc, const double* f, int Nsrc, double Lx,
                double Ly, double xi, double w, double eta, int P, int Mx, int My, double* H1,
                double* H2);

/* Replaces Gather (ewald_tools.cpp:206-271) for one real grid H (My x Mx);
 * out[k] (k < Ntar) receives the gathered, scaled value. */
int se2p_gather(se2p_ctx* ctx, const double* H, const double* ptar, int Ntar, double Lx,
                double Ly, double xi, double w, double eta, int P, int Mx, int My, double* out);

/* Replaces MATLAB fft2/ifft2 on a My x Mx column-major complex grid held as
 * separate real and imaginary arrays (mex_stokes_slp_kspace.cpp:80-81,176-177);
 * in place; the inverse scales by 1/(Mx My) like ifft2. */
int se2p_fft2(se2p_ctx* ctx, double* re, double* im, int My, int Mx, int inverse);

/* The real-space E1 evaluator (mex_stokes_slp_real.cpp:166-205) on n values;
 * tier 0 = full FP64, 1 = reduced (relative error below 1e-12). */
int se2p_e1(se2p_ctx* ctx, const double* x, int n, int tier, double* out);

/* ------------------------------------------------------------------------ */
/* Per-stage timing with HIP events, each stage bracketed on the stream it   */
/* runs on.  on = 0: off; 1: on, the real-space and k-space passes of an     */
/* apply still run concurrently on the context's two streams (a stage's time */
/* includes what the other stream takes from it); 2: on, every stage on one  */
/* stream in order, so each stage is timed alone.                            */
/* ------------------------------------------------------------------------ */
int se2p_prof_enable(se2p_ctx* ctx, int on);
int se2p_prof_reset(se2p_ctx* ctx);
/* Sum of event-measured milliseconds and launch count for a stage name
 * ("bin", "pack", "real", "spread", "fft", "filter", "gather", "free"). */
int se2p_prof_get(se2p_ctx* ctx, const char* stage, double* total_ms, int* count);

#ifdef __cplusplus
}
#endif
#endif
