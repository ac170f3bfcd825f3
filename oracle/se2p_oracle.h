/*
 * CPU oracle for the doubly-periodic 2D Stokes spectral Ewald path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against (and the timed CPU baseline).  It is
 * never linked into, loaded by or called from the product library.
 *
 * Every function restates one piece of the reference algorithm in plain C
 * (FP64, no FMA contraction), citing the reference file:line it follows.
 * Layouts follow the reference: point/density arrays are interleaved
 * 2 x N ([x0,y0,x1,y1,...]); grids are My x Mx column-major, H[ix*My+iy].
 */
#ifndef SE2P_ORACLE_H
#define SE2P_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* a1: cell binning by stable counting sort (ewald_tools.cpp:7-80). */
void orc_assign(const double* p, int n, double Lx, double Ly, int nside_x, int nside_y,
                int* particle_offsets, int* box_offsets, int* n_in_box);

/* a7: start node and in-cell offset of the Gaussian support (ewald_tools.cpp:86-114). */
void orc_find_closest_node(double x, double y, double Lx, double Ly, double h, int P,
                           int* mx, int* my, double* px, double* py);

/* a4: exponential integral E1(x), x > 0, evaluated in long double. */
double orc_e1(double x);

/* a3/a5: SLP real-space cell-list sum (mex_stokes_slp_real.cpp:58-357). */
void orc_slp_real(const double* psrc, const double* ptar, const double* f, int Ns, int Nt,
                  double xi, int nside_x, int nside_y, double Lx, double Ly, double* u);

/* a6: DLP real-space cell-list sum (mex_stokes_dlp_real.cpp:15-211). */
void orc_dlp_real(const double* psrc, const double* ptar, const double* f, const double* n,
                  int Ns, int Nt, double xi, int nside_x, int nside_y, double Lx, double Ly,
                  double* u);

/* a8: FGG spreading of two density components (ewald_tools.cpp:120-199). */
void orc_spread(const double* psrc, const double* f, int Ns, double Lx, double Ly,
                double xi, double w, double eta, int P, int Mx, int My,
                double* H1, double* H2);

/* a9: 2D DFT of a complex My x Mx column-major grid; inverse scales by 1/(Mx My). */
void orc_fft2(double* re, double* im, int My, int Mx, int inverse);

/* a10: SLP k-space filter in place incl. k=0 removal (mex_stokes_slp_kspace.cpp:110-169). */
void orc_slp_filter(double* h1re, double* h1im, double* h2re, double* h2im, int Mx, int My,
                    double Lx, double Ly, double xi, double eta);

/* a11: DLP k-space filter, 4 spectra in, 2 out (mex_stokes_dlp_kspace.cpp:151-232). */
void orc_dlp_filter(double* h1re, double* h1im, double* h2re, double* h2im,
                    const double* h3re, const double* h3im, const double* h4re,
                    const double* h4im, int Mx, int My, double Lx, double Ly, double xi,
                    double eta);

/* a12: FGG gather of one component into out[ncomp*k + comp] (ewald_tools.cpp:206-271). */
void orc_gather(const double* H, int ncomp, int comp, const double* ptar, int Nt,
                double Lx, double Ly, double xi, double w, double eta, int P, int Mx,
                int My, double* out);

/* Whole k-space parts (mex_stokes_slp_kspace.cpp:7-215, mex_stokes_dlp_kspace.cpp:10-282). */
void orc_slp_kspace(const double* psrc, const double* ptar, double xi, double eta,
                    const double* f, int Ns, int Nt, int Mx, int My, double Lx, double Ly,
                    double w, int P, double* uk);
void orc_dlp_kspace(const double* psrc, const double* ptar, double xi, double eta,
                    const double* f, const double* n, int Ns, int Nt, int Mx, int My,
                    double Lx, double Ly, double w, int P, double* uk);

/* Direct (unspread) sums, tests/direct_sums/matlab_sums/stokes_{slp,dlp}_{real,kspace}_ds.m */
void orc_slp_real_ds(const double* psrc, const double* ptar, const double* f, int Ns,
                     int Nt, double Lx, double Ly, double xi, double* u);
void orc_slp_kspace_ds(const double* psrc, const double* ptar, const double* f, int Ns,
                       int Nt, double Lx, double Ly, double xi, int kinf, double* u);
void orc_dlp_real_ds(const double* psrc, const double* ptar, const double* f,
                     const double* n, int Ns, int Nt, double Lx, double Ly, double xi,
                     double* u);
void orc_dlp_kspace_ds(const double* psrc, const double* ptar, const double* f,
                       const double* n, int Ns, int Nt, double Lx, double Ly, double xi,
                       int kinf, double* u);

/* DLP zero mode sum_s (f.n)_s x_s / (Lx Ly) (StokesDLP_ewald_2p.m:149-151). */
void orc_dlp_zero_mode(const double* psrc, const double* f, const double* n, int Ns,
                       double Lx, double Ly, double out[2]);

/* a14: parameter recipe of StokesSLP_ewald_2p.m:85-104 / StokesDLP_ewald_2p.m:105-125.
 * kind 0 = SLP, 1 = DLP.  out: xi, rc, w, eta; iout: nside_x, nside_y, Mx, My, kinf,
 * A, B (the rat() pair).  Returns 0 or a negative error code. */
int orc_params(int kind, int npts, double Q, double Lx, double Ly, int P, double Nb,
               double tol, double* out, int* iout);

/* MATLAB rat(X) with its default tolerance 1e-6*norm(X,1). */
void orc_rat(double X, int* N, int* D);

/* Derived quantity: SLP pressure (mex_stokes_slp_pressure_{real,kspace}.cpp,
 * StokesSLP_pressure_ewald_2p.m, tests/direct_sums/matlab_sums/
 * stokes_slp_pressure_{real,kspace}_ds.m).  One value per target. */
void orc_slp_pressure_real(const double* psrc, const double* ptar, const double* f, int Ns,
                           int Nt, double xi, int nside_x, int nside_y, double Lx, double Ly,
                           double* p);
void orc_slp_pressure_filter(double* h1re, double* h1im, const double* h2re,
                             const double* h2im, int Mx, int My, double Lx, double Ly,
                             double xi, double eta);
void orc_slp_pressure_kspace(const double* psrc, const double* ptar, const double* f,
                             double xi, double eta, int Ns, int Nt, int Mx, int My, double Lx,
                             double Ly, double w, int P, double* pk);
double orc_slp_pressure_zero_mode(const double* psrc, const double* f, int Ns, double Lx,
                                  double Ly);
void orc_slp_pressure_real_ds(const double* psrc, const double* ptar, const double* f, int Ns,
                              int Nt, double Lx, double Ly, double xi, double* p);
void orc_slp_pressure_kspace_ds(const double* psrc, const double* ptar, const double* f,
                                int Ns, int Nt, double Lx, double Ly, double xi, int kinf,
                                double* p);


/* Derived quantity: DLP pressure (mex_stokes_dlp_pressure_{real,kspace}.cpp,
 * StokesDLP_pressure_ewald_2p.m, stokes_dlp_pressure_{real,kspace}_ds.m). */
void orc_dlp_pressure_real(const double* psrc, const double* ptar, const double* f,
                           const double* n, int Ns, int Nt, double xi, int nside_x,
                           int nside_y, double Lx, double Ly, double* p);
void orc_dlp_pressure_filter(double* h1re, double* h1im, const double* h2re,
                             const double* h2im, const double* h3re, const double* h3im,
                             const double* h4re, const double* h4im, int Mx, int My, double Lx,
                             double Ly, double xi, double eta);
void orc_dlp_pressure_kspace(const double* psrc, const double* ptar, const double* f,
                             const double* n, double xi, double eta, int Ns, int Nt, int Mx,
                             int My, double Lx, double Ly, double w, int P, double* pk);
double orc_dlp_pressure_zero_mode(const double* f, const double* n, int Ns, double Lx,
                                  double Ly);
void orc_dlp_pressure_real_ds(const double* psrc, const double* ptar, const double* f,
                              const double* n, int Ns, int Nt, double Lx, double Ly, double xi,
                              double* p);
void orc_dlp_pressure_kspace_ds(const double* psrc, const double* ptar, const double* f,
                                const double* n, int Ns, int Nt, double Lx, double Ly,
                                double xi, int kinf, double* p);

/* Derived quantities, kinds 4..11: velocity gradient, pressure gradient,
 * vorticity, stress of the Stokeslet, then of the stresslet (mex_stokes_{slp,dlp}_
 * {gradient,pressure_grad,vorticity,stress}_{real,kspace}.cpp and the direct
 * sums stokes_*_{real,kspace}_ds.m).  out: nc x Nt with nc = orc_kind_ncomp;
 * n is ignored for Stokeslet kinds.  The direct sums project the gradient and
 * the stress on the per-target vector b (2 x Nt) and write 2 x Nt (1 x Nt for
 * the vorticity). */
int orc_kind_ncomp(int kind);
void orc_derived_real(int kind, const double* psrc, const double* ptar, const double* f,
                      const double* n, int Ns, int Nt, double xi, int nside_x, int nside_y,
                      double Lx, double Ly, double* out);
void orc_derived_kspace(int kind, const double* psrc, const double* ptar, const double* f,
                        const double* n, double xi, double eta, int Ns, int Nt, int Mx, int My,
                        double Lx, double Ly, double w, int P, double* out);
void orc_derived_real_ds(int kind, const double* psrc, const double* ptar, const double* f,
                         const double* n, const double* b, int Ns, int Nt, double Lx, double Ly,
                         double xi, double* out);
void orc_derived_kspace_ds(int kind, const double* psrc, const double* ptar, const double* f,
                           const double* n, const double* b, int Ns, int Nt, double Lx,
                           double Ly, double xi, int kinf, double* out);

/* Free-space direct sums (FMM/tests/stokes_{SLP,DLP}_FMM_test.m): SLP at Nt
 * targets (self: targets are the sources, own term left out), DLP at the
 * sources; u is 2 x Nt interleaved. */
void orc_free_slp(const double* f, const double* psrc, int Ns, const double* ptar, int Nt,
                  int self, double* u);
void orc_free_dlp(const double* f, const double* n, const double* p, int N, double* u);

#ifdef __cplusplus
}
#endif

#endif
