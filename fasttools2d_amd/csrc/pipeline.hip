// The evaluation pipeline split into geometry-dependent set-up (build) and
// density-dependent work (apply).
//
// build: cell binning of sources and targets (stable counting sort, then Morton
// order inside each cell), sorted position records, grid-tile binning and
// start nodes for spreading and gathering, the FGG e1 table, FFT twiddles and
// every grid buffer.  apply: density packing, real-space sum, spreading,
// FFTs, filter, gather -- kernel launches only, no allocation, no host
// synchronisation, so a plan's apply can be captured in a HIP graph.
//
// The reference recomputes all of build on every mex call
// (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_real.cpp:98-100,
// mex/common/ewald_tools.cpp:125-127,152-159); the one-shot entry points here do
// the same, plans keep build across calls.
#include "se2p_internal.h"
#include "e1_eval.h"

namespace se2p {

namespace {

__global__ void k_fill_slp_dens(const int* __restrict__ perm, int n, const double* __restrict__ f,
                                double4* __restrict__ rec)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    rec[i].z = f[2 * o];
    rec[i].w = f[2 * o + 1];
}

// DLP records {x, y, f1, f2, n1, n2, f.n, 0} (8 doubles): fill f1, f2, f.n.
__global__ void k_fill_dlp_dens(const int* __restrict__ perm, int n, const double* __restrict__ f,
                                double* __restrict__ rec)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    double* r = rec + 8 * (size_t)i;
    const double f1 = f[2 * o], f2 = f[2 * o + 1];
    r[2] = f1;
    r[3] = f2;
    r[6] = f1 * r[4] + f2 * r[5];
}

__global__ void k_copy_d(const double* __restrict__ a, double* __restrict__ b, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

}  // namespace

int geom_build(se2p_ctx* ctx, Workspace* ws, Geom* G, int kind, const double* d_psrc,
               const double* d_ptar, const double* d_n, int Ns, int Nt, double Lx, double Ly,
               const se2p_params& prm, bool want_real, bool want_k)
{
    hipStream_t s = ctx->stream;
    *G = Geom{};
    G->kind = kind;
    G->Ns = Ns;
    G->Nt = Nt;
    G->Lx = Lx;
    G->Ly = Ly;
    G->prm = prm;
    G->has_real = want_real;
    G->has_k = want_k;
    G->e1tier = prm.e1_tier == 1 ? kE1Fast : kE1Full;
    void* p;
    if (kind == SE2P_DLP && Ns > 0) {
        // normals are geometry: keep a copy in input order for the spreading passes
        SE2P_TRY(ws_get(ctx, ws, WS_NRM, sizeof(double) * 2 * (size_t)Ns, &p));
        G->nrm = (double*)p;
        const long n2 = 2L * Ns;
        hipLaunchKernelGGL(k_copy_d, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, d_n,
                           G->nrm, n2);
        SE2P_HIP_CHECK(hipGetLastError());
    }
    if (want_real) {
        const int nsx = prm.nside_x, nsy = prm.nside_y;
        const int nbox = nsx * nsy;
        G->nbox = nbox;
        void *skey, *scount, *tkey, *tcount;
        SE2P_TRY(ws_get(ctx, ws, WS_SRC_KEY, sizeof(int) * (size_t)Ns, &skey));
        SE2P_TRY(ws_get(ctx, ws, WS_SRC_COUNT, sizeof(int) * (size_t)nbox, &scount));
        SE2P_TRY(ws_get(ctx, ws, WS_TAR_KEY, sizeof(int) * (size_t)Nt, &tkey));
        SE2P_TRY(ws_get(ctx, ws, WS_TAR_COUNT, sizeof(int) * (size_t)nbox, &tcount));
        SE2P_TRY(ws_get(ctx, ws, WS_SRC_PERM, sizeof(int) * (size_t)Ns, &p));
        G->s_perm = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SRC_START, sizeof(int) * (size_t)(nbox + 1), &p));
        G->s_start = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_TAR_PERM, sizeof(int) * (size_t)Nt, &p));
        G->t_perm = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_TAR_START, sizeof(int) * (size_t)(nbox + 1), &p));
        G->t_start = (int*)p;
        const size_t rec = kind == SE2P_SLP ? 32 : 64;
        SE2P_TRY(ws_get(ctx, ws, WS_SRC_SORTED, rec * (size_t)Ns, &p));
        G->s_rec = (double*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_TAR_SORTED, sizeof(double2) * (size_t)Nt, &p));
        G->t_sorted = (double2*)p;
        StageTimer st(ctx, "bin");
        SE2P_TRY(launch_fill_int(s, (int*)scount, nbox, 0));
        SE2P_TRY(launch_cell_keys(s, d_psrc, Ns, Lx, Ly, nsx, nsy, (int*)skey, (int*)scount));
        SE2P_TRY(sort_by_key(ctx, ws, (int*)skey, (int*)scount, Ns, nbox, G->s_perm, G->s_start,
                             WS_SORT_TMP));
        SE2P_TRY(launch_fill_int(s, (int*)tcount, nbox, 0));
        SE2P_TRY(launch_cell_keys(s, d_ptar, Nt, Lx, Ly, nsx, nsy, (int*)tkey, (int*)tcount));
        SE2P_TRY(sort_by_key(ctx, ws, (int*)tkey, (int*)tcount, Nt, nbox, G->t_perm, G->t_start,
                             WS_SORT_TMP));
        SE2P_TRY(launch_morton_cells(s, G->s_perm, G->s_start, nbox, d_psrc, Lx, Ly, nsx, nsy));
        SE2P_TRY(launch_morton_cells(s, G->t_perm, G->t_start, nbox, d_ptar, Lx, Ly, nsx, nsy));
        // positions (and normals) of the sorted sources; densities are filled per apply
        SE2P_TRY(launch_pack_sources(s, kind, G->s_perm, Ns, d_psrc, d_psrc, G->nrm, G->s_rec));
        SE2P_TRY(launch_pack_targets(s, G->t_perm, Nt, d_ptar, G->t_sorted));
        if (kind == SE2P_DLP) {
            SE2P_TRY(ws_get(ctx, ws, WS_REDUCE, sizeof(double) * 512, &p));
            G->zm_part = (double*)p;
            SE2P_TRY(ws_get(ctx, ws, WS_ZM, sizeof(double) * 2, &p));
            G->zm = (double*)p;
        }
    }
    if (want_k) {
        const int Mx = prm.Mx, My = prm.My, P = prm.P;
        SE2P_TRY(kgeom_init(&G->kg, Mx, My, P, Lx, Ly, prm.xi, prm.eta, prm.w));
        SE2P_TRY(fft_axis_init(&G->ax, Mx));
        SE2P_TRY(fft_axis_init(&G->ay, My));
        SE2P_TRY(fft_twiddles(ctx, ws, Mx, WS_TWIDDLE_X, &G->twx));
        SE2P_TRY(fft_twiddles(ctx, ws, My, WS_TWIDDLE_Y, &G->twy));
        const KGeom& g = G->kg;
        const int ntiles = g.ntx * g.nty;
        const size_t Gn = (size_t)Mx * My;
        void *skey, *scount, *tkey, *tcount;
        SE2P_TRY(ws_get(ctx, ws, WS_E1TAB, sizeof(double) * 64, &p));
        G->e1tab = (double*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_KEY, sizeof(int) * (size_t)Ns, &skey));
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_COUNT, sizeof(int) * (size_t)ntiles, &scount));
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_PERM, sizeof(int) * (size_t)Ns, &p));
        G->sp_perm = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_START, sizeof(int) * (size_t)(ntiles + 1), &p));
        G->sp_start = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_NODE, sizeof(int2) * (size_t)Ns, &p));
        G->sp_node = (int2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_OFF, sizeof(double2) * (size_t)Ns, &p));
        G->sp_off = (double2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_KEY, sizeof(int) * (size_t)Nt, &tkey));
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_COUNT, sizeof(int) * (size_t)ntiles, &tcount));
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_PERM, sizeof(int) * (size_t)Nt, &p));
        G->gt_perm = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_START, sizeof(int) * (size_t)(ntiles + 1), &p));
        G->gt_start = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_NODE, sizeof(int2) * (size_t)Nt, &p));
        G->gt_node = (int2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_OFF, sizeof(double2) * (size_t)Nt, &p));
        G->gt_off = (double2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_PARTIAL, sizeof(double) * 2 * (size_t)g.R * g.R * ntiles, &p));
        G->partial = (double*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GRID_A, sizeof(double2) * Gn, &p));
        G->Z0 = (double2*)p;
        if (kind == SE2P_DLP) {
            SE2P_TRY(ws_get(ctx, ws, WS_GRID_B, sizeof(double2) * Gn, &p));
            G->Z1 = (double2*)p;
        }
        SE2P_TRY(ws_get(ctx, ws, WS_GRID_C, sizeof(double2) * Gn, &p));
        G->W = (double2*)p;
        if (kind == SE2P_DLP) {
            SE2P_TRY(ws_get(ctx, ws, WS_GRID_D, sizeof(double2) * Gn, &p));
            G->D = (double2*)p;
        }
        StageTimer st(ctx, "bin_grid");
        SE2P_TRY(launch_e1_table(s, g, G->e1tab));
        SE2P_TRY(launch_fill_int(s, (int*)scount, ntiles, 0));
        SE2P_TRY(launch_tile_keys(s, d_psrc, Ns, Lx, Ly, g.h, P, Mx, My, g.TX, g.TY, g.ntx,
                                  (int*)skey, (int*)scount, G->sp_node, G->sp_off));
        SE2P_TRY(sort_by_key(ctx, ws, (int*)skey, (int*)scount, Ns, ntiles, G->sp_perm,
                             G->sp_start, WS_SORT_TMP));
        SE2P_TRY(launch_fill_int(s, (int*)tcount, ntiles, 0));
        SE2P_TRY(launch_tile_keys(s, d_ptar, Nt, Lx, Ly, g.h, P, Mx, My, g.TX, g.TY, g.ntx,
                                  (int*)tkey, (int*)tcount, G->gt_node, G->gt_off));
        SE2P_TRY(sort_by_key(ctx, ws, (int*)tkey, (int*)tcount, Nt, ntiles, G->gt_perm,
                             G->gt_start, WS_SORT_TMP));
    }
    return SE2P_OK;
}

int geom_apply_real(se2p_ctx* ctx, const Geom& G, const double* d_f, double* d_u)
{
    hipStream_t s = ctx->stream;
    if (G.Nt == 0) return SE2P_OK;
    {
        StageTimer st(ctx, "pack");
        if (G.Ns > 0) {
            dim3 gr((G.Ns + 255) / 256), b(256);
            if (G.kind == SE2P_SLP)
                hipLaunchKernelGGL(k_fill_slp_dens, gr, b, 0, s, G.s_perm, G.Ns, d_f,
                                   (double4*)G.s_rec);
            else
                hipLaunchKernelGGL(k_fill_dlp_dens, gr, b, 0, s, G.s_perm, G.Ns, d_f, G.s_rec);
            SE2P_HIP_CHECK(hipGetLastError());
        }
        if (G.kind == SE2P_DLP)
            SE2P_TRY(launch_dlp_zero_mode_rec(s, G.s_rec, G.Ns, G.Lx, G.Ly, G.zm_part, G.zm));
    }
    StageTimer st(ctx, "real");
    return launch_real(s, G.kind, G.e1tier, G.prm.nside_x, G.prm.nside_y, G.Lx, G.Ly, G.prm.xi,
                       G.t_start, G.t_perm, G.t_sorted, G.s_start, G.s_rec,
                       (G.Nt + G.nbox - 1) / G.nbox, d_u);
}

int geom_apply_k(se2p_ctx* ctx, const Geom& G, const double* d_f, const double* d_add_first,
                 const double* d_zm, double* d_u)
{
    hipStream_t s = ctx->stream;
    if (G.Nt == 0) return SE2P_OK;
    const KGeom& g = G.kg;
    // Every FFT pass runs on contiguous lines: y lines in the spread layout
    // (Z[ix*My + iy]), then a transpose to Z[iy*Mx + ix] for the x lines, the
    // filter in that layout, and back.
    //   SLP: Z0 -yFFT-> -T-> W -xFFT-> -filter-> Z0 -xIFFT-> -T-> W -yIFFT-> gather(W)
    //   DLP: Z0, Z1 -yFFT-> -T-> W, D -xFFT-> -filter-> Z0 -xIFFT-> -T-> W -yIFFT-> gather(W)
    const int nspread = G.kind == SE2P_SLP ? 1 : 2;
    for (int c = 0; c < nspread; ++c) {
        double2* Zc = c == 0 ? G.Z0 : G.Z1;
        double2* Tc = c == 0 ? G.W : G.D;
        {
            StageTimer st(ctx, "spread");
            SE2P_TRY(launch_spread(s, G.kind == SE2P_SLP ? 0 : 1 + c, g, G.sp_start, G.sp_perm,
                                   G.sp_node, G.sp_off, d_f, G.nrm, G.e1tab, G.partial));
        }
        {
            StageTimer st(ctx, "merge");
            SE2P_TRY(launch_merge(s, 2, g, G.partial, Zc, nullptr));
        }
        {
            StageTimer st(ctx, "fft");
            SE2P_TRY(fft_lines_contig(ctx, Zc, g.My, g.Mx, G.ay, G.twy, 0));
            SE2P_TRY(transpose(ctx, Zc, Tc, g.Mx, g.My));
            SE2P_TRY(fft_lines_contig(ctx, Tc, g.Mx, g.My, G.ax, G.twx, 0));
        }
    }
    {
        StageTimer st(ctx, "filter");
        SE2P_TRY(launch_filter(s, G.kind, g, G.W, G.D, G.Z0, 1));
    }
    {
        StageTimer st(ctx, "fft");
        SE2P_TRY(fft_lines_contig(ctx, G.Z0, g.Mx, g.My, G.ax, G.twx, 1));
        SE2P_TRY(transpose(ctx, G.Z0, G.W, g.My, g.Mx));
        SE2P_TRY(fft_lines_contig(ctx, G.W, g.My, g.Mx, G.ay, G.twy, 1));
    }
    StageTimer st(ctx, "gather");
    return launch_gather(s, g, G.gt_start, G.gt_perm, G.gt_node, G.gt_off, G.W, G.e1tab,
                         d_add_first, d_zm, d_u);
}

int geom_apply(se2p_ctx* ctx, const Geom& G, const double* d_f, double* d_u)
{
    SE2P_TRY(geom_apply_real(ctx, G, d_f, d_u));
    return geom_apply_k(ctx, G, d_f, d_u, G.kind == SE2P_DLP ? G.zm : nullptr, d_u);
}

}  // namespace se2p
