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

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

namespace se2p {

namespace {

__global__ void k_copy_d(const double* __restrict__ a, double* __restrict__ b, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// The strip spreading jobs: each strip's rows cut into segments of at most S
// rows, halved while a segment's records (start rows [y0 - P, y1), the P rows
// before it build up its ring) exceed 1.5 times those of an S-row segment at the
// mean density, so a strip through a dense cluster runs as several shorter jobs
// instead of one long one; never below P rows (or 2), which would repeat each
// dense row's records in more than two segments' build-up.  st[s My + r] = first
// record of strip s, start row r (st[ntx My] = nrec).
std::vector<int4> strip_segments(const std::vector<int>& st, int ntx, int My, int P, int S,
                                 long nrec)
{
    std::vector<int4> out;
    const double mean = (double)nrec / ((double)ntx * My);
    const double cap = std::max(1.5 * mean * (S + P), 128.0);  // at least four chunks
    for (int s = 0; s < ntx; ++s) {
        const int* c = st.data() + (size_t)s * My;
        // records of strip s with start rows in [a, b), rows taken periodically
        auto rows = [&](int a, int b) {
            long n = 0;
            while (a < b) {
                int r0 = a % My;
                r0 = r0 < 0 ? r0 + My : r0;
                const int take = std::min(b - a, My - r0);
                n += c[r0 + take] - c[r0];
                a += take;
            }
            return n;
        };
        for (int y0 = 0; y0 < My;) {
            int len = std::min(S, My - y0);
            const int lmin = std::max(2, P);
            while (len / 2 >= lmin && (double)rows(y0 - P, y0 + len) > cap) len = (len + 1) / 2;
            out.push_back(make_int4(s, y0, y0 + len, 0));
            y0 += len;
        }
    }
    return out;
}

}  // namespace

int geom_build(se2p_ctx* ctx, Workspace* ws, Geom* G, int kind, const double* d_psrc,
               const double* d_ptar, const double* d_n, int Ns, int Nt, double Lx, double Ly,
               const se2p_params& prm, bool want_real, bool want_k)
{
    hipStream_t s = ctx->stream;
    if (prm.free_R > 0 && (kind != SE2P_SLP || !want_real || !want_k))
        return set_error(SE2P_ERR_UNSUPPORTED,
                         "free-space sums (free_R > 0) are Stokeslet velocity sums with both parts");
    *G = Geom{};
    G->kind = kind;
    G->Ns = Ns;
    G->Nt = Nt;
    G->Lx = Lx;
    G->Ly = Ly;
    G->prm = prm;
    G->has_real = want_real;
    G->has_k = want_k;
    // the FP32 tier covers the velocity kinds; the derived kinds stay FP64
    G->prec = prm.precision == SE2P_FP32C && (kind == SE2P_SLP || kind == SE2P_DLP) ? 1 : 0;
    G->e1tier = G->prec ? kF32 : prm.e1_tier == 1 ? kE1Fast : kE1Full;
    void* p;
    if (dlp_family(kind) && Ns > 0) {
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
        // pair-symmetric velocity sums need the targets to be the sources (they
        // then share the sources' order) and a k-space part (its gather adds them)
        const bool sym = prm.symmetric == 1 && (kind == SE2P_SLP || kind == SE2P_DLP) &&
                         d_psrc == d_ptar && Ns == Nt && want_k;
        if (sym) {
            G->t_perm = G->s_perm;
            G->t_start = G->s_start;
            SE2P_TRY(ws_get(ctx, ws, WS_SRC_YKEY, sizeof(int) * (size_t)Ns, &p));
            G->ykey = (int*)p;
            SE2P_TRY(ws_get(ctx, ws, WS_SYM_ACC, sizeof(double2) * (size_t)Ns, &p));
            G->sacc = (double2*)p;
            SE2P_TRY(ws_get(ctx, ws, WS_SYM_POS, sizeof(int) * (size_t)Ns, &p));
            G->spos = (int*)p;
        } else {
            SE2P_TRY(ws_get(ctx, ws, WS_TAR_PERM, sizeof(int) * (size_t)Nt, &p));
            G->t_perm = (int*)p;
            SE2P_TRY(ws_get(ctx, ws, WS_TAR_START, sizeof(int) * (size_t)(nbox + 1), &p));
            G->t_start = (int*)p;
        }
        const size_t rec = rec_layout(kind) == 0 ? 32 : 64;
        SE2P_TRY(ws_get(ctx, ws, WS_SRC_SORTED, rec * (size_t)Ns, &p));
        G->s_rec = (double*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_TAR_SORTED, sizeof(double2) * (size_t)Nt, &p));
        G->t_sorted = (double2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_UNSORTED, sizeof(int), &p));
        G->unsorted = (int*)p;
        StageTimer st(ctx, "bin");
        SE2P_TRY(launch_fill_int(s, (int*)scount, nbox, 0));
        SE2P_TRY(launch_cell_keys(s, d_psrc, Ns, Lx, Ly, nsx, nsy, (int*)skey, (int*)scount));
        SE2P_TRY(sort_by_key(ctx, ws, (int*)skey, (int*)scount, Ns, nbox, G->s_perm, G->s_start,
                             WS_SORT_TMP));
        // the sources per cell on the host: the cells too large for the in-LDS y
        // order are ordered by a second kernel, and (sources = targets) the job
        // lists split heavy blocks
        std::vector<int> sst((size_t)nbox + 1);
        SE2P_HIP_CHECK(hipMemcpyAsync(sst.data(), G->s_start, sizeof(int) * ((size_t)nbox + 1),
                                      hipMemcpyDeviceToHost, s));
        SE2P_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<int> big;
        for (int c = 0; c < nbox; ++c)
            if (sst[c + 1] - sst[c] > kSortCellCap) big.push_back(c);
        if (!sym) {
            SE2P_TRY(launch_fill_int(s, (int*)tcount, nbox, 0));
            SE2P_TRY(launch_cell_keys(s, d_ptar, Nt, Lx, Ly, nsx, nsy, (int*)tkey, (int*)tcount));
            SE2P_TRY(sort_by_key(ctx, ws, (int*)tkey, (int*)tcount, Nt, nbox, G->t_perm,
                                 G->t_start, WS_SORT_TMP));
        }
        // sources by y inside their cells (each stencil column is then sorted by
        // y), targets along a Morton curve (neighbouring lanes, similar stencils)
        SE2P_TRY(launch_fill_int(s, G->unsorted, 1, 0));
        SE2P_TRY(launch_morton_cells(s, G->s_perm, G->s_start, nbox, d_psrc, Lx, Ly, nsx, nsy, 1,
                                     G->unsorted, G->ykey, big.empty() ? 0 : 1));
        if (!big.empty()) {
            void *pc, *pt;
            SE2P_TRY(ws_get(ctx, ws, WS_BIGCELLS, sizeof(int) * big.size(), &pc));
            SE2P_TRY(ws_get(ctx, ws, WS_YSORT, sizeof(int) * 4 * (size_t)Ns, &pt));
            SE2P_HIP_CHECK(hipMemcpyAsync(pc, big.data(), sizeof(int) * big.size(),
                                          hipMemcpyHostToDevice, s));
            SE2P_TRY(launch_ysort_big(s, G->s_perm, G->s_start, (const int*)pc, (int)big.size(),
                                      d_psrc, Ly, nsx, nsy, G->unsorted, G->ykey, (int*)pt, Ns));
            // `big` is read by the copy above before it goes out of scope
            SE2P_HIP_CHECK(hipStreamSynchronize(s));
        }
        if (!sym)
            SE2P_TRY(launch_morton_cells(s, G->t_perm, G->t_start, nbox, d_ptar, Lx, Ly, nsx,
                                         nsy, 0, G->unsorted));
        if (sym) {
            SE2P_TRY(launch_inverse_perm(s, G->s_perm, Ns, G->spos));
            // the symmetric sums rely on the y order; without it (a cell too large
            // to sort, points outside the box) each target sums its own pairs
            int uns = 1;
            SE2P_HIP_CHECK(hipMemcpyAsync(&uns, G->unsorted, sizeof(int), hipMemcpyDeviceToHost, s));
            SE2P_HIP_CHECK(hipStreamSynchronize(s));
            G->sym = uns == 0;
        }
        // the targets per cell, for the job lists: a block of cells holding more
        // targets than a workgroup has lanes is split into one job per 256 targets
        if (sym) {
            G->t_start_h = std::move(sst);
        } else {
            G->t_start_h.resize((size_t)nbox + 1);
            SE2P_HIP_CHECK(hipMemcpyAsync(G->t_start_h.data(), G->t_start,
                                          sizeof(int) * ((size_t)nbox + 1),
                                          hipMemcpyDeviceToHost, s));
            SE2P_HIP_CHECK(hipStreamSynchronize(s));
        }
        // positions (and normals) of the sorted sources; densities are filled per apply
        SE2P_TRY(launch_pack_sources(s, kind, G->s_perm, Ns, d_psrc, d_psrc, G->nrm, G->s_rec));
        SE2P_TRY(launch_pack_targets(s, G->t_perm, Nt, d_ptar, G->t_sorted));
        if (has_zero_mode(kind, prm.free_R)) {
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
        // spreading bins: (strip, start row), strip-major (ring), or tiles of
        // kSB x kSB start nodes (blocks)
        const long nbins = g.ring ? (long)g.ntx * My : (long)g.ntx * g.nty;
        const int gtiles = g.gntx * g.gnty;
        if (nbins >= INT_MAX) return set_error(SE2P_ERR_UNSUPPORTED, "grid too large");
        const size_t Gn = (size_t)Mx * My;
        void *skey, *scount, *tkey, *tcount, *scnt = nullptr, *soffs = nullptr, *rnode = nullptr,
             *owner = nullptr, *rperm;
        SE2P_TRY(ws_get(ctx, ws, WS_E1TAB, sizeof(double) * 64, &p));
        G->e1tab = (double*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_NODE, sizeof(int2) * (size_t)Ns, &p));
        G->sp_node = (int2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_OFF, sizeof(double2) * (size_t)Ns, &p));
        G->sp_off = (double2*)p;
        if (g.ring) {
            SE2P_TRY(ws_get(ctx, ws, WS_SPR_CNT, sizeof(int) * (size_t)std::max(Ns, 1), &scnt));
            SE2P_TRY(ws_get(ctx, ws, WS_SPR_OFFS, sizeof(int) * ((size_t)Ns + 1), &soffs));
        }
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_COUNT, sizeof(int) * (size_t)nbins, &scount));
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_START, sizeof(int) * (size_t)(nbins + 1), &p));
        G->sp_start = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_KEY, sizeof(int) * (size_t)Nt, &tkey));
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_COUNT, sizeof(int) * (size_t)gtiles, &tcount));
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_PERM, sizeof(int) * (size_t)Nt, &p));
        G->gt_perm = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_START, sizeof(int) * (size_t)(gtiles + 1), &p));
        G->gt_start = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_NODE, sizeof(int2) * (size_t)Nt, &p));
        G->gt_node = (int2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_OFF, sizeof(double2) * (size_t)Nt, &p));
        G->gt_off = (double2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GRID_A, sizeof(double2) * Gn, &p));
        G->Z0 = (double2*)p;
        // second grid pair: stresslet spreading (f n2) or a second output grid
        const bool two = dlp_family(kind) || out_grids(kind) == 2;
        if (two) {
            SE2P_TRY(ws_get(ctx, ws, WS_GRID_B, sizeof(double2) * Gn, &p));
            G->Z1 = (double2*)p;
        }
        SE2P_TRY(ws_get(ctx, ws, WS_GRID_C, sizeof(double2) * Gn, &p));
        G->W = (double2*)p;
        if (two) {
            SE2P_TRY(ws_get(ctx, ws, WS_GRID_D, sizeof(double2) * Gn, &p));
            G->D = (double2*)p;
        }
        StageTimer st(ctx, "bin_grid");
        SE2P_TRY(launch_e1_table(s, g, G->e1tab));
        // spreading records (ring: count per point, offsets, then keys; blocks:
        // one per point)
        int nrec = Ns;
        if (g.ring) {
            SE2P_TRY(launch_strip_count(s, d_psrc, Ns, Lx, Ly, g.h, P, Mx, My, g.ntx, (int*)scnt,
                                        G->sp_node, G->sp_off));
            SE2P_TRY(scan_ints(ctx, ws, (const int*)scnt, Ns, (int*)soffs));
            SE2P_HIP_CHECK(hipMemcpyAsync(&nrec, (int*)soffs + Ns, sizeof(int),
                                          hipMemcpyDeviceToHost, s));
            SE2P_HIP_CHECK(hipStreamSynchronize(s));
        }
        G->sp_nrec = nrec;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_KEY, sizeof(int) * (size_t)std::max(nrec, 1), &skey));
        if (g.ring) {
            SE2P_TRY(ws_get(ctx, ws, WS_SPR_RNODE, sizeof(int2) * (size_t)std::max(nrec, 1),
                            &rnode));
            SE2P_TRY(ws_get(ctx, ws, WS_SPR_OWNER, sizeof(int) * (size_t)std::max(nrec, 1),
                            &owner));
        }
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_RPERM, sizeof(int) * (size_t)std::max(nrec, 1), &rperm));
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_PERM, sizeof(int) * (size_t)std::max(nrec, 1), &p));
        G->sp_perm = (int*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_SNODE, sizeof(int2) * (size_t)std::max(nrec, 1), &p));
        G->sp_snode = (int2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_SPR_FAC, sizeof(double4) * (size_t)std::max(nrec, 1), &p));
        G->sp_fac = (double4*)p;
        SE2P_TRY(launch_fill_int(s, (int*)scount, (int)nbins, 0));
        if (g.ring) {
            SE2P_TRY(launch_strip_keys(s, G->sp_node, (const int*)soffs, Ns, P, Mx, My, g.ntx,
                                       (int*)skey, (int*)scount, (int2*)rnode, (int*)owner));
        } else {
            SE2P_TRY(launch_tile_keys(s, d_psrc, Ns, Lx, Ly, g.h, P, Mx, My, g.TX, g.TY, g.ntx,
                                      (int*)skey, (int*)scount, G->sp_node, G->sp_off));
            rnode = G->sp_node;  // one record per point: its grid start node
            owner = nullptr;
        }
        SE2P_TRY(sort_by_key(ctx, ws, (int*)skey, (int*)scount, nrec, (int)nbins, (int*)rperm,
                             G->sp_start, WS_SORT_TMP));
        // the records per strip and start row, for the strip segments
        std::vector<int> sst_h;
        if (g.ring) {
            sst_h.resize((size_t)nbins + 1);
            SE2P_HIP_CHECK(hipMemcpyAsync(sst_h.data(), G->sp_start,
                                          sizeof(int) * ((size_t)nbins + 1),
                                          hipMemcpyDeviceToHost, s));
        }
        SE2P_TRY(launch_spread_records(s, g, G->prec, (const int*)rperm, (const int2*)rnode,
                                       (const int*)owner, G->sp_off, nrec, G->sp_snode,
                                       G->sp_fac, G->sp_perm));
        SE2P_TRY(launch_fill_int(s, (int*)tcount, gtiles, 0));
        SE2P_TRY(launch_tile_keys(s, d_ptar, Nt, Lx, Ly, g.h, P, Mx, My, g.gTX, g.gTY, g.gntx,
                                  (int*)tkey, (int*)tcount, G->gt_node, G->gt_off));
        SE2P_TRY(sort_by_key(ctx, ws, (int*)tkey, (int*)tcount, Nt, gtiles, G->gt_perm,
                             G->gt_start, WS_SORT_TMP));
        SE2P_TRY(launch_gather_order(s, G->gt_perm, G->gt_start, G->gt_node, gtiles, g.gntx,
                                     g.gTX, g.gTY, g.gR));
        // the targets per gathering tile, for the job lists (heavy tiles are split)
        G->gt_start_h.resize((size_t)gtiles + 1);
        SE2P_HIP_CHECK(hipMemcpyAsync(G->gt_start_h.data(), G->gt_start,
                                      sizeof(int) * ((size_t)gtiles + 1), hipMemcpyDeviceToHost,
                                      s));
        SE2P_HIP_CHECK(hipStreamSynchronize(s));
        if (g.ring) {
            const std::vector<int4> segs = strip_segments(sst_h, g.ntx, My, P, g.sseg, nrec);
            if (segs.size() >= (size_t)INT_MAX)
                return set_error(SE2P_ERR_UNSUPPORTED, "too many spreading segments");
            SE2P_TRY(ws_get(ctx, ws, WS_SPR_SEGS, sizeof(int4) * segs.size(), &p));
            G->sp_segs = (int4*)p;
            G->sp_nseg = (int)segs.size();
            SE2P_HIP_CHECK(hipMemcpyAsync(G->sp_segs, segs.data(), sizeof(int4) * segs.size(),
                                          hipMemcpyHostToDevice, s));
            SE2P_HIP_CHECK(hipStreamSynchronize(s));
        }
        // the gather reads its targets' records in binned order (contiguous)
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_BNODE, sizeof(int2) * (size_t)Nt, &p));
        G->gt_bnode = (int2*)p;
        SE2P_TRY(ws_get(ctx, ws, WS_GAT_BOFF, sizeof(double2) * (size_t)Nt, &p));
        G->gt_boff = (double2*)p;
        if (G->sym) {
            SE2P_TRY(ws_get(ctx, ws, WS_GAT_SIDX, sizeof(int) * (size_t)Nt, &p));
            G->gt_sidx = (int*)p;
        }
        SE2P_TRY(launch_gather_records(s, G->gt_perm, G->gt_node, G->gt_off,
                                       G->sym ? G->spos : nullptr, Nt, G->gt_bnode, G->gt_boff,
                                       G->gt_sidx));
    }
    return SE2P_OK;
}

void sysdesc_from_geom(const Geom& G, long f_off, long u_off, SysDesc* d)
{
    *d = SysDesc{};
    d->kind = G.kind;
    d->nc = out_comps(G.kind);
    d->Ns = G.Ns;
    d->Nt = G.Nt;
    d->e1tier = G.e1tier;
    d->prec = G.prec;
    d->f_off = f_off;
    d->u_off = u_off;
    d->Lx = G.Lx;
    d->Ly = G.Ly;
    d->nrm = G.nrm;
    if (G.has_real) {
        const int nsx = G.prm.nside_x, nsy = G.prm.nside_y;
        const double xi = G.prm.xi;
        d->nsx = nsx;
        d->nsy = nsy;
        d->bs = real_block_side(nsx, nsy, (G.Nt + G.nbox - 1) / G.nbox);
        d->nbx = (nsx + d->bs - 1) / d->bs;
        d->xi2 = xi * xi;
        d->cut2 = G.Lx * G.Ly / nsx / nsy;
        // every neighbour candidate within the cutoff and every own-cell source
        // lies within max(rc, cell height) in y; the margin covers the 2^-20
        // resolution of the in-cell y order
        d->trim_w = std::max(std::sqrt(d->cut2), G.Ly / nsy) * (1.0 + 1e-5);
        d->far32_r2 = G.e1tier == kE1Fast && G.prm.far32 > 0 ? G.prm.far32 / (xi * xi)
                                                              : HUGE_VAL;
        d->free_R = G.prm.free_R;
        d->self_r2 = G.prm.free_R > 0 ? 4.9406564584124654e-324 : 1e-15;
        d->unsorted = G.unsorted;
        d->sym = G.sym ? 1 : 0;
        d->ykey = G.ykey;
        d->sacc = G.sacc;
        d->spos = G.spos;
        // -(1 + gamma/2 + ln xi) (mex_stokes_slp_real.cpp:124)
        d->self = -(1.0 + 0.5772156649015328606 / 2 + log(xi));
        d->t_start = G.t_start;
        d->t_perm = G.t_perm;
        d->s_start = G.s_start;
        d->s_perm = G.s_perm;
        d->t_sorted = G.t_sorted;
        d->s_rec = G.s_rec;
        d->zm = G.zm;
        d->zm_part = G.zm_part;
        d->zm_nblk = std::min(128, std::max(1, (G.Ns + 255) / 256));
    }
    if (G.has_k) {
        const KGeom& g = G.kg;
        d->Mx = g.Mx;
        d->My = g.My;
        d->P = g.P;
        d->TX = g.TX;
        d->TY = g.TY;
        d->ntx = g.ntx;
        d->nty = g.nty;
        d->sseg = g.sseg;
        d->snrec = G.sp_nrec;
        d->sp_segs = G.sp_segs;
        d->gTX = g.gTX;
        d->gTY = g.gTY;
        d->gntx = g.gntx;
        d->gnty = g.gnty;
        d->gR = g.gR;
        d->Bx = fft_lines_per_block(g.Mx);
        d->By = fft_lines_per_block(g.My);
        d->t = -2 * g.xi * g.xi / g.eta;
        d->w = g.w;
        d->h = g.h;
        d->xi = g.xi;
        d->eta = g.eta;
        // (4 xi^2/eta)^2 h^2/pi/(4 pi)  (ewald_tools.cpp:267-269)
        double sc = 4 * g.xi * g.xi / g.eta;
        sc = sc * sc * g.h * g.h / kPi;
        d->gscale = sc / (4 * kPi);
        d->sp_start = G.sp_start;
        d->sp_perm = G.sp_perm;
        d->gt_start = G.gt_start;
        d->gt_perm = G.gt_perm;
        d->sp_snode = G.sp_snode;
        d->sp_fac = G.sp_fac;
        d->gt_node = G.gt_bnode;
        d->gt_off = G.gt_boff;
        d->gt_sidx = G.sym ? G.gt_sidx : nullptr;
        d->e1tab = G.e1tab;
        d->grid[G_Z0] = G.Z0;
        d->grid[G_Z1] = G.Z1;
        d->grid[G_W] = G.W;
        d->grid[G_D] = G.D;
        d->out0 = fft_xfilter_fits(G.kind, G.kg.Mx, d->free_R) ? G_Z0 : G_W;
        d->twx = G.twx;
        d->twy = G.twy;
        d->ax = G.ax;
        d->ay = G.ay;
    }
}

int batch_build(se2p_ctx* ctx, Workspace* ws, BatchExec* B, const Geom* const* geoms, int nsys)
{
    *B = BatchExec{};
    B->nsys = nsys;
    std::vector<SysDesc> hs(nsys);
    long foff = 0, uoff = 0;
    for (int i = 0; i < nsys; ++i) {
        const Geom& G = *geoms[i];
        sysdesc_from_geom(G, foff, uoff, &hs[i]);
        foff += 2L * G.Ns;
        uoff += (long)out_comps(G.kind) * G.Nt;
        B->has_real |= G.has_real;
        B->has_k |= G.has_k;
        B->any_dlp |= dlp_family(G.kind);
        B->any_g1 |= G.has_k && out_grids(G.kind) == 2;
        if (G.has_k) {
            size_t& ly = B->lds_y[G.prec][fft_line_class(G.kg.My)];
            size_t& lx = B->lds_x[G.prec][fft_line_class(G.kg.Mx)];
            ly = std::max(ly, fft_lds_bytes(G.kg.My, G.prec));
            lx = std::max(lx, fft_lds_bytes(G.kg.Mx, G.prec));
            B->maxR = std::max(B->maxR, G.kg.gR);
        }
    }
    std::vector<Job> J;
    auto add = [&](JobRange& r, auto&& gen) {
        r.first = (int)J.size();
        for (int i = 0; i < nsys; ++i) gen(i, hs[i]);
        r.count = (int)J.size() - r.first;
    };
    auto blocks = [&](int i, long n) {
        for (long b = 0; b < n; ++b) J.push_back(make_int2(i, (int)b));
    };
    // Jobs 0 .. n-1 in an order where the workgroups that run on one XCD (every
    // 8th of a launch) take consecutive jobs: job k of group g is g*(n/8) + k (the
    // last n mod 8 jobs keep their place).  FFT lines read as the columns of a
    // row-major grid then share each cache line within one L2.
    auto xcd_blocks = [&](int i, long n) {
        const long q = n / 8;
        for (long b = 0; b < n; ++b)
            J.push_back(make_int2(i, (int)(b < 8 * q ? (b % 8) * q + b / 8 : b)));
    };
    // Real space: one job per block of bs x bs cells and 256 of its targets (job y
    // = block + part * nblk), so a dense block runs on as many workgroups as it
    // needs 256-lane passes instead of making one workgroup loop over them.
    bool too_many = false;
    auto real_jobs = [&](int i, const SysDesc& d, const Geom& G) {
        const long nby = (d.nsy + d.bs - 1) / d.bs, nblk = (long)d.nbx * nby;
        for (long b = 0; b < nblk; ++b) {
            const int bx0 = (int)(b % d.nbx) * d.bs, by0 = (int)(b / d.nbx) * d.bs;
            const int bw = std::min(d.bs, d.nsx - bx0), bh = std::min(d.bs, d.nsy - by0);
            long ntar = 0;
            for (int cy = 0; cy < bh; ++cy) {
                const int c0 = (by0 + cy) * d.nsx + bx0;
                ntar += G.t_start_h[c0 + bw] - G.t_start_h[c0];
            }
            const long parts = std::max(1L, (ntar + kRealLanes - 1) / kRealLanes);
            if (b + (parts - 1) * nblk > INT_MAX) too_many = true;
            for (long q = 0; q < parts && !too_many; ++q)
                J.push_back(make_int2(i, (int)(b + q * nblk)));
        }
    };
    // Gathering: one job per tile and kGatherPart of its targets (job y = tile +
    // part * ntiles).
    auto gather_jobs = [&](int i, const Geom& G) {
        const long nt = (long)G.kg.gntx * G.kg.gnty;
        for (long t = 0; t < nt; ++t) {
            const long cnt = G.gt_start_h[t + 1] - G.gt_start_h[t];
            const long parts = std::max(1L, (cnt + kGatherPart - 1) / kGatherPart);
            if (t + (parts - 1) * nt > INT_MAX) too_many = true;
            for (long q = 0; q < parts && !too_many; ++q)
                J.push_back(make_int2(i, (int)(t + q * nt)));
        }
    };
    for (int layout = 0; layout < 2; ++layout)
        add(B->fill[layout], [&](int i, const SysDesc& d) {
            if (geoms[i]->has_real && rec_layout(d.kind) == layout) blocks(i, (d.Ns + 255) / 256);
        });
    for (int kind = 0; kind < kNumKinds; ++kind) {
        for (int prec = 0; prec < 2; ++prec)
            add(B->filter[prec][kind], [&](int i, const SysDesc& d) {
                if (geoms[i]->has_k && d.kind == kind && d.prec == prec &&
                    !fft_xfilter_fits(d.kind, d.Mx, d.free_R))
                    blocks(i, ((long)d.Mx * (d.My / 2 + 1) + 255) / 256);
            });
        for (int tier = 0; tier < 3; ++tier)
            for (int mode = 0; mode < kRealModes; ++mode)
                add(B->real[kind][tier][mode], [&](int i, const SysDesc& d) {
                    const int m = d.sym ? RM_SYM : d.bs == 1 ? RM_ONE_CELL : RM_GENERAL;
                    if (geoms[i]->has_real && d.kind == kind && d.e1tier == tier && m == mode &&
                        d.Nt > 0)
                        real_jobs(i, d, *geoms[i]);
                });
    }

    add(B->zm_part, [&](int i, const SysDesc& d) {
        if (geoms[i]->has_real && has_zero_mode(d.kind, d.free_R)) blocks(i, d.zm_nblk);
    });
    add(B->zm_sys, [&](int i, const SysDesc& d) {
        if (geoms[i]->has_real && has_zero_mode(d.kind, d.free_R)) blocks(i, 1);
    });
    for (int prec = 0; prec < 2; ++prec) {
        for (int c = 0; c < 2; ++c) {
            auto want = [&](int i, const SysDesc& d) {
                return geoms[i]->has_k && d.prec == prec && (c == 0 || dlp_family(d.kind));
            };
            // spreading: list c = 0 the one-pair (Stokeslet) systems, c = 1 the
            // two-pair (stresslet) systems, both pairs in one pass
            for (int q = 0; q < kNumP; ++q)
                add(B->spread[prec][c][q], [&](int i, const SysDesc& d) {
                    if (geoms[i]->has_k && d.prec == prec && dlp_family(d.kind) == (c == 1) &&
                        d.P == 2 * (q + 1))
                        blocks(i, spread_ring(d.P) ? (long)geoms[i]->sp_nseg
                                                   : (long)d.ntx * d.nty);
                });
            for (int cls = 0; cls < kFftClasses; ++cls) {
                add(B->ffty[prec][c][cls], [&](int i, const SysDesc& d) {
                    if (want(i, d) && fft_line_class(d.My) == cls)
                        blocks(i, (d.Mx + d.By - 1) / d.By);
                });
                add(B->fftx[prec][c][cls], [&](int i, const SysDesc& d) {
                    if (want(i, d) && fft_line_class(d.Mx) == cls &&
                        !(c == 0 && fft_xfilter_fits(d.kind, d.Mx, d.free_R)))
                        xcd_blocks(i, (d.My + d.Bx - 1) / d.Bx);
                });
            }
        }
        for (int g = 0; g < 2; ++g) {
            auto want = [&](int i, const SysDesc& d) {
                return geoms[i]->has_k && d.prec == prec && (g == 0 || out_grids(d.kind) == 2);
            };
            for (int cls = 0; cls < kFftClasses; ++cls) {
                add(B->ifftx[prec][g][cls], [&](int i, const SysDesc& d) {
                    if (want(i, d) && fft_line_class(d.Mx) == cls &&
                        !(g == 0 && fft_xfilter_fits(d.kind, d.Mx, d.free_R)))
                        blocks(i, (d.My + d.Bx - 1) / d.Bx);
                });
                add(B->iffty[prec][g][cls], [&](int i, const SysDesc& d) {
                    if (want(i, d) && fft_line_class(d.My) == cls &&
                        !(g == 0 && fft_xfilter_fits(d.kind, d.Mx, d.free_R)))
                        xcd_blocks(i, (d.Mx + d.By - 1) / d.By);
                });
                if (g == 0)
                    add(B->iffty_xf[prec][cls], [&](int i, const SysDesc& d) {
                        if (want(i, d) && fft_line_class(d.My) == cls &&
                            fft_xfilter_fits(d.kind, d.Mx, d.free_R))
                            xcd_blocks(i, (d.Mx + d.By - 1) / d.By);
                    });
            }
            if (g == 0)
                add(B->xfilt[prec], [&](int i, const SysDesc& d) {
                    if (want(i, d) && fft_xfilter_fits(d.kind, d.Mx, d.free_R)) {
                        xcd_blocks(i, d.My / 2);  // row pairs; rows 0 and My/2 share job 0
                        B->lds_xf[prec] = std::max(B->lds_xf[prec],
                                                   (prec ? 8 : 16) * 2 * (size_t)d.Mx);
                    }
                });
            for (int q = 0; q < kNumP; ++q)
                add(B->gather[prec][g][q], [&](int i, const SysDesc& d) {
                    if (want(i, d) && d.Nt > 0 && d.P == 2 * (q + 1)) gather_jobs(i, *geoms[i]);
                });
        }
    }
    if (too_many) return set_error(SE2P_ERR_UNSUPPORTED, "too many blocks for the job lists");
    if (B->has_k) SE2P_TRY(fft_prepare());
    void *psys, *pjobs;
    SE2P_TRY(ws_get(ctx, ws, WS_SYS, sizeof(SysDesc) * (size_t)std::max(1, nsys), &psys));
    SE2P_TRY(ws_get(ctx, ws, WS_JOBS, sizeof(Job) * std::max<size_t>(1, J.size()), &pjobs));
    B->sys = (SysDesc*)psys;
    B->jobs = (Job*)pjobs;
    hipStream_t s = ctx->stream;
    if (nsys)
        SE2P_HIP_CHECK(hipMemcpyAsync(B->sys, hs.data(), sizeof(SysDesc) * (size_t)nsys,
                                      hipMemcpyHostToDevice, s));
    if (!J.empty())
        SE2P_HIP_CHECK(hipMemcpyAsync(B->jobs, J.data(), sizeof(Job) * J.size(),
                                      hipMemcpyHostToDevice, s));
    // the host vectors die on return: wait for the copies
    SE2P_HIP_CHECK(hipStreamSynchronize(s));
    return SE2P_OK;
}

int batch_apply_real(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_u)
{
    hipStream_t s = ctx->stream;
    const Job* J = B.jobs;
    {
        StageTimer st(ctx, "pack");
        for (int layout = 0; layout < 2; ++layout)
            SE2P_TRY(launch_fill_dens(s, layout, B.sys, J + B.fill[layout].first,
                                      B.fill[layout].count, d_f));
        SE2P_TRY(launch_zero_mode(s, B.sys, J + B.zm_part.first, B.zm_part.count,
                                  J + B.zm_sys.first, B.zm_sys.count));
    }
    StageTimer st(ctx, "real");
    for (int kind = 0; kind < kNumKinds; ++kind)
        for (int tier = 0; tier < 3; ++tier)
            for (int mode = 0; mode < kRealModes; ++mode) {
                const JobRange& r = B.real[kind][tier][mode];
                SE2P_TRY(launch_real(s, kind, tier, mode, B.sys, J + r.first, r.count, d_u));
            }
    return SE2P_OK;
}

int batch_spread_grid(se2p_ctx* ctx, hipStream_t s, const BatchExec& B, const double* d_f,
                      int c)
{
    StageTimer st(ctx, s, "spread");
    // every pair is spread with the first: one pass per system over its points
    // for one or two pairs, the densities read from the input
    if (c != 0) return SE2P_OK;
    for (int prec = 0; prec < 2; ++prec)
        for (int pairs = 1; pairs <= 2; ++pairs)
            for (int q = 0; q < kNumP; ++q)
                SE2P_TRY(launch_spread(s, prec, 2 * (q + 1), pairs, B.sys,
                                       B.jobs + B.spread[prec][pairs - 1][q].first,
                                       B.spread[prec][pairs - 1][q].count, d_f));
    return SE2P_OK;
}

int batch_gather(se2p_ctx* ctx, const BatchExec& B, double* d_u)
{
    StageTimer st(ctx, "gather");
    for (int prec = 0; prec < 2; ++prec)
        for (int q = 0; q < kNumP; ++q)
            SE2P_TRY(launch_gather(ctx->stream, prec, B.maxR, 2 * (q + 1), 0, B.sys,
                                   B.jobs + B.gather[prec][0][q].first,
                                   B.gather[prec][0][q].count, d_u, 0, 0));
    return SE2P_OK;
}

int batch_k_grids(se2p_ctx* ctx, hipStream_t s, const BatchExec& B, const double* d_f)
{
    const Job* J = B.jobs;
    // Forward: y lines in place in the spread layout (Z[ix*My + iy]), then x
    // lines read across its columns and written as the rows of W[iy*Mx + ix],
    // where the filter runs.  Inverse: x lines in place, then y lines read across
    // the columns of that grid and written in the layout the gather reads.  The
    // transposes ride on the first stages' loads, 16 bytes at a stride of My or
    // Mx, with the lines of one XCD's workgroups sharing cache lines in its L2.
    //   SLP: Z0 -yFFT-> -(T)xFFT-> W -filter-> Z0 -xIFFT-> -(T)yIFFT-> W -> gather(W)
    //   DLP: Z0, Z1 -yFFT-> -(T)xFFT-> W, D -filter-> Z0 -xIFFT-> -(T)yIFFT-> W -> gather(W)
    // The Stokeslet with x lines of 512-2048 points runs xFFT, filter and xIFFT
    // as one launch:
    //        Z0 -yFFT-> -(T)xFFT-filter-xIFFT-> W -(T)yIFFT-> Z0 -> gather(Z0)
    // Kinds with 4 values per target also filter into Z1 and return it through D.
    // Systems of the FP32 tier run the same passes on float2 grids.
    auto fft = [&](int prec, int axis, int sel, int osel, int tin, int inv, const JobRange* r,
                   const size_t* lds) {
        for (int cls = 0; cls < kFftClasses; ++cls)
            SE2P_TRY(launch_fft_lines(s, prec, axis, sel, osel, tin, inv, cls, lds[cls], B.sys,
                                      J + r[cls].first, r[cls].count));
        return SE2P_OK;
    };
    for (int c = 0; c < (B.any_dlp ? 2 : 1); ++c) {
        SE2P_TRY(batch_spread_grid(ctx, s, B, d_f, c));
        StageTimer st(ctx, s, "fft");
        const int Zc = c ? G_Z1 : G_Z0, Tc = c ? G_D : G_W;
        for (int prec = 0; prec < 2; ++prec) {
            SE2P_TRY(fft(prec, 0, Zc, Zc, 0, 0, B.ffty[prec][c], B.lds_y[prec]));
            SE2P_TRY(fft(prec, 1, Zc, Tc, 1, 0, B.fftx[prec][c], B.lds_x[prec]));
        }
    }
    {
        StageTimer st(ctx, s, "fft");
        for (int prec = 0; prec < 2; ++prec)
            SE2P_TRY(launch_fft_xfilter(s, prec, B.lds_xf[prec], B.sys, J + B.xfilt[prec].first,
                                        B.xfilt[prec].count));
    }
    {
        StageTimer st(ctx, s, "filter");
        for (int prec = 0; prec < 2; ++prec)
            for (int kind = 0; kind < kNumKinds; ++kind)
                SE2P_TRY(launch_filter(s, prec, kind, B.sys, J + B.filter[prec][kind].first,
                                       B.filter[prec][kind].count));
    }
    for (int g = 0; g < (B.any_g1 ? 2 : 1); ++g) {
        StageTimer st(ctx, s, "fft");
        const int Zg = g ? G_Z1 : G_Z0, Tg = g ? G_D : G_W;
        for (int prec = 0; prec < 2; ++prec) {
            SE2P_TRY(fft(prec, 1, Zg, Zg, 0, 1, B.ifftx[prec][g], B.lds_x[prec]));
            SE2P_TRY(fft(prec, 0, Zg, Tg, 1, 1, B.iffty[prec][g], B.lds_y[prec]));
            if (g == 0) SE2P_TRY(fft(prec, 0, G_W, G_Z0, 1, 1, B.iffty_xf[prec], B.lds_y[prec]));
        }
    }
    return SE2P_OK;
}

int batch_k_gather(se2p_ctx* ctx, const BatchExec& B, double* d_u, bool accumulate, bool use_zm)
{
    StageTimer st(ctx, "gather");
    for (int prec = 0; prec < 2; ++prec)
        for (int g = 0; g < (B.any_g1 ? 2 : 1); ++g)
            for (int q = 0; q < kNumP; ++q)
                SE2P_TRY(launch_gather(ctx->stream, prec, B.maxR, 2 * (q + 1), g, B.sys,
                                       B.jobs + B.gather[prec][g][q].first,
                                       B.gather[prec][g][q].count, d_u, accumulate ? 1 : 0,
                                       use_zm && g == 0 ? 1 : 0));
    return SE2P_OK;
}

int batch_apply_k(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_u,
                  bool accumulate, bool use_zm)
{
    SE2P_TRY(batch_k_grids(ctx, ctx->stream, B, d_f));
    return batch_k_gather(ctx, B, d_u, accumulate, use_zm);
}

int batch_apply_parts(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_ur,
                      double* d_uk, bool accumulate, bool use_zm)
{
    if (ctx->prof && ctx->prof_serial) {
        // one stream, so that each stage's events bracket that stage alone
        SE2P_TRY(batch_apply_real(ctx, B, d_f, d_ur));
        return batch_apply_k(ctx, B, d_f, d_uk, accumulate, use_zm);
    }
    // The grid stages read only d_f and geometry and write only the grids; the
    // real-space stages write d_ur and the zero mode.  The gather, which reads
    // the grids and the zero mode and adds into d_uk, waits for both streams, so
    // every value is computed as in the one-stream order.
    hipStream_t s = ctx->stream, side = ctx->side;
    SE2P_HIP_CHECK(hipEventRecord(ctx->ev_fork, s));
    SE2P_HIP_CHECK(hipStreamWaitEvent(side, ctx->ev_fork, 0));
    SE2P_TRY(batch_k_grids(ctx, side, B, d_f));
    SE2P_TRY(batch_apply_real(ctx, B, d_f, d_ur));
    SE2P_HIP_CHECK(hipEventRecord(ctx->ev_join, side));
    SE2P_HIP_CHECK(hipStreamWaitEvent(s, ctx->ev_join, 0));
    return batch_k_gather(ctx, B, d_uk, accumulate, use_zm);
}

int batch_apply(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_u)
{
    if (B.has_real && B.has_k) return batch_apply_parts(ctx, B, d_f, d_u, d_u, true, true);
    if (B.has_real) return batch_apply_real(ctx, B, d_f, d_u);
    if (B.has_k) return batch_apply_k(ctx, B, d_f, d_u, false, false);
    return SE2P_OK;
}

}  // namespace se2p
