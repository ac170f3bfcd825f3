// Internal declarations shared by the HIP translation units of libse2p.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "se2p.h"

namespace se2p {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Status helpers
// ---------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...);

#define SE2P_HIP_CHECK(expr)                                                         \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess)                                                        \
            return ::se2p::set_error(_e == hipErrorOutOfMemory ? SE2P_ERR_NOMEM      \
                                                               : SE2P_ERR_HIP,       \
                                     "%s failed: %s (%s:%d)", #expr,                 \
                                     hipGetErrorString(_e), __FILE__, __LINE__);     \
    } while (0)

#define SE2P_TRY(expr)                 \
    do {                               \
        int _rc = (expr);              \
        if (_rc != SE2P_OK) return _rc; \
    } while (0)

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct ProfRecord {
    std::string name;
    hipEvent_t start, stop;
};

struct ProfAcc {
    std::string name;
    double ms = 0;
    int count = 0;
};

// Grow-only device buffers addressed by slot.
enum Slot {
    WS_SRC_KEY, WS_SRC_PERM, WS_SRC_START, WS_SRC_COUNT,
    WS_TAR_KEY, WS_TAR_PERM, WS_TAR_START, WS_TAR_COUNT,
    WS_SRC_SORTED, WS_TAR_SORTED,
    WS_SPR_KEY, WS_SPR_PERM, WS_SPR_START, WS_SPR_COUNT, WS_SPR_NODE, WS_SPR_OFF,
    WS_GAT_KEY, WS_GAT_PERM, WS_GAT_START, WS_GAT_COUNT, WS_GAT_NODE, WS_GAT_OFF,
    WS_PARTIAL, WS_GRID_A, WS_GRID_B, WS_GRID_C, WS_TWIDDLE_X, WS_TWIDDLE_Y,
    WS_E1TAB, WS_BIGLIST, WS_SORT_TMP, WS_REDUCE, WS_PTS_A, WS_PTS_B, WS_DENS_A,
    WS_DENS_B, WS_OUT_A, WS_NRM, WS_ZM, WS_PLAIN_U, WS_GRID_D,
    WS_NUM_SLOTS
};

// Grow-only device buffers addressed by slot (one set per context, one per plan).
struct Workspace {
    void* ptr[WS_NUM_SLOTS] = {};
    size_t size[WS_NUM_SLOTS] = {};
};

}  // namespace se2p

struct se2p_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    se2p::Workspace ws;
    bool prof = false;
    bool capturing = false;
    std::vector<se2p::ProfRecord> prof_pending;
    std::vector<se2p::ProfAcc> prof_acc;
    std::vector<hipEvent_t> event_pool;
    std::vector<void*> user_allocs;
};

namespace se2p {

// Device buffer of at least `bytes` for a slot; contents are not preserved on growth.
int ws_get(se2p_ctx* ctx, Workspace* ws, Slot slot, size_t bytes, void** out);
void ws_release(Workspace* ws);

// Stage bracketing for per-stage event timing.
int prof_begin(se2p_ctx* ctx, const char* name, ProfRecord* rec);
int prof_end(se2p_ctx* ctx, ProfRecord* rec);
int prof_flush(se2p_ctx* ctx);

struct StageTimer {
    se2p_ctx* ctx;
    ProfRecord rec;
    bool on;
    StageTimer(se2p_ctx* c, const char* name) : ctx(c), on(c->prof && !c->capturing) {
        if (on) on = (prof_begin(ctx, name, &rec) == SE2P_OK);
    }
    ~StageTimer() {
        if (on) prof_end(ctx, &rec);
    }
};

// ---------------------------------------------------------------------------
// Binning (binning.hip)
// ---------------------------------------------------------------------------
// Cell keys of the real-space cell list, bit-identical to the reference's
// floor(nside*(x/L + 0.5)) with clamping.
int launch_cell_keys(hipStream_t s, const double* p, int n, double Lx, double Ly, int nsx,
                     int nsy, int* key, int* count);
// Grid-tile keys of the spreading / gathering tiles; also stores each point's
// wrapped start node (mx, my) and in-cell offset (px, py).
int launch_tile_keys(hipStream_t s, const double* p, int n, double Lx, double Ly, double h,
                     int P, int Mx, int My, int TX, int TY, int ntx, int* key, int* count,
                     int2* node, double2* off);
// Stable counting sort of key[0..n) over nkeys buckets: perm lists the original
// indices bucket by bucket in increasing index order; start has nkeys + 1 entries.
int sort_by_key(se2p_ctx* ctx, Workspace* ws, const int* key, int* count, int n, int nkeys,
                int* perm, int* start, Slot tmp_slot);
int launch_fill_int(hipStream_t s, int* p, int n, int v);
// Morton order inside each cell of a cell-sorted permutation (real-space clusters).
int launch_morton_cells(hipStream_t s, int* perm, const int* start, int nbox, const double* p,
                        double Lx, double Ly, int nsx, int nsy);
// FindClosestNode for n points, unwrapped start nodes.
int launch_closest_nodes(hipStream_t s, const double* p, int n, double Lx, double Ly,
                         double h, int P, int* mx, int* my, double* px, double* py);

// ---------------------------------------------------------------------------
// Real space (realspace.hip)
// ---------------------------------------------------------------------------
// Sorted source records: SLP {x, y, f1, f2}; DLP {x, y, f1, f2, n1, n2, f.n, 0}.
int launch_pack_sources(hipStream_t s, int kind, const int* perm, int n, const double* p,
                        const double* f, const double* nrm, double* out);
int launch_pack_targets(hipStream_t s, const int* perm, int n, const double* p, double2* out);
int launch_real(hipStream_t s, int kind, int tier, int nsx, int nsy, double Lx, double Ly,
                double xi, const int* tstart, const int* tperm, const double2* tsorted,
                const int* sstart, const double* ssorted, int mean_targets, double* u);
// E1 on n values (tier 0 full, 1 reduced).
int launch_e1(hipStream_t s, const double* x, int n, int tier, double* out);

// ---------------------------------------------------------------------------
// k space (kspace.hip)
// ---------------------------------------------------------------------------
struct KGeom {
    int Mx, My, P;
    double Lx, Ly, h, xi, eta, w;
    int TX, TY, ntx, nty, R;  // spread/gather tiles (TX x TY start nodes), region R x R
};
int kgeom_init(KGeom* g, int Mx, int My, int P, double Lx, double Ly, double xi, double eta,
               double w);
int launch_e1_table(hipStream_t s, const KGeom& g, double* tab);
int launch_spread(hipStream_t s, int kind, const KGeom& g, const int* tstart, const int* perm,
                  const int2* node, const double2* off, const double* f, const double* nrm,
                  const double* e1tab, double* partial);
int launch_merge(hipStream_t s, int ncomp, const KGeom& g, const double* partial,
                 double2* Z0, double2* Z1);
// rowmajor = 0: grids are My x Mx column-major (Z[ix*My + iy]); 1: transposed
// (Z[iy*Mx + ix]).
int launch_filter(hipStream_t s, int kind, const KGeom& g, const double2* Z0,
                  const double2* Z1, double2* W, int rowmajor = 0);
int launch_gather(hipStream_t s, const KGeom& g, const int* tstart, const int* perm,
                  const int2* node, const double2* off, const double2* W, const double* e1tab,
                  const double* add_first, const double* zero_mode, double* u);
int launch_dlp_zero_mode(hipStream_t s, const double* p, const double* f, const double* nrm,
                         int n, double Lx, double Ly, double* partial, double* out);
// The same sum from sorted DLP source records {x, y, f1, f2, n1, n2, f.n, 0}.
int launch_dlp_zero_mode_rec(hipStream_t s, const double* rec, int n, double Lx, double Ly,
                             double* partial, double* out);

// ---------------------------------------------------------------------------
// FFT (fft.hip)
// ---------------------------------------------------------------------------
constexpr int kFftMaxLen = 4096;
constexpr int kFftMaxStages = 24;
struct FftAxis {
    int n;
    int nst;
    int radix[kFftMaxStages];
};
int fft_axis_init(FftAxis* a, int n);
int fft_twiddles(se2p_ctx* ctx, Workspace* ws, int n, Slot slot, double2** tw);
// In-place 2D transform of a My x Mx column-major complex grid (H[ix*My + iy]).
// Forward is unnormalised; inverse is unnormalised too (the 1/(Mx My) factor is
// applied by the k-space filter).
// count contiguous lines of length n (line stride n), in place.
int fft_lines_contig(se2p_ctx* ctx, double2* Z, int n, int count, const FftAxis& ax,
                     const double2* tw, int inverse);
// out[c * rows + r] = in[r * cols + c]
int transpose(se2p_ctx* ctx, const double2* in, double2* out, int rows, int cols);
int fft2d(se2p_ctx* ctx, double2* Z, int Mx, int My, const FftAxis& ax, const FftAxis& ay,
          const double2* twx, const double2* twy, int inverse);

// ---------------------------------------------------------------------------
// Pipeline (pipeline.hip): geometry set-up once, density-dependent work per apply
// ---------------------------------------------------------------------------
struct Geom {
    int kind = 0, Ns = 0, Nt = 0;
    double Lx = 1, Ly = 1;
    se2p_params prm{};
    bool has_real = false, has_k = false;
    int e1tier = 0;
    double* nrm = nullptr;  // DLP normals, input order
    // real space
    int nbox = 0;
    int *s_perm = nullptr, *s_start = nullptr, *t_perm = nullptr, *t_start = nullptr;
    double2* t_sorted = nullptr;
    double* s_rec = nullptr;
    double *zm = nullptr, *zm_part = nullptr;
    // k space
    KGeom kg{};
    FftAxis ax{}, ay{};
    double2 *twx = nullptr, *twy = nullptr;
    double* e1tab = nullptr;
    int *sp_perm = nullptr, *sp_start = nullptr;
    int2* sp_node = nullptr;
    double2* sp_off = nullptr;
    int *gt_perm = nullptr, *gt_start = nullptr;
    int2* gt_node = nullptr;
    double2* gt_off = nullptr;
    double* partial = nullptr;
    double2 *Z0 = nullptr, *Z1 = nullptr, *W = nullptr, *D = nullptr;
};
int geom_build(se2p_ctx* ctx, Workspace* ws, Geom* G, int kind, const double* d_psrc,
               const double* d_ptar, const double* d_n, int Ns, int Nt, double Lx, double Ly,
               const se2p_params& prm, bool want_real, bool want_k);
// u = real-space part (set)
int geom_apply_real(se2p_ctx* ctx, const Geom& G, const double* d_f, double* d_u);
// u = add_first + k-space part (+ zero mode when d_zm is given)
int geom_apply_k(se2p_ctx* ctx, const Geom& G, const double* d_f, const double* d_add_first,
                 const double* d_zm, double* d_u);
// u = real + k (+ DLP zero mode)
int geom_apply(se2p_ctx* ctx, const Geom& G, const double* d_f, double* d_u);

}  // namespace se2p
