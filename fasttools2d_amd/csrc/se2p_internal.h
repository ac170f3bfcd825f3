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

// Kinds SE2P_SLP .. SE2P_DLP_STRESS.
constexpr int kNumKinds = SE2P_NUM_KINDS;
// Gaussian support: P even in [2, kMaxP] (P + 1 grid nodes per direction)
constexpr int kMaxP = 40;
constexpr int kNumP = kMaxP / 2;
// spreading: strips of kSB grid columns (the 16 x 16 output of an FP64
// matrix-core step is 16 columns by 16 ring rows) for supports of up to 16
// nodes, blocks of kSB x kSB grid nodes for wider ones
constexpr int kSB = 16;
__host__ __device__ constexpr bool spread_ring(int P) { return P + 1 <= kSB; }

// Spreading records.  A point with start column sx reaches strip s (grid
// columns [kSB s, kSB s + kSB)) through periodic image kx (virtual start column
// vx = sx - kx Mx) when its support [vx, vx + P] meets the strip; it has one
// record per such (image, strip), visited image by image (kx = 0, 1, ... while
// vx + P >= 0), strips in increasing order: f(strip, vx - kSB strip).
template <class F>
__host__ __device__ inline void for_each_strip(int sx, int Mx, int P, int ntx, F f)
{
    for (int kx = 0; sx - kx * Mx + P >= 0; ++kx) {
        const int vx = sx - kx * Mx;
        const int lo = vx > kSB - 1 ? (vx - (kSB - 1) + kSB - 1) / kSB : 0;
        const int hi0 = (vx + P) / kSB, hi = hi0 < ntx - 1 ? hi0 : ntx - 1;
        for (int s = lo; s <= hi; ++s) f(s, vx - kSB * s);
    }
}
// Stresslet kinds: normals, two spreading passes (f n1, f n2).
__host__ __device__ constexpr bool dlp_family(int kind)
{
    return kind == SE2P_DLP || kind == SE2P_DLP_PRESSURE || kind >= SE2P_DLP_GRADIENT;
}
// Values per target: 2 (velocity, pressure gradient), 1 (pressure,
// vorticity) or 4 (velocity gradient, stress).
__host__ __device__ constexpr int out_comps(int kind)
{
    return (kind == SE2P_SLP_PRESSURE || kind == SE2P_DLP_PRESSURE || kind == SE2P_SLP_VORTICITY ||
            kind == SE2P_DLP_VORTICITY)
               ? 1
           : (kind == SE2P_SLP_GRADIENT || kind == SE2P_SLP_STRESS || kind == SE2P_DLP_GRADIENT ||
              kind == SE2P_DLP_STRESS)
               ? 4
               : 2;
}
// Kinds whose full sum carries a zero mode that is the same for every target.
__host__ __device__ constexpr bool has_zero_mode(int kind)
{
    return kind == SE2P_DLP || kind == SE2P_SLP_PRESSURE || kind == SE2P_DLP_PRESSURE ||
           kind == SE2P_SLP_VORTICITY;
}
// Systems with a zero-mode term: the kinds above, and free-space Stokeslet sums
// (the constant sum(f)/(8 pi)).
__host__ __device__ constexpr bool has_zero_mode(int kind, double free_R)
{
    return has_zero_mode(kind) || (kind == SE2P_SLP && free_R > 0);
}
// Sorted source records: stresslet {x, y, f1, f2, n1, n2, f.n, 0}, else {x, y, f1, f2}.
__host__ __device__ constexpr int rec_layout(int kind) { return dlp_family(kind) ? 1 : 0; }
// Complex grids of the k-space output (two real values per grid).
__host__ __device__ constexpr int out_grids(int kind) { return out_comps(kind) == 4 ? 2 : 1; }

// Global memory seen as such.  Pointers read from a descriptor are generic, so
// loads, stores and atomics through them compile to flat_* instructions, which
// also count against the LDS wait counter: every wait for an LDS result then
// waits for the global loads in flight too.  gp(p) is the same address as a
// global-address-space pointer, whose accesses compile to global_* instructions.
// (The host pass of a HIP compilation parses kernel bodies without generating
// them and cannot assign HIP vector types across address spaces, so there the
// alias is the plain type.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using glob = __attribute__((address_space(1))) T;
#else
template <class T>
using glob = T;
#endif
template <class T>
__device__ __forceinline__ glob<T>* gp(T* p)
{
    return (glob<T>*)p;
}

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
    WS_SPR_SNODE, WS_SPR_FAC, WS_UNSORTED, WS_SRC_YKEY, WS_SYM_ACC, WS_SYM_POS,
    WS_GAT_KEY, WS_GAT_PERM, WS_GAT_START, WS_GAT_COUNT, WS_GAT_NODE, WS_GAT_OFF,
    WS_GRID_A, WS_GRID_B, WS_GRID_C, WS_TWIDDLE_X, WS_TWIDDLE_Y,
    WS_E1TAB, WS_BIGLIST, WS_SORT_TMP, WS_REDUCE, WS_PTS_A, WS_PTS_B, WS_DENS_A,
    WS_DENS_B, WS_OUT_A, WS_NRM, WS_ZM, WS_PLAIN_U, WS_GRID_D, WS_SYS, WS_JOBS, WS_FREE_PART,
    WS_MLP_PART, WS_GAT_BNODE, WS_GAT_BOFF, WS_GAT_SIDX, WS_YSORT, WS_BIGCELLS, WS_SPR_CNT, WS_SPR_OFFS, WS_SPR_RNODE,
    WS_SPR_OWNER, WS_SPR_RPERM, WS_SCAN, WS_SPR_SEGS, WS_NUM_SLOTS
};

// Grow-only device buffers addressed by slot (one set per context, one per plan).
// With `arena` set, slots are carved from [arena, arena + arena_cap) instead of
// allocated: a request larger than its slot's region takes a new region past
// the last one (the old one is abandoned), and a request past arena_cap fails
// with SE2P_ERR_UNSUPPORTED.  Nothing is ever freed.
struct Workspace {
    void* ptr[WS_NUM_SLOTS] = {};
    size_t size[WS_NUM_SLOTS] = {};
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
};

}  // namespace se2p

struct se2p_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream for the k-space grid passes, forked from and joined back
    // into `stream` with these two events inside every apply
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    se2p::Workspace ws;
    bool prof = false;
    bool prof_serial = false;  // while profiling: every stage on `stream`, in order
    bool capturing = false;
    std::vector<se2p::ProfRecord> prof_pending;
    std::vector<se2p::ProfAcc> prof_acc;
    std::vector<hipEvent_t> event_pool;
    std::vector<void*> user_allocs;
    // device address of this GPU's scratch array (allocation-free evaluation)
    char* scratch = nullptr;
    // precision of the one-shot entry points (mex equivalents and drivers)
    int precision = SE2P_FP64;
    // free-space Stokeslet method (SE2P_FREE_*)
    int free_method = SE2P_FREE_AUTO;
};

namespace se2p {

// Device buffer of at least `bytes` for a slot; contents are not preserved on growth.
int ws_get(se2p_ctx* ctx, Workspace* ws, Slot slot, size_t bytes, void** out);
void ws_release(Workspace* ws);

// Stage bracketing for per-stage event timing on the stage's stream.
int prof_begin(se2p_ctx* ctx, hipStream_t s, const char* name, ProfRecord* rec);
int prof_end(se2p_ctx* ctx, hipStream_t s, ProfRecord* rec);
int prof_flush(se2p_ctx* ctx);

struct StageTimer {
    se2p_ctx* ctx;
    hipStream_t s;
    ProfRecord rec;
    bool on;
    StageTimer(se2p_ctx* c, const char* name) : StageTimer(c, c->stream, name) {}
    StageTimer(se2p_ctx* c, hipStream_t st, const char* name)
        : ctx(c), s(st), on(c->prof && !c->capturing) {
        if (on) on = (prof_begin(ctx, s, name, &rec) == SE2P_OK);
    }
    ~StageTimer() {
        if (on) prof_end(ctx, s, &rec);
    }
};

// ---------------------------------------------------------------------------
// Binning (binning.hip)
// ---------------------------------------------------------------------------
// Cell keys of the real-space cell list, bit-identical to the reference's
// floor(nside*(x/L + 0.5)) with clamping.
int launch_cell_keys(hipStream_t s, const double* p, int n, double Lx, double Ly, int nsx,
                     int nsy, int* key, int* count);
// Spreading records, first pass: each point's start node (mx, my), offset (px,
// py) and number of records (for_each_strip).
int launch_strip_count(hipStream_t s, const double* p, int n, double Lx, double Ly, double h,
                       int P, int Mx, int My, int ntx, int* cnt, int2* node, double2* off);
// Second pass: point i's records at offs[i] .. offs[i + 1] - 1 get the key
// strip My + my (count[] incremented), rnode = (vx - kSB strip, my) and owner i.
int launch_strip_keys(hipStream_t s, const int2* node, const int* offs, int n, int P, int Mx,
                      int My, int ntx, int* key, int* count, int2* rnode, int* owner);
// Exclusive scan of count[0..n) into start[0..n] (on ctx's stream).
int scan_ints(se2p_ctx* ctx, Workspace* ws, const int* count, int n, int* start);
// Grid-tile keys of the gathering tiles (TX x TY start nodes, ntx per row, row-major);
// also stores each point's wrapped start node (mx, my) and in-cell offset (px, py).
int launch_tile_keys(hipStream_t s, const double* p, int n, double Lx, double Ly, double h,
                     int P, int Mx, int My, int TX, int TY, int ntx, int* key, int* count,
                     int2* node, double2* off);
// Stable counting sort of key[0..n) over nkeys buckets: perm lists the original
// indices bucket by bucket in increasing index order; start has nkeys + 1 entries.
int sort_by_key(se2p_ctx* ctx, Workspace* ws, const int* key, int* count, int n, int nkeys,
                int* perm, int* start, Slot tmp_slot);
int launch_fill_int(hipStream_t s, int* p, int n, int v);
// Morton order (yorder 0) or y order (yorder 1; sets *unsorted where a cell cannot be
// ordered) inside each cell of a cell-sorted permutation.  With yorder and a
// non-null ykey, ykey[k] is the 20-bit in-cell y key of the k-th sorted point.
// Cells of more than kSortCellCap points keep their order here; with yorder
// they set *unsorted unless big_sorted says launch_ysort_big orders them.
constexpr int kSortCellCap = 4096;
int launch_morton_cells(hipStream_t s, int* perm, const int* start, int nbox, const double* p,
                        double Lx, double Ly, int nsx, int nsy, int yorder, int* unsorted,
                        int* ykey = nullptr, int big_sorted = 0);
// The y order of launch_morton_cells for the ncells cells listed in `cells`
// (device array), each larger than kSortCellCap; tmp holds 4 n ints (n >= the
// permutation's length).
int launch_ysort_big(hipStream_t s, int* perm, const int* start, const int* cells, int ncells,
                     const double* p, double Ly, int nsx, int nsy, int* unsorted, int* ykey,
                     int* tmp, int n);
// FindClosestNode for n points, unwrapped start nodes.
int launch_closest_nodes(hipStream_t s, const double* p, int n, double Lx, double Ly,
                         double h, int P, int* mx, int* my, double* px, double* py);

// ---------------------------------------------------------------------------
// k-space tile geometry (kspace.hip)
// ---------------------------------------------------------------------------
struct KGeom {
    int Mx, My, P;
    double Lx, Ly, h, xi, eta, w;
    // spreading: ring (strips of kSB columns x nty segments of sseg rows, records
    // binned by strip and start row) or blocks (ntx x nty blocks of kSB x kSB
    // nodes, points binned by TX x TY tiles of start nodes)
    int TX, TY, ntx, nty;
    int ring, sseg;
    int gTX, gTY, gntx, gnty, gR;  // gathering tiles, region gR x gR
};
int kgeom_init(KGeom* g, int Mx, int My, int P, double Lx, double Ly, double xi, double eta,
               double w);
int launch_e1_table(hipStream_t s, const KGeom& g, double* tab);

// ---------------------------------------------------------------------------
// FFT axes (fft.hip)
// ---------------------------------------------------------------------------
// Lines up to 4096 points use 64 KiB of LDS per workgroup; longer lines (up to
// kFftMaxLen, 160 KiB for FP64) need the device's full LDS per workgroup.
constexpr int kFftMaxLen = 10240;
constexpr int kFftMaxLenSmall = 4096;
constexpr int kFftMaxStages = 24;
struct FftAxis {
    int n;
    int nst;
    int radix[kFftMaxStages];
};
int fft_axis_init(FftAxis* a, int n);
int fft_twiddles(se2p_ctx* ctx, Workspace* ws, int n, Slot slot, double2** tw);
// One-time kernel attributes (large dynamic LDS); not during graph capture.
int fft_prepare();
// Lines of length n per FFT workgroup.
int fft_lines_per_block(int n);

// ---------------------------------------------------------------------------
// Systems and jobs.  Every density-dependent kernel reads a table of system
// descriptors and a job list: job (s, b) is block b of system s, so one launch
// covers every system of a batch (a single plan is a batch of one).
// ---------------------------------------------------------------------------
enum GridSel { G_Z0 = 0, G_Z1 = 1, G_W = 2, G_D = 3 };

struct SysDesc {
    int kind, Ns, Nt, e1tier, nc;  // nc: values per target
    int prec;  // 1: FP32 tier (float2 grids, FP32 far pairs), 0: FP64
    long f_off, u_off;  // offsets (doubles) of this system in the apply's f and u
    double Lx, Ly;
    // real space
    int nsx, nsy, bs, nbx;
    double xi2, cut2, self;
    // half-width in y of the candidates a target keeps from each stencil column
    // (sources y-ordered in their cells); unusable while *unsorted != 0
    double trim_w;
    // fast E1 tier: pairs with r^2 >= far32_r2 (xi^2 r^2 >= se2p_params.far32)
    // are evaluated in FP32; +inf when off
    double far32_r2;
    // free-space Stokeslet sum: truncation radius of the biharmonic kernel in
    // the k-space filter (0: periodic); the zero-mode slot then holds sum(f)/(8 pi)
    double free_R;
    // own-cell Stokeslet pairs with r^2 below this take the self term instead:
    // 1e-15 (the reference's periodic sums), or the least positive double for
    // free-space sums (which leave out exactly coincident pairs only)
    double self_r2;
    const int* unsorted;
    // pair-symmetric sums (sources are the targets, same sorted order): ykey
    // holds the in-cell y key of each sorted source, sacc the sums in sorted order
    int sym;
    const int* ykey;
    double2* sacc;
    const int* spos;  // sorted position of each point (the inverse of t_perm)
    const int *t_start, *t_perm, *s_start, *s_perm;
    const double2* t_sorted;
    double* s_rec;
    double *zm, *zm_part;
    int zm_nblk;
    // k space
    int Mx, My, P, TX, TY, ntx, nty, sseg, snrec, Bx, By;
    int gTX, gTY, gntx, gnty, gR;
    double t, w, h, xi, eta, gscale;
    const int *sp_start, *sp_perm, *gt_start, *gt_perm;
    const int4* sp_segs;  // strip spreading: job j's {strip, first row, end row, 0}
    // spreading records in binned order (snrec of them): sp_perm = their points,
    // sp_snode = (start column - strip's first column, start row)
    const int2 *sp_snode, *gt_node;
    const double4* sp_fac;  // spreading: FGG factors (ex, e3x, ey, e3y) per record
    const double2* gt_off;  // gathering: offsets in binned order (gt_node likewise)
    const int* gt_sidx;     // pair-symmetric sums: sorted position of each binned target
    const double *e1tab, *nrm;
    double2* grid[4];  // GridSel
    int out0;  // the grid holding the first output for the gather: G_W, or G_Z0 for
               // systems on the fused x pass (fft_xfilter_fits)
    const double2 *twx, *twy;
    FftAxis ax, ay;
};
using Job = int2;

// Job lists of one launch: a range of the job array.
struct JobRange {
    int first = 0, count = 0;
};

// Sorted source records: SLP {x, y, f1, f2}; DLP {x, y, f1, f2, n1, n2, f.n, 0}.
int launch_pack_sources(hipStream_t s, int kind, const int* perm, int n, const double* p,
                        const double* f, const double* nrm, double* out);
int launch_pack_targets(hipStream_t s, const int* perm, int n, const double* p, double2* out);
// E1 on n values (tier 0 full, 1 reduced).
int launch_e1(hipStream_t s, const double* x, int n, int tier, double* out);

// Real-space blocks of bs x bs cells for a mean cell occupancy.
int real_block_side(int nsx, int nsy, int mean_targets);
// Lanes of a real-space workgroup: a job takes a block's targets [256 part,
// 256 part + 256).
constexpr int kRealLanes = 256;
// Targets of a gathering tile per job (4 passes of 256 lanes over one staged region).
constexpr int kGatherPart = 1024;
// Real-space launch modes: general blocks, one-cell blocks, pair-symmetric blocks.
enum RealMode { RM_GENERAL = 0, RM_ONE_CELL = 1, RM_SYM = 2, kRealModes = 3 };
// k_real_stencil<kind, tier, mode> over the jobs (block + part * blocks, part
// the block's 256-target slice).  RM_GENERAL
// and RM_ONE_CELL set u; RM_SYM adds into sacc (zeroed by the density fill) with
// atomics, and the gather adds sacc to the k-space part (only systems with a
// k-space part take pair-symmetric sums).
int launch_real(hipStream_t s, int kind, int tier, int mode, const SysDesc* sys, const Job* jobs,
                int njobs, double* u);
// pos[perm[k]] = k for k < n.
int launch_inverse_perm(hipStream_t s, const int* perm, int n, int* pos);
// Spreading of npairs density pairs (Stokeslet kinds: 1, (f1, f2); stresslet
// kinds: 2, (f1 n1, f2 n1) and (f1 n2, f2 n2)), pair c into grid Z0 (c = 0) or Z1
// (c = 1), both pairs from one pass over the points; job b = segment sp_segs[b]
// of a strip of kSB grid columns (spread_ring(P)) or block (b % ntx, b / ntx) of
// kSB x kSB nodes; all jobs' systems of support P and npairs, the densities read
// from f (the apply's input) through each record's point.
int launch_spread(hipStream_t s, int prec, int P, int npairs, const SysDesc* sys,
                  const Job* jobs, int njobs, const double* f);
// Spreading records of a geometry in binned order: record k = rperm[k] of the
// key pass gets snode[k] = rnode[rperm[k]], its point pidx[k] = owner[rperm[k]]
// (rperm[k] itself when owner is null: one record per point) and that point's
// FGG factors.
int launch_spread_records(hipStream_t s, const KGeom& g, int prec, const int* rperm,
                          const int2* rnode, const int* owner, const double2* off, int n,
                          int2* snode, double4* fac, int* pidx);
// The gathering lanes' order: within every chunk of 256 targets of each of the
// ntiles tiles (TX x TY start nodes, ntx per row, region side R), perm is
// reordered in place so the targets are interleaved by their region offset
// modulo 16 (LDS bank groups).
int launch_gather_order(hipStream_t s, int* perm, const int* start, const int2* node, int ntiles,
                        int ntx, int TX, int TY, int R);
// Gathering records in binned order: bnode[k] = node[perm[k]], boff[k] =
// off[perm[k]] and, when spos is given, sidx[k] = spos[perm[k]].
int launch_gather_records(hipStream_t s, const int* perm, const int2* node, const double2* off,
                          const int* spos, int n, int2* bnode, double2* boff, int* sidx);
// Filter (job b: pair modes b*256 .. of rows 0 .. My/2, each with its partner):
// W (and D for stresslet kinds) in transposed layout Z[iy*Mx + ix] ->
// Z0 (and Z1 for kinds with 4 values per target).
int launch_filter(hipStream_t s, int prec, int kind, const SysDesc* sys, const Job* jobs,
                  int njobs);
// Gather of output grid g (W for g = 0, D for g = 1) into values 2g, 2g + 1 of
// each target (+= when accumulate; + the zero mode when use_zm); maxR: the
// largest gathering region side of the jobs' systems, all of support P.
int launch_gather(hipStream_t s, int prec, int maxR, int P, int g, const SysDesc* sys, const Job* jobs,
                  int njobs, double* u, int accumulate, int use_zm);
// Zero modes from the sorted records (DLP: sum_s (f.n)_s x_s / (Lx Ly); SLP
// pressure: -sum_s f_s.x_s / (2 Lx Ly); DLP pressure: -sum_s (f.n)_s / (2 Lx Ly);
// SLP vorticity: -sum_s (f1 y - f2 x)_s / (2 Lx Ly)): partial sums over jobs
// (s, b), b < zm_nblk, then one job per system.
int launch_zero_mode(hipStream_t s, const SysDesc* sys, const Job* part_jobs, int npart,
                     const Job* sys_jobs, int nsys);
// Free-space direct sums (SLP or DLP records, see freespace.hip) of Ns sources
// at Nt targets, split into nslice source slices (free_slices) whose partial
// sums (part: nslice x 2 Nt doubles when nslice > 1) are added in order.
int free_slices(int Ns, int Nt);
int launch_free(hipStream_t s, bool dlp, const double* src, int Ns, const double* tar, int Nt,
                bool self, int nslice, double* part, double* out);
// Density fill of the sorted source records (layout: rec_layout).
int launch_fill_dens(hipStream_t s, int layout, const SysDesc* sys, const Job* jobs, int njobs,
                     const double* f);
// Line FFTs along y (axis 0: My-lines, Mx of them) or x (axis 1: Mx-lines, My
// of them), from grid `sel` to grid `osel` (may be `sel`), which holds them
// contiguously (line l at l*n).  tin 0: so does the input; 1: the input holds
// them as the columns of a row-major grid (point k of line l at k*nlines + l).
// Every job's line length must be of class `cls` (fft_line_class); max_lds from
// fft_lds_bytes.
constexpr int kFftClasses = 4;
int fft_line_class(int n);
int launch_fft_lines(hipStream_t s, int prec, int axis, int sel, int osel, int tin, int inverse,
                     int cls, size_t max_lds, const SysDesc* sys, const Job* jobs, int njobs);
size_t fft_lds_bytes(int n, int prec);
// The Stokeslet's forward x passes, filter and inverse x passes as one launch
// (column-major Z0 rows p and My - p -> row-major W rows p and My - p; job (s, p),
// p = 0 .. My/2 - 1, job 0 taking rows 0 and My/2), for
// periodic systems where fft_xfilter_fits(kind, Mx, free_R); LDS 2 Mx complex values.
bool fft_xfilter_fits(int kind, int Mx, double free_R);
int launch_fft_xfilter(hipStream_t s, int prec, size_t max_lds, const SysDesc* sys,
                       const Job* jobs, int njobs);

// ---------------------------------------------------------------------------
// Parameter selector network (mlp.hip): packed FP32 parameters, device arrays
// ---------------------------------------------------------------------------
int mlp_forward(se2p_ctx* ctx, const float* w, const float* x, int n, float* y);
int mlp_grad(se2p_ctx* ctx, Workspace* ws, const float* w, const float* x, const float* t, int n,
             float* grad, float* loss);
int mlp_adam(se2p_ctx* ctx, float* w, const float* g, float* m, float* v, int step, float lr,
             float b1, float b2, float eps);

// ---------------------------------------------------------------------------
// Pipeline (pipeline.hip): geometry set-up once, density-dependent work per apply
// ---------------------------------------------------------------------------
struct Geom {
    int kind = 0, Ns = 0, Nt = 0;
    double Lx = 1, Ly = 1;
    se2p_params prm{};
    bool has_real = false, has_k = false;
    int e1tier = 0;
    int prec = 0;  // precision in use: SE2P_FP32C only for the velocity kinds
    double* nrm = nullptr;  // DLP normals, input order
    // real space
    int nbox = 0;
    int *s_perm = nullptr, *s_start = nullptr, *t_perm = nullptr, *t_start = nullptr;
    double2* t_sorted = nullptr;
    double* s_rec = nullptr;
    double *zm = nullptr, *zm_part = nullptr;
    int* unsorted = nullptr;  // set by the source ordering when a cell is not y-ordered
    bool sym = false;         // pair-symmetric real space (targets alias the sources)
    int* ykey = nullptr;      // in-cell y keys of the sorted sources (sym only)
    double2* sacc = nullptr;  // pair-symmetric sums in sorted order (sym only)
    int* spos = nullptr;      // sorted position of each point (sym only)
    std::vector<int> t_start_h;  // host copy of t_start: the real-space jobs split heavy blocks
    // k space
    KGeom kg{};
    FftAxis ax{}, ay{};
    double2 *twx = nullptr, *twy = nullptr;
    double* e1tab = nullptr;
    int *sp_perm = nullptr, *sp_start = nullptr;
    int sp_nrec = 0;  // spreading records
    // strip spreading: one job per segment {strip, first row, end row, 0}
    int4* sp_segs = nullptr;
    int sp_nseg = 0;
    int2* sp_node = nullptr;
    double2* sp_off = nullptr;
    int2* sp_snode = nullptr;
    double4* sp_fac = nullptr;
    int *gt_perm = nullptr, *gt_start = nullptr;
    int2* gt_node = nullptr;    // start nodes and offsets in input order
    double2* gt_off = nullptr;
    int2* gt_bnode = nullptr;   // the same in gathering-binned order
    double2* gt_boff = nullptr;
    int* gt_sidx = nullptr;     // pair-symmetric sums: sorted position, binned order
    double2 *Z0 = nullptr, *Z1 = nullptr, *W = nullptr, *D = nullptr;
    std::vector<int> gt_start_h;  // host copy of gt_start: the gathering jobs split heavy tiles
};
int geom_build(se2p_ctx* ctx, Workspace* ws, Geom* G, int kind, const double* d_psrc,
               const double* d_ptar, const double* d_n, int Ns, int Nt, double Lx, double Ly,
               const se2p_params& prm, bool want_real, bool want_k);
void sysdesc_from_geom(const Geom& G, long f_off, long u_off, SysDesc* d);

// The descriptors and job lists of a set of systems (device arrays in `ws`).
struct BatchExec {
    int nsys = 0;
    bool has_real = false, has_k = false, any_dlp = false, any_g1 = false;
    SysDesc* sys = nullptr;
    Job* jobs = nullptr;
    JobRange fill[2];     // by record layout
    JobRange zm_part, zm_sys;
    JobRange real[kNumKinds][3][kRealModes];  // [kind][tier: full, fast, FP32][RealMode]
    // k-space passes, per grid precision (0 FP64, 1 FP32: float2 grids)
    JobRange spread[2][2][kNumP];  // [prec][pairs - 1][P / 2 - 1]
    JobRange ffty[2][2][kFftClasses], fftx[2][2][kFftClasses];  // [prec][c][class]
    JobRange filter[2][kNumKinds];  // [prec][kind]
    // inverse passes of output grid g (g = 1 for kinds with 4 values per target)
    JobRange ifftx[2][2][kFftClasses], iffty[2][2][kFftClasses];  // [prec][g][class]
    JobRange iffty_xf[2][kFftClasses];  // [prec][class]: systems on the fused x pass
    JobRange gather[2][2][kNumP];  // [prec][g][P / 2 - 1]
    JobRange xfilt[2];              // [prec]: fused x passes and filter (Stokeslet)
    size_t lds_xf[2] = {};
    size_t lds_y[2][kFftClasses] = {}, lds_x[2][kFftClasses] = {};  // [prec][line class]
    int maxR = 32;   // largest gathering region side
};
// Builds descriptors (f and u offsets: systems back to back, 2 Ns and 2 Nt
// doubles each) and job lists; synchronises the stream (host staging).
int batch_build(se2p_ctx* ctx, Workspace* ws, BatchExec* B, const Geom* const* geoms, int nsys);
// u = real-space part (set)
int batch_apply_real(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_u);
// u (+)= k-space part (+ DLP zero mode when use_zm)
int batch_apply_k(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_u,
                  bool accumulate, bool use_zm);
// k-space passes on stream s up to the inverse FFTs (no gather)
int batch_k_grids(se2p_ctx* ctx, hipStream_t s, const BatchExec& B, const double* d_f);
// u (+)= gather of the inverse-transformed grids (+ DLP zero mode when use_zm)
int batch_k_gather(se2p_ctx* ctx, const BatchExec& B, double* d_u, bool accumulate, bool use_zm);
// ur = real-space part, uk (+)= k-space part: the two run concurrently on the
// context's two streams (serially while stage profiling is on)
int batch_apply_parts(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_ur,
                      double* d_uk, bool accumulate, bool use_zm);
// u = real + k (+ DLP zero mode)
int batch_apply(se2p_ctx* ctx, const BatchExec& B, const double* d_f, double* d_u);
// Spreading into Z0 (and Z1 for the stresslet's second pair): the whole work at
// c = 0, nothing at c = 1 (both pairs come from one pass).
int batch_spread_grid(se2p_ctx* ctx, hipStream_t s, const BatchExec& B, const double* d_f,
                      int c);
// Gather from W only.
int batch_gather(se2p_ctx* ctx, const BatchExec& B, double* d_u);

}  // namespace se2p
