// k-space part of the spectral Ewald sum: FGG spreading, Stokes filters, FGG
// gather.
//
// Follows PeriodicEwald/src/mex/common/ewald_tools.cpp:120-271 (Spread,
// Gather) and mex_stokes_{slp,dlp}_kspace.cpp (filters).  The Gaussian
// exp(-2 xi^2 |x - x_node|^2 / eta) over the (P+1)^2 nodes that start at a
// point's start node factors into wx[i] wy[j]; both 1D factors come from the
// fast-Gaussian-gridding recurrence wx[i] = exp(t(px^2 + 2 w px)) e3x^i e1[i],
// e3x = exp(-2 t h px), e1[i] = exp(t h^2 (i - P/2)^2), t = -2 xi^2/eta.
//
// Spreading.  For supports of up to 16 nodes (P <= 14), each strip of kSB grid
// columns sums the points whose support reaches it as matrix products H_c =
// A^T diag(d_c) B, A[k][x] = wx_k(x) over the strip's columns, B[k][y] =
// wy_k(y) over a ring of rows that follows the start rows down the strip, on
// the FP64 matrix cores (k_spread below); a point has one record per strip
// (and periodic image) it reaches, binned by (strip, start row) in increasing
// point order.  Wider supports are summed per 16 x 16 block of grid nodes
// (k_spread_block), the points binned by tiles of 16 x 16 start nodes.
//
// Two real grids travel as one complex grid Z = H_0 + i H_1, so the SLP needs
// one forward and one inverse complex FFT; the DLP spreads (f1 n1, f2 n1) and
// (f1 n2, f2 n2) into two complex grids.
#include "se2p_internal.h"
#include "kfilter.h"

#include <climits>
#include <type_traits>

namespace se2p {

namespace {

__device__ __forceinline__ void point_density(int mode, int p,
                                              const glob<const double>* __restrict__ f,
                                              const glob<const double>* __restrict__ nrm,
                                              double& d0, double& d1)
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

// Spreading, supports of up to 16 nodes.  One 64-lane wave per strip of kSB
// grid columns [X0, X0 + kSB) and segment of rows [Y0, Y1).  Its matrix-core
// accumulators (the 16 x 16 output of v_mfma_f64_16x16x4f64, one per component)
// hold the strip's columns over a ring of R = 16 rows: grid row v lives in
// output column v mod R.  (The code is written for R = 16 NT, NT ring tiles,
// but instantiated for NT = 1 only.)  A record (a
// point seen by the strip, with start column o relative to X0 and start row vy)
// adds d wx(x) wy(v) over the rows v = vy .. vy + P; its y weights go to the
// ring slots of those rows (zero elsewhere), so every multiply-add in y falls
// on the support, and its x weights over the strip's columns (zero off the
// support).  Each v_mfma step adds four records, A = their x rows, B = d_c
// times their y rows rotated to the records' start slots.
//
// The strip's records are binned by start row, so the segment's start rows Y0 -
// P .. Y1 - 1 are one or (across the periodic wrap) a few contiguous runs.  The
// ring is the window of rows [F, F + R), F the first row not yet complete: a
// record with start row vy < F + R - P fits, so the records are multiplied in in
// order while they do; at the first that does not, rows F .. F + R - P - 1 are
// complete -- those in [Y0, Y1) are stored, all are cleared to become rows F + R
// .. -- and F advances by R - P.  Rows before Y0 only build up the ring (they
// belong to the segment above).  Records are staged in chunks of kSC: lanes 0-31
// fill their x rows in LDS and lanes 32-63 their y rows (the reference's
// recurrence, e1 table in scalar registers, one code path), and a ballot over
// the chunk's start rows finds how many fit before the next batch of rows must
// be stored.  Records are visited in a fixed order, so every node sum is
// bit-reproducible.
constexpr int kSC = 32;  // records staged per chunk
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// A weight row in LDS, R + 2 entries: the weight of support node i (0 <= i <= P)
// goes to entry o + i -- x rows: o = start column - the strip's first column,
// nodes outside [0, kSB) to the padding slot kSB -- or (o + i) mod R -- y rows:
// o = the start row's ring slot, so the row is already rotated to the ring.
// The first R entries are cleared first, so the stores need no branch.
template <int R>
__device__ __forceinline__ int row_slot(bool ring, int x)
{
    return ring ? (x >= R ? x - R : x) : (x >= 0 && x < kSB ? x : kSB);
}

// FP64: e e3^i e1[i], one factor e3 per node as in the reference's recurrence.
template <int P, int R>
__device__ __forceinline__ void fill_row(double* row, bool ring, int o, double e, double e3,
                                         const double (&e1)[P + 1])
{
#pragma unroll
    for (int x = 0; x < R; x += 2) *reinterpret_cast<double2*>(row + x) = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        row[row_slot<R>(ring, o + i)] = e * e1[i];
        e *= e3;
    }
}

// FP32: exp(t (px - h (i - P/2))^2) with th2 = t h^2 and q = px/h, each weight
// evaluated directly (the FP32 recurrence would accumulate a rounding per node).
template <int P, int R>
__device__ __forceinline__ void fill_row_f32(float* row, bool ring, int o, float q, float th2)
{
#pragma unroll
    for (int x = 0; x < R; x += 2) *reinterpret_cast<float2*>(row + x) = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const float z = q - (float)(i - P / 2);
        row[row_slot<R>(ring, o + i)] = __expf(th2 * z * z);
    }
}

// (hi, lo) += v exactly as a pair (Knuth's two-sum), so a strip's FP32 sums
// carry no rounding from one run of matrix steps to the next.
__device__ __forceinline__ void two_sum_add(float& hi, float& lo, float v)
{
#pragma clang fp contract(off)
    const float s = hi + v;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (v - bb);
    hi = s;
    lo += err;
}

// S = double: the FGG recurrence of the reference and v_mfma_f64_16x16x4f64,
// accumulated straight into the ring.  S = float (the FP32 tier): weights
// evaluated directly, v_mfma_f32_16x16x4f32 into a product local to each run of
// steps, folded into (hi, lo) pairs by two-sum; rows are stored as float2 (hi +
// lo rounded once).
template <int P, typename S, int NC>
__global__ __launch_bounds__(64) void k_spread(const SysDesc* __restrict__ sys,
                                               const Job* __restrict__ jobs,
                                               const double* __restrict__ f_base)
{
    constexpr bool F32 = std::is_same<S, float>::value;
    static_assert(spread_ring(P), "wider supports use k_spread_block");
    constexpr int NT = (P + 16) / 16;  // ring tiles: ceil((P + 1) / 16)
    constexpr int R = 16 * NT;         // ring rows
    constexpr int BR = R - P;          // rows completed per batch
    constexpr int kLen = R + 2;        // LDS row: consecutive rows start in different banks
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const glob<const int>* __restrict__ tstart = gp(d.sp_start);
    const glob<const int2>* __restrict__ snode = gp(d.sp_snode);
    // FGG factors per record as two halves: (ex, e3x), (ey, e3y)
    const glob<const double2>* __restrict__ fac = gp(reinterpret_cast<const double2*>(d.sp_fac));
    // densities straight from the input: record k's point, its (f1, f2) or,
    // for the stresslet's NC = 2 pairs, (f1 n_c, f2 n_c) as the density pass
    // forms them; pair c goes to grid Z0 (c = 0) or Z1 (c = 1)
    const glob<const int>* __restrict__ pidx = gp(d.sp_perm);
    const glob<const double>* __restrict__ fsrc = gp(f_base + d.f_off);
    const glob<const double>* __restrict__ nrm = gp(d.nrm);
    const bool dlp = dlp_family(d.kind);
    const int Mx = d.Mx, My = d.My;
    double e1[P + 1];  // the same for every lane: held in scalar registers
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const double v = gp(d.e1tab)[i];
        const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
        const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
        e1[i] = __hiloint2double(hi, lo);
    }
    __shared__ __align__(16) S W[2][kSC][kLen];  // x rows, y rows (ring order)
    __shared__ double2 DD[NC][kSC];

    const int lane = threadIdx.x;
    // records past a chunk's count are read (and multiplied by a zero x weight):
    // keep every y row and density finite from the start
    for (int i = lane; i < kSC * kLen; i += 64) (&W[1][0][0])[i] = (S)0;
    if (lane < kSC)
        for (int cc = 0; cc < NC; ++cc) DD[cc][lane] = make_double2(0.0, 0.0);
    const int4 sg = d.sp_segs[jb.y];
    const int strip = sg.x;
    const int X0 = strip * kSB;
    const int Y0 = sg.y, Y1 = sg.z;
    const bool xr = lane < kSC;  // staging: lanes 0-31 fill x rows, 32-63 y rows
    const int q = lane & (kSC - 1);
    const int row = lane >> 4, col = lane & 15;  // MFMA operand: record 4 s + row, node col
    f64x4 acc[NC][2][NT];
    float hs[NC][2][NT][4], ls[NC][2][NT][4];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                acc[cc][j][t] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int r = 0; r < 4; ++r) hs[cc][j][t][r] = ls[cc][j][t][r] = 0.f;
            }
    const float th2 = (float)(d.t * d.h * d.h);

    // this lane's output column of register r: f64 product element (x, y) in lane
    // y + 16 (x % 4), register x / 4; f32 product: lane y + 16 (x / 4), register x % 4
    size_t colo[4];  // offset of the column's first node
    bool colok[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ix = F32 ? X0 + 4 * row + r : X0 + 4 * r + row;
        colok[r] = ix < Mx;
        colo[r] = (size_t)min(ix, Mx - 1) * My;
    }
    int F = Y0 - P;  // first row of the ring window
    // rows F .. F + BR - 1 are complete: store those of the segment, clear all
    auto flush = [&]() {
        int fm = F % R;
        fm = fm < 0 ? fm + R : fm;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            int off = 16 * t + col - fm;
            off = off < 0 ? off + R : off;
            const int v = F + off;
            const bool done = off < BR, keep = done && v >= Y0 && v < Y1;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) {
                if constexpr (F32) {
                    glob<float2>* __restrict__ out =
                        gp(reinterpret_cast<float2*>(d.grid[cc ? G_Z1 : G_Z0]));
                    const float z = done ? 0.f : 1.f;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (keep && colok[r])
                            out[colo[r] + v] = make_float2(hs[cc][0][t][r] + ls[cc][0][t][r],
                                                           hs[cc][1][t][r] + ls[cc][1][t][r]);
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            hs[cc][j][t][r] *= z;
                            ls[cc][j][t][r] *= z;
                        }
                    }
                } else {
                    glob<double2>* __restrict__ out = gp(d.grid[cc ? G_Z1 : G_Z0]);
                    const double z = done ? 0.0 : 1.0;  // finite sums: x 1 keeps them exactly
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (keep && colok[r])
                            out[colo[r] + v] = make_double2(acc[cc][0][t][r], acc[cc][1][t][r]);
                    acc[cc][0][t] *= z;
                    acc[cc][1][t] *= z;
                }
            }
        }
        F += BR;
    };
    // the matrix steps of the staged records [lo, hi), two at a time (operands
    // of both read before either multiplies)
    auto steps = [&](int lo, int hi) {
        f32x4 cz[NC][2][NT];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int t = 0; t < NT; ++t) cz[cc][0][t] = cz[cc][1][t] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int st = lo >> 2; 4 * st < hi; st += 2) {
            const bool two = 4 * (st + 1) < hi;
            S a[2], wy[2][NT];
            double2 dd[2][NC];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int kk = (4 * (st + u) + row) & (kSC - 1);
                const S wx = W[0][kk][col];
                a[u] = kk >= lo && kk < hi && (u == 0 || two) ? wx : (S)0;
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) dd[u][cc] = DD[cc][kk];
#pragma unroll
                for (int t = 0; t < NT; ++t) wy[u][t] = W[1][kk][16 * t + col];
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (u == 1 && !two) break;
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc) {
                        if constexpr (F32) {
                            cz[cc][0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[u], (S)dd[u][cc].x * wy[u][t], cz[cc][0][t], 0, 0, 0);
                            cz[cc][1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[u], (S)dd[u][cc].y * wy[u][t], cz[cc][1][t], 0, 0, 0);
                        } else {
                            acc[cc][0][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                a[u], dd[u][cc].x * wy[u][t], acc[cc][0][t], 0, 0, 0);
                            acc[cc][1][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                a[u], dd[u][cc].y * wy[u][t], acc[cc][1][t], 0, 0, 0);
                        }
                    }
            }
        }
        if constexpr (F32) {
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            two_sum_add(hs[cc][j][t][r], ls[cc][j][t][r], cz[cc][j][t][r]);
        }
    };
    // The segment's start rows Y0 - P .. Y1 - 1 as runs within one period, each
    // a contiguous range of records; run() moves to the next non-empty one.
    int a = Y0 - P, kb = 0, ke = 0, shift = 0;
    auto run = [&]() -> bool {
        while (a < Y1) {
            int r0 = a % My;
            r0 = r0 < 0 ? r0 + My : r0;
            const int b = min(Y1, a + (My - r0));
            shift = a - r0;  // start row of a record = its grid row + shift
            kb = tstart[strip * My + r0];
            ke = tstart[strip * My + r0 + (b - a)];
            a = b;
            if (ke > kb) return true;
        }
        return false;
    };
    // chunks of kSC records, the next one's records loaded while the current
    // one is staged and multiplied
    int2 m = make_int2(0, 0), m_n = make_int2(0, 0);
    double2 f = make_double2(0.0, 0.0), f_n = f, dd[NC], dd_n[NC];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) dd[cc] = dd_n[cc] = f;
    auto load = [&](int p0, int end, int2& mm, double2& ff, double2 (&dv)[NC]) {
        const int k = p0 + q;
        if (k < end) {
            mm = snode[k];
            ff = fac[2 * (size_t)k + (xr ? 0 : 1)];
            if (!xr) {
                const int pt = pidx[k];
#pragma unroll
                for (int cc = 0; cc < NC; ++cc)
                    point_density(dlp ? cc + 1 : 0, pt, fsrc, nrm, dv[cc].x, dv[cc].y);
            }
        }
    };
    bool have = run();
    int p0 = kb, end = ke, sh = shift;
    if (have) load(p0, end, m, f, dd);
    while (have) {
        int p0n = p0 + kSC, endn = end, shn = sh;
        bool haven = true;
        if (p0n >= end) {
            haven = run();
            p0n = kb;
            endn = ke;
            shn = shift;
        }
        if (haven) load(p0n, endn, m_n, f_n, dd_n);
        const int n = min(kSC, end - p0);
        const int vy = m.y + sh;  // start row of this lane's record
        __syncthreads();          // the previous chunk's rows have been read
        if (q < n) {
            int sl = vy % R;
            sl = sl < 0 ? sl + R : sl;
            if constexpr (F32)
                fill_row_f32<P, R>(W[xr ? 0 : 1][q], !xr, xr ? m.x : sl, (float)f.x, th2);
            else
                fill_row<P, R>(W[xr ? 0 : 1][q], !xr, xr ? m.x : sl, f.x, f.y, e1);
            if (!xr)
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) DD[cc][q] = dd[cc];
        }
        __syncthreads();
        // records are in start-row order: those that fit the ring are a prefix
        for (int pos = 0;;) {
            const int cnt = __popcll(__ballot(lane < n && lane >= pos && vy < F + BR));
            if (cnt) steps(pos, pos + cnt);
            pos += cnt;
            if (pos >= n) break;
            flush();
        }
        have = haven;
        p0 = p0n;
        end = endn;
        sh = shn;
        m = m_n;
        f = f_n;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) dd[cc] = dd_n[cc];
    }
    while (F < Y1) flush();
}

// Spreading for supports of more than 16 nodes (P >= 16), where the ring would
// need two or three tiles: one 64-lane wave per kSB x kSB block of grid nodes,
// which it owns.  It visits every point whose support reaches the block -- the points
// of the spreading tiles (kSB x kSB start nodes, binned at build) that cover
// the block's window of start nodes [X0 - P, X0 + kSB) x [Y0 - P, Y0 + kSB),
// one periodic image at a time (several on grids narrower than the window) --
// and accumulates
//     H_c(x, y) = sum_k wx_k(x) d_c,k wy_k(y)
// on the FP64 matrix cores: each v_mfma_f64_16x16x4f64 step adds four points,
// A = their wx rows over the block's 16 x nodes, B = their d_c wy rows over its
// 16 y nodes.  The weight rows of a chunk of kSC points are staged in LDS (the
// first 32 lanes fill the x rows, the other 32 the y rows).  The block is
// written to the grid once, zeros included.  Points are visited in a fixed
// order (image, tile, binned order), so every node sum is bit-reproducible.
constexpr int kBRow = kSB + 2;  // padded LDS row: consecutive lanes' rows start in different banks

// row[x], x in [0, kSB): the weight of support node i = x - o (0 <= i <= P) of a
// point whose start node lies at block offset o, e e3^i e1[i] with one factor
// e3 per node as in the reference's recurrence; 0 off the support.  Support
// nodes outside the block go to the row's padding slot kSB, so the stores
// need no branch.
template <int P>
__device__ __forceinline__ void block_row(double* row, int o, double e, double e3,
                                         const double (&e1)[P + 1])
{
#pragma unroll
    for (int x = 0; x < kSB; x += 2) *reinterpret_cast<double2*>(row + x) = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const int x = o + i;
        row[x >= 0 && x < kSB ? x : kSB] = e * e1[i];
        e *= e3;
    }
}

// FP32 rows: the weights exp(t (px - h (i - P/2))^2) of support nodes i, with
// th2 = t h^2 and q = px/h, each evaluated directly (the FP32 recurrence would
// accumulate a rounding per node).
template <int P>
__device__ __forceinline__ void block_row_f32(float* row, int o, float q, float th2)
{
#pragma unroll
    for (int x = 0; x < kSB; x += 2) *reinterpret_cast<float2*>(row + x) = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const int x = o + i;
        const float z = q - (float)(i - P / 2);
        row[x >= 0 && x < kSB ? x : kSB] = __expf(th2 * z * z);
    }
}

// S = double: the FGG recurrence of the reference and v_mfma_f64_16x16x4f64,
// accumulated straight into the block's two 16 x 16 products.  S = float (the
// FP32 tier): weights evaluated directly, v_mfma_f32_16x16x4f32 into a
// chunk-local product, folded into (hi, lo) pairs by two-sum after every chunk;
// the block is stored as float2 (hi + lo rounded once).
template <int P, typename S, int NC>
__global__ __launch_bounds__(64) void k_spread_block(const SysDesc* __restrict__ sys,
                                                     const Job* __restrict__ jobs,
                                                     const double* __restrict__ f_base)
{
    constexpr bool F32 = std::is_same<S, float>::value;
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const glob<const int>* __restrict__ tstart = gp(d.sp_start);
    const glob<const int2>* __restrict__ snode = gp(d.sp_snode);
    const glob<const double4>* __restrict__ fac = gp(d.sp_fac);
    // NC density pairs (the stresslet's two) straight from the input: record k's
    // point, its (f1, f2) or (f1 n_c, f2 n_c); pair c into grid Z0 (c = 0), Z1
    const glob<const int>* __restrict__ pidx = gp(d.sp_perm);
    const glob<const double>* __restrict__ fsrc = gp(f_base + d.f_off);
    const glob<const double>* __restrict__ nrm = gp(d.nrm);
    const bool dlp = dlp_family(d.kind);
    const int Mx = d.Mx, My = d.My, ntx = d.ntx;
    double e1[P + 1];  // the same for every lane: held in scalar registers
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const double v = gp(d.e1tab)[i];
        const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
        const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
        e1[i] = __hiloint2double(hi, lo);
    }
    __shared__ __align__(16) S WX[kSC][kBRow];
    __shared__ __align__(16) S WY[kSC][kBRow];
    __shared__ double4 SF[kSC];
    __shared__ double2 DD[NC][kSC];
    __shared__ int OX[kSC], OY[kSC];

    const int lane = threadIdx.x;
    // rows past a chunk's count are read (and multiplied by a zero x weight) in
    // its last matrix step: keep every y row and density finite from the start
    for (int i = lane; i < kSC * kBRow; i += 64) (&WY[0][0])[i] = (S)0;
    if (lane < kSC)
        for (int cc = 0; cc < NC; ++cc) DD[cc][lane] = make_double2(0.0, 0.0);
    const int X0 = (jb.y % ntx) * kSB, Y0 = (jb.y / ntx) * kSB;
    const int XE = min(X0 + kSB, Mx), YE = min(Y0 + kSB, My);
    const int q = lane & (kSC - 1);
    const int row = lane >> 4, col = lane & 15;  // MFMA operand: point 4 s + row, node col
    f64x4 acc[NC][2];
    float hs[NC][2][4], ls[NC][2][4];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[cc][j] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < 4; ++r) hs[cc][j][r] = ls[cc][j][r] = 0.f;
        }
    const float th2 = (float)(d.t * d.h * d.h);
    // The chunks the block visits, in the order of the loops image kx, image ky,
    // tile ty, tile tx, binned order (image kx: start node m.x - kx Mx; the window
    // reaches images while X0 - P + kx Mx < Mx), generated one ahead: the next
    // chunk's records are loaded while the current one is staged and multiplied.
    // Every quantity of the sequence is the same for the whole wave.
    struct Chunk {
        int kb, ke, sx, sy, lox, hix, loy, hiy;
    };
    int gkx = 0, gky = 0, gtx = 0, gty = 0, gkb = 0, gke = 0;
    int glox = 0, ghix = -1, gloy = 0, ghiy = -1;
    auto image_x = [&]() {
        glox = max(0, X0 - P + gkx * Mx);
        ghix = min(Mx - 1, XE - 1 + gkx * Mx);
    };
    auto image_y = [&]() {
        gloy = max(0, Y0 - P + gky * My);
        ghiy = min(My - 1, YE - 1 + gky * My);
    };
    image_x();
    image_y();
    gtx = glox / kSB;
    gty = gloy / kSB;
    {
        const int tile = gty * ntx + gtx;
        gkb = tstart[tile];
        gke = tstart[tile + 1];
    }
    // the chunk at the generator's position, then the position advanced past it
    // (false when the sequence has ended)
    auto next_chunk = [&](Chunk& ch) -> bool {
        while (gkb >= gke) {  // tile exhausted: the next tile, image row or image
            if (++gtx > ghix / kSB) {
                if (++gty > ghiy / kSB) {
                    ++gky;
                    if (!(Y0 - P + gky * My < My)) {
                        gky = 0;
                        ++gkx;
                        if (!(X0 - P + gkx * Mx < Mx)) return false;
                        image_x();
                    }
                    image_y();
                    gty = gloy / kSB;
                }
                gtx = glox / kSB;
            }
            const int tile = gty * ntx + gtx;
            gkb = tstart[tile];
            gke = tstart[tile + 1];
        }
        ch = Chunk{gkb, gke, gkx * Mx, gky * My, glox, ghix, gloy, ghiy};
        gkb += kSC;
        return true;
    };
    // one round of independent, contiguous loads per chunk
    auto load = [&](const Chunk& ch, int2& m, double4& fk, double2 (&dk)[NC]) {
        const int k = ch.kb + lane;
        m = make_int2(-1, -1);
        fk = make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) dk[cc] = make_double2(0.0, 0.0);
        if (lane < kSC && k < ch.ke) {
            m = snode[k];
            fk = fac[k];
            const int pt = pidx[k];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
                point_density(dlp ? cc + 1 : 0, pt, fsrc, nrm, dk[cc].x, dk[cc].y);
        }
    };
    Chunk cur;
    bool have = next_chunk(cur);
    int2 m;
    double4 fk;
    double2 dk[NC];
    if (have) load(cur, m, fk, dk);
    while (have) {
        Chunk nxt;
        const bool have_n = next_chunk(nxt);
        int2 m_n;
        double4 fk_n;
        double2 dk_n[NC];
        if (have_n) load(nxt, m_n, fk_n, dk_n);
        // lanes past the chunk's end hold start node (-1, -1), outside every window
        const bool keep = m.x >= cur.lox && m.x <= cur.hix && m.y >= cur.loy && m.y <= cur.hiy;
        const unsigned long long bal = __ballot(keep);
        if (bal) {
            const int n = __popcll(bal);
            __syncthreads();  // the previous chunk's rows have been read
            if (keep) {
                const int pos = __builtin_amdgcn_mbcnt_hi(
                    (unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
                OX[pos] = m.x - cur.sx - X0;
                OY[pos] = m.y - cur.sy - Y0;
                SF[pos] = fk;
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) DD[cc][pos] = dk[cc];
            }
            __syncthreads();
            if (q < n) {  // lanes 0-31: x rows, 32-63: y rows, one code path
                const bool xr = lane < kSC;
                const double4 f4 = SF[q];
                if constexpr (F32)
                    block_row_f32<P>(xr ? WX[q] : WY[q], xr ? OX[q] : OY[q],
                                    (float)(xr ? f4.x : f4.z), th2);
                else
                    block_row<P>(xr ? WX[q] : WY[q], xr ? OX[q] : OY[q], xr ? f4.x : f4.z,
                                xr ? f4.y : f4.w, e1);
            }
            __syncthreads();
            // all operands of the chunk first, then the matrix steps back to back
            S a[kSC / 4], wyv[kSC / 4];
            double2 dd[NC][kSC / 4];
#pragma unroll
            for (int st = 0; st < kSC / 4; ++st) {
                const int kk = 4 * st + row;
                const S wx = WX[kk][col];
                a[st] = kk < n ? wx : (S)0;
                wyv[st] = WY[kk][col];
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) dd[cc][st] = DD[cc][kk];
            }
            if constexpr (F32) {
                f32x4 cz[NC][2];
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) cz[cc][0] = cz[cc][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int st = 0; st < kSC / 4; ++st)
                    if (4 * st < n) {
#pragma unroll
                        for (int cc = 0; cc < NC; ++cc) {
                            cz[cc][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[st], (S)dd[cc][st].x * wyv[st], cz[cc][0], 0, 0, 0);
                            cz[cc][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[st], (S)dd[cc][st].y * wyv[st], cz[cc][1], 0, 0, 0);
                        }
                    }
#pragma unroll
                for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        two_sum_add(hs[cc][0][r], ls[cc][0][r], cz[cc][0][r]);
                        two_sum_add(hs[cc][1][r], ls[cc][1][r], cz[cc][1][r]);
                    }
            } else {
#pragma unroll
                for (int st = 0; st < kSC / 4; ++st)
                    if (4 * st < n) {
#pragma unroll
                        for (int cc = 0; cc < NC; ++cc) {
                            acc[cc][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                a[st], dd[cc][st].x * wyv[st], acc[cc][0], 0, 0, 0);
                            acc[cc][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                a[st], dd[cc][st].y * wyv[st], acc[cc][1], 0, 0, 0);
                        }
                    }
            }
        }
        cur = nxt;
        m = m_n;
        fk = fk_n;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) dk[cc] = dk_n[cc];
        have = have_n;
    }
    const int iy = Y0 + col;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        if constexpr (F32) {
            // f32 product: element (x, y) in lane y + 16 (x / 4), register x % 4
            glob<float2>* __restrict__ out =
                gp(reinterpret_cast<float2*>(d.grid[cc ? G_Z1 : G_Z0]));
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ix = X0 + 4 * row + r;
                if (ix < Mx && iy < My)
                    out[(size_t)ix * My + iy] = make_float2(hs[cc][0][r] + ls[cc][0][r],
                                                            hs[cc][1][r] + ls[cc][1][r]);
            }
        } else {
            // f64 product: element (x, y) in lane y + 16 (x % 4), register x / 4
            glob<double2>* __restrict__ out = gp(d.grid[cc ? G_Z1 : G_Z0]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ix = X0 + 4 * r + row;
                if (ix < Mx && iy < My)
                    out[(size_t)ix * My + iy] = make_double2(acc[cc][0][r], acc[cc][1][r]);
            }
        }
    }
}

// Per binned record (a point of a tile for the block kernel, owner null): its
// strip- or grid-relative start node, its point and either the point's FGG factors (ex, e3x, ey, e3y) of the FP64 recurrence or, for the
// FP32 tier (direct), its offsets in grid spacings (px/h, 0, py/h, 0).
__global__ void k_spread_records(const int* __restrict__ rperm, const int2* __restrict__ rnode,
                                 const int* __restrict__ owner,
                                 const double2* __restrict__ off, int n, double t, double w,
                                 double h, int direct, int2* __restrict__ snode,
                                 double4* __restrict__ fac, int* __restrict__ pidx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int r = rperm[k];
    const int p = owner ? owner[r] : r;
    const double2 o = off[p];
    snode[k] = rnode[r];
    pidx[k] = p;
    if (direct)
        fac[k] = make_double4(o.x / h, 0.0, o.y / h, 0.0);
    else
        fac[k] = make_double4(exp(t * (o.x * o.x + 2 * w * o.x)), exp(-2 * t * h * o.x),
                              exp(t * (o.y * o.y + 2 * w * o.y)), exp(-2 * t * h * o.y));
}

// The gathering lanes' order, fixed per geometry: within every chunk of 256
// targets of a tile (the chunks the gather takes, from the tile's first target),
// the targets are reordered so that the k-th of region-offset residue r (b = ox
// gR + oy, r = b mod 16, the LDS bank group of the target's reads) goes to
// position ~16 k + r; the 8- and 16-lane groups an LDS read is served in then
// mostly see distinct residues.  Ranks by wave ballots (in index order within a
// residue), no atomics; one workgroup per tile, perm rewritten in place.
__global__ __launch_bounds__(256) void k_gather_order(int* __restrict__ perm,
                                                      const int* __restrict__ start,
                                                      const int2* __restrict__ node, int ntx,
                                                      int TX, int TY, int R)
{
    __shared__ int wcnt[4][16];
    __shared__ int order[256];
    const int tile = blockIdx.x;
    const int X0 = (tile % ntx) * TX, Y0 = (tile / ntx) * TY;
    const int t0 = start[tile], t1 = start[tile + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int kb = t0; kb < t1; kb += 256) {
        const int cnt = min(256, t1 - kb);
        int key = -1, p = 0;
        if (tid < cnt) {
            p = perm[kb + tid];
            const int2 m = node[p];
            key = ((m.x - X0) * R + (m.y - Y0)) & 15;
        }
        int rank = 0;
        for (int b = 0; b < 16; ++b) {
            const unsigned long long msk = __ballot(key == b);
            if (key == b)
                rank = __builtin_amdgcn_mbcnt_hi((unsigned)(msk >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((unsigned)msk, 0u));
            if (lane == 0) wcnt[wv][b] = __popcll(msk);
        }
        __syncthreads();
        if (key >= 0) {
            for (int q = 0; q < wv; ++q) rank += wcnt[q][key];
            // occupied slots (r, b), r < count of b, that come before (rank, key)
            // in the order 16 r + b
            int pos = 0;
            for (int b = 0; b < 16; ++b) {
                const int tot = wcnt[0][b] + wcnt[1][b] + wcnt[2][b] + wcnt[3][b];
                pos += min(tot, rank + (b < key ? 1 : 0));
            }
            order[pos] = p;
        }
        __syncthreads();
        if (tid < cnt) perm[kb + tid] = order[tid];
        __syncthreads();  // order[] and wcnt[] are rewritten by the next chunk
    }
}

__global__ void k_gather_records(const int* __restrict__ perm, const int2* __restrict__ node,
                                 const double2* __restrict__ off, const int* __restrict__ spos,
                                 int n, int2* __restrict__ bnode, double2* __restrict__ boff,
                                 int* __restrict__ sidx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = perm[k];
    bnode[k] = node[p];
    boff[k] = off[p];
    if (spos) sidx[k] = spos[p];
}

// Each mode and its Hermitian partner (-k) are filtered together from the same
// two loads: rows iy = 0 .. My/2 of the transposed layout Z[iy*Mx + ix] pair
// with rows My - iy, and in the rows that are their own partners (0, My/2) the
// modes ix <= Mx - ix take both.  Each output is the one filter_mode gives for
// its mode alone.
// S = float: the grids hold float2 values; each mode is filtered in FP64.
template <int KIND, typename S>
__global__ void k_filter(const SysDesc* __restrict__ sys, const Job* __restrict__ jobs)
{
    using V = typename std::conditional<std::is_same<S, float>::value, float2, double2>::type;
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int Mx = d.Mx, My = d.My;
    const long gid = (long)jb.y * blockDim.x + threadIdx.x;
    if (gid >= (long)(My / 2 + 1) * Mx) return;
    const int iy = (int)(gid / Mx);
    const int ix = (int)(gid - (long)iy * Mx);
    const int jx = ix ? Mx - ix : 0, jy = iy ? My - iy : 0;
    if (jy == iy && jx < ix) return;  // taken with its partner
    const long gk = gid, gm = (long)jy * Mx + jx;
    const glob<const V>* __restrict__ W = gp(reinterpret_cast<const V*>(d.grid[G_W]));
    const glob<const V>* __restrict__ D = gp(reinterpret_cast<const V*>(d.grid[G_D]));
    glob<V>* __restrict__ Z0 = gp(reinterpret_cast<V*>(d.grid[G_Z0]));
    glob<V>* __restrict__ Z1 = gp(reinterpret_cast<V*>(d.grid[G_Z1]));
    auto ld = [](const glob<const V>* a, long i) {
        const V v = a[i];
        return make_double2(v.x, v.y);
    };
    auto st = [](glob<V>* a, long i, double2 v) {
        V r;
        r.x = v.x;
        r.y = v.y;
        a[i] = r;
    };
    const double2 z = make_double2(0.0, 0.0);
    const bool two_in = dlp_family(KIND);
    const double2 wk = ld(W, gk), wm = ld(W, gm);
    const double2 dk = two_in ? ld(D, gk) : z, dm = two_in ? ld(D, gm) : z;
    double2 o0, o1;
    kf::filter_mode<KIND>(d, ix, iy, wk, wm, dk, dm, o0, o1);
    st(Z0, gk, o0);
    if (out_grids(KIND) == 2) st(Z1, gk, o1);
    if (gm == gk) return;
    kf::filter_mode<KIND>(d, jx, jy, wm, wk, dm, dk, o0, o1);
    st(Z0, gm, o0);
    if (out_grids(KIND) == 2) st(Z1, gm, o1);
}

// Gathering.  One 256-lane workgroup per tile of gTX x gTY start nodes: the
// tile's gR x gR node region of the grid is staged in LDS, then every target of
// the tile sums its (P+1)^2 nodes from there.  At node step (i, j) a lane reads
// region entry b + i gR + j with b = ox gR + oy its start offset, so two lanes
// whose b differ modulo 16 read different 16-byte LDS bank groups at every
// step.  The targets of each chunk of up to 256 are therefore stored
// interleaved by b mod 16 (k_gather_order, once per geometry): lane t takes the
// chunk's t-th target, and the 8- and 16-lane groups an LDS read is served in
// mostly see distinct residues.  Which lane takes a target
// does not change its sum, whose order is fixed: row sums over j with the y
// weights held in registers (P is a template parameter, so the row unrolls and
// its LDS reads are issued together), then the rows over i.
// S = float (the FP32 tier): the region is staged as float2, the weights are
// evaluated directly in FP32 (as in spreading), each row is summed in FP32 and
// the rows are added in FP64.
template <int P, typename S>
__global__ __launch_bounds__(256) void k_gather(const SysDesc* __restrict__ sys,
                                               const Job* __restrict__ jobs,
                                               double* __restrict__ u_base, int g,
                                               int accumulate, int use_zm)
{
    constexpr bool F32 = std::is_same<S, float>::value;
    using V = typename std::conditional<F32, float2, double2>::type;
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const glob<const int>* __restrict__ tstart = gp(d.gt_start);
    const glob<const int>* __restrict__ perm = gp(d.gt_perm);
    const glob<const int2>* __restrict__ node = gp(d.gt_node);    // binned order
    const glob<const double2>* __restrict__ off = gp(d.gt_off);   // binned order
    const double2* __restrict__ W = d.grid[g ? G_D : d.out0];
    const int Mx = d.Mx, My = d.My, TX = d.gTX, TY = d.gTY, ntx = d.gntx, R = d.gR;
    const double t = d.t, w = d.w, h = d.h, scale = d.gscale;
    const glob<const double>* __restrict__ e1tab = gp(d.e1tab);
    const glob<const double>* zm = use_zm ? gp(d.zm) : nullptr;
    const bool sym = d.sym && g == 0;
    glob<double>* u = gp(u_base + d.u_off);
    const glob<const double2>* __restrict__ sacc = gp(d.sacc);
    const glob<const int>* __restrict__ sidx = gp(d.gt_sidx);
    extern __shared__ double2 reg_raw[];
    V* const reg = reinterpret_cast<V*>(reg_raw);
    const glob<const V>* __restrict__ Wg = gp(reinterpret_cast<const V*>(W));
    __shared__ double e1s[P + 1];
    // job: tile, and its targets [part, part + 1) x kGatherPart
    const int ntiles = ntx * d.gnty;
    const int tile = jb.y % ntiles, part = jb.y / ntiles;
    const int tx = tile % ntx, ty = tile / ntx;
    const int X0 = tx * TX, Y0 = ty * TY;
    const int k0 = tstart[tile] + part * kGatherPart;
    const int k1 = min(tstart[tile + 1], k0 + kGatherPart);
    if (k0 >= k1) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i <= P; i += blockDim.x) e1s[i] = e1tab[i];
    // region rows (periodic wrap; a grid may be narrower than the region)
    for (int lx = wv; lx < R; lx += 4) {
        int gx = X0 + lx;
        while (gx >= Mx) gx -= Mx;
        const glob<const V>* __restrict__ col = Wg + (size_t)gx * My;
        for (int ly = lane; ly < R; ly += 64) {
            int gy = Y0 + ly;
            while (gy >= My) gy -= My;
            reg[lx * R + ly] = col[gy];
        }
    }
    const double z0 = zm ? zm[0] : 0.0, z1 = zm ? zm[1] : 0.0;
    __syncthreads();  // the region and e1s stores before their reads
    for (int kb = k0; kb < k1; kb += 256) {
        const int cnt = min(256, k1 - kb);
        if (tid < cnt) {
            const int kk = kb + tid;  // binned index (lane order fixed per geometry)
            const int p = perm[kk];
            const int2 m = node[kk];
            const double2 o = off[kk];
            const int ox = m.x - X0, oy = m.y - Y0;
            double s0 = 0, s1 = 0;
            const V* row = reg + ox * R + oy;
            if constexpr (F32) {
                const float th2 = (float)(t * h * h);
                const float qx = (float)(o.x / h), qy = (float)(o.y / h);
                float wy[P + 1];
#pragma unroll
                for (int j = 0; j <= P; ++j) {
                    const float z = qy - (float)(j - P / 2);
                    wy[j] = __expf(th2 * z * z);
                }
                for (int i = 0; i <= P; ++i, row += R) {
                    float r0 = 0.f, r1 = 0.f;
#pragma unroll
                    for (int j = 0; j <= P; ++j) {
                        const V v = row[j];
                        r0 = fmaf(wy[j], v.x, r0);
                        r1 = fmaf(wy[j], v.y, r1);
                    }
                    const float z = qx - (float)(i - P / 2);
                    const double gx = (double)__expf(th2 * z * z);
                    s0 = fma(gx, (double)r0, s0);
                    s1 = fma(gx, (double)r1, s1);
                }
            } else {
                double wy[P + 1];
                {
                    double ey = exp(t * (o.y * o.y + 2 * w * o.y));
                    const double e3y = exp(-2 * t * h * o.y);
#pragma unroll
                    for (int j = 0; j <= P; ++j) {
                        wy[j] = ey * e1s[j];
                        ey *= e3y;
                    }
                }
                double ex = exp(t * (o.x * o.x + 2 * w * o.x));
                const double e3x = exp(-2 * t * h * o.x);
                for (int i = 0; i <= P; ++i, row += R) {
                    double r0 = 0, r1 = 0;
#pragma unroll
                    for (int j = 0; j <= P; ++j) {
                        const V v = row[j];
                        r0 = fma(wy[j], v.x, r0);
                        r1 = fma(wy[j], v.y, r1);
                    }
                    const double gx = ex * e1s[i];
                    s0 = fma(gx, r0, s0);
                    s1 = fma(gx, r1, s1);
                    ex *= e3x;
                }
            }
            double v0 = s0 * scale + z0, v1 = s1 * scale + z1;
            if (d.nc == 1) {  // one real field: the real part of W
                u[p] = accumulate ? u[p] + v0 : v0;
            } else {
                glob<double>* up = u + (size_t)d.nc * p + 2 * g;
                if (accumulate && sym) {
                    // the pair-symmetric real-space sums, in sorted order
                    const double2 a = sacc[sidx[kk]];
                    v0 = a.x / (4 * kPi) + v0;
                    v1 = a.y / (4 * kPi) + v1;
                } else if (accumulate) {
                    v0 = up[0] + v0;
                    v1 = up[1] + v1;
                }
                up[0] = v0;
                up[1] = v1;
            }
        }
        // the workgroup's waves take the chunks together: measured faster than
        // letting them drift apart (0.103 against 0.130 ms on a stresslet gather)
        __syncthreads();
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
        } else if (d.kind == SE2P_SLP && d.free_R > 0) {  // free-space Stokeslet: f
            const double* r = rec + 4 * (size_t)i;
            ax += r[2];
            ay += r[3];
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
    } else if (d.kind == SE2P_SLP) {
        // free-space Stokeslet: G = (dd - I lap) B + I/(8 pi) with B the
        // biharmonic kernel r^2 (ln r - 1)/(8 pi), and (dd - I lap) B =
        // (dd - I lap) B~ - 2 beta I for the B~ of the filter (B less a
        // quadratic, beta = (2 ln R - 1)/(16 pi)); the k-space sum carries
        // (dd - I lap) B~, this adds the constant, (1 - ln R)/(4 pi) sum(f)
        const double c = (1.0 - log(d.free_R)) / (4 * kPi);
        d.zm[0] = sx * c;
        d.zm[1] = sy * c;
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
    const glob<const double>* __restrict__ f = gp(f_base + d.f_off);
    const int o = gp(d.s_perm)[i];
    const double f1 = f[2 * o], f2 = f[2 * o + 1];
    if (LAYOUT == 0) {
        glob<double2>* rec = gp(reinterpret_cast<double2*>(d.s_rec));
        rec[2 * i + 1] = make_double2(f1, f2);
        // pair-symmetric sums (sources = targets, one sorted order): clear the
        // accumulator the real-space pass adds into
        if (d.sym) gp(d.sacc)[i] = make_double2(0.0, 0.0);
    } else {
        glob<double>* r = gp(d.s_rec + 8 * (size_t)i);
        r[2] = f1;
        r[3] = f2;
        r[6] = f1 * r[4] + f2 * r[5];
        if (d.sym) gp(d.sacc)[i] = make_double2(0.0, 0.0);
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
    g->ntx = (Mx + kSB - 1) / kSB;
    g->ring = spread_ring(P);
    if (g->ring) {
        // one wave per strip of kSB grid columns and segment of sseg rows (records
        // binned by strip and start row); segments as long as a launch of 4096
        // waves allows, at least 32 rows (each starts P rows early to build up
        // its ring)
        g->TX = kSB;
        g->TY = 1;
        g->sseg = 128;
        while (g->sseg > 32 && (long)g->ntx * ((My + g->sseg - 1) / g->sseg) < 4096)
            g->sseg /= 2;
        g->nty = (My + g->sseg - 1) / g->sseg;
    } else {
        // one wave per block of kSB x kSB grid nodes, points binned by tiles of
        // kSB x kSB start nodes
        g->TX = g->TY = kSB;
        g->sseg = 0;
        g->nty = (My + kSB - 1) / kSB;
    }
    // gathering tiles: 32 x 32 start nodes (region up to 56 x 56 for P <= 24) for
    // 256 lanes per tile
    g->gTX = g->gTY = P <= 24 ? 32 : 16;
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

template <typename S, int NC>
static int spread_launch(hipStream_t s, int P, const SysDesc* sys, const Job* jobs, int njobs,
                         const double* f)
{
    switch (P) {
#define SE2P_SPREAD_P(PP)                                                                         \
    case PP:                                                                                      \
        if constexpr (spread_ring(PP))                                                            \
            hipLaunchKernelGGL((k_spread<PP, S, NC>), dim3(njobs), dim3(64), 0, s, sys, jobs, f); \
        else                                                                                      \
            hipLaunchKernelGGL((k_spread_block<PP, S, NC>), dim3(njobs), dim3(64), 0, s, sys,     \
                               jobs, f);                                                          \
        break;
        SE2P_SPREAD_P(2) SE2P_SPREAD_P(4) SE2P_SPREAD_P(6) SE2P_SPREAD_P(8) SE2P_SPREAD_P(10)
        SE2P_SPREAD_P(12) SE2P_SPREAD_P(14) SE2P_SPREAD_P(16) SE2P_SPREAD_P(18) SE2P_SPREAD_P(20)
        SE2P_SPREAD_P(22) SE2P_SPREAD_P(24) SE2P_SPREAD_P(26) SE2P_SPREAD_P(28) SE2P_SPREAD_P(30)
        SE2P_SPREAD_P(32) SE2P_SPREAD_P(34) SE2P_SPREAD_P(36) SE2P_SPREAD_P(38) SE2P_SPREAD_P(40)
#undef SE2P_SPREAD_P
    default: return set_error(SE2P_ERR_UNSUPPORTED, "spread: P = %d", P);
    }
    return SE2P_OK;
}

int launch_spread(hipStream_t s, int prec, int P, int npairs, const SysDesc* sys,
                  const Job* jobs, int njobs, const double* f)
{
    if (njobs <= 0) return SE2P_OK;
    const int rc = npairs == 2 ? (prec ? spread_launch<float, 2>(s, P, sys, jobs, njobs, f)
                                       : spread_launch<double, 2>(s, P, sys, jobs, njobs, f))
                               : (prec ? spread_launch<float, 1>(s, P, sys, jobs, njobs, f)
                                       : spread_launch<double, 1>(s, P, sys, jobs, njobs, f));
    SE2P_TRY(rc);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_spread_records(hipStream_t s, const KGeom& g, int prec, const int* rperm,
                          const int2* rnode, const int* owner, const double2* off, int n,
                          int2* snode, double4* fac, int* pidx)
{
    if (n <= 0) return SE2P_OK;
    const double t = -2 * g.xi * g.xi / g.eta;
    hipLaunchKernelGGL(k_spread_records, dim3((n + 255) / 256), dim3(256), 0, s, rperm, rnode,
                       owner, off, n, t, g.w, g.h, prec, snode, fac, pidx);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_gather_order(hipStream_t s, int* perm, const int* start, const int2* node, int ntiles,
                        int ntx, int TX, int TY, int R)
{
    if (ntiles <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_gather_order, dim3(ntiles), dim3(256), 0, s, perm, start, node, ntx, TX,
                       TY, R);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_gather_records(hipStream_t s, const int* perm, const int2* node, const double2* off,
                          const int* spos, int n, int2* bnode, double2* boff, int* sidx)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_gather_records, dim3((n + 255) / 256), dim3(256), 0, s, perm, node, off,
                       spos, n, bnode, boff, sidx);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

template <typename S>
static void filter_launch(hipStream_t s, int kind, const SysDesc* sys, const Job* jobs, int njobs)
{
#define SE2P_LAUNCH_FILTER(K) \
    hipLaunchKernelGGL((k_filter<K, S>), dim3(njobs), dim3(256), 0, s, sys, jobs)
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
}

int launch_filter(hipStream_t s, int prec, int kind, const SysDesc* sys, const Job* jobs,
                  int njobs)
{
    if (njobs <= 0) return SE2P_OK;
    if (prec)
        filter_launch<float>(s, kind, sys, jobs, njobs);
    else
        filter_launch<double>(s, kind, sys, jobs, njobs);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

template <typename S>
static int gather_launch(hipStream_t s, size_t lds, int P, int g, const SysDesc* sys,
                         const Job* jobs, int njobs, double* u, int accumulate, int use_zm)
{
    switch (P) {
#define SE2P_GATHER_P(PP)                                                                         \
    case PP:                                                                                      \
        hipLaunchKernelGGL((k_gather<PP, S>), dim3(njobs), dim3(256), lds, s, sys, jobs, u, g,   \
                           accumulate, use_zm);                                                   \
        break;
        SE2P_GATHER_P(2) SE2P_GATHER_P(4) SE2P_GATHER_P(6) SE2P_GATHER_P(8) SE2P_GATHER_P(10)
        SE2P_GATHER_P(12) SE2P_GATHER_P(14) SE2P_GATHER_P(16) SE2P_GATHER_P(18) SE2P_GATHER_P(20)
        SE2P_GATHER_P(22) SE2P_GATHER_P(24) SE2P_GATHER_P(26) SE2P_GATHER_P(28) SE2P_GATHER_P(30)
        SE2P_GATHER_P(32) SE2P_GATHER_P(34) SE2P_GATHER_P(36) SE2P_GATHER_P(38) SE2P_GATHER_P(40)
#undef SE2P_GATHER_P
    default: return set_error(SE2P_ERR_UNSUPPORTED, "gather: P = %d", P);
    }
    static_assert(kMaxP == 40, "gather instantiations cover every even P up to kMaxP");
    return SE2P_OK;
}

int launch_gather(hipStream_t s, int prec, int maxR, int P, int g, const SysDesc* sys,
                  const Job* jobs, int njobs, double* u, int accumulate, int use_zm)
{
    if (njobs <= 0) return SE2P_OK;
    const size_t lds = (prec ? sizeof(float2) : sizeof(double2)) * (size_t)maxR * maxR;
    SE2P_TRY(prec ? gather_launch<float>(s, lds, P, g, sys, jobs, njobs, u, accumulate, use_zm)
                  : gather_launch<double>(s, lds, P, g, sys, jobs, njobs, u, accumulate, use_zm));
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
