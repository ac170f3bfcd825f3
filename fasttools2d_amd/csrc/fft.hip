// 2D complex FFT for the spectral Ewald grids.
//
// The reference transforms with MATLAB's fft2/ifft2 called back through
// mexCallMATLAB (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_kspace.cpp:80-81,
// 176-177).  This is the library's own transform: a mixed-radix Stockham FFT
// with one workgroup per line (or per group of short lines).  The first stage
// reads the lines from HBM, the last writes them back, and in between they live
// in one LDS buffer, each stage staging its butterflies through registers.  Any
// length up to kFftMaxLen works: radices 8, 4 and 2 first, then 3, 5, 7 and any
// remaining prime factor (2, 3, 4, 5, 8 as in-register butterflies, the others
// by a direct O(R) sum per output).  Stage s with radix R and Ns = product of the earlier
// radices maps output (j, k), j < n/R, k < R, to
//   out[(j / Ns) Ns R + j % Ns + k Ns] = sum_r in[j + r n/R] W_n^{r ((j % Ns) + k Ns) n/(Ns R)}.
// A pass writes its lines contiguously and reads them contiguously or across
// the columns of the other layout, which transposes the grid in the first
// stage's loads.  Forward: y lines of the column-major grid (Z[ix*My + iy]) in
// place, then x lines read across its columns into the row-major Z[iy*Mx + ix].
// Inverse: x lines in place, then y lines read across the row-major grid's
// columns into the column-major one.
#include "se2p_internal.h"
#include "fft_roots.h"
#include "kfilter.h"

#include <cmath>
#include <vector>

namespace se2p {

namespace {

// Complex values of scalar type S: double2 (FP64 grids) or float2 (FP32 grids).
template <typename S>
struct Cplx;
template <>
struct Cplx<double> {
    using T = double2;
};
template <>
struct Cplx<float> {
    using T = float2;
};
template <typename S>
using cplx = typename Cplx<S>::T;

template <typename V>
__device__ __forceinline__ V mkc(decltype(V::x) x, decltype(V::x) y)
{
    V r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename V>
__device__ __forceinline__ V cmul(V a, V b)
{
    return mkc<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { return mkc<V>(a.x + b.x, a.y + b.y); }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { return mkc<V>(a.x - b.x, a.y - b.y); }
// v * (sgn i)
template <typename V>
__device__ __forceinline__ V rot(V v, decltype(V::x) sgn) { return mkc<V>(-sgn * v.y, sgn * v.x); }

// In-register DFTs of length R with kernel exp(sgn 2 pi i r k / R), sgn = -1 forward.
template <typename V>
__device__ __forceinline__ void dft2(V* v)
{
    const V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename V>
__device__ __forceinline__ void dft4(V& a, V& b, V& c, V& d, decltype(V::x) sgn)
{
    const V t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = rot(csub(b, d), sgn);
    a = cadd(t0, t2);
    c = csub(t0, t2);
    b = cadd(t1, t3);
    d = csub(t1, t3);
}

template <typename V>
__device__ __forceinline__ void dft3(V* v, decltype(V::x) sgn)
{
    using S = decltype(V::x);
    const S h = (S)0.86602540378443864676;  // sin(2 pi/3)
    const V a = v[0], s = cadd(v[1], v[2]), d = rot(csub(v[1], v[2]), sgn * h);
    const V m = mkc<V>(a.x - (S)0.5 * s.x, a.y - (S)0.5 * s.y);
    v[0] = cadd(a, s);
    v[1] = cadd(m, d);
    v[2] = csub(m, d);
}

template <typename V>
__device__ __forceinline__ void dft5(V* v, decltype(V::x) sgn)
{
    using S = decltype(V::x);
    const S c1 = (S)0.30901699437494742410, c2 = (S)-0.80901699437494742410;  // cos(2pi/5), cos(4pi/5)
    const S s1 = (S)0.95105651629515357212, s2 = (S)0.58778525229247312917;   // sin(2pi/5), sin(4pi/5)
    const V a = v[0];
    const V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const V t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const V p1 = mkc<V>(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const V p2 = mkc<V>(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    const V q1 = rot(mkc<V>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sgn);
    const V q2 = rot(mkc<V>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sgn);
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
}

template <typename V>
__device__ __forceinline__ void dft8(V* v, decltype(V::x) sgn)
{
    using S = decltype(V::x);
    const S r = (S)0.70710678118654752440;
    V e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    V o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, sgn);
    dft4(o0, o1, o2, o3, sgn);
    // W8^k = exp(sgn i pi k/4)
    const V w1 = mkc<V>(r * (o1.x - sgn * o1.y), r * (o1.y + sgn * o1.x));
    const V w2 = rot(o2, sgn);
    const V w3 = mkc<V>(-r * (o3.x + sgn * o3.y), r * (sgn * o3.x - o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1);
    v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2);
    v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3);
    v[7] = csub(e3, w3);
}

// Line classes: T threads, at most PPT points per thread (every butterfly's
// inputs are held in registers across the stage's barrier, so PPT sets the
// registers and with them how many lines a CU holds at once):
//   class 0: n <= 2048, T = 256, PPT = 8;  class 1: n <= 2560, T = 256, PPT = 10;
//   class 2: n <= 4096, T = 512, PPT = 8;  class 3: n <= kFftMaxLen (10240),
//   T = 1024, PPT = 10, one line per CU in up to 160 KiB of LDS.

// q = j / d, r = j % d for 0 <= j < 2^24, from a float reciprocal and one
// correction step (no integer division in the butterfly loops).
__device__ __forceinline__ void divmod(int j, int d, float rd, int& q, int& r)
{
    q = (int)((float)j * rd);
    r = j - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
}

template <int R, typename V>
__device__ __forceinline__ void butterfly(V* v, decltype(V::x) sgn);

// DFT of length R = R1 R2 in registers: with r = r1 + R1 r2 and k = k2 + R2 k1,
//   X[k] = sum_r1 W_R1^(r1 k1) W_R^(r1 k2) sum_r2 W_R2^(r2 k2) x[r1 + R1 r2],
// i.e. R1 DFTs of length R2, the twiddles W_R^(r1 k2) (compile-time constants),
// then R2 DFTs of length R1.
template <int R1, int R2, typename V>
__device__ __forceinline__ void dft_split(V* v, decltype(V::x) sgn)
{
    using S = decltype(V::x);
    constexpr int R = R1 * R2;
    V y[R1][R2];
#pragma unroll
    for (int r1 = 0; r1 < R1; ++r1) {
#pragma unroll
        for (int r2 = 0; r2 < R2; ++r2) y[r1][r2] = v[r1 + R1 * r2];
        butterfly<R2, V>(y[r1], sgn);
    }
#pragma unroll
    for (int r1 = 1; r1 < R1; ++r1)
#pragma unroll
        for (int k2 = 1; k2 < R2; ++k2) {
            const int m = (r1 * k2) % R;
            y[r1][k2] = cmul(y[r1][k2], mkc<V>((S)FftRoots<R>::c[m], sgn * (S)FftRoots<R>::s[m]));
        }
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
        V z[R1];
#pragma unroll
        for (int r1 = 0; r1 < R1; ++r1) z[r1] = y[r1][k2];
        butterfly<R1, V>(z, sgn);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) v[k2 + R2 * k1] = z[k1];
    }
}

// In-register DFT of radix R in {2, 3, 4, 5, 8} directly, {6, 9, 10, 12, 15,
// 16, 20, 25} split into two of those.
template <int R, typename V>
__device__ __forceinline__ void butterfly(V* v, decltype(V::x) sgn)
{
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v, sgn);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3], sgn);
    else if constexpr (R == 5) dft5(v, sgn);
    else if constexpr (R == 8) dft8(v, sgn);
    else if constexpr (R == 6) dft_split<2, 3, V>(v, sgn);
    else if constexpr (R == 9) dft_split<3, 3, V>(v, sgn);
    else if constexpr (R == 10) dft_split<2, 5, V>(v, sgn);
    else if constexpr (R == 12) dft_split<4, 3, V>(v, sgn);
    else if constexpr (R == 15) dft_split<3, 5, V>(v, sgn);
    else if constexpr (R == 16) dft_split<4, 4, V>(v, sgn);
    else if constexpr (R == 20) dft_split<4, 5, V>(v, sgn);
    else dft_split<5, 5, V>(v, sgn);
    static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 6 || R == 9 ||
                      R == 10 || R == 12 || R == 15 || R == 16 || R == 20 || R == 25,
                  "unsupported in-register radix");
}

// Pointers into LDS and global memory with their address space spelled out:
// the stage functions are not inlined (each keeps its own register budget,
// which the loop over runtime radices would otherwise merge), so the space
// must travel in the type for LDS to be accessed with ds_* and HBM with
// global_* instructions.
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));
template <typename S>
struct Mem;
template <>
struct Mem<double> {
    typedef __attribute__((address_space(3))) dvec2 lds;
    typedef __attribute__((address_space(1))) dvec2 glb;
};
template <>
struct Mem<float> {
    typedef __attribute__((address_space(3))) fvec2 lds;
    typedef __attribute__((address_space(1))) fvec2 glb;
};
// the twiddle tables hold FP64 values for both grid types
typedef Mem<double>::glb glb_d2;

template <typename S, typename Ptr>
__device__ __forceinline__ cplx<S> ld2(Ptr p, int i)
{
    const auto v = p[i];
    return mkc<cplx<S>>((S)v.x, (S)v.y);
}

template <typename V>
__device__ __forceinline__ V conj_if(V a, int c) { return mkc<V>(a.x, c ? -a.y : a.y); }

template <typename Ptr, typename V>
__device__ __forceinline__ void st2(Ptr p, int i, V a)
{
    auto v = p[i];
    v.x = a.x;
    v.y = a.y;
    p[i] = v;
}

// Where the lines of a transform sit in HBM: point k of line b is at
// b * li + k * ei of the input and b * lo + k * eo of the output.  Contiguous
// lines have (n, 1); lines that are the columns of a row-major grid with
// `nlines` columns have (1, nlines).
struct GStride {
    int li, ei, lo, eo;
};

// One Stockham stage of radix R over nb lines of length n.  Butterflies are
// staged through registers, so the lines stay in one LDS buffer: every thread
// reads its inputs, a barrier, every thread writes its outputs.  FIRST reads the
// lines from HBM (gin) and LAST writes them there (gout, which may be gin); the
// other stages read and write the LDS copy.  Twiddles: one table read per
// butterfly, the other R-2 as powers of it.  S is the grid's scalar type; the
// twiddle table is FP64 for both and rounded on load for FP32 lines.
template <int R, bool FIRST, bool LAST, int T, int PPT, typename S>
__device__ __attribute__((noinline)) void stage_reg(typename Mem<S>::lds* __restrict__ lds,
                                                    typename Mem<S>::glb* gin,
                                                    typename Mem<S>::glb* gout, int n, int nb,
                                                    int Ns, const glb_d2* __restrict__ tw,
                                                    int inverse, GStride gs)
{
    using V = cplx<S>;
    constexpr int BPT = (PPT + R - 1) / R;  // butterflies per thread at most
    const int m = n / R;
    const int base = n / (Ns * R);
    const float rNs = 1.0f / (float)Ns, rm = 1.0f / (float)m;
    const int total = nb * m;
    V v[BPT][R];
    int ob[BPT];
#pragma unroll
    for (int t = 0; t < BPT; ++t) {
        const int q = threadIdx.x + t * T;
        ob[t] = -1;
        if (q < total) {
            int b = 0, j = q;
            if (nb > 1) divmod(q, m, rm, b, j);
            int jq, jm;
            divmod(j, Ns, rNs, jq, jm);
            if constexpr (FIRST) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    v[t][r] = conj_if(ld2<S>(gin, b * gs.li + (j + r * m) * gs.ei), inverse);
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[t][r] = ld2<S>(lds, b * n + j + r * m);
            }
            if (Ns > 1) {
                const V w1 = ld2<S>(tw, jm * base);
                V w = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    v[t][r] = cmul(v[t][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            butterfly<R, V>(v[t], (S)-1.0);
            ob[t] = LAST ? b * gs.lo + (jq * Ns * R + jm) * gs.eo : b * n + jq * Ns * R + jm;
        }
    }
    if (FIRST == LAST) __syncthreads();  // inputs and outputs may share memory
    const int os = LAST ? Ns * gs.eo : Ns;
#pragma unroll
    for (int t = 0; t < BPT; ++t) {
        if (ob[t] >= 0) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if constexpr (LAST)
                    st2(gout, ob[t] + k * os, conj_if(v[t][k], inverse));
                else
                    st2(lds, ob[t] + k * Ns, v[t][k]);
            }
        }
    }
    if (!LAST) __syncthreads();
}

// Any radix: one output per (thread, slot), O(R) work per output.
template <bool FIRST, bool LAST, int T, int PPT, typename S>
__device__ __attribute__((noinline)) void stage_generic(typename Mem<S>::lds* __restrict__ lds,
                                                        typename Mem<S>::glb* gin,
                                                        typename Mem<S>::glb* gout, int n,
                                                        int nb, int Ns, int R,
                                                        const glb_d2* __restrict__ tw, int inverse,
                                                        GStride gs)
{
    using V = cplx<S>;
    constexpr int OPT = PPT;  // outputs per thread at most
    const int m = n / R;
    const int base = n / (Ns * R);
    const float rn = 1.0f / (float)n, rm = 1.0f / (float)m, rNs = 1.0f / (float)Ns;
    const int total = nb * n;
    V res[OPT];
    int ob[OPT];
#pragma unroll
    for (int t = 0; t < OPT; ++t) {
        const int q = threadIdx.x + t * T;
        ob[t] = -1;
        if (q < total) {
            int b, qq, k, j, jq, jm;
            divmod(q, n, rn, b, qq);
            divmod(qq, m, rm, k, j);
            divmod(j, Ns, rNs, jq, jm);
            const int step = (jm + k * Ns) * base;  // < n
            const int i0 = FIRST ? b * gs.li + j * gs.ei : b * n + j;
            const int is = FIRST ? m * gs.ei : m;
            S re = 0, im = 0;
            int e = 0;
            for (int r = 0; r < R; ++r) {
                V x;
                if constexpr (FIRST)
                    x = conj_if(ld2<S>(gin, i0 + r * is), inverse);
                else
                    x = ld2<S>(lds, i0 + r * is);
                const V wv = ld2<S>(tw, e);
                const S wi = wv.y;
                re += x.x * wv.x - x.y * wi;
                im += x.x * wi + x.y * wv.x;
                e += step;
                if (e >= n) e -= n;
            }
            res[t] = mkc<V>(re, im);
            ob[t] = LAST ? b * gs.lo + (jq * Ns * R + jm + k * Ns) * gs.eo
                         : b * n + jq * Ns * R + jm + k * Ns;
        }
    }
    if (FIRST == LAST) __syncthreads();
#pragma unroll
    for (int t = 0; t < OPT; ++t) {
        if (ob[t] >= 0) {
            if constexpr (LAST)
                st2(gout, ob[t], conj_if(res[t], inverse));
            else
                st2(lds, ob[t], res[t]);
        }
    }
    if (!LAST) __syncthreads();
}

// Job (s, b): lines b*B .. b*B + B - 1 of system s, along y (axis 0: My-lines)
// or x (axis 1: Mx-lines), read from grid `sel` and written contiguously (line l
// at l*n) to grid `osel` (which may be `sel`).  The input holds the lines
// contiguously too (tin 0) or as the columns of a row-major grid (tin 1: point k
// of line l at k*nlines + l, so the pass also transposes; a group of lines that
// shares cache lines runs on one XCD, and its 16-byte reads hit in that L2).
// The lines live in one LDS buffer of B*n points between the first and the last
// stage.  T threads with B*n <= PPT * T.  S = float: the grid holds float2
// values (half the bytes of the FP64 grid in the same buffer).
template <int T, int PPT, bool COMP, typename S>
__global__ __launch_bounds__(T) void k_fft_lines(const SysDesc* __restrict__ sys,
                                                 const Job* __restrict__ jobs, int axis, int sel,
                                                 int osel, int tin, int inverse)
{
    using LP = typename Mem<S>::lds;
    using GP = typename Mem<S>::glb;
    extern __shared__ double2 lds[];
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int n = axis ? d.Mx : d.My;
    const int nlines = axis ? d.My : d.Mx;
    const int B = axis ? d.Bx : d.By;
    const FftAxis& ax = axis ? d.ax : d.ay;
    const glb_d2* __restrict__ tw = (const glb_d2*)(axis ? d.twx : d.twy);
    const int l0 = jb.y * B;
    GP* gin = (GP*)d.grid[sel] + (tin ? (size_t)l0 : (size_t)l0 * n);
    GP* gout = (GP*)d.grid[osel] + (size_t)l0 * n;
    const GStride gs{tin ? 1 : n, tin ? nlines : 1, n, 1};
    LP* __restrict__ buf = (LP*)lds;
    const int nb = min(B, nlines - l0);
    // The table holds exp(-2 pi i k/n) and every stage runs the forward
    // transform; the inverse is conj(FFT(conj(x))): the first stage conjugates
    // what it loads and the last what it stores.
    const int nst = ax.nst;
    int Ns = 1;
    for (int st = 0; st < nst; ++st) {
        const int R = ax.radix[st];
        const bool first = st == 0, last = st == nst - 1;
#define SE2P_STAGE(RR)                                                                           \
    if (first && last) stage_reg<RR, true, true, T, PPT, S>(buf, gin, gout, n, nb, Ns, tw, inverse, gs); \
    else if (first) stage_reg<RR, true, false, T, PPT, S>(buf, gin, gout, n, nb, Ns, tw, inverse, gs);  \
    else if (last) stage_reg<RR, false, true, T, PPT, S>(buf, gin, gout, n, nb, Ns, tw, inverse, gs);   \
    else stage_reg<RR, false, false, T, PPT, S>(buf, gin, gout, n, nb, Ns, tw, inverse, gs)
        switch (R) {
            case 2: SE2P_STAGE(2); break;
            case 3: SE2P_STAGE(3); break;
            case 4: SE2P_STAGE(4); break;
            case 5: SE2P_STAGE(5); break;
            case 8: SE2P_STAGE(8); break;
#define SE2P_COMPOSITE(RR)             \
    case RR:                           \
        if constexpr (COMP) {          \
            SE2P_STAGE(RR);            \
        }                              \
        break;
            SE2P_COMPOSITE(6) SE2P_COMPOSITE(9) SE2P_COMPOSITE(10) SE2P_COMPOSITE(12)
            SE2P_COMPOSITE(15) SE2P_COMPOSITE(16)
#undef SE2P_COMPOSITE
            default:
                if (first && last)
                    stage_generic<true, true, T, PPT, S>(buf, gin, gout, n, nb, Ns, R, tw, inverse,
                                                         gs);
                else if (first)
                    stage_generic<true, false, T, PPT, S>(buf, gin, gout, n, nb, Ns, R, tw,
                                                          inverse, gs);
                else if (last)
                    stage_generic<false, true, T, PPT, S>(buf, gin, gout, n, nb, Ns, R, tw,
                                                          inverse, gs);
                else
                    stage_generic<false, false, T, PPT, S>(buf, gin, gout, n, nb, Ns, R, tw,
                                                           inverse, gs);
                break;
        }
#undef SE2P_STAGE
        Ns *= R;
    }
}

// The stages of one transform over nb lines in `buf`.  GIN: the first stage
// reads the lines from HBM (g, laid out as gs.li, gs.ei say), else from `buf`;
// GOUT: the last stage writes them to HBM (g, as gs.lo, gs.eo say), else to `buf`.
template <int T, int PPT, bool GIN, bool GOUT, typename S>
__device__ __forceinline__ void line_stages(typename Mem<S>::lds* __restrict__ buf,
                                            typename Mem<S>::glb* g, GStride gs, int n, int nb,
                                            const FftAxis& ax, const glb_d2* __restrict__ tw,
                                            int inverse)
{
    const int nst = ax.nst;
    int Ns = 1;
    for (int st = 0; st < nst; ++st) {
        const int R = ax.radix[st];
        const bool first = GIN && st == 0, last = GOUT && st == nst - 1;
#define SE2P_STAGE(RR)                                                                          \
    if (first && last)                                                                          \
        stage_reg<RR, true, true, T, PPT, S>(buf, g, g, n, nb, Ns, tw, inverse, gs);          \
    else if (first)                                                                             \
        stage_reg<RR, true, false, T, PPT, S>(buf, g, g, n, nb, Ns, tw, inverse, gs);         \
    else if (last)                                                                              \
        stage_reg<RR, false, true, T, PPT, S>(buf, g, g, n, nb, Ns, tw, inverse, gs);         \
    else                                                                                        \
        stage_reg<RR, false, false, T, PPT, S>(buf, g, g, n, nb, Ns, tw, inverse, gs)
        switch (R) {
            case 2: SE2P_STAGE(2); break;
            case 3: SE2P_STAGE(3); break;
            case 4: SE2P_STAGE(4); break;
            case 5: SE2P_STAGE(5); break;
            case 8: SE2P_STAGE(8); break;
            default:
                if (first && last)
                    stage_generic<true, true, T, PPT, S>(buf, g, g, n, nb, Ns, R, tw, inverse, gs);
                else if (first)
                    stage_generic<true, false, T, PPT, S>(buf, g, g, n, nb, Ns, R, tw, inverse,
                                                          gs);
                else if (last)
                    stage_generic<false, true, T, PPT, S>(buf, g, g, n, nb, Ns, R, tw, inverse,
                                                          gs);
                else
                    stage_generic<false, false, T, PPT, S>(buf, g, g, n, nb, Ns, R, tw, inverse,
                                                           gs);
                break;
        }
#undef SE2P_STAGE
        Ns *= R;
    }
}

// The filter of rows p and q = -p (mod My) held in buf (nb = 1: row p is its own
// partner), in place, each output stored conjugated (the inverse transform runs
// the forward stages on conj(U)).  Not inlined: its constants then live in its
// own register budget instead of the kernel's.  The periodic Stokeslet filter
// only: the free-space one's Bessel functions would not fit beside the stages.
template <int T, typename S>
__device__ __attribute__((noinline)) void filter_rows(typename Mem<S>::lds* __restrict__ buf,
                                                      const SysDesc& d, int n, int nb, int p,
                                                      int q)
{
    using V = cplx<S>;
    for (int ix = threadIdx.x; ix < n; ix += T) {
        const int jx = ix ? n - ix : 0;
        if (nb == 1 && jx < ix) continue;  // taken with its partner
        const int im = (nb - 1) * n + jx;  // buf index of mode (-ix, -p)
        const V a = ld2<S>(buf, ix), b = ld2<S>(buf, im);
        const double2 wk = make_double2(a.x, a.y), wm = make_double2(b.x, b.y);
        const double2 o = kf::slp_filter_periodic(d, ix, p, wk, wm);
        st2(buf, ix, mkc<V>((S)o.x, -(S)o.y));
        if (im != ix) {
            const double2 r = kf::slp_filter_periodic(d, jx, q, wm, wk);
            st2(buf, im, mkc<V>((S)r.x, -(S)r.y));
        }
    }
}

// The Stokeslet's x-direction passes with the filter between them, for x lines
// of 512 to 2048 points: job (s, p) loads the x lines p and My - p of the
// y-transformed spectrum in the column-major Z0 (the rows of modes k and -k,
// read across the columns, which also transposes), transforms them in LDS,
// filters every mode with its partner (unpack, Stokes filter, k = 0 removal,
// 1/(Mx My), repack: the same filter_mode as k_filter), transforms them back and
// writes the rows of the row-major W.  Rows 0 and My/2 are their own partners
// and share one job.
// It runs the same stages and filter arithmetic as the separate passes on values
// that stay in LDS instead of HBM; the results agree with theirs to rounding
// (within 4e-16 relative on FP64 grids, where instruction selection differs in
// the last bits; FP32 grids bit for bit), and the grid makes one HBM round trip
// instead of three.  512 threads, 8 points each: two
// lines of up to 2048 points.
template <typename S>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_fft_xfilter(
    const SysDesc* __restrict__ sys,
                                                     const Job* __restrict__ jobs)
{
    constexpr int T = 512, PPT = 8;
    using LP = typename Mem<S>::lds;
    using GP = typename Mem<S>::glb;
    extern __shared__ double2 lds[];
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int n = d.Mx, My = d.My;
    // job p < My/2: rows p and My - p; job 0 takes rows 0 and My/2, each its own
    // partner (My/2 jobs, a whole number of rounds of two per CU)
    const int p = jb.y, q = p ? My - p : My / 2;
    const int nb = 2;
    const glb_d2* __restrict__ tw = (const glb_d2*)d.twx;
    LP* __restrict__ buf = (LP*)lds;
    // rows p and q of the y-transformed column-major grid Z0 (point ix of row p
    // at ix*My + p); the jobs of one XCD take consecutive p, so their 16-byte
    // loads share cache lines in its L2
    line_stages<T, PPT, true, false, S>(buf, (GP*)d.grid[G_Z0] + p, GStride{q - p, My, 0, 0}, n,
                                        nb, d.ax, tw, 0);
    // the stages end with a barrier: every forward value is in buf
    if (p) {
        filter_rows<T, S>(buf, d, n, 2, p, q);
    } else {
        filter_rows<T, S>(buf, d, n, 1, 0, 0);
        filter_rows<T, S>(buf + n, d, n, 1, q, q);
    }
    __syncthreads();
    // rows p and q of the row-major W
    line_stages<T, PPT, false, true, S>(buf, (GP*)d.grid[G_W] + (size_t)p * n,
                                        GStride{0, 0, (q - p) * n, 1}, n, nb, d.ax, tw, 1);
}

// Lines above kFftMaxLenSmall need more than 64 KiB of LDS per workgroup; set
// when the device accepts that for the class-3 kernels.
bool g_long_lines = false;

}  // namespace

int fft_lines_per_block(int n) { return n >= 256 ? 1 : (256 + n - 1) / n; }

size_t fft_lds_bytes(int n, int prec)
{
    return (prec ? sizeof(float2) : sizeof(double2)) * (size_t)fft_lines_per_block(n) * n;
}

int fft_axis_init(FftAxis* a, int n)
{
    if (n < 1 || n > kFftMaxLen)
        return set_error(SE2P_ERR_UNSUPPORTED, "FFT length %d outside [1, %d]", n, kFftMaxLen);
    if (n > kFftMaxLenSmall) {
        SE2P_TRY(fft_prepare());
        if (!g_long_lines)
            return set_error(SE2P_ERR_UNSUPPORTED,
                             "FFT length %d: lines above %d need 160 KiB of LDS per workgroup, "
                             "which this device refused", n, kFftMaxLenSmall);
    }
    a->n = n;
    a->nst = 0;
    // The 2-3-5-smooth part of n in as few in-register stages as possible (each
    // stage is an LDS round trip and two barriers), larger radices first; the
    // composite radices only in the class whose kernel has them (their
    // registers would cost the other classes a line per CU).  Every other prime
    // factor becomes a stage of its own (7 and up run as direct sums).
    const bool comp = fft_line_class(n) == 1;
    static const int rad_comp[] = {16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};
    static const int rad_base[] = {8, 5, 4, 3, 2};
    int smooth = 1, r = n;
    for (int p : {2, 3, 5})
        while (r % p == 0) {
            smooth *= p;
            r /= p;
        }
    std::vector<int> best(smooth + 1, 1 << 20), pick(smooth + 1, 0);
    best[1] = 0;
    for (int d = 2; d <= smooth; ++d) {
        if (smooth % d) continue;
        auto consider = [&](int q) {
            if (d % q == 0 && best[d / q] + 1 < best[d]) {
                best[d] = best[d / q] + 1;
                pick[d] = q;
            }
        };
        if (comp)
            for (int q : rad_comp) consider(q);
        else
            for (int q : rad_base) consider(q);
    }
    for (int d = smooth; d > 1; d /= pick[d]) a->radix[a->nst++] = pick[d];
    for (int p = 7; r > 1; p += 2)
        while (r % p == 0) {
            if (a->nst >= kFftMaxStages) return set_error(SE2P_ERR_UNSUPPORTED, "FFT length %d", n);
            a->radix[a->nst++] = p;
            r /= p;
        }
    return SE2P_OK;
}

int fft_twiddles(se2p_ctx* ctx, Workspace* ws, int n, Slot slot, double2** tw)
{
    SE2P_TRY(fft_prepare());
    void* d = nullptr;
    SE2P_TRY(ws_get(ctx, ws, slot, sizeof(double2) * (size_t)n, &d));
    std::vector<double2> h(n);
    for (int k = 0; k < n; ++k) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / n;
        h[k] = make_double2((double)cosl(a), (double)sinl(a));
    }
    SE2P_HIP_CHECK(hipMemcpyAsync(d, h.data(), sizeof(double2) * (size_t)n, hipMemcpyHostToDevice,
                                  ctx->stream));
    // The host vector dies on return: wait for the copy.
    SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *tw = (double2*)d;
    return SE2P_OK;
}

int fft_prepare()
{
    static bool attr_set = false;
    if (!attr_set) {
        const void* kernels[] = {(const void*)k_fft_lines<256, 8, false, double>,
                                 (const void*)k_fft_lines<256, 10, true, double>,
                                 (const void*)k_fft_lines<512, 8, false, double>,
                                 (const void*)k_fft_lines<256, 8, false, float>,
                                 (const void*)k_fft_lines<256, 10, true, float>,
                                 (const void*)k_fft_lines<512, 8, false, float>};
        for (const void* k : kernels)
            SE2P_HIP_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               kFftMaxLenSmall * (int)sizeof(double2)));
        const hipError_t e1 = hipFuncSetAttribute(
            (const void*)k_fft_lines<1024, 10, false, double>,
            hipFuncAttributeMaxDynamicSharedMemorySize, kFftMaxLen * (int)sizeof(double2));
        const hipError_t e2 = hipFuncSetAttribute(
            (const void*)k_fft_lines<1024, 10, false, float>,
            hipFuncAttributeMaxDynamicSharedMemorySize, kFftMaxLen * (int)sizeof(float2));
        g_long_lines = e1 == hipSuccess && e2 == hipSuccess;
        if (!g_long_lines) (void)hipGetLastError();  // lines above 4096 stay unsupported
        attr_set = true;
    }
    return SE2P_OK;
}

int fft_line_class(int n) { return n <= 2048 ? 0 : n <= 2560 ? 1 : n <= 4096 ? 2 : 3; }

template <typename S>
static void fft_lines_launch(hipStream_t s, int axis, int sel, int osel, int tin, int inverse,
                             int cls, size_t max_lds, const SysDesc* sys, const Job* jobs,
                             int njobs)
{
    if (cls == 0)
        hipLaunchKernelGGL((k_fft_lines<256, 8, false, S>), dim3(njobs), dim3(256), max_lds, s, sys,
                           jobs, axis, sel, osel, tin, inverse);
    else if (cls == 1)
        hipLaunchKernelGGL((k_fft_lines<256, 10, true, S>), dim3(njobs), dim3(256), max_lds, s, sys,
                           jobs, axis, sel, osel, tin, inverse);
    else if (cls == 2)
        hipLaunchKernelGGL((k_fft_lines<512, 8, false, S>), dim3(njobs), dim3(512), max_lds, s, sys,
                           jobs, axis, sel, osel, tin, inverse);
    else
        hipLaunchKernelGGL((k_fft_lines<1024, 10, false, S>), dim3(njobs), dim3(1024), max_lds, s,
                           sys, jobs, axis, sel, osel, tin, inverse);
}

int launch_fft_lines(hipStream_t s, int prec, int axis, int sel, int osel, int tin, int inverse,
                     int cls, size_t max_lds, const SysDesc* sys, const Job* jobs, int njobs)
{
    if (njobs <= 0) return SE2P_OK;
    if (prec)
        fft_lines_launch<float>(s, axis, sel, osel, tin, inverse, cls, max_lds, sys, jobs, njobs);
    else
        fft_lines_launch<double>(s, axis, sel, osel, tin, inverse, cls, max_lds, sys, jobs,
                                 njobs);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

bool fft_xfilter_fits(int kind, int Mx, double free_R)
{
    return kind == SE2P_SLP && !(free_R > 0) && Mx >= 512 && Mx <= 2048;
}

int launch_fft_xfilter(hipStream_t s, int prec, size_t max_lds, const SysDesc* sys,
                       const Job* jobs, int njobs)
{
    if (njobs <= 0) return SE2P_OK;
    if (prec)
        hipLaunchKernelGGL(k_fft_xfilter<float>, dim3(njobs), dim3(512), max_lds, s, sys, jobs);
    else
        hipLaunchKernelGGL(k_fft_xfilter<double>, dim3(njobs), dim3(512), max_lds, s, sys, jobs);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

}  // namespace se2p
