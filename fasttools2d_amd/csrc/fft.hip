// 2D complex FFT for the spectral Ewald grids.
//
// The reference transforms with MATLAB's fft2/ifft2 called back through
// mexCallMATLAB (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_kspace.cpp:80-81,
// 176-177).  This is the library's own transform: a mixed-radix Stockham FFT
// with one workgroup per line (or per group of lines), the line held in LDS and
// ping-ponged between two LDS buffers stage by stage.  Any length up to
// kFftMaxLen works: radices 4 and 2 first, then 3, 5, 7 and any remaining prime
// factor.  Stage s with radix R and Ns = product of the earlier radices maps
// output (j, k), j < n/R, k < R, to
//   out[(j / Ns) Ns R + j % Ns + k Ns] = sum_r in[j + r n/R] W_n^{r ((j % Ns) + k Ns) n/(Ns R)}.
// The grid is My x Mx column-major (Z[ix*My + iy]): y lines are contiguous,
// x lines have stride My and are loaded B at a time so each global access
// touches B consecutive elements.
#include "se2p_internal.h"

#include <cmath>
#include <vector>

namespace se2p {

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// v * (sgn i)
__device__ __forceinline__ double2 rot(double2 v, double sgn) { return make_double2(-sgn * v.y, sgn * v.x); }

// In-register DFTs of length R with kernel exp(sgn 2 pi i r k / R), sgn = -1 forward.
__device__ __forceinline__ void dft2(double2* v)
{
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

__device__ __forceinline__ void dft4(double2& a, double2& b, double2& c, double2& d, double sgn)
{
    const double2 t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = rot(csub(b, d), sgn);
    a = cadd(t0, t2);
    c = csub(t0, t2);
    b = cadd(t1, t3);
    d = csub(t1, t3);
}

__device__ __forceinline__ void dft3(double2* v, double sgn)
{
    const double h = 0.86602540378443864676;  // sin(2 pi/3)
    const double2 a = v[0], s = cadd(v[1], v[2]), d = rot(csub(v[1], v[2]), sgn * h);
    const double2 m = make_double2(a.x - 0.5 * s.x, a.y - 0.5 * s.y);
    v[0] = cadd(a, s);
    v[1] = cadd(m, d);
    v[2] = csub(m, d);
}

__device__ __forceinline__ void dft5(double2* v, double sgn)
{
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos(2pi/5), cos(4pi/5)
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin(2pi/5), sin(4pi/5)
    const double2 a = v[0];
    const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const double2 p1 = make_double2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const double2 p2 = make_double2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    const double2 q1 = rot(make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sgn);
    const double2 q2 = rot(make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sgn);
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
}

__device__ __forceinline__ void dft8(double2* v, double sgn)
{
    const double r = 0.70710678118654752440;
    double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    double2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, sgn);
    dft4(o0, o1, o2, o3, sgn);
    // W8^k = exp(sgn i pi k/4)
    const double2 w1 = make_double2(r * (o1.x - sgn * o1.y), r * (o1.y + sgn * o1.x));
    const double2 w2 = rot(o2, sgn);
    const double2 w3 = make_double2(-r * (o3.x + sgn * o3.y), r * (sgn * o3.x - o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1);
    v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2);
    v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3);
    v[7] = csub(e3, w3);
}

// Small odd radix with table twiddles: W_R^q = tw[q n/R].
template <int R>
__device__ __forceinline__ void dft_small(double2* v, const double2* __restrict__ tw, int n,
                                          double sgn)
{
    double2 out[R];
    const int step = n / R;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double2 acc = v[0];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const double2 w = tw[((r * k) % R) * step];
            acc = cadd(acc, cmul(v[r], make_double2(w.x, -sgn * w.y)));
        }
        out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
}

// One Stockham stage of radix R over nb lines of length n in LDS.
template <int R>
__device__ __forceinline__ void stage(const double2* __restrict__ src, double2* __restrict__ dst,
                                      int n, int nb, int Ns, const double2* __restrict__ tw,
                                      double sgn)
{
    const int m = n / R;
    const int base = n / (Ns * R);
    for (int q = threadIdx.x; q < nb * m; q += blockDim.x) {
        const int b = q / m;
        const int j = q - b * m;
        const int jm = j % Ns;
        const double2* in = src + b * n + j;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[r * m];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const double2 w = tw[r * jm * base];
                v[r] = cmul(v[r], make_double2(w.x, -sgn * w.y));
            }
        }
        if (R == 2) dft2(v);
        else if (R == 3) dft3(v, sgn);
        else if (R == 4) dft4(v[0], v[1], v[2], v[3], sgn);
        else if (R == 5) dft5(v, sgn);
        else if (R == 8) dft8(v, sgn);
        else dft_small<R>(v, tw, n, sgn);
        double2* out = dst + b * n + (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int k = 0; k < R; ++k) out[k * Ns] = v[k];
    }
}

// Any radix: one output per thread, O(R) work per output.
__device__ void stage_generic(const double2* __restrict__ src, double2* __restrict__ dst, int n,
                              int nb, int Ns, int R, const double2* __restrict__ tw, double sgn)
{
    const int m = n / R;
    const int base = n / (Ns * R);
    for (int q = threadIdx.x; q < nb * n; q += blockDim.x) {
        const int b = q / n;
        const int qq = q - b * n;
        const int k = qq / m;
        const int j = qq - k * m;
        const int jm = j % Ns;
        const int step = (jm + k * Ns) * base;  // < n
        const double2* in = src + b * n + j;
        double re = 0, im = 0;
        int e = 0;
        for (int r = 0; r < R; ++r) {
            const double2 v = in[r * m];
            const double2 wv = tw[e];
            const double wi = -sgn * wv.y;
            re += v.x * wv.x - v.y * wi;
            im += v.x * wi + v.y * wv.x;
            e += step;
            if (e >= n) e -= n;
        }
        dst[b * n + (j / Ns) * Ns * R + jm + k * Ns] = make_double2(re, im);
    }
}

// Job (s, b): lines b*B .. b*B + B - 1 of grid `sel` of system s, along y
// (axis 0: My-lines, the column-major grid) or x (axis 1: Mx-lines, the
// transposed grid); every line is contiguous.
__global__ __launch_bounds__(256) void k_fft_lines(const SysDesc* __restrict__ sys,
                                                   const Job* __restrict__ jobs, int axis, int sel,
                                                   int inverse)
{
    extern __shared__ double2 lds[];
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int n = axis ? d.Mx : d.My;
    const int nlines = axis ? d.My : d.Mx;
    const int B = axis ? d.Bx : d.By;
    const FftAxis& ax = axis ? d.ax : d.ay;
    const double2* __restrict__ tw = axis ? d.twx : d.twy;
    double2* __restrict__ data = d.grid[sel] + (size_t)jb.y * B * n;
    const int nb = min(B, nlines - jb.y * B);
    const int total = nb * n;
    double2* src = lds;
    double2* dst = lds + (size_t)B * n;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) src[idx] = data[idx];
    __syncthreads();
    // the table holds exp(-2 pi i k/n); sgn is the sign of the transform's exponent
    const double sgn = inverse ? 1.0 : -1.0;
    int Ns = 1;
    for (int st = 0; st < ax.nst; ++st) {
        const int R = ax.radix[st];
        switch (R) {
            case 2: stage<2>(src, dst, n, nb, Ns, tw, sgn); break;
            case 3: stage<3>(src, dst, n, nb, Ns, tw, sgn); break;
            case 4: stage<4>(src, dst, n, nb, Ns, tw, sgn); break;
            case 5: stage<5>(src, dst, n, nb, Ns, tw, sgn); break;
            case 7: stage<7>(src, dst, n, nb, Ns, tw, sgn); break;
            case 8: stage<8>(src, dst, n, nb, Ns, tw, sgn); break;
            case 11: stage<11>(src, dst, n, nb, Ns, tw, sgn); break;
            case 13: stage<13>(src, dst, n, nb, Ns, tw, sgn); break;
            default: stage_generic(src, dst, n, nb, Ns, R, tw, sgn); break;
        }
        __syncthreads();
        double2* tmp = src;
        src = dst;
        dst = tmp;
        Ns *= R;
    }
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) data[idx] = src[idx];
}

// out[c * rows + r] = in[r * cols + c], 32 x 32 tiles through LDS; dir 0: rows
// = Mx, cols = My (column-major grid to transposed), dir 1 the reverse.
__global__ __launch_bounds__(256) void k_transpose(const SysDesc* __restrict__ sys,
                                                   const Job* __restrict__ jobs, int in_sel,
                                                   int out_sel, int dir)
{
    __shared__ double2 tile[32][33];
    const Job jb = jobs[blockIdx.x];
    const SysDesc& d = sys[jb.x];
    const int rows = dir ? d.My : d.Mx, cols = dir ? d.Mx : d.My;
    const double2* __restrict__ in = d.grid[in_sel];
    double2* __restrict__ out = d.grid[out_sel];
    const int ntc = (cols + 31) / 32;
    const int r0 = (jb.y / ntc) * 32, c0 = (jb.y % ntc) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[tx][k];
    }
}

}  // namespace

int fft_lines_per_block(int n) { return n >= 256 ? 1 : (256 + n - 1) / n; }

size_t fft_lds_bytes(int n) { return 2 * sizeof(double2) * (size_t)fft_lines_per_block(n) * n; }

int fft_axis_init(FftAxis* a, int n)
{
    if (n < 1 || n > kFftMaxLen)
        return set_error(SE2P_ERR_UNSUPPORTED, "FFT length %d outside [1, %d]", n, kFftMaxLen);
    a->n = n;
    a->nst = 0;
    int r = n;
    const int small[] = {8, 4, 2, 3, 5, 7};
    for (int p : small)
        while (r % p == 0) {
            a->radix[a->nst++] = p;
            r /= p;
        }
    for (int p = 11; r > 1; p += 2)
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
        SE2P_HIP_CHECK(hipFuncSetAttribute((const void*)k_fft_lines,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           2 * kFftMaxLen * (int)sizeof(double2)));
        attr_set = true;
    }
    return SE2P_OK;
}

int launch_fft_lines(hipStream_t s, int axis, int sel, int inverse, size_t max_lds,
                     const SysDesc* sys, const Job* jobs, int njobs)
{
    if (njobs <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_fft_lines, dim3(njobs), dim3(256), max_lds, s, sys, jobs, axis, sel,
                       inverse);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int transpose_tiles(int rows, int cols) { return ((rows + 31) / 32) * ((cols + 31) / 32); }

int launch_transpose(hipStream_t s, int in, int out, int dir, const SysDesc* sys,
                     const Job* jobs, int njobs)
{
    if (njobs <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_transpose, dim3(njobs), dim3(256), 0, s, sys, jobs, in, out, dir);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

}  // namespace se2p
