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

__global__ __launch_bounds__(256) void k_fft_lines(double2* __restrict__ data, int n, int nlines,
                                                   long elem_stride, long line_stride, int B,
                                                   FftAxis ax, const double2* __restrict__ tw,
                                                   int inverse)
{
    extern __shared__ double2 lds[];
    const int line0 = blockIdx.x * B;
    const int nb = min(B, nlines - line0);
    const int total = nb * n;
    double2* src = lds;
    double2* dst = lds + (size_t)B * n;
    const bool contiguous = (elem_stride == 1);
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        int b, e;
        if (contiguous) {
            b = idx / n;
            e = idx % n;
        } else {
            b = idx % nb;
            e = idx / nb;
        }
        src[b * n + e] = data[(line0 + b) * line_stride + e * elem_stride];
    }
    __syncthreads();
    const double sgn = inverse ? -1.0 : 1.0;  // table holds exp(-2 pi i k/n)
    int Ns = 1;
    for (int s = 0; s < ax.nst; ++s) {
        const int R = ax.radix[s];
        const int m = n / R;
        const int base = n / (Ns * R);
        for (int q = threadIdx.x; q < total; q += blockDim.x) {
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
                const double wi = sgn * wv.y;
                re += v.x * wv.x - v.y * wi;
                im += v.x * wi + v.y * wv.x;
                e += step;
                if (e >= n) e -= n;
            }
            dst[b * n + (j / Ns) * Ns * R + jm + k * Ns] = make_double2(re, im);
        }
        __syncthreads();
        double2* tmp = src;
        src = dst;
        dst = tmp;
        Ns *= R;
    }
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        int b, e;
        if (contiguous) {
            b = idx / n;
            e = idx % n;
        } else {
            b = idx % nb;
            e = idx / nb;
        }
        data[(line0 + b) * line_stride + e * elem_stride] = src[b * n + e];
    }
}

int lines_per_block(int n, bool contiguous)
{
    if (contiguous) return n >= 256 ? 1 : (256 + n - 1) / n;
    int B = kFftMaxLen / n;
    if (B > 16) B = 16;
    return B < 1 ? 1 : B;
}

}  // namespace

int fft_prepare();

int fft_axis_init(FftAxis* a, int n)
{
    if (n < 1 || n > kFftMaxLen)
        return set_error(SE2P_ERR_UNSUPPORTED, "FFT length %d outside [1, %d]", n, kFftMaxLen);
    a->n = n;
    a->nst = 0;
    int r = n;
    const int small[] = {4, 2, 3, 5, 7};
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

int fft2d(se2p_ctx* ctx, double2* Z, int Mx, int My, const FftAxis& ax, const FftAxis& ay,
          const double2* twx, const double2* twy, int inverse)
{
    if (!ctx->capturing) SE2P_TRY(fft_prepare());
    hipStream_t s = ctx->stream;
    // y lines: contiguous, Mx of them
    {
        const int B = lines_per_block(My, true);
        const size_t lds = 2 * sizeof(double2) * (size_t)B * My;
        hipLaunchKernelGGL(k_fft_lines, dim3((Mx + B - 1) / B), dim3(256), lds, s, Z, My, Mx, 1L,
                           (long)My, B, ay, twy, inverse);
        SE2P_HIP_CHECK(hipGetLastError());
    }
    // x lines: stride My, My of them
    {
        const int B = lines_per_block(Mx, false);
        const size_t lds = 2 * sizeof(double2) * (size_t)B * Mx;
        hipLaunchKernelGGL(k_fft_lines, dim3((My + B - 1) / B), dim3(256), lds, s, Z, Mx, My,
                           (long)My, 1L, B, ax, twx, inverse);
        SE2P_HIP_CHECK(hipGetLastError());
    }
    return SE2P_OK;
}

}  // namespace se2p
