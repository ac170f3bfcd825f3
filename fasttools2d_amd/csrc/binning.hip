// Binning of points by key with a deterministic, stable counting sort.
//
// Cell keys follow Assign (PeriodicEwald/src/mex/common/ewald_tools.cpp:7-80);
// grid-tile keys use FindClosestNode (ewald_tools.cpp:86-114).  The
// reference's serial counting sort visits points in input order, so each
// bucket lists its points by increasing index.  Here the histogram and the
// scatter run with integer atomics (the bucket *sets* are order independent)
// and every bucket is then sorted by index, which yields exactly that order.
#include "se2p_internal.h"

#include <climits>

namespace se2p {

namespace {

constexpr int kSortCap = kSortCellCap;  // largest bucket sorted in LDS by one workgroup

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

__global__ void k_cell_keys(const double* __restrict__ p, int n, double Lx, double Ly, int nsx,
                            int nsy, int* __restrict__ key, int* __restrict__ count)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = p[2 * i], y = p[2 * i + 1];
    int bx = clampi((int)floor(nsx * (x / Lx + 0.5)), 0, nsx - 1);
    int by = clampi((int)floor(nsy * (y / Ly + 0.5)), 0, nsy - 1);
    int k = by * nsx + bx;
    key[i] = k;
    atomicAdd(&count[k], 1);
}

// FindClosestNode with the same operation order and no contraction, so the
// start node and the snap decisions are bit-identical to the reference.
__device__ __forceinline__ void closest_node(double x, double y, double Lx, double Ly, double h,
                                             int P, int* mx, int* my, double* px, double* py)
{
#pragma clang fp contract(off)
    const double snap = 1e-13;
    double ox = x - h * floor(x / h);
    double oy = y - h * floor(y / h);
    double hx = Lx / 2.0, hy = Ly / 2.0;
    int half = P / 2;
    int sx = (int)(floor((x + hx) / h) - half);
    int sy = (int)(floor((y + hy) / h) - half);
    if (fabs(ox) < snap || fabs(ox - h) < snap) ox = 0;
    if (fabs(oy) < snap || fabs(oy - h) < snap) oy = 0;
    if (ox == 0 && remainder((x + hx) / h - half, 1.0) < 0) sx++;
    if (oy == 0 && remainder((y + hy) / h - half, 1.0) < 0) sy++;
    *mx = sx;
    *my = sy;
    *px = ox;
    *py = oy;
}

__global__ void k_tile_keys(const double* __restrict__ p, int n, double Lx, double Ly, double h,
                            int P, int Mx, int My, int TX, int TY, int ntx,
                            int* __restrict__ key, int* __restrict__ count,
                            int2* __restrict__ node, double2* __restrict__ off)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int mx, my;
    double px, py;
    closest_node(p[2 * i], p[2 * i + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
    int wx = ((mx % Mx) + Mx) % Mx;
    int wy = ((my % My) + My) % My;
    int k = (wy / TY) * ntx + (wx / TX);
    key[i] = k;
    node[i] = make_int2(wx, wy);
    off[i] = make_double2(px, py);
    atomicAdd(&count[k], 1);
}

__global__ void k_strip_count(const double* __restrict__ p, int n, double Lx, double Ly,
                              double h, int P, int Mx, int My, int ntx, int* __restrict__ cnt,
                              int2* __restrict__ node, double2* __restrict__ off)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int mx, my;
    double px, py;
    closest_node(p[2 * i], p[2 * i + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
    const int wx = ((mx % Mx) + Mx) % Mx;
    const int wy = ((my % My) + My) % My;
    node[i] = make_int2(wx, wy);
    off[i] = make_double2(px, py);
    int c = 0;
    for_each_strip(wx, Mx, P, ntx, [&](int, int) { ++c; });
    cnt[i] = c;
}

__global__ void k_strip_keys(const int2* __restrict__ node, const int* __restrict__ offs, int n,
                             int P, int Mx, int My, int ntx, int* __restrict__ key,
                             int* __restrict__ count, int2* __restrict__ rnode,
                             int* __restrict__ owner)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 m = node[i];
    int o = offs[i];
    for_each_strip(m.x, Mx, P, ntx, [&](int s, int ox) {
        const int k = s * My + m.y;
        key[o] = k;
        rnode[o] = make_int2(ox, m.y);
        owner[o] = i;
        ++o;
        atomicAdd(&count[k], 1);
    });
}

__global__ void k_closest_nodes(const double* __restrict__ p, int n, double Lx, double Ly,
                                double h, int P, int* __restrict__ mx, int* __restrict__ my,
                                double* __restrict__ px, double* __restrict__ py)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    closest_node(p[2 * i], p[2 * i + 1], Lx, Ly, h, P, &mx[i], &my[i], &px[i], &py[i]);
}

__global__ void k_fill_int(int* p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Exclusive scan of count[0..n) into start[0..n]; one workgroup of 1024 threads,
// each owning a contiguous segment.
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ count, int n,
                                               int* __restrict__ start)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int seg = (n + 1023) / 1024;
    const int b = t * seg, e = min(n, b + seg);
    int s = 0;
    for (int i = b; i < e; ++i) s += count[i];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        int v = (t >= d) ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;  // exclusive prefix of this segment
    for (int i = b; i < e; ++i) {
        start[i] = run;
        run += count[i];
    }
    if (t == 1023) start[n] = part[1023];
}

// Longer scans in three passes: per-tile sums (kScanTile entries per 256-lane
// workgroup, 16 contiguous per lane), a k_scan of the tile sums, then each tile
// rescanned from its offset.
constexpr int kScanTile = 4096;

__device__ __forceinline__ int block_exclusive_scan(int v, int* part, int& total)
{
    const int t = threadIdx.x;
    part[t] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int u = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += u;
        __syncthreads();
    }
    total = part[255];
    const int r = part[t] - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_scan_tiles(const int* __restrict__ count, int n,
                                                    int* __restrict__ tsum)
{
    __shared__ int part[256];
    const int b = blockIdx.x * kScanTile + threadIdx.x * 16;
    int s = 0;
    for (int i = b; i < min(n, b + 16); ++i) s += count[i];
    int total;
    block_exclusive_scan(s, part, total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_apply(const int* __restrict__ count, int n,
                                                    const int* __restrict__ toff,
                                                    int* __restrict__ start)
{
    __shared__ int part[256];
    const int b = blockIdx.x * kScanTile + threadIdx.x * 16;
    int s = 0;
    for (int i = b; i < min(n, b + 16); ++i) s += count[i];
    int total;
    int run = toff[blockIdx.x] + block_exclusive_scan(s, part, total);
    for (int i = b; i < min(n, b + 16); ++i) {
        start[i] = run;
        run += count[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[n] = toff[gridDim.x];
}

__global__ void k_copy_int(const int* __restrict__ a, int* __restrict__ b, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_scatter(const int* __restrict__ key, int n, int* __restrict__ cursor,
                          int* __restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int pos = atomicAdd(&cursor[key[i]], 1);
    perm[pos] = i;
}

// One workgroup per bucket: bitonic sort of the bucket's indices in LDS.
// Buckets larger than kSortCap are queued in biglist (biglist[0] = count).
__global__ __launch_bounds__(256) void k_sort_buckets(int* __restrict__ perm,
                                                      const int* __restrict__ start,
                                                      int* __restrict__ biglist)
{
    __shared__ int s[kSortCap];
    const int b = blockIdx.x;
    const int lo = start[b], size = start[b + 1] - lo;
    if (size <= 1) return;
    if (size > kSortCap) {
        if (threadIdx.x == 0) {
            int slot = atomicAdd(&biglist[0], 1);
            biglist[1 + slot] = b;
        }
        return;
    }
    int np = 2;
    while (np < size) np <<= 1;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s[i] = i < size ? perm[lo + i] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                int ij = i ^ j;
                if (ij > i) {
                    int a = s[i], c = s[ij];
                    bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        s[i] = c;
                        s[ij] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < size; i += blockDim.x) perm[lo + i] = s[i];
}

// Buckets beyond the LDS capacity: each element's final position is its rank
// among the (distinct) indices of its bucket.  O(size^2) per bucket; only
// reached by heavily clustered inputs.
__global__ __launch_bounds__(256) void k_sort_big(int* __restrict__ perm,
                                                  const int* __restrict__ start,
                                                  const int* __restrict__ biglist,
                                                  int* __restrict__ tmp)
{
    __shared__ int tile[1024];
    const int nbig = biglist[0];
    for (int q = blockIdx.x; q < nbig; q += gridDim.x) {
        const int b = biglist[1 + q];
        const int lo = start[b], size = start[b + 1] - lo;
        for (int i0 = 0; i0 < size; i0 += blockDim.x) {
            int i = i0 + threadIdx.x;
            int v = i < size ? perm[lo + i] : INT_MAX;
            int rank = 0;
            for (int j0 = 0; j0 < size; j0 += 1024) {
                int cnt = min(1024, size - j0);
                __syncthreads();
                for (int j = threadIdx.x; j < cnt; j += blockDim.x) tile[j] = perm[lo + j0 + j];
                __syncthreads();
                for (int j = 0; j < cnt; ++j) rank += tile[j] < v;
            }
            if (i < size) tmp[lo + rank] = v;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < size; i += blockDim.x) perm[lo + i] = tmp[lo + i];
        __syncthreads();
    }
}

// Reorders each cell's indices along a Morton curve over a 16 x 16 sub-grid of
// the cell (ties keep index order), so consecutive runs of 8 points are
// spatially compact -- or, with yorder, by y to 2^-20 of the cell height, so a
// column of stacked cells is ordered by y.  Cells larger than kSortCap keep
// index order (with yorder they, and points outside their cell, set *unsorted).
__global__ __launch_bounds__(256) void k_morton_cells(int* __restrict__ perm,
                                                      const int* __restrict__ start,
                                                      const double* __restrict__ p, double Lx,
                                                      double Ly, int nsx, int nsy, int yorder,
                                                      int* __restrict__ unsorted,
                                                      int* __restrict__ ykey, int big_sorted)
{
    __shared__ unsigned key[kSortCap];
    __shared__ int old[kSortCap];
    const int b = blockIdx.x;
    const int lo = start[b], size = start[b + 1] - lo;
    if (size > kSortCap && yorder && !big_sorted && threadIdx.x == 0) *unsorted = 1;
    // a single point is in order, but its y key is still wanted
    if (size > kSortCap || size == 0 || (size == 1 && !(yorder && ykey))) return;
    const int bx = b % nsx, by = b / nsx;
    int np = 2;
    while (np < size) np <<= 1;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        if (i < size) {
            const int o = perm[lo + i];
            old[i] = o;
            double fx = (p[2 * o] / Lx + 0.5) * nsx - bx;
            double fy = (p[2 * o + 1] / Ly + 0.5) * nsy - by;
            if (yorder) {
                // y inside the cell to 2^-20 of its height; a point outside its
                // cell (outside the box) leaves the cell's order unusable for
                // trimming by y
                if (!(fy >= 0.0 && fy < 1.0)) *unsorted = 1;
                const unsigned qy = (unsigned)min(1048575, max(0, (int)(fy * 1048576.0)));
                key[i] = (qy << 12) | (unsigned)i;
            } else {
                int qx = min(15, max(0, (int)(fx * 16))), qy = min(15, max(0, (int)(fy * 16)));
                unsigned code = 0;
                for (int k = 0; k < 4; ++k)
                    code |= (((qx >> k) & 1u) << (2 * k)) | (((qy >> k) & 1u) << (2 * k + 1));
                key[i] = (code << 12) | (unsigned)i;
            }
        } else {
            key[i] = 0xFFFFFFFFu;
        }
    }
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                int ij = i ^ j;
                if (ij > i) {
                    unsigned a = key[i], c = key[ij];
                    bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        key[i] = c;
                        key[ij] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < size; i += blockDim.x) {
        perm[lo + i] = old[key[i] & 0xFFFu];
        if (yorder && ykey) ykey[lo + i] = (int)(key[i] >> 12);
    }
}

// y order inside the cells larger than kSortCap (one workgroup per listed cell):
// the same order as k_morton_cells with yorder -- by the 20-bit in-cell y key,
// ties in index order -- by a stable LSD radix sort on 4-bit digits through
// global scratch (keys and indices, two buffers each, at the cell's offset).
__global__ __launch_bounds__(256) void k_ysort_big(int* __restrict__ perm,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ cells,
                                                   const double* __restrict__ p, double Ly,
                                                   int nsx, int nsy, int* __restrict__ unsorted,
                                                   int* __restrict__ ykey, int* __restrict__ kA,
                                                   int* __restrict__ vA, int* __restrict__ kB,
                                                   int* __restrict__ vB)
{
    __shared__ int hist[16], base[16], wcnt[4][16];
    const int b = cells[blockIdx.x];
    const int lo = start[b], size = start[b + 1] - lo;
    const int by = b / nsx;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i < size; i += 256) {
        const int o = perm[lo + i];
        const double fy = (p[2 * o + 1] / Ly + 0.5) * nsy - by;
        if (!(fy >= 0.0 && fy < 1.0)) *unsorted = 1;
        kA[lo + i] = min(1048575, max(0, (int)(fy * 1048576.0)));
        vA[lo + i] = o;
    }
    int *ks = kA, *vs = vA, *kd = kB, *vd = vB;
    for (int shift = 0; shift < 20; shift += 4) {
        __syncthreads();
        if (tid < 16) hist[tid] = 0;
        __syncthreads();
        for (int i = tid; i < size; i += 256) atomicAdd(&hist[(ks[lo + i] >> shift) & 15], 1);
        __syncthreads();
        if (tid == 0) {
            int acc = 0;
            for (int d = 0; d < 16; ++d) {
                base[d] = acc;
                acc += hist[d];
            }
        }
        __syncthreads();
        // chunks of 256 in order; inside a chunk, ranks by wave ballots per digit
        for (int c0 = 0; c0 < size; c0 += 256) {
            const int i = c0 + tid;
            const bool valid = i < size;
            const int kv = valid ? ks[lo + i] : 0, vv = valid ? vs[lo + i] : 0;
            const int dg = valid ? (kv >> shift) & 15 : 16;
            int rank = 0;
            for (int d = 0; d < 16; ++d) {
                const unsigned long long m = __ballot(dg == d);
                if (dg == d)
                    rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                if (lane == 0) wcnt[wv][d] = __popcll(m);
            }
            __syncthreads();
            if (valid) {
                int pos = base[dg] + rank;
                for (int w = 0; w < wv; ++w) pos += wcnt[w][dg];
                kd[lo + pos] = kv;
                vd[lo + pos] = vv;
            }
            __syncthreads();
            if (tid < 16) base[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        }
        int* t = ks;
        ks = kd;
        kd = t;
        t = vs;
        vs = vd;
        vd = t;
    }
    __syncthreads();
    for (int i = tid; i < size; i += 256) {
        perm[lo + i] = vs[lo + i];
        if (ykey) ykey[lo + i] = ks[lo + i];
    }
}

inline int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

}  // namespace

int launch_fill_int(hipStream_t s, int* p, int n, int v)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_closest_nodes(hipStream_t s, const double* p, int n, double Lx, double Ly,
                         double h, int P, int* mx, int* my, double* px, double* py)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_closest_nodes, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, Lx, Ly, h,
                       P, mx, my, px, py);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_cell_keys(hipStream_t s, const double* p, int n, double Lx, double Ly, int nsx,
                     int nsy, int* key, int* count)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, Lx, Ly, nsx,
                       nsy, key, count);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_tile_keys(hipStream_t s, const double* p, int n, double Lx, double Ly, double h,
                     int P, int Mx, int My, int TX, int TY, int ntx, int* key, int* count,
                     int2* node, double2* off)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_tile_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, Lx, Ly, h, P,
                       Mx, My, TX, TY, ntx, key, count, node, off);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_strip_count(hipStream_t s, const double* p, int n, double Lx, double Ly, double h,
                       int P, int Mx, int My, int ntx, int* cnt, int2* node, double2* off)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_strip_count, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, Lx, Ly, h, P,
                       Mx, My, ntx, cnt, node, off);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_strip_keys(hipStream_t s, const int2* node, const int* offs, int n, int P, int Mx,
                      int My, int ntx, int* key, int* count, int2* rnode, int* owner)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_strip_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, node, offs, n, P,
                       Mx, My, ntx, key, count, rnode, owner);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int scan_ints(se2p_ctx* ctx, Workspace* ws, const int* count, int n, int* start)
{
    hipStream_t s = ctx->stream;
    if (n <= 8 * kScanTile) {
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, count, n, start);
        SE2P_HIP_CHECK(hipGetLastError());
        return SE2P_OK;
    }
    const int nt = (n + kScanTile - 1) / kScanTile;
    void* tmp = nullptr;
    SE2P_TRY(ws_get(ctx, ws, WS_SCAN, sizeof(int) * (2 * (size_t)nt + 1), &tmp));
    int* tsum = (int*)tmp;
    int* toff = tsum + nt;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, s, count, n, tsum);
    SE2P_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, tsum, nt, toff);
    SE2P_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(256), 0, s, count, n, toff, start);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int sort_by_key(se2p_ctx* ctx, Workspace* ws, const int* key, int* count, int n, int nkeys,
                int* perm, int* start, Slot tmp_slot)
{
    hipStream_t s = ctx->stream;
    // count[] was filled by the key kernel; start = exclusive scan.
    SE2P_TRY(scan_ints(ctx, ws, count, nkeys, start));
    if (n <= 0) return SE2P_OK;
    // count[] is reused as the scatter cursor.
    hipLaunchKernelGGL(k_copy_int, dim3(grid_for(nkeys, 256)), dim3(256), 0, s, start, count,
                       nkeys);
    SE2P_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 256)), dim3(256), 0, s, key, n, count, perm);
    SE2P_HIP_CHECK(hipGetLastError());
    // biglist: [count, bucket ids...]; tmp: n ints.
    void* big = nullptr;
    void* tmp = nullptr;
    SE2P_TRY(ws_get(ctx, ws, WS_BIGLIST, sizeof(int) * (size_t)(nkeys + 1), &big));
    SE2P_TRY(ws_get(ctx, ws, tmp_slot, sizeof(int) * (size_t)n, &tmp));
    SE2P_TRY(launch_fill_int(s, (int*)big, 1, 0));
    hipLaunchKernelGGL(k_sort_buckets, dim3(nkeys), dim3(256), 0, s, perm, start, (int*)big);
    SE2P_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sort_big, dim3(64), dim3(256), 0, s, perm, start, (const int*)big,
                       (int*)tmp);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_morton_cells(hipStream_t s, int* perm, const int* start, int nbox, const double* p,
                        double Lx, double Ly, int nsx, int nsy, int yorder, int* unsorted,
                        int* ykey, int big_sorted)
{
    hipLaunchKernelGGL(k_morton_cells, dim3(nbox), dim3(256), 0, s, perm, start, p, Lx, Ly, nsx,
                       nsy, yorder, unsorted, ykey, big_sorted);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int launch_ysort_big(hipStream_t s, int* perm, const int* start, const int* cells, int ncells,
                     const double* p, double Ly, int nsx, int nsy, int* unsorted, int* ykey,
                     int* tmp, int n)
{
    if (ncells <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_ysort_big, dim3(ncells), dim3(256), 0, s, perm, start, cells, p, Ly, nsx,
                       nsy, unsorted, ykey, tmp, tmp + n, tmp + 2 * (size_t)n,
                       tmp + 3 * (size_t)n);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

}  // namespace se2p
