// Host side of libse2p.so: the C ABI, contexts, workspaces, stage timing, the
// mex-equivalent entry points, plans and the driver (parameter recipe,
// wrapping, zero mode).  Reference paths are relative to the reference
// repository root.
#include "se2p_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

// A captured apply for one (f, u) pair.
struct GraphCache {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double* f = nullptr;
    double* u = nullptr;
};

struct se2p_plan {
    se2p_ctx* ctx = nullptr;
    se2p::Workspace ws;
    se2p::Geom G;
    se2p::BatchExec B;
    GraphCache gc;
};

struct se2p_batch {
    se2p_ctx* ctx = nullptr;
    std::vector<se2p::Geom> G;
    std::vector<se2p::Workspace> ws;  // one per system
    se2p::Workspace bws;              // descriptors and job lists
    se2p::BatchExec B;
    long long nf = 0, nu = 0;
    GraphCache gc;
};

// Scratch memory of the allocation-free evaluation (se2p_eval_device_noalloc):
// one array per GPU, held until the process ends.  Host code never passes this
// name to a kernel or a copy: on the host it is only a stand-in for the symbol.
// The device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device at context creation with that device current.
static constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace se2p {

namespace {
thread_local char g_err[1024] = "";
}

int set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int ws_get(se2p_ctx* ctx, Workspace* ws, Slot slot, size_t bytes, void** out)
{
    if (bytes < 256) bytes = 256;
    if (ws->arena) {
        if (ws->size[slot] < bytes) {
            const size_t off = (ws->arena_used + 255) & ~(size_t)255;
            if (off > ws->arena_cap || bytes > ws->arena_cap - off)
                return set_error(SE2P_ERR_UNSUPPORTED,
                                 "scratch of %zu bytes exhausted (%zu more needed)",
                                 ws->arena_cap, off + bytes - ws->arena_cap);
            ws->ptr[slot] = ws->arena + off;
            ws->size[slot] = bytes;
            ws->arena_used = off + bytes;
        }
        *out = ws->ptr[slot];
        return SE2P_OK;
    }
    if (ws->size[slot] < bytes) {
        if (ctx->capturing)
            return set_error(SE2P_ERR_ARG, "allocation requested while a graph is captured");
        if (ws->ptr[slot]) {
            // The stream may still read the old buffer.
            SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            SE2P_HIP_CHECK(hipFree(ws->ptr[slot]));
            ws->ptr[slot] = nullptr;
            ws->size[slot] = 0;
        }
        size_t sz = bytes + bytes / 8;  // headroom for slowly growing inputs
        void* p = nullptr;
        SE2P_HIP_CHECK(hipMalloc(&p, sz));
        ws->ptr[slot] = p;
        ws->size[slot] = sz;
    }
    *out = ws->ptr[slot];
    return SE2P_OK;
}

void ws_release(Workspace* ws)
{
    for (int i = 0; i < WS_NUM_SLOTS; ++i) {
        if (ws->ptr[i] && !ws->arena) (void)hipFree(ws->ptr[i]);
        ws->ptr[i] = nullptr;
        ws->size[i] = 0;
    }
}

int prof_begin(se2p_ctx* ctx, hipStream_t s, const char* name, ProfRecord* rec)
{
    auto take = [&](hipEvent_t* e) -> int {
        if (!ctx->event_pool.empty()) {
            *e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
            return SE2P_OK;
        }
        SE2P_HIP_CHECK(hipEventCreate(e));
        return SE2P_OK;
    };
    rec->name = name;
    SE2P_TRY(take(&rec->start));
    SE2P_TRY(take(&rec->stop));
    SE2P_HIP_CHECK(hipEventRecord(rec->start, s));
    return SE2P_OK;
}

int prof_end(se2p_ctx* ctx, hipStream_t s, ProfRecord* rec)
{
    SE2P_HIP_CHECK(hipEventRecord(rec->stop, s));
    ctx->prof_pending.push_back(*rec);
    return SE2P_OK;
}

int prof_flush(se2p_ctx* ctx)
{
    if (ctx->prof_pending.empty()) return SE2P_OK;
    SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto& r : ctx->prof_pending) {
        float ms = 0;
        SE2P_HIP_CHECK(hipEventElapsedTime(&ms, r.start, r.stop));
        ProfAcc* a = nullptr;
        for (auto& x : ctx->prof_acc)
            if (x.name == r.name) a = &x;
        if (!a) {
            ctx->prof_acc.push_back(ProfAcc{r.name, 0.0, 0});
            a = &ctx->prof_acc.back();
        }
        a->ms += ms;
        a->count += 1;
        ctx->event_pool.push_back(r.start);
        ctx->event_pool.push_back(r.stop);
    }
    ctx->prof_pending.clear();
    return SE2P_OK;
}

namespace {

// ---------------------------------------------------------------------------
// Validation
// ---------------------------------------------------------------------------
bool valid_kind(int kind) { return kind >= 0 && kind < SE2P_NUM_KINDS; }

// Recipe of a kind's MATLAB driver: the SLP one for the Stokeslet quantities and
// the stresslet vorticity (StokesDLP_vorticity_ewald_2p.m:113-131 and its
// find_xi/find_kinfb at :170-224 are the SLP recipe), the DLP one otherwise.
int recipe_of(int kind)
{
    return dlp_family(kind) && kind != SE2P_DLP_VORTICITY ? SE2P_DLP : SE2P_SLP;
}

int check_common(se2p_ctx* ctx, int Ns, int Nt, double Lx, double Ly)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (Ns < 0 || Nt < 0) return set_error(SE2P_ERR_ARG, "negative point count");
    if (!(Lx > 0) || !(Ly > 0) || !std::isfinite(Lx) || !std::isfinite(Ly))
        return set_error(SE2P_ERR_ARG, "box lengths must be positive and finite");
    return SE2P_OK;
}

int check_real(int nsx, int nsy, double xi)
{
    if (nsx < 1 || nsy < 1) return set_error(SE2P_ERR_ARG, "nside_x and nside_y must be >= 1");
    if ((long)nsx * nsy > (1L << 26)) return set_error(SE2P_ERR_UNSUPPORTED, "too many cells");
    if (!(xi > 0) || !std::isfinite(xi)) return set_error(SE2P_ERR_ARG, "xi must be positive");
    return SE2P_OK;
}

int check_kspace(int Mx, int My, double Lx, double Ly, double xi, double eta, double w, int P)
{
    if (Mx < 2 || My < 2 || (Mx & 1) || (My & 1))
        return set_error(SE2P_ERR_ARG, "Mx and My must be even and >= 2 (got %d, %d)", Mx, My);
    if (std::fabs(Lx / Mx - Ly / My) > 1e-12 * (Lx / Mx))
        return set_error(SE2P_ERR_ARG, "grid spacing must agree: Lx/Mx = %.17g, Ly/My = %.17g",
                         Lx / Mx, Ly / My);
    if (P < 2 || (P & 1)) return set_error(SE2P_ERR_ARG, "P must be even and >= 2 (got %d)", P);
    if (!(xi > 0) || !(eta > 0) || !std::isfinite(w))
        return set_error(SE2P_ERR_ARG, "xi and eta must be positive");
    if (Mx > kFftMaxLen || My > kFftMaxLen)
        return set_error(SE2P_ERR_UNSUPPORTED, "grid %d x %d beyond %d per side", Mx, My,
                         kFftMaxLen);
    return SE2P_OK;
}

int check_params(const se2p_params* prm, double Lx, double Ly)
{
    if (!prm) return set_error(SE2P_ERR_ARG, "null parameters");
    SE2P_TRY(check_real(prm->nside_x, prm->nside_y, prm->xi));
    SE2P_TRY(check_kspace(prm->Mx, prm->My, Lx, Ly, prm->xi, prm->eta, prm->w, prm->P));
    if (prm->e1_tier != 0 && prm->e1_tier != 1)
        return set_error(SE2P_ERR_ARG, "e1_tier must be 0 or 1");
    if (prm->symmetric != 0 && prm->symmetric != 1)
        return set_error(SE2P_ERR_ARG, "symmetric must be 0 or 1");
    if (prm->precision != SE2P_FP64 && prm->precision != SE2P_FP32C)
        return set_error(SE2P_ERR_ARG, "precision must be SE2P_FP64 or SE2P_FP32C");
    if (!(prm->far32 >= 0)) return set_error(SE2P_ERR_ARG, "far32 must be >= 0");
    if (!(prm->free_R >= 0) || !std::isfinite(prm->free_R))
        return set_error(SE2P_ERR_ARG, "free_R must be finite and >= 0");
    return SE2P_OK;
}

// Copies host inputs into a workspace slot of the context.
int upload(se2p_ctx* ctx, Slot slot, const double* h, size_t n, double** d)
{
    void* p = nullptr;
    SE2P_TRY(ws_get(ctx, &ctx->ws, slot, sizeof(double) * n, &p));
    if (n)
        SE2P_HIP_CHECK(hipMemcpyAsync(p, h, sizeof(double) * n, hipMemcpyHostToDevice,
                                      ctx->stream));
    *d = (double*)p;
    return SE2P_OK;
}

int download(se2p_ctx* ctx, double* h, const double* d, size_t n)
{
    if (n)
        SE2P_HIP_CHECK(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost,
                                      ctx->stream));
    SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SE2P_OK;
}

int need_ptr(const void* p, int n, const char* what)
{
    if (n > 0 && !p) return set_error(SE2P_ERR_ARG, "%s is NULL", what);
    return SE2P_OK;
}

// Parameters carrying only what one half of a mex call needs.
se2p_params real_only(double xi, int nsx, int nsy)
{
    se2p_params p{};
    p.xi = xi;
    p.nside_x = nsx;
    p.nside_y = nsy;
    return p;
}

se2p_params k_only(double xi, double eta, int Mx, int My, double w, int P)
{
    se2p_params p{};
    p.xi = xi;
    p.eta = eta;
    p.Mx = Mx;
    p.My = My;
    p.w = w;
    p.P = P;
    return p;
}

// ---------------------------------------------------------------------------
// MATLAB rat() with its default tolerance 1e-6 |X|: continued-fraction
// convergents of X.
void rat(double X, int* N, int* D)
{
    const double tol = 1e-6 * std::fabs(X);
    double nn = std::round(X), dd = 1, lastn = 1, lastd = 0;
    double frac = X - nn;
    int guard = 0;
    while (std::fabs(X - nn / dd) >= tol && guard++ < 64) {
        const double flip = 1 / frac;
        const double step = std::round(flip);
        frac = flip - step;
        const double sn = nn, sd = dd;
        nn = nn * step + lastn;
        dd = dd * step + lastd;
        lastn = sn;
        lastd = sd;
    }
    *N = (int)(nn * (dd < 0 ? -1 : 1));
    *D = (int)std::fabs(dd);
}

// Newton solve with the drivers' step rule: step limited to 0.9 x, at most 101
// steps, stop once |f| <= tol.  which = 0: xi from rc; which = 1: k_inf from xi.
double newton(int kind, int which, double x, double Q, double Lx, double Ly, double rc, double xi,
              double tol, bool* ok)
{
    *ok = true;
    for (int it = 0;; ++it) {
        double fv, fp;
        if (which == 0) {
            const double g = std::exp(-x * x * rc * rc);
            if (kind == SE2P_SLP) {
                fv = std::sqrt(Q * kPi / (4 * Lx * Ly * x)) * g - tol;
                fp = -std::sqrt(Q * kPi / (4 * Lx * Ly)) * g *
                     (0.5 * std::pow(x, -1.5) + 2 * rc * rc * x * std::sqrt(1 / x));
            } else {
                const double c = std::sqrt(2 * kPi * Q / (Lx * Ly));
                fv = g * x * rc * c - tol;
                fp = (1 - 2 * x * x * rc * rc) * g * c * rc;
            }
        } else {
            const double Lm = Lx > Ly ? Lx : Ly;
            const double g = std::exp(-x * x / (4 * xi * xi));
            if (kind == SE2P_SLP) {
                const double c = 4 * Q * kPi * Lm / (std::pow(Lx, 3) * std::pow(Ly, 3));
                fv = std::sqrt(c / x) * g - tol;
                fp = -std::sqrt(c) * g *
                     (0.5 * std::pow(x, -1.5) + std::sqrt(1 / x) * 2 * x / (4 * xi * xi));
            } else {
                // StokesDLP_ewald_2p.m:177: MATLAB evaluates .../Lx^3*Ly^3 left to right.
                const double c = 8 * kPi * Q * Lm / std::pow(Lx, 3) * std::pow(Ly, 3);
                fv = std::sqrt(x * c) * g - tol;
                fp = std::sqrt(c) * g * (0.5 * std::pow(x, -0.5) - 2 * std::pow(x, 1.5) / (4 * xi * xi));
            }
        }
        if (!(std::fabs(fv) > tol)) break;
        if (it > 100) {
            *ok = false;
            break;
        }
        double d = fv / fp;
        if (!(d < 0.9 * x)) d = 0.9 * x;
        x -= d;
    }
    return x;
}

double default_nb(double npts, int P)
{
    // min(P*round(log2(npts)), npts/4), MATLAB round = half away from zero
    const double a = P * std::round(std::log2(npts));
    const double b = npts / 4;
    return a < b ? a : b;
}

int params_ref(int kind, int Ns, int Nt, double Q, double Lx, double Ly, const se2p_options* o,
               se2p_params* out)
{
    se2p_options d;
    se2p_options_default(&d);
    if (!o) o = &d;
    const int P = o->P;
    if (P < 2 || (P & 1)) return set_error(SE2P_ERR_ARG, "P must be even and >= 2");
    if (!(o->tol > 0)) return set_error(SE2P_ERR_ARG, "tol must be positive");
    const double npts0 = (double)Ns + Nt;
    if (npts0 < 1) return set_error(SE2P_ERR_ARG, "no points");
    const double Nb = o->Nb > 0 ? o->Nb : default_nb(npts0, P);
    // The DLP driver recomputes npts from the 2 x N arrays (StokesDLP_ewald_2p.m:107).
    const double npts = kind == SE2P_SLP ? npts0 : (double)(Ns > 2 ? Ns : 2) + (Nt > 2 ? Nt : 2);
    int A, B;
    rat(Lx / Ly, &A, &B);
    if (A <= 0 || B <= 0) return set_error(SE2P_ERR_ARG, "rat(Lx/Ly) failed");
    const double a = std::ceil(std::sqrt(npts / (Nb * A * B)));
    const double m = 0.95 * std::sqrt(kPi * P);
    const int nsx = (int)(a * A), nsy = (int)(a * B);
    const double rc = Lx / nsx;
    bool ok1, ok2;
    const double xi = newton(kind, 0, kind == SE2P_SLP ? 4 / rc : 1 / rc, Q, Lx, Ly, rc, 0,
                             o->tol, &ok1);
    // Both drivers pass Lx for Ly here (StokesSLP_ewald_2p.m:96).
    const double kinf = newton(kind, 1, std::round(5 * xi), Q, Lx, Lx, rc, xi, o->tol, &ok2);
    const int kx = (int)std::round(Lx * kinf / (2 * kPi));
    const int M0 = 2 * kx < 10000 ? 2 * kx : 10000;
    *out = se2p_params{};  // every field this recipe does not set takes its default (0)
    out->P = P;
    out->nside_x = nsx;
    out->nside_y = nsy;
    out->rc = rc;
    out->xi = xi;
    out->kinf = kx;
    out->Mx = A * M0;
    out->My = B * M0;
    out->w = P * Lx / out->Mx / 2;
    out->eta = std::pow(2 * xi * out->w / m, 2);
    out->converged = ok1 && ok2;
    out->e1_tier = 0;  // the reference evaluates E1 to full FP64 accuracy
    out->symmetric = 0;
    out->precision = SE2P_FP64;
    if (o->verbose)
        fprintf(stderr,
                "se2p %s: xi %.6g rc %.6g kinf %d Mx %d My %d nside %d x %d w %.6g eta %.6g\n",
                kind == SE2P_SLP ? "SLP" : "DLP", xi, rc, kx, out->Mx, out->My, nsx, nsy, out->w,
                out->eta);
    return SE2P_OK;
}

// mod(x + L/2, L) - L/2 with MATLAB's mod (StokesSLP_ewald_2p.m:75-78).
double wrap1(double x, double L)
{
#pragma clang fp contract(off)
    double v = x + L / 2;
    v = v - std::floor(v / L) * L;
    return v - L / 2;
}

// Q = sum(sum(f.^2)) + 1 with f = [f1'; f2'] (2 x N, StokesSLP_ewald_2p.m:82,87):
// sum over a matrix adds down each column, f1(i)^2 + f2(i)^2, and the outer sum
// adds those N column sums in index order.
double density_q(const double* f1, const double* f2, int n)
{
#pragma clang fp contract(off)
    double q = 0;
    for (int i = 0; i < n; ++i) q += f1[i] * f1[i] + f2[i] * f2[i];
    return q + 1;
}

int ewald_driver(se2p_ctx* ctx, int kind, const double* xsrc, const double* ysrc,
                 const double* xtar, const double* ytar, const double* n1, const double* n2,
                 const double* f1, const double* f2, int Ns, int Nt, double Lx, double Ly,
                 const se2p_options* opt, double* u1, double* u2, double* ur_out,
                 double* uk_out, double* xi_out)
{
    SE2P_TRY(check_common(ctx, Ns, Nt, Lx, Ly));
    SE2P_TRY(need_ptr(xsrc, Ns, "xsrc"));
    SE2P_TRY(need_ptr(ysrc, Ns, "ysrc"));
    SE2P_TRY(need_ptr(f1, Ns, "f1"));
    SE2P_TRY(need_ptr(f2, Ns, "f2"));
    SE2P_TRY(need_ptr(xtar, Nt, "xtar"));
    SE2P_TRY(need_ptr(ytar, Nt, "ytar"));
    SE2P_TRY(need_ptr(u1, Nt, "u1"));
    SE2P_TRY(need_ptr(u2, Nt, "u2"));
    if (kind == SE2P_DLP) {
        SE2P_TRY(need_ptr(n1, Ns, "n1"));
        SE2P_TRY(need_ptr(n2, Ns, "n2"));
    }
    std::vector<double> ps(2 * (size_t)Ns), pt(2 * (size_t)Nt), f(2 * (size_t)Ns), nn;
    for (int i = 0; i < Ns; ++i) {
        ps[2 * i] = wrap1(xsrc[i], Lx);
        ps[2 * i + 1] = wrap1(ysrc[i], Ly);
        f[2 * i] = f1[i];
        f[2 * i + 1] = f2[i];
    }
    const double Q = density_q(f1, f2, Ns);
    for (int i = 0; i < Nt; ++i) {
        pt[2 * i] = wrap1(xtar[i], Lx);
        pt[2 * i + 1] = wrap1(ytar[i], Ly);
    }
    if (kind == SE2P_DLP) {
        nn.resize(2 * (size_t)Ns);
        for (int i = 0; i < Ns; ++i) {
            nn[2 * i] = n1[i];
            nn[2 * i + 1] = n2[i];
        }
    }
    se2p_params prm;
    SE2P_TRY(params_ref(kind, Ns, Nt, Q, Lx, Ly, opt, &prm));
    prm.precision = ctx->precision;
    SE2P_TRY(check_params(&prm, Lx, Ly));
    if (xi_out) *xi_out = prm.xi;
    if (Nt == 0) return SE2P_OK;

    double *dps, *dpt, *df, *dn = nullptr;
    SE2P_TRY(upload(ctx, WS_PTS_A, ps.data(), ps.size(), &dps));
    SE2P_TRY(upload(ctx, WS_PTS_B, pt.data(), pt.size(), &dpt));
    SE2P_TRY(upload(ctx, WS_DENS_A, f.data(), f.size(), &df));
    if (kind == SE2P_DLP) SE2P_TRY(upload(ctx, WS_DENS_B, nn.data(), nn.size(), &dn));
    void* out = nullptr;
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, sizeof(double) * 4 * (size_t)Nt, &out));
    double* dur = (double*)out;
    double* duk = dur + 2 * (size_t)Nt;
    Geom G;
    SE2P_TRY(geom_build(ctx, &ctx->ws, &G, kind, dps, dpt, dn, Ns, Nt, Lx, Ly, prm, true, true));
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ctx->ws, &B, &gp, 1));
    SE2P_TRY(batch_apply_parts(ctx, B, df, dur, duk, false, kind == SE2P_DLP));
    std::vector<double> h(4 * (size_t)Nt);
    SE2P_TRY(download(ctx, h.data(), dur, 4 * (size_t)Nt));
    for (int i = 0; i < Nt; ++i) {
        const double r1 = h[2 * i], r2 = h[2 * i + 1];
        const double k1 = h[2 * (size_t)Nt + 2 * i], k2 = h[2 * (size_t)Nt + 2 * i + 1];
        u1[i] = r1 + k1;
        u2[i] = r2 + k2;
        if (ur_out) {
            ur_out[2 * i] = r1;
            ur_out[2 * i + 1] = r2;
        }
        if (uk_out) {
            uk_out[2 * i] = k1;
            uk_out[2 * i + 1] = k2;
        }
    }
    return SE2P_OK;
}

// StokesSLP_pressure_ewald_2p.m:1-138 / StokesDLP_pressure_ewald_2p.m: wrap, the
// SLP or DLP recipe, pr + pk with the zero mode added to pk (SLP: -sum(f1 x +
// f2 y)/(2 Lx Ly), :127-128; DLP: -sum(n1 f1 + n2 f2)/(2 Lx Ly), :156-157).
// Hash of a coordinate pair; -0.0 is folded into +0.0 by the callers (+ 0.0).
struct PairHash {
    size_t operator()(const std::pair<double, double>& p) const
    {
        return std::hash<double>()(p.first) * 1000003u ^ std::hash<double>()(p.second);
    }
};

int pressure_driver(se2p_ctx* ctx, int kind, const double* xsrc, const double* ysrc,
                    const double* xtar, const double* ytar, const double* n1,
                    const double* n2, const double* f1, const double* f2, int Ns, int Nt,
                    double Lx, double Ly, const se2p_options* opt, double* p_out,
                    double* pr_out, double* pk_out, double* xi_out)
{
    SE2P_TRY(check_common(ctx, Ns, Nt, Lx, Ly));
    SE2P_TRY(need_ptr(xsrc, Ns, "xsrc"));
    SE2P_TRY(need_ptr(ysrc, Ns, "ysrc"));
    SE2P_TRY(need_ptr(f1, Ns, "f1"));
    SE2P_TRY(need_ptr(f2, Ns, "f2"));
    SE2P_TRY(need_ptr(xtar, Nt, "xtar"));
    SE2P_TRY(need_ptr(ytar, Nt, "ytar"));
    SE2P_TRY(need_ptr(p_out, Nt, "p"));
    if (kind == SE2P_DLP_PRESSURE) {
        SE2P_TRY(need_ptr(n1, Ns, "n1"));
        SE2P_TRY(need_ptr(n2, Ns, "n2"));
    }
    std::vector<double> ps(2 * (size_t)Ns), pt(2 * (size_t)Nt), f(2 * (size_t)Ns), nn;
    if (kind == SE2P_DLP_PRESSURE) {
        nn.resize(2 * (size_t)Ns);
        for (int i = 0; i < Ns; ++i) {
            nn[2 * i] = n1[i];
            nn[2 * i + 1] = n2[i];
        }
    }
    for (int i = 0; i < Ns; ++i) {
        ps[2 * i] = wrap1(xsrc[i], Lx);
        ps[2 * i + 1] = wrap1(ysrc[i], Ly);
        f[2 * i] = f1[i];
        f[2 * i + 1] = f2[i];
    }
    for (int i = 0; i < Nt; ++i) {
        pt[2 * i] = wrap1(xtar[i], Lx);
        pt[2 * i + 1] = wrap1(ytar[i], Ly);
    }
    se2p_params prm;
    SE2P_TRY(params_ref(kind == SE2P_DLP_PRESSURE ? SE2P_DLP : SE2P_SLP, Ns, Nt,
                        density_q(f1, f2, Ns), Lx, Ly, opt, &prm));
    SE2P_TRY(check_params(&prm, Lx, Ly));
    if (xi_out) *xi_out = prm.xi;
    if (Nt == 0) return SE2P_OK;
    double *dps, *dpt, *df, *dn = nullptr;
    SE2P_TRY(upload(ctx, WS_PTS_A, ps.data(), ps.size(), &dps));
    SE2P_TRY(upload(ctx, WS_PTS_B, pt.data(), pt.size(), &dpt));
    SE2P_TRY(upload(ctx, WS_DENS_A, f.data(), f.size(), &df));
    if (kind == SE2P_DLP_PRESSURE) SE2P_TRY(upload(ctx, WS_DENS_B, nn.data(), nn.size(), &dn));
    void* out = nullptr;
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, sizeof(double) * 2 * (size_t)Nt, &out));
    double* dpr = (double*)out;
    double* dpk = dpr + Nt;
    Geom G;
    SE2P_TRY(geom_build(ctx, &ctx->ws, &G, kind, dps, dpt, dn, Ns, Nt, Lx, Ly, prm, true, true));
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ctx->ws, &B, &gp, 1));
    SE2P_TRY(batch_apply_parts(ctx, B, df, dpr, dpk, false, true));
    std::vector<double> h(2 * (size_t)Nt);
    SE2P_TRY(download(ctx, h.data(), dpr, 2 * (size_t)Nt));
    for (int i = 0; i < Nt; ++i) {
        p_out[i] = h[i] + h[Nt + i];
        if (pr_out) pr_out[i] = h[i];
        if (pk_out) pk_out[i] = h[Nt + i];
    }
    if (kind == SE2P_DLP_PRESSURE) {
        // StokesDLP_pressure_ewald_2p.m:64-70,165-172: a source that is also a
        // target (same raw coordinates) adds xi^2 (f.n)/(2 pi) to p there.  The
        // reference adds the k-th coincident source's term to the k-th target,
        // the coincident one when the targets are the sources.
        std::unordered_map<std::pair<double, double>, std::vector<int>, PairHash> where;
        for (int j = 0; j < Nt; ++j) where[{xtar[j] + 0.0, ytar[j] + 0.0}].push_back(j);
        const double c = prm.xi * prm.xi / (2 * kPi);
        for (int i = 0; i < Ns; ++i) {
            auto it = where.find({xsrc[i] + 0.0, ysrc[i] + 0.0});
            if (it == where.end()) continue;
            for (int j : it->second) p_out[j] += c * (f1[i] * n1[i] + f2[i] * n2[i]);
        }
    }
    return SE2P_OK;
}

// One mex call: upload, build the half that is needed, apply, download.
int mex_call(se2p_ctx* ctx, int kind, bool real, const double* psrc, const double* ptar,
             const double* f, const double* n, int Ns, int Nt, double Lx, double Ly,
             const se2p_params& prm_in, double* out)
{
    se2p_params prm = prm_in;
    prm.precision = ctx->precision;
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (Nt == 0) return SE2P_OK;
    double *dps, *dpt, *df, *dn = nullptr;
    void* du;
    SE2P_TRY(upload(ctx, WS_PTS_A, psrc, 2 * (size_t)Ns, &dps));
    SE2P_TRY(upload(ctx, WS_PTS_B, ptar, 2 * (size_t)Nt, &dpt));
    SE2P_TRY(upload(ctx, WS_DENS_A, f, 2 * (size_t)Ns, &df));
    if (dlp_family(kind)) SE2P_TRY(upload(ctx, WS_DENS_B, n, 2 * (size_t)Ns, &dn));
    const size_t nout = (size_t)out_comps(kind) * Nt;
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, sizeof(double) * nout, &du));
    Geom G;
    SE2P_TRY(geom_build(ctx, &ctx->ws, &G, kind, dps, dpt, dn, Ns, Nt, Lx, Ly, prm, real, !real));
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ctx->ws, &B, &gp, 1));
    if (real)
        SE2P_TRY(batch_apply_real(ctx, B, df, (double*)du));
    else
        SE2P_TRY(batch_apply_k(ctx, B, df, (double*)du, false, false));
    return download(ctx, out, (double*)du, nout);
}

void graph_release(GraphCache* gc)
{
    if (gc->exec) (void)hipGraphExecDestroy(gc->exec);
    if (gc->graph) (void)hipGraphDestroy(gc->graph);
    *gc = GraphCache{};
}

// batch_apply replayed from a graph captured on the first call for (d_f, d_u).
int graph_apply(se2p_ctx* ctx, GraphCache* gc, const BatchExec& B, const double* d_f,
                double* d_u)
{
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (!gc->exec || gc->f != d_f || gc->u != d_u) {
        graph_release(gc);
        SE2P_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
        int rc = batch_apply(ctx, B, d_f, d_u);
        ctx->capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc != SE2P_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return set_error(SE2P_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        gc->graph = g;
        SE2P_HIP_CHECK(hipGraphInstantiate(&gc->exec, g, nullptr, nullptr, 0));
        gc->f = d_f;
        gc->u = d_u;
    }
    SE2P_HIP_CHECK(hipGraphLaunch(gc->exec, ctx->stream));
    return SE2P_OK;
}

int derived_real(se2p_ctx* ctx, int kind, const double* psrc, const double* ptar,
                 const double* f, const double* n, int Nsrc, int Ntar, double xi, int nside_x,
                 int nside_y, double Lx, double Ly, double* out)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_real(nside_x, nside_y, xi));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    if (dlp_family(kind)) SE2P_TRY(need_ptr(n, Nsrc, "n"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(out, Ntar, "out"));
    return mex_call(ctx, kind, true, psrc, ptar, f, n, Nsrc, Ntar, Lx, Ly,
                    real_only(xi, nside_x, nside_y), out);
}

int derived_kspace(se2p_ctx* ctx, int kind, const double* psrc, const double* ptar,
                   const double* f, const double* n, double xi, double eta, int Nsrc, int Ntar,
                   int Mx, int My, double Lx, double Ly, double w, int P, double* out)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_kspace(Mx, My, Lx, Ly, xi, eta, w, P));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    if (dlp_family(kind)) SE2P_TRY(need_ptr(n, Nsrc, "n"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(out, Ntar, "out"));
    return mex_call(ctx, kind, false, psrc, ptar, f, n, Nsrc, Ntar, Lx, Ly,
                    k_only(xi, eta, Mx, My, w, P), out);
}

int free_eval(se2p_ctx* ctx, bool dlp, const double* f1, const double* f2, const double* xs,
              const double* ys, const double* n1, const double* n2, int Ns, const double* xt,
              const double* yt, int Nt, bool self, double* u1, double* u2)
{
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (Nt == 0) return SE2P_OK;
    const int rec = dlp ? 6 : 4;
    std::vector<double> hs((size_t)rec * std::max(Ns, 1)), ht(2 * (size_t)Nt);
    for (int i = 0; i < Ns; ++i) {
        double* r = hs.data() + (size_t)rec * i;
        r[0] = xs[i];
        r[1] = ys[i];
        r[2] = f1[i];
        r[3] = f2[i];
        if (dlp) {
            r[4] = n1[i];
            r[5] = n2[i];
        }
    }
    for (int i = 0; i < Nt; ++i) {
        ht[2 * i] = xt[i];
        ht[2 * i + 1] = yt[i];
    }
    double *ds, *dt;
    SE2P_TRY(upload(ctx, WS_PTS_A, hs.data(), (size_t)rec * Ns, &ds));
    SE2P_TRY(upload(ctx, WS_PTS_B, ht.data(), ht.size(), &dt));
    const int nslice = free_slices(Ns, Nt);
    void *out, *part = nullptr;
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, sizeof(double) * 2 * (size_t)Nt, &out));
    if (nslice > 1)
        SE2P_TRY(ws_get(ctx, &ctx->ws, WS_FREE_PART, sizeof(double) * 2 * (size_t)Nt * nslice,
                        &part));
    {
        StageTimer st(ctx, "free");
        SE2P_TRY(launch_free(ctx->stream, dlp, ds, Ns, dt, Nt, self, nslice, (double*)part,
                             (double*)out));
    }
    std::vector<double> h(2 * (size_t)Nt);
    SE2P_TRY(download(ctx, h.data(), (double*)out, h.size()));
    for (int i = 0; i < Nt; ++i) {
        u1[i] = h[2 * i];
        u2[i] = h[2 * i + 1];
    }
    return SE2P_OK;
}

void destroy_streams(se2p_ctx* c)
{
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

}  // namespace
}  // namespace se2p

using namespace se2p;

namespace se2p {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

// This GPU's scratch address, resolved on the first call for the device (the
// caller has made the device current).
int scratch_address(int device, char** out)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(SE2P_ERR_UNSUPPORTED, "device %d: scratch table holds %d devices", device,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(SE2P_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s",
                             device, hipGetErrorString(e));
        g_scratch[device] = static_cast<char*>(p);
    }
    *out = g_scratch[device];
    return SE2P_OK;
}
}  // namespace
}  // namespace se2p

extern "C" {

int se2p_version(void) { return 2; }

const char* se2p_last_error(void) { return se2p::g_err; }

int se2p_device_count(int* count)
{
    if (!count) return set_error(SE2P_ERR_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return SE2P_OK;
}


int se2p_ctx_create(int device, se2p_ctx** out)
{
    if (!out) return set_error(SE2P_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_error(SE2P_ERR_NODEV, "no HIP device available");
    if (device < 0 || device >= n)
        return set_error(SE2P_ERR_ARG, "device %d outside [0, %d)", device, n);
    SE2P_HIP_CHECK(hipSetDevice(device));
    se2p_ctx* c = new (std::nothrow) se2p_ctx();
    if (!c) return set_error(SE2P_ERR_NOMEM, "context allocation failed");
    c->device = device;
    // both streams at the default priority
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        destroy_streams(c);
        delete c;
        return set_error(SE2P_ERR_HIP, "stream set-up: %s", hipGetErrorString(e));
    }
    int rc = scratch_address(device, &c->scratch);
    if (rc != SE2P_OK) {
        destroy_streams(c);
        delete c;
        return rc;
    }
    *out = c;
    return SE2P_OK;
}

int se2p_ctx_destroy(se2p_ctx* ctx)
{
    if (!ctx) return SE2P_OK;
    // Teardown is best effort: every resource is released even if one call fails.
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ws_release(&ctx->ws);
    for (void* p : ctx->user_allocs) (void)hipFree(p);
    for (auto& r : ctx->prof_pending) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    destroy_streams(ctx);
    delete ctx;
    return SE2P_OK;
}

int se2p_slp_real(se2p_ctx* ctx, const double* psrc, const double* ptar, const double* f,
                  int Nsrc, int Ntar, double xi, int nside_x, int nside_y, double Lx, double Ly,
                  double* ur)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_real(nside_x, nside_y, xi));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(ur, Ntar, "ur"));
    return mex_call(ctx, SE2P_SLP, true, psrc, ptar, f, nullptr, Nsrc, Ntar, Lx, Ly,
                    real_only(xi, nside_x, nside_y), ur);
}

int se2p_dlp_real(se2p_ctx* ctx, const double* psrc, const double* ptar, const double* f,
                  const double* n, int Nsrc, int Ntar, double xi, int nside_x, int nside_y,
                  double Lx, double Ly, double* ur)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_real(nside_x, nside_y, xi));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(n, Nsrc, "n"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(ur, Ntar, "ur"));
    return mex_call(ctx, SE2P_DLP, true, psrc, ptar, f, n, Nsrc, Ntar, Lx, Ly,
                    real_only(xi, nside_x, nside_y), ur);
}

int se2p_slp_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                    double eta, const double* f, int Nsrc, int Ntar, int Mx, int My, double Lx,
                    double Ly, double w, int P, double* uk)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_kspace(Mx, My, Lx, Ly, xi, eta, w, P));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(uk, Ntar, "uk"));
    return mex_call(ctx, SE2P_SLP, false, psrc, ptar, f, nullptr, Nsrc, Ntar, Lx, Ly,
                    k_only(xi, eta, Mx, My, w, P), uk);
}

int se2p_dlp_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                    double eta, const double* f, const double* n, int Nsrc, int Ntar, int Mx,
                    int My, double Lx, double Ly, double w, int P, double* uk)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_kspace(Mx, My, Lx, Ly, xi, eta, w, P));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(n, Nsrc, "n"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(uk, Ntar, "uk"));
    return mex_call(ctx, SE2P_DLP, false, psrc, ptar, f, n, Nsrc, Ntar, Lx, Ly,
                    k_only(xi, eta, Mx, My, w, P), uk);
}

int se2p_slp_pressure_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                           const double* f, int Nsrc, int Ntar, double xi, int nside_x,
                           int nside_y, double Lx, double Ly, double* pr)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_real(nside_x, nside_y, xi));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(pr, Ntar, "pr"));
    return mex_call(ctx, SE2P_SLP_PRESSURE, true, psrc, ptar, f, nullptr, Nsrc, Ntar, Lx, Ly,
                    real_only(xi, nside_x, nside_y), pr);
}

int se2p_slp_pressure_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                             const double* f, double xi, double eta, int Nsrc, int Ntar, int Mx,
                             int My, double Lx, double Ly, double w, int P, double* pk)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_kspace(Mx, My, Lx, Ly, xi, eta, w, P));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(pk, Ntar, "pk"));
    return mex_call(ctx, SE2P_SLP_PRESSURE, false, psrc, ptar, f, nullptr, Nsrc, Ntar, Lx, Ly,
                    k_only(xi, eta, Mx, My, w, P), pk);
}

int se2p_slp_pressure_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc,
                            const double* xtar, const double* ytar, const double* f1,
                            const double* f2, int Nsrc, int Ntar, double Lx, double Ly,
                            const se2p_options* opt, double* p, double* pr, double* pk,
                            double* xi)
{
    if (ctx) SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    return pressure_driver(ctx, SE2P_SLP_PRESSURE, xsrc, ysrc, xtar, ytar, nullptr, nullptr, f1,
                           f2, Nsrc, Ntar, Lx, Ly, opt, p, pr, pk, xi);
}

int se2p_dlp_pressure_real(se2p_ctx* ctx, const double* psrc, const double* ptar,
                           const double* f, const double* n, int Nsrc, int Ntar, double xi,
                           int nside_x, int nside_y, double Lx, double Ly, double* pr)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_real(nside_x, nside_y, xi));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(n, Nsrc, "n"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(pr, Ntar, "pr"));
    return mex_call(ctx, SE2P_DLP_PRESSURE, true, psrc, ptar, f, n, Nsrc, Ntar, Lx, Ly,
                    real_only(xi, nside_x, nside_y), pr);
}

int se2p_dlp_pressure_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                             const double* f, const double* n, double xi, double eta, int Nsrc,
                             int Ntar, int Mx, int My, double Lx, double Ly, double w, int P,
                             double* pk)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    SE2P_TRY(check_kspace(Mx, My, Lx, Ly, xi, eta, w, P));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    SE2P_TRY(need_ptr(n, Nsrc, "n"));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(pk, Ntar, "pk"));
    return mex_call(ctx, SE2P_DLP_PRESSURE, false, psrc, ptar, f, n, Nsrc, Ntar, Lx, Ly,
                    k_only(xi, eta, Mx, My, w, P), pk);
}

// ---------------------------------------------------------------------------
// Derived quantities: velocity gradient, pressure gradient, vorticity, stress
// ---------------------------------------------------------------------------

#define SE2P_SLP_DERIVED(NAME, KIND)                                                           \
    int se2p_slp_##NAME##_real(se2p_ctx* ctx, const double* psrc, const double* ptar,        \
                               const double* f, int Nsrc, int Ntar, double xi, int nside_x,  \
                               int nside_y, double Lx, double Ly, double* out)               \
    {                                                                                        \
        return derived_real(ctx, KIND, psrc, ptar, f, nullptr, Nsrc, Ntar, xi, nside_x, \
                                  nside_y, Lx, Ly, out);                                     \
    }
#define SE2P_DLP_DERIVED(NAME, KIND)                                                           \
    int se2p_dlp_##NAME##_real(se2p_ctx* ctx, const double* psrc, const double* ptar,        \
                               const double* f, const double* n, int Nsrc, int Ntar,         \
                               double xi, int nside_x, int nside_y, double Lx, double Ly,    \
                               double* out)                                                  \
    {                                                                                        \
        return derived_real(ctx, KIND, psrc, ptar, f, n, Nsrc, Ntar, xi, nside_x,      \
                                  nside_y, Lx, Ly, out);                                     \
    }
SE2P_SLP_DERIVED(gradient, SE2P_SLP_GRADIENT)
SE2P_SLP_DERIVED(pressure_grad, SE2P_SLP_PRESSURE_GRAD)
SE2P_SLP_DERIVED(vorticity, SE2P_SLP_VORTICITY)
SE2P_SLP_DERIVED(stress, SE2P_SLP_STRESS)
SE2P_DLP_DERIVED(gradient, SE2P_DLP_GRADIENT)
SE2P_DLP_DERIVED(pressure_grad, SE2P_DLP_PRESSURE_GRAD)
SE2P_DLP_DERIVED(vorticity, SE2P_DLP_VORTICITY)
SE2P_DLP_DERIVED(stress, SE2P_DLP_STRESS)
#undef SE2P_SLP_DERIVED
#undef SE2P_DLP_DERIVED

// k-space parts, each with its mex function's argument order
int se2p_slp_gradient_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                             double eta, const double* f, int Nsrc, int Ntar, int Mx, int My,
                             double Lx, double Ly, double w, int P, double* out)
{
    return derived_kspace(ctx, SE2P_SLP_GRADIENT, psrc, ptar, f, nullptr, xi, eta, Nsrc,
                                Ntar, Mx, My, Lx, Ly, w, P, out);
}
int se2p_slp_pressure_grad_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                                  const double* f, double xi, double eta, int Nsrc, int Ntar,
                                  int Mx, int My, double Lx, double Ly, double w, int P,
                                  double* out)
{
    return derived_kspace(ctx, SE2P_SLP_PRESSURE_GRAD, psrc, ptar, f, nullptr, xi, eta,
                                Nsrc, Ntar, Mx, My, Lx, Ly, w, P, out);
}
int se2p_slp_vorticity_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                              double eta, const double* f, int Nsrc, int Ntar, int Mx, int My,
                              double Lx, double Ly, double w, int P, double* out)
{
    return derived_kspace(ctx, SE2P_SLP_VORTICITY, psrc, ptar, f, nullptr, xi, eta, Nsrc,
                                Ntar, Mx, My, Lx, Ly, w, P, out);
}
int se2p_slp_stress_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                           double eta, const double* f, int Nsrc, int Ntar, int Mx, int My,
                           double Lx, double Ly, double w, int P, double* out)
{
    return derived_kspace(ctx, SE2P_SLP_STRESS, psrc, ptar, f, nullptr, xi, eta, Nsrc,
                                Ntar, Mx, My, Lx, Ly, w, P, out);
}
int se2p_dlp_gradient_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                             double eta, const double* f, const double* n, int Nsrc, int Ntar,
                             int Mx, int My, double Lx, double Ly, double w, int P, double* out)
{
    return derived_kspace(ctx, SE2P_DLP_GRADIENT, psrc, ptar, f, n, xi, eta, Nsrc, Ntar,
                                Mx, My, Lx, Ly, w, P, out);
}
int se2p_dlp_pressure_grad_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                                  const double* f, const double* n, double xi, double eta,
                                  int Nsrc, int Ntar, int Mx, int My, double Lx, double Ly,
                                  double w, int P, double* out)
{
    return derived_kspace(ctx, SE2P_DLP_PRESSURE_GRAD, psrc, ptar, f, n, xi, eta, Nsrc,
                                Ntar, Mx, My, Lx, Ly, w, P, out);
}
int se2p_dlp_vorticity_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar,
                              const double* f, const double* n, double xi, double eta, int Nsrc,
                              int Ntar, int Mx, int My, double Lx, double Ly, double w, int P,
                              double* out)
{
    return derived_kspace(ctx, SE2P_DLP_VORTICITY, psrc, ptar, f, n, xi, eta, Nsrc, Ntar,
                                Mx, My, Lx, Ly, w, P, out);
}
int se2p_dlp_stress_kspace(se2p_ctx* ctx, const double* psrc, const double* ptar, double xi,
                           double eta, const double* f, const double* n, int Nsrc, int Ntar,
                           int Mx, int My, double Lx, double Ly, double w, int P, double* out)
{
    return derived_kspace(ctx, SE2P_DLP_STRESS, psrc, ptar, f, n, xi, eta, Nsrc, Ntar, Mx,
                                My, Lx, Ly, w, P, out);
}

// ---------------------------------------------------------------------------
// Free-space layer potentials (the FMM mex functions, computed directly)
// ---------------------------------------------------------------------------

// Smallest even 2-3-5-smooth integer >= n.
static int even_smooth_up(int n)
{
    for (int m = std::max(2, n + (n & 1));; m += 2) {
        int r = m;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

int se2p_free_params(int Nsrc, int Ntar, double wx, double wy, double tol, se2p_params* prm,
                     double* Lx, double* Ly)
{
    if (!prm || !Lx || !Ly) return set_error(SE2P_ERR_ARG, "null output");
    if (Nsrc < 0 || Ntar < 0 || Nsrc + Ntar < 1) return set_error(SE2P_ERR_ARG, "no points");
    if (!(wx >= 0) || !(wy >= 0) || !std::isfinite(wx) || !std::isfinite(wy))
        return set_error(SE2P_ERR_ARG, "box sides must be finite and >= 0");
    if (!(tol >= 1e-13) || !(tol < 1)) return set_error(SE2P_ERR_ARG, "tol must be in [1e-13, 1)");
    // a degenerate extent (points on a line, or one point) gets a box of its own
    double W = std::max(wx, wy);
    if (!(W > 0)) W = 1;
    wx = std::max(wx, 1e-3 * W);
    wy = std::max(wy, 1e-3 * W);
    // the periodic tuner's tolerance rules (tune.rule_params) with a margin of
    // 100 rather than 10: their real-space truncation exp(-(xi rc)^2) alone is
    // e^{3.2}/margin of the tolerance, which points on curves (few neighbours
    // within rc, sums that do not grow with the density's mean) reach
    const double L = std::log(100 / tol);
    const double cr = std::sqrt(std::max(L - 3.2, 1.0));
    const int P = std::min(24, std::max(8, 2 * (int)std::ceil(0.4 * L + 0.6)));
    const double npts = std::max(Nsrc, Ntar);
    // cells of about 22 points of the point box, at most the box
    double rc = std::min(std::sqrt(22.0 * wx * wy / npts), std::max(wx, wy));
    double xi, h, delta, R;
    int Mx, My;
    for (;;) {
        xi = cr / rc;
        // the screened truncated kernel differs from the free-space one by the
        // jump of the truncated kernel at R times the screening function's
        // tail, ~ (xi delta)^2 exp(-(xi delta)^2) at distance delta from it
        const double Dg = std::hypot(wx, wy);
        delta = std::sqrt(L + 2 * std::log(std::max(xi * (Dg + 1.0 / xi), 1.0))) / xi;
        R = Dg + delta;
        // k_inf = ck xi = pi / h.  The truncated kernel's transform keeps an
        // oscillating tail, R J1(K R)/(2 K^3) times K^2, of relative size
        // ~ (R K)^{1/2} against the periodic 1/K^2, so the Gaussian must fall
        // further at k_inf than the periodic rule's exp(-(L - 2.4)): by
        // (R k_inf)^{1/2}.
        double ck = 2 * std::sqrt(std::max(L - 2.4, 1.0));
        for (int it = 0; it < 3; ++it)
            ck = 2 * std::sqrt(std::max(L - 2.4 + 0.5 * std::log(std::max(R * ck * xi, 1.0)), 1.0));
        h = kPi / (ck * xi);
        Mx = even_smooth_up((int)std::ceil((wx + R + delta) / h));
        My = even_smooth_up((int)std::ceil((wy + R + delta) / h));
        if (std::max(Mx, My) <= kFftMaxLen) break;
        rc *= 1.25;  // coarser grid, more pairs in real space
    }
    *Lx = Mx * h;
    *Ly = My * h;
    const int nsx = std::max(1, (int)std::floor(*Lx / rc)), nsy = std::max(1, (int)std::floor(*Ly / rc));
    *prm = se2p_params{};
    prm->P = P;
    prm->nside_x = nsx;
    prm->nside_y = nsy;
    prm->rc = *Lx / nsx;
    prm->xi = xi;
    prm->kinf = Mx / 2;
    prm->Mx = Mx;
    prm->My = My;
    prm->w = P * h / 2;
    prm->eta = std::pow(2 * xi * prm->w / (0.95 * std::sqrt(kPi * P)), 2);
    prm->converged = 1;
    prm->e1_tier = tol >= 1e-10 ? 1 : 0;
    prm->symmetric = 0;
    prm->precision = SE2P_FP64;
    prm->far32 = prm->e1_tier == 1 && tol < 1e-6 ? std::log(1e-6 / tol) : 0;
    prm->free_R = R;
    return SE2P_OK;
}

// FMM accuracy flag -> relative tolerance (cfmm2dpart.f:351-361), kept within
// what the spectral Ewald sum reaches.
static double free_tol(int iprec)
{
    static const double eps[] = {5e-3, 5e-3, 5e-3, 5e-4, 5e-7, 5e-10, 5e-13};
    return eps[std::min(std::max(iprec, -2), 4) + 2];
}

// The free-space Stokeslet sum by spectral Ewald: the points centred in the box
// of se2p_free_params, one evaluation (pair-symmetric real space when the
// targets are the sources).
static int free_slp_ewald(se2p_ctx* ctx, const double* f1, const double* f2, const double* xs,
                          const double* ys, int Ns, const double* xt, const double* yt, int Nt,
                          bool same, double tol, double* u1, double* u2)
{
    double x0 = HUGE_VAL, x1 = -HUGE_VAL, y0 = HUGE_VAL, y1 = -HUGE_VAL;
    auto extend = [&](const double* x, const double* y, int n) {
        for (int i = 0; i < n; ++i) {
            x0 = std::min(x0, x[i]);
            x1 = std::max(x1, x[i]);
            y0 = std::min(y0, y[i]);
            y1 = std::max(y1, y[i]);
        }
    };
    extend(xs, ys, Ns);
    if (!same) extend(xt, yt, Nt);
    if (!std::isfinite(x1 - x0) || !std::isfinite(y1 - y0))
        return set_error(SE2P_ERR_ARG, "non-finite point coordinates");
    se2p_params prm;
    double Lx, Ly;
    SE2P_TRY(se2p_free_params(Ns, Nt, x1 - x0, y1 - y0, tol, &prm, &Lx, &Ly));
    prm.symmetric = same ? 1 : 0;
    const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
    std::vector<double> ps(2 * (size_t)Ns), f(2 * (size_t)Ns), pt(same ? 0 : 2 * (size_t)Nt);
    for (int i = 0; i < Ns; ++i) {
        ps[2 * i] = xs[i] - cx;
        ps[2 * i + 1] = ys[i] - cy;
        f[2 * i] = f1[i];
        f[2 * i + 1] = f2[i];
    }
    for (int i = 0; i < (same ? 0 : Nt); ++i) {
        pt[2 * i] = xt[i] - cx;
        pt[2 * i + 1] = yt[i] - cy;
    }
    double *dps, *dpt, *df;
    SE2P_TRY(upload(ctx, WS_PTS_A, ps.data(), ps.size(), &dps));
    if (same)
        dpt = dps;  // one pointer: the pair-symmetric sums apply
    else
        SE2P_TRY(upload(ctx, WS_PTS_B, pt.data(), pt.size(), &dpt));
    SE2P_TRY(upload(ctx, WS_DENS_A, f.data(), f.size(), &df));
    void* out = nullptr;
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, sizeof(double) * 2 * (size_t)Nt, &out));
    Geom G;
    SE2P_TRY(geom_build(ctx, &ctx->ws, &G, SE2P_SLP, dps, dpt, nullptr, Ns, Nt, Lx, Ly, prm, true,
                        true));
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ctx->ws, &B, &gp, 1));
    SE2P_TRY(batch_apply(ctx, B, df, (double*)out));
    std::vector<double> h(2 * (size_t)Nt);
    SE2P_TRY(download(ctx, h.data(), (double*)out, h.size()));
    for (int i = 0; i < Nt; ++i) {
        u1[i] = h[2 * i];
        u2[i] = h[2 * i + 1];
    }
    return SE2P_OK;
}

int se2p_free_slp(se2p_ctx* ctx, const double* f1, const double* f2, const double* xsrc,
                  const double* ysrc, int Nsrc, const double* xtar, const double* ytar, int Ntar,
                  int iequal, int iprec, double* u1, double* u2)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (Nsrc < 0 || Ntar < 0) return set_error(SE2P_ERR_ARG, "negative size");
    if (iequal != 0 && iequal != 1) return set_error(SE2P_ERR_ARG, "iequal must be 0 or 1");
    if (iprec < -2 || iprec > 5) return set_error(SE2P_ERR_ARG, "iprec must be in -2..5");
    if (iequal && Ntar != Nsrc)
        return set_error(SE2P_ERR_ARG, "iequal = 1 needs as many targets as sources");
    for (const double* p : {f1, f2, xsrc, ysrc})
        if (Nsrc > 0 && !p) return set_error(SE2P_ERR_ARG, "null source array");
    for (const double* p : {(const double*)u1, (const double*)u2})
        if (Ntar > 0 && !p) return set_error(SE2P_ERR_ARG, "null output array");
    // with iequal = 1 the targets are the sources (stokesSLP.f: rfmm2dpart at the sources)
    const double* xt = iequal ? xsrc : xtar;
    const double* yt = iequal ? ysrc : ytar;
    if (Ntar > 0 && (!xt || !yt)) return set_error(SE2P_ERR_ARG, "null target array");
    const bool ewald = iprec <= 4 && Nsrc > 0 && Ntar > 0 &&
                       (ctx->free_method == SE2P_FREE_EWALD ||
                        (ctx->free_method == SE2P_FREE_AUTO && (double)Nsrc * Ntar >= 1e9));
    if (ewald) {
        SE2P_HIP_CHECK(hipSetDevice(ctx->device));
        return free_slp_ewald(ctx, f1, f2, xsrc, ysrc, Nsrc, xt, yt, Ntar, iequal == 1,
                              free_tol(iprec), u1, u2);
    }
    return free_eval(ctx, false, f1, f2, xsrc, ysrc, nullptr, nullptr, Nsrc, xt, yt, Ntar,
                     iequal == 1, u1, u2);
}

int se2p_free_dlp(se2p_ctx* ctx, const double* f1, const double* f2, const double* xsrc,
                  const double* ysrc, const double* n1, const double* n2, int N, double* u1,
                  double* u2)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (N < 0) return set_error(SE2P_ERR_ARG, "negative size");
    for (const double* p : {f1, f2, xsrc, ysrc, n1, n2})
        if (N > 0 && !p) return set_error(SE2P_ERR_ARG, "null array");
    if (N > 0 && (!u1 || !u2)) return set_error(SE2P_ERR_ARG, "null output array");
    // stokesDLP.f:31: a density with sum |f1| + |f2| <= 1e-14 gives zero
    double zc = 0;
    for (int i = 0; i < N; ++i) zc += std::fabs(f1[i]) + std::fabs(f2[i]);
    if (zc <= 1e-14) {
        for (int i = 0; i < N; ++i) u1[i] = u2[i] = 0;
        return SE2P_OK;
    }
    return free_eval(ctx, true, f1, f2, xsrc, ysrc, n1, n2, N, xsrc, ysrc, N, true, u1, u2);
}

int se2p_dlp_pressure_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc,
                            const double* xtar, const double* ytar, const double* n1,
                            const double* n2, const double* f1, const double* f2, int Nsrc,
                            int Ntar, double Lx, double Ly, const se2p_options* opt, double* p,
                            double* pr, double* pk, double* xi)
{
    if (ctx) SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    return pressure_driver(ctx, SE2P_DLP_PRESSURE, xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Nsrc,
                           Ntar, Lx, Ly, opt, p, pr, pk, xi);
}

void se2p_options_default(se2p_options* opt)
{
    if (!opt) return;
    opt->P = 24;
    opt->Nb = 0;
    opt->tol = 1e-16;
    opt->verbose = 0;
}

int se2p_params_ref(int kind, int Nsrc, int Ntar, double Q, double Lx, double Ly,
                    const se2p_options* opt, se2p_params* out)
{
    if (!out) return set_error(SE2P_ERR_ARG, "null output");
    if (!valid_kind(kind)) return set_error(SE2P_ERR_ARG, "bad kind");
    if (!(Lx > 0) || !(Ly > 0)) return set_error(SE2P_ERR_ARG, "box lengths must be positive");
    // e.g. the pressure driver uses the velocity driver's recipe
    // (StokesSLP_pressure_ewald_2p.m:83-104 = StokesSLP_ewald_2p.m:83-104)
    return params_ref(recipe_of(kind), Nsrc, Ntar, Q, Lx, Ly, opt, out);
}

int se2p_slp_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc, const double* xtar,
                   const double* ytar, const double* f1, const double* f2, int Nsrc, int Ntar,
                   double Lx, double Ly, const se2p_options* opt, double* u1, double* u2,
                   double* ur, double* uk, double* xi)
{
    if (ctx) SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    return ewald_driver(ctx, SE2P_SLP, xsrc, ysrc, xtar, ytar, nullptr, nullptr, f1, f2, Nsrc,
                        Ntar, Lx, Ly, opt, u1, u2, ur, uk, xi);
}

int se2p_dlp_ewald(se2p_ctx* ctx, const double* xsrc, const double* ysrc, const double* xtar,
                   const double* ytar, const double* n1, const double* n2, const double* f1,
                   const double* f2, int Nsrc, int Ntar, double Lx, double Ly,
                   const se2p_options* opt, double* u1, double* u2, double* ur, double* uk,
                   double* xi)
{
    if (ctx) SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    return ewald_driver(ctx, SE2P_DLP, xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Nsrc, Ntar, Lx,
                        Ly, opt, u1, u2, ur, uk, xi);
}

int se2p_dev_alloc(se2p_ctx* ctx, size_t bytes, void** dptr)
{
    if (!ctx || !dptr) return set_error(SE2P_ERR_ARG, "null argument");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    void* p = nullptr;
    SE2P_HIP_CHECK(hipMalloc(&p, bytes ? bytes : 1));
    ctx->user_allocs.push_back(p);
    *dptr = p;
    return SE2P_OK;
}

int se2p_dev_free(se2p_ctx* ctx, void* dptr)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    for (size_t i = 0; i < ctx->user_allocs.size(); ++i)
        if (ctx->user_allocs[i] == dptr) {
            SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            SE2P_HIP_CHECK(hipFree(dptr));
            ctx->user_allocs.erase(ctx->user_allocs.begin() + i);
            return SE2P_OK;
        }
    return set_error(SE2P_ERR_ARG, "pointer was not allocated by se2p_dev_alloc");
}

int se2p_memcpy_h2d(se2p_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    SE2P_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SE2P_OK;
}

int se2p_memcpy_d2h(se2p_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    SE2P_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SE2P_OK;
}

int se2p_memcpy_d2d(se2p_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    SE2P_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SE2P_OK;
}

int se2p_ctx_sync(se2p_ctx* ctx)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SE2P_OK;
}

int se2p_eval_device(se2p_ctx* ctx, int kind, const double* d_psrc, const double* d_ptar,
                     const double* d_f, const double* d_n, int Nsrc, int Ntar, double Lx,
                     double Ly, const se2p_params* prm, double* d_u)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    if (!valid_kind(kind)) return set_error(SE2P_ERR_ARG, "bad kind");
    SE2P_TRY(check_params(prm, Lx, Ly));
    if (dlp_family(kind)) SE2P_TRY(need_ptr(d_n, Nsrc, "d_n"));
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (Ntar == 0) return SE2P_OK;
    Geom G;
    SE2P_TRY(geom_build(ctx, &ctx->ws, &G, kind, d_psrc, d_ptar, d_n, Nsrc, Ntar, Lx, Ly, *prm,
                        true, true));
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ctx->ws, &B, &gp, 1));
    return batch_apply(ctx, B, d_f, d_u);
}

// An upper bound of the scratch bytes one evaluation takes: every workspace
// request of geom_build, sort_by_key, fft_twiddles and batch_build, each counted
// as a region of its own rounded up to 256 bytes.
static size_t noalloc_bytes(int kind, int Ns, int Nt, const se2p_params& prm)
{
    auto r = [](double b) { return ((size_t)std::max(256.0, std::ceil(b)) + 255) & ~(size_t)255; };
    const double ns = Ns, nt = Nt;
    const double nbox = (double)prm.nside_x * prm.nside_y;
    const int P = prm.P;
    const double ntile = std::ceil(prm.Mx / 16.0) * std::ceil(prm.My / 16.0);
    // spreading: strips (records per point and strip, bins per strip and start
    // row, a segment table) or blocks (one record per point, 16 x 16 tiles)
    const bool ring = spread_ring(P);
    const int ntx = (prm.Mx + kSB - 1) / kSB;
    int dmax = 1;  // records per point at most
    if (ring)
        for (int sx = 0; sx < prm.Mx; ++sx) {
            int c = 0;
            for_each_strip(sx, prm.Mx, P, ntx, [&](int, int) { ++c; });
            dmax = std::max(dmax, c);
        }
    const double nrec = ns * dmax;
    const double nbins = ring ? (double)ntx * prm.My : ntile;
    const double nseg = ring ? (double)ntx * (std::ceil((double)prm.My / std::max(2, P)) + 1) : 0;
    const int gT = P <= 24 ? 32 : 16;
    const double gtile = std::ceil((double)prm.Mx / gT) * std::ceil((double)prm.My / gT);
    const double G = (double)prm.Mx * prm.My;
    const bool dlp = dlp_family(kind), two = dlp || out_grids(kind) == 2;
    size_t b = 0;
    if (dlp) b += r(16 * ns);  // normals
    // real space: keys, counts, permutations, starts, y keys, accumulator,
    // positions, sorted records and targets, flag, zero-mode partials
    b += r(4 * ns) + r(4 * nbox) + r(4 * nt) + r(4 * nbox) + r(4 * ns) + r(4 * (nbox + 1)) +
         r(4 * nt) + r(4 * (nbox + 1)) + r(4 * ns) + r(16 * ns) + r(4 * ns) +
         r((dlp ? 64 : 32) * ns) + r(16 * nt) + r(4) + r(8 * 512) + r(16);
    // k space: table, points' nodes and offsets, record counts and offsets
    // (strips), bin counts and starts, gathering keys, counts, permutations,
    // starts, nodes and offsets, grids, twiddles
    b += r(512) + r(8 * ns) + r(16 * ns) + (ring ? r(4 * ns) + r(4 * (ns + 1)) : 0) +
         r(4 * nbins) + r(4 * (nbins + 1)) + r(4 * nt) + r(4 * gtile) + r(4 * nt) +
         r(4 * (gtile + 1)) + r(8 * nt) + r(16 * nt) + r(16 * G) * (two ? 4 : 2) +
         r(16.0 * prm.Mx) + r(16.0 * prm.My);
    // spreading records: keys, strip nodes and owners (strips), sort order, points,
    // start nodes, FGG factors, segment table (strips)
    b += r(4 * nrec) + (ring ? r(8 * nrec) + r(4 * nrec) : 0) + r(4 * nrec) + r(4 * nrec) +
         r(8 * nrec) + r(32 * nrec) + (ring ? r(16 * nseg) : 0);
    b += r(8 * nt) + r(16 * nt) + r(4 * nt);  // binned gathering records
    // four stable sorts (source cells, target cells, spreading records and
    // gathering tiles): scratch of n and a list of nkeys + 1 each; the scans of
    // long counts (each may take a fresh region) in three passes
    for (double n : {ns, nt, nrec, nt}) b += r(4 * n);
    for (double k : {nbox, nbox, nbins, gtile}) b += r(4 * (k + 1));
    for (double n : {ns, nbox, nbox, nbins, gtile}) b += r(4 * (2 * std::ceil(n / 4096) + 1));
    // descriptor and job list (every job range of one system, generously)
    b += r(sizeof(SysDesc)) +
         r(sizeof(Job) * (8 * (ns + nt) / 256 + 8 * nbox + 8 * ntile + 2 * nseg + 4 * gtile +
                          4 * nrec / 256 + 4 * G / 256 + 4.0 * (prm.Mx + prm.My) + 64));
    return b;
}

int se2p_eval_device_noalloc(se2p_ctx* ctx, int kind, const double* d_psrc, const double* d_ptar,
                             const double* d_f, const double* d_n, int Nsrc, int Ntar, double Lx,
                             double Ly, const se2p_params* prm, double* d_u)
{
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    if (!valid_kind(kind)) return set_error(SE2P_ERR_ARG, "bad kind");
    SE2P_TRY(check_params(prm, Lx, Ly));
    if (dlp_family(kind)) SE2P_TRY(need_ptr(d_n, Nsrc, "d_n"));
    const size_t need = noalloc_bytes(kind, Nsrc, Ntar, *prm);
    if (need > kScratchBytes)
        return set_error(SE2P_ERR_UNSUPPORTED,
                         "se2p_eval_device_noalloc: Nsrc %d, Ntar %d, grid %d x %d need up to "
                         "%zu bytes of scratch, more than its fixed %zu; use se2p_eval_device",
                         Nsrc, Ntar, prm->Mx, prm->My, need, kScratchBytes);
    if (!ctx->scratch) return set_error(SE2P_ERR_HIP, "scratch address not resolved");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (Ntar == 0) return SE2P_OK;
    // Every workspace request is carved from this GPU's scratch array: no
    // allocation.  The kernels clear what they need cleared (counts, flags,
    // accumulators) themselves, as with fresh allocations.
    Workspace ws;
    ws.arena = ctx->scratch;
    ws.arena_cap = kScratchBytes;
    Geom G;
    SE2P_TRY(geom_build(ctx, &ws, &G, kind, d_psrc, d_ptar, d_n, Nsrc, Ntar, Lx, Ly, *prm, true,
                        true));
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ws, &B, &gp, 1));
    return batch_apply(ctx, B, d_f, d_u);
}

int se2p_plan_create(se2p_ctx* ctx, int kind, const double* d_psrc, const double* d_ptar,
                     const double* d_n, int Nsrc, int Ntar, double Lx, double Ly,
                     const se2p_params* prm, se2p_plan** out)
{
    if (!out) return set_error(SE2P_ERR_ARG, "null output pointer");
    *out = nullptr;
    SE2P_TRY(check_common(ctx, Nsrc, Ntar, Lx, Ly));
    if (!valid_kind(kind)) return set_error(SE2P_ERR_ARG, "bad kind");
    SE2P_TRY(check_params(prm, Lx, Ly));
    SE2P_TRY(need_ptr(d_psrc, Nsrc, "d_psrc"));
    SE2P_TRY(need_ptr(d_ptar, Ntar, "d_ptar"));
    if (dlp_family(kind)) SE2P_TRY(need_ptr(d_n, Nsrc, "d_n"));
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    se2p_plan* pl = new (std::nothrow) se2p_plan();
    if (!pl) return set_error(SE2P_ERR_NOMEM, "plan allocation failed");
    pl->ctx = ctx;
    int rc = geom_build(ctx, &pl->ws, &pl->G, kind, d_psrc, d_ptar, d_n, Nsrc, Ntar, Lx, Ly,
                        *prm, true, true);
    if (rc == SE2P_OK) {
        const Geom* gp = &pl->G;
        rc = batch_build(ctx, &pl->ws, &pl->B, &gp, 1);
    }
    if (rc == SE2P_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = set_error(SE2P_ERR_HIP, "plan build: %s", hipGetErrorString(e));
    }
    if (rc != SE2P_OK) {
        ws_release(&pl->ws);
        delete pl;
        return rc;
    }
    *out = pl;
    return SE2P_OK;
}

int se2p_plan_apply(se2p_plan* plan, const double* d_f, double* d_u)
{
    if (!plan) return set_error(SE2P_ERR_ARG, "null plan");
    SE2P_TRY(need_ptr(d_f, plan->G.Ns, "d_f"));
    SE2P_TRY(need_ptr(d_u, plan->G.Nt, "d_u"));
    SE2P_HIP_CHECK(hipSetDevice(plan->ctx->device));
    return batch_apply(plan->ctx, plan->B, d_f, d_u);
}

int se2p_plan_apply_graph(se2p_plan* plan, const double* d_f, double* d_u)
{
    if (!plan) return set_error(SE2P_ERR_ARG, "null plan");
    SE2P_TRY(need_ptr(d_f, plan->G.Ns, "d_f"));
    SE2P_TRY(need_ptr(d_u, plan->G.Nt, "d_u"));
    return graph_apply(plan->ctx, &plan->gc, plan->B, d_f, d_u);
}

int se2p_plan_destroy(se2p_plan* plan)
{
    if (!plan) return SE2P_OK;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    graph_release(&plan->gc);
    ws_release(&plan->ws);
    delete plan;
    return SE2P_OK;
}

// Parameters of a geometry as its evaluations run them: symmetric reports
// whether pair-symmetric real-space sums are in use.
static void params_in_use(const Geom& G, se2p_params* used)
{
    *used = G.prm;
    used->symmetric = G.sym ? 1 : 0;
    used->precision = G.prec ? SE2P_FP32C : SE2P_FP64;
}

int se2p_plan_params(const se2p_plan* plan, se2p_params* used)
{
    if (!plan || !used) return set_error(SE2P_ERR_ARG, "null plan or output");
    params_in_use(plan->G, used);
    return SE2P_OK;
}

int se2p_batch_create(se2p_ctx* ctx, int nsys, const se2p_system* systems, se2p_batch** out)
{
    if (!out) return set_error(SE2P_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (nsys < 0 || (nsys > 0 && !systems)) return set_error(SE2P_ERR_ARG, "bad system list");
    for (int i = 0; i < nsys; ++i) {
        const se2p_system& y = systems[i];
        if (!valid_kind(y.kind)) return set_error(SE2P_ERR_ARG, "system %d: bad kind", i);
        int rc = check_common(ctx, y.Nsrc, y.Ntar, y.Lx, y.Ly);
        if (rc == SE2P_OK) rc = check_params(&y.prm, y.Lx, y.Ly);
        if (rc == SE2P_OK) rc = need_ptr(y.d_psrc, y.Nsrc, "d_psrc");
        if (rc == SE2P_OK) rc = need_ptr(y.d_ptar, y.Ntar, "d_ptar");
        if (rc == SE2P_OK && dlp_family(y.kind)) rc = need_ptr(y.d_n, y.Nsrc, "d_n");
        if (rc != SE2P_OK)
            return set_error(rc, "system %d: %s", i, std::string(se2p_last_error()).c_str());
    }
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    se2p_batch* b = new (std::nothrow) se2p_batch();
    if (!b) return set_error(SE2P_ERR_NOMEM, "batch allocation failed");
    b->ctx = ctx;
    b->G.resize(nsys);
    b->ws.resize(nsys);
    int rc = SE2P_OK;
    std::vector<const Geom*> gp(nsys);
    for (int i = 0; i < nsys && rc == SE2P_OK; ++i) {
        const se2p_system& y = systems[i];
        rc = geom_build(ctx, &b->ws[i], &b->G[i], y.kind, y.d_psrc, y.d_ptar, y.d_n, y.Nsrc,
                        y.Ntar, y.Lx, y.Ly, y.prm, true, true);
        gp[i] = &b->G[i];
        b->nf += 2LL * y.Nsrc;
        b->nu += (long long)out_comps(y.kind) * y.Ntar;
    }
    if (rc == SE2P_OK) rc = batch_build(ctx, &b->bws, &b->B, gp.data(), nsys);
    if (rc != SE2P_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& w : b->ws) ws_release(&w);
        ws_release(&b->bws);
        delete b;
        return rc;
    }
    *out = b;
    return SE2P_OK;
}

int se2p_batch_sizes(const se2p_batch* batch, long long* nf, long long* nu)
{
    if (!batch) return set_error(SE2P_ERR_ARG, "null batch");
    if (nf) *nf = batch->nf;
    if (nu) *nu = batch->nu;
    return SE2P_OK;
}

int se2p_batch_apply(se2p_batch* batch, const double* d_f, double* d_u)
{
    if (!batch) return set_error(SE2P_ERR_ARG, "null batch");
    if ((batch->nf && !d_f) || (batch->nu && !d_u)) return set_error(SE2P_ERR_ARG, "null array");
    SE2P_HIP_CHECK(hipSetDevice(batch->ctx->device));
    return batch_apply(batch->ctx, batch->B, d_f, d_u);
}

int se2p_batch_apply_graph(se2p_batch* batch, const double* d_f, double* d_u)
{
    if (!batch) return set_error(SE2P_ERR_ARG, "null batch");
    if ((batch->nf && !d_f) || (batch->nu && !d_u)) return set_error(SE2P_ERR_ARG, "null array");
    return graph_apply(batch->ctx, &batch->gc, batch->B, d_f, d_u);
}

int se2p_batch_params(const se2p_batch* batch, int sys, se2p_params* used)
{
    if (!batch || !used) return set_error(SE2P_ERR_ARG, "null batch or output");
    if (sys < 0 || sys >= (int)batch->G.size())
        return set_error(SE2P_ERR_ARG, "system %d of %d", sys, (int)batch->G.size());
    params_in_use(batch->G[sys], used);
    return SE2P_OK;
}

int se2p_batch_destroy(se2p_batch* batch)
{
    if (!batch) return SE2P_OK;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipStreamSynchronize(batch->ctx->stream);
    graph_release(&batch->gc);
    for (auto& w : batch->ws) ws_release(&w);
    ws_release(&batch->bws);
    delete batch;
    return SE2P_OK;
}

int se2p_mlp_forward(se2p_ctx* ctx, const float* d_w, const float* d_x, int n, float* d_y)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (n < 0 || (n > 0 && (!d_w || !d_x || !d_y))) return set_error(SE2P_ERR_ARG, "bad arrays");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    return mlp_forward(ctx, d_w, d_x, n, d_y);
}

int se2p_mlp_loss_grad(se2p_ctx* ctx, const float* d_w, const float* d_x, const float* d_t,
                       int n, float* d_grad, float* d_loss)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (n <= 0 || !d_w || !d_x || !d_t || !d_grad)
        return set_error(SE2P_ERR_ARG, "bad arrays or no samples");
    if (ctx->capturing) return set_error(SE2P_ERR_ARG, "not under graph capture");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    return mlp_grad(ctx, &ctx->ws, d_w, d_x, d_t, n, d_grad, d_loss);
}

int se2p_mlp_adam(se2p_ctx* ctx, float* d_w, const float* d_g, float* d_m, float* d_v, int step,
                  float lr, float beta1, float beta2, float eps)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (!d_w || !d_g || !d_m || !d_v || step < 1) return set_error(SE2P_ERR_ARG, "bad arguments");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    return mlp_adam(ctx, d_w, d_g, d_m, d_v, step, lr, beta1, beta2, eps);
}

int se2p_ctx_set_precision(se2p_ctx* ctx, int precision)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (precision != SE2P_FP64 && precision != SE2P_FP32C)
        return set_error(SE2P_ERR_ARG, "precision must be SE2P_FP64 or SE2P_FP32C");
    ctx->precision = precision;
    return SE2P_OK;
}

int se2p_ctx_set_free_method(se2p_ctx* ctx, int method)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (method != SE2P_FREE_AUTO && method != SE2P_FREE_DIRECT && method != SE2P_FREE_EWALD)
        return set_error(SE2P_ERR_ARG, "free-space method must be SE2P_FREE_AUTO, _DIRECT or _EWALD");
    ctx->free_method = method;
    return SE2P_OK;
}

int se2p_prof_enable(se2p_ctx* ctx, int on)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (on < 0 || on > 2) return set_error(SE2P_ERR_ARG, "profiling mode %d: need 0, 1 or 2", on);
    ctx->prof = on != 0;
    ctx->prof_serial = on == 2;
    return SE2P_OK;
}

int se2p_prof_reset(se2p_ctx* ctx)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    SE2P_TRY(prof_flush(ctx));
    ctx->prof_acc.clear();
    return SE2P_OK;
}

int se2p_prof_get(se2p_ctx* ctx, const char* stage, double* total_ms, int* count)
{
    if (!ctx || !stage || !total_ms || !count) return set_error(SE2P_ERR_ARG, "null argument");
    SE2P_TRY(prof_flush(ctx));
    *total_ms = 0;
    *count = 0;
    for (auto& a : ctx->prof_acc)
        if (a.name == stage) {
            *total_ms = a.ms;
            *count = a.count;
        }
    return SE2P_OK;
}

int se2p_assign(se2p_ctx* ctx, const double* p, int n, double Lx, double Ly, int nside_x,
                int nside_y, int* particle_offsets, int* box_offsets, int* n_in_box)
{
    SE2P_TRY(check_common(ctx, n, 0, Lx, Ly));
    SE2P_TRY(check_real(nside_x, nside_y, 1.0));
    SE2P_TRY(need_ptr(p, n, "p"));
    SE2P_TRY(need_ptr(particle_offsets, n, "particle_offsets"));
    if (!box_offsets || !n_in_box) return set_error(SE2P_ERR_ARG, "null box arrays");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    const int nbox = nside_x * nside_y;
    double* dp;
    void *key, *perm, *start, *count;
    SE2P_TRY(upload(ctx, WS_PTS_A, p, 2 * (size_t)n, &dp));
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_SRC_KEY, sizeof(int) * (size_t)n, &key));
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_SRC_PERM, sizeof(int) * (size_t)n, &perm));
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_SRC_START, sizeof(int) * (size_t)(nbox + 1), &start));
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_SRC_COUNT, sizeof(int) * (size_t)nbox, &count));
    SE2P_TRY(launch_fill_int(ctx->stream, (int*)count, nbox, 0));
    SE2P_TRY(launch_cell_keys(ctx->stream, dp, n, Lx, Ly, nside_x, nside_y, (int*)key,
                              (int*)count));
    SE2P_TRY(sort_by_key(ctx, &ctx->ws, (int*)key, (int*)count, n, nbox, (int*)perm, (int*)start,
                         WS_SORT_TMP));
    if (n) SE2P_HIP_CHECK(hipMemcpyAsync(particle_offsets, perm, sizeof(int) * (size_t)n,
                                         hipMemcpyDeviceToHost, ctx->stream));
    SE2P_HIP_CHECK(hipMemcpyAsync(box_offsets, start, sizeof(int) * (size_t)(nbox + 1),
                                  hipMemcpyDeviceToHost, ctx->stream));
    SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < nbox; ++b) n_in_box[b] = box_offsets[b + 1] - box_offsets[b];
    return SE2P_OK;
}

int se2p_find_closest_node(se2p_ctx* ctx, const double* p, int n, double Lx, double Ly,
                           double h, int P, int* mx, int* my, double* px, double* py)
{
    SE2P_TRY(check_common(ctx, n, 0, Lx, Ly));
    if (!(h > 0)) return set_error(SE2P_ERR_ARG, "h must be positive");
    SE2P_TRY(need_ptr(p, n, "p"));
    if (n > 0 && (!mx || !my || !px || !py)) return set_error(SE2P_ERR_ARG, "null output");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (n == 0) return SE2P_OK;
    double* dp;
    void* out;
    SE2P_TRY(upload(ctx, WS_PTS_A, p, 2 * (size_t)n, &dp));
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, (2 * sizeof(int) + 2 * sizeof(double)) * (size_t)n, &out));
    double* dpx = (double*)out;
    double* dpy = dpx + n;
    int* dmx = (int*)(dpy + n);
    int* dmy = dmx + n;
    SE2P_TRY(launch_closest_nodes(ctx->stream, dp, n, Lx, Ly, h, P, dmx, dmy, dpx, dpy));
    hipStream_t s = ctx->stream;
    SE2P_HIP_CHECK(hipMemcpyAsync(mx, dmx, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
    SE2P_HIP_CHECK(hipMemcpyAsync(my, dmy, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
    SE2P_HIP_CHECK(hipMemcpyAsync(px, dpx, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s));
    SE2P_HIP_CHECK(hipMemcpyAsync(py, dpy, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s));
    SE2P_HIP_CHECK(hipStreamSynchronize(s));
    return SE2P_OK;
}

int se2p_spread(se2p_ctx* ctx, const double* psrc, const double* f, int Nsrc, double Lx,
                double Ly, double xi, double w, double eta, int P, int Mx, int My, double* H1,
                double* H2)
{
    SE2P_TRY(check_common(ctx, Nsrc, 0, Lx, Ly));
    SE2P_TRY(check_kspace(Mx, My, Lx, Ly, xi, eta, w, P));
    SE2P_TRY(need_ptr(psrc, Nsrc, "psrc"));
    SE2P_TRY(need_ptr(f, Nsrc, "f"));
    if (!H1 || !H2) return set_error(SE2P_ERR_ARG, "null grid");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    double *dps, *df;
    SE2P_TRY(upload(ctx, WS_PTS_A, psrc, 2 * (size_t)Nsrc, &dps));
    SE2P_TRY(upload(ctx, WS_DENS_A, f, 2 * (size_t)Nsrc, &df));
    // k-space geometry of the sources only (no targets); spreading of (f1, f2)
    Geom G;
    se2p_params kp = k_only(xi, eta, Mx, My, w, P);
    kp.precision = ctx->precision;
    SE2P_TRY(geom_build(ctx, &ctx->ws, &G, SE2P_SLP, dps, dps, nullptr, Nsrc, 0, Lx, Ly, kp, false,
                        true));
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ctx->ws, &B, &gp, 1));
    SE2P_TRY(batch_spread_grid(ctx, ctx->stream, B, df, 0));
    const size_t Gn = (size_t)Mx * My;
    if (G.prec) {  // FP32 tier: a float2 grid
        std::vector<float2> h(Gn);
        SE2P_HIP_CHECK(hipMemcpyAsync(h.data(), G.Z0, sizeof(float2) * Gn, hipMemcpyDeviceToHost,
                                      ctx->stream));
        SE2P_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < Gn; ++i) {
            H1[i] = h[i].x;
            H2[i] = h[i].y;
        }
        return SE2P_OK;
    }
    std::vector<double> h(2 * Gn);
    SE2P_TRY(download(ctx, h.data(), (double*)G.Z0, 2 * Gn));
    for (size_t i = 0; i < Gn; ++i) {
        H1[i] = h[2 * i];
        H2[i] = h[2 * i + 1];
    }
    return SE2P_OK;
}

int se2p_gather(se2p_ctx* ctx, const double* H, const double* ptar, int Ntar, double Lx,
                double Ly, double xi, double w, double eta, int P, int Mx, int My, double* out)
{
    SE2P_TRY(check_common(ctx, 0, Ntar, Lx, Ly));
    SE2P_TRY(check_kspace(Mx, My, Lx, Ly, xi, eta, w, P));
    SE2P_TRY(need_ptr(ptar, Ntar, "ptar"));
    SE2P_TRY(need_ptr(out, Ntar, "out"));
    if (!H) return set_error(SE2P_ERR_ARG, "null grid");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (Ntar == 0) return SE2P_OK;
    double* dpt;
    SE2P_TRY(upload(ctx, WS_PTS_B, ptar, 2 * (size_t)Ntar, &dpt));
    // k-space geometry of the targets only (no sources); the grid goes where the
    // gather reads it
    Geom G;
    se2p_params kp = k_only(xi, eta, Mx, My, w, P);
    kp.precision = ctx->precision;
    SE2P_TRY(geom_build(ctx, &ctx->ws, &G, SE2P_SLP, dpt, dpt, nullptr, 0, Ntar, Lx, Ly, kp, false,
                        true));
    const size_t Gn = (size_t)Mx * My;
    double2* const gdst = fft_xfilter_fits(SE2P_SLP, Mx, kp.free_R) ? G.Z0 : G.W;
    std::vector<double> wz(2 * Gn, 0.0);
    std::vector<float2> wf;
    if (G.prec) {  // FP32 tier: the grid travels as float2
        wf.assign(Gn, make_float2(0.f, 0.f));
        for (size_t i = 0; i < Gn; ++i) wf[i].x = (float)H[i];
        SE2P_HIP_CHECK(hipMemcpyAsync(gdst, wf.data(), sizeof(float2) * Gn,
                                      hipMemcpyHostToDevice, ctx->stream));
    } else {
        for (size_t i = 0; i < Gn; ++i) wz[2 * i] = H[i];
        SE2P_HIP_CHECK(hipMemcpyAsync(gdst, wz.data(), sizeof(double) * 2 * Gn,
                                      hipMemcpyHostToDevice, ctx->stream));
    }
    BatchExec B;
    const Geom* gp = &G;
    SE2P_TRY(batch_build(ctx, &ctx->ws, &B, &gp, 1));  // synchronises: wz may die after
    void* u;
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, sizeof(double) * 2 * (size_t)Ntar, &u));
    SE2P_TRY(batch_gather(ctx, B, (double*)u));
    std::vector<double> h(2 * (size_t)Ntar);
    SE2P_TRY(download(ctx, h.data(), (double*)u, 2 * (size_t)Ntar));
    for (int i = 0; i < Ntar; ++i) out[i] = h[2 * i];
    return SE2P_OK;
}

int se2p_fft2(se2p_ctx* ctx, double* re, double* im, int My, int Mx, int inverse)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (Mx < 1 || My < 1 || !re || !im) return set_error(SE2P_ERR_ARG, "bad grid");
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    SysDesc d{};
    d.Mx = Mx;
    d.My = My;
    d.Bx = fft_lines_per_block(Mx);
    d.By = fft_lines_per_block(My);
    SE2P_TRY(fft_axis_init(&d.ax, Mx));
    SE2P_TRY(fft_axis_init(&d.ay, My));
    double2 *twx, *twy;
    SE2P_TRY(fft_twiddles(ctx, &ctx->ws, Mx, WS_TWIDDLE_X, &twx));
    SE2P_TRY(fft_twiddles(ctx, &ctx->ws, My, WS_TWIDDLE_Y, &twy));
    d.twx = twx;
    d.twy = twy;
    const size_t Gn = (size_t)Mx * My;
    const int prec = ctx->precision == SE2P_FP32C ? 1 : 0;
    std::vector<double> z(2 * Gn);
    for (size_t i = 0; i < Gn; ++i) {
        z[2 * i] = re[i];
        z[2 * i + 1] = im[i];
    }
    double* dz;
    void* tmp;
    std::vector<float2> zf;
    if (prec) {  // FP32 tier: float2 lines
        zf.resize(Gn);
        for (size_t i = 0; i < Gn; ++i) zf[i] = make_float2((float)re[i], (float)im[i]);
        void* p = nullptr;
        SE2P_TRY(ws_get(ctx, &ctx->ws, WS_GRID_A, sizeof(double2) * Gn, &p));
        dz = (double*)p;
        SE2P_HIP_CHECK(hipMemcpyAsync(dz, zf.data(), sizeof(float2) * Gn, hipMemcpyHostToDevice, s));
    } else {
        SE2P_TRY(upload(ctx, WS_GRID_A, z.data(), 2 * Gn, &dz));
    }
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_GRID_C, sizeof(double2) * Gn, &tmp));
    d.grid[G_Z0] = (double2*)dz;
    d.grid[G_W] = (double2*)tmp;
    // x lines read as the columns of the column-major Z0 into the row-major W,
    // then y lines read as the columns of W back into Z0 (the pipeline runs
    // both line kernels on contiguous lines too, and the second read mode)
    const int ny = (Mx + d.By - 1) / d.By, nx = (My + d.Bx - 1) / d.Bx;
    std::vector<Job> J;
    for (int b = 0; b < nx; ++b) J.push_back(make_int2(0, b));
    for (int b = 0; b < ny; ++b) J.push_back(make_int2(0, b));
    void *psys, *pjobs;
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_SYS, sizeof(SysDesc), &psys));
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_JOBS, sizeof(Job) * J.size(), &pjobs));
    SE2P_HIP_CHECK(hipMemcpyAsync(psys, &d, sizeof(SysDesc), hipMemcpyHostToDevice, s));
    SE2P_HIP_CHECK(hipMemcpyAsync(pjobs, J.data(), sizeof(Job) * J.size(), hipMemcpyHostToDevice, s));
    SE2P_TRY(fft_prepare());
    const SysDesc* ds = (const SysDesc*)psys;
    const Job* dj = (const Job*)pjobs;
    const int inv = inverse ? 1 : 0;
    SE2P_TRY(launch_fft_lines(s, prec, 1, G_Z0, G_W, 1, inv, fft_line_class(Mx),
                              fft_lds_bytes(Mx, prec), ds, dj, nx));
    SE2P_TRY(launch_fft_lines(s, prec, 0, G_W, G_Z0, 1, inv, fft_line_class(My),
                              fft_lds_bytes(My, prec), ds, dj + nx, ny));
    if (prec) {
        SE2P_HIP_CHECK(hipMemcpyAsync(zf.data(), dz, sizeof(float2) * Gn, hipMemcpyDeviceToHost, s));
        SE2P_HIP_CHECK(hipStreamSynchronize(s));
        for (size_t i = 0; i < Gn; ++i) {
            z[2 * i] = zf[i].x;
            z[2 * i + 1] = zf[i].y;
        }
    } else {
        SE2P_TRY(download(ctx, z.data(), dz, 2 * Gn));
    }
    const double sc = inverse ? 1.0 / ((double)Mx * (double)My) : 1.0;
    for (size_t i = 0; i < Gn; ++i) {
        re[i] = z[2 * i] * sc;
        im[i] = z[2 * i + 1] * sc;
    }
    return SE2P_OK;
}

int se2p_e1(se2p_ctx* ctx, const double* x, int n, int tier, double* out)
{
    if (!ctx) return set_error(SE2P_ERR_ARG, "null context");
    if (n < 0) return set_error(SE2P_ERR_ARG, "negative count");
    if (tier != 0 && tier != 1) return set_error(SE2P_ERR_ARG, "tier must be 0 or 1");
    SE2P_TRY(need_ptr(x, n, "x"));
    SE2P_TRY(need_ptr(out, n, "out"));
    SE2P_HIP_CHECK(hipSetDevice(ctx->device));
    if (n == 0) return SE2P_OK;
    double* dx;
    void* dout;
    SE2P_TRY(upload(ctx, WS_PTS_A, x, (size_t)n, &dx));
    SE2P_TRY(ws_get(ctx, &ctx->ws, WS_OUT_A, sizeof(double) * (size_t)n, &dout));
    SE2P_TRY(launch_e1(ctx->stream, dx, n, tier, (double*)dout));
    return download(ctx, out, (double*)dout, (size_t)n);
}

}  // extern "C"
