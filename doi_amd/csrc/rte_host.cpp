// rte_host.cpp -- host side of librte: the C ABI of include/rte.h, device
// memory of contexts and media, per-medium tables and the Neumann driver.
//
// The driver replaces the reference's Neumann()
// (GPU package/src/neumann.cu:23-181, FisherInfo_IO/src/neumann.cu:23-120) and
// the CPU main loop (CPU package/src/RTE_uniform.c:79-115): the field stays on
// the device for the whole series, every term is one propagation launch, one
// combine launch (self block, scatter, exitance increment, stop-rule sums) and
// one per-solve bottom-face reduction; only two scalars per solve come back to
// the host per term, where the stop rule runs in FP64.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <tuple>
#include <unordered_set>
#include <vector>

#include "../../include/rte.h"
#include "rte_internal.hpp"

using namespace rte;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(...)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (__VA_ARGS__);                                                         \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? RTE_E_NOMEM : RTE_E_DEVICE, "%s: %s (%s:%d)", \
                        #__VA_ARGS__, hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return e;
        }
        bytes = n;
        return hipSuccess;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// ---------------------------------------------------------------------------
// host-side math: SH recurrence constants, exitance weights, self-block grid
// ---------------------------------------------------------------------------
static double nlm(int l, int m)
{
    double f = 1.0;  // (l+m)!/(l-m)!
    for (int i = l - m + 1; i <= l + m; ++i) f *= (double)i;
    return std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI * f));
}

// n_a of the scaled coefficients (the monic recurrences' gamma is sh_gamma)
static double sh_scale(int l, int m)
{
    double c = ((m & 1) ? -1.0 : 1.0);
    for (int i = 2 * m - 1; i > 0; i -= 2) c *= (double)i;        // c_mm = (-1)^m (2m-1)!!
    for (int j = m + 1; j <= l; ++j) c *= (2.0 * j - 1.0) / (double)(j - m);
    return (m > 0 ? std::sqrt(2.0) : 1.0) * nlm(l, m) * c;
}

static double sh_scale_a(int a)
{
    int l = 0;
    while ((l + 1) * (l + 1) <= a) ++l;
    const int m = a - l * l - l;
    return sh_scale(l, m < 0 ? -m : m);
}

// real SH R_a(x,y,z) for a unit vector, with the recurrences of the kernels
static void real_sh(int L, double x, double y, double z, double *out)
{
    double C[kMaxL + 1], S[kMaxL + 1], P[kMaxL + 1][kMaxL + 1];
    C[0] = 1.0;
    S[0] = 0.0;
    for (int m = 1; m <= L; ++m) {
        C[m] = C[m - 1] * x - S[m - 1] * y;
        S[m] = S[m - 1] * x + C[m - 1] * y;
    }
    for (int m = 0; m <= L; ++m) {
        P[m][m] = 1.0;
        if (m + 1 <= L) P[m + 1][m] = z;
        for (int l = m + 2; l <= L; ++l) P[l][m] = z * P[l - 1][m] - sh_gamma<double>(l, m) * P[l - 2][m];
    }
    for (int l = 0; l <= L; ++l) {
        out[l * l + l] = sh_scale(l, 0) * P[l][0];
        for (int m = 1; m <= l; ++m) {
            out[l * l + l + m] = sh_scale(l, m) * P[l][m] * C[m];
            out[l * l + l - m] = sh_scale(l, m) * P[l][m] * S[m];
        }
    }
}

static double legendre(int l, double x)
{
    if (l == 0) return 1.0;
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= l; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// Q_l = int_0^1 P_l(u) u du: the reference's 100-point left Riemann sum
// (CPU package/src/Dist.c:190-193,211-218) or exact (Gauss-Legendre, 32 nodes)
static void quad_weights(int L, int quad, double *Q);

// per-solve pixel buffers: dE, bottom sums, E, dR, R
constexpr int kPixBufs = 5;
template <typename T, typename TA>
static void set_pixels(SolveDev<T, TA> &s, TA *base, int P)
{
    s.dE = base;
    s.bsum = base + P;
    s.E = base + 2 * P;
    s.dR = base + 3 * P;
    s.R = base + 4 * P;
}

// R_l = int_{-1}^{0} P_l(u) u du: the reference's left Riemann sum over u = p/100 - 1
// (CPU package/src/Dist.c:146-160), or exactly (-1)^(l+1) int_0^1 P_l(u) u du
static void quad_weights_refl(int L, int quad, double *R)
{
    if (quad == RTE_QUAD_REF) {
        for (int l = 0; l <= L; ++l) {
            double s = 0.0;
            for (int p = 0; p < 100; ++p) {
                const double u = (double)p / 100.0 - 1.0;
                s += legendre(l, u) * u / 100.0;
            }
            R[l] = s;
        }
        return;
    }
    quad_weights(L, quad, R);
    for (int l = 0; l <= L; ++l) R[l] *= (l & 1) ? 1.0 : -1.0;
}

// weights of w_l0 in the bottom exitance (CPU package/src/Dist.c:211-218) and
// the top reflectance (Dist.c:154-158, with its minus sign), in c_m-free units
static void face_coefs(int L, int quad, double dx, double dy, double *cE, double *cR)
{
    double Q[kMaxL + 1], R[kMaxL + 1];
    quad_weights(L, quad, Q);
    quad_weights_refl(L, quad, R);
    for (int l = 0; l <= kMaxL; ++l) {
        const double f = l <= L ? 2.0 * M_PI * dx * dy / std::sqrt(4.0 * M_PI) * std::sqrt(2.0 * l + 1.0) : 0.0;
        cE[l] = l <= L ? f * Q[l] : 0.0;
        cR[l] = l <= L ? -f * R[l] : 0.0;
    }
}

static void quad_weights(int L, int quad, double *Q)
{
    if (quad == RTE_QUAD_REF) {
        for (int l = 0; l <= L; ++l) {
            double s = 0.0;
            for (int p = 0; p < 100; ++p) {
                const double u = (double)p / 100.0;
                s += legendre(l, u) * u / 100.0;
            }
            Q[l] = s;
        }
        return;
    }
    const int NG = 32;
    double xg[NG], wg[NG];
    for (int i = 0; i < NG; ++i) {
        double x = std::cos(M_PI * (i + 0.75) / (NG + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= NG; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = NG * (x * p1 - p0) / (x * x - 1.0);
            const double step = p1 / dp;
            x -= step;
            if (std::fabs(step) < 1e-16) break;
        }
        xg[i] = x;
        wg[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    for (int l = 0; l <= L; ++l) {
        double s = 0.0;
        for (int i = 0; i < NG; ++i) {
            const double u = 0.5 * (xg[i] + 1.0);
            s += 0.5 * wg[i] * legendre(l, u) * u;
        }
        Q[l] = s;
    }
}

// Centre-to-face distance beta(omega) of the self-voxel quadrature on the
// 200x200 (theta, phi) grid, following CPU package/src/Kernels_selfsub.c:815-948
// with one sub-voxel at the centre (selfSubVox = 1, CPU package/src/RTE_uniform.c:63):
// direction cosines and face intersections are single precision there, the
// faces are tried x, y, z with the last accepted hit kept (also across
// directions), and the same is done here so that beta agrees to the bit.
static void self_beta(double dx, double dy, double dz, std::vector<double> &beta,
                      std::vector<double> &theta, std::vector<double> &phi)
{
    beta.resize(kNOmega);
    theta.resize(kNOmega);
    phi.resize(kNOmega);
    float hx = 0.f, hy = 0.f, hz = 0.f;  // last accepted intersection
    int w = 0;
    for (int t = 0; t < kAngRes; ++t) {
        for (int p = 0; p < kAngRes; ++p, ++w) {
            const double th = t * M_PI / kAngRes;
            const double ph = p * 2.0 * M_PI / kAngRes;
            theta[w] = th;
            phi[w] = ph;
            const float u[3] = {(float)(std::cos(ph) * std::sin(th)),
                                (float)(std::sin(ph) * std::sin(th)), (float)std::cos(th)};
            for (int ax = 0; ax < 3; ++ax) {
                if (u[ax] == 0.0f) continue;
                for (int f = 0; f < 2; ++f) {
                    float h[3];
                    h[ax] = f == 0 ? 1.0f : -1.0f;
                    // the two other coordinates, in the reference's order
                    const int o1 = ax == 0 ? 1 : (ax == 1 ? 2 : 0);
                    const int o2 = ax == 0 ? 2 : (ax == 1 ? 0 : 1);
                    h[o1] = (float)((double)h[ax] * (double)u[o1] / (double)u[ax]);
                    h[o2] = (float)((double)h[ax] * (double)u[o2] / (double)u[ax]);
                    bool inside = true;
                    for (int c = 0; c < 3; ++c) inside = inside && h[c] <= 1.0f && h[c] >= -1.0f;
                    if (inside && u[ax] * h[ax] >= 0.0f) {
                        hx = h[0];
                        hy = h[1];
                        hz = h[2];
                    }
                }
            }
            const double ex = hx, ey = hy, ez = hz;
            beta[w] = std::sqrt(ex * ex * (dx / 2.0) * (dx / 2.0) + ey * ey * (dy / 2.0) * (dy / 2.0) +
                                ez * ez * (dz / 2.0) * (dz / 2.0));
        }
    }
}

// ---------------------------------------------------------------------------
// context and media
// ---------------------------------------------------------------------------
struct MomentSet {
    DevBuf md, mf;  // [kNMom][K][K] in double and float
};

struct PendingTiming {
    int kind;  // 0 prop, 1 ktab, 2 other
    hipEvent_t a, b;
    double pairs, flops;
    int group;  // propagation kernel (LaunchTimer::group)
};

struct rte_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool prof = false;
    int max_rhs = 5;  // beams per propagation group (1..kMaxNR); 5 measured fastest at 32^3 L=5
    rte_prof pc{};
    std::vector<hipEvent_t> ev_pool;
    std::vector<PendingTiming> pending;
    std::map<std::tuple<int, double, double, double>, std::unique_ptr<MomentSet>> moments;
    // solve workspace (grow-only)
    DevBuf ws_W, ws_Wp, ws_part, ws_pix, ws_sv, ws_beams, ws_red;
    std::set<rte_medium *> media;  // live media of this context
    std::vector<char> host_red;
};

struct rte_medium {
    rte_ctx *ctx = nullptr;
    rte_grid grid{};
    double g = 0, n = 1, c = 1;
    int N = 0;
    std::vector<double> mua_h, mus_h;
    DevBuf mua_d, mus_d;          // FP64 maps
    DevBuf mut_f, mus_f;          // FP32 mu_t, mu_s
    DevBuf mut_d;                 // FP64 mu_t
    DevBuf ktab_f, ktab_d;        // cached e^{-tau}/r^2 tables
    double ktab_f_thr = -1, ktab_d_thr = -1;
};

static int set_device(rte_ctx *ctx)
{
    HIP_TRY(hipSetDevice(ctx->device));
    return RTE_OK;
}

// ---- launch timing (only while profiling is on) ----------------------------
static hipEvent_t pool_event(rte_ctx *ctx)
{
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct LaunchTimer {
    rte_ctx *ctx;
    int kind;
    hipEvent_t a = nullptr;
    int group = -1;  // propagation kernel: group size k (0: pairs across media), -1 other
    LaunchTimer(rte_ctx *c, int k, int grp = -1) : ctx(c), kind(k), group(grp)
    {
        if (ctx->prof && (a = pool_event(ctx))) (void)hipEventRecord(a, ctx->stream);
    }
    void done(double pairs = 0, double flops = 0)
    {
        if (!a) return;
        hipEvent_t b = pool_event(ctx);
        if (!b) return;
        (void)hipEventRecord(b, ctx->stream);
        ctx->pending.push_back(PendingTiming{kind, a, b, pairs, flops, group});
        a = nullptr;
    }
};

// folds completed launch timings into the counters (stream must be idle)
static void drain_timings(rte_ctx *ctx)
{
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess) {
            (void)hipGetLastError();
            ms = 0.f;
        }
        if (p.kind == 0) {
            ctx->pc.prop_launches++;
            ctx->pc.prop_ms += ms;
            ctx->pc.prop_pairs += p.pairs;
            ctx->pc.prop_flops += p.flops;
            if (p.group >= 0 && p.group <= kMaxNR) {
                ctx->pc.group_launches[p.group]++;
                ctx->pc.group_ms[p.group] += ms;
                ctx->pc.group_flops[p.group] += p.flops;
            }
        } else if (p.kind == 1) {
            ctx->pc.ktab_launches++;
            ctx->pc.ktab_ms += ms;
            ctx->pc.ktab_pairs += p.pairs;
        } else {
            ctx->pc.other_ms += ms;
        }
        ctx->ev_pool.push_back(p.a);
        ctx->ev_pool.push_back(p.b);
    }
    ctx->pending.clear();
}

// pairs (d != s) with r < thr on a grid, and the sum of |dx|+|dy|+|dz| (in
// voxels) over them: the work that defines a propagation's algorithmic flops
static void coupled_pairs(const rte_grid &g, double thr, double *pairs, double *l1)
{
    const double dx = (g.xmax - g.xmin) / g.nx, dy = (g.ymax - g.ymin) / g.ny, dz = (g.zmax - g.zmin) / g.nz;
    double P = 0, S = 0;
    for (int k = -(g.nz - 1); k <= g.nz - 1; ++k)
        for (int j = -(g.ny - 1); j <= g.ny - 1; ++j)
            for (int i = -(g.nx - 1); i <= g.nx - 1; ++i) {
                if (!i && !j && !k) continue;
                const double r2 = (i * dx) * (i * dx) + (j * dy) * (j * dy) + (k * dz) * (k * dz);
                if (thr > 0 && !(r2 < thr * thr)) continue;
                const double c = (double)(g.nx - std::abs(i)) * (g.ny - std::abs(j)) * (g.nz - std::abs(k));
                P += c;
                S += c * (std::abs(i) + std::abs(j) + std::abs(k));
            }
    *pairs = P;
    *l1 = S;
}

extern "C" int rte_version(void) { return 1; }

extern "C" const char *rte_last_error(void) { return g_err.c_str(); }

extern "C" int rte_ctx_create(int device, rte_ctx **out)
{
    if (!out) return fail(RTE_E_ARG, "rte_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(RTE_E_ARG, "rte_ctx_create: device %d of %d", device, ndev);
    auto ctx = std::make_unique<rte_ctx>();
    ctx->device = device;
    HIP_TRY(hipSetDevice(device));
    // once per device: the address of rte_propagate_noalloc's scratch array
    HIP_TRY(scratch_init(device));
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&ctx->ev0));
    HIP_TRY(hipEventCreate(&ctx->ev1));
    *out = ctx.release();
    return RTE_OK;
}

extern "C" int rte_ctx_set_group_size(rte_ctx *ctx, int max_rhs)
{
    if (!ctx) return fail(RTE_E_ARG, "rte_ctx_set_group_size: NULL");
    if (max_rhs < 1 || max_rhs > kMaxNR) return fail(RTE_E_ARG, "group size %d outside 1..%d", max_rhs, kMaxNR);
    ctx->max_rhs = max_rhs;
    return RTE_OK;
}

extern "C" int rte_ctx_profile(rte_ctx *ctx, int on)
{
    if (!ctx) return fail(RTE_E_ARG, "rte_ctx_profile: NULL");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drain_timings(ctx);
    if (on) ctx->pc = rte_prof{};
    ctx->prof = on != 0;
    return RTE_OK;
}

extern "C" int rte_ctx_get_profile(rte_ctx *ctx, rte_prof *out)
{
    if (!ctx || !out) return fail(RTE_E_ARG, "rte_ctx_get_profile: NULL");
    (void)hipSetDevice(ctx->device);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    drain_timings(ctx);
    *out = ctx->pc;
    return RTE_OK;
}

extern "C" int rte_ctx_destroy(rte_ctx *ctx)
{
    if (!ctx) return RTE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drain_timings(ctx);
    // media still attached become invalid with their context
    for (rte_medium *m : std::vector<rte_medium *>(ctx->media.begin(), ctx->media.end())) rte_medium_destroy(m);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    ctx->ev_pool.clear();
    ctx->moments.clear();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RTE_OK;
}

static int check_grid(const rte_grid *g)
{
    if (!g) return fail(RTE_E_ARG, "grid is NULL");
    if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(RTE_E_ARG, "grid dims must be >= 1");
    if ((long long)g->nx * g->ny * g->nz > (1LL << 24)) return fail(RTE_E_ARG, "grid too large");
    if (!(g->xmax > g->xmin) || !(g->ymax > g->ymin) || !(g->zmax > g->zmin))
        return fail(RTE_E_ARG, "grid extents must be increasing");
    return RTE_OK;
}

static int check_optics(const rte_grid *g, const rte_optics *o)
{
    if (!o || !o->mu_a || !o->mu_s) return fail(RTE_E_ARG, "optics maps are NULL");
    if (!(std::fabs(o->g) <= 1.0)) return fail(RTE_E_ARG, "|g| must be <= 1");
    if (!(o->n > 0.0) || !(o->c > 0.0)) return fail(RTE_E_ARG, "n and c must be > 0");
    const long long N = (long long)g->nx * g->ny * g->nz;
    for (long long v = 0; v < N; ++v) {
        const double a = o->mu_a[v], s = o->mu_s[v];
        if (!(a >= 0.0) || !(s >= 0.0) || !std::isfinite(a) || !std::isfinite(s) || !(a + s > 0.0))
            return fail(RTE_E_ARG, "invalid optics at voxel %lld (mu_a=%g, mu_s=%g)", v, a, s);
    }
    return RTE_OK;
}

template <typename T>
static GridDev<T> make_grid(const rte_grid &g)
{
    GridDev<T> d;
    d.nx = g.nx;
    d.ny = g.ny;
    d.nz = g.nz;
    d.N = g.nx * g.ny * g.nz;
    d.bx = (g.nx + kBrick - 1) / kBrick;
    d.by = (g.ny + kBrick - 1) / kBrick;
    d.bz = (g.nz + kBrick - 1) / kBrick;
    d.nblk = d.bx * d.by * d.bz;
    d.npad = d.nblk * kWave;
    const double dx = (g.xmax - g.xmin) / g.nx, dy = (g.ymax - g.ymin) / g.ny,
                 dz = (g.zmax - g.zmin) / g.nz;
    d.dx = T(dx);
    d.dy = T(dy);
    d.dz = T(dz);
    d.dV = T(dx * dy * dz);
    return d;
}

extern "C" int rte_medium_create(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics,
                                 rte_medium **out)
{
    if (!ctx || !out) return fail(RTE_E_ARG, "rte_medium_create: NULL argument");
    *out = nullptr;
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_optics(grid, optics))) return rc;
    if ((rc = set_device(ctx))) return rc;
    auto m = std::make_unique<rte_medium>();
    m->ctx = ctx;
    m->grid = *grid;
    m->g = optics->g;
    m->n = optics->n;
    m->c = optics->c;
    m->N = grid->nx * grid->ny * grid->nz;
    const size_t N = (size_t)m->N;
    m->mua_h.assign(optics->mu_a, optics->mu_a + N);
    m->mus_h.assign(optics->mu_s, optics->mu_s + N);
    std::vector<float> mt_f(N), ms_f(N);
    std::vector<double> mt_d(N);
    for (size_t v = 0; v < N; ++v) {
        mt_d[v] = m->mua_h[v] + m->mus_h[v];
        mt_f[v] = (float)mt_d[v];
        ms_f[v] = (float)m->mus_h[v];
    }
    HIP_TRY(m->mua_d.alloc(N * 8));
    HIP_TRY(m->mus_d.alloc(N * 8));
    HIP_TRY(m->mut_d.alloc(N * 8));
    HIP_TRY(m->mut_f.alloc(N * 4));
    HIP_TRY(m->mus_f.alloc(N * 4));
    HIP_TRY(hipMemcpy(m->mua_d.p, m->mua_h.data(), N * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->mus_d.p, m->mus_h.data(), N * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->mut_d.p, mt_d.data(), N * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->mut_f.p, mt_f.data(), N * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->mus_f.p, ms_f.data(), N * 4, hipMemcpyHostToDevice));
    *out = m.release();
    ctx->media.insert(*out);
    return RTE_OK;
}

extern "C" int rte_medium_destroy(rte_medium *m)
{
    if (!m) return RTE_OK;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    m->ctx->media.erase(m);
    delete m;
    return RTE_OK;
}

extern "C" int rte_medium_reset_cache(rte_medium *m)
{
    if (!m) return fail(RTE_E_ARG, "rte_medium_reset_cache: NULL");
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    m->ktab_f.release();
    m->ktab_d.release();
    m->ktab_f_thr = m->ktab_d_thr = -1;
    return RTE_OK;
}

// ---- per-(L, voxel shape) self-block moments --------------------------------
static int get_moments(rte_ctx *ctx, int L, double dx, double dy, double dz, MomentSet **out)
{
    auto key = std::make_tuple(L, dx, dy, dz);
    auto it = ctx->moments.find(key);
    if (it != ctx->moments.end()) {
        *out = it->second.get();
        return RTE_OK;
    }
    const int K = (L + 1) * (L + 1);
    std::vector<double> beta, th, ph;
    self_beta(dx, dy, dz, beta, th, ph);
    std::vector<double> R((size_t)kNOmega * K), sn(kNOmega);
    for (int w = 0; w < kNOmega; ++w) {
        const double st = std::sin(th[w]);
        real_sh(L, std::cos(ph[w]) * st, std::sin(ph[w]) * st, std::cos(th[w]), &R[(size_t)w * K]);
        sn[w] = st;
    }
    DevBuf dR, dsn, dbeta;
    auto ms = std::make_unique<MomentSet>();
    HIP_TRY(dR.alloc(R.size() * 8));
    HIP_TRY(dsn.alloc(sn.size() * 8));
    HIP_TRY(dbeta.alloc(beta.size() * 8));
    HIP_TRY(ms->md.alloc((size_t)kNMom * K * K * 8));
    HIP_TRY(ms->mf.alloc((size_t)kNMom * K * K * 4));
    HIP_TRY(hipMemcpyAsync(dR.p, R.data(), R.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dsn.p, sn.data(), sn.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dbeta.p, beta.data(), beta.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_moments(K, dR.as<double>(), dsn.as<double>(), dbeta.as<double>(), ms->md.as<double>(),
                           ctx->stream));
    std::vector<double> Mh((size_t)kNMom * K * K);
    HIP_TRY(hipMemcpyAsync(Mh.data(), ms->md.p, Mh.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<float> Mf(Mh.size());
    for (size_t i = 0; i < Mh.size(); ++i) Mf[i] = (float)Mh[i];
    HIP_TRY(hipMemcpy(ms->mf.p, Mf.data(), Mf.size() * 4, hipMemcpyHostToDevice));
    *out = ms.get();
    ctx->moments.emplace(key, std::move(ms));
    return RTE_OK;
}

// ---- per-medium path tables ------------------------------------------------
static int pick_chunks(int nblk, int nsolve, int N)
{
    const long long target = 4096;
    long long c = (target + (long long)nblk * nsolve - 1) / ((long long)nblk * nsolve);
    long long cmax = N / 32 > 0 ? N / 32 : 1;
    if (c > cmax) c = cmax;
    if (c < 1) c = 1;
    return (int)c;
}

template <typename T>
static int ensure_ktab(rte_medium *m, double thr, hipStream_t on = nullptr)
{
    hipStream_t st = on ? on : m->ctx->stream;
    DevBuf &tab = sizeof(T) == 4 ? m->ktab_f : m->ktab_d;
    double &tthr = sizeof(T) == 4 ? m->ktab_f_thr : m->ktab_d_thr;
    if (tab.p && tthr == thr) return RTE_OK;
    const GridDev<T> g = make_grid<T>(m->grid);
    HIP_TRY(tab.alloc((size_t)g.N * g.npad * sizeof(T)));
    const T *mut = sizeof(T) == 4 ? (const T *)m->mut_f.p : (const T *)m->mut_d.p;
    const T thr2 = thr > 0 ? T(thr * thr) : T(sizeof(T) == 4 ? 1e30 : 1e300);
    if (st == m->ctx->stream) {
        LaunchTimer tm(m->ctx, 1);
        HIP_TRY(launch_ktab<T>(g, thr2, mut, tab.as<T>(), st));
        tm.done((double)g.N * g.N);
    } else {
        HIP_TRY(launch_ktab<T>(g, thr2, mut, tab.as<T>(), st));
    }
    tthr = thr;
    return RTE_OK;
}

// ---------------------------------------------------------------------------
// the Neumann driver over a batch of solves of one shape
// ---------------------------------------------------------------------------
struct HostSolve {
    rte_medium *m;
    int state;  // 0 term, 2 final, 3 done
    int n;
    double T, sumOut, fi, lam, rel, ball;
    double Ti = 0.0;  // imaginary part of the total (frequency-domain solves)
};

static Scales make_scales(int L)
{
    Scales sc;
    for (int a = 0; a < kMaxK; ++a) sc.n[a] = 1.0;
    for (int a = 0; a < (L + 1) * (L + 1); ++a) sc.n[a] = sh_scale_a(a);
    return sc;
}

static Scales unit_scales()
{
    Scales sc;
    for (int a = 0; a < kMaxK; ++a) sc.n[a] = 1.0;
    return sc;
}

static double stop_weight(int a)
{
    int l = 0;
    while ((l + 1) * (l + 1) <= a) ++l;
    const int m = a - l * l - l;
    if (m == 0) return 1.0;
    if (m > 0 && (m % 2) == 0) return std::sqrt(2.0);
    return 0.0;
}

// The same sums for the imaginary half of a complex field: with
// c_lm = (r_lm - i r_l,-m)/sqrt2 and c_l,-m = (-1)^m (r_lm + i r_l,-m)/sqrt2,
// Re c_lm + Re c_l,-m = sqrt2 Im r_l,-m for odd m (CPU package/src/
// Kernels_selfsub.c:1194-1206 takes the real part of every complex coefficient).
static double stop_weight_imag(int a)
{
    int l = 0;
    while ((l + 1) * (l + 1) <= a) ++l;
    const int m = a - l * l - l;
    return (m < 0 && ((-m) % 2) == 1) ? std::sqrt(2.0) : 0.0;
}

// One term of the Neumann series has been propagated: fold its exitance
// increment and stop-rule sum into the series state and decide what comes
// next (0 another term, 2 the final exitance-only propagation, 3 done).
static void advance_series(HostSolve &h, double dT, double sumW, int rule, double tol, int max_terms,
                           double dTi = 0.0)
{
    h.T += dT;
    h.Ti += dTi;
    if (h.state == 2) {
        h.state = 3;
        return;
    }
    h.n += 1;
    // |dT| / |T| of the complex total (Ti = dTi = 0 for a continuous-wave solve)
    const double aT = std::hypot(h.T, h.Ti);
    h.rel = aT != 0.0 ? std::hypot(dT, dTi) / aT : 0.0;
    if (rule == RTE_STOP_DELTA) {
        if (h.rel < tol || h.n >= max_terms) h.state = 3;
        return;
    }
    // CPU package/src/Kernels_selfsub.c:1208-1229 at n + 1
    h.sumOut += sumW;
    bool stop;
    if (h.sumOut == 0.0) {
        stop = true;
        h.lam = 0.0;
    } else {
        const double gam = sumW / h.sumOut;
        double hi = 1.0 - gam * (1.0 / (1.0 - gam / h.fi));
        h.fi = gam;
        if (hi > 1.0 || hi < 0.0) hi = 0.0;
        h.lam = hi;
        stop = hi > 0.9999;
    }
    if (stop || h.n >= max_terms) h.state = 2;
}

// First terms given on the device instead of beam sources: the reference's
// derivative solves (one block parameter per solve, see launch_ref_source).
template <typename TA>
struct RefSrc {
    const int *part = nullptr;   // device, N: parameter of each voxel (-1 none)
    const int *param = nullptr;  // device, nb: parameter of each solve
    int wrt = 0;                 // 0 mu_a, 1 mu_s
    const TA *phi = nullptr;     // device, [N][K]: sum_q A W_q of the forward solve
    const TA *wsum = nullptr;    // device, K: stop-rule weights
    double *cplx_field = nullptr; // host, [N][K] complex: the final field in the reference's layout
};

template <typename T, typename TA>
static int run_batch(rte_ctx *ctx, int nb, rte_medium *const *media, const rte_beam *beams,
                     const rte_params *p, double *exitance, rte_stats *stats, const RefSrc<TA> *rs = nullptr,
                     double *reflectance = nullptr, double *field = nullptr, const double *kim = nullptr)
{
    // frequency-domain solves (kim != nullptr) run as two real halves each,
    // solve i = (sv[2i], sv[2i+1]) = (real, imaginary) part of its complex field
    const int hv = kim ? 2 : 1, ns = nb * hv;
    if (kim && rs) return fail(RTE_E_UNSUPP, "derivative solves are continuous-wave only");
    const rte_grid &grid = media[0]->grid;
    const int L = p->L, K = (L + 1) * (L + 1);
    const GridDev<T> g = make_grid<T>(grid);
    const int N = g.N, P = grid.nx * grid.ny;
    const double dx = (grid.xmax - grid.xmin) / grid.nx, dy = (grid.ymax - grid.ymin) / grid.ny,
                 dz = (grid.zmax - grid.zmin) / grid.nz;
    const int path = p->path_mode;
    const double thr = p->prop_thresh > 0 ? p->prop_thresh : -1.0;
    const int max_terms = p->max_terms > 0 ? p->max_terms : 1000;
    hipStream_t st = ctx->stream;

    const auto t_start = std::chrono::steady_clock::now();
    MomentSet *mom = nullptr;
    int rc;
    if ((rc = get_moments(ctx, L, dx, dy, dz, &mom))) return rc;
    double cp_pairs = 0, cp_l1 = 0;
    coupled_pairs(grid, thr, &cp_pairs, &cp_l1);
    // algorithmic flops per pair: 4K for the projection into and out of the
    // pair direction, 16 for r^2, rsqrt, 1/r^2 and scaling, plus 2 per voxel
    // crossed when the path integral is marched in the term
    // (beams sharing one medium in a launch share the 16 direction flops)
    const double flops_dir = cp_pairs * 16.0 + (path == RTE_PATH_MARCH ? 2.0 * cp_l1 : 0.0);
    const double flops_rhs = cp_pairs * 4.0 * K;
    // beams on one medium are propagated together on the mixed cached path
    const bool multi = sizeof(T) == 4 && sizeof(TA) == 8 && path == RTE_PATH_CACHED;
    std::vector<double> lat(nb, 0.0);

    HIP_TRY(hipEventRecord(ctx->ev0, st));
    if (path == RTE_PATH_CACHED)
        for (int i = 0; i < nb; ++i)
            if ((rc = ensure_ktab<T>(media[i], thr))) return rc;

    // the source chunking depends on the grid only, so a solve's partial sums,
    // and so its result bits, are the same alone and in any batch
    const int nchunk = pick_chunks(g.nblk, 1, N);
    const int clen = (N + nchunk - 1) / nchunk;
    const int nch = (N + clen - 1) / clen;
    const size_t szW = (size_t)N * K, szPart = (size_t)nch * N * K;
    HIP_TRY(ctx->ws_W.alloc(szW * sizeof(T) * ns));
    HIP_TRY(ctx->ws_Wp.alloc(szW * sizeof(T) * ns));
    HIP_TRY(ctx->ws_part.alloc(szPart * sizeof(T) * ns));
    HIP_TRY(ctx->ws_pix.alloc((size_t)kPixBufs * P * sizeof(TA) * ns));
    HIP_TRY(ctx->ws_red.alloc((size_t)2 * sizeof(TA) * ns));
    HIP_TRY(ctx->ws_sv.alloc(sizeof(SolveDev<T, TA>) * ns * 2));
    HIP_TRY(ctx->ws_beams.alloc(sizeof(BeamDev) * nb));

    std::vector<SolveDev<T, TA>> sv(ns);
    std::vector<BeamDev> bd(nb);
    std::vector<HostSolve> hs(nb);
    for (int i = 0; i < nb; ++i) {
        rte_medium *m = media[i];
        for (int hh = 0; hh < hv; ++hh) {
            const int q = i * hv + hh;
            SolveDev<T, TA> &s = sv[q];
            s.W = ctx->ws_W.as<T>() + szW * q;
            s.Wp = ctx->ws_Wp.as<T>() + szW * q;
            s.part = ctx->ws_part.as<T>() + szPart * q;
            s.ktab = sizeof(T) == 4 ? (const T *)m->ktab_f.p : (const T *)m->ktab_d.p;
            s.mut = sizeof(T) == 4 ? (const T *)m->mut_f.p : (const T *)m->mut_d.p;
            s.mus = sizeof(T) == 4 ? (const T *)m->mus_f.p : (const T *)m->mus_d.p;
            s.g = T(m->g);
            set_pixels(s, ctx->ws_pix.as<TA>() + (size_t)kPixBufs * P * q, P);
            s.red = ctx->ws_red.as<TA>() + 2 * q;
            s.mode = 1;
            s.kim = kim ? T(kim[i]) : T(0);
        }
        hs[i] = HostSolve{m, 0, 0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
        if (rs) continue;
        // beam columns and power per column (CPU package/src/Dist.c:1090-1114, Inits.c:56)
        const double half = (beams[i].diam + 1.0e-16) / 2.0;
        int cols = 0;
        for (int j = 0; j < grid.ny; ++j)
            for (int ii = 0; ii < grid.nx; ++ii) {
                const double xc = (double)ii * dx + grid.xmin + dx / 2.0;
                const double yc = (double)j * dy + grid.ymin + dy / 2.0;
                const double rad = std::sqrt((xc - beams[i].pos_x) * (xc - beams[i].pos_x) +
                                             (yc - beams[i].pos_y) * (yc - beams[i].pos_y));
                if (rad <= half) ++cols;
            }
        if (cols == 0) return fail(RTE_E_ARG, "solve %d: beam hits no voxel column", i);
        bd[i] = BeamDev{beams[i].pos_x, beams[i].pos_y, half, beams[i].mag / (double)cols,
                        m->mua_d.as<double>(), m->mus_d.as<double>(), kim ? kim[i] : 0.0};
    }
    // the final field sum_q A W_q of solve 0 (both halves when complex), if asked for
    DevBuf dphi;
    const bool want_cplx = rs && rs->cplx_field;
    if (field || want_cplx) {
        HIP_TRY(dphi.alloc(szW * hv * sizeof(TA)));
        HIP_TRY(launch_fill<TA>(dphi.as<TA>(), szW * hv, TA(0), st));
        for (int hh = 0; hh < hv; ++hh) sv[hh].phi = dphi.as<TA>() + szW * hh;
    }
    // FP64 frequency-domain solves read the pair phase (cos, sin)(kim r) from a
    // table over the displacements, one per distinct kim, filled with the C
    // library's sine and cosine (the FP32 kernel evaluates it in hardware)
    std::vector<std::unique_ptr<DevBuf>> dphase;
    if (kim && sizeof(T) == 8) {
        const int ex = 2 * grid.nx - 1, ey = 2 * grid.ny - 1, ez = 2 * grid.nz - 1;
        const size_t nd = (size_t)ex * ey * ez;
        std::vector<double> done, tab(2 * nd);
        for (int i = 0; i < nb; ++i) {
            size_t t = 0;
            while (t < done.size() && done[t] != kim[i]) ++t;
            if (t == done.size()) {
                size_t j = 0;
                for (int iz = 0; iz < ez; ++iz)
                    for (int iy = 0; iy < ey; ++iy)
                        for (int ix = 0; ix < ex; ++ix, ++j) {
                            const double rx = (ix - (grid.nx - 1)) * dx, ry = (iy - (grid.ny - 1)) * dy,
                                         rz = (iz - (grid.nz - 1)) * dz;
                            const double ph = kim[i] * std::sqrt(rx * rx + ry * ry + rz * rz);
                            tab[2 * j] = std::cos(ph);
                            tab[2 * j + 1] = std::sin(ph);
                        }
                dphase.emplace_back(new DevBuf);
                HIP_TRY(dphase.back()->alloc(sizeof(double) * tab.size()));
                HIP_TRY(hipMemcpy(dphase.back()->p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
                done.push_back(kim[i]);
            }
            for (int hh = 0; hh < hv; ++hh) sv[i * hv + hh].phase = (const T *)dphase[t]->p;
        }
    }
    SolveDev<T, TA> *svd = ctx->ws_sv.as<SolveDev<T, TA>>();
    HIP_TRY(hipMemcpyAsync(svd, sv.data(), sizeof(SolveDev<T, TA>) * ns, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->ws_beams.p, bd.data(), sizeof(BeamDev) * nb, hipMemcpyHostToDevice, st));
    const Scales sc = make_scales(L);
    {
        LaunchTimer tm(ctx, 2);
        if (rs) {
            if constexpr (sizeof(TA) == 8) {
                HIP_TRY(launch_ref_source<T, TA>(L, g, sc, rs->part, rs->param, rs->wrt, TA(T(media[0]->g)),
                                                 TA(media[0]->n), rs->phi, rs->wsum, svd, nb, st));
            } else {
                return fail(RTE_E_UNSUPP, "derivative solves need FP64 running sums");
            }
        } else {
            HIP_TRY(launch_fill<T>(ctx->ws_W.as<T>(), szW * ns, T(0), st));
            if (kim)
                HIP_TRY(launch_source_c<T, TA>(L, g, dx, dy, dz, grid.xmin, grid.ymin, sc,
                                               ctx->ws_beams.as<BeamDev>(), svd, nb, st));
            else
                HIP_TRY(launch_source<T, TA>(L, g, dx, dy, dz, grid.xmin, grid.ymin, sc,
                                             ctx->ws_beams.as<BeamDev>(), svd, nb, st));
        }
        HIP_TRY(launch_bottom<T, TA>(P, svd, ns, st));
        tm.done();
    }
    ctx->host_red.resize(sizeof(TA) * 2 * ns);
    TA *red = reinterpret_cast<TA *>(ctx->host_red.data());
    HIP_TRY(hipMemcpyAsync(red, ctx->ws_red.p, sizeof(TA) * 2 * ns, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < nb; ++i) {
        HostSolve &h = hs[i];
        const int q = i * hv;
        h.ball = (double)red[2 * q];
        h.T = h.ball;
        h.Ti = hv == 2 ? (double)red[2 * q + 2] : 0.0;
        if (p->stop_rule == RTE_STOP_LAMBDA) {
            h.n = 1;
            h.sumOut = (double)red[2 * q + 1] + (hv == 2 ? (double)red[2 * q + 3] : 0.0);
            h.state = (max_terms <= 1) ? 2 : 0;
        } else {
            h.n = 0;
            h.state = 0;
        }
    }

    PropArgs<T> pa;
    pa.g = g;
    pa.nchunk = nch;
    pa.chunk_len = clen;
    pa.thr2 = thr > 0 ? T(thr * thr) : T(sizeof(T) == 4 ? 1e30 : 1e300);
    CombArgs<T, TA> ca;
    std::memset(&ca, 0, sizeof ca);
    ca.g = g;
    ca.nchunk = nch;
    ca.op = 0;
    ca.mom = sizeof(TA) == 4 ? (const TA *)mom->mf.p : (const TA *)mom->md.p;
    ca.momT = sizeof(T) == 4 ? (const T *)mom->mf.p : (const T *)mom->md.p;
    ca.self_scale = TA((M_PI / kAngRes) * (2.0 * M_PI / kAngRes) / (dx * dy * dz));
    ca.dV = TA(dx * dy * dz);
    {
        double cE[kMaxL + 1], cR[kMaxL + 1];
        face_coefs(L, p->exit_quad, dx, dy, cE, cR);
        for (int l = 0; l <= L; ++l) {
            ca.coefE[l] = TA(cE[l]);
            ca.coefR[l] = TA(cR[l]);
        }
        for (int a = 0; a < K; ++a) {
            ca.wsum[a] = TA(stop_weight(a));
            ca.wsumI[a] = TA(stop_weight_imag(a));
        }
        for (int a = 0; a < kMaxK; ++a) ca.nsc[a] = TA(sc.n[a]);
    }

    std::vector<int> act;
    std::vector<SolveDev<T, TA>> svact, svgrp;
    SolveDev<T, TA> *svd_grp = svd + nb;
    for (;;) {
        act.clear();
        for (int i = 0; i < nb; ++i)
            if (hs[i].state != 3) act.push_back(i);
        if (act.empty()) break;
        svact.resize(act.size() * hv);
        for (size_t j = 0; j < act.size(); ++j)
            for (int hh = 0; hh < hv; ++hh) {
                svact[j * hv + hh] = sv[act[j] * hv + hh];
                svact[j * hv + hh].mode = hs[act[j]].state;
            }
        const int na = (int)act.size();
        HIP_TRY(hipMemcpyAsync(svd, svact.data(), sizeof(SolveDev<T, TA>) * na * hv, hipMemcpyHostToDevice, st));
        if (kim) {
            LaunchTimer tm(ctx, 0);
            HIP_TRY(launch_prop_c<T, TA>(L, path, pa, svd, na, st));
            tm.done(cp_pairs * 2 * na, (flops_dir + 2 * flops_rhs) * na);
        } else if (!multi) {
            LaunchTimer tm(ctx, 0, 1);
            HIP_TRY(launch_prop<T, TA>(L, path, 1, pa, svd, na, st));
            tm.done(cp_pairs * na, (flops_dir + flops_rhs) * na);
        } else {
            // groups of up to kMaxNR active solves on the same medium, launched
            // by group size (the arithmetic of a solve does not depend on it)
            std::vector<std::vector<int>> by_size(kMaxNR + 1);  // indices into svact
            std::vector<char> taken(na, 0);
            for (int j = 0; j < na; ++j) {
                if (taken[j]) continue;
                std::vector<int> same;
                for (int q = j; q < na; ++q)
                    if (!taken[q] && svact[q].ktab == svact[j].ktab) same.push_back(q);
                // above L = 5 a group's accumulators outgrow the register file: at most
                // 3 at L = 6 and 2 from L = 7 (measured fastest at 24^3)
                const size_t gs = (size_t)std::min(ctx->max_rhs, L <= 5 ? kMaxNR : (L == 6 ? 3 : 2));
                for (size_t o = 0; o < same.size(); o += gs) {
                    const size_t e = std::min(same.size(), o + gs);
                    for (size_t q = o; q < e; ++q) {
                        by_size[e - o].push_back(same[q]);
                        taken[same[q]] = 1;
                    }
                }
            }
            // solves left alone on their medium go in pairs across media: they share
            // the pair geometry and read their own tables (their bits do not change)
            std::vector<int> sep;
            const size_t npair = by_size[1].size() / 2;
            sep.assign(by_size[1].begin(), by_size[1].begin() + 2 * npair);
            by_size[1].erase(by_size[1].begin(), by_size[1].begin() + 2 * npair);
            svgrp.clear();
            std::vector<std::pair<int, int>> launches;  // (nr, offset)
            for (int k = kMaxNR; k >= 1; --k) {
                if (by_size[k].empty()) continue;
                launches.emplace_back(k, (int)svgrp.size());
                for (int q : by_size[k]) svgrp.push_back(svact[q]);
            }
            const int sep_off = (int)svgrp.size();
            for (int q : sep) svgrp.push_back(svact[q]);
            HIP_TRY(hipMemcpyAsync(svd_grp, svgrp.data(), sizeof(SolveDev<T, TA>) * svgrp.size(),
                                   hipMemcpyHostToDevice, st));
            for (auto &ln : launches) {
                const int k = ln.first, ng = (int)by_size[k].size() / k;
                LaunchTimer tm(ctx, 0, k);
                HIP_TRY(launch_prop<T, TA>(L, path, k, pa, svd_grp + ln.second, ng, st));
                tm.done(cp_pairs * k * ng, (flops_dir + k * flops_rhs) * ng);
            }
            if (npair) {
                LaunchTimer tm(ctx, 0, 0);
                HIP_TRY(launch_prop_sep<T, TA>(L, 2, pa, svd_grp + sep_off, (int)npair, st));
                tm.done(cp_pairs * 2 * npair, (flops_dir + 2 * flops_rhs) * npair);
            }
        }
        {
            LaunchTimer tm(ctx, 2);
            if (kim)
                HIP_TRY(launch_combine_c<T, TA>(L, ca, svd, na, st));
            else
                HIP_TRY(launch_combine<T, TA>(L, ca, svd, na, st));
            HIP_TRY(launch_bottom<T, TA>(P, svd, na * hv, st));
            tm.done();
        }
        std::vector<TA> r(2 * ns);
        HIP_TRY(hipMemcpyAsync(r.data(), ctx->ws_red.p, sizeof(TA) * 2 * ns, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const double now_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        for (int j = 0; j < na; ++j) {
            HostSolve &h = hs[act[j]];
            const int q = act[j] * hv;
            const double dT = (double)r[2 * q];
            const double sumW = (double)r[2 * q + 1] + (hv == 2 ? (double)r[2 * q + 3] : 0.0);
            const double dTi = hv == 2 ? (double)r[2 * q + 2] : 0.0;
            lat[act[j]] = now_ms;
            advance_series(h, dT, sumW, p->stop_rule, p->tol, max_terms, dTi);
        }
        // the rows of svd hold the active solves; stale entries are rebuilt next pass
    }
    HIP_TRY(hipEventRecord(ctx->ev1, st));
    // gather exitance images (a complex solve: real plane, then imaginary plane)
    std::vector<TA> img((size_t)P);
    for (int q0 = 0; q0 < ns; ++q0) {
        HIP_TRY(hipMemcpyAsync(img.data(), sv[q0].E, sizeof(TA) * P, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int q = 0; q < P; ++q) exitance[(size_t)q0 * P + q] = (double)img[q];
        if (reflectance) {
            HIP_TRY(hipMemcpyAsync(img.data(), sv[q0].R, sizeof(TA) * P, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (int q = 0; q < P; ++q) reflectance[(size_t)q0 * P + q] = (double)img[q];
        }
    }
    if (want_cplx) {
        // the reference's complex basis and layout, converted on the device
        DevBuf dC;
        HIP_TRY(dC.alloc(sizeof(double) * 2 * szW));
        HIP_TRY(launch_real_to_cplx<TA>(L, N, media[0]->n / media[0]->c, dphi.as<TA>(), dC.as<double>(), st));
        HIP_TRY(hipMemcpyAsync(rs->cplx_field, dC.p, sizeof(double) * 2 * szW, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (field) {
        // coefficient-major, in the reference's units (A carries 1/c_m there)
        std::vector<TA> f(szW * hv);
        HIP_TRY(hipMemcpyAsync(f.data(), dphi.p, sizeof(TA) * szW * hv, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const double inv_cm = media[0]->n / media[0]->c;
        for (int hh = 0; hh < hv; ++hh)
            for (int v = 0; v < N; ++v)
                for (int a = 0; a < K; ++a)
                    field[szW * hh + (size_t)a * N + v] = (double)f[szW * hh + (size_t)v * K + a] * inv_cm;
    }
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (stats) {
        for (int i = 0; i < nb; ++i) {
            rte_stats &s = stats[i];
            s.terms = hs[i].n;
            s.lambda_last = hs[i].lam;
            s.rel_change_last = hs[i].rel;
            s.total_exitance = hs[i].T;
            s.ballistic_total = hs[i].ball;
            s.gpu_ms = ms;
            s.latency_ms = lat[i];
        }
    }
    if (ctx->prof) {
        ctx->pc.path_mode = path;
        ctx->pc.L = L;
        ctx->pc.precision = sizeof(T) == 8 ? RTE_PREC_FP64 : (sizeof(TA) == 8 ? RTE_PREC_MIXED : RTE_PREC_FP32);
        drain_timings(ctx);
    }
    return RTE_OK;
}

static int check_params(const rte_params *p)
{
    if (!p) return fail(RTE_E_ARG, "params is NULL");
    if (p->L < 0 || p->L > RTE_MAX_L) return fail(RTE_E_UNSUPP, "L=%d outside 0..%d", p->L, RTE_MAX_L);
    if (p->stop_rule != RTE_STOP_LAMBDA && p->stop_rule != RTE_STOP_DELTA)
        return fail(RTE_E_ARG, "unknown stop_rule %d", p->stop_rule);
    if (p->stop_rule == RTE_STOP_DELTA && !(p->tol >= 0.0)) return fail(RTE_E_ARG, "tol must be >= 0");
    if (p->precision < RTE_PREC_FP64 || p->precision > RTE_PREC_FP32)
        return fail(RTE_E_ARG, "unknown precision %d", p->precision);
    if (p->exit_quad != RTE_QUAD_REF && p->exit_quad != RTE_QUAD_EXACT)
        return fail(RTE_E_ARG, "unknown exit_quad %d", p->exit_quad);
    if (p->path_mode != RTE_PATH_CACHED && p->path_mode != RTE_PATH_MARCH)
        return fail(RTE_E_ARG, "unknown path_mode %d", p->path_mode);
    return RTE_OK;
}

extern "C" int rte_solve_images(rte_ctx *ctx, int nb, rte_medium *const *media, const rte_beam *beams,
                                const rte_params *params, double *exitance, double *reflectance, rte_stats *stats)
{
    if (!ctx || !media || !beams || !exitance) return fail(RTE_E_ARG, "rte_solve_media: NULL argument");
    if (nb < 1) return fail(RTE_E_ARG, "rte_solve_media: nb < 1");
    int rc;
    if ((rc = check_params(params))) return rc;
    for (int i = 0; i < nb; ++i) {
        if (!media[i]) return fail(RTE_E_ARG, "media[%d] is NULL", i);
        if (media[i]->ctx != ctx) return fail(RTE_E_ARG, "media[%d] belongs to another context", i);
        const rte_grid &a = media[i]->grid, &b = media[0]->grid;
        if (a.nx != b.nx || a.ny != b.ny || a.nz != b.nz || a.xmin != b.xmin || a.xmax != b.xmax ||
            a.ymin != b.ymin || a.ymax != b.ymax || a.zmin != b.zmin || a.zmax != b.zmax)
            return fail(RTE_E_ARG, "media[%d] has a different grid than media[0]", i);
        if (!std::isfinite(beams[i].pos_x) || !std::isfinite(beams[i].pos_y) || !(beams[i].diam >= 0.0))
            return fail(RTE_E_ARG, "beam %d invalid", i);
    }
    if ((rc = set_device(ctx))) return rc;
    switch (params->precision) {
    case RTE_PREC_FP64:
        return run_batch<double, double>(ctx, nb, media, beams, params, exitance, stats, nullptr, reflectance);
    case RTE_PREC_MIXED:
        return run_batch<float, double>(ctx, nb, media, beams, params, exitance, stats, nullptr, reflectance);
    default:
        return run_batch<float, float>(ctx, nb, media, beams, params, exitance, stats, nullptr, reflectance);
    }
}

extern "C" int rte_solve_media(rte_ctx *ctx, int nb, rte_medium *const *media, const rte_beam *beams,
                               const rte_params *params, double *exitance, rte_stats *stats)
{
    return rte_solve_images(ctx, nb, media, beams, params, exitance, nullptr, stats);
}

extern "C" int rte_solve_field(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                               double *field, double *exitance, double *reflectance, rte_stats *stats)
{
    if (!ctx || !m || !beam || !field || !exitance) return fail(RTE_E_ARG, "rte_solve_field: NULL argument");
    if (m->ctx != ctx) return fail(RTE_E_ARG, "rte_solve_field: medium belongs to another context");
    int rc;
    if ((rc = check_params(params))) return rc;
    if (!std::isfinite(beam->pos_x) || !std::isfinite(beam->pos_y) || !(beam->diam >= 0.0))
        return fail(RTE_E_ARG, "rte_solve_field: invalid beam");
    if ((rc = set_device(ctx))) return rc;
    rte_medium *const media[1] = {m};
    switch (params->precision) {
    case RTE_PREC_FP64:
        return run_batch<double, double>(ctx, 1, media, beam, params, exitance, stats, nullptr, reflectance, field);
    case RTE_PREC_MIXED:
        return run_batch<float, double>(ctx, 1, media, beam, params, exitance, stats, nullptr, reflectance, field);
    default:
        return run_batch<float, float>(ctx, 1, media, beam, params, exitance, stats, nullptr, reflectance, field);
    }
}

// The series from a given first term (FisherInfo_IO/src/neumann.cu:23-120, flag 1):
// the source enters as the derivative solves' first term does (k_ref_source, wrt 2:
// factor 1 on every voxel); the first term is the same in the reference's units and
// this library's (A carries 1/c_m there and D carries c_m), so it is uploaded as given.
template <typename T, typename TA>
static int neumann_run(rte_ctx *ctx, rte_medium *m, const rte_params *p, const double *src, double *field,
                       double *exitance, double *reflectance, rte_stats *stats)
{
    const int L = p->L, K = (L + 1) * (L + 1), N = m->N, P = m->grid.nx * m->grid.ny;
    hipStream_t st = ctx->stream;
    std::vector<TA> vm((size_t)N * K);  // voxel-major, as k_ref_source reads phi
    for (int a = 0; a < K; ++a)
        for (int v = 0; v < N; ++v) vm[(size_t)v * K + a] = TA(src[(size_t)a * N + v]);
    std::vector<TA> ws(K);
    for (int a = 0; a < K; ++a) ws[a] = TA(stop_weight(a));
    const int param0 = 0;
    DevBuf dSrc, dParam, dWs;
    HIP_TRY(dSrc.alloc(sizeof(TA) * vm.size()));
    HIP_TRY(dParam.alloc(sizeof(int)));
    HIP_TRY(dWs.alloc(sizeof(TA) * K));
    HIP_TRY(hipMemcpyAsync(dSrc.p, vm.data(), sizeof(TA) * vm.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dParam.p, &param0, sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dWs.p, ws.data(), sizeof(TA) * K, hipMemcpyHostToDevice, st));
    RefSrc<TA> rs;
    rs.part = nullptr;
    rs.param = dParam.as<int>();
    rs.wrt = 2;
    rs.phi = dSrc.as<TA>();
    rs.wsum = dWs.as<TA>();
    std::vector<double> img;
    if (!exitance) {
        img.resize(P);
        exitance = img.data();
    }
    rte_medium *const media[1] = {m};
    int rc = run_batch<T, TA>(ctx, 1, media, nullptr, p, exitance, stats, &rs, reflectance, field);
    HIP_TRY(hipStreamSynchronize(st));  // the uploads' host buffers end here
    return rc;
}

extern "C" int rte_neumann(rte_ctx *ctx, rte_medium *m, const rte_params *params, const double *src,
                           double *field, double *exitance, double *reflectance, rte_stats *stats)
{
    if (!ctx || !m || !src || !field) return fail(RTE_E_ARG, "rte_neumann: NULL argument");
    if (m->ctx != ctx) return fail(RTE_E_ARG, "rte_neumann: medium belongs to another context");
    int rc;
    if ((rc = check_params(params))) return rc;
    if (params->precision != RTE_PREC_FP64 && params->precision != RTE_PREC_MIXED)
        return fail(RTE_E_UNSUPP, "rte_neumann: precision must be RTE_PREC_FP64 or RTE_PREC_MIXED");
    const size_t n = (size_t)(params->L + 1) * (params->L + 1) * m->N;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(src[i])) return fail(RTE_E_ARG, "rte_neumann: src[%zu] is not finite", i);
    if ((rc = set_device(ctx))) return rc;
    if (params->precision == RTE_PREC_FP64)
        return neumann_run<double, double>(ctx, m, params, src, field, exitance, reflectance, stats);
    return neumann_run<float, double>(ctx, m, params, src, field, exitance, reflectance, stats);
}

template <typename T, typename TA>
static int propagate_impl(rte_ctx *ctx, rte_medium *m, int L, double prop_thresh, int path,
                          const double *W, double *out, bool cplx = false);

// Neumann(src, out, flag) on the reference's own field layout: the complex field
// is converted to the real basis on the device, run through the series of
// neumann_run (flag 1) or one propagation (flag 0), and converted back there.
template <typename T, typename TA>
static int neumann_ref_run(rte_ctx *ctx, rte_medium *m, const rte_params *p, const double *src, double *out,
                           rte_stats *stats)
{
    const int L = p->L, K = (L + 1) * (L + 1), N = m->N, P = m->grid.nx * m->grid.ny;
    hipStream_t st = ctx->stream;
    std::vector<TA> ws(K);
    for (int a = 0; a < K; ++a) ws[a] = TA(stop_weight(a));
    const int param0 = 0;
    DevBuf dC, dSrc, dParam, dWs;
    HIP_TRY(dC.alloc(sizeof(double) * 2 * K * (size_t)N));
    HIP_TRY(dSrc.alloc(sizeof(TA) * K * (size_t)N));
    HIP_TRY(dParam.alloc(sizeof(int)));
    HIP_TRY(dWs.alloc(sizeof(TA) * K));
    HIP_TRY(hipMemcpyAsync(dC.p, src, sizeof(double) * 2 * K * (size_t)N, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dParam.p, &param0, sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dWs.p, ws.data(), sizeof(TA) * K, hipMemcpyHostToDevice, st));
    // the first term is the same in the reference's units and this library's
    HIP_TRY(launch_cplx_to_real<TA>(L, N, 1.0, unit_scales(), dC.as<double>(), dSrc.as<TA>(), st));
    RefSrc<TA> rs;
    rs.part = nullptr;
    rs.param = dParam.as<int>();
    rs.wrt = 2;
    rs.phi = dSrc.as<TA>();
    rs.wsum = dWs.as<TA>();
    rs.cplx_field = out;
    std::vector<double> img(P);
    rte_medium *const media[1] = {m};
    int rc = run_batch<T, TA>(ctx, 1, media, nullptr, p, img.data(), stats, &rs);
    HIP_TRY(hipStreamSynchronize(st));  // the uploads' host buffers end here
    return rc;
}

extern "C" int rte_neumann_ref(rte_ctx *ctx, rte_medium *m, const rte_params *params, const double *src,
                               double *out, int flag, rte_stats *stats)
{
    if (!ctx || !m || !out) return fail(RTE_E_ARG, "rte_neumann_ref: NULL argument");
    if (m->ctx != ctx) return fail(RTE_E_ARG, "rte_neumann_ref: medium belongs to another context");
    if (flag != 0 && flag != 1) return fail(RTE_E_ARG, "rte_neumann_ref: flag must be 0 or 1");
    if (flag == 1 && !src) return fail(RTE_E_ARG, "rte_neumann_ref: flag 1 needs src");
    int rc;
    if ((rc = check_params(params))) return rc;
    if (params->precision != RTE_PREC_FP64 && params->precision != RTE_PREC_MIXED)
        return fail(RTE_E_UNSUPP, "rte_neumann_ref: precision must be RTE_PREC_FP64 or RTE_PREC_MIXED");
    const int L = params->L, K = (L + 1) * (L + 1), N = m->N;
    const double *f = flag == 1 ? src : out;
    // a real radiance: Im c_l0 = 0, c_l,-m = (-1)^m conj c_lm (to rounding)
    double big = 0.0, bad = 0.0;
    for (size_t i = 0; i < (size_t)2 * K * N; ++i) {
        if (!std::isfinite(f[i])) return fail(RTE_E_ARG, "rte_neumann_ref: field value %zu is not finite", i);
        big = std::max(big, std::fabs(f[i]));
    }
    for (int v = 0; v < N; ++v)
        for (int l = 0; l <= L; ++l) {
            const double *c = f + (size_t)v * K * 2;
            bad = std::max(bad, std::fabs(c[2 * (l * l + l) + 1]));
            for (int mm = 1; mm <= l; ++mm) {
                const double sg = (mm & 1) ? -1.0 : 1.0;
                const int ap = l * l + l + mm, an = l * l + l - mm;
                bad = std::max(bad, std::fabs(c[2 * an] - sg * c[2 * ap]));
                bad = std::max(bad, std::fabs(c[2 * an + 1] + sg * c[2 * ap + 1]));
            }
        }
    if (bad > 1e-12 * big)
        return fail(RTE_E_ARG, "rte_neumann_ref: the field is not that of a real radiance "
                               "(c_l,-m != (-1)^m conj c_lm by %g of max %g)", bad, big);
    if (flag == 1 && out != src)
        for (size_t i = 0; i < (size_t)2 * K * N; ++i)
            if (out[i] != src[i])
                return fail(RTE_E_ARG, "rte_neumann_ref: flag 1 needs out preset to src, as the reference's "
                                       "callers do");
    if ((rc = set_device(ctx))) return rc;
    if (flag == 0) {
        std::vector<double> in(f, f + (size_t)2 * K * N);
        const double thr = params->prop_thresh;
        if (params->precision == RTE_PREC_FP64)
            rc = propagate_impl<double, double>(ctx, m, L, thr, params->path_mode, in.data(), out, true);
        else
            rc = propagate_impl<float, double>(ctx, m, L, thr, params->path_mode, in.data(), out, true);
        if (rc == RTE_OK && stats) std::memset(stats, 0, sizeof *stats);  // no series ran
        return rc;
    }
    if (params->precision == RTE_PREC_FP64) return neumann_ref_run<double, double>(ctx, m, params, src, out, stats);
    return neumann_ref_run<float, double>(ctx, m, params, src, out, stats);
}

extern "C" int rte_solve_fd(rte_ctx *ctx, int nb, rte_medium *const *media, const rte_beam *beams,
                            const double *omega, const rte_params *params, double *exitance, double *reflectance,
                            double *field, rte_stats *stats)
{
    if (!ctx || !media || !beams || !omega || !exitance) return fail(RTE_E_ARG, "rte_solve_fd: NULL argument");
    if (nb < 1) return fail(RTE_E_ARG, "rte_solve_fd: nb < 1");
    int rc;
    if ((rc = check_params(params))) return rc;
    std::vector<double> kim(nb);
    for (int i = 0; i < nb; ++i) {
        if (!media[i]) return fail(RTE_E_ARG, "media[%d] is NULL", i);
        if (media[i]->ctx != ctx) return fail(RTE_E_ARG, "media[%d] belongs to another context", i);
        const rte_grid &a = media[i]->grid, &b = media[0]->grid;
        if (a.nx != b.nx || a.ny != b.ny || a.nz != b.nz || a.xmin != b.xmin || a.xmax != b.xmax ||
            a.ymin != b.ymin || a.ymax != b.ymax || a.zmin != b.zmin || a.zmax != b.zmax)
            return fail(RTE_E_ARG, "media[%d] has a different grid than media[0]", i);
        if (!std::isfinite(beams[i].pos_x) || !std::isfinite(beams[i].pos_y) || !(beams[i].diam >= 0.0))
            return fail(RTE_E_ARG, "beam %d invalid", i);
        if (!std::isfinite(omega[i])) return fail(RTE_E_ARG, "omega[%d] is not finite", i);
        // CPU package/src/RTE_uniform.c:126: mu_a + i modulation / (c / n)
        kim[i] = omega[i] * media[i]->n / media[i]->c;
    }
    if ((rc = set_device(ctx))) return rc;
    switch (params->precision) {
    case RTE_PREC_FP64:
        return run_batch<double, double>(ctx, nb, media, beams, params, exitance, stats, nullptr, reflectance, field,
                                         kim.data());
    case RTE_PREC_MIXED:
        return run_batch<float, double>(ctx, nb, media, beams, params, exitance, stats, nullptr, reflectance, field,
                                        kim.data());
    default:
        return run_batch<float, float>(ctx, nb, media, beams, params, exitance, stats, nullptr, reflectance, field,
                                       kim.data());
    }
}

extern "C" int rte_solve(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics,
                         const rte_beam *beam, const rte_params *params, double *exitance,
                         rte_stats *stats)
{
    rte_medium *m = nullptr;
    int rc = rte_medium_create(ctx, grid, optics, &m);
    if (rc) return rc;
    rc = rte_solve_media(ctx, 1, &m, beam, params, exitance, stats);
    std::string keep = g_err;
    rte_medium_destroy(m);
    g_err = keep;
    return rc;
}

extern "C" long long rte_solve_footprint(const rte_grid *grid, const rte_params *params, int path_mode)
{
    if (!grid || !params || params->L < 0 || params->L > RTE_MAX_L) return -1;
    const GridDev<float> g = make_grid<float>(*grid);
    const long long K = (params->L + 1) * (params->L + 1);
    const long long es = params->precision == RTE_PREC_FP64 ? 8 : 4;
    const long long N = g.N;
    const int nch = pick_chunks(g.nblk, 1, g.N);
    long long bytes = N * 8 * 3 + N * 4 * 2;             // maps
    bytes += N * K * es * (2 + nch);                      // W, W in the monomial basis, partials
    bytes += (long long)grid->nx * grid->ny * kPixBufs * 8 + 64; // pixel buffers
    if (path_mode == RTE_PATH_CACHED) bytes += N * (long long)g.npad * es;
    return bytes;
}

// ---------------------------------------------------------------------------
// reverse mode: vector-Jacobian products of the exitance (rte_adjoint.hip)
// ---------------------------------------------------------------------------
// One forward series (A W_q kept at the checkpoints), then any number of
// backward passes, one per cotangent of the image.
template <typename T, typename TA>
struct AdjointRun {
    rte_ctx *ctx = nullptr;
    rte_medium *m = nullptr;
    const rte_params *p = nullptr;
    int stride = 1, L = 0, K = 0, N = 0, P = 0, n = 0, nch = 0, clen = 0;
    double dx = 0, dy = 0, dz = 0, dV = 0;
    size_t F = 0;
    GridDev<T> g{};
    Scales sc{};
    BeamDev bd{};
    PropArgs<T> pa{};
    CombArgs<T, TA> ca{};
    double coefE[kMaxL + 1] = {};
    HostSolve h{};
    SolveDev<T, TA> sv{};
    std::vector<TA> E;
    DevBuf dW, dWp, dS0, dPart, dPix, dRed, dSv, dBeam, dCoefE, dWq, dWqz, dBp;
    // per cotangent of a backward pass (up to kMaxNC share its propagation launches;
    // the first one's chunk partials reuse the forward's dPart)
    DevBuf dCot[kMaxNC], dX[kMaxNC], dXz[kMaxNC], dPartC[kMaxNC], dPair[kMaxNC], dGt[kMaxNC], dGs[kMaxNC],
        dGa2[kMaxNC], dGs2[kMaxNC], dLam0[kMaxNC];
    int nc_ready = 0;
    std::vector<std::unique_ptr<DevBuf>> ckpt, seg;
    bool backward_ready = false;

    int forward(rte_ctx *c, rte_medium *med, const rte_beam *beam, const rte_params *prm, int strd)
    {
        ctx = c;
        m = med;
        p = prm;
        stride = strd;
        const rte_grid &grid = m->grid;
        L = p->L;
        K = (L + 1) * (L + 1);
        g = make_grid<T>(grid);
        N = g.N;
        P = grid.nx * grid.ny;
        dx = (grid.xmax - grid.xmin) / grid.nx;
        dy = (grid.ymax - grid.ymin) / grid.ny;
        dz = (grid.zmax - grid.zmin) / grid.nz;
        dV = dx * dy * dz;
        const double thr = p->prop_thresh > 0 ? p->prop_thresh : -1.0;
        const int max_terms = p->max_terms > 0 ? p->max_terms : 1000;
        hipStream_t st = ctx->stream;
        int rc;
        MomentSet *mom = nullptr;
        if ((rc = get_moments(ctx, L, dx, dy, dz, &mom))) return rc;
        HIP_TRY(hipEventRecord(ctx->ev0, st));
        if ((rc = ensure_ktab<T>(m, thr))) return rc;
        const int nch0 = pick_chunks(g.nblk, 1, N);
        clen = (N + nch0 - 1) / nch0;
        nch = (N + clen - 1) / clen;
        F = (size_t)N * K;
        HIP_TRY(dW.alloc(F * sizeof(T)));
        HIP_TRY(dWp.alloc(F * sizeof(T)));
        HIP_TRY(dS0.alloc(F * sizeof(T)));
        HIP_TRY(dPart.alloc((size_t)nch * F * sizeof(T)));
        HIP_TRY(dPix.alloc((size_t)kPixBufs * P * sizeof(TA)));
        HIP_TRY(dRed.alloc(2 * sizeof(TA)));
        HIP_TRY(dSv.alloc(sizeof(SolveDev<T, TA>)));
        HIP_TRY(dBeam.alloc(sizeof(BeamDev)));

        const double half = (beam->diam + 1.0e-16) / 2.0;
        int cols = 0;
        for (int j = 0; j < grid.ny; ++j)
            for (int i = 0; i < grid.nx; ++i) {
                const double xc = (double)i * dx + grid.xmin + dx / 2.0, yc = (double)j * dy + grid.ymin + dy / 2.0;
                if (std::sqrt((xc - beam->pos_x) * (xc - beam->pos_x) + (yc - beam->pos_y) * (yc - beam->pos_y)) <=
                    half)
                    ++cols;
            }
        if (cols == 0) return fail(RTE_E_ARG, "beam hits no voxel column");
        bd = BeamDev{beam->pos_x, beam->pos_y, half, beam->mag / cols, m->mua_d.as<double>(), m->mus_d.as<double>()};
        std::memset(&sv, 0, sizeof sv);
        sv.W = dW.as<T>();
        sv.Wp = dWp.as<T>();
        sv.part = dPart.as<T>();
        sv.ktab = sizeof(T) == 4 ? (const T *)m->ktab_f.p : (const T *)m->ktab_d.p;
        sv.mut = sizeof(T) == 4 ? (const T *)m->mut_f.p : (const T *)m->mut_d.p;
        sv.mus = sizeof(T) == 4 ? (const T *)m->mus_f.p : (const T *)m->mus_d.p;
        sv.g = T(m->g);
        set_pixels(sv, dPix.as<TA>(), P);
        sv.red = dRed.as<TA>();
        sv.mode = 1;
        SolveDev<T, TA> *svd = dSv.as<SolveDev<T, TA>>();
        sc = make_scales(L);
        HIP_TRY(hipMemcpyAsync(svd, &sv, sizeof sv, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dBeam.p, &bd, sizeof bd, hipMemcpyHostToDevice, st));
        HIP_TRY(launch_fill<T>(dW.as<T>(), F, T(0), st));
        HIP_TRY(launch_source<T, TA>(L, g, dx, dy, dz, grid.xmin, grid.ymin, sc, dBeam.as<BeamDev>(), svd, 1, st));
        HIP_TRY(launch_bottom<T, TA>(P, svd, 1, st));
        HIP_TRY(hipMemcpyAsync(dS0.p, dW.p, F * sizeof(T), hipMemcpyDeviceToDevice, st));
        TA red[2];
        HIP_TRY(hipMemcpyAsync(red, dRed.p, sizeof red, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        h = HostSolve{m, 0, 0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
        h.ball = (double)red[0];
        h.T = h.ball;
        if (p->stop_rule == RTE_STOP_LAMBDA) {
            h.n = 1;
            h.sumOut = (double)red[1];
            h.state = (max_terms <= 1) ? 2 : 0;
        }
        pa.g = g;
        pa.nchunk = nch;
        pa.chunk_len = clen;
        pa.thr2 = thr > 0 ? T(thr * thr) : T(sizeof(T) == 4 ? 1e30 : 1e300);
        std::memset(&ca, 0, sizeof ca);
        ca.g = g;
        ca.nchunk = nch;
        ca.mom = sizeof(TA) == 4 ? (const TA *)mom->mf.p : (const TA *)mom->md.p;
        ca.momT = sizeof(T) == 4 ? (const T *)mom->mf.p : (const T *)mom->md.p;
        ca.self_scale = TA((M_PI / kAngRes) * (2.0 * M_PI / kAngRes) / dV);
        ca.dV = TA(dV);
        double cR[kMaxL + 1];
        face_coefs(L, p->exit_quad, dx, dy, coefE, cR);
        for (int l = 0; l <= L; ++l) {
            ca.coefE[l] = TA(coefE[l]);
            ca.coefR[l] = TA(cR[l]);
        }
        for (int a = 0; a < K; ++a) ca.wsum[a] = TA(stop_weight(a));
        for (int a = 0; a < kMaxK; ++a) ca.nsc[a] = TA(sc.n[a]);

        // A W_q is kept when (q + 1) % stride == 0 (every term for stride 1)
        int q = 0;
        while (h.state != 3) {
            sv.mode = h.state;
            sv.aw_store = nullptr;
            if ((q + 1) % stride == 0) {
                while (ckpt.size() <= (size_t)q) ckpt.emplace_back(nullptr);
                ckpt[q].reset(new DevBuf());
                HIP_TRY(ckpt[q]->alloc(F * sizeof(TA)));
                sv.aw_store = ckpt[q]->as<TA>();
            }
            HIP_TRY(hipMemcpyAsync(svd, &sv, sizeof sv, hipMemcpyHostToDevice, st));
            HIP_TRY(launch_prop<T, TA>(L, RTE_PATH_CACHED, 1, pa, svd, 1, st));
            HIP_TRY(launch_combine<T, TA>(L, ca, svd, 1, st));
            HIP_TRY(launch_bottom<T, TA>(P, svd, 1, st));
            HIP_TRY(hipMemcpyAsync(red, dRed.p, sizeof red, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            advance_series(h, (double)red[0], (double)red[1], p->stop_rule, p->tol, max_terms);
            ++q;
        }
        n = q;  // W_0 .. W_{n-1} entered the exitance
        E.resize(P);
        HIP_TRY(hipMemcpy(E.data(), sv.E, sizeof(TA) * P, hipMemcpyDeviceToHost));
        return RTE_OK;
    }

    int prepare_backward(int nc)
    {
        if (!backward_ready) {
            HIP_TRY(dCoefE.alloc(sizeof(TA) * (kMaxL + 1)));
            TA cE[kMaxL + 1];
            for (int l = 0; l <= kMaxL; ++l) cE[l] = l <= L ? TA(coefE[l]) : TA(0);
            HIP_TRY(hipMemcpy(dCoefE.p, cE, sizeof cE, hipMemcpyHostToDevice));
            HIP_TRY(dWq.alloc(F * sizeof(T)));
            HIP_TRY(dWqz.alloc(F * sizeof(T)));
            for (int i = 0; i + 1 < stride; ++i) {
                seg.emplace_back(new DevBuf());
                HIP_TRY(seg.back()->alloc(F * sizeof(TA)));
            }
            backward_ready = true;
        }
        for (int c = nc_ready; c < nc; ++c) {
            HIP_TRY(dCot[c].alloc(sizeof(TA) * P));
            HIP_TRY(dX[c].alloc(F * sizeof(T)));
            HIP_TRY(dXz[c].alloc(F * sizeof(T)));
            if (c > 0) HIP_TRY(dPartC[c].alloc((size_t)nch * F * sizeof(T)));
            HIP_TRY(dPair[c].alloc((size_t)N * g.npad * sizeof(T)));
            HIP_TRY(dGt[c].alloc(N * sizeof(TA)));
            HIP_TRY(dGs[c].alloc(N * sizeof(TA)));
            HIP_TRY(dGa2[c].alloc(N * sizeof(TA)));
            HIP_TRY(dGs2[c].alloc(N * sizeof(TA)));
            HIP_TRY(dLam0[c].alloc((size_t)N * (L + 1) * sizeof(TA)));
            nc_ready = c + 1;
        }
        return RTE_OK;
    }

    // Optional output of a backward pass: the back-projection's mu_t gradient summed
    // over parameter blocks (part: device [N], block of each voxel or -1) into bp[c]
    // (host [nparam]) instead of per voxel.  used reports whether it was taken (the
    // blocks' accumulators fit in LDS); then dmu_a and dmu_s leave that part out.
    struct BlockOut {
        const int *part = nullptr;
        int nparam = 0;
        double *bp[kMaxNC] = {};
        bool used = false;
    };

    // d(cot . E)/d mu_a, d mu_s for every voxel, for nc <= kMaxNC cotangents at once:
    // their adjoint series share every propagation launch
    int backward_multi(int nc, const double *const *cot_host, double *const *dmu_a, double *const *dmu_s,
                       BlockOut *blk = nullptr)
    {
        if (nc < 1 || nc > kMaxNC) return fail(RTE_E_ARG, "backward: %d cotangents per pass (1..%d)", nc, kMaxNC);
        int rc;
        if ((rc = prepare_backward(nc))) return rc;
        hipStream_t st = ctx->stream;
        const rte_grid &grid = m->grid;
        // the anisotropy exactly as the forward scatter used it (compute precision)
        const TA gan = TA(T(m->g));
        AdjCombArgs<T, TA> aa[kMaxNC];
        const T *Xc[kMaxNC];
        T *Xzc[kMaxNC], *partc[kMaxNC], *pairc[kMaxNC];
        for (int c = 0; c < nc; ++c) {
            std::vector<TA> cot(P);
            for (int i = 0; i < P; ++i) cot[i] = TA(cot_host[c][i]);
            HIP_TRY(hipMemcpyAsync(dCot[c].p, cot.data(), sizeof(TA) * P, hipMemcpyHostToDevice, st));
            HIP_TRY(hipStreamSynchronize(st));  // cot lives on this frame
            HIP_TRY(launch_fill<T>(dPair[c].as<T>(), (size_t)N * g.npad, T(0), st));
            HIP_TRY(launch_fill<TA>(dGt[c].as<TA>(), N, TA(0), st));
            HIP_TRY(launch_fill<TA>(dGs[c].as<TA>(), N, TA(0), st));
            HIP_TRY(launch_fill<TA>(dGa2[c].as<TA>(), N, TA(0), st));
            HIP_TRY(launch_fill<TA>(dGs2[c].as<TA>(), N, TA(0), st));
            HIP_TRY(launch_adj_init<T, TA>(L, g, dCot[c].as<TA>(), sc, dCoefE.as<TA>(), dX[c].as<T>(), st));
            AdjCombArgs<T, TA> &A = aa[c];
            std::memset(&A, 0, sizeof A);
            A.g = g;
            A.nchunk = nch;
            A.mom = ca.mom;
            A.momT = ca.momT;
            A.self_scale = ca.self_scale;
            A.dV = ca.dV;
            for (int l = 0; l <= kMaxL; ++l) A.coefE[l] = ca.coefE[l];
            for (int a = 0; a < kMaxK; ++a) A.nsc[a] = ca.nsc[a];
            partc[c] = c == 0 ? dPart.as<T>() : dPartC[c].as<T>();
            A.X = dX[c].as<T>();
            A.Wq = dWq.as<T>();
            A.part = partc[c];
            A.mut = sv.mut;
            A.mus = sv.mus;
            A.g_aniso = gan;
            A.cot = dCot[c].as<TA>();
            A.gt = dGt[c].as<TA>();
            A.gs = dGs[c].as<TA>();
            Xc[c] = dX[c].as<T>();
            Xzc[c] = dXz[c].as<T>();
            pairc[c] = dPair[c].as<T>();
        }
        SolveDev<T, TA> *svd = dSv.as<SolveDev<T, TA>>();
        const int nseg = (n + stride - 1) / stride;
        for (int j = nseg - 1; j >= 0; --j) {
            const int jc = j * stride, end = std::min(n, jc + stride);
            // A W_q for q in [jc, end-2]: stored (stride 1) or recomputed from W_jc
            auto aw_of = [&](int qq) -> const TA * {
                if (qq < 0) return nullptr;
                if (stride == 1 || (qq + 1) % stride == 0) return ckpt[qq]->as<TA>();
                return seg[qq - jc]->as<TA>();
            };
            if (stride > 1 && end - 1 > jc) {
                if (jc == 0)
                    HIP_TRY(hipMemcpyAsync(dW.p, dS0.p, F * sizeof(T), hipMemcpyDeviceToDevice, st));
                else
                    HIP_TRY(launch_make_wq<T, TA>(L, g, sc, gan, sv.mus, aw_of(jc - 1), dW.as<T>(), st));
                for (int qq = jc; qq <= end - 2; ++qq) {
                    SolveDev<T, TA> s2 = sv;
                    s2.mode = 3;
                    s2.aw_store = seg[qq - jc]->as<TA>();
                    HIP_TRY(hipMemcpyAsync(svd, &s2, sizeof s2, hipMemcpyHostToDevice, st));
                    HIP_TRY(launch_prop<T, TA>(L, RTE_PATH_CACHED, 1, pa, svd, 1, st));
                    HIP_TRY(launch_combine<T, TA>(L, ca, svd, 1, st));
                }
            }
            for (int qq = end - 1; qq >= jc; --qq) {
                const TA *awp = aw_of(qq - 1);
                if (qq == 0)
                    HIP_TRY(hipMemcpyAsync(dWq.p, dS0.p, F * sizeof(T), hipMemcpyDeviceToDevice, st));
                else
                    HIP_TRY(launch_make_wq<T, TA>(L, g, sc, gan, sv.mus, awp, dWq.as<T>(), st));
                HIP_TRY(launch_adj_prop<T, TA>(L, pa, nc, Xc, Xzc, dWq.as<T>(), dWqz.as<T>(), partc, pairc, sv.ktab,
                                               st));
                for (int c = 0; c < nc; ++c) {
                    aa[c].aw_prev = awp;
                    aa[c].lam0 = (qq == 0) ? dLam0[c].as<TA>() : nullptr;
                    HIP_TRY(launch_adj_combine<T, TA>(L, aa[c], st));
                }
            }
        }
        std::vector<TA> gt(N), gs(N), ga2(N), gs2(N);
        if (blk) {
            blk->used = false;
            BlockBp<T, TA> B{};
            B.part = blk->part;
            B.nparam = blk->nparam;
            B.scale = TA(-dV);
            HIP_TRY(dBp.alloc(sizeof(TA) * kMaxNC * (size_t)blk->nparam));
            HIP_TRY(launch_fill<TA>(dBp.as<TA>(), (size_t)kMaxNC * blk->nparam, TA(0), st));
            for (int c = 0; c < nc; ++c) {
                B.pairbuf[c] = dPair[c].as<T>();
                B.out[c] = dBp.as<TA>() + (size_t)c * blk->nparam;
            }
            const hipError_t e = launch_backproject_blocks<T, TA>(g, pa.thr2, B, nc, st);
            if (e == hipSuccess) {
                std::vector<TA> bp((size_t)nc * blk->nparam);
                HIP_TRY(hipMemcpyAsync(bp.data(), dBp.p, sizeof(TA) * bp.size(), hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
                for (int c = 0; c < nc; ++c)
                    for (int q = 0; q < blk->nparam; ++q) blk->bp[c][q] = (double)bp[(size_t)c * blk->nparam + q];
                blk->used = true;
            } else if (e != hipErrorInvalidValue) {
                HIP_TRY(e);
            }
        }
        for (int c = 0; c < nc; ++c) {
            if (!(blk && blk->used))
                HIP_TRY(launch_backproject<T, TA>(g, nch, clen, pa.thr2, dPair[c].as<T>(), TA(-dV), dGt[c].as<TA>(),
                                                  st));
            HIP_TRY(launch_src_grad<TA>(L, grid.nx, grid.ny, grid.nz, dx, dy, dz, grid.xmin, grid.ymin, m->g, bd,
                                        dLam0[c].as<TA>(), dCot[c].as<TA>(), dGa2[c].as<TA>(), dGs2[c].as<TA>(), st));
            HIP_TRY(hipMemcpyAsync(gt.data(), dGt[c].p, sizeof(TA) * N, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(gs.data(), dGs[c].p, sizeof(TA) * N, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(ga2.data(), dGa2[c].p, sizeof(TA) * N, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(gs2.data(), dGs2[c].p, sizeof(TA) * N, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (int v = 0; v < N; ++v) {
                if (dmu_a && dmu_a[c]) dmu_a[c][v] = (double)gt[v] + (double)ga2[v];
                if (dmu_s && dmu_s[c]) dmu_s[c][v] = (double)gt[v] + (double)gs[v] + (double)gs2[v];
            }
        }
        return RTE_OK;
    }

    int backward(const double *cot_host, double *dmu_a, double *dmu_s)
    {
        return backward_multi(1, &cot_host, &dmu_a, &dmu_s);
    }

    void fill_stats(rte_stats *stats)
    {
        if (!stats) return;
        float ms = 0.f;
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        (void)hipEventSynchronize(ctx->ev1);
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) ms = 0.f;
        std::memset(stats, 0, sizeof *stats);
        stats->terms = h.n;
        stats->lambda_last = h.lam;
        stats->rel_change_last = h.rel;
        stats->total_exitance = h.T;
        stats->ballistic_total = h.ball;
        stats->gpu_ms = ms;
    }
};

static int grad_check(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                      const rte_grad_opts *opts, int *stride)
{
    if (!ctx || !m || !beam || !params) return fail(RTE_E_ARG, "gradient: NULL argument");
    if (m->ctx != ctx) return fail(RTE_E_ARG, "gradient: medium belongs to another context");
    int rc;
    if ((rc = check_params(params))) return rc;
    if (params->path_mode != RTE_PATH_CACHED)
        return fail(RTE_E_UNSUPP, "gradient: needs the cached path tables (path_mode RTE_PATH_CACHED)");
    if (params->precision != RTE_PREC_FP64 && params->precision != RTE_PREC_MIXED)
        return fail(RTE_E_UNSUPP, "gradient: precision must be RTE_PREC_FP64 or RTE_PREC_MIXED");
    *stride = opts ? opts->checkpoint_stride : 1;
    if (*stride < 1) return fail(RTE_E_ARG, "checkpoint_stride must be >= 1");
    return set_device(ctx);
}

template <typename T, typename TA>
static int vjp_run(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *p, int stride,
                   const double *cot_in, const double *target, double *loss, double *exitance, double *dmu_a,
                   double *dmu_s, rte_stats *stats)
{
    AdjointRun<T, TA> run;
    int rc;
    if ((rc = run.forward(ctx, m, beam, p, stride))) return rc;
    std::vector<double> cot(run.P);
    double lsum = 0.0;
    for (int i = 0; i < run.P; ++i) {
        const double e = (double)run.E[i];
        if (exitance) exitance[i] = e;
        if (target) {
            lsum += (e - target[i]) * (e - target[i]);
            cot[i] = 2.0 * (e - target[i]);
        } else {
            cot[i] = cot_in[i];
        }
    }
    if (loss) *loss = lsum;
    if ((rc = run.backward(cot.data(), dmu_a, dmu_s))) return rc;
    run.fill_stats(stats);
    return RTE_OK;
}

extern "C" int rte_exitance_vjp(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                                const rte_grad_opts *opts, const double *cotangent, double *exitance, double *dmu_a,
                                double *dmu_s, rte_stats *stats)
{
    if (!cotangent) return fail(RTE_E_ARG, "rte_exitance_vjp: cotangent is NULL");
    int stride, rc;
    if ((rc = grad_check(ctx, m, beam, params, opts, &stride))) return rc;
    if (params->precision == RTE_PREC_FP64)
        return vjp_run<double, double>(ctx, m, beam, params, stride, cotangent, nullptr, nullptr, exitance, dmu_a,
                                       dmu_s, stats);
    return vjp_run<float, double>(ctx, m, beam, params, stride, cotangent, nullptr, nullptr, exitance, dmu_a, dmu_s,
                                  stats);
}

extern "C" int rte_gradient(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                            const rte_grad_opts *opts, const double *target, double *loss, double *exitance,
                            double *dmu_a, double *dmu_s, rte_stats *stats)
{
    if (!target) return fail(RTE_E_ARG, "rte_gradient: target is NULL");
    int stride, rc;
    if ((rc = grad_check(ctx, m, beam, params, opts, &stride))) return rc;
    if (params->precision == RTE_PREC_FP64)
        return vjp_run<double, double>(ctx, m, beam, params, stride, nullptr, target, loss, exitance, dmu_a, dmu_s,
                                       stats);
    return vjp_run<float, double>(ctx, m, beam, params, stride, nullptr, target, loss, exitance, dmu_a, dmu_s, stats);
}

static int gram_mfma(rte_ctx *ctx, int P, int n, const float *J, const float *w, float *G);

// ---- Fisher information ---------------------------------------------------
// The reference's forward-mode Jacobian of the exitance with respect to block
// parameters (FisherInfo_IO/src/compute_grad_rte.cu:273-301 mu_a, 341-371 mu_s;
// GPU package/src/inverse_rte.cu:290-433): one forward solve keeping
// phi = sum_q A W_q, then one Neumann series per parameter from phi restricted
// to the parameter's voxels (times -n for mu_a, n (g^l - 1) for mu_s: the
// reference's -1.0*C and (g^l - 1) C on its field phi / c_m), all
// parameters as one batch on the shared medium; column a of J is that series'
// exitance without the ballistic term.  Not the exact derivative of the
// discrete model: the reference approximates dA/dmu by -A chi A.
template <typename T, typename TA>
static int ref_jacobian_run(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *p, const int *part,
                            int nparam, int wrt, std::vector<double> &J, std::vector<double> &E, rte_stats *stats)
{
    AdjointRun<T, TA> run;
    int rc;
    if ((rc = run.forward(ctx, m, beam, p, 1))) return rc;
    const int P = run.P, N = run.N, K = run.K;
    hipStream_t st = ctx->stream;
    DevBuf dPhi, dPart, dParam, dWs;
    HIP_TRY(dPhi.alloc(run.F * sizeof(TA)));
    HIP_TRY(dPart.alloc(sizeof(int) * N));
    HIP_TRY(dParam.alloc(sizeof(int) * nparam));
    HIP_TRY(dWs.alloc(sizeof(TA) * K));
    HIP_TRY(launch_fill<TA>(dPhi.as<TA>(), run.F, TA(0), st));
    for (int q = 0; q < run.n; ++q) HIP_TRY(launch_accumulate<TA>(run.F, run.ckpt[q]->template as<TA>(), dPhi.template as<TA>(), st));
    std::vector<int> prm(nparam);
    for (int a = 0; a < nparam; ++a) prm[a] = a;
    std::vector<TA> ws(K);
    for (int a = 0; a < K; ++a) ws[a] = TA(stop_weight(a));
    HIP_TRY(hipMemcpyAsync(dPart.p, part, sizeof(int) * N, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dParam.p, prm.data(), sizeof(int) * nparam, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dWs.p, ws.data(), sizeof(TA) * K, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    RefSrc<TA> rs;
    rs.part = dPart.as<int>();
    rs.param = dParam.as<int>();
    rs.wrt = wrt;
    rs.phi = dPhi.as<TA>();
    rs.wsum = dWs.as<TA>();
    std::vector<rte_medium *> med(nparam, m);
    std::vector<double> img((size_t)nparam * P);
    if ((rc = run_batch<T, TA>(ctx, nparam, med.data(), nullptr, p, img.data(), nullptr, &rs))) return rc;
    J.assign((size_t)P * nparam, 0.0);
    for (int a = 0; a < nparam; ++a)
        for (int pix = 0; pix < P; ++pix) J[(size_t)pix * nparam + a] = img[(size_t)a * P + pix];
    E.assign(run.E.begin(), run.E.end());
    run.fill_stats(stats);
    return RTE_OK;
}

// J[p][a] = d E_p / d theta_a with theta_a = mu_a (or mu_s) on the voxels of
// partition a: exact (method 0, one backward pass per detector) or the
// reference's forward-mode derivative (method 1); F = J^T diag(1/E) J on FP32
// MFMA (detectors with E = 0 skipped, FisherInfo_IO/src/rte_fim.cu:318-324);
// lower Cholesky factor of F in FP64.
template <typename T, typename TA>
static int fisher_run(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *p, int stride,
                      int method, const int *part, int nparam, int wrt, double *Fout, double *chol, double *Jout,
                      double *exitance, rte_stats *stats)
{
    int rc;
    const int P = m->grid.nx * m->grid.ny, N = m->N;
    std::vector<double> J((size_t)P * nparam, 0.0), E(P, 0.0);
    if (method == 1) {
        if ((rc = ref_jacobian_run<T, TA>(ctx, m, beam, p, part, nparam, wrt, J, E, stats))) return rc;
    } else {
        // exact reverse mode: one adjoint series per detector, kMaxNC detectors per
        // backward pass (their series share the propagation launches)
        AdjointRun<T, TA> run;
        if ((rc = run.forward(ctx, m, beam, p, stride))) return rc;
        std::vector<int> pixels;
        for (int pix = 0; pix < P; ++pix) {
            E[pix] = (double)run.E[pix];
            if (E[pix] != 0.0) pixels.push_back(pix);
        }
        std::vector<std::vector<double>> cot(kMaxNC, std::vector<double>(P, 0.0)), ga(kMaxNC, std::vector<double>(N)),
            gs(kMaxNC, std::vector<double>(N)), bp(kMaxNC, std::vector<double>(nparam));
        // the back-projection straight into the blocks (one march for both passes)
        DevBuf dpart;
        HIP_TRY(dpart.alloc(sizeof(int) * N));
        HIP_TRY(hipMemcpy(dpart.p, part, sizeof(int) * N, hipMemcpyHostToDevice));
        typename AdjointRun<T, TA>::BlockOut blk;
        blk.part = dpart.as<int>();
        blk.nparam = nparam;
        for (size_t o = 0; o < pixels.size(); o += kMaxNC) {
            const int nc = (int)std::min<size_t>(kMaxNC, pixels.size() - o);
            const double *cp[kMaxNC];
            double *ap[kMaxNC], *sp[kMaxNC];
            for (int c = 0; c < nc; ++c) {
                cot[c].assign(P, 0.0);
                cot[c][pixels[o + c]] = 1.0;
                cp[c] = cot[c].data();
                ap[c] = ga[c].data();
                sp[c] = gs[c].data();
                blk.bp[c] = bp[c].data();
            }
            if ((rc = run.backward_multi(nc, cp, ap, sp, &blk))) return rc;
            for (int c = 0; c < nc; ++c) {
                const std::vector<double> &gv = wrt == 0 ? ga[c] : gs[c];
                const int pix = pixels[o + c];
                for (int v = 0; v < N; ++v)
                    if (part[v] >= 0) J[(size_t)pix * nparam + part[v]] += gv[v];
                if (blk.used)
                    for (int q = 0; q < nparam; ++q) J[(size_t)pix * nparam + q] += bp[c][q];
            }
        }
        run.fill_stats(stats);
    }
    if (exitance) std::memcpy(exitance, E.data(), sizeof(double) * P);
    if (Jout) std::memcpy(Jout, J.data(), sizeof(double) * J.size());
    // Gram on MFMA (FP32 inputs, FP32 accumulation)
    std::vector<float> Jf((size_t)P * nparam), wf(P);
    for (size_t i = 0; i < Jf.size(); ++i) Jf[i] = (float)J[i];
    for (int pix = 0; pix < P; ++pix) wf[pix] = E[pix] != 0.0 ? (float)(1.0 / E[pix]) : 0.f;
    std::vector<float> Ff((size_t)nparam * nparam);
    if ((rc = gram_mfma(ctx, P, nparam, Jf.data(), wf.data(), Ff.data()))) return rc;
    // the two triangles come from different MFMA tiles: symmetrise
    for (int a = 0; a < nparam; ++a)
        for (int b = 0; b < nparam; ++b)
            Fout[(size_t)a * nparam + b] = 0.5 * ((double)Ff[(size_t)a * nparam + b] + (double)Ff[(size_t)b * nparam + a]);
    if (chol) {
        // FP64 Cholesky F = C C^T (lower); fails if F is not positive definite
        std::vector<double> C((size_t)nparam * nparam, 0.0);
        for (int i = 0; i < nparam; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = Fout[(size_t)i * nparam + j];
                for (int k = 0; k < j; ++k) s -= C[(size_t)i * nparam + k] * C[(size_t)j * nparam + k];
                if (i == j) {
                    if (!(s > 0.0)) return fail(RTE_E_ARG, "Fisher matrix is not positive definite (pivot %d)", i);
                    C[(size_t)i * nparam + i] = std::sqrt(s);
                } else {
                    C[(size_t)i * nparam + j] = s / C[(size_t)j * nparam + j];
                }
            }
        std::memcpy(chol, C.data(), sizeof(double) * C.size());
    }
    return RTE_OK;
}

extern "C" int rte_fisher(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                          const rte_grad_opts *opts, const int *partition, int nparam, int wrt, double *F,
                          double *chol, double *J, double *exitance, rte_stats *stats)
{
    if (!partition || !F || nparam < 1) return fail(RTE_E_ARG, "rte_fisher: bad partition/nparam/F");
    if (wrt != 0 && wrt != 1) return fail(RTE_E_ARG, "rte_fisher: wrt must be 0 (mu_a) or 1 (mu_s)");
    int stride, rc;
    if ((rc = grad_check(ctx, m, beam, params, opts, &stride))) return rc;
    for (int v = 0; v < m->N; ++v)
        if (partition[v] >= nparam) return fail(RTE_E_ARG, "rte_fisher: partition[%d] = %d >= nparam", v, partition[v]);
    const int method = opts ? opts->jacobian : 0;
    if (method != 0 && method != 1) return fail(RTE_E_ARG, "rte_fisher: jacobian must be 0 (exact) or 1 (reference)");
    if (params->precision == RTE_PREC_FP64)
        return fisher_run<double, double>(ctx, m, beam, params, stride, method, partition, nparam, wrt, F, chol, J,
                                          exitance, stats);
    return fisher_run<float, double>(ctx, m, beam, params, stride, method, partition, nparam, wrt, F, chol, J,
                                     exitance, stats);
}

extern "C" int rte_fisher_inverse(int nparam, const double *chol, double *Finv)
{
    if (nparam < 1 || !chol || !Finv) return fail(RTE_E_ARG, "rte_fisher_inverse: bad nparam/chol/Finv");
    const int n = nparam;
    for (int i = 0; i < n; ++i)
        if (!(chol[(size_t)i * n + i] > 0.0) || !std::isfinite(chol[(size_t)i * n + i]))
            return fail(RTE_E_ARG, "rte_fisher_inverse: pivot %d is not positive", i);
    // Cinv = C^-1 (lower triangular) by forward substitution, column by column
    std::vector<double> Ci((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        Ci[(size_t)j * n + j] = 1.0 / chol[(size_t)j * n + j];
        for (int i = j + 1; i < n; ++i) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s += chol[(size_t)i * n + k] * Ci[(size_t)k * n + j];
            Ci[(size_t)i * n + j] = -s / chol[(size_t)i * n + i];
        }
    }
    // F^-1 = Cinv^T Cinv, symmetric
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = i; k < n; ++k) s += Ci[(size_t)k * n + i] * Ci[(size_t)k * n + j];
            Finv[(size_t)i * n + j] = Finv[(size_t)j * n + i] = s;
        }
    return RTE_OK;
}

extern "C" int rte_ref_jacobian(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                                const int *partition, int nparam, int wrt, double *J, double *exitance,
                                rte_stats *stats)
{
    if (!partition || !J || nparam < 1) return fail(RTE_E_ARG, "rte_ref_jacobian: bad partition/nparam/J");
    if (wrt != 0 && wrt != 1) return fail(RTE_E_ARG, "rte_ref_jacobian: wrt must be 0 (mu_a) or 1 (mu_s)");
    int stride, rc;
    if ((rc = grad_check(ctx, m, beam, params, nullptr, &stride))) return rc;
    for (int v = 0; v < m->N; ++v)
        if (partition[v] >= nparam)
            return fail(RTE_E_ARG, "rte_ref_jacobian: partition[%d] = %d >= nparam", v, partition[v]);
    std::vector<double> Jv, E;
    if (params->precision == RTE_PREC_FP64)
        rc = ref_jacobian_run<double, double>(ctx, m, beam, params, partition, nparam, wrt, Jv, E, stats);
    else
        rc = ref_jacobian_run<float, double>(ctx, m, beam, params, partition, nparam, wrt, Jv, E, stats);
    if (rc) return rc;
    std::memcpy(J, Jv.data(), sizeof(double) * Jv.size());
    if (exitance) std::memcpy(exitance, E.data(), sizeof(double) * E.size());
    return RTE_OK;
}

// Gram G = J^T diag(w) J of a P x n matrix on MFMA (rte_kernels.hip: k_gram)
static int gram_mfma(rte_ctx *ctx, int P, int n, const float *J, const float *w, float *G)
{
    DevBuf dJ, dw, dG;
    HIP_TRY(dJ.alloc(sizeof(float) * (size_t)P * n));
    HIP_TRY(dw.alloc(sizeof(float) * P));
    HIP_TRY(dG.alloc(sizeof(float) * (size_t)n * n));
    HIP_TRY(hipMemcpyAsync(dJ.p, J, sizeof(float) * (size_t)P * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dw.p, w, sizeof(float) * P, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_gram(P, n, dJ.as<float>(), dw.as<float>(), dG.as<float>(), ctx->stream));
    HIP_TRY(hipMemcpyAsync(G, dG.p, sizeof(float) * (size_t)n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTE_OK;
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
template <typename T, typename TA>
static int propagate_impl(rte_ctx *ctx, rte_medium *m, int L, double prop_thresh, int path,
                          const double *W, double *out, bool cplx)
{
    // cplx: W and out in the reference's layout (complex basis, voxel-major, 2 K N doubles)
    const rte_grid &grid = m->grid;
    const int K = (L + 1) * (L + 1);
    const GridDev<T> g = make_grid<T>(grid);
    const int N = g.N;
    const double dx = (grid.xmax - grid.xmin) / grid.nx, dy = (grid.ymax - grid.ymin) / grid.ny,
                 dz = (grid.zmax - grid.zmin) / grid.nz;
    const double thr = prop_thresh > 0 ? prop_thresh : -1.0;
    hipStream_t st = ctx->stream;
    MomentSet *mom = nullptr;
    int rc;
    if ((rc = get_moments(ctx, L, dx, dy, dz, &mom))) return rc;
    if (path == RTE_PATH_CACHED && (rc = ensure_ktab<T>(m, thr))) return rc;
    const int nch0 = pick_chunks(g.nblk, 1, N);
    const int clen = (N + nch0 - 1) / nch0;
    const int nch = (N + clen - 1) / clen;
    DevBuf dIn, dW, dWp, dPart, dAW, dOut, dSv;
    const size_t nio = (size_t)K * N * (cplx ? 2 : 1);
    HIP_TRY(dIn.alloc(nio * 8));
    HIP_TRY(dW.alloc((size_t)K * N * sizeof(T)));
    HIP_TRY(dWp.alloc((size_t)K * N * sizeof(T)));
    HIP_TRY(dPart.alloc((size_t)nch * N * K * sizeof(T)));
    HIP_TRY(dAW.alloc((size_t)K * N * sizeof(TA)));
    HIP_TRY(dOut.alloc(nio * 8));
    HIP_TRY(dSv.alloc(sizeof(SolveDev<T, TA>)));
    HIP_TRY(hipMemcpyAsync(dIn.p, W, nio * 8, hipMemcpyHostToDevice, st));
    const Scales sc = make_scales(L);
    if (cplx)
        HIP_TRY(launch_cplx_to_real<T>(L, N, 1.0, sc, dIn.as<double>(), dW.as<T>(), st));
    else
        HIP_TRY(launch_to_voxel_major<T>(K, N, sc, dIn.as<double>(), dW.as<T>(), st));
    SolveDev<T, TA> s;
    std::memset(&s, 0, sizeof s);
    s.W = dW.as<T>();
    s.Wp = dWp.as<T>();
    s.part = dPart.as<T>();
    s.ktab = sizeof(T) == 4 ? (const T *)m->ktab_f.p : (const T *)m->ktab_d.p;
    s.mut = sizeof(T) == 4 ? (const T *)m->mut_f.p : (const T *)m->mut_d.p;
    s.mus = sizeof(T) == 4 ? (const T *)m->mus_f.p : (const T *)m->mus_d.p;
    s.g = T(m->g);
    HIP_TRY(hipMemcpyAsync(dSv.p, &s, sizeof s, hipMemcpyHostToDevice, st));
    PropArgs<T> pa;
    pa.g = g;
    pa.nchunk = nch;
    pa.chunk_len = clen;
    pa.thr2 = thr > 0 ? T(thr * thr) : T(sizeof(T) == 4 ? 1e30 : 1e300);
    HIP_TRY(launch_prop<T, TA>(L, path, 1, pa, dSv.as<SolveDev<T, TA>>(), 1, st));
    CombArgs<T, TA> ca;
    std::memset(&ca, 0, sizeof ca);
    ca.g = g;
    ca.nchunk = nch;
    ca.op = 1;
    ca.mom = sizeof(TA) == 4 ? (const TA *)mom->mf.p : (const TA *)mom->md.p;
    ca.momT = sizeof(T) == 4 ? (const T *)mom->mf.p : (const T *)mom->md.p;
    ca.self_scale = TA((M_PI / kAngRes) * (2.0 * M_PI / kAngRes) / (dx * dy * dz));
    ca.dV = TA(dx * dy * dz);
    ca.aw_out = dAW.as<TA>();
    for (int a = 0; a < kMaxK; ++a) ca.nsc[a] = TA(sc.n[a]);
    HIP_TRY(launch_combine<T, TA>(L, ca, dSv.as<SolveDev<T, TA>>(), 1, st));
    // the reference's A carries 1/c_m (CPU package/src/Kernels_selfsub.c:948,1046)
    const double cm = m->c / m->n;
    if (cplx) {
        HIP_TRY(launch_real_to_cplx<TA>(L, N, 1.0 / cm, dAW.as<TA>(), dOut.as<double>(), st));
        HIP_TRY(hipMemcpyAsync(out, dOut.p, nio * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return RTE_OK;
    }
    HIP_TRY(launch_to_coef_major<TA>(K, N, unit_scales(), dAW.as<TA>(), dOut.as<double>(), st));
    HIP_TRY(hipMemcpyAsync(out, dOut.p, (size_t)K * N * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < (size_t)K * N; ++i) out[i] /= cm;
    return RTE_OK;
}

extern "C" int rte_propagate(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int L,
                             double prop_thresh, int precision, int path_mode, const double *W,
                             double *out)
{
    if (!ctx || !W || !out) return fail(RTE_E_ARG, "rte_propagate: NULL argument");
    if (L < 0 || L > RTE_MAX_L) return fail(RTE_E_UNSUPP, "L=%d outside 0..%d", L, RTE_MAX_L);
    if (path_mode != RTE_PATH_CACHED && path_mode != RTE_PATH_MARCH)
        return fail(RTE_E_ARG, "unknown path_mode %d", path_mode);
    rte_medium *m = nullptr;
    int rc = rte_medium_create(ctx, grid, optics, &m);
    if (rc) return rc;
    switch (precision) {
    case RTE_PREC_FP64: rc = propagate_impl<double, double>(ctx, m, L, prop_thresh, path_mode, W, out); break;
    case RTE_PREC_MIXED: rc = propagate_impl<float, double>(ctx, m, L, prop_thresh, path_mode, W, out); break;
    case RTE_PREC_FP32: rc = propagate_impl<float, float>(ctx, m, L, prop_thresh, path_mode, W, out); break;
    default: rc = fail(RTE_E_ARG, "unknown precision %d", precision);
    }
    std::string keep = g_err;
    rte_medium_destroy(m);
    g_err = keep;
    return rc;
}

// ---- rte_propagate without device allocation --------------------------------
// Buffers are carved out of the device's scratch array in a fixed order, each
// on a 256-byte boundary.  size() before any device work tells whether the
// call fits.
struct ScratchCarve {
    unsigned char *base;
    size_t used = 0;
    explicit ScratchCarve(void *p) : base(static_cast<unsigned char *>(p)) {}
    template <typename X>
    X *take(size_t n)
    {
        used = (used + 255) & ~size_t(255);
        X *p = base ? reinterpret_cast<X *>(base + used) : nullptr;
        used += n * sizeof(X);
        return p;
    }
};

// The same launches as propagate_impl with the marched path integral, every
// device buffer in the scratch: the self-block moments (and their quadrature
// inputs), mu_t and mu_s, the field, the chunk partials, A W and the output.
// Nothing needs to start at zero: k_prop writes every partial, k_combine every
// value of A W.
template <typename T, typename TA>
static int propagate_noalloc_impl(rte_ctx *ctx, const rte_grid &grid, const rte_optics &opt, int L,
                                  double prop_thresh, const double *W, double *out)
{
    const int K = (L + 1) * (L + 1);
    const GridDev<T> g = make_grid<T>(grid);
    const int N = g.N;
    const double dx = (grid.xmax - grid.xmin) / grid.nx, dy = (grid.ymax - grid.ymin) / grid.ny,
                 dz = (grid.zmax - grid.zmin) / grid.nz;
    const double thr = prop_thresh > 0 ? prop_thresh : -1.0;
    const int nch0 = pick_chunks(g.nblk, 1, N);
    const int clen = (N + nch0 - 1) / nch0;
    const int nch = (N + clen - 1) / clen;
    const size_t nmom = (size_t)kNMom * K * K;
    auto carve = [&](ScratchCarve &c, double *&R, double *&sn, double *&beta, double *&md, float *&mf, T *&mut,
                     T *&mus, double *&in, T *&w, T *&wp, T *&part, TA *&aw, double *&o, SolveDev<T, TA> *&sv) {
        R = c.take<double>((size_t)kNOmega * K);
        sn = c.take<double>(kNOmega);
        beta = c.take<double>(kNOmega);
        md = c.take<double>(nmom);
        mf = c.take<float>(nmom);
        mut = c.take<T>(N);
        mus = c.take<T>(N);
        in = c.take<double>((size_t)K * N);
        w = c.take<T>((size_t)K * N);
        wp = c.take<T>((size_t)K * N);
        part = c.take<T>((size_t)nch * N * K);
        aw = c.take<TA>((size_t)K * N);
        o = c.take<double>((size_t)K * N);
        sv = c.take<SolveDev<T, TA>>(1);
    };
    double *dR, *dsn, *dbeta, *dmd, *dIn, *dOut;
    float *dmf;
    T *dmut, *dmus, *dW, *dWp, *dPart;
    TA *dAW;
    SolveDev<T, TA> *dSv;
    {
        ScratchCarve probe(nullptr);
        carve(probe, dR, dsn, dbeta, dmd, dmf, dmut, dmus, dIn, dW, dWp, dPart, dAW, dOut, dSv);
        if (probe.used > kScratchBytes)
            return fail(RTE_E_UNSUPP,
                        "rte_propagate_noalloc: %d^3-voxel-equivalent grid (N = %d) at L = %d needs %zu bytes of "
                        "device scratch, more than its %zu; use rte_propagate",
                        (int)std::cbrt((double)N), N, L, probe.used, kScratchBytes);
    }
    void *base = scratch_ptr(ctx->device);
    if (!base) return fail(RTE_E_DEVICE, "rte_propagate_noalloc: no scratch address for device %d", ctx->device);
    ScratchCarve c(base);
    carve(c, dR, dsn, dbeta, dmd, dmf, dmut, dmus, dIn, dW, dWp, dPart, dAW, dOut, dSv);
    hipStream_t st = ctx->stream;
    // self-block moments, as get_moments computes them
    std::vector<double> beta, th, ph;
    self_beta(dx, dy, dz, beta, th, ph);
    std::vector<double> R((size_t)kNOmega * K), sn(kNOmega);
    for (int q = 0; q < kNOmega; ++q) {
        const double s = std::sin(th[q]);
        real_sh(L, std::cos(ph[q]) * s, std::sin(ph[q]) * s, std::cos(th[q]), &R[(size_t)q * K]);
        sn[q] = s;
    }
    HIP_TRY(hipMemcpyAsync(dR, R.data(), R.size() * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dsn, sn.data(), sn.size() * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dbeta, beta.data(), beta.size() * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(launch_moments(K, dR, dsn, dbeta, dmd, st));
    HIP_TRY(launch_narrow(nmom, dmd, dmf, st));
    // the optical maps in the compute precision, as rte_medium_create forms them
    std::vector<T> mt(N), ms(N);
    for (int v = 0; v < N; ++v) {
        mt[v] = T(opt.mu_a[v] + opt.mu_s[v]);
        ms[v] = T(opt.mu_s[v]);
    }
    HIP_TRY(hipMemcpyAsync(dmut, mt.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dmus, ms.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dIn, W, (size_t)K * N * 8, hipMemcpyHostToDevice, st));
    const Scales sc = make_scales(L);
    HIP_TRY(launch_to_voxel_major<T>(K, N, sc, dIn, dW, st));
    SolveDev<T, TA> s;
    std::memset(&s, 0, sizeof s);
    s.W = dW;
    s.Wp = dWp;
    s.part = dPart;
    s.mut = dmut;
    s.mus = dmus;
    s.g = T(opt.g);
    HIP_TRY(hipMemcpyAsync(dSv, &s, sizeof s, hipMemcpyHostToDevice, st));
    PropArgs<T> pa;
    pa.g = g;
    pa.nchunk = nch;
    pa.chunk_len = clen;
    pa.thr2 = thr > 0 ? T(thr * thr) : T(sizeof(T) == 4 ? 1e30 : 1e300);
    HIP_TRY(launch_prop<T, TA>(L, RTE_PATH_MARCH, 1, pa, dSv, 1, st));
    CombArgs<T, TA> ca;
    std::memset(&ca, 0, sizeof ca);
    ca.g = g;
    ca.nchunk = nch;
    ca.op = 1;
    ca.mom = sizeof(TA) == 4 ? (const TA *)(const void *)dmf : (const TA *)(const void *)dmd;
    ca.momT = sizeof(T) == 4 ? (const T *)(const void *)dmf : (const T *)(const void *)dmd;
    ca.self_scale = TA((M_PI / kAngRes) * (2.0 * M_PI / kAngRes) / (dx * dy * dz));
    ca.dV = TA(dx * dy * dz);
    ca.aw_out = dAW;
    for (int a = 0; a < kMaxK; ++a) ca.nsc[a] = TA(sc.n[a]);
    HIP_TRY(launch_combine<T, TA>(L, ca, dSv, 1, st));
    HIP_TRY(launch_to_coef_major<TA>(K, N, unit_scales(), dAW, dOut, st));
    HIP_TRY(hipMemcpyAsync(out, dOut, (size_t)K * N * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // the reference's A carries 1/c_m (CPU package/src/Kernels_selfsub.c:948,1046)
    const double cm = opt.c / opt.n;
    for (size_t i = 0; i < (size_t)K * N; ++i) out[i] /= cm;
    return RTE_OK;
}

extern "C" int rte_propagate_noalloc(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int L,
                                     double prop_thresh, int precision, int path_mode, const double *W,
                                     double *out)
{
    if (!ctx || !grid || !optics || !W || !out) return fail(RTE_E_ARG, "rte_propagate_noalloc: NULL argument");
    if (L < 0 || L > RTE_MAX_L) return fail(RTE_E_UNSUPP, "L=%d outside 0..%d", L, RTE_MAX_L);
    if (path_mode != RTE_PATH_MARCH)
        return fail(RTE_E_UNSUPP, "rte_propagate_noalloc: only RTE_PATH_MARCH (a cached path table is a "
                                  "per-medium allocation of N^2 entries)");
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_optics(grid, optics)) || (rc = set_device(ctx))) return rc;
    switch (precision) {
    case RTE_PREC_FP64: return propagate_noalloc_impl<double, double>(ctx, *grid, *optics, L, prop_thresh, W, out);
    case RTE_PREC_MIXED: return propagate_noalloc_impl<float, double>(ctx, *grid, *optics, L, prop_thresh, W, out);
    case RTE_PREC_FP32: return propagate_noalloc_impl<float, float>(ctx, *grid, *optics, L, prop_thresh, W, out);
    default: return fail(RTE_E_ARG, "unknown precision %d", precision);
    }
}

extern "C" int rte_scatter(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int L, double *W)
{
    if (!ctx || !W) return fail(RTE_E_ARG, "rte_scatter: NULL argument");
    if (L < 0 || L > RTE_MAX_L) return fail(RTE_E_UNSUPP, "L=%d outside 0..%d", L, RTE_MAX_L);
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_optics(grid, optics)) || (rc = set_device(ctx))) return rc;
    const int N = grid->nx * grid->ny * grid->nz, K = (L + 1) * (L + 1);
    DevBuf dW, dmus;
    HIP_TRY(dW.alloc((size_t)K * N * 8));
    HIP_TRY(dmus.alloc((size_t)N * 8));
    HIP_TRY(hipMemcpyAsync(dW.p, W, (size_t)K * N * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dmus.p, optics->mu_s, (size_t)N * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_scatter_op(L, N, optics->g, dmus.as<double>(), dW.as<double>(), ctx->stream));
    HIP_TRY(hipMemcpyAsync(W, dW.p, (size_t)K * N * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // the reference's D carries c_m (CPU package/src/Kernels_selfsub.c:42)
    const double cm = optics->c / optics->n;
    for (size_t i = 0; i < (size_t)K * N; ++i) W[i] *= cm;
    return RTE_OK;
}

extern "C" int rte_source(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics,
                          const rte_beam *beam, int L, double *s0, double *ballistic, int *ncols)
{
    if (!ctx || !beam || !s0) return fail(RTE_E_ARG, "rte_source: NULL argument");
    if (L < 0 || L > RTE_MAX_L) return fail(RTE_E_UNSUPP, "L=%d outside 0..%d", L, RTE_MAX_L);
    rte_medium *m = nullptr;
    int rc = rte_medium_create(ctx, grid, optics, &m);
    if (rc) return rc;
    const int N = m->N, K = (L + 1) * (L + 1), P = grid->nx * grid->ny;
    const double dx = (grid->xmax - grid->xmin) / grid->nx, dy = (grid->ymax - grid->ymin) / grid->ny,
                 dz = (grid->zmax - grid->zmin) / grid->nz;
    const double half = (beam->diam + 1.0e-16) / 2.0;
    int cols = 0;
    for (int j = 0; j < grid->ny; ++j)
        for (int i = 0; i < grid->nx; ++i) {
            const double xc = (double)i * dx + grid->xmin + dx / 2.0;
            const double yc = (double)j * dy + grid->ymin + dy / 2.0;
            if (std::sqrt((xc - beam->pos_x) * (xc - beam->pos_x) + (yc - beam->pos_y) * (yc - beam->pos_y)) <=
                half)
                ++cols;
        }
    if (ncols) *ncols = cols;
    auto run = [&]() -> int {
        if (cols == 0) return fail(RTE_E_ARG, "beam hits no voxel column");
        hipStream_t st = ctx->stream;
        DevBuf dW, dpix, dred, dsv, dbeam, dout;
        HIP_TRY(dW.alloc((size_t)K * N * 8));
        HIP_TRY(dpix.alloc((size_t)kPixBufs * P * 8));
        HIP_TRY(dred.alloc(16));
        HIP_TRY(dsv.alloc(sizeof(SolveDev<double, double>)));
        HIP_TRY(dbeam.alloc(sizeof(BeamDev)));
        HIP_TRY(dout.alloc((size_t)K * N * 8));
        HIP_TRY(launch_fill<double>(dW.as<double>(), (size_t)K * N, 0.0, st));
        SolveDev<double, double> s;
        std::memset(&s, 0, sizeof s);
        s.W = dW.as<double>();
        s.g = m->g;
        set_pixels(s, dpix.as<double>(), P);
        s.red = dred.as<double>();
        BeamDev b{beam->pos_x, beam->pos_y, half, beam->mag / cols, m->mua_d.as<double>(), m->mus_d.as<double>()};
        HIP_TRY(hipMemcpyAsync(dsv.p, &s, sizeof s, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dbeam.p, &b, sizeof b, hipMemcpyHostToDevice, st));
        const GridDev<double> g = make_grid<double>(*grid);
        const Scales sc = make_scales(L);
        HIP_TRY(launch_source<double, double>(L, g, dx, dy, dz, grid->xmin, grid->ymin, sc, dbeam.as<BeamDev>(),
                                              dsv.as<SolveDev<double, double>>(), 1, st));
        HIP_TRY(launch_to_coef_major<double>(K, N, sc, dW.as<double>(), dout.as<double>(), st));
        HIP_TRY(hipMemcpyAsync(s0, dout.p, (size_t)K * N * 8, hipMemcpyDeviceToHost, st));
        if (ballistic)
            HIP_TRY(hipMemcpyAsync(ballistic, s.E, (size_t)P * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return RTE_OK;
    };
    rc = run();
    std::string keep = g_err;
    rte_medium_destroy(m);
    g_err = keep;
    return rc;
}

extern "C" int rte_exitance(rte_ctx *ctx, const rte_grid *grid, int L, int exit_quad, double c_m,
                            const double *w, double *img)
{
    if (!ctx || !w || !img) return fail(RTE_E_ARG, "rte_exitance: NULL argument");
    if (!(c_m > 0.0)) return fail(RTE_E_ARG, "rte_exitance: c_m must be > 0");
    if (L < 0 || L > RTE_MAX_L) return fail(RTE_E_UNSUPP, "L=%d outside 0..%d", L, RTE_MAX_L);
    if (exit_quad != RTE_QUAD_REF && exit_quad != RTE_QUAD_EXACT) return fail(RTE_E_ARG, "unknown exit_quad");
    int rc;
    if ((rc = check_grid(grid)) || (rc = set_device(ctx))) return rc;
    const int N = grid->nx * grid->ny * grid->nz, K = (L + 1) * (L + 1), P = grid->nx * grid->ny;
    const double dx = (grid->xmax - grid->xmin) / grid->nx, dy = (grid->ymax - grid->ymin) / grid->ny;
    double Q[kMaxL + 1], coef[kMaxL + 1];
    quad_weights(L, exit_quad, Q);
    for (int l = 0; l <= L; ++l)
        coef[l] = 2.0 * M_PI * dx * dy * c_m / std::sqrt(4.0 * M_PI) * std::sqrt(2.0 * l + 1.0) * Q[l];
    DevBuf dw, dc, dimg;
    HIP_TRY(dw.alloc((size_t)K * N * 8));
    HIP_TRY(dc.alloc(sizeof coef));
    HIP_TRY(dimg.alloc((size_t)P * 8));
    HIP_TRY(hipMemcpyAsync(dw.p, w, (size_t)K * N * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dc.p, coef, sizeof coef, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_exitance_op(L, grid->nx, grid->ny, grid->nz, dc.as<double>(), dw.as<double>(),
                               dimg.as<double>(), ctx->stream));
    HIP_TRY(hipMemcpyAsync(img, dimg.p, (size_t)P * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTE_OK;
}

extern "C" int rte_path_integrals(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int npairs,
                                  const int *dest, const int *src, double *tau)
{
    if (!ctx || !dest || !src || !tau || npairs < 0) return fail(RTE_E_ARG, "rte_path_integrals: bad argument");
    rte_medium *m = nullptr;
    int rc = rte_medium_create(ctx, grid, optics, &m);
    if (rc) return rc;
    for (int i = 0; i < npairs; ++i)
        if (dest[i] < 0 || dest[i] >= m->N || src[i] < 0 || src[i] >= m->N) {
            rte_medium_destroy(m);
            return fail(RTE_E_ARG, "pair %d outside the grid", i);
        }
    auto run = [&]() -> int {
        if (npairs == 0) return RTE_OK;
        DevBuf dd, ds, dt;
        HIP_TRY(dd.alloc((size_t)npairs * 4));
        HIP_TRY(ds.alloc((size_t)npairs * 4));
        HIP_TRY(dt.alloc((size_t)npairs * 8));
        HIP_TRY(hipMemcpyAsync(dd.p, dest, (size_t)npairs * 4, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ds.p, src, (size_t)npairs * 4, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(launch_tau_list(make_grid<float>(*grid), m->mut_f.as<float>(), npairs, dd.as<int>(), ds.as<int>(),
                                dt.as<double>(), ctx->stream));
        HIP_TRY(hipMemcpyAsync(tau, dt.p, (size_t)npairs * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return RTE_OK;
    };
    rc = run();
    std::string keep = g_err;
    rte_medium_destroy(m);
    g_err = keep;
    return rc;
}

// ===========================================================================
// Serving: continuous batching.
//
// A worker thread owns the device while the server exists.  Requests wait in
// a FIFO and are granted a solve slot of the bucket of their shape (grid, L,
// precision, path, stop rule) while the device memory of all slots fits the
// HBM budget, first come first served.  A slot holds one solve's whole state
// (optical maps, path table, Neumann field, partial sums, images) and is
// pooled: when its solve leaves, the next request of that shape reuses it, and
// idle slots of other shapes are freed only when memory is needed.
//
// Each bucket has its own stream and at most one step in flight: joining
// solves get their maps, path table, source and initial sums, then every
// resident solve advances one Neumann term in a HIP graph (propagation,
// self block/scatter/exitance, bottom-layer sums) that is captured once per
// resident count and replayed; the per-solve descriptors live in a device
// array at a fixed address that is refreshed before every replay.  The worker
// polls the buckets' events, so a bucket of small grids runs its terms while a
// bucket of large grids is still in its own, and the work of all buckets
// overlaps on the device.  A solve's arithmetic does not depend on the other
// solves of its step, so its result is the same alone and in any mix.
// ===========================================================================
#include <atomic>
#include <condition_variable>
#include <deque>
#include <list>
#include <mutex>
#include <thread>

namespace {

enum { REQ_QUEUED = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_FAILED = 3, REQ_CANCELLED = 4 };

struct ServReq {
    long long id = 0;
    rte_grid grid{};
    std::vector<double> mua, mus, mutd;  // FP64 maps, mu_t = mu_a + mu_s
    std::vector<float> mutf, musf;       // FP32 mu_t, mu_s
    double g = 0;
    rte_beam beam{};
    int beam_cols = 0;
    rte_params params{};
    std::string key;
    double t_submit = 0, t_admit = 0;
    int state = REQ_QUEUED;  // read and written under the server lock
    int outcome = REQ_DONE;  // set by the worker, published as state under the lock
    std::atomic<bool> cancel{false};
    int rc = 0;
    std::string err;
    long long footprint = 0;
    std::vector<double> img;
    rte_stats st{};
};

struct SlotBase {
    ServReq *r = nullptr;  // null: idle
    int idx = 0;           // row of the bucket's reduction and image arrays
    long long bytes = 0;
    DevBuf mua, mus, mutd, mutf, musf, ktab, W, Wp, part, pix;
    HostSolve h{};
    bool fresh = false, ran = false;
    void free_memory()
    {
        for (DevBuf *b : {&mua, &mus, &mutd, &mutf, &musf, &ktab, &W, &Wp, &part, &pix}) b->release();
        bytes = 0;
    }
};

struct BucketBase {
    hipStream_t stream = nullptr;
    hipEvent_t ev_step = nullptr, ev_drain = nullptr;
    bool step_inflight = false, drain_inflight = false;
    int init_err = RTE_OK;
    int cap = 0;
    long long slot_bytes = 0;
    std::vector<std::unique_ptr<SlotBase>> pool;  // allocated slots, index = idx
    std::vector<SlotBase *> joining, live, draining;
    virtual ~BucketBase() {}
    virtual int launch(std::vector<ServReq *> &finished) = 0;   // enqueue one step (stream idle)
    virtual void complete(std::vector<ServReq *> &finished) = 0; // after ev_step
    virtual void drained(std::vector<ServReq *> &finished) = 0;  // after ev_drain
    virtual int alloc_slot(SlotBase &s) = 0;
    int n_busy() const { return (int)(joining.size() + live.size() + draining.size()); }
    SlotBase *idle_slot()
    {
        for (auto &s : pool)
            if (!s->r && s->bytes) return s.get();
        return nullptr;
    }
    // frees one idle slot's memory; returns the bytes released
    long long evict_one()
    {
        for (auto &s : pool)
            if (!s->r && s->bytes) {
                const long long b = s->bytes;
                s->free_memory();
                return b;
            }
        return 0;
    }
    int n_allocated() const
    {
        int n = 0;
        for (auto &s : pool) n += s->bytes ? 1 : 0;
        return n;
    }
};

template <typename T, typename TA>
struct Bucket : BucketBase {
    rte_ctx *ctx = nullptr;
    rte_grid grid{};
    rte_params p{};
    int L = 0, K = 0, N = 0, P = 0, nch = 0, clen = 0, max_terms = 0;
    double dx = 0, dy = 0, dz = 0, thr = -1;
    GridDev<T> g{};
    Scales sc{};
    PropArgs<T> pa{};
    CombArgs<T, TA> ca{};
    std::vector<SolveDev<T, TA>> sv;  // per slot index
    DevBuf d_init, d_beam, d_all, d_red;
    // pinned host staging: [init sv | all sv | beams], the sums, the images
    char *h_stage = nullptr;
    TA *h_red = nullptr;  // [cap][2] term sums, then [cap][2] initial sums
    TA *h_img = nullptr;  // [cap][P]
    std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;  // by resident count

    Bucket(rte_ctx *c, const rte_grid &gr, const rte_params &pr, int capacity, MomentSet *mom)
        : ctx(c), grid(gr), p(pr)
    {
        cap = capacity;
        L = p.L;
        K = (L + 1) * (L + 1);
        g = make_grid<T>(grid);
        N = g.N;
        P = grid.nx * grid.ny;
        dx = (grid.xmax - grid.xmin) / grid.nx;
        dy = (grid.ymax - grid.ymin) / grid.ny;
        dz = (grid.zmax - grid.zmin) / grid.nz;
        thr = p.prop_thresh > 0 ? p.prop_thresh : -1.0;
        max_terms = p.max_terms > 0 ? p.max_terms : 1000;
        slot_bytes = rte_solve_footprint(&grid, &p, p.path_mode);
        // the same source chunking as rte_solve_media: it depends on the grid only
        const int nch0 = pick_chunks(g.nblk, 1, N);
        clen = (N + nch0 - 1) / nch0;
        nch = (N + clen - 1) / clen;
        sc = make_scales(L);
        pa.g = g;
        pa.nchunk = nch;
        pa.chunk_len = clen;
        pa.thr2 = thr > 0 ? T(thr * thr) : T(sizeof(T) == 4 ? 1e30 : 1e300);
        std::memset(&ca, 0, sizeof ca);
        ca.g = g;
        ca.nchunk = nch;
        ca.op = 0;
        ca.mom = sizeof(TA) == 4 ? (const TA *)mom->mf.p : (const TA *)mom->md.p;
        ca.momT = sizeof(T) == 4 ? (const T *)mom->mf.p : (const T *)mom->md.p;
        ca.self_scale = TA((M_PI / kAngRes) * (2.0 * M_PI / kAngRes) / (dx * dy * dz));
        ca.dV = TA(dx * dy * dz);
        double cE[kMaxL + 1], cR[kMaxL + 1];
        face_coefs(L, p.exit_quad, dx, dy, cE, cR);
        for (int l = 0; l <= L; ++l) {
            ca.coefE[l] = TA(cE[l]);
            ca.coefR[l] = TA(cR[l]);
        }
        for (int a = 0; a < K; ++a) ca.wsum[a] = TA(stop_weight(a));
        for (int a = 0; a < kMaxK; ++a) ca.nsc[a] = TA(sc.n[a]);
        sv.resize(cap);
        for (int i = 0; i < cap; ++i) {
            pool.emplace_back(new SlotBase);
            pool.back()->idx = i;
        }
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
            stream = nullptr;
            init_err = RTE_E_DEVICE;
            return;
        }
        const size_t svb = sizeof(SolveDev<T, TA>) * cap;
        if (hipEventCreateWithFlags(&ev_step, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_drain, hipEventDisableTiming) != hipSuccess) {
            init_err = RTE_E_DEVICE;
            return;
        }
        if (d_init.alloc(svb) != hipSuccess || d_all.alloc(svb) != hipSuccess ||
            d_beam.alloc(sizeof(BeamDev) * cap) != hipSuccess || d_red.alloc(sizeof(TA) * 2 * cap) != hipSuccess ||
            hipHostMalloc((void **)&h_stage, 2 * svb + sizeof(BeamDev) * cap, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&h_red, sizeof(TA) * 4 * cap, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&h_img, sizeof(TA) * (size_t)P * cap, hipHostMallocDefault) != hipSuccess)
            init_err = RTE_E_NOMEM;
    }

    ~Bucket() override
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &kv : graphs) {
            (void)hipGraphExecDestroy(kv.second.second);
            (void)hipGraphDestroy(kv.second.first);
        }
        for (auto &s : pool) s->free_memory();
        if (h_img) (void)hipHostFree(h_img);
        if (h_red) (void)hipHostFree(h_red);
        if (h_stage) (void)hipHostFree(h_stage);
        if (ev_step) (void)hipEventDestroy(ev_step);
        if (ev_drain) (void)hipEventDestroy(ev_drain);
        if (stream) (void)hipStreamDestroy(stream);
    }

    SolveDev<T, TA> *h_init() { return reinterpret_cast<SolveDev<T, TA> *>(h_stage); }
    SolveDev<T, TA> *h_all() { return h_init() + cap; }
    BeamDev *h_beam() { return reinterpret_cast<BeamDev *>(h_init() + 2 * cap); }

    int alloc_slot(SlotBase &s) override
    {
        const size_t n = (size_t)N, F = (size_t)N * K;
        const bool cached = p.path_mode == RTE_PATH_CACHED;
        if (s.mua.alloc(n * 8) || s.mus.alloc(n * 8) || s.mutd.alloc(n * 8) || s.mutf.alloc(n * 4) ||
            s.musf.alloc(n * 4) || (cached && s.ktab.alloc((size_t)N * g.npad * sizeof(T))) ||
            s.W.alloc(F * sizeof(T)) || s.Wp.alloc(F * sizeof(T)) || s.part.alloc((size_t)nch * F * sizeof(T)) ||
            s.pix.alloc((size_t)kPixBufs * P * sizeof(TA))) {
            s.free_memory();
            return fail(RTE_E_NOMEM, "server: device memory for a %dx%dx%d L=%d solve", grid.nx, grid.ny, grid.nz, L);
        }
        s.bytes = slot_bytes;
        return RTE_OK;
    }

    // maps, path table and descriptors of a joining solve (enqueued on the stream)
    int setup(SlotBase &s, int k)
    {
        ServReq *r = s.r;
        const size_t n = (size_t)N;
        HIP_TRY(hipMemcpyAsync(s.mua.p, r->mua.data(), n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(s.mus.p, r->mus.data(), n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(s.mutd.p, r->mutd.data(), n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(s.mutf.p, r->mutf.data(), n * 4, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(s.musf.p, r->musf.data(), n * 4, hipMemcpyHostToDevice, stream));
        const T *mut = sizeof(T) == 4 ? (const T *)s.mutf.p : (const T *)s.mutd.p;
        if (p.path_mode == RTE_PATH_CACHED)
            HIP_TRY(launch_ktab<T>(g, pa.thr2, mut, s.ktab.template as<T>(), stream));
        const size_t F = (size_t)N * K;
        HIP_TRY(launch_fill<T>(s.W.template as<T>(), F, T(0), stream));
        SolveDev<T, TA> &v = sv[s.idx];
        std::memset(&v, 0, sizeof v);
        v.W = s.W.template as<T>();
        v.Wp = s.Wp.template as<T>();
        v.part = s.part.template as<T>();
        v.ktab = p.path_mode == RTE_PATH_CACHED ? s.ktab.template as<T>() : nullptr;
        v.mut = mut;
        v.mus = sizeof(T) == 4 ? (const T *)s.musf.p : (const T *)s.mus.p;
        v.g = T(r->g);
        set_pixels(v, s.pix.template as<TA>(), P);
        v.red = d_red.template as<TA>() + 2 * s.idx;
        v.mode = 1;
        h_init()[k] = v;
        // beam columns and power per column (CPU package/src/Dist.c:1090-1114, Inits.c:56)
        h_beam()[k] = BeamDev{r->beam.pos_x, r->beam.pos_y, (r->beam.diam + 1.0e-16) / 2.0,
                              r->beam.mag / (double)r->beam_cols, s.mua.template as<double>(),
                              s.mus.template as<double>()};
        s.h = HostSolve{nullptr, 0, 0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
        if (p.stop_rule == RTE_STOP_LAMBDA) {
            s.h.n = 1;
            s.h.state = (max_terms <= 1) ? 2 : 0;
        }
        s.fresh = true;
        return RTE_OK;
    }

    int launch(std::vector<ServReq *> &finished) override
    {
        int nj = 0;
        for (SlotBase *s : joining) {
            if (s->r->cancel.load()) {
                s->r->outcome = REQ_CANCELLED;
                finished.push_back(s->r);
                s->r = nullptr;
                continue;
            }
            if (const int rc = setup(*s, nj)) return rc;
            ++nj;
            live.push_back(s);
        }
        joining.clear();
        if (nj) {
            HIP_TRY(hipMemcpyAsync(d_init.p, h_init(), sizeof(SolveDev<T, TA>) * nj, hipMemcpyHostToDevice, stream));
            HIP_TRY(hipMemcpyAsync(d_beam.p, h_beam(), sizeof(BeamDev) * nj, hipMemcpyHostToDevice, stream));
            HIP_TRY(launch_source<T, TA>(L, g, dx, dy, dz, grid.xmin, grid.ymin, sc, d_beam.as<BeamDev>(),
                                         d_init.as<SolveDev<T, TA>>(), nj, stream));
            HIP_TRY(launch_bottom<T, TA>(P, d_init.as<SolveDev<T, TA>>(), nj, stream));
            HIP_TRY(hipMemcpyAsync(h_red + 2 * cap, d_red.p, sizeof(TA) * 2 * cap, hipMemcpyDeviceToHost, stream));
        }
        if (!live.empty()) {
            // one term of every resident solve
            const int na = (int)live.size();
            SolveDev<T, TA> *all = h_all();
            for (int j = 0; j < na; ++j) {
                all[j] = sv[live[j]->idx];
                all[j].mode = live[j]->h.state;
                live[j]->ran = true;
            }
            HIP_TRY(hipMemcpyAsync(d_all.p, all, sizeof(SolveDev<T, TA>) * na, hipMemcpyHostToDevice, stream));
            auto it = graphs.find(na);
            if (it == graphs.end()) {
                HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                const SolveDev<T, TA> *d = d_all.as<SolveDev<T, TA>>();
                // resident solves in pairs across media on the mixed cached path (they
                // share the pair geometry, each reads its own table), the odd one alone
                hipError_t e = hipSuccess;
                const int npair = (sizeof(T) == 4 && sizeof(TA) == 8 && p.path_mode == RTE_PATH_CACHED) ? na / 2 : 0;
                if (npair) e = launch_prop_sep<T, TA>(L, 2, pa, d, npair, stream);
                if (e == hipSuccess && na > 2 * npair)
                    e = launch_prop<T, TA>(L, p.path_mode, 1, pa, d + 2 * npair, na - 2 * npair, stream);
                if (e == hipSuccess) e = launch_combine<T, TA>(L, ca, d, na, stream);
                if (e == hipSuccess) e = launch_bottom<T, TA>(P, d, na, stream);
                hipGraph_t gr = nullptr;
                const hipError_t e2 = hipStreamEndCapture(stream, &gr);
                if (e != hipSuccess || e2 != hipSuccess) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return fail(RTE_E_DEVICE, "server: graph capture failed (%s)",
                                hipGetErrorString(e != hipSuccess ? e : e2));
                }
                hipGraphExec_t ex = nullptr;
                const hipError_t e3 = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                if (e3 != hipSuccess) {
                    (void)hipGraphDestroy(gr);
                    return fail(RTE_E_DEVICE, "server: graph instantiation failed (%s)", hipGetErrorString(e3));
                }
                it = graphs.emplace(na, std::make_pair(gr, ex)).first;
            }
            HIP_TRY(hipGraphLaunch(it->second.second, stream));
            HIP_TRY(hipMemcpyAsync(h_red, d_red.p, sizeof(TA) * 2 * cap, hipMemcpyDeviceToHost, stream));
        }
        HIP_TRY(hipEventRecord(ev_step, stream));
        step_inflight = true;
        return RTE_OK;
    }

    void complete(std::vector<ServReq *> &finished) override
    {
        std::vector<SlotBase *> keep;
        bool copies = false;
        for (SlotBase *s : live) {
            HostSolve &h = s->h;
            if (s->fresh) {
                h.ball = (double)h_red[2 * cap + 2 * s->idx];
                h.T = h.ball;
                if (p.stop_rule == RTE_STOP_LAMBDA) h.sumOut = (double)h_red[2 * cap + 2 * s->idx + 1];
                s->fresh = false;
            }
            if (s->ran) {
                advance_series(h, (double)h_red[2 * s->idx], (double)h_red[2 * s->idx + 1], p.stop_rule, p.tol,
                               max_terms);
                s->ran = false;
            }
            if (s->r->cancel.load()) {
                s->r->outcome = REQ_CANCELLED;
                finished.push_back(s->r);
                s->r = nullptr;
            } else if (h.state == 3) {
                // the image is read back asynchronously; the slot stays busy until it lands
                (void)hipMemcpyAsync(h_img + (size_t)s->idx * P, sv[s->idx].E, sizeof(TA) * P, hipMemcpyDeviceToHost,
                                     stream);
                draining.push_back(s);
                copies = true;
            } else {
                keep.push_back(s);
            }
        }
        live.swap(keep);
        step_inflight = false;
        if (copies) {
            (void)hipEventRecord(ev_drain, stream);
            drain_inflight = true;
        }
    }

    void drained(std::vector<ServReq *> &finished) override
    {
        for (SlotBase *s : draining) {
            ServReq *r = s->r;
            const TA *im = h_img + (size_t)s->idx * P;
            r->img.assign(im, im + P);
            std::memset(&r->st, 0, sizeof r->st);
            r->st.terms = s->h.n;
            r->st.lambda_last = s->h.lam;
            r->st.rel_change_last = s->h.rel;
            r->st.total_exitance = s->h.T;
            r->st.ballistic_total = s->h.ball;
            r->outcome = REQ_DONE;
            finished.push_back(r);
            s->r = nullptr;
        }
        draining.clear();
        drain_inflight = false;
    }
};

}  // namespace

struct rte_server {
    rte_ctx *ctx = nullptr;
    rte_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<ServReq *> queue;                 // submitted, not yet seen by the worker
    std::map<long long, std::unique_ptr<ServReq>> reqs;
    long long next_id = 1;
    // published by the worker under the lock
    long long waiting = 0, running = 0, hbm_used = 0;
    long long n_done = 0, n_failed = 0, n_cancelled = 0, steps = 0;
    int n_buckets = 0;
    int worker_rc = RTE_OK;
    bool stop = false, failed = false;
    std::thread worker;
    // worker only
    std::map<std::string, std::unique_ptr<BucketBase>> buckets;
    std::list<ServReq *> fifo;
    long long reserved = 0;  // bytes of allocated slots, idle or busy
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

static std::string bucket_key(const rte_grid &g, const rte_params &p)
{
    char b[512];
    snprintf(b, sizeof b, "%d,%d,%d,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g|%d,%d,%d,%.17g,%.17g,%d,%d,%d", g.nx, g.ny, g.nz,
             g.xmin, g.xmax, g.ymin, g.ymax, g.zmin, g.zmax, p.L, p.max_terms, p.stop_rule, p.tol, p.prop_thresh,
             p.precision, p.exit_quad, p.path_mode);
    return b;
}

static BucketBase *make_bucket(rte_server *S, const ServReq *r, int *rc)
{
    const rte_grid &gr = r->grid;
    MomentSet *mom = nullptr;
    if ((*rc = get_moments(S->ctx, r->params.L, (gr.xmax - gr.xmin) / gr.nx, (gr.ymax - gr.ymin) / gr.ny,
                           (gr.zmax - gr.zmin) / gr.nz, &mom)))
        return nullptr;
    const int cap = S->opts.max_slots_per_bucket;
    std::unique_ptr<BucketBase> b;
    switch (r->params.precision) {
    case RTE_PREC_FP64: b.reset(new Bucket<double, double>(S->ctx, gr, r->params, cap, mom)); break;
    case RTE_PREC_MIXED: b.reset(new Bucket<float, double>(S->ctx, gr, r->params, cap, mom)); break;
    default: b.reset(new Bucket<float, float>(S->ctx, gr, r->params, cap, mom)); break;
    }
    if (b->init_err) {
        *rc = fail(b->init_err, "server: resources for a bucket of %dx%dx%d L=%d", gr.nx, gr.ny, gr.nz,
                   r->params.L);
        return nullptr;
    }
    return b.release();
}

// Grants slots to waiting requests, first come first served.  A request whose
// bucket has all its slots busy is passed over; one that needs memory the
// budget cannot give, even after freeing idle slots, stops the scan.
static void admit(rte_server *S, std::vector<ServReq *> &finished, std::vector<ServReq *> &admitted)
{
    for (auto it = S->fifo.begin(); it != S->fifo.end();) {
        ServReq *r = *it;
        if (r->cancel.load()) {
            r->outcome = REQ_CANCELLED;
            finished.push_back(r);
            it = S->fifo.erase(it);
            continue;
        }
        auto bi = S->buckets.find(r->key);
        if (bi == S->buckets.end()) {
            int rc = RTE_OK;
            BucketBase *b = make_bucket(S, r, &rc);
            if (!b) {
                r->rc = rc;
                r->err = g_err;
                r->outcome = REQ_FAILED;
                finished.push_back(r);
                it = S->fifo.erase(it);
                continue;
            }
            bi = S->buckets.emplace(r->key, std::unique_ptr<BucketBase>(b)).first;
        }
        BucketBase *B = bi->second.get();
        if (B->n_busy() >= B->cap) {
            ++it;
            continue;
        }
        SlotBase *s = B->idle_slot();
        if (!s) {
            // a new slot: make room by freeing idle slots of other shapes
            for (auto &ob : S->buckets) {
                while (S->reserved + B->slot_bytes > S->opts.hbm_budget && ob.second.get() != B) {
                    const long long freed = ob.second->evict_one();
                    if (!freed) break;
                    S->reserved -= freed;
                }
            }
            if (S->reserved + B->slot_bytes > S->opts.hbm_budget) break;
            for (auto &ps : B->pool)
                if (!ps->r && !ps->bytes) {
                    s = ps.get();
                    break;
                }
            const int rc = B->alloc_slot(*s);  // a slot without memory exists: fewer than cap are busy
            if (rc) {
                r->rc = rc;
                r->err = g_err;
                r->outcome = REQ_FAILED;
                finished.push_back(r);
                it = S->fifo.erase(it);
                continue;
            }
            S->reserved += B->slot_bytes;
        }
        s->r = r;
        B->joining.push_back(s);
        admitted.push_back(r);
        it = S->fifo.erase(it);
    }
}

static void server_loop(rte_server *S)
{
    (void)hipSetDevice(S->ctx->device);
    std::vector<ServReq *> finished, admitted;
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->stop) break;
            while (!S->queue.empty()) {
                S->fifo.push_back(S->queue.front());
                S->queue.pop_front();
            }
        }
        admit(S, finished, admitted);
        const double t_admit = S->now_ms();
        for (ServReq *r : admitted) r->t_admit = t_admit;
        int rc = RTE_OK;
        bool progressed = false, inflight = false;
        for (auto &kv : S->buckets) {
            BucketBase *B = kv.second.get();
            if (!B->step_inflight && (!B->joining.empty() || !B->live.empty())) {
                if ((rc = B->launch(finished))) break;
            }
        }
        for (auto &kv : S->buckets) {
            if (rc) break;
            BucketBase *B = kv.second.get();
            if (B->step_inflight) {
                const hipError_t e = hipEventQuery(B->ev_step);
                if (e == hipSuccess) {
                    B->complete(finished);
                    progressed = true;
                } else if (e != hipErrorNotReady) {
                    rc = fail(RTE_E_DEVICE, "server: device error (%s)", hipGetErrorString(e));
                }
            }
            if (!rc && B->drain_inflight) {
                const hipError_t e = hipEventQuery(B->ev_drain);
                if (e == hipSuccess) {
                    B->drained(finished);
                    progressed = true;
                } else if (e != hipErrorNotReady) {
                    rc = fail(RTE_E_DEVICE, "server: device error (%s)", hipGetErrorString(e));
                }
            }
            inflight = inflight || B->step_inflight || B->drain_inflight;
        }
        {
            std::unique_lock<std::mutex> lk(S->mu);
            const double now = S->now_ms();
            for (ServReq *r : admitted) {
                r->state = REQ_RUNNING;
                ++S->running;
            }
            admitted.clear();
            if (rc) {
                // a device error ends the service: every request not yet finished fails
                // (those this pass already finished keep their outcome and are counted once)
                S->worker_rc = rc;
                S->failed = true;
                const std::unordered_set<ServReq *> done_now(finished.begin(), finished.end());
                for (auto &kv : S->reqs)
                    if ((kv.second->state == REQ_RUNNING || kv.second->state == REQ_QUEUED) &&
                        !done_now.count(kv.second.get())) {
                        kv.second->rc = rc;
                        kv.second->err = g_err;
                        kv.second->outcome = REQ_FAILED;
                        finished.push_back(kv.second.get());
                    }
                S->queue.clear();
                S->fifo.clear();
            }
            for (ServReq *r : finished) {
                if (r->state == REQ_RUNNING) --S->running;
                r->state = r->outcome;
                r->st.latency_ms = now - r->t_submit;
                r->st.gpu_ms = r->t_admit > 0 ? now - r->t_admit : 0.0;
                if (r->outcome == REQ_DONE) ++S->n_done;
                else if (r->outcome == REQ_CANCELLED) ++S->n_cancelled;
                else ++S->n_failed;
            }
            if (progressed) ++S->steps;
            S->waiting = (long long)S->fifo.size();
            S->hbm_used = S->reserved;
            S->n_buckets = (int)S->buckets.size();
            const bool notify = !finished.empty();
            finished.clear();
            if (notify) S->cv_done.notify_all();
            if (rc) break;
            if (!progressed) {
                // sleep until a submission, or poll the device again shortly
                if (inflight)
                    S->cv.wait_for(lk, std::chrono::microseconds(50));
                else if (S->fifo.empty())
                    S->cv.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            }
        }
    }
}

extern "C" int rte_server_create(rte_ctx *ctx, const rte_server_opts *opts, rte_server **out)
{
    if (!ctx || !out) return fail(RTE_E_ARG, "rte_server_create: NULL argument");
    *out = nullptr;
    auto S = std::make_unique<rte_server>();
    S->ctx = ctx;
    if (opts) S->opts = *opts;
    if (S->opts.hbm_budget <= 0) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        S->opts.hbm_budget = (long long)(0.8 * (double)fr);
    }
    if (S->opts.max_slots_per_bucket <= 0) S->opts.max_slots_per_bucket = 64;
    if (S->opts.max_queue <= 0) S->opts.max_queue = 100000;
    rte_server *raw = S.release();
    raw->worker = std::thread(server_loop, raw);
    *out = raw;
    return RTE_OK;
}

extern "C" int rte_server_destroy(rte_server *S)
{
    if (!S) return RTE_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    S->buckets.clear();  // waits for each bucket's stream, then frees its slots
    delete S;
    return RTE_OK;
}

static int beam_columns(const rte_grid *grid, const rte_beam *beam)
{
    const double dx = (grid->xmax - grid->xmin) / grid->nx, dy = (grid->ymax - grid->ymin) / grid->ny;
    const double half = (beam->diam + 1.0e-16) / 2.0;
    int cols = 0;
    for (int j = 0; j < grid->ny; ++j)
        for (int i = 0; i < grid->nx; ++i) {
            const double xc = (double)i * dx + grid->xmin + dx / 2.0, yc = (double)j * dy + grid->ymin + dy / 2.0;
            if (std::sqrt((xc - beam->pos_x) * (xc - beam->pos_x) + (yc - beam->pos_y) * (yc - beam->pos_y)) <= half)
                ++cols;
        }
    return cols;
}

extern "C" int rte_submit(rte_server *S, const rte_grid *grid, const rte_optics *optics, const rte_beam *beam,
                          const rte_params *params, long long *id)
{
    if (!S || !grid || !optics || !beam || !params || !id) return fail(RTE_E_ARG, "rte_submit: NULL argument");
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_optics(grid, optics)) || (rc = check_params(params))) return rc;
    if (!std::isfinite(beam->pos_x) || !std::isfinite(beam->pos_y) || !(beam->diam >= 0.0) ||
        !std::isfinite(beam->mag))
        return fail(RTE_E_ARG, "rte_submit: invalid beam");
    auto r = std::make_unique<ServReq>();
    if (!(r->beam_cols = beam_columns(grid, beam))) return fail(RTE_E_ARG, "rte_submit: beam hits no voxel column");
    r->grid = *grid;
    const size_t N = (size_t)grid->nx * grid->ny * grid->nz;
    r->mua.assign(optics->mu_a, optics->mu_a + N);
    r->mus.assign(optics->mu_s, optics->mu_s + N);
    r->mutd.resize(N);
    r->mutf.resize(N);
    r->musf.resize(N);
    for (size_t v = 0; v < N; ++v) {
        r->mutd[v] = r->mua[v] + r->mus[v];
        r->mutf[v] = (float)r->mutd[v];
        r->musf[v] = (float)r->mus[v];
    }
    r->g = optics->g;
    r->beam = *beam;
    r->params = *params;
    r->key = bucket_key(*grid, *params);
    r->footprint = rte_solve_footprint(grid, params, params->path_mode);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->failed) return fail(RTE_E_DEVICE, "rte_submit: the server stopped after a device error");
    if (r->footprint > S->opts.hbm_budget)
        return fail(RTE_E_NOMEM, "rte_submit: a solve of this shape needs %lld bytes, the budget is %lld",
                    r->footprint, S->opts.hbm_budget);
    if ((long long)S->queue.size() + S->waiting >= S->opts.max_queue)
        return fail(RTE_E_BUSY, "rte_submit: queue full");
    r->id = S->next_id++;
    r->t_submit = S->now_ms();
    *id = r->id;
    S->queue.push_back(r.get());
    S->reqs.emplace(r->id, std::move(r));
    S->cv.notify_all();
    return RTE_OK;
}

static int poll_locked(rte_server *S, long long id, double *exitance, rte_stats *stats, int *status)
{
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(RTE_E_NOTFOUND, "request %lld unknown or already collected", id);
    ServReq *r = it->second.get();
    if (status) *status = r->state;
    if (r->state == REQ_DONE) {
        if (exitance) std::memcpy(exitance, r->img.data(), sizeof(double) * r->img.size());
        if (stats) *stats = r->st;
        S->reqs.erase(it);
        return RTE_OK;
    }
    if (r->state == REQ_FAILED) {
        const int rc = r->rc ? r->rc : RTE_E_DEVICE;
        g_err = r->err;
        S->reqs.erase(it);
        return rc;
    }
    if (r->state == REQ_CANCELLED) S->reqs.erase(it);
    return RTE_OK;
}

extern "C" int rte_poll(rte_server *S, long long id, double *exitance, rte_stats *stats, int *status)
{
    if (!S) return fail(RTE_E_ARG, "rte_poll: NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    return poll_locked(S, id, exitance, stats, status);
}

extern "C" int rte_wait(rte_server *S, long long id, double timeout_ms, double *exitance, rte_stats *stats,
                        int *status)
{
    if (!S) return fail(RTE_E_ARG, "rte_wait: NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto ready = [&] {
        auto it = S->reqs.find(id);
        return it == S->reqs.end() || it->second->state >= REQ_DONE;
    };
    if (timeout_ms < 0) {
        S->cv_done.wait(lk, ready);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        if (status) *status = S->reqs[id]->state;
        return fail(RTE_E_BUSY, "rte_wait: request %lld not finished after %.0f ms", id, timeout_ms);
    }
    return poll_locked(S, id, exitance, stats, status);
}

extern "C" int rte_cancel(rte_server *S, long long id)
{
    if (!S) return fail(RTE_E_ARG, "rte_cancel: NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(RTE_E_NOTFOUND, "request %lld unknown or already collected", id);
    it->second->cancel.store(true);
    S->cv.notify_all();
    return RTE_OK;
}

extern "C" int rte_server_info(rte_server *S, rte_server_stats *out)
{
    if (!S || !out) return fail(RTE_E_ARG, "rte_server_info: NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(out, 0, sizeof *out);
    out->queued = (long long)S->queue.size() + S->waiting;
    out->running = S->running;
    out->done = S->n_done;
    out->failed = S->n_failed;
    out->cancelled = S->n_cancelled;
    out->steps = S->steps;
    out->hbm_used = S->hbm_used;
    out->hbm_budget = S->opts.hbm_budget;
    out->buckets = S->n_buckets;
    out->worker_status = S->worker_rc;
    return RTE_OK;
}
