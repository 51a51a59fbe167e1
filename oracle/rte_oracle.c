/*
 * rte_oracle.c -- FP64 CPU restatement of the reference Neumann-series RTE
 * solve.  TEST INFRASTRUCTURE ONLY (the checker of the HIP product path).
 *
 * Every function names the reference lines it restates.  Heterogeneous media
 * follow the GPU package's per-voxel generalisation (Siddon path integral,
 * column-attenuated source and ballistic term); in the uniform limit every
 * function reduces to the CPU package's form, which the tests pin against the
 * reference binary built from CPU package/src.
 *
 * Internally the field is held in the real SH basis (see rte_oracle.h); the
 * change of basis is unitary within each l, so every operator here equals
 * conj(U) M U^T of the reference's complex-basis operator M and the m = 0
 * coefficients (all the exitance reads) are identical.
 */
#include "rte_oracle.h"

#include <omp.h>

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_SQRT2
#define M_SQRT2 1.41421356237309504880
#endif
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ANG_RES 200 /* CPU package/src/Kernels_selfsub.c:737 */
#define NMOM 16     /* moment-series terms for the self block */

static int ncoef(int L) { return (L + 1) * (L + 1); }

/* ---------------------------------------------------------------------------
 * Spherical harmonics
 * ------------------------------------------------------------------------- */

/* P_l^m(x), m >= 0, Condon-Shortley phase included, by the upward three-term
 * recurrence in l from P_m^m = (-1)^m (2m-1)!! (1-x^2)^(m/2)
 * (CPU package/src/SphHarm.c:115-160, Abramowitz & Stegun 8.5.3). */
double orc_assoc_legendre(double x, int l, int m)
{
    double pmm = 1.0;
    if (m > 0) {
        double dfact = 1.0;
        for (int i = 2 * m - 1; i > 0; i -= 2) dfact *= (double)i;
        pmm = ((m & 1) ? -1.0 : 1.0) * dfact * pow(1.0 - x * x, 0.5 * (double)m);
    }
    if (l == m) return pmm;
    double prev = 0.0, cur = pmm;
    for (int ll = m + 1; ll <= l; ++ll) {
        double nxt = (x * (2.0 * ll - 1.0) * cur - (double)(ll + m - 1) * prev) / (double)(ll - m);
        prev = cur;
        cur = nxt;
    }
    return cur;
}

/* N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)  (CPU package/src/SphHarm.c:84-88) */
static double ylm_norm(int l, int m)
{
    double f = 1.0;
    for (int i = 0; i < 2 * m; ++i) f *= (double)(m + l - i);
    return sqrt((2.0 * l + 1.0) / (4.0 * M_PI * f));
}

/* Complex orthonormal Y_lm with Y_l,-m = (-1)^m conj(Y_lm)
 * (CPU package/src/SphHarm.c:75-103). */
void orc_ylm_complex(int L, double theta, double phi, double *re, double *im)
{
    double x = cos(theta);
    for (int l = 0; l <= L; ++l) {
        for (int m = 0; m <= l; ++m) {
            double v = ylm_norm(l, m) * orc_assoc_legendre(x, l, m);
            double c = v * cos((double)m * phi), s = v * sin((double)m * phi);
            re[l * l + l + m] = c;
            im[l * l + l + m] = s;
            if (m > 0) {
                double sg = (m & 1) ? -1.0 : 1.0;
                re[l * l + l - m] = sg * c;
                im[l * l + l - m] = -sg * s;
            }
        }
    }
}

/* Real orthonormal basis: R_l0 = Y_l0, R_lm = sqrt2 Re Y_lm, R_l,-m = sqrt2 Im Y_lm */
void orc_ylm_real(int L, double theta, double phi, double *out)
{
    double x = cos(theta);
    for (int l = 0; l <= L; ++l) {
        out[l * l + l] = ylm_norm(l, 0) * orc_assoc_legendre(x, l, 0);
        for (int m = 1; m <= l; ++m) {
            double v = M_SQRT2 * ylm_norm(l, m) * orc_assoc_legendre(x, l, m);
            out[l * l + l + m] = v * cos((double)m * phi);
            out[l * l + l - m] = v * sin((double)m * phi);
        }
    }
}

/* ---------------------------------------------------------------------------
 * Self-voxel block
 * ------------------------------------------------------------------------- */

/* Centre-to-face distance beta(omega) on the 200x200 (theta, phi) grid,
 * restating CPU package/src/Kernels_selfsub.c:815-948 with selfSubVox = 1
 * (forced by CPU package/src/RTE_uniform.c:63): one sub-voxel at the centre of
 * a [-1,1]^3 cube.  The reference holds the direction cosines and the face
 * intersection in float, keeps the last accepted intersection across
 * directions, and tests the x, y, z faces in that order; all of that is kept,
 * since beta enters the block at float precision. */
void orc_self_beta(double dx, double dy, double dz, double *beta)
{
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    const float face = 1.0f;
    int w = 0;
    for (int t = 0; t < ANG_RES; ++t) {
        for (int p = 0; p < ANG_RES; ++p, ++w) {
            double th = t * M_PI / ANG_RES;
            double ph = p * 2.0 * M_PI / ANG_RES;
            float ux = (float)(cos(ph) * sin(th));
            float uy = (float)(sin(ph) * sin(th));
            float uz = (float)cos(th);
            if (ux != 0.0f) {
                for (int f = 0; f < 2; ++f) {
                    float fx = f == 0 ? face : -face;
                    float fy = (float)((double)fx * (double)uy / (double)ux);
                    float fz = (float)((double)fx * (double)uz / (double)ux);
                    if (fx <= face && fx >= -face && fy <= face && fy >= -face && fz <= face &&
                        fz >= -face && ux * fx >= 0.0f) {
                        cx = fx; cy = fy; cz = fz;
                    }
                }
            }
            if (uy != 0.0f) {
                for (int f = 0; f < 2; ++f) {
                    float fy = f == 0 ? face : -face;
                    float fz = (float)((double)fy * (double)uz / (double)uy);
                    float fx = (float)((double)fy * (double)ux / (double)uy);
                    if (fx <= face && fx >= -face && fy <= face && fy >= -face && fz <= face &&
                        fz >= -face && uy * fy >= 0.0f) {
                        cx = fx; cy = fy; cz = fz;
                    }
                }
            }
            if (uz != 0.0f) {
                for (int f = 0; f < 2; ++f) {
                    float fz = f == 0 ? face : -face;
                    float fx = (float)((double)fz * (double)ux / (double)uz);
                    float fy = (float)((double)fz * (double)uy / (double)uz);
                    if (fx <= face && fx >= -face && fy <= face && fy >= -face && fz <= face &&
                        fz >= -face && uz * fz >= 0.0f) {
                        cx = fx; cy = fy; cz = fz;
                    }
                }
            }
            double ex = (double)cx, ey = (double)cy, ez = (double)cz;
            beta[w] = sqrt(ex * ex * (dx / 2.0) * (dx / 2.0) + ey * ey * (dy / 2.0) * (dy / 2.0) +
                           ez * ez * (dz / 2.0) * (dz / 2.0));
        }
    }
}

/* R_a(omega) sin(theta) on the 200x200 grid, K rows of 40000 */
static double *self_grid_harmonics(int L, double **sin_out)
{
    int K = ncoef(L);
    double *R = malloc(sizeof(double) * (size_t)K * ANG_RES * ANG_RES);
    double *sn = malloc(sizeof(double) * ANG_RES * ANG_RES);
    double *tmp = malloc(sizeof(double) * K);
    int w = 0;
    for (int t = 0; t < ANG_RES; ++t) {
        for (int p = 0; p < ANG_RES; ++p, ++w) {
            double th = t * M_PI / ANG_RES;
            double ph = p * 2.0 * M_PI / ANG_RES;
            orc_ylm_real(L, th, ph, tmp);
            for (int a = 0; a < K; ++a) R[(size_t)a * ANG_RES * ANG_RES + w] = tmp[a];
            sn[w] = sin(th);
        }
    }
    free(tmp);
    *sin_out = sn;
    return R;
}

/* M_k[a][b] = sum_w R_a R_b beta^k sin(theta), k = 1..nmom (layout [k-1][a][b]) */
int orc_self_moments(int L, double dx, double dy, double dz, int nmom, double *M)
{
    int K = ncoef(L), NW = ANG_RES * ANG_RES;
    double *beta = malloc(sizeof(double) * NW), *sn;
    orc_self_beta(dx, dy, dz, beta);
    double *R = self_grid_harmonics(L, &sn);
#pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < K; ++a) {
        double *acc = calloc((size_t)nmom * K, sizeof(double));
        for (int w = 0; w < NW; ++w) {
            double ra = R[(size_t)a * NW + w] * sn[w];
            if (ra == 0.0) continue;
            double bk = 1.0;
            for (int k = 0; k < nmom; ++k) {
                bk *= beta[w];
                double f = ra * bk;
                for (int b = 0; b < K; ++b) acc[k * K + b] += f * R[(size_t)b * NW + w];
            }
        }
        for (int k = 0; k < nmom; ++k)
            for (int b = 0; b < K; ++b) M[((size_t)k * K + a) * K + b] = acc[k * K + b];
        free(acc);
    }
    free(R); free(sn); free(beta);
    return 0;
}

/* S_ab(mu) = dtheta dphi / (c_m mu dV) sum_w R_a R_b (1 - exp(-mu beta)) sin(theta)
 * (CPU package/src/Kernels_selfsub.c:948-959).  method 1 evaluates the same sum
 * through (1 - e^{-mu beta})/mu = sum_k (-mu)^{k-1} beta^k / k!. */
int orc_self_block(int L, double dx, double dy, double dz, double mu_t, double c_m,
                   int method, double *S)
{
    int K = ncoef(L), NW = ANG_RES * ANG_RES;
    double dth = M_PI / ANG_RES, dph = 2.0 * M_PI / ANG_RES;
    double dV = dx * dy * dz;
    if (method == 0) {
        double *beta = malloc(sizeof(double) * NW), *sn;
        orc_self_beta(dx, dy, dz, beta);
        double *R = self_grid_harmonics(L, &sn);
        double *wt = malloc(sizeof(double) * NW);
        for (int w = 0; w < NW; ++w) wt[w] = (1.0 - exp(-mu_t * beta[w])) * sn[w];
        double scale = dth * dph / (c_m * mu_t * dV);
#pragma omp parallel for schedule(dynamic)
        for (int a = 0; a < K; ++a) {
            for (int b = 0; b < K; ++b) {
                double s = 0.0;
                for (int w = 0; w < NW; ++w) s += R[(size_t)a * NW + w] * R[(size_t)b * NW + w] * wt[w];
                S[a * K + b] = s * scale;
            }
        }
        free(R); free(sn); free(beta); free(wt);
        return 0;
    }
    double *M = malloc(sizeof(double) * (size_t)NMOM * K * K);
    orc_self_moments(L, dx, dy, dz, NMOM, M);
    double scale = dth * dph / (c_m * dV);
    for (int ab = 0; ab < K * K; ++ab) S[ab] = 0.0;
    double coef = 1.0; /* (-mu)^{k-1}/k! */
    for (int k = 0; k < NMOM; ++k) {
        coef = (k == 0) ? 1.0 : coef * (-mu_t) / (double)(k + 1);
        for (int ab = 0; ab < K * K; ++ab) S[ab] += coef * M[(size_t)k * K * K + ab];
    }
    for (int ab = 0; ab < K * K; ++ab) S[ab] *= scale;
    free(M);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Geometry helpers
 * ------------------------------------------------------------------------- */

typedef struct {
    int nx, ny, nz, N;
    double dx, dy, dz, dV;
    double *xs, *ys, *zs;
} geo;

/* CPU package/src/Inits.c:106-122 */
static void geo_init(geo *G, const orc_grid *g)
{
    G->nx = g->nx; G->ny = g->ny; G->nz = g->nz;
    G->N = g->nx * g->ny * g->nz;
    G->dx = (g->xmax - g->xmin) / (double)g->nx;
    G->dy = (g->ymax - g->ymin) / (double)g->ny;
    G->dz = (g->zmax - g->zmin) / (double)g->nz;
    G->dV = G->dx * G->dy * G->dz;
    G->xs = malloc(sizeof(double) * g->nx);
    G->ys = malloc(sizeof(double) * g->ny);
    G->zs = malloc(sizeof(double) * g->nz);
    for (int i = 0; i < g->nx; ++i) G->xs[i] = (double)i * G->dx + g->xmin + G->dx / 2.0;
    for (int i = 0; i < g->ny; ++i) G->ys[i] = (double)i * G->dy + g->ymin + G->dy / 2.0;
    for (int i = 0; i < g->nz; ++i) G->zs[i] = (double)i * G->dz + g->zmin + G->dz / 2.0;
}

static void geo_free(geo *G) { free(G->xs); free(G->ys); free(G->zs); }

/* ---------------------------------------------------------------------------
 * Siddon path integral
 * ------------------------------------------------------------------------- */

/* tau(d, s) = sum over the voxels crossed by the segment from the centre of s
 * to the centre of d of mu_t(v) * length(v), restating mu_int
 * (GPU package/src/kernels.cu:49-145; FP64 here, FP32 alphas there):
 *   - the first crossing along an axis with |delta| steps is at alpha = 0.5/|delta|,
 *     later ones every 1/|delta| (kernels.cu:65-77, 93-94);
 *   - the next crossing is taken along z if it is <= both others, else along y
 *     if < z and <= x, else along x (kernels.cu:89, 108, 126);
 *   - entering d adds the half step mu_t(d) * r * 0.5/|delta| (kernels.cu:100-105). */
static double tau_pair(const geo *G, const double *mut, int di, int dj, int dk, int si, int sj,
                       int sk, double r)
{
    int n[3] = {abs(di - si), abs(dj - sj), abs(dk - sk)};  /* x, y, z */
    int st[3] = {di > si ? 1 : -1, dj > sj ? 1 : -1, dk > sk ? 1 : -1};
    double nxt[3], inc[3];
    for (int a = 0; a < 3; ++a) {
        nxt[a] = n[a] ? 0.5 / (double)n[a] : INFINITY;
        inc[a] = n[a] ? 1.0 / (double)n[a] : INFINITY;
    }
    int c[3] = {si, sj, sk};
    double prev = 0.0, tau = 0.0;
    int nsteps = n[0] + n[1] + n[2], ax = 0;
    for (int s = 0; s < nsteps; ++s) {
        if (nxt[2] <= nxt[0] && nxt[2] <= nxt[1]) ax = 2;
        else if (nxt[1] < nxt[2] && nxt[1] <= nxt[0]) ax = 1;
        else ax = 0;
        int v = c[0] + G->nx * (c[1] + G->ny * c[2]);
        tau += mut[v] * r * (nxt[ax] - prev);
        prev = nxt[ax];
        nxt[ax] += inc[ax];
        c[ax] += st[ax];
    }
    /* now c == d: the half step inside the destination voxel, 0.5/|delta| along
     * the axis of the last crossing */
    int vd = di + G->nx * (dj + G->ny * dk);
    tau += mut[vd] * r * 0.5 * inc[ax];
    return tau;
}

double orc_tau(const orc_grid *g, const double *mu_t, int d, int s)
{
    geo G;
    geo_init(&G, g);
    int di = d % G.nx, dj = (d / G.nx) % G.ny, dk = d / (G.nx * G.ny);
    int si = s % G.nx, sj = (s / G.nx) % G.ny, sk = s / (G.nx * G.ny);
    double ddx = G.xs[di] - G.xs[si], ddy = G.ys[dj] - G.ys[sj], ddz = G.zs[dk] - G.zs[sk];
    double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    double t = (d == s) ? 0.0 : tau_pair(&G, mu_t, di, dj, dk, si, sj, sk, r);
    geo_free(&G);
    return t;
}

/* ---------------------------------------------------------------------------
 * Source and ballistic term
 * ------------------------------------------------------------------------- */

/* Beam columns: rad <= diam/2 with diam + 1e-16 as LoadSource stores it
 * (CPU package/src/Inits.c:56, Dist.c:1090-1092); power normalised over the
 * columns (Dist.c:1113-1114).  Column-attenuated first-scatter source
 *   s0_l0(i,j,k) = fact g^l (mu_s/mu_t) sqrt((2l+1)/4pi) T(k) (1 - e^{-mu_t dz}) / dV,
 *   T(k) = prod_{k'<k} e^{-mu_t(k') dz}
 * (GPU package/src/dist.cu:83-91), which is Dist.c:1097-1103 for uniform media;
 * ballistic exitance fact * T(nz) on the beam columns (GPU package/src/
 * dist.cu:129-141; Dist.c:203-209). */
int orc_source(const orc_grid *g, const orc_optics *o, const orc_beam *b, int L, double *s0,
               double *ballistic)
{
    geo G;
    geo_init(&G, g);
    int K = ncoef(L), N = G.N;
    double diam = b->diam + 1.0e-16;
    memset(s0, 0, sizeof(double) * (size_t)K * N);
    int cols = 0;
    for (int j = 0; j < G.ny; ++j)
        for (int i = 0; i < G.nx; ++i) {
            double rad = sqrt((G.xs[i] - b->pos_x) * (G.xs[i] - b->pos_x) +
                              (G.ys[j] - b->pos_y) * (G.ys[j] - b->pos_y));
            if (rad <= diam / 2.0) ++cols;
        }
    if (cols == 0) { geo_free(&G); return 0; }
    double fact = b->mag / (double)cols;
    for (int j = 0; j < G.ny; ++j)
        for (int i = 0; i < G.nx; ++i) {
            double rad = sqrt((G.xs[i] - b->pos_x) * (G.xs[i] - b->pos_x) +
                              (G.ys[j] - b->pos_y) * (G.ys[j] - b->pos_y));
            if (ballistic) ballistic[i + G.nx * j] = 0.0;
            if (!(rad <= diam / 2.0)) continue;
            double T = 1.0;
            for (int k = 0; k < G.nz; ++k) {
                int v = i + G.nx * (j + G.ny * k);
                double mt = o->mu_a[v] + o->mu_s[v];
                double att = exp(-mt * G.dz);
                double base = fact * (o->mu_s[v] / mt) * T * (1.0 - att) / G.dV;
                double gl = 1.0;
                for (int l = 0; l <= L; ++l) {
                    s0[(size_t)(l * l + l) * N + v] = gl * sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) * base;
                    gl *= o->g;
                }
                T *= att;
            }
            if (ballistic) ballistic[i + G.nx * j] = fact * T;
        }
    geo_free(&G);
    return cols;
}

/* ---------------------------------------------------------------------------
 * Propagation A
 * ------------------------------------------------------------------------- */

static int is_uniform(const double *x, int n)
{
    for (int i = 1; i < n; ++i)
        if (x[i] != x[0]) return 0;
    return 1;
}

/* (A W)_a(d) = dV [ sum_b S_ab(mu_t(d)) W_b(d)
 *   + sum_{s != d, r < R} e^{-tau(d,s)} / (c_m r^2) R_a(u) sum_b R_b(u) W_b(s) ],
 * u = (r_d - r_s)/r, theta = acos(u_z), phi = atan2(u_y, u_x)
 * (CPU package/src/Kernels_selfsub.c:450-538 with the table of :806-818,1029-1053;
 * the sphere r < R as in GPU package/src/kernels.cu:327; self term
 * GPU package/src/kernels.cu:148-175 evaluated at the destination's mu_t).
 * The CPU package's table holds 1/(c r^2) e^{-mu_t r} conj(Y) Y; the pair
 * kernel here is the same product, factored through the real basis. */
int orc_propagate(const orc_grid *g, const orc_optics *o, int L, double prop_thresh, const double *W,
                  double *out)
{
    return orc_propagate_range(g, o, L, prop_thresh, W, out, 0, g->nx * g->ny * g->nz);
}

static int propagate_range_impl(const orc_grid *g, const orc_optics *o, int L, double prop_thresh,
                                const double *W, double *out, int d_begin, int d_end, double *loop_s);

int orc_propagate_range(const orc_grid *g, const orc_optics *o, int L, double prop_thresh,
                        const double *W, double *out, int d_begin, int d_end)
{
    return propagate_range_impl(g, o, L, prop_thresh, W, out, d_begin, d_end, NULL);
}

/* as orc_propagate_range; *loop_s receives the wall time of the pair loop alone
 * (the tables built before it excluded): the CPU baseline's timer */
int orc_propagate_range_timed(const orc_grid *g, const orc_optics *o, int L, double prop_thresh,
                              const double *W, double *out, int d_begin, int d_end, double *loop_s)
{
    return propagate_range_impl(g, o, L, prop_thresh, W, out, d_begin, d_end, loop_s);
}

static int propagate_range_impl(const orc_grid *g, const orc_optics *o, int L, double prop_thresh,
                                const double *W, double *out, int d_begin, int d_end, double *loop_s)
{
    geo G;
    geo_init(&G, g);
    int K = ncoef(L), N = G.N;
    int NX = 2 * G.nx - 1, NY = 2 * G.ny - 1, NZ = 2 * G.nz - 1;
    size_t ND = (size_t)NX * NY * NZ;
    double c_m = o->c / o->n;
    double *mut = malloc(sizeof(double) * N);
    for (int v = 0; v < N; ++v) mut[v] = o->mu_a[v] + o->mu_s[v];
    int uniform = is_uniform(mut, N);

    /* displacement tables: R(u) and 1/(c_m r^2), r */
    double *Ytab = malloc(sizeof(double) * ND * K);
    double *rtab = malloc(sizeof(double) * ND);
#pragma omp parallel for
    for (long q = 0; q < (long)ND; ++q) {
        int ix = (int)(q % NX), iy = (int)((q / NX) % NY), iz = (int)(q / ((size_t)NX * NY));
        double ddx = G.dx * (ix - (G.nx - 1)), ddy = G.dy * (iy - (G.ny - 1)),
               ddz = G.dz * (iz - (G.nz - 1));
        double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        rtab[q] = r;
        if (r > 0.0) orc_ylm_real(L, acos(ddz / r), atan2(ddy, ddx), Ytab + q * K);
        else for (int a = 0; a < K; ++a) Ytab[q * K + a] = 0.0;
    }

    /* self blocks */
    double *S = malloc(sizeof(double) * (size_t)K * K * (uniform ? 1 : N));
    if (uniform) {
        orc_self_block(L, G.dx, G.dy, G.dz, mut[0], c_m, 0, S);
    } else {
        double *M = malloc(sizeof(double) * (size_t)NMOM * K * K);
        orc_self_moments(L, G.dx, G.dy, G.dz, NMOM, M);
        double scale = (M_PI / ANG_RES) * (2.0 * M_PI / ANG_RES) / (c_m * G.dV);
#pragma omp parallel for
        for (int v = 0; v < N; ++v) {
            double *Sv = S + (size_t)v * K * K;
            for (int ab = 0; ab < K * K; ++ab) Sv[ab] = 0.0;
            double coef = 1.0;
            for (int k = 0; k < NMOM; ++k) {
                coef = (k == 0) ? 1.0 : coef * (-mut[v]) / (double)(k + 1);
                for (int ab = 0; ab < K * K; ++ab) Sv[ab] += coef * M[(size_t)k * K * K + ab];
            }
            for (int ab = 0; ab < K * K; ++ab) Sv[ab] *= scale;
        }
        free(M);
    }

    const double t_loop = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 4)
    for (int d = d_begin; d < d_end; ++d) {
        int di = d % G.nx, dj = (d / G.nx) % G.ny, dk = d / (G.nx * G.ny);
        double acc[64];
        for (int a = 0; a < K; ++a) acc[a] = 0.0;
        for (int s = 0; s < N; ++s) {
            if (s == d) continue;
            int si = s % G.nx, sj = (s / G.nx) % G.ny, sk = s / (G.nx * G.ny);
            size_t q = (size_t)(di - si + G.nx - 1) +
                       (size_t)NX * ((size_t)(dj - sj + G.ny - 1) + (size_t)NY * (dk - sk + G.nz - 1));
            double r = rtab[q];
            if (!(r < prop_thresh)) continue;
            double tau = uniform ? mut[0] * r : tau_pair(&G, mut, di, dj, dk, si, sj, sk, r);
            double kf = exp(-tau) / (c_m * r * r);
            const double *y = Ytab + q * K;
            double p = 0.0;
            for (int b = 0; b < K; ++b) p += y[b] * W[(size_t)b * N + s];
            p *= kf;
            for (int a = 0; a < K; ++a) acc[a] += p * y[a];
        }
        const double *Sd = S + (uniform ? 0 : (size_t)d * K * K);
        for (int a = 0; a < K; ++a) {
            double s = 0.0;
            for (int b = 0; b < K; ++b) s += Sd[a * K + b] * W[(size_t)b * N + d];
            out[(size_t)a * N + d] = (acc[a] + s) * G.dV;
        }
    }
    if (loop_s) *loop_s = omp_get_wtime() - t_loop;
    free(Ytab); free(rtab); free(S); free(mut);
    geo_free(&G);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Scatter, exitance, stop rule
 * ------------------------------------------------------------------------- */

/* W_a(v) <- c_m g^l mu_s(v) W_a(v)  (CPU package/src/Kernels_selfsub.c:30-51) */
void orc_scatter(const orc_grid *g, const orc_optics *o, int L, double *W)
{
    int N = g->nx * g->ny * g->nz;
    double c_m = o->c / o->n;
    for (int l = 0; l <= L; ++l) {
        double gl = pow(o->g, (double)l);
        for (int m = -l; m <= l; ++m) {
            double *row = W + (size_t)(l * l + l + m) * N;
            for (int v = 0; v < N; ++v) row[v] = c_m * gl * o->mu_s[v] * row[v];
        }
    }
}

/* Q_l = int_0^1 P_l(u) u du: the reference's 100-point left Riemann sum
 * (CPU package/src/Dist.c:190-193,211-218), or exact (32-point Gauss-Legendre
 * on [0,1], exact for l <= 61). */
void orc_quad_weights(int L, int quad, double *Q)
{
    if (quad == ORC_QUAD_REF) {
        for (int l = 0; l <= L; ++l) {
            double s = 0.0;
            for (int p = 0; p < 100; ++p) {
                double u = (double)p / 100.0;
                s += orc_assoc_legendre(u, l, 0) * u / 100.0;
            }
            Q[l] = s;
        }
        return;
    }
    enum { NG = 32 };
    double xg[NG], wg[NG];
    for (int i = 0; i < NG; ++i) {
        double x = cos(M_PI * (i + 0.75) / (NG + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= NG; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            pp = NG * (x * p1 - p0) / (x * x - 1.0);
            double dxn = p1 / pp;
            x -= dxn;
            if (fabs(dxn) < 1e-16) break;
        }
        xg[i] = x;
        wg[i] = 2.0 / ((1.0 - x * x) * pp * pp);
    }
    for (int l = 0; l <= L; ++l) {
        double s = 0.0;
        for (int i = 0; i < NG; ++i) {
            double u = 0.5 * (xg[i] + 1.0);
            s += 0.5 * wg[i] * orc_assoc_legendre(u, l, 0) * u;
        }
        Q[l] = s;
    }
}

/* R_l = int_{-1}^{0} P_l(u) u du: the reference's left Riemann sum over
 * u = p/100 - 1, p = 0..99 (CPU package/src/Dist.c:146-160), or exactly,
 * R_l = (-1)^(l+1) int_0^1 P_l(u) u du since P_l(-u) = (-1)^l P_l(u). */
void orc_quad_weights_refl(int L, int quad, double *R)
{
    if (quad == ORC_QUAD_REF) {
        for (int l = 0; l <= L; ++l) {
            double s = 0.0;
            for (int p = 0; p < 100; ++p) {
                double u = (double)p / 100.0 - 1.0;
                s += orc_assoc_legendre(u, l, 0) * u / 100.0;
            }
            R[l] = s;
        }
        return;
    }
    orc_quad_weights(L, quad, R);
    for (int l = 0; l <= L; ++l) R[l] *= (l & 1) ? 1.0 : -1.0;
}

/* Top-face reflectance R_ij = -sum_l 2 pi dx dy c_m / sqrt(4 pi) sqrt(2l+1) R_l
 * w_l0(i, j, 0) (CPU package/src/Dist.c:134-171); no ballistic term. */
void orc_reflectance(const orc_grid *g, const orc_optics *o, int L, int quad, const double *w,
                     double *img)
{
    geo G;
    geo_init(&G, g);
    double R[128];
    orc_quad_weights_refl(L, quad, R);
    double c_m = o->c / o->n;
    for (int j = 0; j < G.ny; ++j)
        for (int i = 0; i < G.nx; ++i) {
            int v = i + G.nx * j;
            double val = 0.0;
            for (int l = 0; l <= L; ++l)
                val -= 2.0 * M_PI * G.dx * G.dy * c_m / sqrt(4.0 * M_PI) * sqrt(2.0 * l + 1.0) * R[l] *
                       w[(size_t)(l * l + l) * G.N + v];
            img[i + G.nx * j] = val;
        }
    geo_free(&G);
}

/* E_ij = sum_l 2 pi dx dy c_m / sqrt(4 pi) sqrt(2l+1) Q_l w_l0(i, j, nz-1)
 * (CPU package/src/Dist.c:211-218), ballistic term excluded. */
void orc_exitance(const orc_grid *g, const orc_optics *o, int L, int quad, const double *w,
                  double *img)
{
    geo G;
    geo_init(&G, g);
    double Q[128];
    orc_quad_weights(L, quad, Q);
    double c_m = o->c / o->n;
    int k = G.nz - 1;
    for (int j = 0; j < G.ny; ++j)
        for (int i = 0; i < G.nx; ++i) {
            int v = i + G.nx * (j + G.ny * k);
            double val = 0.0;
            for (int l = 0; l <= L; ++l)
                val += 2.0 * M_PI * G.dx * G.dy * c_m / sqrt(4.0 * M_PI) * sqrt(2.0 * l + 1.0) * Q[l] *
                       w[(size_t)(l * l + l) * G.N + v];
            img[i + G.nx * j] = val;
        }
    geo_free(&G);
}

/* Sum over the bottom layer of Re of every complex-basis coefficient
 * (CPU package/src/Kernels_selfsub.c:1194-1206).  For a real radiance
 * Re c_l0 = R_l0, Re c_lm + Re c_l,-m = (1 + (-1)^m) R_lm / sqrt2 (m > 0). */
double orc_bottom_sum(const orc_grid *g, int L, const double *W)
{
    int N = g->nx * g->ny * g->nz, base = g->nx * g->ny * (g->nz - 1);
    double s = 0.0;
    for (int l = 0; l <= L; ++l)
        for (int m = -l; m <= l; ++m) {
            double wgt = (m == 0) ? 1.0 : ((m > 0 && (m % 2) == 0) ? M_SQRT2 : 0.0);
            if (wgt == 0.0) continue;
            const double *row = W + (size_t)(l * l + l + m) * N + base;
            double t = 0.0;
            for (int v = 0; v < g->nx * g->ny; ++v) t += row[v];
            s += wgt * t;
        }
    return s;
}

/* lambda rule (CPU package/src/Kernels_selfsub.c:1184-1233, NSTOP of RTE.h:94) */
static int stop_lambda(double *fi, int n, double sumW, double sumOut, double *lam)
{
    double hi = 0.0;
    if (n > 1) {
        if (sumOut == 0.0) { *lam = 0.0; return 1; }
        double gam = sumW / sumOut;
        hi = 1.0 - gam * (1.0 / (1.0 - gam / *fi));
        *fi = gam;
        if (hi > 1.0 || hi < 0.0) hi = 0.0;
    }
    *lam = hi;
    return hi > 0.9999;
}

/* The Neumann series from the first term s0 (K*N): exitance = ball + the
 * bottom exitance of A sum_q (DA)^q s0 (ball may be NULL: no ballistic term);
 * w_out (may be NULL) receives A sum_q (DA)^q s0. */
static int neumann_from(const orc_grid *g, const orc_optics *o, const orc_params *p, const double *s0,
                        const double *ball, double *exitance, double *w_out, double *refl, orc_stats *st)
{
    int L = p->L, K = ncoef(L), N = g->nx * g->ny * g->nz, P = g->nx * g->ny;
    double *W = malloc(sizeof(double) * (size_t)K * N);
    double *AW = malloc(sizeof(double) * (size_t)K * N);
    double *out = malloc(sizeof(double) * (size_t)K * N);
    double *img = malloc(sizeof(double) * P);
    double btot = 0.0;
    if (ball)
        for (int i = 0; i < P; ++i) btot += ball[i];
    memcpy(W, s0, sizeof(double) * (size_t)K * N);
    memcpy(out, s0, sizeof(double) * (size_t)K * N);
    /* with g == 0 only l = 0 survives the scatter, and the reference propagates
     * l = 0 alone inside the loop (CPU package/src/RTE_uniform.c:98-106) */
    int Lloop = (o->g != 0.0) ? L : 0;
    int n, max_terms = p->max_terms > 0 ? p->max_terms : 1000;
    double lam = 0.0, rel = 0.0;
    if (p->stop_rule == ORC_STOP_LAMBDA) {
        /* CPU package/src/RTE_uniform.c:95-111 */
        double fi = 1.0;
        n = 1;
        while (!stop_lambda(&fi, n, orc_bottom_sum(g, L, W), orc_bottom_sum(g, L, out), &lam)) {
            if (n >= max_terms) break;
            orc_propagate(g, o, Lloop, p->prop_thresh, W, AW);
            if (Lloop < L) memset(AW + (size_t)ncoef(Lloop) * N, 0, sizeof(double) * (size_t)(K - ncoef(Lloop)) * N);
            orc_scatter(g, o, L, AW);
            memcpy(W, AW, sizeof(double) * (size_t)K * N);
            for (size_t i = 0; i < (size_t)K * N; ++i) out[i] += W[i];
            ++n;
        }
        orc_propagate(g, o, L, p->prop_thresh, out, AW);
        orc_exitance(g, o, L, p->exit_quad, AW, img);
        for (int i = 0; i < P; ++i) exitance[i] = (ball ? ball[i] : 0.0) + img[i];
        if (refl) orc_reflectance(g, o, L, p->exit_quad, AW, refl);
        if (w_out) memcpy(w_out, AW, sizeof(double) * (size_t)K * N);
    } else {
        /* relative change of the total exitance: by linearity of A the exitance
         * of A out_n is the running sum of the exitances of A W_q */
        double T = btot;
        for (int i = 0; i < P; ++i) exitance[i] = ball ? ball[i] : 0.0;
        if (w_out) memset(w_out, 0, sizeof(double) * (size_t)K * N);
        if (refl) memset(refl, 0, sizeof(double) * P);
        n = 0;
        for (;;) {
            orc_propagate(g, o, L, p->prop_thresh, W, AW);
            orc_exitance(g, o, L, p->exit_quad, AW, img);
            double dT = 0.0;
            for (int i = 0; i < P; ++i) { exitance[i] += img[i]; dT += img[i]; }
            if (w_out) for (size_t i = 0; i < (size_t)K * N; ++i) w_out[i] += AW[i];
            if (refl) {
                orc_reflectance(g, o, L, p->exit_quad, AW, img);
                for (int i = 0; i < P; ++i) refl[i] += img[i];
                orc_exitance(g, o, L, p->exit_quad, AW, img);
            }
            T += dT;
            ++n;
            rel = (T != 0.0) ? fabs(dT) / fabs(T) : 0.0;
            if (rel < p->tol || n >= max_terms) break;
            orc_scatter(g, o, L, AW);
            memcpy(W, AW, sizeof(double) * (size_t)K * N);
        }
    }
    if (st) {
        st->terms = n;
        st->lambda_last = lam;
        st->rel_change_last = rel;
        st->ballistic_total = btot;
    }
    free(W); free(AW); free(out); free(img);
    return 0;
}

static int solve_impl(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p,
                      double *exitance, double *w_out, double *refl, orc_stats *st)
{
    int L = p->L, K = ncoef(L), N = g->nx * g->ny * g->nz, P = g->nx * g->ny;
    double *s0 = malloc(sizeof(double) * (size_t)K * N);
    double *ball = malloc(sizeof(double) * P);
    int rc = -1;
    if (orc_source(g, o, b, L, s0, ball)) rc = neumann_from(g, o, p, s0, ball, exitance, w_out, refl, st);
    free(s0);
    free(ball);
    return rc;
}

int orc_solve(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p,
              double *exitance, orc_stats *st)
{
    return solve_impl(g, o, b, p, exitance, NULL, NULL, st);
}

int orc_solve_images(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p,
                     double *exitance, double *reflectance, orc_stats *st)
{
    return solve_impl(g, o, b, p, exitance, NULL, reflectance, st);
}

/* FisherInfo_IO/src/compute_grad_rte.cu:241-301 (mu_a) and :341-371 (mu_s):
 * out2 = the forward field A sum_q (DA)^q s0; the derivative source is out2
 * restricted to the region and scaled by -C (mu_a) or by -C + g^l (mu_s,
 * PropScatmu1, GPU package/src/dist.cu:675-686); Neumann(src) and the
 * transmitted image without the ballistic term (generate_trans_image(out3, 0)).
 * C is the vacuum speed of light (#define C 1, FisherInfo_IO/src/rte.h:15; the
 * optics' c here, in whatever units the optics use), so the source is
 * -c phi|_R (mu_a) and (g^l - 1) c phi|_R (mu_s: PropScatmu1's g^l times the
 * field, plus -C times it), phi in the reference's units (A carrying 1/c_m).
 * For n != 1 this is n times the source -c_m phi|_R, c_m = c/n.  The region is
 * applied as intended in both branches (the reference's mu_s branch indexes the
 * region with a stale voxel index, compute_grad_rte.cu:357-366). */
int orc_ref_derivative(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p,
                       const int *region, int which, double *dimg, orc_stats *st)
{
    int L = p->L, K = ncoef(L), N = g->nx * g->ny * g->nz, P = g->nx * g->ny;
    double *phi = malloc(sizeof(double) * (size_t)K * N);
    double *img = malloc(sizeof(double) * P);
    int rc = solve_impl(g, o, b, p, img, phi, NULL, NULL);
    if (rc == 0) {
        for (int a = 0; a < K; ++a) {
            const int l = (int)sqrt((double)a + 0.5);
            const double f = o->c * (which == 0 ? -1.0 : pow(o->g, (double)l) - 1.0);
            for (int v = 0; v < N; ++v) phi[(size_t)a * N + v] = region[v] ? f * phi[(size_t)a * N + v] : 0.0;
        }
        rc = neumann_from(g, o, p, phi, NULL, dimg, NULL, NULL, st);
    }
    free(phi);
    free(img);
    return rc;
}

int orc_solve_field(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p,
                    double *exitance, double *w_out, orc_stats *st)
{
    return solve_impl(g, o, b, p, exitance, w_out, NULL, st);
}

/* FisherInfo_IO/src/neumann.cu:23-120 with flag 1, called with out preset to src
 * (FisherInfo_IO/src/rte_fim.cu:235-237, 275-276): the series of neumann_from
 * from the given first term, no ballistic term */
int orc_neumann(const orc_grid *g, const orc_optics *o, const orc_params *p, const double *src,
                double *w_out, double *exitance, double *refl, orc_stats *st)
{
    return neumann_from(g, o, p, src, NULL, exitance, w_out, refl, st);
}

/* ---------------------------------------------------------------------------
 * Frequency domain: complex mu_t = mu_a + mu_s + i kim, kim = omega n / c
 * (CPU package/src/RTE_uniform.c:126).  Fields are complex (interleaved
 * re, im doubles), still in the real SH basis: A and S become complex, D
 * stays real.  Pair factor e^{-mu_t r}/(c r^2) with complex mu_t
 * (Kernels_selfsub.c:1046; e^{-tau - i kim r} for a heterogeneous medium),
 * self block (1 - e^{-mu beta}) sin(theta) / (c mu dV) (Kernels_selfsub.c:927,
 * 948), source and ballistic with the complex column attenuation
 * (Dist.c:1097-1103 evaluated with complex mu_t).
 * ------------------------------------------------------------------------- */

typedef double complex dcx;

/* S(mu) for complex mu, interleaved K x K; method 0 direct 200x200 quadrature
 * (the reference's), 1 the moment series with complex coefficients. */
int orc_self_block_c(int L, double dx, double dy, double dz, double mu_re, double mu_im, double c_m, int method,
                     double *S_out)
{
    int K = ncoef(L), NW = ANG_RES * ANG_RES;
    double dth = M_PI / ANG_RES, dph = 2.0 * M_PI / ANG_RES;
    double dV = dx * dy * dz;
    dcx mu = CMPLX(mu_re, mu_im);
    dcx *S = (dcx *)S_out;
    if (method == 0) {
        double *beta = malloc(sizeof(double) * NW), *sn;
        orc_self_beta(dx, dy, dz, beta);
        double *R = self_grid_harmonics(L, &sn);
        dcx *wt = malloc(sizeof(dcx) * NW);
        for (int w = 0; w < NW; ++w) wt[w] = (1.0 - cexp(-mu * beta[w])) * sn[w];
        dcx scale = dth * dph / (c_m * mu * dV);
#pragma omp parallel for schedule(dynamic)
        for (int a = 0; a < K; ++a) {
            for (int b = 0; b < K; ++b) {
                dcx s = 0.0;
                for (int w = 0; w < NW; ++w) s += R[(size_t)a * NW + w] * R[(size_t)b * NW + w] * wt[w];
                S[a * K + b] = s * scale;
            }
        }
        free(R); free(sn); free(beta); free(wt);
        return 0;
    }
    double *M = malloc(sizeof(double) * (size_t)NMOM * K * K);
    orc_self_moments(L, dx, dy, dz, NMOM, M);
    double scale = dth * dph / (c_m * dV);
    for (int ab = 0; ab < K * K; ++ab) S[ab] = 0.0;
    dcx coef = 1.0;
    for (int k = 0; k < NMOM; ++k) {
        coef = (k == 0) ? 1.0 : coef * (-mu) / (double)(k + 1);
        for (int ab = 0; ab < K * K; ++ab) S[ab] += coef * M[(size_t)k * K * K + ab];
    }
    for (int ab = 0; ab < K * K; ++ab) S[ab] *= scale;
    free(M);
    return 0;
}

/* complex A on a complex K*N field (interleaved), every destination */
int orc_propagate_c(const orc_grid *g, const orc_optics *o, int L, double prop_thresh, double kim,
                    const double *W_in, double *out_)
{
    return orc_propagate_c_range(g, o, L, prop_thresh, kim, W_in, out_, 0, g->nx * g->ny * g->nz);
}

/* the same restricted to destinations d_begin <= d < d_end (other rows untouched) */
int orc_propagate_c_range(const orc_grid *g, const orc_optics *o, int L, double prop_thresh, double kim,
                          const double *W_in, double *out_, int d_begin, int d_end)
{
    geo G;
    geo_init(&G, g);
    int K = ncoef(L), N = G.N;
    const dcx *W = (const dcx *)W_in;
    dcx *out = (dcx *)out_;
    int NX = 2 * G.nx - 1, NY = 2 * G.ny - 1, NZ = 2 * G.nz - 1;
    size_t ND = (size_t)NX * NY * NZ;
    double c_m = o->c / o->n;
    double *mut = malloc(sizeof(double) * N);
    for (int v = 0; v < N; ++v) mut[v] = o->mu_a[v] + o->mu_s[v];
    int uniform = is_uniform(mut, N);
    double *Ytab = malloc(sizeof(double) * ND * K);
    double *rtab = malloc(sizeof(double) * ND);
#pragma omp parallel for
    for (long q = 0; q < (long)ND; ++q) {
        int ix = (int)(q % NX), iy = (int)((q / NX) % NY), iz = (int)(q / ((size_t)NX * NY));
        double ddx = G.dx * (ix - (G.nx - 1)), ddy = G.dy * (iy - (G.ny - 1)), ddz = G.dz * (iz - (G.nz - 1));
        double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        rtab[q] = r;
        if (r > 0.0) orc_ylm_real(L, acos(ddz / r), atan2(ddy, ddx), Ytab + q * K);
        else for (int a = 0; a < K; ++a) Ytab[q * K + a] = 0.0;
    }
    dcx *S = malloc(sizeof(dcx) * (size_t)K * K * (uniform ? 1 : N));
    if (uniform) {
        orc_self_block_c(L, G.dx, G.dy, G.dz, mut[0], kim, c_m, 0, (double *)S);
    } else {
        double *M = malloc(sizeof(double) * (size_t)NMOM * K * K);
        orc_self_moments(L, G.dx, G.dy, G.dz, NMOM, M);
        double scale = (M_PI / ANG_RES) * (2.0 * M_PI / ANG_RES) / (c_m * G.dV);
#pragma omp parallel for
        for (int v = 0; v < N; ++v) {
            dcx *Sv = S + (size_t)v * K * K, mu = CMPLX(mut[v], kim), coef = 1.0;
            for (int ab = 0; ab < K * K; ++ab) Sv[ab] = 0.0;
            for (int k = 0; k < NMOM; ++k) {
                coef = (k == 0) ? 1.0 : coef * (-mu) / (double)(k + 1);
                for (int ab = 0; ab < K * K; ++ab) Sv[ab] += coef * M[(size_t)k * K * K + ab];
            }
            for (int ab = 0; ab < K * K; ++ab) Sv[ab] *= scale;
        }
        free(M);
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int d = d_begin; d < d_end; ++d) {
        int di = d % G.nx, dj = (d / G.nx) % G.ny, dk = d / (G.nx * G.ny);
        dcx acc[64];
        for (int a = 0; a < K; ++a) acc[a] = 0.0;
        for (int s = 0; s < N; ++s) {
            if (s == d) continue;
            int si = s % G.nx, sj = (s / G.nx) % G.ny, sk = s / (G.nx * G.ny);
            size_t q = (size_t)(di - si + G.nx - 1) +
                       (size_t)NX * ((size_t)(dj - sj + G.ny - 1) + (size_t)NY * (dk - sk + G.nz - 1));
            double r = rtab[q];
            if (!(r < prop_thresh)) continue;
            double tau = uniform ? mut[0] * r : tau_pair(&G, mut, di, dj, dk, si, sj, sk, r);
            dcx kf = cexp(CMPLX(-tau, -kim * r)) / (c_m * r * r);
            const double *y = Ytab + q * K;
            dcx p = 0.0;
            for (int b = 0; b < K; ++b) p += y[b] * W[(size_t)b * N + s];
            p *= kf;
            for (int a = 0; a < K; ++a) acc[a] += p * y[a];
        }
        const dcx *Sd = S + (uniform ? 0 : (size_t)d * K * K);
        for (int a = 0; a < K; ++a) {
            dcx s = 0.0;
            for (int b = 0; b < K; ++b) s += Sd[a * K + b] * W[(size_t)b * N + d];
            out[(size_t)a * N + d] = (acc[a] + s) * G.dV;
        }
    }
    free(Ytab); free(rtab); free(S); free(mut);
    geo_free(&G);
    return 0;
}

/* complex first-scatter source and ballistic image (interleaved); returns the
 * number of beam columns (0 = error) */
int orc_source_c(const orc_grid *g, const orc_optics *o, const orc_beam *b, int L, double kim, double *s0_,
                 double *ball_)
{
    geo G;
    geo_init(&G, g);
    int K = ncoef(L), N = G.N;
    dcx *s0 = (dcx *)s0_, *ball = (dcx *)ball_;
    double diam = b->diam + 1.0e-16;
    for (size_t i = 0; i < (size_t)K * N; ++i) s0[i] = 0.0;
    int cols = 0;
    for (int j = 0; j < G.ny; ++j)
        for (int i = 0; i < G.nx; ++i) {
            double rad = sqrt((G.xs[i] - b->pos_x) * (G.xs[i] - b->pos_x) + (G.ys[j] - b->pos_y) * (G.ys[j] - b->pos_y));
            if (rad <= diam / 2.0) ++cols;
        }
    if (cols == 0) { geo_free(&G); return 0; }
    double fact = b->mag / (double)cols;
    for (int j = 0; j < G.ny; ++j)
        for (int i = 0; i < G.nx; ++i) {
            double rad = sqrt((G.xs[i] - b->pos_x) * (G.xs[i] - b->pos_x) + (G.ys[j] - b->pos_y) * (G.ys[j] - b->pos_y));
            if (ball) ball[i + G.nx * j] = 0.0;
            if (!(rad <= diam / 2.0)) continue;
            dcx T = 1.0;
            for (int k = 0; k < G.nz; ++k) {
                int v = i + G.nx * (j + G.ny * k);
                dcx mt = CMPLX(o->mu_a[v] + o->mu_s[v], kim);
                dcx att = cexp(-mt * G.dz);
                dcx base = fact * (o->mu_s[v] / mt) * T * (1.0 - att) / G.dV;
                double gl = 1.0;
                for (int l = 0; l <= L; ++l) {
                    s0[(size_t)(l * l + l) * N + v] = gl * sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) * base;
                    gl *= o->g;
                }
                T *= att;
            }
            if (ball) ball[i + G.nx * j] = fact * T;
        }
    geo_free(&G);
    return cols;
}

/* Sum over the bottom layer of Re of every complex-basis coefficient of a
 * complex field given in the real basis (Kernels_selfsub.c:1194-1206): with
 * c_lm = (r_lm - i r_l,-m)/sqrt2 and c_l,-m = (-1)^m (r_lm + i r_l,-m)/sqrt2,
 * Re c_lm + Re c_l,-m = sqrt2 Re r_lm for even m and sqrt2 Im r_l,-m for odd m. */
double orc_bottom_sum_c(const orc_grid *g, int L, const double *W_)
{
    const dcx *W = (const dcx *)W_;
    int N = g->nx * g->ny * g->nz, base = g->nx * g->ny * (g->nz - 1);
    double s = 0.0;
    for (int l = 0; l <= L; ++l)
        for (int m = -l; m <= l; ++m) {
            double wgt;
            int use_im;
            if (m == 0) { wgt = 1.0; use_im = 0; }
            else if (m > 0 && (m % 2) == 0) { wgt = M_SQRT2; use_im = 0; }
            else if (m < 0 && (-m % 2) == 1) { wgt = M_SQRT2; use_im = 1; }
            else continue;
            const dcx *row = W + (size_t)(l * l + l + m) * N + base;
            double t = 0.0;
            for (int v = 0; v < g->nx * g->ny; ++v) t += use_im ? cimag(row[v]) : creal(row[v]);
            s += wgt * t;
        }
    return s;
}

/* bottom exitance and top reflectance of a complex field (no ballistic) */
static void images_c(const orc_grid *g, const orc_optics *o, int L, int quad, const dcx *w, dcx *ex, dcx *rf)
{
    int K = ncoef(L), N = g->nx * g->ny * g->nz, P = g->nx * g->ny;
    double *part = malloc(sizeof(double) * (size_t)K * N), *img = malloc(sizeof(double) * P);
    for (int h = 0; h < 2; ++h) {
        for (size_t i = 0; i < (size_t)K * N; ++i) part[i] = h ? cimag(w[i]) : creal(w[i]);
        orc_exitance(g, o, L, quad, part, img);
        for (int i = 0; i < P; ++i) ex[i] = h ? ex[i] + CMPLX(0.0, img[i]) : img[i];
        if (rf) {
            orc_reflectance(g, o, L, quad, part, img);
            for (int i = 0; i < P; ++i) rf[i] = h ? rf[i] + CMPLX(0.0, img[i]) : img[i];
        }
    }
    free(part);
    free(img);
}

/* Frequency-domain solve: the Neumann series of the complex operators,
 * exitance = ballistic + bottom exitance of A sum_q (D A)^q s0 (complex, nx*ny
 * interleaved); refl (or NULL) the top reflectance; w_out (or NULL) the final
 * complex field.  Stop rules as orc_solve: lambda on the real parts of the
 * complex-basis bottom sums, delta on |dE| / |E| of the complex total. */
int orc_solve_fd(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p, double kim,
                 double *exitance_, double *refl_, double *w_out_, orc_stats *st)
{
    int L = p->L, K = ncoef(L), N = g->nx * g->ny * g->nz, P = g->nx * g->ny;
    size_t KN = (size_t)K * N;
    dcx *exitance = (dcx *)exitance_, *refl = (dcx *)refl_, *w_out = (dcx *)w_out_;
    dcx *s0 = malloc(sizeof(dcx) * KN), *ball = malloc(sizeof(dcx) * P);
    if (!orc_source_c(g, o, b, L, kim, (double *)s0, (double *)ball)) { free(s0); free(ball); return -1; }
    dcx *W = malloc(sizeof(dcx) * KN), *AW = malloc(sizeof(dcx) * KN), *out = malloc(sizeof(dcx) * KN);
    dcx *img = malloc(sizeof(dcx) * P), *rimg = malloc(sizeof(dcx) * P);
    double *mus = (double *)o->mu_s;
    dcx btot = 0.0;
    for (int i = 0; i < P; ++i) btot += ball[i];
    memcpy(W, s0, sizeof(dcx) * KN);
    memcpy(out, s0, sizeof(dcx) * KN);
    int Lloop = (o->g != 0.0) ? L : 0;
    int n, max_terms = p->max_terms > 0 ? p->max_terms : 1000;
    double lam = 0.0, rel = 0.0, c_m = o->c / o->n;
    /* D: c_m g^l mu_s, real */
#define SCATTER_C(F)                                                                 \
    for (int l = 0; l <= L; ++l) {                                                   \
        double gl = pow(o->g, (double)l);                                            \
        for (int m = -l; m <= l; ++m) {                                              \
            dcx *row = (F) + (size_t)(l * l + l + m) * N;                            \
            for (int v = 0; v < N; ++v) row[v] = c_m * gl * mus[v] * row[v];         \
        }                                                                            \
    }
    if (p->stop_rule == ORC_STOP_LAMBDA) {
        double fi = 1.0;
        n = 1;
        while (!stop_lambda(&fi, n, orc_bottom_sum_c(g, L, (double *)W), orc_bottom_sum_c(g, L, (double *)out),
                            &lam)) {
            if (n >= max_terms) break;
            orc_propagate_c(g, o, Lloop, p->prop_thresh, kim, (double *)W, (double *)AW);
            if (Lloop < L)
                for (size_t i = (size_t)ncoef(Lloop) * N; i < KN; ++i) AW[i] = 0.0;
            SCATTER_C(AW)
            memcpy(W, AW, sizeof(dcx) * KN);
            for (size_t i = 0; i < KN; ++i) out[i] += W[i];
            ++n;
        }
        orc_propagate_c(g, o, L, p->prop_thresh, kim, (double *)out, (double *)AW);
        images_c(g, o, L, p->exit_quad, AW, img, refl ? rimg : NULL);
        for (int i = 0; i < P; ++i) exitance[i] = ball[i] + img[i];
        if (refl) memcpy(refl, rimg, sizeof(dcx) * P);
        if (w_out) memcpy(w_out, AW, sizeof(dcx) * KN);
    } else {
        dcx T = btot;
        for (int i = 0; i < P; ++i) exitance[i] = ball[i];
        if (w_out) for (size_t i = 0; i < KN; ++i) w_out[i] = 0.0;
        if (refl) for (int i = 0; i < P; ++i) refl[i] = 0.0;
        n = 0;
        for (;;) {
            orc_propagate_c(g, o, L, p->prop_thresh, kim, (double *)W, (double *)AW);
            images_c(g, o, L, p->exit_quad, AW, img, refl ? rimg : NULL);
            dcx dT = 0.0;
            for (int i = 0; i < P; ++i) { exitance[i] += img[i]; dT += img[i]; }
            if (w_out) for (size_t i = 0; i < KN; ++i) w_out[i] += AW[i];
            if (refl) for (int i = 0; i < P; ++i) refl[i] += rimg[i];
            T += dT;
            ++n;
            rel = (cabs(T) != 0.0) ? cabs(dT) / cabs(T) : 0.0;
            if (rel < p->tol || n >= max_terms) break;
            SCATTER_C(AW)
            memcpy(W, AW, sizeof(dcx) * KN);
        }
    }
#undef SCATTER_C
    if (st) {
        st->terms = n;
        st->lambda_last = lam;
        st->rel_change_last = rel;
        st->ballistic_total = creal(btot);
    }
    free(s0); free(ball); free(W); free(AW); free(out); free(img); free(rimg);
    return 0;
}

/* Threads of the OpenMP loops (the CPU baseline is timed on one core and on all). */
void orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

int orc_max_threads(void) { return omp_get_max_threads(); }
