/*
 * rte.h -- C ABI of the MI355X Neumann-series RTE solver (librte.so).
 *
 * Drop-in boundary for the reference's solver seam: the reference has no
 * plugin API; its solve is the C function
 *     void Neumann(complex_double *src, complex_double *out, int flag)
 * (FisherInfo_IO/src/rte.h:187, FisherInfo_IO/src/neumann.cu:23; GPU package/
 * src/neumann.cu:23) over process globals set up by each main
 * (GPU package/src/rte.cu:11-18), plus the CPU package's operator functions
 * PropAbs2 / PropScat / GenerateSourceBeam / ComputeTransImage / StopIteration
 * (CPU package/src/RTE.h:98-128).  This header replaces them with re-entrant
 * entry points over POD structs.  Every function returns RTE_OK (0) or a
 * negative RTE_E* code and never exits the process (the reference calls
 * exit() on any failure, GPU package/src/defines.h:24-44); rte_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Units: cm and cm^-1.  Exitance is photons per incident photon on the bottom
 * face (GPU package/GPUManual.pdf p.12).
 *
 * Conventions:
 *   voxel v = x + nx*(y + ny*z), voxel centre x_i = xmin + (i + 1/2) dx,
 *     dx = (xmax - xmin)/nx                  (CPU package/src/Inits.c:106-122)
 *   fields (rte_propagate & co.) are coefficient-major W[a*N + v], N = nx*ny*nz,
 *     a = l*l + l + m, in the REAL orthonormal SH basis
 *     R_l0 = Y_l0, R_lm = sqrt2 Re Y_lm, R_l,-m = sqrt2 Im Y_lm (m > 0);
 *     for a real radiance this is a unitary change of the reference's complex
 *     coefficients (CPU package/src/Dist.c:1207-1216) that leaves every m = 0
 *     coefficient, and so the exitance, unchanged
 *   exitance images are img[x + nx*y]  (CPU package/src/Dist.c:24-36 order)
 *
 * Threading: a context is used by one host thread at a time.  Device memory
 * belongs to the context or medium that allocated it; host buffers to the
 * caller.
 */
#ifndef RTE_H
#define RTE_H

#ifdef __cplusplus
extern "C" {
#endif

#define RTE_OK 0
#define RTE_E_ARG -1      /* invalid argument (sizes, NaN/negative optics, ...) */
#define RTE_E_DEVICE -2   /* HIP runtime error */
#define RTE_E_NOMEM -3    /* device or host allocation failed */
#define RTE_E_UNSUPP -4   /* combination not supported (e.g. L > RTE_MAX_L) */
#define RTE_E_BUSY -5     /* server queue full (back-pressure) */
#define RTE_E_NOTFOUND -6 /* unknown request id */

#define RTE_MAX_L 7

/* Geometry (CPU package/src/RTE.h:44-75; CPU package/src/Inits.c:76-134). */
typedef struct {
    int nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
} rte_grid;

/* Optical maps (reference: uniform Phantom, CPU package/src/RTE.h:86-92, or a
 * tissue byte map with <= 3 mu pairs, GPU package/src/rte.h:54-60; here one
 * value per voxel).  c_m = c/n cancels from the exitance and is not used in
 * arithmetic; n and c are accepted for interface fidelity and validated. */
typedef struct {
    const double *mu_a; /* N values, >= 0 */
    const double *mu_s; /* N values, >= 0, mu_a + mu_s > 0 */
    double g;           /* Henyey-Greenstein anisotropy, |g| <= 1 */
    double n;           /* refractive index, > 0 */
    double c;           /* speed of light, > 0 */
} rte_optics;

/* Collimated pencil beam on the top face (CPU package/src/RTE.h:77-86).  The
 * columns with |centre - pos| <= (diam + 1e-16)/2 carry mag/#columns each
 * (CPU package/src/Inits.c:56, Dist.c:1090-1114). */
typedef struct {
    double pos_x, pos_y, diam, mag;
} rte_beam;

#define RTE_STOP_LAMBDA 0 /* reference rule: lambda > 0.9999 (CPU package/src/Kernels_selfsub.c:1184-1233) */
#define RTE_STOP_DELTA 1  /* |change of total exitance| < tol * |total exitance| */

#define RTE_PREC_FP64 0  /* everything in FP64 */
#define RTE_PREC_MIXED 1 /* propagation/scatter FP32; running sums, exitance FP64 */
#define RTE_PREC_FP32 2  /* everything FP32 */

#define RTE_QUAD_REF 0   /* reference 100-point left Riemann sum (CPU package/src/Dist.c:190-218) */
#define RTE_QUAD_EXACT 1 /* exact int_0^1 P_l(u) u du */

#define RTE_PATH_CACHED 0 /* path integrals computed once per medium, reused by every term */
#define RTE_PATH_MARCH 1  /* path integral marched per pair in every term (reference structure) */

typedef struct {
    int L;              /* SH order, 0 <= L <= RTE_MAX_L */
    int max_terms;      /* cap on Neumann terms; <= 0 means 1000 */
    int stop_rule;      /* RTE_STOP_* */
    double tol;         /* for RTE_STOP_DELTA */
    double prop_thresh; /* pairs with r >= prop_thresh are not coupled; <= 0 = no limit */
    int precision;      /* RTE_PREC_* */
    int exit_quad;      /* RTE_QUAD_* */
    int path_mode;      /* RTE_PATH_* */
} rte_params;

typedef struct {
    int terms;              /* the reference's "Neumann series: n" */
    double lambda_last;     /* last lambda of the reference stop rule */
    double rel_change_last; /* last relative change of total exitance */
    double total_exitance;  /* sum of the image incl. ballistic */
    double ballistic_total; /* sum of the ballistic image */
    double gpu_ms;          /* device time of the whole call (events on the solve stream) */
    double latency_ms;      /* host time from the call's start until this solve's result was final */
} rte_stats;

/* Kernel timing of a context, collected with HIP events on the solve stream
 * around every launch while profiling is on (rte_ctx_profile). */
typedef struct {
    long long prop_launches; /* propagation launches (all solves of a term share one) */
    double prop_ms;          /* summed device time of those launches */
    double prop_pairs;       /* voxel pairs (d != s, r < prop_thresh) x solves processed */
    double prop_flops;       /* algorithmic FP flops of those pairs (see DESIGN: F_pair) */
    long long ktab_launches; /* path-table launches */
    double ktab_ms;
    double ktab_pairs;
    double other_ms;         /* combine, bottom, source launches */
    int path_mode, L, precision; /* of the last propagation launch */
    /* the real propagation launches by kernel: [k] groups of k right-hand sides on one
     * medium (k = 1..5), [0] pairs of solves on different media; launches, summed device
     * time and algorithmic flops of each (the complex groups are in the totals only) */
    long long group_launches[6];
    double group_ms[6];
    double group_flops[6];
} rte_prof;

typedef struct rte_ctx rte_ctx;
typedef struct rte_medium rte_medium;

/* ---- context ------------------------------------------------------------ */
int rte_ctx_create(int device, rte_ctx **out);
/* also destroys the media still attached to the context (their handles become invalid) */
int rte_ctx_destroy(rte_ctx *ctx);
/* at most max_rhs (1..5, default 5; at most 3 at L = 6 and 2 from L = 7) beams on one
 * medium share a propagation launch on the mixed-precision cached path; the result bits
 * do not depend on it */
int rte_ctx_set_group_size(rte_ctx *ctx, int max_rhs);
/* on != 0: time every launch from now on (resets the counters); 0: stop */
int rte_ctx_profile(rte_ctx *ctx, int on);
int rte_ctx_get_profile(rte_ctx *ctx, rte_prof *out);
const char *rte_last_error(void);
int rte_version(void);

/* ---- device-resident media ---------------------------------------------
 * Uploads the optical maps once; solves on a medium reuse its device copy and
 * its per-medium tables (path integrals, self blocks) across calls. */
int rte_medium_create(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics,
                      rte_medium **out);
int rte_medium_destroy(rte_medium *m);
/* drops the per-medium tables (the next solve rebuilds them) */
int rte_medium_reset_cache(rte_medium *m);
/* bytes of device memory a solve of this shape needs (admission by HBM footprint) */
long long rte_solve_footprint(const rte_grid *grid, const rte_params *params, int path_mode);

/* ---- solves --------------------------------------------------------------
 * Source -> Neumann series -> bottom-face exitance, replacing the reference
 * main's GenerateSourceBeam / Neumann / ComputeTransImage sequence
 * (CPU package/src/RTE_uniform.c:79-181; GPU package/src/rte.cu:136-203). */
int rte_solve(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics,
              const rte_beam *beam, const rte_params *params, double *exitance /* nx*ny */,
              rte_stats *stats);

/* nb solves on device-resident media of one shape (same grid dims and L);
 * media[i] may repeat (several beams on one medium share its tables and their
 * propagations).  exitance is nb*nx*ny, stats nb entries (may be NULL). */
int rte_solve_media(rte_ctx *ctx, int nb, rte_medium *const *media, const rte_beam *beams,
                    const rte_params *params, double *exitance, rte_stats *stats);

/* as rte_solve_media, with each solve's top-face reflectance image as well
 * (ComputeReflImage, CPU package/src/Dist.c:134-171, written to ReflectFile by
 * CPU package/src/RTE_uniform.c:183-184; no ballistic term); reflectance is
 * nb*nx*ny and may be NULL */
int rte_solve_images(rte_ctx *ctx, int nb, rte_medium *const *media, const rte_beam *beams,
                     const rte_params *params, double *exitance, double *reflectance, rte_stats *stats);

/* one solve that also returns the final radiance field, the reference's
 * out2 = A (s0 + sum W) (CPU package/src/RTE_uniform.c:95-115, OutputDist's DistFile),
 * coefficient-major field[a*N + v] in the real SH basis and the reference's units
 * (A including 1/c_m); the boundary-face images and the photon count of the
 * reference (CPU package/src/Dist.c:349-989) are quadratures of it.
 * reflectance (nx*ny) may be NULL. */
int rte_solve_field(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                    double *field, double *exitance, double *reflectance, rte_stats *stats);

/* The reference's Neumann(src, out, 1) on a given first term, called with out
 * preset to src as its mains do (FisherInfo_IO/src/neumann.cu:23-120,
 * rte_fim.cu:235-237, 275-276; flag 0, out = A out, is rte_propagate): W_0 = src,
 * W_{q+1} = D A W_q until the stop rule, field = A sum_q W_q.  src and field
 * are K*N coefficient-major in the reference's units (rte_source's s0 is such a
 * first term, rte_solve_field's field such a result); exitance and reflectance
 * (nx*ny, may be NULL) are the face images of the field, with no ballistic
 * term.  RTE_PREC_FP64 or RTE_PREC_MIXED (FP64 running sums). */
int rte_neumann(rte_ctx *ctx, rte_medium *m, const rte_params *params, const double *src, double *field,
                double *exitance, double *reflectance, rte_stats *stats);

/* The reference's Neumann(src, out, flag) on its own field layout
 * (FisherInfo_IO/src/neumann.cu:23-120, FisherInfo_IO/src/rte.h:187; GPU package/src/
 * neumann.cu:23-181): src and out are K*N complex_double (2*K*N doubles, re then im) of
 * the reference's complex SH basis (CPU package/src/SphHarm.c:113-160: orthonormal,
 * Condon-Shortley phase, Y_l,-m = (-1)^m conj Y_lm), voxel-major W[v*K + cnt],
 * cnt = l*l + l + m (GPU package/src/defines.h:59), in the reference's units.  The
 * field must be that of a real radiance (c_l,-m = (-1)^m conj c_lm, Im c_l0 = 0, to
 * 1e-12 of its largest value), as every field of the reference's continuous-wave
 * solves is; others are rejected (RTE_E_ARG).  Both layouts are converted on the device.
 * flag 1: the whole series, out <- A sum_q (D A)^q src until the stop rule, with out
 *   preset to src on entry as every reference call site does (GPU package/src/
 *   rte.cu:140-174, FisherInfo_IO/src/rte_fim.cu:235-237, 275-276,
 *   compute_grad_rte.cu:244-249, 297-299; anything else is RTE_E_ARG); src is left
 *   unchanged (the reference leaves the last term in it);
 * flag 0: the final absorption alone, out <- A out (src unused, may be NULL).
 * RTE_PREC_FP64 or RTE_PREC_MIXED. */
int rte_neumann_ref(rte_ctx *ctx, rte_medium *m, const rte_params *params, const double *src, double *out,
                    int flag, rte_stats *stats);

/* Frequency-domain solves of intensity-modulated beams: omega[i] (rad/s) is
 * the reference's [Source] modulation, which turns mu_a into
 * mu_a + i omega / (c/n) (CPU package/src/RTE_uniform.c:120-177, Inits.c:60).
 * Source, ballistic term, propagation A (pair factor e^{-tau - i omega n r / c},
 * self block at the complex mu_t) and the Neumann series are complex; the
 * scatter and the image functionals are real.  Outputs per solve: exitance
 * and reflectance (may be NULL) as 2 planes of nx*ny (real, imaginary), i.e.
 * nb*2*nx*ny doubles; field (may be NULL) the final complex field of solve 0
 * as 2 planes of K*N (real, imaginary), coefficient-major, reference units.
 * Stop rules as rte_solve_media: lambda on the real parts of the complex-basis
 * bottom sums (StopIteration, CPU package/src/Kernels_selfsub.c:1184-1233,
 * on a complex field), delta on |dE| / |E| of the complex total exitance.
 * stats report the real parts.  This is the series the reference's modulated
 * branch sets out to compute; that branch itself accumulates into the
 * continuous-wave field, normalises the source twice and propagates the
 * source once (RTE_uniform.c:128-160), which rte_solve_fd does not repeat. */
int rte_solve_fd(rte_ctx *ctx, int nb, rte_medium *const *media, const rte_beam *beams, const double *omega,
                 const rte_params *params, double *exitance, double *reflectance, double *field,
                 rte_stats *stats);

/* ---- derivatives ----------------------------------------------------------
 * Reverse mode through the Neumann series as it ran (same term count): one
 * adjoint series with the transpose of A (reciprocity) and one back-projection
 * of the pair weights along the rays give the derivative with respect to
 * every voxel's mu_a and mu_s, including the source and ballistic terms.  The
 * reference instead runs one extra forward solve per parameter with the source
 * -c w restricted to the parameter's voxels (FisherInfo_IO/src/compute_grad_rte.cu:
 * 273-385, GPU package/src/inverse_rte.cu:290-433).  Needs RTE_PATH_CACHED and
 * RTE_PREC_FP64 or RTE_PREC_MIXED. */
typedef struct {
    int checkpoint_stride; /* keep A W_q of every stride-th term of the forward pass and
                              recompute the others in the backward pass (>= 1; 1 keeps all) */
    int jacobian;          /* rte_fisher: 0 the exact derivative (one adjoint series per
                              detector), 1 the reference's forward-mode derivative
                              (rte_ref_jacobian: one series per parameter) */
} rte_grad_opts;

/* dmu_a[v] = sum_p cotangent[p] dE_p/dmu_a(v), same for dmu_s (N each, may be NULL);
 * exitance (nx*ny, may be NULL) receives E */
int rte_exitance_vjp(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                     const rte_grad_opts *opts, const double *cotangent, double *exitance,
                     double *dmu_a, double *dmu_s, rte_stats *stats);
/* loss = sum_p (E_p - target_p)^2 and its gradient (FisherInfo_IO/src/compute_grad_rte.cu:315) */
int rte_gradient(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                 const rte_grad_opts *opts, const double *target, double *loss, double *exitance,
                 double *dmu_a, double *dmu_s, rte_stats *stats);

/* Fisher information of Poisson detectors for block parameters
 * (FisherInfo_IO/src/rte_fim.cu:316-331, FisherInfo_IO/src/inverse_rte.cu:301-316):
 * theta_a = mu_a (wrt = 0) or mu_s (wrt = 1) on the voxels with partition[v] == a
 * (-1: not a parameter); J[p][a] = dE_p/dtheta_a (nx*ny x nparam, may be NULL),
 * F = J^T diag(1/E) J (nparam x nparam, detectors with E = 0 skipped) formed on FP32
 * MFMA, chol = its lower Cholesky factor in FP64 (may be NULL; RTE_E_ARG if F is not
 * positive definite). */
int rte_fisher(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
               const rte_grad_opts *opts, const int *partition, int nparam, int wrt, double *F,
               double *chol, double *J, double *exitance, rte_stats *stats);

/* F^-1 (nparam x nparam, FP64) from rte_fisher's lower Cholesky factor chol (F = C C^T):
 * F^-1 = C^-T C^-1 by forward and back substitution; its diagonal is the Cramer-Rao
 * bound on the parameters' variances.  The reference forms F and stops there
 * (FisherInfo_IO/src/rte_fim.cu:316-331, inverse_rte.cu:301-316).  Host arithmetic only;
 * RTE_E_ARG for a non-positive pivot. */
int rte_fisher_inverse(int nparam, const double *chol, double *Finv);

/* The reference's own derivative of the exitance with respect to block parameters
 * (FisherInfo_IO/src/compute_grad_rte.cu:273-301 mu_a, 341-371 mu_s; GPU package/src/
 * inverse_rte.cu:290-433): phi = A sum_q W_q of the forward solve in the reference's units;
 * per parameter a, one Neumann series from -C phi (wrt 0) or C (g^l - 1) phi (wrt 1) on the
 * voxels with partition[v] == a, C = optics.c, the reference's literal -1.0*C with its C the
 * vacuum speed (FisherInfo_IO/src/rte.h:15, compute_grad_rte.cu:292,350-353), and
 * J[p][a] = its exitance without the ballistic term (nx*ny x nparam).  For n != 1 that is n
 * times the series a source of -c_m phi would give (c_m = c/n, the speed in the medium).
 * The reference approximates dA/dmu by -A chi A and ignores the ballistic and source terms,
 * so this is not the exact derivative rte_exitance_vjp gives; it reproduces the reference's
 * gradients (the mu_s region applied as intended: the reference indexes it with a stale
 * voxel index, compute_grad_rte.cu:357-366).  All parameters run as one batch. */
int rte_ref_jacobian(rte_ctx *ctx, rte_medium *m, const rte_beam *beam, const rte_params *params,
                     const int *partition, int nparam, int wrt, double *J, double *exitance,
                     rte_stats *stats);

/* ---- serving --------------------------------------------------------------
 * Continuous batching (no counterpart in the reference, which runs one solve
 * per process, GPU package/src/rte.cu:89-227).  A worker thread owns the
 * context while the server exists (make no other call on it until
 * rte_server_destroy): requests are granted a solve slot of the bucket of
 * their shape (grid, L, precision, path, stop rule) first come first served
 * while the slots' device memory fits hbm_budget, and advance one Neumann
 * term per step of their bucket with every other resident solve of it;
 * buckets step independently and overlap on the device.  Finished solves
 * leave their slot to the next request of the shape (idle slots are freed
 * when another shape needs the memory).  A solve's result is bit-identical to
 * rte_solve_media's for it, whatever it was batched with. */
typedef struct rte_server rte_server;
typedef struct {
    long long hbm_budget;     /* bytes for resident solves; <= 0: 80 % of free device memory */
    int max_slots_per_bucket; /* resident solves per shape bucket; <= 0: 64 */
    int max_queue;            /* queued requests before rte_submit returns RTE_E_BUSY; <= 0: 100000 */
} rte_server_opts;
typedef struct {
    long long queued, running, done, failed, cancelled, steps;
    long long hbm_used; /* device memory of the solve slots, busy and pooled idle ones */
    long long hbm_budget;
    int buckets;
    int worker_status; /* last device error of the worker (RTE_OK if none) */
} rte_server_stats;

int rte_server_create(rte_ctx *ctx, const rte_server_opts *opts, rte_server **out);
int rte_server_destroy(rte_server *s);
/* validates and queues a solve (the maps are copied); RTE_E_BUSY if the queue is
 * full, RTE_E_NOMEM if the solve can never fit the budget */
int rte_submit(rte_server *s, const rte_grid *grid, const rte_optics *optics, const rte_beam *beam,
               const rte_params *params, long long *id);
/* *status: 0 queued, 1 running, 2 done (exitance/stats written, id released),
 * 3 failed (the solve's error code is returned), 4 cancelled; RTE_E_NOTFOUND for an
 * unknown or released id.  stats->latency_ms is submit-to-result, gpu_ms admit-to-result. */
int rte_poll(rte_server *s, long long id, double *exitance, rte_stats *stats, int *status);
/* as rte_poll, blocking until the request leaves the queue/device or timeout_ms passes
 * (< 0: no limit; RTE_E_BUSY on timeout) */
int rte_wait(rte_server *s, long long id, double timeout_ms, double *exitance, rte_stats *stats,
             int *status);
int rte_cancel(rte_server *s, long long id);
int rte_server_info(rte_server *s, rte_server_stats *out);

/* ---- operators (for operator-level parity) ------------------------------ */
/* A: propagation incl. self-voxel block and dV (CPU package/src/Kernels_selfsub.c:450-538,
 * GPU package/src/kernels.cu:148-175,286-381).  W, out: K*N doubles. */
int rte_propagate(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int L,
                  double prop_thresh, int precision, int path_mode, const double *W,
                  double *out);
/* rte_propagate without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and result as
 * rte_propagate with path_mode RTE_PATH_MARCH (the only mode accepted: a cached path
 * table is a per-medium allocation of N^2 entries); every device buffer, the
 * self-block moments and maps included, lives in one static device array of 64 MiB
 * per GPU, held until the process ends.  Calls whose buffers would not fit return
 * RTE_E_UNSUPP before any device work (e.g. 12^3 at L = 5 in FP64 fits, 16^3 at L = 5
 * in FP64 does not).  Calls on one GPU share that array: calls that overlap on one
 * GPU (two contexts, two threads) race; one context's calls are ordered. */
int rte_propagate_noalloc(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int L,
                          double prop_thresh, int precision, int path_mode, const double *W,
                          double *out);
/* D: Henyey-Greenstein scatter g^l mu_s (CPU package/src/Kernels_selfsub.c:30-51), in place */
int rte_scatter(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int L, double *W);
/* s0 (K*N) and the ballistic image (nx*ny) (CPU package/src/Dist.c:1077-1128, 203-209;
 * GPU package/src/dist.cu:66-101, 129-141); returns #beam columns in *ncols */
int rte_source(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, const rte_beam *beam,
               int L, double *s0, double *ballistic, int *ncols);
/* bottom-face exitance of a field w (K*N), without the ballistic term; c_m =
 * c/n is the speed of light in the medium (CPU package/src/Dist.c:178-229) */
int rte_exitance(rte_ctx *ctx, const rte_grid *grid, int L, int exit_quad, double c_m,
                 const double *w, double *img);
/* path integral tau(d, s) for a list of voxel pairs (GPU package/src/kernels.cu:49-145) */
int rte_path_integrals(rte_ctx *ctx, const rte_grid *grid, const rte_optics *optics, int npairs,
                       const int *dest, const int *src, double *tau);

#ifdef __cplusplus
}
#endif
#endif
