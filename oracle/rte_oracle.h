/*
 * rte_oracle.h -- FP64 CPU restatement of the reference Neumann-series RTE
 * solve.  TEST INFRASTRUCTURE ONLY: this library is the checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * Conventions follow the reference CPU package (CPU package/src):
 *   grid      voxel centre x_i = xmin + (i + 1/2) dx, dx = (xmax - xmin)/nx
 *             (CPU package/src/Inits.c:106-122)
 *   voxel id  v = x + nx*(y + ny*z)                   (CPU package/src/Dist.c:1207)
 *   fields    coefficient-major W[a*N + v], a = l*l + l + m, N = nx*ny*nz
 *             (CPU package/src/Dist.c:1207-1216), REAL spherical harmonics:
 *             R_l0 = Y_l0, R_lm = sqrt2 Re Y_lm (m>0), R_l,-m = sqrt2 Im Y_lm.
 *             Exitance only reads m = 0, which is identical in both bases.
 */
#ifndef RTE_ORACLE_H
#define RTE_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
} orc_grid;

typedef struct {
    const double *mu_a;   /* N values, cm^-1 */
    const double *mu_s;   /* N values, cm^-1 */
    double g;             /* Henyey-Greenstein anisotropy */
    double n;             /* refractive index */
    double c;             /* speed of light, cm/s; c_m = c/n cancels in exitance */
} orc_optics;

typedef struct {
    double pos_x, pos_y;  /* beam centre on the top face, cm */
    double diam;          /* beam diameter as written in the .par file, cm */
    double mag;           /* incident power */
} orc_beam;

enum { ORC_STOP_LAMBDA = 0, ORC_STOP_DELTA = 1 };
enum { ORC_QUAD_REF = 0, ORC_QUAD_EXACT = 1 };

typedef struct {
    int L;              /* SH order */
    int max_terms;      /* cap on Neumann terms (W_0 .. W_{n-1}) */
    int stop_rule;      /* ORC_STOP_LAMBDA (reference) or ORC_STOP_DELTA */
    double tol;         /* relative change in total exitance for ORC_STOP_DELTA */
    double prop_thresh; /* pairs with r >= prop_thresh are not coupled */
    int exit_quad;      /* ORC_QUAD_REF (100-point left Riemann) or ORC_QUAD_EXACT */
} orc_params;

typedef struct {
    int terms;          /* the reference's printed "Neumann series: n" */
    double lambda_last; /* last lambda of the stop rule */
    double rel_change_last;
    double ballistic_total;
} orc_stats;

/* spherical harmonics, all a = l*l+l+m <= (L+1)^2-1, at one direction */
void orc_ylm_complex(int L, double theta, double phi, double *re, double *im);
void orc_ylm_real(int L, double theta, double phi, double *out);
double orc_assoc_legendre(double x, int l, int m);

/* self-voxel block S (K x K, row a = output) in the real basis, including the
 * reference's 1/(c_m mu_t dV) dtheta dphi factor.  method 0: direct 200x200
 * quadrature, method 1: moment series. */
int orc_self_block(int L, double dx, double dy, double dz, double mu_t, double c_m,
                   int method, double *S);
/* raw angular moments M_k[a][b] = sum_w R_a R_b beta^k sin(theta), k = 1..nmom */
int orc_self_moments(int L, double dx, double dy, double dz, int nmom, double *M);
/* beta(omega) for the 200x200 grid, in the reference's float arithmetic */
void orc_self_beta(double dx, double dy, double dz, double *beta /*40000*/);

/* Siddon radiological path between voxel centres of s (source) and d (dest) */
double orc_tau(const orc_grid *g, const double *mu_t, int d, int s);

/* first-scatter source s0 (K*N) and ballistic exitance image (nx*ny); returns
 * the number of beam columns (0 = error) */
int orc_source(const orc_grid *g, const orc_optics *o, const orc_beam *b, int L,
               double *s0, double *ballistic);

/* A (propagation incl. self term and dV) on a K*N field */
int orc_propagate(const orc_grid *g, const orc_optics *o, int L, double prop_thresh,
                  const double *W, double *out);
/* A restricted to destination voxels d_begin <= d < d_end (other rows of out untouched) */
int orc_propagate_range(const orc_grid *g, const orc_optics *o, int L, double prop_thresh,
                        const double *W, double *out, int d_begin, int d_end);
/* D (Henyey-Greenstein scatter) in place on a K*N field */
void orc_scatter(const orc_grid *g, const orc_optics *o, int L, double *W);
/* exitance of a field on the bottom face, without the ballistic term */
void orc_exitance(const orc_grid *g, const orc_optics *o, int L, int quad,
                  const double *w, double *img);
void orc_quad_weights(int L, int quad, double *Q);
/* top-face reflectance of a field (nx*ny) and its quadrature int_{-1}^0 P_l(u) u du */
void orc_reflectance(const orc_grid *g, const orc_optics *o, int L, int quad,
                     const double *w, double *img);
void orc_quad_weights_refl(int L, int quad, double *R);

/* full solve: source -> Neumann -> bottom exitance (nx*ny) incl. ballistic */
int orc_solve(const orc_grid *g, const orc_optics *o, const orc_beam *b,
              const orc_params *p, double *exitance, orc_stats *st);

/* full solve with the top-face reflectance image as well (nx*ny each) */
int orc_solve_images(const orc_grid *g, const orc_optics *o, const orc_beam *b,
                     const orc_params *p, double *exitance, double *reflectance, orc_stats *st);

/* full solve keeping the final radiance w = A out (K*N) for gradient checks */
int orc_solve_field(const orc_grid *g, const orc_optics *o, const orc_beam *b,
                    const orc_params *p, double *exitance, double *w_out,
                    orc_stats *st);

/* the reference's forward-mode derivative of the exitance with respect to
 * mu_a (which = 0) or mu_s (which = 1) of the voxels with region[v] != 0:
 * one Neumann series from the forward field restricted to the region, its
 * exitance without the ballistic term (nx*ny) */
int orc_ref_derivative(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p,
                       const int *region, int which, double *dimg, orc_stats *st);

/* the reference's Neumann(src, out, 1) on a given first term src (K*N):
 * w_out = A sum_q (DA)^q src; exitance (nx*ny) and refl (may be NULL) are its
 * face images, with no ballistic term */
int orc_neumann(const orc_grid *g, const orc_optics *o, const orc_params *p, const double *src,
                double *w_out, double *exitance, double *refl, orc_stats *st);

/* Frequency domain (complex mu_t = mu_a + mu_s + i kim, kim = omega n / c).
 * Complex arrays are interleaved (re, im) doubles, fields coefficient-major. */
int orc_self_block_c(int L, double dx, double dy, double dz, double mu_re, double mu_im, double c_m,
                     int method, double *S);
int orc_propagate_c(const orc_grid *g, const orc_optics *o, int L, double prop_thresh, double kim,
                    const double *W, double *out);
int orc_propagate_c_range(const orc_grid *g, const orc_optics *o, int L, double prop_thresh, double kim,
                          const double *W, double *out, int d_begin, int d_end);
int orc_source_c(const orc_grid *g, const orc_optics *o, const orc_beam *b, int L, double kim,
                 double *s0, double *ballistic);
double orc_bottom_sum_c(const orc_grid *g, int L, const double *W);
int orc_solve_fd(const orc_grid *g, const orc_optics *o, const orc_beam *b, const orc_params *p, double kim,
                 double *exitance, double *reflectance, double *w_out, orc_stats *st);

/* bottom-layer sums of the stop rule (complex-basis real parts) */
double orc_bottom_sum(const orc_grid *g, int L, const double *W);

/* orc_propagate_range that also returns the pair loop's wall time (tables excluded) */
int orc_propagate_range_timed(const orc_grid *g, const orc_optics *o, int L, double prop_thresh,
                              const double *W, double *out, int d_begin, int d_end, double *loop_s);
/* threads of the OpenMP loops */
void orc_set_threads(int n);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
