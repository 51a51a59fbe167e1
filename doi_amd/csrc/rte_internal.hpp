// rte_internal.hpp -- structures shared by the host library and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rte {

constexpr int kMaxL = 7;
constexpr int kMaxK = (kMaxL + 1) * (kMaxL + 1);
constexpr int kNMom = 16;      // moment-series terms of the self-voxel block
constexpr int kAngRes = 200;   // self-block quadrature grid (CPU package/src/Kernels_selfsub.c:737)
constexpr int kNOmega = kAngRes * kAngRes;
constexpr int kBrick = 4;      // destination bricks of 4x4x4 voxels = one wave64
constexpr int kWave = 64;
constexpr int kMaxNR = 5;      // right-hand sides (beams on one medium) per propagation wave

// Device scratch of rte_propagate_noalloc (rte_noalloc.hip): a static array per GPU.
// scratch_init runs once per device with that device current (rte_ctx_create) and
// looks up the array's device address; scratch_ptr returns it (nullptr before).
constexpr size_t kScratchBytes = size_t(64) << 20;
hipError_t scratch_init(int device);
void *scratch_ptr(int device);

// Geometry in device form.  Voxel v = x + nx*(y + ny*z); destination bricks
// are numbered bx fastest; slot = brick*64 + lane, lane = lx + 4*(ly + 4*lz).
template <typename T>
struct GridDev {
    int nx, ny, nz, N;
    int bx, by, bz, nblk;  // brick counts per axis and total
    int npad;              // nblk * 64
    T dx, dy, dz, dV;
};

// Real SH through monic recurrences.  With P_l^m = (1-z^2)^{m/2} Q_l^m (Condon-
// Shortley phase in Q), Q_l^m = c_lm Pt_l^m where
//   Pt_m^m = 1,  Pt_{m+1}^m = z,  Pt_l^m = z Pt_{l-1}^m - gamma[l][m] Pt_{l-2}^m,
//   gamma_lm = (l+m-1)(l-m-1) / ((2l-1)(2l-3)),
// so that R_l0 = n_l0 Pt_l^0, R_lm = n_lm Pt_l^m Re((x+iy)^m),
// R_l,-m = n_lm Pt_l^m Im((x+iy)^m), n_lm = s_m N_lm c_lm (s_m = sqrt2 for m>0).
// Device fields hold the scaled coefficients W'_a = n_a W_a, so a pair costs
// no multiplication by n; partial sums come out scaled by 1/n_a.
// gamma_lm is a compile-time constant (rounded once from double): in the
// unrolled recurrences it folds into the instructions and takes no registers.
template <typename T>
__host__ __device__ constexpr T sh_gamma(int l, int m)
{
    return T(double(l + m - 1) * double(l - m - 1) / ((2.0 * l - 1.0) * (2.0 * l - 3.0)));
}

// Per-solve device state of one Neumann series.
template <typename T, typename TA>
struct SolveDev {
    T *W;              // [N][K] current term (input of the propagation)
    T *Wp;             // [N][K] the same in the propagation's z-power basis, written by each launch
    T *part;           // [nchunk][N][K] partial pair sums
    const T *ktab;     // [N][npad] e^{-tau}/r^2 per (source, destination slot), cached mode
    const T *mut;      // [N] mu_t
    const T *mus;      // [N] mu_s
    T g;               // anisotropy
    TA *dE;            // [nx*ny] exitance increment of this term
    TA *bsum;          // [nx*ny] per-pixel bottom sums of the next term (stop rule)
    TA *E;             // [nx*ny] accumulated exitance (ballistic included)
    TA *dR;            // [nx*ny] top-face reflectance increment of this term
    TA *R;             // [nx*ny] accumulated reflectance
    TA *red;           // [2] reduced (sum dE, sum bsum) of this term
    TA *aw_store;      // [N][K] if set, k_combine also writes A W (unscaled) here
    TA *phi;           // [N][K] if set, k_combine adds A W (unscaled) here: the final field
    int mode;          // 0 term, 1 init, 2 final, 3 recompute (see k_bottom / k_combine)
    T kim;             // frequency-domain solves: Im mu_t = omega n / c (uniform), 0 otherwise
    // FP64 frequency-domain solves: (cos, sin)(kim r) per displacement
    // (ix, iy, iz) in [1-nx, nx-1] x [1-ny, ny-1] x [1-nz, nz-1], index
    // (ix + nx-1) + (2nx-1) ((iy + ny-1) + (2ny-1) (iz + nz-1)); else nullptr
    const T *phase;
};

template <typename T>
struct PropArgs {
    GridDev<T> g;
    int nchunk, chunk_len;
    T thr2;  // prop_thresh^2; pairs with r^2 >= thr2 are dropped
};

// combine launch: A.op = 1 writes A W to aw_out; otherwise the per-solve mode
// decides (0 term: scatter + exitance + stop sums, 2 final: exitance only)
template <typename T, typename TA>
struct CombArgs {
    GridDev<T> g;
    int nchunk;
    int op;
    const TA *mom;        // [kNMom][K][K] self-block moments
    const T *momT;        // the same in the compute precision (FP32 self block)
    TA self_scale;        // dtheta dphi / dV (S's 1/dV; the result is scaled by dV)
    TA dV;
    TA coefE[kMaxL + 1];  // exitance weight of w_l0 (bottom face)
    TA coefR[kMaxL + 1];  // reflectance weight of w_l0 (top face)
    TA wsum[kMaxK];       // stop-rule weight of coefficient a
    TA wsumI[kMaxK];      // the same for the imaginary half of a complex field
    TA nsc[kMaxK];        // n_a of the scaled coefficients
    TA *aw_out;           // op = 1
};

// n_a of the scaled coefficients, by value
struct Scales {
    double n[kMaxK];
};

struct BeamDev {
    double pos_x, pos_y, half_diam, fact;  // fact = power per beam column
    const double *mua, *mus;               // FP64 maps
    double kim;                            // Im mu_t of a frequency-domain solve (0: continuous wave)
};

// launchers (rte_kernels.hip)
template <typename T>
hipError_t launch_fill(T *p, size_t n, T v, hipStream_t st);
// sv: ngroups * nr solves, group z = sv[z*nr .. z*nr+nr-1] on one medium
template <typename T, typename TA>
hipError_t launch_prop(int L, int path, int nr, const PropArgs<T> &A, const SolveDev<T, TA> *sv, int ngroups,
                       hipStream_t st);
// groups of nr (2 or 3) solves on different media of one grid (mixed precision, cached
// path): they share the pair geometry, each reads its own path table
template <typename T, typename TA>
hipError_t launch_prop_sep(int L, int nr, const PropArgs<T> &A, const SolveDev<T, TA> *sv, int ngroups,
                           hipStream_t st);
// frequency-domain solves: group z = (sv[2z], sv[2z+1]) = (real, imaginary) half of one complex field
template <typename T, typename TA>
hipError_t launch_prop_c(int L, int path, const PropArgs<T> &A, const SolveDev<T, TA> *sv, int ncomplex,
                         hipStream_t st);
template <typename T>
hipError_t launch_ktab(const GridDev<T> &g, T thr2, const T *mut, T *ktab, hipStream_t st);
template <typename T, typename TA>
hipError_t launch_combine(int L, const CombArgs<T, TA> &A, const SolveDev<T, TA> *sv, int nsolve,
                          hipStream_t st);
template <typename T, typename TA>
hipError_t launch_combine_c(int L, const CombArgs<T, TA> &A, const SolveDev<T, TA> *sv, int ncomplex,
                            hipStream_t st);
template <typename T, typename TA>
hipError_t launch_bottom(int P, const SolveDev<T, TA> *sv, int nsolve, hipStream_t st);
template <typename T, typename TA>
hipError_t launch_source(int L, const GridDev<T> &g, double dx, double dy, double dz, double xmin,
                         double ymin, const Scales &sc, const BeamDev *beams, const SolveDev<T, TA> *sv,
                         int nsolve, hipStream_t st);
// complex source of ncomplex frequency-domain solves (halves as in launch_prop_c)
template <typename T, typename TA>
hipError_t launch_source_c(int L, const GridDev<T> &g, double dx, double dy, double dz, double xmin,
                           double ymin, const Scales &sc, const BeamDev *beams, const SolveDev<T, TA> *sv,
                           int ncomplex, hipStream_t st);
hipError_t launch_moments(int K, const double *R, const double *sn, const double *beta, double *M,
                          hipStream_t st);
// in: coefficient-major double [K][N] -> out: voxel-major T [N][K], times sc.n[a]
template <typename T>
hipError_t launch_to_voxel_major(int K, int N, const Scales &sc, const double *in, T *out, hipStream_t st);
// in: voxel-major TA [N][K] -> out: coefficient-major double [K][N], divided by sc.n[a]
template <typename TA>
hipError_t launch_to_coef_major(int K, int N, const Scales &sc, const TA *in, double *out, hipStream_t st);
// the reference's field layout (complex SH basis, voxel-major [N][K], re/im interleaved,
// a real radiance) <-> the real basis: out = voxel-major T [N][K], times s sc.n[a]
template <typename T>
hipError_t launch_cplx_to_real(int L, int N, double s, const Scales &sc, const double *c, T *out, hipStream_t st);
// voxel-major TA [N][K] real basis, times s -> the reference's layout
template <typename TA>
hipError_t launch_real_to_cplx(int L, int N, double s, const TA *w, double *c, hipStream_t st);
hipError_t launch_narrow(size_t n, const double *in, float *out, hipStream_t st);
hipError_t launch_scatter_op(int L, int N, double g, const double *mus, double *W, hipStream_t st);
hipError_t launch_exitance_op(int L, int nx, int ny, int nz, const double *coefE, const double *w,
                              double *img, hipStream_t st);
// G = J^T diag(w) J on MFMA (J: P x n row-major FP32)
hipError_t launch_gram(int P, int n, const float *J, const float *w, float *G, hipStream_t st);
hipError_t launch_tau_list(const GridDev<float> &g, const float *mut, int np, const int *dst,
                           const int *src, double *tau, hipStream_t st);

// adjoint launchers (rte_adjoint.hip)
template <typename T, typename TA>
struct AdjCombArgs {
    GridDev<T> g;
    int nchunk;
    const TA *mom;        // [kNMom][K][K]
    const T *momT;        // the same in the compute precision
    TA self_scale;        // dtheta dphi / dV
    TA dV;
    TA coefE[kMaxL + 1];
    TA nsc[kMaxK];
    T *X;                 // in: n (-1)^l xi_q; out: n (-1)^l xi_{q-1}
    const T *Wq;          // n W_q
    const T *part;
    const TA *aw_prev;    // A W_{q-1} (unscaled) or nullptr when q == 0
    const T *mut, *mus;
    TA g_aniso;
    const TA *cot;        // [nx*ny] cotangent of the image
    TA *gt;               // [N] d/d mu_t accumulator (self block)
    TA *gs;               // [N] d/d mu_s accumulator (scatter)
    TA *lam0;             // [N][L+1] Lambda_0 m = 0 coefficients when q == 0, else nullptr
};
// adjoint series (cotangents) propagated by one launch
constexpr int kMaxNC = 2;
template <typename T, typename TA>
hipError_t launch_adj_prop(int L, const PropArgs<T> &A, int nc, const T *const *X, T *const *Xz, const T *Wq, T *Wqz,
                           T *const *part, T *const *pairbuf, const T *ktab, hipStream_t st);
template <typename T, typename TA>
hipError_t launch_adj_combine(int L, const AdjCombArgs<T, TA> &A, hipStream_t st);
template <typename T, typename TA>
hipError_t launch_adj_init(int L, const GridDev<T> &g, const TA *cot, const Scales &sc, const TA *coefE, T *X,
                           hipStream_t st);
template <typename T, typename TA>
hipError_t launch_ref_source(int L, const GridDev<T> &g, const Scales &sc, const int *part, const int *param, int wrt,
                             TA gan, TA fs, const TA *phi, const TA *wsum, const SolveDev<T, TA> *sv, int nsolve,
                             hipStream_t st);
template <typename TA>
hipError_t launch_accumulate(size_t n, const TA *x, TA *y, hipStream_t st);
template <typename T, typename TA>
hipError_t launch_make_wq(int L, const GridDev<T> &g, const Scales &sc, TA gan, const T *mus, const TA *aw_prev,
                          T *Wq, hipStream_t st);
// back-projection onto parameter blocks for up to kMaxNC backward passes
template <typename T, typename TA>
struct BlockBp {
    const T *pairbuf[kMaxNC];  // [N][npad] pair weights of each pass
    const int *part;           // [N] block of each voxel, -1 for none
    int nparam;
    TA scale;
    TA *out[kMaxNC];           // [nparam] block sums, added to
};
template <typename T, typename TA>
hipError_t launch_backproject_blocks(const GridDev<T> &g, T thr2, const BlockBp<T, TA> &B, int nc, hipStream_t st);
template <typename T, typename TA>
hipError_t launch_backproject(const GridDev<T> &g, int nchunk, int chunk_len, T thr2, const T *pairbuf, TA scale,
                              TA *gt, hipStream_t st);
template <typename TA>
hipError_t launch_src_grad(int L, int nx, int ny, int nz, double dx, double dy, double dz, double xmin, double ymin,
                           double gan, const BeamDev &B, const TA *lam0, const TA *cot, TA *ga, TA *gs, hipStream_t st);

}  // namespace rte
