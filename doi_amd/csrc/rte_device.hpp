// rte_device.hpp -- device helpers shared by the forward and adjoint kernels.
#pragma once

#include "rte_internal.hpp"

#include <utility>

namespace rte {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T dev_rsqrt(T x);
template <>
__device__ __forceinline__ float dev_rsqrt<float>(float x) { return __frsqrt_rn(x); }
template <>
__device__ __forceinline__ double dev_rsqrt<double>(double x) { return 1.0 / sqrt(x); }

// a * b + c with one rounding.  The propagation's pair arithmetic is written
// with it wherever a product meets a sum, and compiled without contraction, so
// every kernel that evaluates a pair (any group size, alone or across media)
// rounds the same operations the same way: a solve's bits do not depend on the
// launch it is in, whatever the compiler schedules around them.
template <typename T>
__host__ __device__ __forceinline__ T fmad(T a, T b, T c);
template <>
__host__ __device__ __forceinline__ float fmad<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__host__ __device__ __forceinline__ double fmad<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ T dev_exp(T x);
template <>
__device__ __forceinline__ float dev_exp<float>(float x) { return expf(x); }
template <>
__device__ __forceinline__ double dev_exp<double>(double x) { return exp(x); }

// sin and cos of the FP32 pair phase k r: the hardware sine/cosine (absolute
// error ~1e-6 at arguments of order 10, below the FP32 path's tolerance).  The
// FP64 kernel reads the phase from a per-displacement table instead: a library
// sincos in its loop branches per lane to a large-argument reduction, and in that
// heavily spilled kernel the register allocator placed copies of values live in
// every lane into the branch's flow block, under its partial EXEC mask.
template <typename T>
__device__ __forceinline__ void dev_sincos(T x, T *s, T *c);
template <>
__device__ __forceinline__ void dev_sincos<float>(float x, float *s, float *c) { __sincosf(x, s, c); }

template <typename T>
__device__ __forceinline__ T big_value();
template <>
__device__ __forceinline__ float big_value<float>() { return 1.0e30f; }
template <>
__device__ __forceinline__ double big_value<double>() { return 1.0e300; }


// ---------------------------------------------------------------------------
// Siddon path integral between the centres of source s = (si,sj,sk) and
// destination d = s + (ix,iy,iz):  tau = r * sum_v mu_t(v) dalpha(v).
// Crossings along an axis with n steps sit at alpha = (j + 1/2)/n; ties go to
// z, then y, then x; the destination contributes the half step 0.5/n of the
// axis it was entered through (GPU package/src/kernels.cu:49-145).
// ---------------------------------------------------------------------------
// The next voxel's mu_t is loaded one step ahead of its use (the voxel
// sequence does not depend on the loaded values), so the load latency
// overlaps the crossing arithmetic.  P: a global or LDS pointer.
template <typename T, typename P>
__device__ __forceinline__ T march_tau(P mut, int v, int ix, int iy, int iz, int nx, int nxy, T r)
{
    const int n0 = abs(ix), n1 = abs(iy), n2 = abs(iz);
    const T BIG = big_value<T>();
    const T inc0 = n0 ? T(1) / T(n0) : BIG;
    const T inc1 = n1 ? T(1) / T(n1) : BIG;
    const T inc2 = n2 ? T(1) / T(n2) : BIG;
    T a0 = T(0.5) * inc0, a1 = T(0.5) * inc1, a2 = T(0.5) * inc2;
    const int st0 = ix > 0 ? 1 : -1, st1 = iy > 0 ? nx : -nx, st2 = iz > 0 ? nxy : -nxy;
    const int nsteps = n0 + n1 + n2;
    T prev = T(0), acc = T(0), last = T(0);
    T mu = mut[v];
    for (int t = 0; t < nsteps; ++t) {
        const bool pz = (a2 <= a0) && (a2 <= a1);
        const bool py = !pz && (a1 <= a0);
        const T a = pz ? a2 : (py ? a1 : a0);
        v += pz ? st2 : (py ? st1 : st0);
        const T mu_next = mut[v];
        acc += mu * (a - prev);
        prev = a;
        last = pz ? inc2 : (py ? inc1 : inc0);
        a2 = pz ? a2 + inc2 : a2;
        a1 = py ? a1 + inc1 : a1;
        a0 = (!pz && !py) ? a0 + inc0 : a0;
        mu = mu_next;
    }
    acc += mu * T(0.5) * last;
    return acc * r;
}

// ---------------------------------------------------------------------------
// One (destination, source) pair for the propagation kernels (forward, every
// group, real or complex, and adjoint), in the basis of the scaled real
// harmonics with their Legendre factor replaced by a power of z:
//   Z_l,m = C_m z^(l-m),  Z_l,-m = S_m z^(l-m)  (C_m + i S_m = (x + i y)^m),
// stored where Rt_l,+-m = C_m Pt_l^m, S_m Pt_l^m sit (a = l^2 + l +- m).  For
// each m the powers z^0 .. z^(L-m) span the same polynomials as the monic
// Legendre factors Pt_m^m .. Pt_L^m, so
//   Rt_l,m = sum_{l' <= l} P^m_l,l'-m Z_l',m,  P^m_l,k the coefficient of z^k
// in Pt_l^m (ZPowBasis below).  A pair's 35 values cost one product each
// beyond the powers (x + i y)^m (43 VALU instructions at L = 5 against 55 with
// the Legendre recurrence); the propagation reads W in this basis,
// W'_l',m = sum_{l >= l'} P^m_l,l'-m W_l,m (k_to_zpow, before each
// propagation launch), sums t Z(u) per pair, and k_combine turns the summed
// partials back, acc_l,m = sum_{l' <= l} P^m_l,l'-m acc'_l',m.  The map is
// well conditioned (condition number 3.4 at L = 5, 6.8 at L = 7), so the FP32
// projection rounds about as the Legendre form does.  (A basis of plain
// monomials x^a y^b z^c would cost 32 instructions, but its map's condition
// number, 18 at L = 5 and 88 at L = 7, shows in FP32 results.)  Every
// right-hand side costs one K-term dot and one K-term axpy of FMAs.
// ---------------------------------------------------------------------------

// M_aj: Rt_a = sum_j M_aj Z_j; nonzero only between entries of the same m and
// sign, l' <= l, l' = l mod 2.
template <int L>
struct ZPowBasis {
    static constexpr int K = (L + 1) * (L + 1);
    double m[K][K];
    constexpr ZPowBasis() : m{}
    {
        for (int mm = 0; mm <= L; ++mm) {
            double P[L + 2][L + 2] = {};  // P[l][k]: coefficient of z^k in Pt_l^mm (one spare row, column)
            P[mm][0] = 1.0;
            if (mm + 1 <= L) P[mm + 1][1] = 1.0;
            for (int l = mm + 2; l <= L; ++l)
                for (int k = 0; k <= l - mm; ++k)
                    P[l][k] = (k ? P[l - 1][k - 1] : 0.0) - sh_gamma<double>(l, mm) * P[l - 2][k];
            for (int sn = 0; sn < (mm ? 2 : 1); ++sn)
                for (int l = mm; l <= L; ++l)
                    for (int lp = mm; lp <= l; ++lp)
                        this->m[l * l + l + (sn ? -mm : mm)][lp * lp + lp + (sn ? -mm : mm)] = P[l][lp - mm];
        }
    }
};

template <int L>
inline constexpr ZPowBasis<L> kZPow{};

// M's nonzero entries as compile-time lists, in the two orders the conversions
// accumulate in: by z-power index j, then harmonic index a (for M^T x), and by a,
// then j (for M x).  A fold over the list expands every term with its constant,
// whatever the optimiser's unrolling limits.
template <int L>
struct ZPowTerms {
    static constexpr int K = (L + 1) * (L + 1);
    static constexpr int count()
    {
        int c = 0;
        for (int a = 0; a < K; ++a)
            for (int j = 0; j < K; ++j) c += kZPow<L>.m[a][j] != 0.0;
        return c;
    }
    static constexpr int NZ = count();
    int a[NZ], j[NZ];
    double v[NZ];
    constexpr ZPowTerms(bool by_j) : a{}, j{}, v{}
    {
        int c = 0;
        for (int p = 0; p < K; ++p)
            for (int q = 0; q < K; ++q) {
                const int aa = by_j ? q : p, jj = by_j ? p : q;
                if (kZPow<L>.m[aa][jj] != 0.0) {
                    a[c] = aa;
                    j[c] = jj;
                    v[c] = kZPow<L>.m[aa][jj];
                    ++c;
                }
            }
    }
};

template <int L>
inline constexpr ZPowTerms<L> kZPowByJ{true};
template <int L>
inline constexpr ZPowTerms<L> kZPowByA{false};

template <typename TA, int L, int... I>
__device__ __forceinline__ void zpow_from_terms(const TA *x, TA *y, std::integer_sequence<int, I...>)
{
    ((y[kZPowByJ<L>.j[I]] = fmad(TA(kZPowByJ<L>.v[I]), x[kZPowByJ<L>.a[I]], y[kZPowByJ<L>.j[I]])), ...);
}

template <typename TA, int L, int... I>
__device__ __forceinline__ void zpow_to_terms(const TA *x, TA *y, std::integer_sequence<int, I...>)
{
    ((y[kZPowByA<L>.a[I]] = fmad(TA(kZPowByA<L>.v[I]), x[kZPowByA<L>.j[I]], y[kZPowByA<L>.a[I]])), ...);
}

// harmonic-basis coefficients into the z-power basis, y_j = sum_a M_aj x_a
// (a ascending for each j)
template <typename TA, int L>
__device__ __forceinline__ void zpow_from_harmonic(const TA (&x)[(L + 1) * (L + 1)], TA (&y)[(L + 1) * (L + 1)])
{
#pragma unroll
    for (int q = 0; q < (L + 1) * (L + 1); ++q) y[q] = TA(0);
    zpow_from_terms<TA, L>(x, y, std::make_integer_sequence<int, ZPowTerms<L>::NZ>{});
}

// sums of z-power terms back to the harmonic basis: x_a <- sum_j M_aj x_j
// (j ascending for each a)
template <typename TA, int L>
__device__ __forceinline__ void zpow_to_harmonic(TA (&x)[(L + 1) * (L + 1)])
{
    constexpr int K = (L + 1) * (L + 1);
    TA y[K];
#pragma unroll
    for (int q = 0; q < K; ++q) y[q] = TA(0);
    zpow_to_terms<TA, L>(x, y, std::make_integer_sequence<int, ZPowTerms<L>::NZ>{});
#pragma unroll
    for (int q = 0; q < K; ++q) x[q] = y[q];
}

template <typename T, int L>
struct PairZ {
    T R[(L + 1) * (L + 1)];  // R[0] = 1 is not used
};

// the pair's Z values: the powers of z, then (x + i y)^m by complex products
// and their products with the powers
template <typename T, int L>
__host__ __device__ __forceinline__ void pair_zpow(PairZ<T, L> &G, T x, T y, T z)
{
#pragma clang fp contract(off)
    G.R[0] = T(1);
    if constexpr (L >= 1) {
        T zp[L + 1];
        zp[0] = T(1);
        zp[1] = z;
#pragma unroll
        for (int k = 2; k <= L; ++k) zp[k] = zp[k - 1] * z;
#pragma unroll
        for (int l = 1; l <= L; ++l) G.R[l * l + l] = zp[l];
        T c = x, s = y;
#pragma unroll
        for (int m = 1; m <= L; ++m) {
            if (m > 1) {
                const T cn = fmad(c, x, -(s * y)), sn = fmad(s, x, c * y);
                c = cn;
                s = sn;
            }
#pragma unroll
            for (int l = m; l <= L; ++l) {
                G.R[l * l + l + m] = l == m ? c : c * zp[l - m];
                G.R[l * l + l - m] = l == m ? s : s * zp[l - m];
            }
        }
    }
}

// p = sum_j Z_j w'_j, in two interleaved chains (odd and even j)
template <typename T, int L>
__device__ __forceinline__ T pair_dot_z(const T (&w)[(L + 1) * (L + 1)], const PairZ<T, L> &G)
{
#pragma clang fp contract(off)
    constexpr int K = (L + 1) * (L + 1);
    if constexpr (K == 1) {
        return w[0];
    } else if constexpr (K < 3) {
        return fmad(G.R[1], w[1], w[0]);
    } else {
        T p = fmad(G.R[1], w[1], w[0]), q = G.R[2] * w[2];
#pragma unroll
        for (int a = 3; a < K; a += 2) {
            p = fmad(G.R[a], w[a], p);
            if (a + 1 < K) q = fmad(G.R[a + 1], w[a + 1], q);
        }
        return p + q;
    }
}

// the chains of pair_dot_z over coefficients [A0, A0 + N) of a row, read from
// h[a - A0]: a projection taken in parts, each chain in the same order, so that
// parts covering [0, K) in turn leave p + q equal to pair_dot_z's value
template <typename T, int L, int A0, int N>
__device__ __forceinline__ void pair_dot_z_part(T &p, T &q, const T *h, const PairZ<T, L> &G)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int a = A0; a < A0 + N; ++a) {
        if (a == 0) continue;
        if (a == 1)
            p = fmad(G.R[1], h[1 - A0], h[0 - A0]);
        else if (a == 2)
            q = G.R[2] * h[2 - A0];
        else if (a & 1)
            p = fmad(G.R[a], h[a - A0], p);
        else
            q = fmad(G.R[a], h[a - A0], q);
    }
}

// acc' += t Z
template <typename T, int L>
__device__ __forceinline__ void pair_axpy_z(T (&acc)[(L + 1) * (L + 1)], T t, const PairZ<T, L> &G)
{
#pragma clang fp contract(off)
    acc[0] += t;
#pragma unroll
    for (int a = 1; a < (L + 1) * (L + 1); ++a) acc[a] = fmad(t, G.R[a], acc[a]);
}

// Walks the same Siddon crossings as march_tau and calls f(v, length) for
// every voxel v crossed (length = r * dalpha, the destination's half step
// included): the transpose of the path integral, used to back-project.
template <typename T, typename F>
__device__ __forceinline__ void march_visit(int v, int ix, int iy, int iz, int nx, int nxy, T r, F &&f)
{
    const int n0 = abs(ix), n1 = abs(iy), n2 = abs(iz);
    const T BIG = big_value<T>();
    const T inc0 = n0 ? T(1) / T(n0) : BIG;
    const T inc1 = n1 ? T(1) / T(n1) : BIG;
    const T inc2 = n2 ? T(1) / T(n2) : BIG;
    T a0 = T(0.5) * inc0, a1 = T(0.5) * inc1, a2 = T(0.5) * inc2;
    const int st0 = ix > 0 ? 1 : -1, st1 = iy > 0 ? nx : -nx, st2 = iz > 0 ? nxy : -nxy;
    const int nsteps = n0 + n1 + n2;
    T prev = T(0), last = T(0);
    for (int t = 0; t < nsteps; ++t) {
        const bool pz = (a2 <= a0) && (a2 <= a1);
        const bool py = !pz && (a1 <= a0);
        const T a = pz ? a2 : (py ? a1 : a0);
        f(v, r * (a - prev));
        prev = a;
        v += pz ? st2 : (py ? st1 : st0);
        last = pz ? inc2 : (py ? inc1 : inc0);
        a2 = pz ? a2 + inc2 : a2;
        a1 = py ? a1 + inc1 : a1;
        a0 = (!pz && !py) ? a0 + inc0 : a0;
    }
    f(v, r * T(0.5) * last);
}

// address-space views: wave-uniform reads through the scalar cache, per-lane
// reads as global (not flat) loads
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
using gptr = const __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) { return (cptr<T>)(p); }
template <typename T>
__device__ __forceinline__ gptr<T> as_global(const T *p) { return (gptr<T>)(p); }

// Whether the self-block moments M_k[a][b] = sum_w R_a R_b beta^k sin(theta)
// can couple real-basis coefficients a and b.  Reflecting the voxel in x, y
// or z multiplies R_lm by (-1)^|m| (x; once more -1 for the sin type), by +1
// for the cos type and -1 for the sin type (y), and by (-1)^(l+|m|) (z); the
// angular grid and beta share those symmetries, so an entry survives only if
// R_a R_b is even under all three: the same type, |m| and l of equal parity.
// (With beta in the reference's float geometry the others are ~1e-7 of the
// largest entry rather than 0: the FP32 path skips them, FP64 keeps them.)
__host__ __device__ constexpr bool self_couples(int a, int b)
{
    int la = 0, lb = 0;
    while ((la + 1) * (la + 1) <= a) ++la;
    while ((lb + 1) * (lb + 1) <= b) ++lb;
    const int ma = a - la * la - la, mb = b - lb * lb - lb;
    const bool sin_a = ma < 0, sin_b = mb < 0;
    const int am = ma < 0 ? -ma : ma, bm = mb < 0 ? -mb : mb;
    return sin_a == sin_b && ((am - bm) & 1) == 0 && ((la - lb) & 1) == 0;
}

// s += M[a][b] w[b] over the b that self_couples(a, b) keeps, and the same for
// a second vector; unrolled at compile time, so the skipped entries cost nothing
template <int a, int b, int K, typename T>
__device__ __forceinline__ void self_term(T &s, const T *__restrict__ M, const T (&w)[K])
{
    if constexpr (self_couples(a, b)) s += M[a * K + b] * w[b];
}
template <int a, int K, typename T, int... bs>
__device__ __forceinline__ T self_row(const T *__restrict__ M, const T (&w)[K], std::integer_sequence<int, bs...>)
{
    T s = T(0);
    (self_term<a, bs, K>(s, M, w), ...);
    return s;
}
// st[a] += coef * sum_b M[a][b] w[b] for every a (the kept entries)
template <int K, typename T, int... as>
__device__ __forceinline__ void self_apply(T (&st)[K], T coef, const T *__restrict__ M, const T (&w)[K],
                                           std::integer_sequence<int, as...>)
{
    ((st[as] += coef * self_row<as, K>(M, w, std::make_integer_sequence<int, K>{})), ...);
}
// as self_apply, and gk += sum_a x[a] (M w)[a] (the mu-derivative term of the adjoint)
template <int K, typename T, int... as>
__device__ __forceinline__ void self_apply_adj(T (&st)[K], T &gk, T coef, const T *__restrict__ M,
                                               const T (&x)[K], const T (&w)[K], std::integer_sequence<int, as...>)
{
    ((st[as] += coef * self_row<as, K>(M, x, std::make_integer_sequence<int, K>{}),
      gk += x[as] * self_row<as, K>(M, w, std::make_integer_sequence<int, K>{})),
     ...);
}

// destination of a lane: brick blk, lane -> (di, dj, dk)
template <typename T>
__device__ __forceinline__ void lane_dest(const GridDev<T> &g, int blk, int lane, int &di, int &dj,
                                          int &dk)
{
    const int bxi = blk % g.bx, byi = (blk / g.bx) % g.by, bzi = blk / (g.bx * g.by);
    di = bxi * kBrick + (lane & 3);
    dj = byi * kBrick + ((lane >> 2) & 3);
    dk = bzi * kBrick + (lane >> 4);
}


}  // namespace rte
