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
__device__ __forceinline__ T fmad(T a, T b, T c);
template <>
__device__ __forceinline__ float fmad<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ __forceinline__ double fmad<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

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
// One (destination, source) pair in the factored real-SH form, on scaled
// coefficients (see sh_gamma):  p = sum_a Rt_a(u) w'_a ;  acc' += (kf p) Rt(u)
// with Rt_l0 = Pt_l^0, Rt_lm = Pt_l^m C_m, Rt_l,-m = Pt_l^m S_m,
// C_m + i S_m = (x + i y)^m.  The adjoint propagation (k_prop_adj) uses this
// form; the forward propagation uses the expanded one below.
// ---------------------------------------------------------------------------
template <typename T, int L>
struct PairGeom {
    T C[L + 1], S[L + 1];
    T P[L + 1][L + 1];  // only l >= m used
};

template <typename T, int L>
__device__ __forceinline__ void pair_geom(PairGeom<T, L> &G, T x, T y, T z)
{
    G.C[0] = T(1);
    G.S[0] = T(0);
    if constexpr (L >= 1) {
        G.C[1] = x;
        G.S[1] = y;
    }
#pragma unroll
    for (int m = 2; m <= L; ++m) {
        G.C[m] = G.C[m - 1] * x - G.S[m - 1] * y;
        G.S[m] = G.S[m - 1] * x + G.C[m - 1] * y;
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) {
        G.P[m][m] = T(1);
        if (m + 1 <= L) G.P[m + 1][m] = z;
#pragma unroll
        for (int l = m + 2; l <= L; ++l) G.P[l][m] = z * G.P[l - 1][m] - sh_gamma<T>(l, m) * G.P[l - 2][m];
    }
}

// p = sum_a Rt_a(u) w'_a
template <typename T, int L>
__device__ __forceinline__ T pair_dot(const T (&w)[(L + 1) * (L + 1)], const PairGeom<T, L> &G)
{
    // two partial sums: halves the dependent chain of the dot product
    T p = w[0], q = T(0);
#pragma unroll
    for (int l = 1; l <= L; ++l) p += G.P[l][0] * w[l * l + l];
#pragma unroll
    for (int m = 1; m <= L; ++m) {
        T A = w[m * m + m + m], B = w[m * m + m - m];
#pragma unroll
        for (int l = m + 1; l <= L; ++l) {
            A += G.P[l][m] * w[l * l + l + m];
            B += G.P[l][m] * w[l * l + l - m];
        }
        q += G.C[m] * A;
        q += G.S[m] * B;
    }
    return p + q;
}

// acc' += t Rt(u)
template <typename T, int L>
__device__ __forceinline__ void pair_axpy(T (&acc)[(L + 1) * (L + 1)], T t, const PairGeom<T, L> &G)
{
    acc[0] += t;
#pragma unroll
    for (int l = 1; l <= L; ++l) acc[l * l + l] += t * G.P[l][0];
#pragma unroll
    for (int m = 1; m <= L; ++m) {
        const T u = t * G.C[m], vv = t * G.S[m];
        acc[m * m + m + m] += u;
        acc[m * m + m - m] += vv;
#pragma unroll
        for (int l = m + 1; l <= L; ++l) {
            acc[l * l + l + m] += u * G.P[l][m];
            acc[l * l + l - m] += vv * G.P[l][m];
        }
    }
}

// ---------------------------------------------------------------------------
// The same pair in expanded form for the propagation kernel (every group,
// real or complex): the direction's scaled real harmonics Rt_a(u) are formed
// once per pair,
//   Rt_l0 = Pt_l^0,  Rt_l,m = C_m Pt_l^m,  Rt_l,-m = S_m Pt_l^m   (Rt_00 = 1),
// the products by the Legendre recurrence in l itself (it is linear, so it
// holds for C_m Pt_l^m with the same gamma):
//   X_m,m = C_m,  X_m+1,m = z C_m,  X_l,m = z X_l-1,m - gamma_lm X_l-2,m,
// and every right-hand side then costs one 36-term dot and one axpy of FMAs,
// with no per-side products C_m t, S_m t or partial-sum set-up.
// ---------------------------------------------------------------------------
template <typename T, int L>
struct PairRt {
    T R[(L + 1) * (L + 1)];  // R[0] = 1 is not used
};

template <typename T, int L>
__device__ __forceinline__ void pair_rt(PairRt<T, L> &G, T x, T y, T z)
{
#pragma clang fp contract(off)
    if constexpr (L >= 1) {
        G.R[2] = z;
        T p2 = T(1), p1 = z;
#pragma unroll
        for (int l = 2; l <= L; ++l) {
            const T p = fmad(z, p1, -(sh_gamma<T>(l, 0) * p2));
            G.R[l * l + l] = p;
            p2 = p1;
            p1 = p;
        }
        T c = x, s = y;  // C_1 + i S_1 = x + i y
#pragma unroll
        for (int m = 1; m <= L; ++m) {
            if (m > 1) {
                const T cn = fmad(c, x, -(s * y)), sn = fmad(s, x, c * y);
                c = cn;
                s = sn;
            }
            G.R[m * m + m + m] = c;
            G.R[m * m + m - m] = s;
            // the Legendre factor Pt_l^m (Pt_m^m = 1, Pt_m+1^m = z), then its
            // products with C_m and S_m
            T q2 = T(1), q1 = z;
#pragma unroll
            for (int l = m + 1; l <= L; ++l) {
                const T q = (l == m + 1) ? z : fmad(z, q1, -(sh_gamma<T>(l, m) * q2));
                if (l > m + 1) {
                    q2 = q1;
                    q1 = q;
                }
                G.R[l * l + l + m] = c * q;
                G.R[l * l + l - m] = s * q;
            }
        }
    }
}

// p = sum_a Rt_a w'_a, in two interleaved chains (odd and even a)
template <typename T, int L>
__device__ __forceinline__ T pair_dot_rt(const T (&w)[(L + 1) * (L + 1)], const PairRt<T, L> &G)
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

// the chains of pair_dot_rt over coefficients [A0, A0 + N) of a row, read from
// h[a - A0]: a projection taken in parts, each chain in the same order, so that
// parts covering [0, K) in turn leave p + q equal to pair_dot_rt's value
template <typename T, int L, int A0, int N>
__device__ __forceinline__ void pair_dot_part(T &p, T &q, const T *h, const PairRt<T, L> &G)
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

// acc' += t Rt
template <typename T, int L>
__device__ __forceinline__ void pair_axpy_rt(T (&acc)[(L + 1) * (L + 1)], T t, const PairRt<T, L> &G)
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
