// rte_adjoint.hip -- reverse mode of the truncated Neumann series: the
// vector-Jacobian product of the bottom-face exitance with respect to the
// per-voxel mu_a and mu_s (the reference differentiates by one extra forward
// Neumann solve per parameter, FisherInfo_IO/src/compute_grad_rte.cu:273-385;
// here one adjoint series serves all voxels).
//
// With W_0 = s0, W_{q+1} = D A W_q (q < n-1) and E = ball + e^T A sum_q W_q,
// for a cotangent c on the image (eps = sum_p c_p e_p):
//   xi_{n-1} = eps,  Lambda_q = A^T xi_q,  xi_{q-1} = eps + D Lambda_q,
//   dE.c = sum_q xi_q^T dA W_q + sum_{q>=1} Lambda_q^T dD (A W_{q-1})
//          + Lambda_0^T ds0 + c^T dball.
// A^T applies the off-diagonal part through reciprocity, A_off^T = P A_off P
// (tau and r symmetric, R(-u) = (-1)^l R(u)), and the self block S = S^T as is.
// The mu_t-derivative of A_off goes through k(d,s) = e^{-tau}/r^2:
//   d k / d mu_t(v) = -len_v(d,s) k, len_v the Siddon length of the ray in v,
// so the adjoint propagation also accumulates per pair
//   w(d,s) = sum_q k (xi_q(d).R) (R.W_q(s))
// which one back-projection along the rays turns into the gradient.
#include "rte_device.hpp"

namespace rte {

// Up to kMaxNC adjoint series (cotangents) share each propagation launch: the
// pair geometry, its z-power values and the projection of the forward term are
// formed once per pair for all of them.
template <typename T, typename TA>
struct AdjDev {
    const T *Xz[kMaxNC];  // [N][K] adjoint inputs n_a (-1)^l xi_a, in the z-power basis
    const T *Wqz;         // [N][K] forward term n_a W_a, in the z-power basis
    T *part[kMaxNC];      // [nchunk][N][K] (z-power basis; k_adj_combine turns the sum back)
    T *pairbuf[kMaxNC];   // [N][npad] pair weights w(d = source slot row, s = lane)
    const T *ktab;        // [N][npad]
};

template <int L>
__device__ __forceinline__ int l_of(int a)
{
    int l = 0;
#pragma unroll
    for (int ll = 1; ll <= L; ++ll) l += (a >= ll * ll) ? 1 : 0;
    return l;
}

// ---------------------------------------------------------------------------
// A_off applied to P xi_c (uniform input u, lane output o) for NC cotangents c,
// plus the pair weights pairbuf_c[u][slot(o)] += k (Z.P xi_c(u)) (Z.P W_q(o)).
// In the z-power basis Z(u) of the forward kernel: P acts on it as on the
// harmonics, (-1)^l on entry (l, m), since Z_l,m(-u) = (-1)^l Z_l,m(u), and the
// pair weight, a product of two projections, does not depend on the basis.
// grid = (nblk, nchunk), block = 64.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L, int NC>
__global__ __launch_bounds__(kWave) void k_prop_adj(PropArgs<T> A, AdjDev<T, TA> D)
{
#pragma clang fp contract(off)
    constexpr int K = (L + 1) * (L + 1);
    const GridDev<T> &g = A.g;
    const int lane = threadIdx.x, blk = blockIdx.x, chunk = blockIdx.y;
    int di, dj, dk;
    lane_dest(g, blk, lane, di, dj, dk);
    const bool valid = di < g.nx && dj < g.ny && dk < g.nz;
    const int cdi = min(di, g.nx - 1), cdj = min(dj, g.ny - 1), cdk = min(dk, g.nz - 1);
    const int vo = cdi + g.nx * (cdj + g.ny * cdk);
    const int slot = blk * kWave + lane;
    const int s0 = chunk * A.chunk_len;
    const int s1 = min(g.N, s0 + A.chunk_len);
    const int nxy = g.nx * g.ny;

    T acc[NC][K], wl[K];
#pragma unroll
    for (int a = 0; a < K; ++a) {
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c][a] = T(0);
        const T sgn = (l_of<L>(a) & 1) ? T(-1) : T(1);
        wl[a] = sgn * D.Wqz[(size_t)vo * K + a];
    }
    cptr<T> X[NC];
    T *pb[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        X[c] = as_const<T>(D.Xz[c]);
        pb[c] = D.pairbuf[c] + slot;
    }
    const gptr<T> ktab = as_global<T>(D.ktab);
    const T dxc = T(cdi) * g.dx, dyc = T(cdj) * g.dy, dzc = T(cdk) * g.dz;
    int si = s0 % g.nx, sj = (s0 / g.nx) % g.ny, sk = s0 / nxy;
    // the table entry and the pair weights of the next source are loaded one source
    // ahead (the last source's again at the chunk's end: in bounds, unused); a lane
    // outside the grid keeps its weights in registers and never stores them
    T kf_next = s0 < s1 ? __builtin_nontemporal_load(ktab + (size_t)s0 * g.npad + slot) : T(0), pw_next[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) pw_next[c] = s0 < s1 ? pb[c][(size_t)s0 * g.npad] : T(0);
    for (int s = s0; s < s1; ++s) {
        const T ddx = fmad(-T(si), g.dx, dxc), ddy = fmad(-T(sj), g.dy, dyc), ddz = fmad(-T(sk), g.dz, dzc);
        // the self pair's table entry is 0 and its direction (0, 0, 0): r^2 clamped as in k_prop
        const T r2 = max(fmad(ddx, ddx, fmad(ddy, ddy, ddz * ddz)), T(1e-30));
        const T kf = kf_next;
        T pw[NC];
        const int sn = (s + 1 < s1) ? s + 1 : s;
        kf_next = __builtin_nontemporal_load(ktab + (size_t)sn * g.npad + slot);  // read once, as in k_prop
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            pw[c] = pw_next[c];
            pw_next[c] = pb[c][(size_t)sn * g.npad];
        }
        const T rinv = dev_rsqrt<T>(r2);
        PairZ<T, L> G;
        pair_zpow<T, L>(G, ddx * rinv, ddy * rinv, ddz * rinv);
        const T dl = pair_dot_z<T, L>(wl, G);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T w[K];
#pragma unroll
            for (int a = 0; a < K; ++a) w[a] = X[c][(size_t)s * K + a];
            const T t = kf * pair_dot_z<T, L>(w, G);
            pair_axpy_z<T, L>(acc[c], t, G);
            if (valid) pb[c][(size_t)s * g.npad] = fmad(t, dl, pw[c]);
        }
        if (++si == g.nx) {
            si = 0;
            if (++sj == g.ny) {
                sj = 0;
                ++sk;
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T *__restrict__ out = D.part[c] + ((size_t)chunk * g.N + vo) * K;
#pragma unroll
            for (int a = 0; a < K; ++a) out[a] = acc[c][a];
        }
    }
}

// [N][K] buffers into the z-power basis, out_i = M^T in_i (one thread per voxel
// and buffer; grid = (ceil(N / 256), n), block = 256)
template <typename T>
struct ZBufs {
    const T *in[kMaxNC + 1];
    T *out[kMaxNC + 1];
};

template <typename T, int L>
__global__ __launch_bounds__(256) void k_bufs_to_zpow(int N, ZBufs<T> B)
{
    constexpr int K = (L + 1) * (L + 1);
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const T *__restrict__ w = B.in[blockIdx.y] + (size_t)v * K;
    T *__restrict__ o = B.out[blockIdx.y] + (size_t)v * K;
    double x[K], y[K];
#pragma unroll
    for (int a = 0; a < K; ++a) x[a] = double(w[a]);
    zpow_from_harmonic<double, L>(x, y);
#pragma unroll
    for (int j = 0; j < K; ++j) o[j] = T(y[j]);
}

// ---------------------------------------------------------------------------
// Per voxel: Lambda_q = A^T xi_q, the self-block and scatter gradient terms,
// and the next adjoint input xi_{q-1} = eps + D Lambda_q (written over X).
// grid = ceil(N/64), block = 64.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L>
__global__ __launch_bounds__(kWave) void k_adj_combine(AdjCombArgs<T, TA> A)
{
    constexpr int K = (L + 1) * (L + 1);
    const GridDev<T> &g = A.g;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.N) return;
    TA acc[K], xi[K], w[K];
#pragma unroll
    for (int a = 0; a < K; ++a) {
        const TA sgn = (l_of<L>(a) & 1) ? TA(-1) : TA(1);
        xi[a] = sgn * TA(A.X[(size_t)v * K + a]) / A.nsc[a];
        w[a] = TA(A.Wq[(size_t)v * K + a]) / A.nsc[a];
        acc[a] = TA(0);
    }
    for (int c = 0; c < A.nchunk; ++c) {
        const T *__restrict__ p = A.part + ((size_t)c * g.N + v) * K;
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] += TA(p[a]);
    }
    zpow_to_harmonic<TA, L>(acc);
#pragma unroll
    for (int a = 0; a < K; ++a) {
        const TA sgn = (l_of<L>(a) & 1) ? TA(-1) : TA(1);
        acc[a] = sgn * acc[a] * A.nsc[a];
    }
    // self block S(mu) = self_scale sum_k c_k M_k, c_k = (-mu)^{k-1}/k!, and its mu-derivative;
    // the FP32 path uses the same FP32 entries as the forward combine (S symmetric)
    TA gself = TA(0);
    if constexpr (sizeof(T) == 4) {
        T xt[K], wt[K], st[K];
#pragma unroll
        for (int a = 0; a < K; ++a) {
            xt[a] = T(xi[a]);
            wt[a] = T(w[a]);
            st[a] = T(0);
        }
        const T mu = A.mut[v];
        T coef = T(1), dcoef = T(0), pw = T(1), fact = T(1), gs = T(0);
        for (int k = 0; k < kNMom; ++k) {
            fact *= T(k + 1);
            dcoef = (k == 0) ? T(0) : -T(k) * pw / fact;
            pw = (k == 0) ? T(1) : pw * (-mu);
            coef = pw / fact;
            T gk = T(0);
            self_apply_adj<K>(st, gk, coef, A.momT + (size_t)k * K * K, xt, wt, std::make_integer_sequence<int, K>{});
            gs += dcoef * gk;
        }
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] += A.self_scale * TA(st[a]);
        gself = TA(gs);
    } else {
        const TA mu = TA(A.mut[v]);
        TA coef = TA(1), dcoef = TA(0), pw = TA(1);  // pw = (-mu)^{k-1}
        TA fact = TA(1);
        for (int k = 0; k < kNMom; ++k) {
            // k here is moment index k+1: c = (-mu)^k/(k+1)!, dc/dmu = -k (-mu)^{k-1}/(k+1)!
            fact *= TA(k + 1);
            dcoef = (k == 0) ? TA(0) : -TA(k) * pw / fact;
            pw = (k == 0) ? TA(1) : pw * (-mu);
            coef = pw / fact;
            const TA *__restrict__ M = A.mom + (size_t)k * K * K;
            TA gk = TA(0);
#pragma unroll
            for (int a = 0; a < K; ++a) {
                TA sx = TA(0), sw = TA(0);
#pragma unroll
                for (int b = 0; b < K; ++b) {
                    sx += M[a * K + b] * xi[b];
                    sw += M[a * K + b] * w[b];
                }
                acc[a] += A.self_scale * coef * sx;
                gk += xi[a] * sw;
            }
            gself += dcoef * gk;
        }
    }
    A.gt[v] += A.self_scale * A.dV * gself;
#pragma unroll
    for (int a = 0; a < K; ++a) acc[a] *= A.dV;   // acc = Lambda_q(v)
    const TA ms = TA(A.mus[v]);
    if (A.aw_prev) {
        TA gsv = TA(0), gl = TA(1);
#pragma unroll
        for (int l = 0; l <= L; ++l) {
#pragma unroll
            for (int m = -l; m <= l; ++m) gsv += acc[l * l + l + m] * gl * A.aw_prev[(size_t)v * K + l * l + l + m];
            gl *= A.g_aniso;
        }
        A.gs[v] += gsv;
    }
    if (A.lam0) {
#pragma unroll
        for (int l = 0; l <= L; ++l) A.lam0[(size_t)v * (L + 1) + l] = acc[l * l + l];
    }
    const int zk = v / (g.nx * g.ny);
    const int pix = v - zk * g.nx * g.ny;
    const bool bottom = (zk == g.nz - 1);
    TA gl = TA(1);
#pragma unroll
    for (int l = 0; l <= L; ++l) {
        const TA sgn = (l & 1) ? TA(-1) : TA(1);
#pragma unroll
        for (int m = -l; m <= l; ++m) {
            const int a = l * l + l + m;
            TA xn = gl * ms * acc[a];
            if (m == 0 && bottom) xn += A.cot[pix] * A.coefE[l];
            A.X[(size_t)v * K + a] = T(sgn * xn * A.nsc[a]);
        }
        gl *= A.g_aniso;
    }
}

// X <- n (-1)^l eps: the adjoint input of the last term.  grid = ceil(N/64).
template <typename T, typename TA, int L>
__global__ void k_adj_init(GridDev<T> g, const TA *cot, Scales sc, const TA *coefE, T *X)
{
    constexpr int K = (L + 1) * (L + 1);
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.N) return;
    const int zk = v / (g.nx * g.ny);
    const int pix = v - zk * g.nx * g.ny;
#pragma unroll
    for (int l = 0; l <= L; ++l) {
        const double sgn = (l & 1) ? -1.0 : 1.0;
#pragma unroll
        for (int m = -l; m <= l; ++m) {
            const int a = l * l + l + m;
            double x = 0.0;
            if (m == 0 && zk == g.nz - 1) x = (double)cot[pix] * (double)coefE[l];
            X[(size_t)v * K + a] = T(sgn * x * sc.n[a]);
        }
    }
}

// Wq <- n (g^l mu_s A W_{q-1}): the forward term W_q from the stored A W_{q-1}.
template <typename T, typename TA, int L>
__global__ void k_make_wq(GridDev<T> g, Scales sc, TA gan, const T *mus, const TA *aw_prev, T *Wq)
{
    constexpr int K = (L + 1) * (L + 1);
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.N) return;
    const TA ms = TA(mus[v]);
    TA gl = TA(1);
#pragma unroll
    for (int l = 0; l <= L; ++l) {
#pragma unroll
        for (int m = -l; m <= l; ++m) {
            const int a = l * l + l + m;
            Wq[(size_t)v * K + a] = T(gl * ms * aw_prev[(size_t)v * K + a] * TA(sc.n[a]));
        }
        gl *= gan;
    }
}

// ---------------------------------------------------------------------------
// Back-projection: gt(v) += scale * sum_pairs w(pair) len_v(pair), each pair's
// ray walked as the path table was built, accumulated in LDS when the grid fits
// and in global memory otherwise, over unordered pairs: the ray between two voxel
// centres crosses the same voxels with the same Siddon lengths both ways (the
// crossings alpha = (j + 1/2)/n mirror into each other; only zero-length
// segments at exact ties may fall in different voxels), so the pair weights of
// (d, s) and (s, d) are added and the ray is marched once.  A wave owns a
// brick pair (bd, bs), bs <= bd (the path table's tiling), lane = destination
// d in bd, and walks the 64 sources of bs; on the diagonal tile only the pairs
// with slot(s) < slot(d).  grid = ceil(ntiles / waves per block), block =
// 1024 (LDS accumulator) or 256.
// ---------------------------------------------------------------------------
template <typename T, typename TA, bool LDS>
__global__ __launch_bounds__(1024) void k_backproject_sym(GridDev<T> g, int ntiles, T thr2, const T *pairbuf,
                                                          TA scale, TA *gt)
{
    extern __shared__ unsigned char smem_raw[];
    TA *sacc = reinterpret_cast<TA *>(smem_raw);
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    if constexpr (LDS) {
        for (int i = tid; i < g.N; i += blockDim.x) sacc[i] = TA(0);
        __syncthreads();
    }
    const int tile = blockIdx.x * (blockDim.x / kWave) + tid / kWave;
    if (tile < ntiles) {
        int bd = (int)((sqrtf(8.0f * (float)tile + 1.0f) - 1.0f) * 0.5f);
        while (bd > 0 && bd * (bd + 1) / 2 > tile) --bd;
        while ((bd + 1) * (bd + 2) / 2 <= tile) ++bd;
        const int bs = tile - bd * (bd + 1) / 2;
        int di, dj, dk;
        lane_dest(g, bd, lane, di, dj, dk);
        const bool dvalid = di < g.nx && dj < g.ny && dk < g.nz;
        const int cdi = min(di, g.nx - 1), cdj = min(dj, g.ny - 1), cdk = min(dk, g.nz - 1);
        const int nxy = g.nx * g.ny;
        const int dslot = bd * kWave + lane;
        const int dlin = cdi + g.nx * cdj + nxy * cdk;
        const int bxi = bs % g.bx, byi = (bs / g.bx) % g.by, bzi = bs / (g.bx * g.by);
        for (int q = 0; q < kWave; ++q) {
            const int si = bxi * kBrick + (q & 3), sj = byi * kBrick + ((q >> 2) & 3), sk = bzi * kBrick + (q >> 4);
            if (si >= g.nx || sj >= g.ny || sk >= g.nz) continue;  // wave-uniform
            const int sslot = bs * kWave + q;
            const int ix = cdi - si, iy = cdj - sj, iz = cdk - sk;
            const T ddx = T(ix) * g.dx, ddy = T(iy) * g.dy, ddz = T(iz) * g.dz;
            const T r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const bool self = (ix == 0) && (iy == 0) && (iz == 0);
            const bool once = bs != bd || sslot < dslot;
            if (!(dvalid && !self && once && r2 < thr2)) continue;
            const int slin = si + g.nx * sj + nxy * sk;
            const TA w = TA(pairbuf[(size_t)slin * g.npad + dslot]) + TA(pairbuf[(size_t)dlin * g.npad + sslot]);
            if (w == TA(0)) continue;
            const TA sw = scale * w;
            march_visit<T>(slin, ix, iy, iz, g.nx, nxy, sqrt(r2), [&](int v, T len) {
                if constexpr (LDS)
                    atomicAdd(&sacc[v], sw * TA(len));
                else
                    atomicAdd(&gt[v], sw * TA(len));
            });
        }
    }
    if constexpr (LDS) {
        __syncthreads();
        for (int i = tid; i < g.N; i += blockDim.x)
            if (sacc[i] != TA(0)) atomicAdd(&gt[i], sacc[i]);
    }
}

// ---------------------------------------------------------------------------
// Back-projection onto parameter blocks (the exact Fisher information needs only
// sums of the mu_t gradient over each block, part[v] in [0, nparam) or -1): for
// NC backward passes at once, out_c[p] += scale * sum_pairs w_c(pair) len_p(pair),
// len_p the length of the pair's ray inside block p.  The tiling and the march are
// those of k_backproject_sym (unordered pairs, w_c = both directions' weights), one
// march for all NC passes; a lane adds the lengths of consecutive voxels of one
// block in a register and adds w_c times that run to the block's LDS accumulator
// when the ray leaves the block.  The part map is read from LDS.
// grid = ceil(ntiles / 16), block = 1024; LDS = N ints + NC nparam accumulators.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int NC>
__global__ __launch_bounds__(1024) void k_backproject_blocks(GridDev<T> g, int ntiles, T thr2, BlockBp<T, TA> B)
{
    extern __shared__ unsigned char smem_raw[];
    int *spart = reinterpret_cast<int *>(smem_raw);
    TA *sacc = reinterpret_cast<TA *>(smem_raw + (((size_t)g.N * sizeof(int) + 15) & ~size_t(15)));
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    for (int i = tid; i < g.N; i += blockDim.x) spart[i] = B.part[i];
    for (int i = tid; i < NC * B.nparam; i += blockDim.x) sacc[i] = TA(0);
    __syncthreads();
    const int tile = blockIdx.x * (blockDim.x / kWave) + tid / kWave;
    if (tile < ntiles) {
        int bd = (int)((sqrtf(8.0f * (float)tile + 1.0f) - 1.0f) * 0.5f);
        while (bd > 0 && bd * (bd + 1) / 2 > tile) --bd;
        while ((bd + 1) * (bd + 2) / 2 <= tile) ++bd;
        const int bs = tile - bd * (bd + 1) / 2;
        int di, dj, dk;
        lane_dest(g, bd, lane, di, dj, dk);
        const bool dvalid = di < g.nx && dj < g.ny && dk < g.nz;
        const int cdi = min(di, g.nx - 1), cdj = min(dj, g.ny - 1), cdk = min(dk, g.nz - 1);
        const int nxy = g.nx * g.ny;
        const int dslot = bd * kWave + lane;
        const int dlin = cdi + g.nx * cdj + nxy * cdk;
        const int bxi = bs % g.bx, byi = (bs / g.bx) % g.by, bzi = bs / (g.bx * g.by);
        for (int q = 0; q < kWave; ++q) {
            const int si = bxi * kBrick + (q & 3), sj = byi * kBrick + ((q >> 2) & 3), sk = bzi * kBrick + (q >> 4);
            if (si >= g.nx || sj >= g.ny || sk >= g.nz) continue;  // wave-uniform
            const int sslot = bs * kWave + q;
            const int ix = cdi - si, iy = cdj - sj, iz = cdk - sk;
            const T ddx = T(ix) * g.dx, ddy = T(iy) * g.dy, ddz = T(iz) * g.dz;
            const T r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const bool self = (ix == 0) && (iy == 0) && (iz == 0);
            const bool once = bs != bd || sslot < dslot;
            if (!(dvalid && !self && once && r2 < thr2)) continue;
            const int slin = si + g.nx * sj + nxy * sk;
            TA sw[NC];
            bool any = false;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                sw[c] = B.scale * (TA(B.pairbuf[c][(size_t)slin * g.npad + dslot]) +
                                   TA(B.pairbuf[c][(size_t)dlin * g.npad + sslot]));
                any = any || sw[c] != TA(0);
            }
            if (!any) continue;
            int cur = -1;
            TA run = TA(0);
            auto flush = [&]() {
                if (cur >= 0) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) atomicAdd(&sacc[c * B.nparam + cur], sw[c] * run);
                }
            };
            march_visit<T>(slin, ix, iy, iz, g.nx, nxy, sqrt(r2), [&](int v, T len) {
                const int p = spart[v];
                if (p != cur) {
                    flush();
                    cur = p;
                    run = TA(0);
                }
                run += TA(len);
            });
            flush();
        }
    }
    __syncthreads();
    for (int i = tid; i < NC * B.nparam; i += blockDim.x)
        if (sacc[i] != TA(0)) atomicAdd(&B.out[i / B.nparam][i % B.nparam], sacc[i]);
}

// ---------------------------------------------------------------------------
// Source and ballistic derivatives along each beam column (column products of
// k_source): s0_l0(k) = F_l h(k), h(k) = (mu_s/mu_t)(k) T(k) (1 - a_k),
// a_k = e^{-mu_t(k) dz}, T(k) = prod_{k'<k} a_k', ballistic = fact T(nz).
// ga/gs receive d(Lambda_0.s0 + c.ball)/d mu_a, d mu_s.  grid = ceil(P/64).
// ---------------------------------------------------------------------------
template <typename TA, int L>
__global__ void k_src_grad(int nx, int ny, int nz, double dxd, double dyd, double dzd, double xmin,
                           double ymin, double gan, BeamDev B, const TA *lam0, const TA *cot, TA *ga, TA *gs)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = nx * ny;
    if (pix >= P) return;
    const int i = pix % nx, j = pix / nx;
    const double xc = (double)i * dxd + xmin + dxd / 2.0;
    const double yc = (double)j * dyd + ymin + dyd / 2.0;
    const double rad = sqrt((xc - B.pos_x) * (xc - B.pos_x) + (yc - B.pos_y) * (yc - B.pos_y));
    if (!(rad <= B.half_diam)) return;
    const double dV = dxd * dyd * dzd;
    // F_l folded into Y(k) = sum_l Lambda0_l0(k) fact g^l sqrt((2l+1)/4pi) / dV
    double ball = 1.0;
    for (int k = 0; k < nz; ++k) ball *= exp(-(B.mua[pix + P * k] + B.mus[pix + P * k]) * dzd);
    ball *= B.fact;
    // suffix sum of Y(k) h(k) for k > k''
    double suffix = 0.0;
    for (int k = nz - 1; k >= 0; --k) {
        const int v = pix + P * k;
        double Tprod = 1.0;
        for (int kk = 0; kk < k; ++kk) Tprod *= exp(-(B.mua[pix + P * kk] + B.mus[pix + P * kk]) * dzd);
        const double ma = B.mua[v], ms = B.mus[v], mt = ma + ms;
        const double ak = exp(-mt * dzd);
        double Y = 0.0, gl = 1.0;
        for (int l = 0; l <= L; ++l) {
            Y += (double)lam0[(size_t)v * (L + 1) + l] * B.fact * gl * sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) / dV;
            gl *= gan;
        }
        const double h = (ms / mt) * Tprod * (1.0 - ak);
        // d h / d mu_t(k) at fixed mu_s, and the mu_s/mu_t factor
        const double dh_dmt = ms * Tprod * (-(1.0 - ak) / (mt * mt) + dzd * ak / mt);
        const double dh_dms_extra = Tprod * (1.0 - ak) / mt;
        const double up = -dzd * suffix;                 // later voxels of the column
        const double bal = -dzd * (double)cot[pix] * ball;
        ga[v] += TA(up + Y * dh_dmt + bal);
        gs[v] += TA(up + Y * (dh_dmt + dh_dms_extra) + bal);
        suffix += Y * h;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// The reference's forward-mode derivative (FisherInfo_IO/src/compute_grad_rte.cu:
// 273-301 mu_a, 341-371 mu_s): the first term of solve i is its parameter's
// region of the forward field phi = sum_q A W_q times -fs (mu_a) or
// fs (g^l - 1) (mu_s), in scaled coefficients; its pixel buffers start at zero
// (no ballistic term) with the bottom-layer stop-rule sums of that first term.
// fs = n: the reference scales its field out2 = phi / c_m by -1.0*C and
// (g^l - 1) C with C the vacuum speed (FisherInfo_IO/src/rte.h:15,
// compute_grad_rte.cu:292,350-353), which is phi times C / c_m = n here.
// wrt 2 with part == nullptr: phi itself, every voxel (a given first term,
// the reference's Neumann(src, out, 1), FisherInfo_IO/src/neumann.cu:23).
// grid = (ceil(N/64), nsolve), block = 64.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L>
__global__ __launch_bounds__(kWave) void k_ref_source(GridDev<T> g, Scales sc, const int *__restrict__ part,
                                                     const int *__restrict__ param, int wrt, TA gan, TA fs,
                                                     const TA *__restrict__ phi, const TA *__restrict__ wsum,
                                                     const SolveDev<T, TA> *__restrict__ sv)
{
    constexpr int K = (L + 1) * (L + 1);
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.N) return;
    const SolveDev<T, TA> S = sv[blockIdx.y];
    // part == nullptr: every voxel (a given first term, wrt 2)
    const bool in = part ? part[v] == param[blockIdx.y] : true;
    TA gl = TA(1), bs = TA(0);
#pragma unroll
    for (int l = 0; l <= L; ++l) {
        const TA f = wrt == 0 ? -fs : wrt == 1 ? fs * (gl - TA(1)) : TA(1);
#pragma unroll
        for (int m = -l; m <= l; ++m) {
            const int a = l * l + l + m;
            const T wt = in ? T(f * phi[(size_t)v * K + a] * TA(sc.n[a])) : T(0);
            S.W[(size_t)v * K + a] = wt;
            bs += wsum[a] * (TA(wt) / TA(sc.n[a]));
        }
        gl *= gan;
    }
    const int P = g.nx * g.ny;
    if (v >= (g.nz - 1) * P) {
        const int pix = v - (g.nz - 1) * P;
        S.E[pix] = TA(0);
        S.dE[pix] = TA(0);
        S.bsum[pix] = bs;
    }
    if (v < P) {
        S.R[v] = TA(0);
        S.dR[v] = TA(0);
    }
}

// y += x over n elements
template <typename TA>
__global__ void k_accumulate(size_t n, const TA *__restrict__ x, TA *__restrict__ y)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

#define RTE_DISPATCH_L2(L, ...)                          \
    switch (L) {                                        \
    case 0: { constexpr int LL = 0; __VA_ARGS__; } break; \
    case 1: { constexpr int LL = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int LL = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int LL = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int LL = 4; __VA_ARGS__; } break; \
    case 5: { constexpr int LL = 5; __VA_ARGS__; } break; \
    case 6: { constexpr int LL = 6; __VA_ARGS__; } break; \
    case 7: { constexpr int LL = 7; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;                \
    }

template <typename T, typename TA>
hipError_t launch_adj_prop(int L, const PropArgs<T> &A, int nc, const T *const *X, T *const *Xz, const T *Wq, T *Wqz,
                           T *const *part, T *const *pairbuf, const T *ktab, hipStream_t st)
{
    if (nc < 1 || nc > kMaxNC) return hipErrorInvalidValue;
    // the inputs into the z-power basis: the forward term and each cotangent's xi
    ZBufs<T> B{};
    B.in[0] = Wq;
    B.out[0] = Wqz;
    AdjDev<T, TA> D{};
    D.Wqz = Wqz;
    D.ktab = ktab;
    for (int c = 0; c < nc; ++c) {
        B.in[1 + c] = X[c];
        B.out[1 + c] = Xz[c];
        D.Xz[c] = Xz[c];
        D.part[c] = part[c];
        D.pairbuf[c] = pairbuf[c];
    }
    RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_bufs_to_zpow<T, LL>), dim3((A.g.N + 255) / 256, 1 + nc), dim3(256), 0,
                                          st, A.g.N, B));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 grid(A.g.nblk, A.nchunk), block(kWave);
    if (nc == 1) {
        RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_prop_adj<T, TA, LL, 1>), grid, block, 0, st, A, D));
    } else {
        RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_prop_adj<T, TA, LL, 2>), grid, block, 0, st, A, D));
    }
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_adj_combine(int L, const AdjCombArgs<T, TA> &A, hipStream_t st)
{
    dim3 grid((A.g.N + kWave - 1) / kWave), block(kWave);
    RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_adj_combine<T, TA, LL>), grid, block, 0, st, A));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_adj_init(int L, const GridDev<T> &g, const TA *cot, const Scales &sc, const TA *coefE, T *X,
                           hipStream_t st)
{
    dim3 grid((g.N + 63) / 64), block(64);
    RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_adj_init<T, TA, LL>), grid, block, 0, st, g, cot, sc, coefE, X));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_make_wq(int L, const GridDev<T> &g, const Scales &sc, TA gan, const T *mus, const TA *aw_prev,
                          T *Wq, hipStream_t st)
{
    dim3 grid((g.N + 63) / 64), block(64);
    RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_make_wq<T, TA, LL>), grid, block, 0, st, g, sc, gan, mus, aw_prev, Wq));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_backproject(const GridDev<T> &g, int nchunk, int chunk_len, T thr2, const T *pairbuf, TA scale,
                              TA *gt, hipStream_t st)
{
    static int lds_max = -1;
    if (lds_max < 0) {
        int dev = 0, v = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) return e;
        lds_max = v;
    }
    const size_t lds = (size_t)g.N * sizeof(TA);
    const int ntiles = g.nblk * (g.nblk + 1) / 2;
    (void)nchunk;
    (void)chunk_len;
    if (lds <= (size_t)lds_max) {
        // a large accumulator leaves room for one block per CU: make it 16 waves
        const int threads = lds > 64 * 1024 ? 1024 : 256;
        if (lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_backproject_sym<T, TA, true>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        const int waves = threads / kWave;
        hipLaunchKernelGGL((k_backproject_sym<T, TA, true>), dim3((ntiles + waves - 1) / waves), dim3(threads), lds,
                           st, g, ntiles, thr2, pairbuf, scale, gt);
    } else {
        constexpr int threads = 256, waves = threads / kWave;
        hipLaunchKernelGGL((k_backproject_sym<T, TA, false>), dim3((ntiles + waves - 1) / waves), dim3(threads), 0,
                           st, g, ntiles, thr2, pairbuf, scale, gt);
    }
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_backproject_blocks(const GridDev<T> &g, T thr2, const BlockBp<T, TA> &B, int nc, hipStream_t st)
{
    static int lds_max = -1;
    if (lds_max < 0) {
        int dev = 0, v = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) return e;
        lds_max = v;
    }
    if (nc < 1 || nc > kMaxNC || B.nparam < 1) return hipErrorInvalidValue;
    const size_t lds = (((size_t)g.N * sizeof(int) + 15) & ~size_t(15)) + (size_t)nc * B.nparam * sizeof(TA);
    if (lds > (size_t)lds_max) return hipErrorInvalidValue;  // the caller back-projects per voxel
    const int ntiles = g.nblk * (g.nblk + 1) / 2;
    constexpr int threads = 1024, waves = threads / kWave;
    const void *fn = nc == 1 ? reinterpret_cast<const void *>(&k_backproject_blocks<T, TA, 1>)
                             : reinterpret_cast<const void *>(&k_backproject_blocks<T, TA, 2>);
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (nc == 1)
        hipLaunchKernelGGL((k_backproject_blocks<T, TA, 1>), dim3((ntiles + waves - 1) / waves), dim3(threads), lds,
                           st, g, ntiles, thr2, B);
    else
        hipLaunchKernelGGL((k_backproject_blocks<T, TA, 2>), dim3((ntiles + waves - 1) / waves), dim3(threads), lds,
                           st, g, ntiles, thr2, B);
    return hipGetLastError();
}

template <typename TA>
hipError_t launch_src_grad(int L, int nx, int ny, int nz, double dx, double dy, double dz, double xmin, double ymin,
                           double gan, const BeamDev &B, const TA *lam0, const TA *cot, TA *ga, TA *gs, hipStream_t st)
{
    const int P = nx * ny;
    dim3 grid((P + 63) / 64), block(64);
    RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_src_grad<TA, LL>), grid, block, 0, st, nx, ny, nz, dx, dy, dz, xmin,
                                          ymin, gan, B, lam0, cot, ga, gs));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_ref_source(int L, const GridDev<T> &g, const Scales &sc, const int *part, const int *param, int wrt,
                             TA gan, TA fs, const TA *phi, const TA *wsum, const SolveDev<T, TA> *sv, int nsolve,
                             hipStream_t st)
{
    dim3 grid((g.N + kWave - 1) / kWave, nsolve), block(kWave);
    RTE_DISPATCH_L2(L, hipLaunchKernelGGL((k_ref_source<T, TA, LL>), grid, block, 0, st, g, sc, part, param, wrt,
                                          gan, fs, phi, wsum, sv));
    return hipGetLastError();
}

template <typename TA>
hipError_t launch_accumulate(size_t n, const TA *x, TA *y, hipStream_t st)
{
    hipLaunchKernelGGL(k_accumulate<TA>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, x, y);
    return hipGetLastError();
}

template hipError_t launch_ref_source<float, double>(int, const GridDev<float> &, const Scales &, const int *,
                                                     const int *, int, double, double, const double *, const double *,
                                                     const SolveDev<float, double> *, int, hipStream_t);
template hipError_t launch_ref_source<double, double>(int, const GridDev<double> &, const Scales &, const int *,
                                                      const int *, int, double, double, const double *, const double *,
                                                      const SolveDev<double, double> *, int, hipStream_t);
template hipError_t launch_accumulate<double>(size_t, const double *, double *, hipStream_t);
template hipError_t launch_adj_prop<float, double>(int, const PropArgs<float> &, int, const float *const *,
                                                   float *const *, const float *, float *, float *const *,
                                                   float *const *, const float *, hipStream_t);
template hipError_t launch_adj_prop<double, double>(int, const PropArgs<double> &, int, const double *const *,
                                                    double *const *, const double *, double *, double *const *,
                                                    double *const *, const double *, hipStream_t);
template hipError_t launch_adj_combine<float, double>(int, const AdjCombArgs<float, double> &, hipStream_t);
template hipError_t launch_adj_combine<double, double>(int, const AdjCombArgs<double, double> &, hipStream_t);
template hipError_t launch_adj_init<float, double>(int, const GridDev<float> &, const double *, const Scales &,
                                                   const double *, float *, hipStream_t);
template hipError_t launch_adj_init<double, double>(int, const GridDev<double> &, const double *, const Scales &,
                                                    const double *, double *, hipStream_t);
template hipError_t launch_make_wq<float, double>(int, const GridDev<float> &, const Scales &, double,
                                                  const float *, const double *, float *, hipStream_t);
template hipError_t launch_make_wq<double, double>(int, const GridDev<double> &, const Scales &, double,
                                                   const double *, const double *, double *, hipStream_t);
template hipError_t launch_backproject<float, double>(const GridDev<float> &, int, int, float, const float *,
                                                      double, double *, hipStream_t);
template hipError_t launch_backproject<double, double>(const GridDev<double> &, int, int, double, const double *,
                                                       double, double *, hipStream_t);
template hipError_t launch_backproject_blocks<float, double>(const GridDev<float> &, float,
                                                             const BlockBp<float, double> &, int, hipStream_t);
template hipError_t launch_backproject_blocks<double, double>(const GridDev<double> &, double,
                                                              const BlockBp<double, double> &, int, hipStream_t);
template hipError_t launch_src_grad<double>(int, int, int, int, double, double, double, double, double, double,
                                            const BeamDev &, const double *, const double *, double *, double *,
                                            hipStream_t);

}  // namespace rte
