// rte_kernels.hip -- CDNA4 (gfx950) kernels of the Neumann-series RTE solve.
//
// Propagation A (the hot path, CPU package/src/Kernels_selfsub.c:450-538 and
// GPU package/src/kernels.cu:286-364) is evaluated pair by pair in the real SH
// basis, factored through the direction of the pair:
//   (A W)_a(d) = dV [ sum_{s != d} k(d,s) R_a(u) (R(u) . W(s)) + (S(mu_t(d)) W(d))_a ],
//   k(d,s) = e^{-tau(d,s)} / r^2,   u = (r_d - r_s)/r,
// i.e. 2K FMAs per pair instead of the reference's K^2 complex MACs.
//
// Work decomposition: one wave64 owns a 4x4x4 brick of destination voxels (one
// per lane) and walks a contiguous chunk of source voxels in lockstep, so the
// source's coefficients W(s) are wave-uniform (scalar loads into SGPRs) and
// every lane is a valid pair; R(u) is taken per lane in the basis C_m z^k,
// S_m z^k (the harmonics with their Legendre factor replaced by powers of z:
// one product each; W is converted into that basis before the launch and the
// summed partials back in k_combine), then one dot product and one axpy per
// right-hand side.  The path integral
// either comes from a per-medium table (cached mode, written once by
// k_ktab) or is marched per pair in every term (march mode, the reference's
// structure, GPU package/src/kernels.cu:49-145).  Partial sums per source
// chunk are combined in a fixed order by k_combine, so results do not depend
// on how many solves share a launch.
#include "rte_device.hpp"

namespace rte {

template <typename T>
__global__ void k_fill(T *__restrict__ p, size_t n, T v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}


// ---------------------------------------------------------------------------
// Propagation: partial pair sums over one chunk of sources.
//   grid = (ceil(nblk / kPropWaves), nchunk, ngroups), block = 64 kPropWaves:
//   wave i of workgroup x owns brick x kPropWaves + i.  The waves of a
//   workgroup share one CU and walk the same sources of the same solves at
//   the same time, so one scalar-cache fill of a W row serves all of them.
//   Group z holds the NR solves sv[z*NR .. z*NR+NR-1], which share one medium
//   (same path table and mu_t): the pair direction, its z-power values and kf
//   are computed once and applied to every right-hand side.
// PATH 0: kf from the per-medium table; PATH 1: marched per pair.
// ---------------------------------------------------------------------------
constexpr int kPropWaves = 4;
template <typename T, typename TA, int L, int PATH, int NR, bool CPLX = false, bool SEP = false>
__global__ __launch_bounds__(kWave * kPropWaves) void k_prop(PropArgs<T> A, const SolveDev<T, TA> *__restrict__ sv)
{
#pragma clang fp contract(off)
    static_assert(!CPLX || NR == 2, "a complex field is propagated as its two real halves");
    // SEP: the group's right-hand sides are solves on different media of one grid; they
    // share the pair geometry and each reads its own path table
    static_assert(!SEP || (PATH == 0 && !CPLX), "separate media need the cached path");
    constexpr int NT = SEP ? NR : 1;  // path tables read per pair
    constexpr int K = (L + 1) * (L + 1);
    const GridDev<T> &g = A.g;
    const int lane = threadIdx.x % kWave, blk = blockIdx.x * kPropWaves + threadIdx.x / kWave;
    const int chunk = blockIdx.y;
    if (blk >= g.nblk) return;  // the last workgroup's waves past the last brick (no barriers follow)
    const SolveDev<T, TA> *__restrict__ grp = sv + (size_t)blockIdx.z * NR;
    int di, dj, dk;
    lane_dest(g, blk, lane, di, dj, dk);
    const bool valid = di < g.nx && dj < g.ny && dk < g.nz;
    // keep marching lanes inside the grid; their kf is zeroed below
    const int cdi = min(di, g.nx - 1), cdj = min(dj, g.ny - 1), cdk = min(dk, g.nz - 1);
    const int slot = blk * kWave + lane;
    const int s0 = chunk * A.chunk_len;
    const int s1 = min(g.N, s0 + A.chunk_len);
    const int nxy = g.nx * g.ny;

    T acc[NR][K];
    cptr<T> Wsrc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        Wsrc[r] = as_const<T>(grp[r].Wp);
#pragma unroll
        for (int a = 0; a < K; ++a) acc[r][a] = T(0);
    }
    gptr<T> ktab[NT];
#pragma unroll
    for (int r = 0; r < NT; ++r) ktab[r] = as_global<T>(grp[r].ktab);
    const gptr<T> mut = as_global<T>(grp[0].mut);
    const T kim = grp[0].kim;
    // destination coordinates; source coordinates advance by dx along x
    const T dxc = T(cdi) * g.dx, dyc = T(cdj) * g.dy, dzc = T(cdk) * g.dz;
    int si = s0 % g.nx, sj = (s0 / g.nx) % g.ny, sk = s0 / nxy;
    T kf_next[NT];
#pragma unroll
    for (int r = 0; r < NT; ++r) kf_next[r] = T(0);
    if constexpr (PATH == 0) {
#pragma unroll
        for (int r = 0; r < NT; ++r)
            if (s0 < s1) kf_next[r] = __builtin_nontemporal_load(ktab[r] + (size_t)s0 * g.npad + slot);
    }
    // FP64 complex group: the pair phase (cos, sin)(kim r) from the per-displacement
    // table (read where it is used: requested a source ahead, the two values live
    // across the loop cost this SGPR-bound kernel over 300 spill moves per pair)
    constexpr bool PHASE_TAB = CPLX && sizeof(T) == 8;
    const gptr<T> phase = as_global<T>(grp[0].phase);
    const int ex = 2 * g.nx - 1, exy = ex * (2 * g.ny - 1);
    const int dbase = (cdi + g.nx - 1) + ex * (cdj + g.ny - 1) + exy * (cdk + g.nz - 1);
    T w[K];
#pragma unroll
    for (int a = 0; a < K; ++a) w[a] = (s0 < s1) ? Wsrc[0][(size_t)s0 * K + a] : T(0);
    // The W rows of a chunk's sources are read by every brick, but through the scalar
    // cache, whose misses wait for all scalar loads in flight; fetched from beyond L2
    // they stalled the waves of the groups of three to five.  So each of their waves also
    // touches its rows PFD sources ahead with one vector load (lane l < 4 NR: float
    // {0, 12, 24, 35}[l % 4] of right-hand side l / 4's row, which reaches every 64-byte
    // line of it), bringing them into L2; the loaded values are kept alive, unused.
    // Groups of three or more on one medium only: the group of two (which waited
    // little) and single beams gained nothing from it, and the separate-media and
    // complex groups would lose a wave per SIMD to the two address registers.
    constexpr bool PF = PATH == 0 && !SEP && !CPLX && sizeof(T) == 4 && NR >= 3;
    constexpr int PFD = 2;
    const gptr<T> pfrow = PF ? as_global<T>(grp[min(lane / 4, NR - 1)].Wp + min((lane % 4) * 12, K - 1)) : nullptr;
    T pf_keep = T(0);
    for (int s = s0; s < s1; ++s) {
        const T ddx = fmad(-T(si), g.dx, dxc), ddy = fmad(-T(sj), g.dy, dyc), ddz = fmad(-T(sk), g.dz, dzc);
        T r2 = fmad(ddx, ddx, fmad(ddy, ddy, ddz * ddz));
        T kfs[NT];
        if constexpr (PATH == 0) {
            // the self pair's table entry is 0 and its direction (0, 0, 0) whatever r is:
            // clamping r^2 keeps rsqrt finite without testing the indices
            r2 = max(r2, T(1e-30));
            // the next source's entry (the last source's again at the end: in bounds, unused).
            // Each entry is read once per launch and group: a nontemporal load, so the
            // table's stream does not push the W rows, which every brick reads, out of L2
            const int sn = (s + 1 < s1) ? s + 1 : s;
#pragma unroll
            for (int r = 0; r < NT; ++r) {
                kfs[r] = kf_next[r];
                kf_next[r] = __builtin_nontemporal_load(ktab[r] + (size_t)sn * g.npad + slot);
            }
        } else {
            const bool self = (cdi == si) && (cdj == sj) && (cdk == sk);
            const int ix = cdi - si, iy = cdj - sj, iz = cdk - sk;
            const bool use = valid && !self && (r2 < A.thr2);
            const T r = sqrt(self ? T(1) : r2);
            const T tau = use ? march_tau<T>(mut, s, ix, iy, iz, g.nx, nxy, r) : T(0);
            kfs[0] = use ? dev_exp<T>(-tau) / r2 : T(0);
            r2 = self ? T(1) : r2;
        }
        const T kf = kfs[0];
        T pfx = T(0);
        if constexpr (PF) pfx = pfrow[(size_t)((s + PFD < s1) ? s + PFD : s1 - 1) * K];
        const T rinv = dev_rsqrt<T>(r2);
        // the pair's Z values, formed once and shared by every right-hand side
        PairZ<T, L> G;
        pair_zpow<T, L>(G, ddx * rinv, ddy * rinv, ddz * rinv);
        // w holds right-hand side r of source s; the next one (r+1, or 0 of
        // s+1) is requested before r is applied, hiding the scalar-load latency
        const int snext = (s + 1 < s1) ? s + 1 : s;
        if constexpr (CPLX) {
            // the real and imaginary halves of one complex field share the pair:
            // t = kf e^{-i kim r} (p0 + i p1), both projections taken first.
            // One set of scalar registers holds W (see below): the imaginary
            // half is loaded after the real projection, the next source's real
            // half after the imaginary one, covered by the two axpys.
            const T p0 = pair_dot_z<T, L>(w, G);
#pragma unroll
            for (int a = 0; a < K; ++a) w[a] = Wsrc[1][(size_t)s * K + a];
            T sn, cs;
            if constexpr (PHASE_TAB) {
                const int d = dbase - si - ex * sj - exy * sk;
                cs = phase[2 * d];
                sn = phase[2 * d + 1];
            } else {
                dev_sincos<T>(kim * (r2 * rinv), &sn, &cs);
            }
            const T kc = kf * cs, ks = kf * sn;
            __builtin_amdgcn_sched_barrier(0);
            const T p1 = pair_dot_z<T, L>(w, G);
#pragma unroll
            for (int a = 0; a < K; ++a) w[a] = Wsrc[0][(size_t)snext * K + a];
            __builtin_amdgcn_sched_barrier(0);
            pair_axpy_z<T, L>(acc[0], fmad(kc, p0, ks * p1), G);
            pair_axpy_z<T, L>(acc[1], fmad(kc, p1, -(ks * p0)), G);
            __builtin_amdgcn_sched_barrier(0);
        } else if constexpr (sizeof(T) == 4 && K <= 36 && NR == 2) {
            // Groups of two: two sets of scalar registers hold W (2 x 36 floats
            // fit beside the loop's other scalars; with three right-hand sides
            // they do not, and the spills cost more than the latency saved).
            // The next right-hand side's row, or the next source's first, is
            // requested before this side's projection, so a projection and an
            // axpy cover each scalar load (74 and 151 VALU instructions, against
            // 53 and 70).  Scalar loads return out of order, so a wait is always
            // for all of them: this side's row is waited for first (lgkmcnt(0);
            // vmcnt and expcnt left at their maxima), then the next requested.
            T wn[K];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                __builtin_amdgcn_s_waitcnt(0xC07F);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int a = 0; a < K; ++a)
                    wn[a] = (r + 1 < NR) ? Wsrc[r + 1][(size_t)s * K + a] : Wsrc[0][(size_t)snext * K + a];
                __builtin_amdgcn_sched_barrier(0);
                const T t = kfs[SEP ? r : 0] * pair_dot_z<T, L>(w, G);
                pair_axpy_z<T, L>(acc[r], t, G);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int a = 0; a < K; ++a) w[a] = wn[a];
            }
        } else if constexpr (sizeof(T) == 4 && NR == 3 && K % 2 == 0 && K >= 16 && K <= 36) {
            // Groups of three: W rows in halves, three half-row sets of scalar
            // registers (w's two halves and c; two whole rows do not fit beside the
            // loop's other scalars).  Scalar loads return out of order, so the wait
            // before a side's projection is for every load in flight: the next
            // side's halves are requested only after that wait (the first term of
            // the projection reads the row), the first half at once and the second
            // into this side's first half once the projection has read it, so
            // they are covered by most of a projection and an axpy (2K - 3 and
            // 3K/2 FMAs, against K with one row).  The halves rotate through the
            // three sets with period three sides: a source starts again with its
            // first side's row in w.  The empty asm statements order each load
            // after the terms that free its registers, the scheduling barriers
            // keep the terms after it from moving ahead of the request.
            constexpr int H = K / 2;
            T c[H];
            T p, q, t;
            cptr<T> wp;
            // the pair values complete before the first projection, so the wait for
            // this source's first row (requested during the last side of the
            // previous source) comes after them
            __builtin_amdgcn_sched_barrier(0);
            // side 0: halves in w[0, H) and w[H, K); c free
            pair_dot_z_part<T, L, 0, 3>(p, q, w, G);
            wp = Wsrc[1] + (size_t)s * K;
            asm volatile("" : "+v"(p), "+v"(q), "+s"(wp));
#pragma unroll
            for (int a = 0; a < H; ++a) c[a] = wp[a];
            __builtin_amdgcn_sched_barrier(0);
            pair_dot_z_part<T, L, 3, H - 3>(p, q, w + 3, G);
            wp = Wsrc[1] + (size_t)s * K + H;
            asm volatile("" : "+v"(p), "+v"(q), "+s"(wp));
#pragma unroll
            for (int a = 0; a < H; ++a) w[a] = wp[a];
            __builtin_amdgcn_sched_barrier(0);
            pair_dot_z_part<T, L, H, H>(p, q, w + H, G);
            t = kfs[0] * (p + q);
            pair_axpy_z<T, L>(acc[0], t, G);
            __builtin_amdgcn_sched_barrier(0);
            // side 1: halves in c and w[0, H); w[H, K) free
            pair_dot_z_part<T, L, 0, 3>(p, q, c, G);
            wp = Wsrc[2] + (size_t)s * K;
            asm volatile("" : "+v"(p), "+v"(q), "+s"(wp));
#pragma unroll
            for (int a = 0; a < H; ++a) w[H + a] = wp[a];
            __builtin_amdgcn_sched_barrier(0);
            pair_dot_z_part<T, L, 3, H - 3>(p, q, c + 3, G);
            wp = Wsrc[2] + (size_t)s * K + H;
            asm volatile("" : "+v"(p), "+v"(q), "+s"(wp));
#pragma unroll
            for (int a = 0; a < H; ++a) c[a] = wp[a];
            __builtin_amdgcn_sched_barrier(0);
            pair_dot_z_part<T, L, H, H>(p, q, w, G);
            t = kfs[SEP ? 1 : 0] * (p + q);
            pair_axpy_z<T, L>(acc[1], t, G);
            __builtin_amdgcn_sched_barrier(0);
            // side 2: halves in w[H, K) and c; w[0, H) free
            pair_dot_z_part<T, L, 0, 3>(p, q, w + H, G);
            wp = Wsrc[0] + (size_t)snext * K;
            asm volatile("" : "+v"(p), "+v"(q), "+s"(wp));
#pragma unroll
            for (int a = 0; a < H; ++a) w[a] = wp[a];
            __builtin_amdgcn_sched_barrier(0);
            pair_dot_z_part<T, L, 3, H - 3>(p, q, w + H + 3, G);
            wp = Wsrc[0] + (size_t)snext * K + H;
            asm volatile("" : "+v"(p), "+v"(q), "+s"(wp));
#pragma unroll
            for (int a = 0; a < H; ++a) w[H + a] = wp[a];
            __builtin_amdgcn_sched_barrier(0);
            pair_dot_z_part<T, L, H, H>(p, q, c, G);
            t = kfs[SEP ? 2 : 0] * (p + q);
            pair_axpy_z<T, L>(acc[2], t, G);
            __builtin_amdgcn_sched_barrier(0);
        } else {
            // the pair values complete before the first projection: interleaved with
            // it, they pulled the wait for the row requested during the previous
            // source's last axpy up to 69 VALU instructions after the request (120 now)
            __builtin_amdgcn_sched_barrier(0);
            // One set of scalar registers holds W: the projection reads it, the
            // next right-hand side (or source) is then loaded into the same
            // registers, and the axpy, which needs no W, covers the load.  The
            // scheduling barriers keep the next projection from being hoisted
            // above the axpy (two live sets of W do not fit the SGPR file).
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const T t = kfs[SEP ? r : 0] * pair_dot_z<T, L>(w, G);
#pragma unroll
                for (int a = 0; a < K; ++a)
                    w[a] = (r + 1 < NR) ? Wsrc[r + 1][(size_t)s * K + a] : Wsrc[0][(size_t)snext * K + a];
                __builtin_amdgcn_sched_barrier(0);
                pair_axpy_z<T, L>(acc[r], t, G);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if constexpr (PF) pf_keep += pfx;  // at the source's end: the load is waited for there
        if (++si == g.nx) {
            si = 0;
            if (++sj == g.ny) {
                sj = 0;
                ++sk;
            }
        }
    }
    asm volatile("" ::"v"(pf_keep));
    if (valid) {
        // partial sums in the Z basis (k_combine turns their chunk sum back)
        const int v = di + g.nx * (dj + g.ny * dk);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            T *__restrict__ out = grp[r].part + ((size_t)chunk * g.N + v) * K;
#pragma unroll
            for (int a = 0; a < K; ++a) out[a] = acc[r][a];
        }
    }
}

// ---------------------------------------------------------------------------
// A propagation's input in the Z basis: W'_j = sum_a M_aj W_a for every
// solve of the launch (their SolveDev entries are consecutive), from W into Wp.
// grid = (ceil(N / 256), nsolve), block = 256: one thread per voxel and solve.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L>
__global__ __launch_bounds__(256) void k_to_zpow(int N, const SolveDev<T, TA> *__restrict__ sv)
{
    constexpr int K = (L + 1) * (L + 1);
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const SolveDev<T, TA> &s = sv[blockIdx.y];
    const T *__restrict__ w = s.W + (size_t)v * K;
    T *__restrict__ o = s.Wp + (size_t)v * K;
    double x[K], y[K];
#pragma unroll
    for (int a = 0; a < K; ++a) x[a] = double(w[a]);
    zpow_from_harmonic<double, L>(x, y);
#pragma unroll
    for (int j = 0; j < K; ++j) o[j] = T(y[j]);
}

// ---------------------------------------------------------------------------
// Per-medium table kf(d, s) = e^{-tau(d,s)}/r^2 (0 for d == s, r >= thresh and
// padding), layout [s][slot(d)].  The ray between two voxel centres is the
// same both ways, so one wave marches the pairs of a brick pair (bd, bs) with
// bs <= bd once (from s to d) and writes both entries: the direct one
// (row s, the wave's 64 destination slots: one coalesced row segment per
// source) and, off the diagonal, the mirrored one (row d, the source brick's
// slots), held in registers and written as one 256-byte run per lane after the
// brick's march (FP32; FP64 one x row of 4 sources at a time).  mu_t is read
// from LDS when the whole map fits.
// grid = ceil(ntiles / waves per block), block = KT threads.
// ---------------------------------------------------------------------------
template <typename T, bool LDS>
__global__ __launch_bounds__(1024) void k_ktab(GridDev<T> g, int ntiles, T thr2, const T *__restrict__ mut,
                                               T *__restrict__ ktab)
{
    extern __shared__ __align__(16) unsigned char ktab_lds[];
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    if constexpr (LDS) {
        T *sm = reinterpret_cast<T *>(ktab_lds);
        for (int i = tid; i < g.N; i += blockDim.x) sm[i] = mut[i];
        __syncthreads();
    }
    const int tile = blockIdx.x * (blockDim.x / kWave) + tid / kWave;
    if (tile >= ntiles) return;
    // tile -> (bd, bs), bs <= bd
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
    const bool mirror = bs != bd;
    // kf of source (si, sj, sk) of the brick for this lane's destination, written
    // to the direct entry (row s, this wave's 64 slots: one coalesced segment)
    auto entry = [&](int jj, int q) -> T {
        const int si = bxi * kBrick + jj, sj = byi * kBrick + (q & 3), sk = bzi * kBrick + (q >> 2);
        T kf = T(0);
        if (si < g.nx && sj < g.ny && sk < g.nz) {  // wave-uniform
            const int ix = cdi - si, iy = cdj - sj, iz = cdk - sk;
            const T ddx = T(ix) * g.dx, ddy = T(iy) * g.dy, ddz = T(iz) * g.dz;
            const T r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const bool self = (ix == 0) && (iy == 0) && (iz == 0);
            const bool use = dvalid && !self && (r2 < thr2);
            const T r = sqrt(self ? T(1) : r2);
            const int slin = si + g.nx * sj + nxy * sk;
            T tau = T(0);
            if (use) {
                if constexpr (LDS)
                    tau = march_tau<T>((const __attribute__((address_space(3))) T *)ktab_lds, slin, ix, iy, iz, g.nx,
                                       nxy, r);
                else
                    tau = march_tau<T>(as_global<T>(mut), slin, ix, iy, iz, g.nx, nxy, r);
            }
            kf = use ? dev_exp<T>(-tau) / r2 : T(0);
            ktab[(size_t)slin * g.npad + dslot] = kf;
        }
        return kf;
    };
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int per = sizeof(V) / sizeof(T);
    V *row = reinterpret_cast<V *>(ktab + (size_t)dlin * g.npad + bs * kWave);
    if constexpr (sizeof(T) == 4) {
        // FP32: the brick's 64 mirrored entries stay in registers and go out as one
        // 256-byte run per lane after the march (written piecewise, the rows' lines
        // reached HBM partly filled: 2.1x the table's bytes)
        T m[kWave];
#pragma unroll
        for (int q = 0; q < 16; ++q)  // the source brick's x rows, 4 voxels each
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) m[q * 4 + jj] = entry(jj, q);
        if (mirror && dvalid) {
#pragma unroll
            for (int u = 0; u < kWave / per; ++u) {
                V x;
                T *xs = reinterpret_cast<T *>(&x);
#pragma unroll
                for (int e = 0; e < per; ++e) xs[e] = m[u * per + e];
                row[u] = x;
            }
        }
    } else {
        // FP64: one x row of 4 sources at a time (the whole brick would spill)
        for (int q = 0; q < 16; ++q) {
            T m[4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) m[jj] = entry(jj, q);
            if (mirror && dvalid) {
#pragma unroll
                for (int u = 0; u < 4 / per; ++u) {
                    V x;
                    T *xs = reinterpret_cast<T *>(&x);
#pragma unroll
                    for (int e = 0; e < per; ++e) xs[e] = m[u * per + e];
                    row[q * 4 / per + u] = x;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Combine: chunk partials (fixed order) + self block, times dV, then
//   operator launch (A.op = 1): aw_out <- A W (TA, [v][K])
//   per-solve mode 0 (term):  W <- g^l mu_s A W, dE = exitance(A W), bsum = stop-rule sums of W
//   per-solve mode 2 (final): dE only
// grid = (ceil(N/64), nsolve), block = 64.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L>
__device__ __forceinline__ void combine_out(const CombArgs<T, TA> &A, const SolveDev<T, TA> &S, int v,
                                            const TA (&acc)[(L + 1) * (L + 1)], const TA *__restrict__ wsum);


template <typename T, typename TA, int L>
__global__ __launch_bounds__(kWave) void k_combine(CombArgs<T, TA> A, const SolveDev<T, TA> *__restrict__ sv)
{
    constexpr int K = (L + 1) * (L + 1);
    const GridDev<T> &g = A.g;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const SolveDev<T, TA> S = sv[blockIdx.y];
    if (v >= g.N) return;
    TA acc[K], w[K];
#pragma unroll
    for (int a = 0; a < K; ++a) {
        acc[a] = TA(0);
        w[a] = TA(S.W[(size_t)v * K + a]) / A.nsc[a];
    }
    for (int c = 0; c < A.nchunk; ++c) {
        const T *__restrict__ p = S.part + ((size_t)c * g.N + v) * K;
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] += TA(p[a]);
    }
    zpow_to_harmonic<TA, L>(acc);
#pragma unroll
    for (int a = 0; a < K; ++a) acc[a] *= A.nsc[a];
    // self block: sum_k coef_k (M_k w), coef_k = (-mu)^{k-1}/k!
    if constexpr (sizeof(T) == 4) {
        // in FP32, on the entries the voxel's reflection symmetries leave
        T wt[K], st[K];
#pragma unroll
        for (int a = 0; a < K; ++a) {
            wt[a] = T(w[a]);
            st[a] = T(0);
        }
        const T mu = S.mut[v];
        T coef = T(1);
        for (int k = 0; k < kNMom; ++k) {
            if (k > 0) coef = coef * (-mu) / T(k + 1);
            self_apply<K>(st, coef, A.momT + (size_t)k * K * K, wt, std::make_integer_sequence<int, K>{});
        }
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] += A.self_scale * TA(st[a]);
    } else {
        const TA mu = TA(S.mut[v]);
        TA coef = TA(1);
        for (int k = 0; k < kNMom; ++k) {
            if (k > 0) coef = coef * (-mu) / TA(k + 1);
            const TA *__restrict__ M = A.mom + (size_t)k * K * K;
#pragma unroll
            for (int a = 0; a < K; ++a) {
                TA s = TA(0);
#pragma unroll
                for (int b = 0; b < K; ++b) s += M[a * K + b] * w[b];
                acc[a] += A.self_scale * coef * s;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < K; ++a) acc[a] *= A.dV;
    if (A.op) {
#pragma unroll
        for (int a = 0; a < K; ++a) A.aw_out[(size_t)v * K + a] = acc[a];
        return;
    }
    combine_out<T, TA, L>(A, S, v, acc, A.wsum);
}

// The per-solve outputs of one voxel's A W (acc, unscaled, TA): stored or
// accumulated copies, the exitance and reflectance increments, and in a term
// the scattered next term with its stop-rule sums (weights wsum).
template <typename T, typename TA, int L>
__device__ __forceinline__ void combine_out(const CombArgs<T, TA> &A, const SolveDev<T, TA> &S, int v,
                                            const TA (&acc)[(L + 1) * (L + 1)], const TA *__restrict__ wsum)
{
    constexpr int K = (L + 1) * (L + 1);
    const GridDev<T> &g = A.g;
    const int zk = v / (g.nx * g.ny);
    const int pix = v - zk * g.nx * g.ny;
    const bool bottom = (zk == g.nz - 1);
    if (S.aw_store) {
#pragma unroll
        for (int a = 0; a < K; ++a) S.aw_store[(size_t)v * K + a] = acc[a];
    }
    if (S.phi && S.mode != 3) {
#pragma unroll
        for (int a = 0; a < K; ++a) S.phi[(size_t)v * K + a] += acc[a];
    }
    if (bottom && S.mode != 3) {
        TA e = TA(0);
#pragma unroll
        for (int l = 0; l <= L; ++l) e += A.coefE[l] * acc[l * l + l];
        S.dE[pix] = e;
    }
    if (zk == 0 && S.mode != 3) {
        TA r = TA(0);
#pragma unroll
        for (int l = 0; l <= L; ++l) r += A.coefR[l] * acc[l * l + l];
        S.dR[pix] = r;
    }
    if (S.mode == 0 || S.mode == 3) {
        const TA ms = TA(S.mus[v]);
        TA gl = TA(1), bs = TA(0);
#pragma unroll
        for (int l = 0; l <= L; ++l) {
#pragma unroll
            for (int m = -l; m <= l; ++m) {
                const int a = l * l + l + m;
                const TA wn = gl * ms * acc[a];
                const T wt = T(wn * A.nsc[a]);
                S.W[(size_t)v * K + a] = wt;
                bs += wsum[a] * (TA(wt) / A.nsc[a]);
            }
            gl *= TA(S.g);
        }
        if (bottom && S.mode == 0) S.bsum[pix] = bs;
    }
}

// ---------------------------------------------------------------------------
// Combine of frequency-domain solves: halves (sv[2y], sv[2y+1]) = (real,
// imaginary) of one complex field.  The self block has the complex series
// coefficients c_k = (-mu)^{k-1}/k!, mu = mu_t + i kim:
//   S (w0 + i w1) = sum_k (Re c_k M_k w0 - Im c_k M_k w1) + i (Re c_k M_k w1 + Im c_k M_k w0);
// everything else acts on each half alone (D, the exitance and reflectance
// functionals and the chunk sums are real).  Both halves' W are read before
// either is overwritten by the scattered next term.
// grid = (ceil(N/64), ncomplex), block = 64.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L>
__global__ __launch_bounds__(kWave) void k_combine_c(CombArgs<T, TA> A, const SolveDev<T, TA> *__restrict__ sv)
{
    constexpr int K = (L + 1) * (L + 1);
    const GridDev<T> &g = A.g;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const SolveDev<T, TA> S0 = sv[2 * blockIdx.y], S1 = sv[2 * blockIdx.y + 1];
    if (v >= g.N) return;
    T w0[K], w1[K], st0[K], st1[K];
#pragma unroll
    for (int a = 0; a < K; ++a) {
        w0[a] = T(TA(S0.W[(size_t)v * K + a]) / A.nsc[a]);
        w1[a] = T(TA(S1.W[(size_t)v * K + a]) / A.nsc[a]);
        st0[a] = st1[a] = T(0);
    }
    const T mr = S0.mut[v], mi = S0.kim;
    T cr = T(1), ci = T(0);
    for (int k = 0; k < kNMom; ++k) {
        if (k > 0) {
            const T nr = -(cr * mr - ci * mi) / T(k + 1);
            const T ni = -(cr * mi + ci * mr) / T(k + 1);
            cr = nr;
            ci = ni;
        }
        if constexpr (sizeof(T) == 4) {
            const T *__restrict__ M = A.momT + (size_t)k * K * K;
            self_apply<K>(st0, cr, M, w0, std::make_integer_sequence<int, K>{});
            self_apply<K>(st0, -ci, M, w1, std::make_integer_sequence<int, K>{});
            self_apply<K>(st1, cr, M, w1, std::make_integer_sequence<int, K>{});
            self_apply<K>(st1, ci, M, w0, std::make_integer_sequence<int, K>{});
        } else {
            const TA *__restrict__ M = A.mom + (size_t)k * K * K;
#pragma unroll
            for (int a = 0; a < K; ++a) {
                T s0 = T(0), s1 = T(0);
#pragma unroll
                for (int b = 0; b < K; ++b) {
                    s0 += T(M[a * K + b]) * w0[b];
                    s1 += T(M[a * K + b]) * w1[b];
                }
                st0[a] += cr * s0 - ci * s1;
                st1[a] += cr * s1 + ci * s0;
            }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const SolveDev<T, TA> &S = h ? S1 : S0;
        TA acc[K];
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] = TA(0);
        for (int c = 0; c < A.nchunk; ++c) {
            const T *__restrict__ p = S.part + ((size_t)c * g.N + v) * K;
#pragma unroll
            for (int a = 0; a < K; ++a) acc[a] += TA(p[a]);
        }
        // the steps of k_combine, in its order
        zpow_to_harmonic<TA, L>(acc);
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] *= A.nsc[a];
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] += A.self_scale * TA(h ? st1[a] : st0[a]);
#pragma unroll
        for (int a = 0; a < K; ++a) acc[a] *= A.dV;
        combine_out<T, TA, L>(A, S, v, acc, h ? A.wsumI : A.wsum);
    }
}

// ---------------------------------------------------------------------------
// Per-solve reduction of the bottom face in a fixed order (deterministic,
// independent of the batch).  Per-solve mode:
//   0 (term):  E += dE, R += dR; red = (sum dE, sum bsum)
//   1 (init):  red = (sum E, sum bsum)          (after the source)
//   2 (final): E += dE, R += dR; red = (sum dE, 0)
// grid = nsolve, block = 256.
// ---------------------------------------------------------------------------
template <typename T, typename TA>
__global__ __launch_bounds__(256) void k_bottom(int P, const SolveDev<T, TA> *__restrict__ sv)
{
    __shared__ TA s1[256], s2[256];
    const SolveDev<T, TA> S = sv[blockIdx.x];
    TA a = TA(0), b = TA(0);
    for (int p = threadIdx.x; p < P; p += 256) {
        if (S.mode == 1) {
            a += S.E[p];
        } else {
            const TA d = S.dE[p];
            S.E[p] += d;
            S.R[p] += S.dR[p];
            a += d;
        }
        if (S.mode != 2) b += S.bsum[p];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            s1[threadIdx.x] += s1[threadIdx.x + h];
            s2[threadIdx.x] += s2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S.red[0] = s1[0];
        S.red[1] = s2[0];
    }
}

// ---------------------------------------------------------------------------
// First-scatter source s0 and ballistic image (column products, FP64):
//   s0_l0(i,j,k) = fact g^l sqrt((2l+1)/4pi) (mu_s/mu_t) T(k) (1 - e^{-mu_t dz}) / dV
// (GPU package/src/dist.cu:66-101; CPU package/src/Dist.c:1077-1128),
// ballistic = fact T(nz) (GPU package/src/dist.cu:129-141).  W must be zero.
// grid = (ceil(P/64), nsolve), block = 64.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L>
__global__ __launch_bounds__(kWave) void k_source(GridDev<T> g, double dxd, double dyd, double dzd,
                                                 double xmin, double ymin, Scales sc,
                                                 const BeamDev *__restrict__ beams,
                                                 const SolveDev<T, TA> *__restrict__ sv)
{
    constexpr int K = (L + 1) * (L + 1);
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = g.nx * g.ny;
    if (pix >= P) return;
    const BeamDev B = beams[blockIdx.y];
    const SolveDev<T, TA> S = sv[blockIdx.y];
    const int i = pix % g.nx, j = pix / g.nx;
    const double xc = (double)i * dxd + xmin + dxd / 2.0;
    const double yc = (double)j * dyd + ymin + dyd / 2.0;
    const double rad = sqrt((xc - B.pos_x) * (xc - B.pos_x) + (yc - B.pos_y) * (yc - B.pos_y));
    const bool in = rad <= B.half_diam;
    const double dV = dxd * dyd * dzd;
    const double gg = (double)S.g;
    double Tc = 1.0;
    double bs = 0.0;
    if (in) {
        for (int k = 0; k < g.nz; ++k) {
            const int v = pix + P * k;
            const double mt = B.mua[v] + B.mus[v];
            const double att = exp(-mt * dzd);
            const double base = B.fact * (B.mus[v] / mt) * Tc * (1.0 - att) / dV;
            double gl = 1.0;
#pragma unroll
            for (int l = 0; l <= L; ++l) {
                const double s = gl * sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) * base;
                const T st = T(s * sc.n[l * l + l]);
                S.W[(size_t)v * K + l * l + l] = st;
                if (k == g.nz - 1) bs += (double)st / sc.n[l * l + l];
                gl *= gg;
            }
            Tc *= att;
        }
    }
    S.E[pix] = TA(in ? B.fact * Tc : 0.0);
    S.bsum[pix] = TA(bs);
    S.dE[pix] = TA(0);
    S.R[pix] = TA(0);
    S.dR[pix] = TA(0);
}

// ---------------------------------------------------------------------------
// Complex first-scatter source and ballistic image of frequency-domain solves:
// k_source with mu_t -> mu_t + i kim (the column attenuation e^{-mu_t dz}
// picks up the phase e^{-i kim dz}, and mu_s/mu_t is a complex ratio); real
// part to sv[2y], imaginary part to sv[2y+1].  W must be zero.
// grid = (ceil(P/64), ncomplex), block = 64.
// ---------------------------------------------------------------------------
template <typename T, typename TA, int L>
__global__ __launch_bounds__(kWave) void k_source_c(GridDev<T> g, double dxd, double dyd, double dzd,
                                                   double xmin, double ymin, Scales sc,
                                                   const BeamDev *__restrict__ beams,
                                                   const SolveDev<T, TA> *__restrict__ sv)
{
    constexpr int K = (L + 1) * (L + 1);
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = g.nx * g.ny;
    if (pix >= P) return;
    const BeamDev B = beams[blockIdx.y];
    const SolveDev<T, TA> S0 = sv[2 * blockIdx.y], S1 = sv[2 * blockIdx.y + 1];
    const int i = pix % g.nx, j = pix / g.nx;
    const double xc = (double)i * dxd + xmin + dxd / 2.0;
    const double yc = (double)j * dyd + ymin + dyd / 2.0;
    const double rad = sqrt((xc - B.pos_x) * (xc - B.pos_x) + (yc - B.pos_y) * (yc - B.pos_y));
    const bool in = rad <= B.half_diam;
    const double dV = dxd * dyd * dzd;
    const double gg = (double)S0.g;
    const double ki = B.kim;
    const double ph_c = cos(ki * dzd), ph_s = sin(ki * dzd);
    double Tr = 1.0, Ti = 0.0;
    double bs = 0.0;
    if (in) {
        for (int k = 0; k < g.nz; ++k) {
            const int v = pix + P * k;
            const double mt = B.mua[v] + B.mus[v];
            const double e = exp(-mt * dzd);
            const double ar = e * ph_c, ai = -e * ph_s;               // e^{-(mt + i ki) dz}
            // mu_s / (mt + i ki) = (mu_s/mt) / (1 + i t), t = ki/mt; the products in
            // k_source's order, so that ki = 0 gives its bits
            const double t = ki / mt, q0 = B.mus[v] / mt, d = 1.0 + t * t;
            const double fr = B.fact * (q0 / d), fi = B.fact * (-q0 * t / d);
            const double ftr = fr * Tr - fi * Ti, fti = fr * Ti + fi * Tr;
            const double omr = 1.0 - ar, omi = -ai;
            const double br = (ftr * omr - fti * omi) / dV;
            const double bi = (ftr * omi + fti * omr) / dV;
            double gl = 1.0;
#pragma unroll
            for (int l = 0; l <= L; ++l) {
                const double f = gl * sqrt((2.0 * l + 1.0) / (4.0 * M_PI));
                const int a = l * l + l;
                const T sr = T(f * br * sc.n[a]);
                S0.W[(size_t)v * K + a] = sr;
                S1.W[(size_t)v * K + a] = T(f * bi * sc.n[a]);
                if (k == g.nz - 1) bs += (double)sr / sc.n[a];
                gl *= gg;
            }
            const double nTr = Tr * ar - Ti * ai, nTi = Tr * ai + Ti * ar;
            Tr = nTr;
            Ti = nTi;
        }
    }
    S0.E[pix] = TA(in ? B.fact * Tr : 0.0);
    S1.E[pix] = TA(in ? B.fact * Ti : 0.0);
    // the stop-rule weights of the imaginary half vanish on m = 0
    S0.bsum[pix] = TA(bs);
    S1.bsum[pix] = TA(0);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const SolveDev<T, TA> &S = h ? S1 : S0;
        S.dE[pix] = TA(0);
        S.R[pix] = TA(0);
        S.dR[pix] = TA(0);
    }
}

// ---------------------------------------------------------------------------
// Self-block moments M_k[a][b] = sum_w R_a(w) R_b(w) sin(theta_w) beta_w^k,
// k = 1..kNMom, on the reference's 200x200 grid.  grid = K*K, block = 256.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_moments(int K, const double *__restrict__ R /*[NW][K]*/,
                                                 const double *__restrict__ sn,
                                                 const double *__restrict__ beta,
                                                 double *__restrict__ M)
{
    __shared__ double red[kNMom][256];
    const int a = blockIdx.x / K, b = blockIdx.x % K;
    double acc[kNMom];
#pragma unroll
    for (int k = 0; k < kNMom; ++k) acc[k] = 0.0;
    for (int w = threadIdx.x; w < kNOmega; w += 256) {
        const double f = R[(size_t)w * K + a] * R[(size_t)w * K + b] * sn[w];
        double bk = beta[w];
#pragma unroll
        for (int k = 0; k < kNMom; ++k) {
            acc[k] += f * bk;
            bk *= beta[w];
        }
    }
#pragma unroll
    for (int k = 0; k < kNMom; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
#pragma unroll
            for (int k = 0; k < kNMom; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x < kNMom) M[((size_t)threadIdx.x * K + a) * K + b] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// Layout conversions for the operator entry points and small operators
// ---------------------------------------------------------------------------
// coefficient-major double [K][N]  ->  voxel-major T [N][K]
template <typename T>
__global__ void k_to_voxel_major(int K, int N, Scales sc, const double *__restrict__ in, T *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * N) return;
    const int v = (int)(i / K), a = (int)(i % K);
    out[i] = T(in[(size_t)a * N + v] * sc.n[a]);
}

template <typename TA>
__global__ void k_to_coef_major(int K, int N, Scales sc, const TA *__restrict__ in, double *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * N) return;
    const int a = (int)(i / N), v = (int)(i % N);
    out[i] = (double)in[(size_t)v * K + a] / sc.n[a];
}

// The reference's fields are complex coefficients of the complex orthonormal
// harmonics (Condon-Shortley phase, Y_l,-m = (-1)^m conj Y_lm), voxel-major
// W[v K + cnt], cnt = l^2 + l + m (GPU package/src/defines.h:59).  For a real
// radiance c_l,-m = (-1)^m conj c_lm, and in the real basis R_lm = sqrt2 Re Y_lm,
// R_l,-m = sqrt2 Im Y_lm (m > 0) its coefficients are
//   r_l0 = c_l0,  r_lm = sqrt2 Re c_lm,  r_l,-m = -sqrt2 Im c_lm,
// and back  c_lm = (r_lm - i r_l,-m)/sqrt2,  c_l,-m = (-1)^m (r_lm + i r_l,-m)/sqrt2.
// One thread per (voxel, coefficient) of the output.
template <typename T>
__global__ void k_cplx_to_real(int L, int N, double s, Scales sc, const double *__restrict__ c,
                               T *__restrict__ out)
{
    const int K = (L + 1) * (L + 1);
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * N) return;
    const int v = (int)(i / K), a = (int)(i % K);
    int l = 0;
    while ((l + 1) * (l + 1) <= a) ++l;
    const int m = a - l * l - l;
    const double *cv = c + (size_t)v * K * 2;
    double r;
    if (m == 0)
        r = cv[2 * a];
    else if (m > 0)
        r = M_SQRT2 * cv[2 * a];
    else
        r = -M_SQRT2 * cv[2 * (l * l + l - m) + 1];  // Im c_l,|m|
    out[i] = T(r * s * sc.n[a]);
}

template <typename TA>
__global__ void k_real_to_cplx(int L, int N, double s, const TA *__restrict__ w, double *__restrict__ c)
{
    const int K = (L + 1) * (L + 1);
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * N) return;
    const int v = (int)(i / K), a = (int)(i % K);
    int l = 0;
    while ((l + 1) * (l + 1) <= a) ++l;
    const int m = a - l * l - l, am = m < 0 ? -m : m;
    const TA *wv = w + (size_t)v * K;
    double re, im;
    if (m == 0) {
        re = (double)wv[a];
        im = 0.0;
    } else {
        const double rc = (double)wv[l * l + l + am], rs = (double)wv[l * l + l - am];
        const double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;  // (-1)^m for the negative order
        re = sg * rc * M_SQRT1_2;
        im = (m > 0 ? -rs : sg * rs) * M_SQRT1_2;
    }
    c[2 * i] = re * s;
    c[2 * i + 1] = im * s;
}

// out[i] = float(in[i]) (round to nearest, as the host's conversion)
__global__ void k_narrow(size_t n, const double *__restrict__ in, float *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// D on a coefficient-major double field: W_a(v) *= g^l mu_s(v)
__global__ void k_scatter_op(int L, int N, double g, const double *__restrict__ mus,
                             double *__restrict__ W)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int K = (L + 1) * (L + 1);
    if (i >= (size_t)K * N) return;
    const int a = (int)(i / N), v = (int)(i % N);
    int l = 0;
    while ((l + 1) * (l + 1) <= a) ++l;
    W[i] *= pow(g, (double)l) * mus[v];
}

// exitance of a coefficient-major double field (ballistic excluded)
__global__ void k_exitance_op(int L, int nx, int ny, int nz, const double *__restrict__ coefE,
                              const double *__restrict__ w, double *__restrict__ img)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = nx * ny, N = P * nz;
    if (pix >= P) return;
    const int v = pix + P * (nz - 1);
    double e = 0.0;
    for (int l = 0; l <= L; ++l) e += coefE[l] * w[(size_t)(l * l + l) * N + v];
    img[pix] = e;
}

// tau for a list of pairs, marched in FP32 as the kernels do
__global__ void k_tau_list(GridDev<float> g, const float *__restrict__ mut, int np,
                           const int *__restrict__ dst, const int *__restrict__ src,
                           double *__restrict__ tau)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int d = dst[i], s = src[i];
    const int nxy = g.nx * g.ny;
    const int di = d % g.nx, dj = (d / g.nx) % g.ny, dk = d / nxy;
    const int si = s % g.nx, sj = (s / g.nx) % g.ny, sk = s / nxy;
    const int ix = di - si, iy = dj - sj, iz = dk - sk;
    const float ddx = float(ix) * g.dx, ddy = float(iy) * g.dy, ddz = float(iz) * g.dz;
    const float r = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
    tau[i] = (d == s) ? 0.0 : (double)march_tau<float>(as_global<float>(mut), s, ix, iy, iz, g.nx, nxy, r);
}

// ---------------------------------------------------------------------------
// Fisher Gram G = J^T diag(w) J (J: P x n, row-major) on MFMA:
// v_mfma_f32_16x16x4f32, one wave per 16x16 tile of G, four detectors per
// step.  Operand layout of the instruction (wave64): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; the 4 results of lane l are
// D[4 (l / 16) + r][l % 16].  grid = (ceil(n/16), ceil(n/16)), block = 64.
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kWave) void k_gram(int P, int n, const float *__restrict__ J,
                                               const float *__restrict__ w, float *__restrict__ G)
{
    const int lane = threadIdx.x;
    const int r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int ia = r0 + (lane & 15), kk = lane >> 4;
    const int jb = c0 + (lane & 15);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int p0 = 0; p0 < P; p0 += 4) {
        const int pp = p0 + kk;
        const bool okp = pp < P;
        const float a = (okp && ia < n) ? J[(size_t)pp * n + ia] : 0.f;
        const float b = (okp && jb < n) ? w[pp] * J[(size_t)pp * n + jb] : 0.f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = r0 + 4 * (lane >> 4) + r, col = c0 + (lane & 15);
        if (row < n && col < n) G[(size_t)row * n + col] = acc[r];
    }
}

hipError_t launch_gram(int P, int n, const float *J, const float *w, float *G, hipStream_t st)
{
    dim3 grid((n + 15) / 16, (n + 15) / 16), block(kWave);
    hipLaunchKernelGGL(k_gram, grid, block, 0, st, P, n, J, w, G);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host-side launchers (dispatch over L)
// ---------------------------------------------------------------------------
#define RTE_DISPATCH_L(L, ...)                          \
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

template <typename T>
hipError_t launch_fill(T *p, size_t n, T v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_to_zpow(int L, int N, const SolveDev<T, TA> *sv, int nsolve, hipStream_t st)
{
    if (nsolve < 1 || nsolve > 65535) return hipErrorInvalidValue;
    const dim3 grid((N + 255) / 256, nsolve), block(256);
    RTE_DISPATCH_L(L, hipLaunchKernelGGL((k_to_zpow<T, TA, LL>), grid, block, 0, st, N, sv));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_prop(int L, int path, int nr, const PropArgs<T> &A, const SolveDev<T, TA> *sv, int ngroups,
                       hipStream_t st)
{
    dim3 grid((A.g.nblk + kPropWaves - 1) / kPropWaves, A.nchunk, ngroups), block(kWave * kPropWaves);
    // several right-hand sides per group only on the mixed-precision cached path
    constexpr bool multi = sizeof(T) == 4 && sizeof(TA) == 8;
    if (nr < 1 || nr > kMaxNR || (nr > 1 && (!multi || path != 0))) return hipErrorInvalidValue;
    hipError_t e = launch_to_zpow<T, TA>(L, A.g.N, sv, ngroups * nr, st);
    if (e != hipSuccess) return e;
    RTE_DISPATCH_L(L, {
        if (path == 1) {
            hipLaunchKernelGGL((k_prop<T, TA, LL, 1, 1>), grid, block, 0, st, A, sv);
        } else if constexpr (multi) {
            switch (nr) {
            case 1: hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 1>), grid, block, 0, st, A, sv); break;
            case 2: hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 2>), grid, block, 0, st, A, sv); break;
            case 3: hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 3>), grid, block, 0, st, A, sv); break;
            case 4: hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 4>), grid, block, 0, st, A, sv); break;
            default: hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 5>), grid, block, 0, st, A, sv); break;
            }
        } else {
            hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 1>), grid, block, 0, st, A, sv);
        }
    });
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_prop_sep(int L, int nr, const PropArgs<T> &A, const SolveDev<T, TA> *sv, int ngroups,
                           hipStream_t st)
{
    dim3 grid((A.g.nblk + kPropWaves - 1) / kPropWaves, A.nchunk, ngroups), block(kWave * kPropWaves);
    if constexpr (sizeof(T) == 4 && sizeof(TA) == 8) {
        if (nr != 2 && nr != 3) return hipErrorInvalidValue;
        hipError_t e = launch_to_zpow<T, TA>(L, A.g.N, sv, ngroups * nr, st);
        if (e != hipSuccess) return e;
        RTE_DISPATCH_L(L, {
            if (nr == 2)
                hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 2, false, true>), grid, block, 0, st, A, sv);
            else
                hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 3, false, true>), grid, block, 0, st, A, sv);
        });
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

template <typename T, typename TA>
hipError_t launch_prop_c(int L, int path, const PropArgs<T> &A, const SolveDev<T, TA> *sv, int ncomplex,
                         hipStream_t st)
{
    dim3 grid((A.g.nblk + kPropWaves - 1) / kPropWaves, A.nchunk, ncomplex), block(kWave * kPropWaves);
    hipError_t e = launch_to_zpow<T, TA>(L, A.g.N, sv, 2 * ncomplex, st);
    if (e != hipSuccess) return e;
    RTE_DISPATCH_L(L, {
        if (path == 1)
            hipLaunchKernelGGL((k_prop<T, TA, LL, 1, 2, true>), grid, block, 0, st, A, sv);
        else
            hipLaunchKernelGGL((k_prop<T, TA, LL, 0, 2, true>), grid, block, 0, st, A, sv);
    });
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ktab(const GridDev<T> &g, T thr2, const T *mut, T *ktab, hipStream_t st)
{
    static int lds_max = -1;
    if (lds_max < 0) {
        int dev = 0, v = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) return e;
        lds_max = v;
    }
    const int ntiles = g.nblk * (g.nblk + 1) / 2;
    const size_t lds = (size_t)g.N * sizeof(T);
    if (lds <= (size_t)lds_max) {
        constexpr int threads = 1024;
        if (lds > 65536) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ktab<T, true>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        const int blocks = (ntiles + threads / kWave - 1) / (threads / kWave);
        hipLaunchKernelGGL((k_ktab<T, true>), dim3(blocks), dim3(threads), lds, st, g, ntiles, thr2, mut, ktab);
    } else {
        constexpr int threads = 256;
        const int blocks = (ntiles + threads / kWave - 1) / (threads / kWave);
        hipLaunchKernelGGL((k_ktab<T, false>), dim3(blocks), dim3(threads), 0, st, g, ntiles, thr2, mut, ktab);
    }
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_combine(int L, const CombArgs<T, TA> &A, const SolveDev<T, TA> *sv, int nsolve,
                          hipStream_t st)
{
    dim3 grid((A.g.N + kWave - 1) / kWave, nsolve), block(kWave);
    RTE_DISPATCH_L(L, hipLaunchKernelGGL((k_combine<T, TA, LL>), grid, block, 0, st, A, sv));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_combine_c(int L, const CombArgs<T, TA> &A, const SolveDev<T, TA> *sv, int ncomplex,
                            hipStream_t st)
{
    dim3 grid((A.g.N + kWave - 1) / kWave, ncomplex), block(kWave);
    RTE_DISPATCH_L(L, hipLaunchKernelGGL((k_combine_c<T, TA, LL>), grid, block, 0, st, A, sv));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_bottom(int P, const SolveDev<T, TA> *sv, int nsolve, hipStream_t st)
{
    hipLaunchKernelGGL((k_bottom<T, TA>), dim3(nsolve), dim3(256), 0, st, P, sv);
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_source(int L, const GridDev<T> &g, double dx, double dy, double dz, double xmin,
                         double ymin, const Scales &sc, const BeamDev *beams, const SolveDev<T, TA> *sv,
                         int nsolve, hipStream_t st)
{
    dim3 grid((g.nx * g.ny + kWave - 1) / kWave, nsolve), block(kWave);
    RTE_DISPATCH_L(L, hipLaunchKernelGGL((k_source<T, TA, LL>), grid, block, 0, st, g, dx, dy, dz,
                                         xmin, ymin, sc, beams, sv));
    return hipGetLastError();
}

template <typename T, typename TA>
hipError_t launch_source_c(int L, const GridDev<T> &g, double dx, double dy, double dz, double xmin,
                           double ymin, const Scales &sc, const BeamDev *beams, const SolveDev<T, TA> *sv,
                           int ncomplex, hipStream_t st)
{
    dim3 grid((g.nx * g.ny + kWave - 1) / kWave, ncomplex), block(kWave);
    RTE_DISPATCH_L(L, hipLaunchKernelGGL((k_source_c<T, TA, LL>), grid, block, 0, st, g, dx, dy, dz,
                                         xmin, ymin, sc, beams, sv));
    return hipGetLastError();
}

hipError_t launch_moments(int K, const double *R, const double *sn, const double *beta, double *M,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_moments, dim3(K * K), dim3(256), 0, st, K, R, sn, beta, M);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_to_voxel_major(int K, int N, const Scales &sc, const double *in, T *out, hipStream_t st)
{
    size_t n = (size_t)K * N;
    hipLaunchKernelGGL(k_to_voxel_major<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, K, N,
                       sc, in, out);
    return hipGetLastError();
}

template <typename TA>
hipError_t launch_to_coef_major(int K, int N, const Scales &sc, const TA *in, double *out, hipStream_t st)
{
    size_t n = (size_t)K * N;
    hipLaunchKernelGGL(k_to_coef_major<TA>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, K, N,
                       sc, in, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cplx_to_real(int L, int N, double s, const Scales &sc, const double *c, T *out, hipStream_t st)
{
    const size_t n = (size_t)(L + 1) * (L + 1) * N;
    hipLaunchKernelGGL(k_cplx_to_real<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, N, s, sc, c, out);
    return hipGetLastError();
}

template <typename TA>
hipError_t launch_real_to_cplx(int L, int N, double s, const TA *w, double *c, hipStream_t st)
{
    const size_t n = (size_t)(L + 1) * (L + 1) * N;
    hipLaunchKernelGGL(k_real_to_cplx<TA>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, N, s, w, c);
    return hipGetLastError();
}

hipError_t launch_narrow(size_t n, const double *in, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_narrow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, in, out);
    return hipGetLastError();
}

hipError_t launch_scatter_op(int L, int N, double g, const double *mus, double *W, hipStream_t st)
{
    size_t n = (size_t)(L + 1) * (L + 1) * N;
    hipLaunchKernelGGL(k_scatter_op, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, N, g, mus,
                       W);
    return hipGetLastError();
}

hipError_t launch_exitance_op(int L, int nx, int ny, int nz, const double *coefE, const double *w,
                              double *img, hipStream_t st)
{
    int P = nx * ny;
    hipLaunchKernelGGL(k_exitance_op, dim3((P + 255) / 256), dim3(256), 0, st, L, nx, ny, nz, coefE, w,
                       img);
    return hipGetLastError();
}

hipError_t launch_tau_list(const GridDev<float> &g, const float *mut, int np, const int *dst,
                           const int *src, double *tau, hipStream_t st)
{
    hipLaunchKernelGGL(k_tau_list, dim3((np + 255) / 256), dim3(256), 0, st, g, mut, np, dst, src, tau);
    return hipGetLastError();
}

// explicit instantiations used by the host library
template hipError_t launch_prop_sep<float, double>(int, int, const PropArgs<float> &, const SolveDev<float, double> *,
                                                   int, hipStream_t);
template hipError_t launch_prop_sep<double, double>(int, int, const PropArgs<double> &,
                                                    const SolveDev<double, double> *, int, hipStream_t);
template hipError_t launch_prop_sep<float, float>(int, int, const PropArgs<float> &, const SolveDev<float, float> *,
                                                  int, hipStream_t);
template hipError_t launch_prop_c<float, double>(int, int, const PropArgs<float> &, const SolveDev<float, double> *, int,
                                                     hipStream_t);
template hipError_t launch_combine_c<float, double>(int, const CombArgs<float, double> &, const SolveDev<float, double> *, int,
                                                        hipStream_t);
template hipError_t launch_source_c<float, double>(int, const GridDev<float> &, double, double, double, double, double,
                                                       const Scales &, const BeamDev *, const SolveDev<float, double> *, int,
                                                       hipStream_t);
template hipError_t launch_prop_c<double, double>(int, int, const PropArgs<double> &, const SolveDev<double, double> *, int,
                                                     hipStream_t);
template hipError_t launch_combine_c<double, double>(int, const CombArgs<double, double> &, const SolveDev<double, double> *, int,
                                                        hipStream_t);
template hipError_t launch_source_c<double, double>(int, const GridDev<double> &, double, double, double, double, double,
                                                       const Scales &, const BeamDev *, const SolveDev<double, double> *, int,
                                                       hipStream_t);
template hipError_t launch_prop_c<float, float>(int, int, const PropArgs<float> &, const SolveDev<float, float> *, int,
                                                     hipStream_t);
template hipError_t launch_combine_c<float, float>(int, const CombArgs<float, float> &, const SolveDev<float, float> *, int,
                                                        hipStream_t);
template hipError_t launch_source_c<float, float>(int, const GridDev<float> &, double, double, double, double, double,
                                                       const Scales &, const BeamDev *, const SolveDev<float, float> *, int,
                                                       hipStream_t);
template hipError_t launch_fill<float>(float *, size_t, float, hipStream_t);
template hipError_t launch_fill<double>(double *, size_t, double, hipStream_t);
template hipError_t launch_prop<float, double>(int, int, int, const PropArgs<float> &,
                                               const SolveDev<float, double> *, int, hipStream_t);
template hipError_t launch_prop<double, double>(int, int, int, const PropArgs<double> &,
                                                const SolveDev<double, double> *, int, hipStream_t);
template hipError_t launch_prop<float, float>(int, int, int, const PropArgs<float> &,
                                              const SolveDev<float, float> *, int, hipStream_t);
template hipError_t launch_ktab<float>(const GridDev<float> &, float, const float *, float *, hipStream_t);
template hipError_t launch_ktab<double>(const GridDev<double> &, double, const double *, double *, hipStream_t);
template hipError_t launch_combine<float, double>(int, const CombArgs<float, double> &,
                                                  const SolveDev<float, double> *, int, hipStream_t);
template hipError_t launch_combine<double, double>(int, const CombArgs<double, double> &,
                                                   const SolveDev<double, double> *, int, hipStream_t);
template hipError_t launch_combine<float, float>(int, const CombArgs<float, float> &,
                                                 const SolveDev<float, float> *, int, hipStream_t);
template hipError_t launch_bottom<float, double>(int, const SolveDev<float, double> *, int,
                                                 hipStream_t);
template hipError_t launch_bottom<double, double>(int, const SolveDev<double, double> *, int,
                                                  hipStream_t);
template hipError_t launch_bottom<float, float>(int, const SolveDev<float, float> *, int,
                                                hipStream_t);
template hipError_t launch_source<float, double>(int, const GridDev<float> &, double, double, double,
                                                 double, double, const Scales &, const BeamDev *,
                                                 const SolveDev<float, double> *, int, hipStream_t);
template hipError_t launch_source<double, double>(int, const GridDev<double> &, double, double, double,
                                                  double, double, const Scales &, const BeamDev *,
                                                  const SolveDev<double, double> *, int, hipStream_t);
template hipError_t launch_source<float, float>(int, const GridDev<float> &, double, double, double,
                                                double, double, const Scales &, const BeamDev *,
                                                const SolveDev<float, float> *, int, hipStream_t);
template hipError_t launch_to_voxel_major<float>(int, int, const Scales &, const double *, float *, hipStream_t);
template hipError_t launch_to_voxel_major<double>(int, int, const Scales &, const double *, double *,
                                                  hipStream_t);
template hipError_t launch_cplx_to_real<float>(int, int, double, const Scales &, const double *, float *,
                                               hipStream_t);
template hipError_t launch_cplx_to_real<double>(int, int, double, const Scales &, const double *, double *,
                                                hipStream_t);
template hipError_t launch_real_to_cplx<float>(int, int, double, const float *, double *, hipStream_t);
template hipError_t launch_real_to_cplx<double>(int, int, double, const double *, double *, hipStream_t);
template hipError_t launch_to_coef_major<float>(int, int, const Scales &, const float *, double *, hipStream_t);
template hipError_t launch_to_coef_major<double>(int, int, const Scales &, const double *, double *,
                                                 hipStream_t);

}  // namespace rte
