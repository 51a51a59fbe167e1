"""Seeded synthetic media and beams for the benchmark and parity workloads.

RNG: SplitMix64 counter stream -> uniform doubles in [0,1) -> Box-Muller
normals; lognormal factor exp(sigma*Z) (median 1).  Every workload is a pure
function of its seed.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n consecutive SplitMix64 outputs starting from state `seed`."""
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = state
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    def __init__(self, seed):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def uniform(self, n):
        z = splitmix64((self.seed + self.pos * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF, n)
        self.pos += n
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def cube_bounds(n, dx=0.1):
    """[-n dx/2, n dx/2]^2 x [0, n dx] (top face at z = 0)."""
    h = n * dx / 2.0
    return (-h, h, -h, h, 0.0, n * dx)


def voxel_centres(n, dx=0.1):
    return (np.arange(n) + 0.5) * dx - n * dx / 2.0


def heterogeneous(n, seed, mua=0.01, mus=1.0, sigma=0.2, n_incl=1, radius=(3, 3),
                  factor=(2.0, 2.0), margin=3):
    """Lognormal per-voxel mu_a and mu_s plus spherical mu_a inclusions
    (radius in voxels, centre uniform over voxel centres at least `margin`
    voxels from each face).  Arrays are x fastest (N = n^3)."""
    st = Stream(seed)
    N = n ** 3
    ma = mua * np.exp(sigma * st.normal(N))
    ms = mus * np.exp(sigma * st.normal(N))
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ii, jj, kk = i.ravel(), j.ravel(), k.ravel()
    if isinstance(n_incl, tuple):
        n_incl = int(n_incl[0] + np.floor(st.uniform(1)[0] * (n_incl[1] - n_incl[0] + 1)))
    lo, hi = margin, n - 1 - margin
    for _ in range(n_incl):
        c = lo + np.floor(st.uniform(3) * (hi - lo + 1))
        r = radius[0] + st.uniform(1)[0] * (radius[1] - radius[0])
        f = factor[0] + st.uniform(1)[0] * (factor[1] - factor[0])
        inside = (ii - c[0]) ** 2 + (jj - c[1]) ** 2 + (kk - c[2]) ** 2 <= r * r
        ma = np.where(inside, ma * f, ma)
    return ma, ms


def uniform_8cube():
    """8^3 uniform cube, 0.1 cm voxels, mu_a 0.01, mu_s 1, g 0.8, L 3, centred beam diam 0.2."""
    return dict(shape=(8, 8, 8), bounds=cube_bounds(8), mu_a=0.01, mu_s=1.0, g=0.8, L=3,
                beam=(0.0, 0.0, 0.2, 1.0))


def lognormal_16cube(seed=2):
    """16^3 heterogeneous (lognormal sigma 0.2, one radius-3 sphere with mu_a x2), L 5, g 0.8."""
    ma, ms = heterogeneous(16, seed, sigma=0.2, n_incl=1, radius=(3, 3), factor=(2.0, 2.0))
    return dict(shape=(16, 16, 16), bounds=cube_bounds(16), mu_a=ma, mu_s=ms, g=0.8, L=5,
                beam=(0.0, 0.0, 0.2, 1.0))


def lognormal_media_with_inclusions(n_media=32, seed=5, n=32):
    """32 media of 32^3 (lognormal sigma 0.3, 1-3 spheres radius 2-5, mu_a x1.5-3), g 0.9."""
    out = []
    for m in range(n_media):
        ma, ms = heterogeneous(n, seed * 1000003 + m, sigma=0.3, n_incl=(1, 3), radius=(2, 5),
                               factor=(1.5, 3.0), margin=2)
        out.append((ma, ms))
    return out


def top_face_beams(seed=5, n=32, per_medium=5, dx=0.1):
    """Pencil beams (diam = one voxel) at seeded top-face voxel centres."""
    st = Stream(seed * 7919 + 1)
    xs = voxel_centres(n, dx)
    idx = np.floor(st.uniform(2 * per_medium) * n).astype(int)
    return [(xs[idx[2 * b]], xs[idx[2 * b + 1]], dx, 1.0) for b in range(per_medium)]


def serving_requests(n_req=2000, seed=4, edges=(8, 12, 16, 24, 32), orders=(3, 5), rate=200.0, dx=0.1,
                     tol=1e-6, max_terms=30, precision=None):
    """Mixed serving stream: Poisson arrivals at `rate` per second (all at t=0
    if rate is 0), cubic grid edge uniform over `edges`, L over `orders`,
    g ~ U[0.7, 0.9], background mu_s ~ U[0.5, 1.5] and mu_a ~ U[0.005, 0.05]
    per cm, each with a lognormal per-voxel factor (sigma 0.2), and a pencil
    beam one voxel wide on a random top-face voxel column.  Solves run in
    mixed precision to a 1e-6 relative change of total exitance (cap 30)."""
    from . import rte
    prec = rte.PREC_MIXED if precision is None else precision
    st = Stream(seed * 104729 + 17)
    t, out = 0.0, []
    for k in range(n_req):
        u = st.uniform(8)
        if rate > 0:
            t += -np.log1p(-u[0]) / rate
        n = int(edges[min(int(u[1] * len(edges)), len(edges) - 1)])
        L = int(orders[min(int(u[2] * len(orders)), len(orders) - 1)])
        g = 0.7 + 0.2 * u[3]
        mus = 0.5 + 1.0 * u[4]
        mua = 0.005 + 0.045 * u[5]
        N = n ** 3
        ma = mua * np.exp(0.2 * st.normal(N))
        ms = mus * np.exp(0.2 * st.normal(N))
        xs = voxel_centres(n, dx)
        bx = xs[min(int(u[6] * n), n - 1)]
        by = xs[min(int(u[7] * n), n - 1)]
        out.append(dict(t=t, shape=(n, n, n), bounds=cube_bounds(n, dx), mu_a=ma, mu_s=ms, g=g,
                        beam=(bx, by, dx, 1.0), L=L,
                        params=rte.make_params(L, precision=prec, stop_rule=rte.STOP_DELTA, tol=tol,
                                               max_terms=max_terms)))
    return out
