"""rte_propagate_noalloc (every device buffer in the library's static scratch
array) against rte_propagate with the marched path integral: the same kernels,
chunking and moments, so the same bits, on two calls in a row (the second on a
smaller grid, so the scratch holds the first call's larger contents), in each
precision; oversized calls and the cached path are refused before any device
work."""
import numpy as np
import pytest

from doi_amd import rte
from tests.helpers import bounds_for, hetero_maps

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [rte.PREC_FP64, rte.PREC_MIXED, rte.PREC_FP32])
def test_noalloc_propagate_equals_propagate(prec):
    ctx = rte.Context(0)
    try:
        for shape, L, seed in (((12, 12, 12), 5, 3), ((6, 5, 4), 3, 4)):
            print("noalloc", prec, shape, L, flush=True)
            bounds = bounds_for(shape)
            ma, ms = hetero_maps(shape, seed=seed)
            K, N = (L + 1) ** 2, int(np.prod(shape))
            W = np.random.default_rng(seed).standard_normal((K, N))
            ref = ctx.propagate(shape, bounds, ma, ms, 0.85, W, L, precision=prec, path_mode=rte.PATH_MARCH,
                                n=1.3)
            got = ctx.propagate_noalloc(shape, bounds, ma, ms, 0.85, W, L, precision=prec, n=1.3)
            assert np.max(np.abs(ref)) > 0
            assert np.array_equal(got, ref), np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    finally:
        ctx.close()


def test_noalloc_refuses_what_does_not_fit():
    ctx = rte.Context(0)
    try:
        shape, L = (16, 16, 16), 5
        bounds = bounds_for(shape)
        ma, ms = hetero_maps(shape, seed=1)
        W = np.zeros(((L + 1) ** 2, int(np.prod(shape))))
        with pytest.raises(rte.RteError) as ei:
            ctx.propagate_noalloc(shape, bounds, ma, ms, 0.8, W, L, precision=rte.PREC_FP64)
        assert ei.value.code == rte.E_UNSUPP and "scratch" in str(ei.value)
        with pytest.raises(rte.RteError) as ei:
            ctx.propagate_noalloc(shape, bounds, ma, ms, 0.8, W, L, precision=rte.PREC_MIXED,
                                  path_mode=rte.PATH_CACHED)
        assert ei.value.code == rte.E_UNSUPP
    finally:
        ctx.close()
