"""The SNR study (doi_amd.snr) on the GPU: the signal-present and -absent images
of every depth against the oracle's (FP64 path, 1e-9) for the three readouts,
SNR^2 from them, and the small-signal limit: the two-solve SNR^2 against
delta^2 F from the exact Fisher information of the inclusion's mu_a
(rte_fisher, an independent adjoint computation)."""
import numpy as np
import pytest

from doi_amd import rte, snr
from oracle import oracle as O
from tests.helpers import rel_max

pytestmark = pytest.mark.gpu

SHAPE, BOUNDS = (8, 8, 8), (-0.4, 0.4, -0.4, 0.4, 0.0, 0.8)
L, G = 3, 0.8
BEAM = (0.0, 0.0, 0.2, 1.0)


def _oracle_image(mask, mu_inc, image):
    N = int(np.prod(SHAPE))
    P = O.Problem(SHAPE, BOUNDS, np.where(mask, mu_inc, 0.01), np.full(N, 1.0), G)
    if image == "reflectance":
        return P.solve(BEAM, L, reflectance=True)["reflectance"]
    r = P.solve(BEAM, L, field=True)
    if image == "transmission":
        return r["exitance"]
    return snr.fim_top_image(r["w"], SHAPE, BOUNDS, L, mag=BEAM[3])


def test_depth_scan_against_oracle(ctx):
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_LAMBDA)
    for image in snr.IMAGES:
        print("snr scan", image, flush=True)
        res = snr.depth_scan(ctx, SHAPE, BOUNDS, 0.01, 1.0, G, BEAM, p, image=image)
        assert res["depths"] == list(range(2, 8))
        for z, s2, g1, g0 in zip(res["depths"], res["snr2"], res["g1"], res["g0"]):
            mask = snr.inclusion_mask(SHAPE, BOUNDS, (4, 4, z), 0.15)
            o1 = _oracle_image(mask, 0.35, image)
            o0 = _oracle_image(mask, 0.25, image)
            assert rel_max(g1, o1) < 1e-9 and rel_max(g0, o0) < 1e-9, (image, z)
            assert abs(s2 - snr.snr2(o1, o0)) <= 1e-7 * snr.snr2(o1, o0), (image, z)
            assert s2 > 0


def test_small_signal_limit_is_the_fisher_information(ctx):
    p = rte.make_params(L, precision=rte.PREC_FP64, stop_rule=rte.STOP_DELTA, tol=1e-12, max_terms=40)
    z, delta = 4, 1e-3
    mask = snr.inclusion_mask(SHAPE, BOUNDS, (4, 4, z), 0.15)
    res = snr.depth_scan(ctx, SHAPE, BOUNDS, 0.01, 1.0, G, BEAM, p, depths=[z], mua_absent=0.25, delta=delta,
                         image="transmission")
    N = int(np.prod(SHAPE))
    lin = snr.linearized_snr2(ctx, SHAPE, BOUNDS, np.where(mask, 0.25, 0.01), np.full(N, 1.0), G, BEAM, p, mask,
                              delta)
    print("snr2", res["snr2"][0], "linearized", lin, flush=True)
    assert lin > 0
    assert abs(res["snr2"][0] - lin) < 1e-2 * lin
