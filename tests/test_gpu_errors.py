"""Error behaviour through the C ABI on the device: every invalid input is
refused with its status code and message before any kernel runs, and the
context keeps working afterwards (the reference exits the process instead,
GPU package/src/defines.h:24-44)."""
import numpy as np
import pytest

from doi_amd import rte
from tests.helpers import bounds_for, hetero_maps

pytestmark = pytest.mark.gpu

SHAPE = (6, 5, 4)


def _ok(ctx):
    bounds = bounds_for(SHAPE)
    ma, ms = hetero_maps(SHAPE, seed=3)
    img, st = ctx.solve(SHAPE, bounds, ma, ms, 0.8, (0.0, 0.0, 0.2, 1.0), rte.make_params(2))
    assert np.all(np.isfinite(img)) and st["terms"] > 1


def _code(fn):
    with pytest.raises(rte.RteError) as e:
        fn()
    return e.value.code, str(e.value)


def test_invalid_inputs_are_refused(ctx):
    bounds = bounds_for(SHAPE)
    N = int(np.prod(SHAPE))
    ma, ms = hetero_maps(SHAPE, seed=3)
    beam = (0.0, 0.0, 0.2, 1.0)
    p = rte.make_params(2)
    cases = []
    # optics and grid
    bad = ma.copy()
    bad[7] = -0.01
    cases.append(("negative mu_a", lambda: ctx.medium(SHAPE, bounds, bad, ms, 0.8), rte.E_ARG, "voxel 7"))
    nan = ms.copy()
    nan[3] = np.nan
    cases.append(("NaN mu_s", lambda: ctx.medium(SHAPE, bounds, ma, nan, 0.8), rte.E_ARG, "voxel 3"))
    cases.append(("|g| > 1", lambda: ctx.medium(SHAPE, bounds, ma, ms, 1.5), rte.E_ARG, "|g|"))
    cases.append(("n <= 0", lambda: ctx.medium(SHAPE, bounds, ma, ms, 0.8, n=0.0), rte.E_ARG, "n and c"))
    cases.append(("empty grid", lambda: ctx.medium((0, 5, 4), bounds, ma, ms, 0.8), rte.E_ARG, "dims"))
    flipped = (bounds[1], bounds[0]) + bounds[2:]
    cases.append(("decreasing extent", lambda: ctx.medium(SHAPE, flipped, ma, ms, 0.8), rte.E_ARG, "increasing"))
    m = ctx.medium(SHAPE, bounds, ma, ms, 0.8)
    # parameters
    cases.append(("L too large", lambda: ctx.solve_media([m], [beam], rte.make_params(rte.MAX_L + 1)),
                  rte.E_UNSUPP, "L="))
    cases.append(("stop rule", lambda: ctx.solve_media([m], [beam], rte.make_params(2, stop_rule=7)), rte.E_ARG,
                  "stop_rule"))
    cases.append(("negative tol", lambda: ctx.solve_media([m], [beam], rte.make_params(2, stop_rule=rte.STOP_DELTA,
                                                                                       tol=-1.0)),
                  rte.E_ARG, "tol"))
    cases.append(("precision", lambda: ctx.solve_media([m], [beam], rte.make_params(2, precision=9)), rte.E_ARG,
                  "precision"))
    # beams and batches
    cases.append(("beam off the grid", lambda: ctx.solve_media([m], [(5.0, 5.0, 0.1, 1.0)], p), rte.E_ARG,
                  "no voxel column"))
    cases.append(("NaN beam", lambda: ctx.solve_media([m], [(np.nan, 0.0, 0.2, 1.0)], p), rte.E_ARG, "beam"))
    other = ctx.medium((6, 5, 5), bounds_for((6, 5, 5)), np.full(150, 0.01), np.full(150, 1.0), 0.8)
    cases.append(("mixed grids", lambda: ctx.solve_media([m, other], [beam, beam], p), rte.E_ARG, "different grid"))
    cases.append(("NaN omega", lambda: ctx.solve_fd([m], [beam], [np.nan], p), rte.E_ARG, "omega"))
    with rte.Context(0) as ctx2:
        foreign = ctx2.medium(SHAPE, bounds, ma, ms, 0.8)
        cases.append(("medium of another context", lambda: ctx.solve_media([foreign], [beam], p), rte.E_ARG,
                      "another context"))
        for name, fn, code, text in cases:
            print("error case:", name, flush=True)
            got, msg = _code(fn)
            assert got == code, (name, got, msg)
            assert text in msg, (name, msg)
        foreign.close()
    other.close()
    m.close()
    _ok(ctx)
