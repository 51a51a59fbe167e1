"""Writes tests/golden/oracle_lognormal_16cube_L5.json: the FP64 oracle's exitance for the
16^3 heterogeneous L=5 workload (seeded maps from doi_amd.media.lognormal_16cube), with
the reference lambda stop rule and with the 1e-6 total-exitance rule.  The GPU
tests compare against this file (the oracle would take minutes at this size).

    python tests/golden/gen_oracle_fixtures.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from doi_amd import media  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    c = media.lognormal_16cube(seed=2)
    P = O.Problem(c["shape"], c["bounds"], c["mu_a"], c["mu_s"], c["g"])
    out = dict(workload="16^3, 0.1 cm voxels, L=5, g=0.8, lognormal sigma 0.2 (seed 2), one "
                        "radius-3 sphere with mu_a x2, beam at the top-face centre (diam 0.2)",
               shape=list(c["shape"]), bounds=list(c["bounds"]), L=c["L"], g=c["g"], beam=list(c["beam"]),
               seed=2, cases=[])
    for rule, name in ((O.STOP_DELTA, "delta_1e-6_cap30"), (O.STOP_LAMBDA, "lambda_ref")):
        t = time.time()
        r = P.solve(c["beam"], c["L"], stop_rule=rule, tol=1e-6, max_terms=30)
        print(name, r["terms"], time.time() - t, flush=True)
        out["cases"].append(dict(name=name, stop_rule=rule, tol=1e-6, max_terms=30, terms=r["terms"],
                                 exitance=r["exitance"].tolist()))
    with open(os.path.join(HERE, "oracle_lognormal_16cube_L5.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
