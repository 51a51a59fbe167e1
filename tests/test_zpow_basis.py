"""The propagation's z-power basis on the CPU: the conversion table ZPowBasis<L>
(doi_amd/csrc/rte_device.hpp) must reproduce the scaled real harmonics from
C_m z^k, S_m z^k on the unit sphere for every supported order, with C_m, S_m
from the polar form and Pt_l^m from its recurrence, and pair_zpow (the
kernel's products) must form those values.  Compiles tests/zpow_basis_check.hip
as a host program with hipcc."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.mark.skipif(HIPCC is None, reason="hipcc not found")
def test_zpow_basis_reproduces_the_harmonics():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "zpow_basis_check")
        subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", "-o", exe,
                        os.path.join(HERE, "zpow_basis_check.hip")], check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [l.split() for l in out if l.startswith("L ")]
    assert [int(r[1]) for r in rows] == list(range(8))
    for r in rows:
        assert float(r[3]) < 1e-13, r  # sum_j M_aj Z_j against the harmonics
        assert float(r[5]) < 1e-13, r  # pair_zpow against the direct products
