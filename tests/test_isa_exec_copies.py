"""The gfx950 code of every kernel is free of one miscompile seen in the FP64
complex propagation kernel: inside an if/else, the register allocator put
copies of registers that every lane needs into the flow block between the two
branches, where EXEC holds only the then-lanes; the else-lanes then read a
register that nothing wrote for them (here the path-table address, whose next
load faulted).  The check compiles the kernels to assembly (CPU only) and
flags any copy in a flow block whose destination is written nowhere outside
the then-branch and its flow block yet read elsewhere; two synthetic snippets
pin the check itself (the faulting shape is flagged, a phi merge is not)."""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from doi_amd import build

COPY = re.compile(r'^\s+(v_accvgpr_write_b32|v_accvgpr_read_b32|v_accvgpr_mov_b32|v_mov_b32_e32|v_mov_b64_e32|'
                  r'v_mov_b32_e64|v_cndmask_b32_e32|v_cndmask_b32_e64)\s+(\S+),')
STORES = ("global_store", "flat_store", "scratch_store", "buffer_store")


def _regs(tok):
    tok = tok.strip().lstrip("-|").rstrip("|,")
    m = re.match(r"([va])\[(\d+):(\d+)\]$", tok)
    if m:
        return {m.group(1) + str(i) for i in range(int(m.group(2)), int(m.group(3)) + 1)}
    return {tok} if re.match(r"[va]\d+$", tok) else set()


def _writes(line):
    t = line.split()
    if len(t) < 2 or not t[0].startswith(("v_", "global_load", "flat_load", "scratch_load", "buffer_load")):
        return set()
    if t[0].startswith(("v_writelane", "v_cmp", "v_cmpx")):
        return set()
    return _regs(t[1])


def _reads(line):
    t = line.strip().split(None, 1)
    if len(t) < 2:
        return set()
    ops = t[1].split(",")
    out = set()
    for o in (ops if t[0].startswith(STORES) else ops[1:]):
        out |= _regs(o)
    return out


def lost_lane_copies(asm):
    """[(kernel, line, register)] of copies into then-lane-only flow blocks
    whose register no instruction writes for the else-lanes."""
    found = []
    for m in re.finditer(r"^(_Z\S+):\s*;", asm, re.M):
        body = asm[m.end():asm.index(".Lfunc_end", m.end())].split("\n")
        writes = {}
        for i, l in enumerate(body):
            for r in _writes(l):
                writes.setdefault(r, []).append(i)
        for i, l in enumerate(body):
            br = re.match(r"\s+s_cbranch_execz (\S+)", l)
            if not (br and i and "s_xor_b64" in body[i - 1]):
                continue
            j = next((q for q in range(i, len(body)) if body[q].startswith(br.group(1) + ":")), None)
            if j is None:
                continue
            k = j + 1
            while k < len(body) and not body[k].startswith(".LBB") and "s_andn2_saveexec" not in body[k]:
                k += 1
            if k == len(body) or "s_andn2_saveexec" not in body[k]:
                continue
            for q in range(j + 1, k):
                c = COPY.match(body[q])
                if not c:
                    continue
                for r in _regs(c.group(2)):
                    if all(i <= w <= k for w in writes[r]) and \
                            any(r in _reads(body[z]) for z in range(len(body)) if not i <= z <= k):
                        found.append((m.group(1), q, r))
    return found


SNIPPET = """_Zkernel: ; @_Zkernel
\tv_mov_b32_e32 v2, 7
\ts_and_saveexec_b64 s[4:5], vcc
\ts_xor_b64 s[4:5], exec, s[4:5]
\ts_cbranch_execz .LBB0_2
\tv_mov_b32_e32 v3, 1
.LBB0_2:
\t{flow}
\ts_andn2_saveexec_b64 s[0:1], s[4:5]
\ts_cbranch_execz .LBB0_4
\t{other}
.LBB0_4:
\ts_or_b64 exec, exec, s[0:1]
\tv_accvgpr_read_b32 v2, a7
\tglobal_load_dword v4, v[2:3], off
\ts_endpgm
.Lfunc_end0:
"""


def test_check_flags_the_faulting_shape_and_not_a_phi():
    # the address v2 parked in a7 under the then-lanes' EXEC only: flagged
    bad = SNIPPET.format(flow="v_accvgpr_write_b32 a7, v2", other="v_mov_b32_e32 v3, 2")
    assert lost_lane_copies(bad) == [("_Zkernel", 7, "a7")]
    # a7 also written by the else branch (a phi of the two results): clean
    phi = SNIPPET.format(flow="v_accvgpr_write_b32 a7, v2", other="v_accvgpr_write_b32 a7, v3")
    assert lost_lane_copies(phi) == []
    # a7 written before the if (the else-lanes keep that value): clean
    pre = bad.replace("\tv_mov_b32_e32 v2, 7\n", "\tv_mov_b32_e32 v2, 7\n\tv_accvgpr_write_b32 a7, v2\n")
    assert lost_lane_copies(pre) == []


@pytest.mark.skipif(shutil.which(os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")) is None, reason="no hipcc")
def test_kernels_have_no_lost_lane_copies(tmp_path):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    flags = [f for f in build.HIPCC_FLAGS if f != "-shared"]
    inc = os.path.join(build.HERE, os.pardir, "include")

    def asm(src):
        out = str(tmp_path / (src + ".s"))
        subprocess.run([hipcc] + flags + ["-I", inc, "--cuda-device-only", "-S", os.path.join(build.CSRC, src),
                                          "-o", out], check=True, capture_output=True)
        with open(out) as f:
            return f.read()

    srcs = [s for s in build.SOURCES if s.endswith(".hip")]
    with ThreadPoolExecutor(len(srcs)) as ex:
        texts = list(ex.map(asm, srcs))
    found = [x for t in texts for x in lost_lane_copies(t)]
    assert found == [], found[:5]
