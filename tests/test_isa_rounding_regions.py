"""The round-down DDA loops run with the wave's FP64 rounding mode set toward
-inf (nm_cast.hip: fp64_round_down_begin / fp64_round_nearest).  Inside such
a region every rounded FP64 operation must be one whose round-down result is
wanted: the crossing-time adds, plus operations that are exact in any mode.
A divide or square root there would come out wrong (their expansions assume
round-to-nearest, and the compiler rewrites the mode bits around them), and a
branch that enters or leaves the region without passing a mode write would
run code in the wrong mode.  This test compiles nm_cast.hip for gfx950 and
checks every region in the generated assembly.  (CPU only: hipcc
cross-compiles without a GPU.)
"""
import os
import re
import subprocess
import tempfile

import pytest

from nanomap_amd import build as nb

BEGIN = re.compile(r"s_setreg_imm32_b32\s+hwreg\(HW_REG_MODE,\s*2,\s*2\),\s*2\b")
END = re.compile(r"s_setreg_imm32_b32\s+hwreg\(HW_REG_MODE,\s*2,\s*2\),\s*0\b")
# floating-point mnemonics allowed inside a region: the adds, exact scalings
# and floors of the in-binade jump, min/max, compares and selects
ALLOWED_FP = re.compile(
    r"^v_(add_f64|mul_f64|floor_f64|min_f64|max_f64|cmp\w*_f64|cmpx\w*_f64)")
FP_OP = re.compile(r"^v_\w*?_(f16|f32|f64|bf16)(_|$)|^v_cvt\w*")
FORBIDDEN = re.compile(
    r"^(v_div\w*|v_sqrt\w*|v_rsq\w*|v_rcp\w*|s_setreg\w*|s_round_mode|"
    r"s_denorm_mode|s_setreg_imm32\w*)")


@pytest.fixture(scope="module")
def cast_asm():
    if not os.path.exists(nb.HIPCC):
        pytest.skip("hipcc not available")
    src = os.path.join(nb.CSRC, "nm_cast.hip")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "cast.s")
        cmd = ([nb.HIPCC] + nb.COMMON +
               ["--cuda-device-only", "-S", "-o", out, src])
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        return open(out).read().split("\n")


def _functions(lines):
    """(name, [instruction or label lines]) per kernel body."""
    out, cur = [], None
    for ln in lines:
        m = re.match(r"^(_Z\w+):\s*(;.*)?$", ln)
        if m:
            cur = (m.group(1), [])
            out.append(cur)
            continue
        if cur is None:
            continue
        t = ln.strip()
        if not t or t.startswith(";") or (t.startswith(".") and not
                                          re.match(r"^\.LBB\w+:", t)):
            continue
        cur[1].append(t.split(";")[0].strip())
        if t.startswith("s_endpgm"):
            cur = None
    return out


def _regions(body):
    regions, start = [], None
    for k, t in enumerate(body):
        if BEGIN.search(t):
            assert start is None, "nested mode write"
            start = k
        elif END.search(t):
            assert start is not None, "mode restored without being set"
            regions.append((start, k))
            start = None
    assert start is None, "mode set and never restored"
    return regions


def test_round_down_regions_hold_only_intended_operations(cast_asm):
    funcs = _functions(cast_asm)
    seen = {}
    for name, body in funcs:
        regs = _regions(body)
        if not regs:
            continue
        kind = "accumulate" if "k_accumulate" in name else (
            "discover" if "k_discover" in name else name)
        seen[kind] = seen.get(kind, 0) + len(regs)
        labels = {t[:-1]: k for k, t in enumerate(body) if t.endswith(":")}
        for a, b in regs:
            adds = 0
            for k in range(a + 1, b):
                t = body[k]
                op = t.split()[0] if t else ""
                assert not FORBIDDEN.match(op), (name, t)
                if FP_OP.match(op):
                    assert ALLOWED_FP.match(op), (name, t)
                adds += op == "v_add_f64"
                if op.startswith("s_cbranch") or op == "s_branch":
                    tgt = t.split()[-1]
                    assert a < labels.get(tgt, -1) < b, (
                        "branch out of region", name, t)
            assert adds >= 1, (name, "region without the crossing add")
            # nothing outside jumps into the region
            inside = {lab for lab, k in labels.items() if a < k < b}
            for k, t in enumerate(body):
                if a <= k <= b:
                    continue
                op = t.split()[0] if t else ""
                if op.startswith("s_cbranch") or op == "s_branch":
                    assert t.split()[-1] not in inside, (
                        "branch into region", name, t)
    # both loops use the hardware mode in their round-down instantiations
    # (one k_discover, one k_accumulate per LDS table size)
    assert seen.get("discover", 0) >= 1 and seen.get("accumulate", 0) >= 2, seen
