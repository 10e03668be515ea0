"""The round-down DDA loops run with the wave's FP64 rounding mode set toward
-inf (nm_cast.hip: fp64_round_down_begin / fp64_round_nearest).  Inside such
a region every rounded FP64 operation must be one whose round-down result is
wanted: the crossing-time adds, plus operations that are exact in any mode.
A divide or square root there would come out wrong (their expansions assume
round-to-nearest, and the compiler rewrites the mode bits around them), and
code reachable both with and without the mode set would run in an unknown
mode.  This test compiles nm_cast.hip for gfx950, builds each kernel's
control-flow graph, propagates the mode from the kernel entry and checks
every instruction against the mode it runs in.  (CPU only: hipcc
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


def _blocks(body):
    """Basic blocks of a kernel body: (label, instructions, successors)."""
    blocks, cur = [], ["<entry>", [], []]
    blocks.append(cur)
    for t in body:
        if t.endswith(":"):
            cur = [t[:-1], [], []]
            blocks.append(cur)
            continue
        cur[1].append(t)
        op = t.split()[0]
        if op == "s_branch" or op.startswith("s_cbranch") or op == "s_endpgm":
            cur = ["<after %d>" % len(blocks), [], []]
            blocks.append(cur)
    index = {b[0]: k for k, b in enumerate(blocks)}
    for k, b in enumerate(blocks):
        last = b[1][-1].split() if b[1] else [""]
        op = last[0]
        if op == "s_branch":
            b[2] = [index[last[-1]]]
        elif op.startswith("s_cbranch"):
            b[2] = [index[last[-1]]] + ([k + 1] if k + 1 < len(blocks) else [])
        elif op == "s_endpgm":
            b[2] = []
        else:
            b[2] = [k + 1] if k + 1 < len(blocks) else []
    return blocks


def _mode_states(blocks):
    """FP64 rounding mode at each instruction, by forward dataflow from the
    kernel entry (round to nearest): 'rn', 'rd' or 'both'."""
    def join(a, b):
        return a if b is None or a == b else (b if a is None else "both")

    def transfer(state, insns):
        out = []
        for t in insns:
            out.append(state)
            if BEGIN.search(t):
                state = "rd"
            elif END.search(t):
                state = "rn"
        return state, out

    entry = [None] * len(blocks)
    entry[0] = "rn"
    work = [0]
    while work:
        k = work.pop()
        st, _ = transfer(entry[k], blocks[k][1])
        for n in blocks[k][2]:
            j = join(entry[n], st)
            if j != entry[n]:
                entry[n] = j
                work.append(n)
    return [transfer(entry[k], b[1])[1] if entry[k] is not None
            else [None] * len(b[1]) for k, b in enumerate(blocks)]


# The rounded FP64 operations of each round-down region, counted for this
# compiler (ROCm 7.2 hipcc, -O3).  The accumulate region is the voxel loop:
# its one add is the crossing-time add c + delta, and everything else is a
# compare.  The discover region holds the axis steps' adds (single rounded
# steps, and the in-binade jumps t + 8 D, exact), the jump's exact scalings
# by powers of two and its floor.  Any other count means the compiler moved
# code into or out of a region (an emulated set-up add, say, would run in
# the wrong mode and still pass the mnemonic check above): re-inspect the
# region before changing the numbers.
EXPECTED_RD_OPS = {
    "accumulate": {"v_add_f64": 1, "v_mul_f64": 0, "v_floor_f64": 0,
                   "v_min_f64": 0, "v_max_f64": 0},
    "discover": {"v_add_f64": 14, "v_mul_f64": 3, "v_floor_f64": 1,
                 "v_min_f64": 2, "v_max_f64": 5},
}


def test_round_down_regions_hold_only_intended_operations(cast_asm):
    funcs = _functions(cast_asm)
    seen = {}
    for name, body in funcs:
        if not any(BEGIN.search(t) for t in body):
            continue
        kind = "accumulate" if "k_accumulate" in name else (
            "discover" if "k_discover" in name else name)
        seen[kind] = seen.get(kind, 0) + 1
        blocks = _blocks(body)
        states = _mode_states(blocks)
        rd_adds = 0
        counts = {op: 0 for op in EXPECTED_RD_OPS.get(kind, {})}
        for b, st in zip(blocks, states):
            for t, m in zip(b[1], st):
                op = t.split()[0]
                if m == "both":
                    # reached both with and without the mode set
                    assert not FP_OP.match(op) and not FORBIDDEN.match(op), (
                        "mode-ambiguous", name, b[0], t)
                    assert not (BEGIN.search(t) or END.search(t)), (
                        "mode-ambiguous mode write", name, b[0], t)
                elif m == "rd":
                    assert not FORBIDDEN.match(op) or END.search(t), (
                        name, b[0], t)
                    assert not op.startswith("s_swappc") and not \
                        op.startswith("s_setpc"), ("call in region", name, t)
                    if FP_OP.match(op):
                        assert ALLOWED_FP.match(op), (name, b[0], t)
                    rd_adds += op == "v_add_f64"
                    base = re.sub(r"_e(32|64)$", "", op)
                    if base in counts:
                        counts[base] += 1
                elif m == "rn":
                    assert not END.search(t), ("restore outside region", name, t)
        assert rd_adds >= 1, (name, "no crossing add in round-down mode")
        if kind in EXPECTED_RD_OPS:
            assert counts == EXPECTED_RD_OPS[kind], (name, counts)
    # both loops use the hardware mode in their round-down instantiations
    # (one k_discover, one k_accumulate per LDS table size)
    assert seen.get("discover", 0) >= 1 and seen.get("accumulate", 0) >= 2, seen
