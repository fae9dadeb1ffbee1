"""The Fortran drop-in boundary on the GPU.

(1) Link substitution: the reference's own driver MPI/src/main.F90, compiled unchanged
    against hpccg-f90_amd/fortran's modules (oracle/_ref/dropin/HPCCG, built from the reference
    sources by oracle/Makefile), prints exactly the reference's stdout (timer lines aside).
(2) Every drop-in procedure called from Fortran (tests/fortran/dropin_selftest.F90) returns
    the reference's bits on the golden 6x5x4 inputs.
"""
import json
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLD, ROOT, load_golden

pytestmark = pytest.mark.gpu

DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin", "HPCCG")
DROPIN3 = os.path.join(ROOT, "oracle", "_ref", "dropin3", "HPCCG")   # both built with -DTIMINGLEV=3
REF3 = os.path.join(ROOT, "oracle", "_ref", "serial3", "HPCCG")
DRIVER = os.path.join(ROOT, "hpccg-f90_amd", "lib", "hpccg")
SELFTEST = os.path.join(ROOT, "hpccg-f90_amd", "lib", "dropin_selftest")


def run(exe, args, mode, keep_timers=False):
    env = dict(os.environ, HPCCG_DOT_MODE=mode)
    p = subprocess.run([exe] + args, capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return [ln for ln in p.stdout.splitlines() if keep_timers or "TIMER" not in ln]


def mask(lines):
    """Timer and MFLOPS values replaced as in tests/golden/ref_stdout_t3_*.txt (make_golden.py)."""
    out = []
    for ln in lines:
        if ln.strip().startswith("TIMER:"):
            ln = re.sub(r"(TIMER:\s+\S+\s+)\S+(\s+sec)", r"\1<t>\2", ln)
        elif ln.strip().startswith("MFLOPS SPARSEMV"):
            ln = re.sub(r"(MFLOPS SPARSEMV\s+)\S+", r"\1<r>", ln)
        out.append(ln)
    return out


def test_timinglev3_stdout_matches_the_reference_build():
    """The reference main with every timer region on (-DTIMINGLEV=3), linked against the
    drop-in modules, prints the reference's TIMINGLEV=3 lines in the reference's order
    (numerics.F90:91-156): TIMER: Prepare before the Initial Residual, TIMER: Iteration after
    every iteration, TIMER: Total, MFLOPS SPARSEMV; every other line identical (SEQUENTIAL dots).
    The values are device times: positive, the iterations within Total, and the MFLOPS line
    equal to niters * 2 * 27 N over the summed SpMV times -- so above the rate of the whole
    iterations."""
    if not os.path.exists(DROPIN3):
        pytest.skip("oracle/_ref/dropin3 not built (needs the reference sources at build time)")
    out = run(DROPIN3, ["16", "16", "16"], "sequential", keep_timers=True)
    with open(os.path.join(GOLD, "ref_stdout_t3_16x16x16.txt")) as f:
        ref = f.read().splitlines()
    assert mask(out) == ref
    if os.path.exists(REF3):  # the reference binary itself, run now, has the same shape
        assert mask(run(REF3, ["16", "16", "16"], "sequential", keep_timers=True)) == ref
    t = {"Iteration": [], "Prepare": [], "Total": []}
    for ln in out:
        m = re.match(r"\s*TIMER:\s+(\S+)\s+(\S+)\s+sec", ln)
        if m:
            t[m.group(1)].append(float(m.group(2)))
    assert len(t["Iteration"]) == 150 and len(t["Prepare"]) == len(t["Total"]) == 1
    assert min(t["Iteration"]) > 0 and t["Prepare"][0] > 0
    assert sum(t["Iteration"]) <= t["Total"][0] * 1.001
    mflops = float(out[-1].split()[-1])
    n = 16 ** 3
    assert mflops > 150 * 2.0 * 27 * n / sum(t["Iteration"]) / 1e6


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("level", [1, 2])
def test_timed_cg_equals_cg_bitwise(mode, level):
    """hpccg_cg_timed (level 1: graph-replayed loop; level 2: one iteration at a time with HIP
    events) gives hpccg_cg's niters, normr, history and x bit for bit, and times that add up."""
    import hpccg_amd
    import oracle
    A = oracle.Problem(20, 18, 16)
    xs = oracle.xstar_sines(20, 18, 16, 3)
    b = oracle.spmv(A, xs)
    with hpccg_amd.Context(0, mode) as ctx:
        ctx.upload(A.rowindex, A.colindex, A.values)
        tol = 1e-10 * float(np.sqrt(oracle.ddot1(b)))
        x, nit, normr, hist, st = ctx.cg(b, max_iter=300, tolerance=tol)
        r = ctx.cg_timed(b, max_iter=300, tolerance=tol, level=level)
    assert (r["niters"], r["normr"], r["status"]) == (nit, normr, st)
    assert r["hist"].tolist() == hist.tolist() and np.array_equal(r["x"], x)
    assert r["t_total"] > 0 and r["t_prepare"] > 0
    if level >= 2:
        assert len(r["t_iter"]) == nit and np.all(r["t_iter"] > 0) and np.all(r["t_spmv"] > 0)
        assert np.all(r["t_spmv"] < r["t_iter"])
        assert r["t_iter"].sum() <= r["t_total"] * 1.001


def golden_stdout():
    with open(os.path.join(GOLD, "ref_stdout_32x32x32.txt")) as f:
        return f.read().splitlines()


@pytest.mark.parametrize("exe", [DROPIN, DRIVER], ids=["reference-main", "own-driver"])
def test_stdout_identical_sequential_mode(exe):
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (oracle/_ref needs the reference sources at build time)")
    assert run(exe, ["32", "32", "32"], "sequential") == golden_stdout()


def test_stdout_tree_mode_matches_while_residual_is_resolved():
    exe = DROPIN if os.path.exists(DROPIN) else DRIVER
    out, ref = run(exe, ["32", "32", "32"], "tree"), golden_stdout()
    assert len(out) == len(ref)
    for a, b in zip(out, ref):
        if a.startswith("ITER"):
            k, va, vb = int(a.split()[1]), float(a.split()[-1]), float(b.split()[-1])
            if vb / 813.855 > 1e-10:  # printed to 10 digits; below ~1e-10 the tail is rounding noise
                assert abs(va - vb) <= 1e-9 * vb, (a, b)
        else:
            assert a == b


def _hex(s):
    return struct.unpack(">d", bytes.fromhex(s))[0]


@pytest.mark.parametrize("mode", ["sequential", "tree"])
def test_fortran_procedures_bitwise(mode):
    if not os.path.exists(SELFTEST):
        pytest.skip("dropin_selftest not built")
    case = [c for c in load_golden("matrix_and_kernels.json")["kernels"] if c["grid"] == [6, 5, 4]][0]
    lines = run(SELFTEST, [], mode)
    vecs = {ln.split()[1]: np.array([_hex(h) for h in ln.split()[2:]]) for ln in lines if ln.startswith("VEC")}
    for name in ("spmv", "waxpby", "wwpby", "wxbw", "vcopy"):
        assert np.array_equal(vecs[name], np.array(case[name])), name
    d = [ln for ln in lines if ln.startswith("DOTS")][0].split()
    if mode == "sequential":
        assert (_hex(d[1]), _hex(d[2])) == (case["ddot"], case["ddot1"])
    else:
        assert abs(_hex(d[1]) - case["ddot"]) <= 1e-13 and abs(_hex(d[2]) - case["ddot1"]) <= 1e-13
    cg = [ln for ln in lines if ln.startswith("CG")][0].split()
    assert int(cg[1]) == 40
    assert _hex(cg[3]) < 1e-12  # ||x - 1||_inf after 40 iterations on 6x5x4
