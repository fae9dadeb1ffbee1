"""GPU tests of the mixed-precision CG (FP16 values, FP32 vectors, FP64 dots / x, residual
replacement).  It has no reference counterpart; it is checked by the FP64 oracle: the true
residual ||b - A x|| / ||b|| of its answer (computed by the oracle in FP64) must meet the
tolerance, and its error against x* must be as small as the FP64 reference solve's."""
import math

import numpy as np
import pytest

import hpccg_amd
import oracle
from conftest import load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


def test_spmv_mixed_is_fp64_sum_rounded_once(ctx):
    """FP16 values x FP32 vector: products exact in FP64, row sums in the reference's order,
    one rounding to FP32 -- bitwise equal to the oracle's FP64 spmv rounded to FP32."""
    for dims in [(16, 8, 12), (33, 17, 5)]:
        A = oracle.Problem(*dims)
        ctx.upload(A.rowindex, A.colindex, A.values)
        x32 = oracle.xstar_uniform(A.n, 77).astype(np.float32)
        xd = hpccg_amd.DeviceArray.from_numpy(ctx, x32)
        yd = hpccg_amd.DeviceArray(ctx, A.n, np.float32)
        ctx.spmv_mixed_dev(xd, yd)
        ref = oracle.spmv(A, x32.astype(np.float64)).astype(np.float32)
        assert np.array_equal(yd.numpy(), ref)


@pytest.mark.parametrize("grid,seed", [((32, 32, 32), 11), ((64, 64, 64), 12), ((40, 24, 16), 3)])
def test_cg_mixed_reaches_fp64_true_residual(ctx, grid, seed):
    A = oracle.Problem(*grid)
    xs = oracle.xstar_sines(*grid, seed)
    b = oracle.spmv(A, xs)
    normb = math.sqrt(oracle.ddot1(b))
    tol = 1e-10 * normb
    ctx.upload(A.rowindex, A.colindex, A.values)
    x, nit, tn, hr, ht, st = ctx.cg_mixed(b, max_iter=1000, tolerance=tol, replace_every=25)
    assert st == 0
    r = b - oracle.spmv(A, x)
    true_rel = math.sqrt(oracle.ddot1(r)) / normb
    assert true_rel <= 1e-10, true_rel
    assert abs(tn / normb - true_rel) <= 1e-3 * true_rel  # the reported norm is the true one
    _, nit64, _, _, _ = oracle.cg(A, b, max_iter=1000, tolerance=tol)
    assert nit <= 1.3 * nit64 + 10, (nit, nit64)
    xo, _, _, _, _ = oracle.cg(A, b, max_iter=1000, tolerance=tol)
    e_mixed = float(np.max(np.abs(x - xs)))
    e_fp64 = float(np.max(np.abs(xo - xs)))
    assert e_mixed <= 10 * e_fp64 + 1e-12, (e_mixed, e_fp64)
    # replacement points carry a true residual; Prepare's is ||b|| (of b - A*0 stored in FP32)
    assert ht[0] == pytest.approx(normb, rel=1e-7)
    assert all(ht[j] > 0 for j in range(25, nit, 25))


def test_cg_mixed_device_api_matches_host_api(ctx):
    A = oracle.Problem(24, 20, 16)
    xs = oracle.xstar_sines(24, 20, 16, 4)
    b = oracle.spmv(A, xs)
    tol = 1e-10 * math.sqrt(oracle.ddot1(b))
    ctx.upload(A.rowindex, A.colindex, A.values)
    xh, nit, tn, hr, ht, _ = ctx.cg_mixed(b, max_iter=500, tolerance=tol)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
    res = ctx.cg_mixed_dev(bd, xd, max_iter=500, tolerance=tol)
    assert res["niters"] == nit and np.array_equal(xd.numpy(), xh)  # deterministic
    xd.upload(np.zeros(A.n))
    res2 = ctx.cg_mixed_dev(bd, xd, max_iter=500, tolerance=tol)  # cached graph replay
    assert res2["niters"] == nit and np.array_equal(xd.numpy(), xh)


def test_cg_mixed_forms_a_true_residual_after_a_1e5_fall(ctx):
    """Without periodic replacement (replace_every=0), the FP32 recursive residual alone would
    drift from b - A x by ~1e-7 of ||b||.  So the solver forms the FP64 true residual again once
    the recursive norm has fallen 1e5-fold since the last one: at least one true residual lies
    strictly between Prepare and the final iteration.  The solve still meets 1e-10, within a
    few iterations of FP64 CG."""
    grid = (48, 48, 48)
    A = oracle.Problem(*grid)
    b = oracle.spmv(A, oracle.xstar_sines(*grid, 21))
    normb = math.sqrt(oracle.ddot1(b))
    ctx.upload(A.rowindex, A.colindex, A.values)
    x, nit, tn, hr, ht, st = ctx.cg_mixed(b, max_iter=1000, tolerance=1e-10 * normb, replace_every=0)
    assert st == 0
    inner = [j for j in range(1, nit) if ht[j] > 0]
    assert inner, (nit, ht[:nit + 1])
    j = inner[0]
    assert hr[j] <= 1e-5 * normb * 1.0000001  # it fired on the 1e5 fall of the recursive norm
    rel = math.sqrt(oracle.ddot1(b - oracle.spmv(A, x))) / normb
    assert rel <= 1e-10, rel
    _, nit64, _, _, _ = oracle.cg(A, b, max_iter=1000, tolerance=1e-10 * normb)
    assert nit <= nit64 + 3, (nit, nit64)


def test_cg_mixed_edges(ctx):
    A = oracle.Problem(10, 9, 8)
    ctx.upload(A.rowindex, A.colindex, A.values)
    x, nit, tn, hr, ht, st = ctx.cg_mixed(A.b, max_iter=50, tolerance=1e9)
    assert nit == 0 and np.all(x == 0) and tn == pytest.approx(math.sqrt(oracle.ddot1(A.b)))
    x, nit, tn, hr, ht, st = ctx.cg_mixed(A.b, max_iter=7, tolerance=0.0, replace_every=0)
    assert nit == 7 and ht[7] > 0  # max_iter forces a final true residual
    r = A.b - oracle.spmv(A, x)
    assert tn == pytest.approx(math.sqrt(oracle.ddot1(r)), rel=1e-6)
    x, nit, tn, hr, ht, st = ctx.cg_mixed(A.b, max_iter=0)
    assert nit == 0


def test_cg_mixed_rejects_values_not_in_fp16(ctx):
    A = oracle.Problem(5, 5, 5)
    vals = A.values.copy()
    vals[3] = 0.1
    ctx.upload(A.rowindex, A.colindex, vals)
    with pytest.raises(hpccg_amd.HpccgError):
        ctx.cg_mixed(A.b, max_iter=5)


def test_mixed_cg_over_int32_columns(ctx):
    """A symmetric matrix with thousands of distinct offsets col - row: the FP16 sliced copy keeps
    int32 columns.  Values -1 off the diagonal and (row degree + 1) on it are FP16 numbers and
    make the matrix SPD (strictly diagonally dominant).  The FP16/FP32 SpMV is the oracle's FP64
    spmv rounded once to FP32, and the mixed solve meets 1e-10 in the oracle's FP64 true residual."""
    rng = np.random.default_rng(17)
    n = 5000
    i = rng.integers(0, n, 40000)
    j = rng.integers(0, n, 40000)
    keep = i != j
    pairs = set(zip(i[keep].tolist(), j[keep].tolist())) | set(zip(j[keep].tolist(), i[keep].tolist()))
    rows = [[] for _ in range(n)]
    for a, c in pairs:
        rows[a].append(c)
    cols, vals, rowindex = [], [], [1]
    for a in range(n):
        cs = sorted(rows[a] + [a])
        cols += [c + 1 for c in cs]
        vals += [float(len(cs)) if c == a else -1.0 for c in cs]
        rowindex.append(len(cols) + 1)
    rowindex = np.array(rowindex, np.int32)
    cols = np.array(cols, np.int32)
    vals = np.array(vals)
    ctx.upload(rowindex, cols, vals)
    assert ctx.matrix_format()["ncodes"] == 0
    A = type("P", (), {})()  # minimal oracle.Problem stand-in for oracle.spmv
    A.n, A.rowindex, A.colindex, A.values = n, rowindex, cols, vals
    x32 = oracle.xstar_uniform(n, 41).astype(np.float32)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, x32)
    yd = hpccg_amd.DeviceArray(ctx, n, np.float32)
    ctx.spmv_mixed_dev(xd, yd)
    assert np.array_equal(yd.numpy(), oracle.spmv(A, x32.astype(np.float64)).astype(np.float32))
    xs = oracle.xstar_uniform(n, 42)
    b = oracle.spmv(A, xs)
    normb = math.sqrt(oracle.ddot1(b))
    x, nit, tn, hr, ht, st = ctx.cg_mixed(b, max_iter=1000, tolerance=1e-10 * normb, replace_every=25)
    assert st == 0
    rel = math.sqrt(oracle.ddot1(b - oracle.spmv(A, x))) / normb
    assert rel <= 1e-10, rel


def test_cg_mixed_from_a_nonzero_start(ctx):
    """Prepare forms b - A x0 with the SpMV when x0 != 0 (it skips it only for x0 = 0): its true
    residual is the oracle's ||b - A x0|| in FP32 steps, and the solve still meets 1e-10."""
    A = oracle.Problem(24, 20, 16)
    xs = oracle.xstar_sines(24, 20, 16, 8)
    b = oracle.spmv(A, xs)
    normb = math.sqrt(oracle.ddot1(b))
    ctx.upload(A.rowindex, A.colindex, A.values)
    x0 = oracle.xstar_uniform(A.n, 31)
    r0 = (b - oracle.spmv(A, x0)).astype(np.float32).astype(np.float64)
    x, nit, tn, hr, ht, st = ctx.cg_mixed(b, x0=x0, max_iter=1000, tolerance=1e-10 * normb, replace_every=25)
    assert st == 0
    assert ht[0] == pytest.approx(math.sqrt(float(r0 @ r0)), rel=1e-12)
    r = b - oracle.spmv(A, x)
    assert math.sqrt(oracle.ddot1(r)) / normb <= 1e-10
    # a start of negative zeros is a zero start: the same iterates as from +0
    xz, nz, _, hz, _, _ = ctx.cg_mixed(b, x0=np.full(A.n, -0.0), max_iter=40, tolerance=0.0)
    xp, npos, _, hp, _, _ = ctx.cg_mixed(b, max_iter=40, tolerance=0.0)
    assert nz == npos and np.array_equal(xz, xp) and np.array_equal(hz, hp)
