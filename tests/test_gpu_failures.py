"""GPU tests of failure handling: a NaN or Inf in the right-hand side.  The reference guards
nothing (numerics.F90:137-138 forms alpha = rtrans / (p.Ap) whatever p.Ap is, and the loop runs
to max_iter with NaN); here every solver ends within its iteration limit and says so: the FP64
and mixed solvers return the breakdown warning with the reference's NaN iterates, the flexible
solver hands over to FP64 CG, and in a batch only the system with the bad entry stops."""
import math

import numpy as np
import pytest

import hpccg_amd
import oracle
from hpccg_amd import learned

pytestmark = pytest.mark.gpu

WARN = hpccg_amd._lib.WARN_BREAKDOWN


@pytest.fixture(scope="module")
def ctx():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


def _bad_rhs(A, value, seed=3):
    b = oracle.spmv(A, oracle.xstar_sines(*A.dims, seed))
    b[A.n // 3] = value
    return b


@pytest.mark.parametrize("value", [np.nan, np.inf])
@pytest.mark.parametrize("mode", ["sequential", "tree"])
def test_fp64_cg_reports_breakdown_and_keeps_the_reference_iterates(ctx, value, mode):
    A = oracle.Problem(12, 10, 9)
    ctx.upload(A.rowindex, A.colindex, A.values)
    ctx.set_dot_mode(hpccg_amd.DOT_SEQUENTIAL if mode == "sequential" else hpccg_amd.DOT_TREE)
    b = _bad_rhs(A, value)
    x, nit, normr, hist, st = ctx.cg(b, max_iter=20, tolerance=1e-10)
    ctx.set_dot_mode(hpccg_amd.DOT_TREE)
    assert st == WARN
    assert nit == 20  # the reference's loop runs to max_iter: `normr < tolerance` is false for NaN
    assert not math.isfinite(normr) and np.isnan(x).any()
    if mode == "sequential":
        xo, nito, normro, histo, _ = oracle.cg(A, b, max_iter=20, tolerance=1e-10)
        assert nito == nit and np.array_equal(np.isnan(x), np.isnan(xo))


@pytest.mark.parametrize("value", [np.nan, np.inf])
def test_mixed_cg_reports_breakdown_within_max_iter(ctx, value):
    A = oracle.Problem(12, 10, 9)
    ctx.upload(A.rowindex, A.colindex, A.values)
    x, nit, tn, hr, ht, st = ctx.cg_mixed(_bad_rhs(A, value), max_iter=30, tolerance=1e-10)
    assert st == WARN
    assert nit <= 30 and not math.isfinite(tn)


def test_flexible_cg_hands_a_nan_rhs_to_fp64_cg(ctx):
    w = learned.load()
    assert w is not None, "stored preconditioner weights are part of the package"
    ctx.cnn_set_weights(w)
    n = 16
    A = oracle.Problem(n, n, n)
    ctx.upload(A.rowindex, A.colindex, A.values)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, _bad_rhs(A, np.nan))
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
    res = ctx.fcg_learned_dev((n, n, n), bd, xd, max_iter=40, tolerance=1e-10)
    assert res["fell_back"], res
    assert res["fcg_iters"] + res["cg_iters"] <= 40
    assert np.isnan(xd.numpy()).any()


def test_batch_nan_rhs_stops_only_its_system(ctx):
    dims = np.array([[12, 10, 9], [9, 9, 9], [14, 8, 11]], np.int32)
    info = ctx.batch_setup(dims)
    off = info["row_off"]
    bs = []
    for s, d in enumerate(dims):
        A = oracle.Problem(*map(int, d))
        bs.append(oracle.spmv(A, oracle.xstar_uniform(A.n, 40 + s)))
    bs[1][5] = np.nan
    b = np.concatenate(bs)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(info["rows"]))
    res = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=300)
    assert res["status"] == WARN
    assert list(res["converged"]) == [1, 0, 1]
    x = xd.numpy()
    for s in (0, 2):
        A = oracle.Problem(*map(int, dims[s]))
        xs, bs_ = x[off[s]:off[s + 1]], b[off[s]:off[s + 1]]
        rel = math.sqrt(oracle.ddot1(bs_ - oracle.spmv(A, xs))) / math.sqrt(oracle.ddot1(bs_))
        assert rel <= 1e-8, (s, rel)
