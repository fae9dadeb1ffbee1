"""Every reduction on the device is a per-workgroup partial summed in a fixed order by a finish
kernel (no floating-point atomics), so a solve repeated on the same inputs gives the same bits:
the iterate, the iteration count and the residual history.  The reference's serial loops
(numerics.F90:41-162) are deterministic; these tests hold the GPU solvers to the same."""
import numpy as np
import pytest

import hpccg_amd
import oracle
from hpccg_amd import learned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


def _twice(fn):
    return fn(), fn()


def test_fp64_cg_tree_mode_repeats_bitwise(ctx):
    A = oracle.Problem(40, 24, 36)
    ctx.upload(A.rowindex, A.colindex, A.values)
    ctx.set_dot_mode(hpccg_amd.DOT_TREE)
    b = oracle.spmv(A, oracle.xstar_sines(40, 24, 36, 8))
    tol = 1e-10 * np.linalg.norm(b)
    (x1, n1, r1, h1, _), (x2, n2, r2, h2, _) = _twice(lambda: ctx.cg(b, max_iter=300, tolerance=tol))
    assert n1 == n2 and r1 == r2
    assert np.array_equal(x1, x2) and np.array_equal(h1, h2)


def test_mixed_cg_repeats_bitwise(ctx):
    A = oracle.Problem(40, 24, 36)
    ctx.upload(A.rowindex, A.colindex, A.values)
    b = oracle.spmv(A, oracle.xstar_sines(40, 24, 36, 9))
    (x1, n1, t1, hr1, ht1, _), (x2, n2, t2, hr2, ht2, _) = _twice(
        lambda: ctx.cg_mixed(b, max_iter=300, tolerance=1e-10 * np.linalg.norm(b)))
    assert n1 == n2 and t1 == t2
    assert np.array_equal(x1, x2) and np.array_equal(hr1, hr2) and np.array_equal(ht1, ht2)


def test_learned_fcg_repeats_bitwise(ctx):
    ctx.cnn_set_weights(learned.load())
    dims = (48, 40, 33)
    A = oracle.Problem(*dims)
    ctx.upload(A.rowindex, A.colindex, A.values)
    b = oracle.spmv(A, oracle.xstar_sines(*dims, 10))
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray(ctx, A.n)

    def solve():
        xd.upload(np.zeros(A.n))
        res = ctx.fcg_learned_dev(dims, bd, xd, max_iter=300, tolerance=1e-10 * np.linalg.norm(b), want_hist=True)
        return res, xd.numpy()

    (r1, x1), (r2, x2) = _twice(solve)
    assert (r1["fcg_iters"], r1["cg_iters"], r1["fell_back"]) == (r2["fcg_iters"], r2["cg_iters"], r2["fell_back"])
    assert r1["true_normr"] == r2["true_normr"]
    assert np.array_equal(x1, x2) and np.array_equal(r1["hist"], r2["hist"])


def test_batched_solve_repeats_bitwise(ctx):
    dims = np.array([[20, 17, 9], [12, 30, 16], [25, 25, 25], [9, 9, 40]], np.int32)
    info = ctx.batch_setup(dims)
    rng = np.random.default_rng(12)
    bs = [oracle.spmv(oracle.Problem(*map(int, d)), rng.uniform(-1.0, 1.0, int(np.prod(d)))) for d in dims]
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, np.concatenate(bs))
    xd = hpccg_amd.DeviceArray(ctx, info["rows"])

    def solve():
        xd.upload(np.zeros(info["rows"]))
        res = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=300)
        return res, xd.numpy()

    (r1, x1), (r2, x2) = _twice(solve)
    assert np.array_equal(r1["niters"], r2["niters"]) and np.array_equal(r1["relres"], r2["relres"])
    assert list(r1["converged"]) == [1, 1, 1, 1]
    assert np.array_equal(x1, x2)
