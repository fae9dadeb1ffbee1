"""The no-allocation FP64 CG entry point (hpccg_cg_dev_noalloc) against hpccg_cg_dev, through the
C ABI: same niters, normr, per-iteration history and x, bit for bit, on two calls in a row with
the second on a smaller system (the static workspace keeps the first call's contents), in both
dot modes; sizes beyond the workspace's limits are rejected with an error."""
import math

import numpy as np
import pytest

import hpccg_amd
import oracle

pytestmark = pytest.mark.gpu


def _solve_both(ctx, dims, seed, max_iter, tol_rel, mode):
    A = oracle.Problem(*dims)
    xs = oracle.xstar_sines(*dims, seed)
    b = oracle.spmv(A, xs)
    ctx.upload(A.rowindex, A.colindex, A.values)
    ctx.set_dot_mode(mode)
    tol = tol_rel * math.sqrt(oracle.ddot1(b))
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
    na = ctx.cg_dev_noalloc(bd, xd, max_iter=max_iter, tolerance=tol)
    x_na = xd.numpy()
    xd.upload(np.zeros(A.n))
    ref = ctx.cg_dev(bd, xd, max_iter=max_iter, tolerance=tol)
    x_ref = xd.numpy()
    ctx.set_dot_mode(hpccg_amd.DOT_TREE)
    return A, b, na, x_na, ref, x_ref


@pytest.mark.parametrize("mode", [hpccg_amd.DOT_TREE, hpccg_amd.DOT_SEQUENTIAL])
def test_noalloc_cg_equals_cg_dev_on_two_calls(mode):
    with hpccg_amd.Context(0) as ctx:
        for dims, seed, max_iter, tol in [((40, 36, 32), 5, 500, 1e-10), ((17, 12, 9), 6, 300, 1e-10)]:
            A, b, na, x_na, ref, x_ref = _solve_both(ctx, dims, seed, max_iter, tol, mode)
            assert na["status"] == ref["status"] == 0
            assert (na["niters"], na["normr"]) == (ref["niters"], ref["normr"])
            assert na["hist"].tolist() == ref["hist"].tolist()
            assert np.array_equal(x_na, x_ref)
            rel = math.sqrt(oracle.ddot1(b - oracle.spmv(A, x_na))) / math.sqrt(oracle.ddot1(b))
            assert rel <= 2e-10, rel


def test_noalloc_cg_fixed_iteration_count_and_history():
    """tolerance 0 (the reference driver's setting): max_iter iterations, hist of every one."""
    with hpccg_amd.Context(0) as ctx:
        A = oracle.Problem(16, 8, 12)
        ctx.upload(A.rowindex, A.colindex, A.values)
        bd = hpccg_amd.DeviceArray.from_numpy(ctx, A.b)
        xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
        na = ctx.cg_dev_noalloc(bd, xd, max_iter=150)
        x_na = xd.numpy()
        xd.upload(np.zeros(A.n))
        ref = ctx.cg_dev(bd, xd, max_iter=150)
        assert na["niters"] == ref["niters"] == 150
        assert na["hist"].tolist() == ref["hist"].tolist() and np.array_equal(x_na, xd.numpy())


def test_noalloc_cg_rejects_sizes_beyond_the_workspace():
    with hpccg_amd.Context(0) as ctx:
        A = oracle.Problem(6, 5, 4)
        ctx.upload(A.rowindex, A.colindex, A.values)
        bd = hpccg_amd.DeviceArray.from_numpy(ctx, A.b)
        xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
        with pytest.raises(hpccg_amd.HpccgError, match="max_iter"):
            ctx.cg_dev_noalloc(bd, xd, max_iter=hpccg_amd._lib.NOALLOC_MAX_ITER + 1)
        ctx.generate(129, 128, 128)  # 2,113,536 rows > 2,097,152
        big = hpccg_amd.DeviceArray(ctx, 129 * 128 * 128)
        with pytest.raises(hpccg_amd.HpccgError, match="rows exceed"):
            ctx.cg_dev_noalloc(big, big, max_iter=10)
        # the limit itself is accepted: 128^3 rows
        ctx.generate(128, 128, 128)
        b = hpccg_amd.DeviceArray.from_numpy(ctx, np.ones(128 ** 3))
        x = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(128 ** 3))
        r = ctx.cg_dev_noalloc(b, x, max_iter=5, want_hist=False)
        assert r["niters"] == 5 and r["normr"] > 0
