"""GPU tests of the batched solve (64 uneven 27-point systems in one graph launch).

Each system is a generate_matrix grid; the oracle checks each independently: the batched
SpMV equals the oracle's per-system FP64 spmv rounded once to FP32 (bitwise), and every
system's answer meets rel_tol in the FP64 true residual computed by the oracle."""
import math

import numpy as np
import pytest

import hpccg_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


def _batch(ctx, dims, seed):
    info = ctx.batch_setup(dims)
    off = info["row_off"]
    xs = np.concatenate([oracle.xstar_uniform(int(off[s + 1] - off[s]), seed + s) for s in range(len(dims))])
    bs = []
    for s, d in enumerate(dims):
        A = oracle.Problem(*map(int, d))
        bs.append(oracle.spmv(A, xs[off[s]:off[s + 1]]))
    return info, xs, np.concatenate(bs)


def test_batch_spmv_bitwise_per_system(ctx):
    # [130, 5, 4] and [64, 3, 3] hold uniform-code slices (whole 64-row runs of interior rows)
    dims = np.array([[5, 4, 3], [16, 9, 7], [1, 1, 1], [33, 2, 5], [12, 12, 12], [130, 5, 4], [64, 3, 3]], np.int32)
    info = ctx.batch_setup(dims)
    off = info["row_off"]
    x32 = oracle.xstar_uniform(info["rows"], 5).astype(np.float32)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, x32)
    yd = hpccg_amd.DeviceArray(ctx, info["rows"], np.float32)
    ctx.batch_spmv_dev(xd, yd)
    y = yd.numpy()
    for s, d in enumerate(dims):
        A = oracle.Problem(*map(int, d))
        ref = oracle.spmv(A, x32[off[s]:off[s + 1]].astype(np.float64)).astype(np.float32)
        assert np.array_equal(y[off[s]:off[s + 1]], ref), s
    assert info["nnz"] == sum((3 * a - 2) * (3 * b - 2) * (3 * c - 2) for a, b, c in dims)


@pytest.mark.parametrize("variant", [-1, 0, 2, 4, 7])
def test_batch_solve_every_system_meets_tolerance(ctx, variant):
    dims = oracle.batch_dims(12, 8, 40, 99)
    info, xs, b = _batch(ctx, dims, 1000)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(info["rows"]))
    res = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=300, variant=variant)
    x = xd.numpy()
    off = info["row_off"]
    assert res["converged"].all()
    for s, d in enumerate(dims):
        A = oracle.Problem(*map(int, d))
        bs, xs_ = b[off[s]:off[s + 1]], x[off[s]:off[s + 1]]
        r = bs - oracle.spmv(A, xs_)
        rel = math.sqrt(oracle.ddot1(r)) / math.sqrt(oracle.ddot1(bs))
        assert rel <= 1e-8, (s, rel)
        assert res["relres"][s] == pytest.approx(rel, rel=1e-4, abs=1e-14)
        _, nit64, _, _, _ = oracle.cg(A, bs, max_iter=1000, tolerance=1e-8 * math.sqrt(oracle.ddot1(bs)))
        assert res["niters"][s] <= 1.3 * nit64 + 10


def test_batch_autotune_and_graph_replay(ctx):
    dims = oracle.batch_dims(16, 16, 48, 7)
    info, xs, b = _batch(ctx, dims, 2000)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(info["rows"]))
    best, ms = ctx.batch_autotune(bd, xd, probe_iters=4)
    assert 0 <= best < len(ms) and ms[best] == ms.min() and np.all(ms > 0)
    assert np.all(xd.numpy() == 0)  # restored
    r1 = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=200)
    x1 = xd.numpy()
    xd.upload(np.zeros(info["rows"]))
    r2 = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=200)  # same graph, replayed
    assert np.array_equal(x1, xd.numpy()) and np.array_equal(r1["niters"], r2["niters"])
    assert r1["converged"].all()


def test_batch_edges(ctx):
    dims = np.array([[1, 1, 1], [3, 1, 2]], np.int32)
    info, xs, b = _batch(ctx, dims, 3)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(info["rows"]))
    res = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=50)
    assert res["converged"].all()
    assert np.max(np.abs(xd.numpy() - xs)) < 1e-7
    xd.upload(np.zeros(info["rows"]))
    res = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=0)
    assert (res["niters"] == 0).all() and not res["converged"].any()


def test_batch_zero_rhs_system_breaks_down_alone(ctx):
    """b_s = 0 makes p.Ap = 0 in that system's first iteration: it stops there (x_s stays 0,
    reported unconverged, the call returns the breakdown warning) while the other systems
    converge as usual and the solve ends once they have."""
    dims = np.array([[6, 5, 4], [7, 3, 5], [8, 8, 2]], np.int32)
    info, xs, b = _batch(ctx, dims, 11)
    off = info["row_off"]
    b[off[1]:off[2]] = 0.0
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(info["rows"]))
    res = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=2000)
    assert res["status"] == hpccg_amd._lib.WARN_BREAKDOWN
    assert res["converged"].tolist() == [1, 0, 1]
    assert res["niters"][1] == 0
    x = xd.numpy()
    assert np.all(x[off[1]:off[2]] == 0.0)
    for s in (0, 2):
        A = oracle.Problem(*map(int, dims[s]))
        r = b[off[s]:off[s + 1]] - oracle.spmv(A, x[off[s]:off[s + 1]])
        assert math.sqrt(oracle.ddot1(r)) <= 1e-8 * math.sqrt(oracle.ddot1(b[off[s]:off[s + 1]]))
        assert res["niters"][s] <= 60


def test_batch_max_iter_2000_plans_its_launches(ctx):
    """max_iter = 2000 (the other solvers' default): launches are sized from the residual rate,
    so the solve costs about what the converged systems need, and a repeated solve (which
    enqueues the previous count at once) gives the same answer bit for bit."""
    dims = oracle.batch_dims(10, 8, 32, 31)
    info, xs, b = _batch(ctx, dims, 3000)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(info["rows"]))
    r1 = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=60)
    x1 = xd.numpy()
    assert r1["converged"].all()
    xd.upload(np.zeros(info["rows"]))
    r2 = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=2000)
    assert r2["converged"].all() and np.array_equal(r1["niters"], r2["niters"])
    assert np.array_equal(xd.numpy(), x1)
    xd.upload(np.zeros(info["rows"]))
    r3 = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=2000)
    assert np.array_equal(xd.numpy(), x1)
    # 2000 captured iterations would cost ~50 us each even when idle
    assert r3["ms"] < 20.0 + 2.0 * r1["ms"], (r1["ms"], r3["ms"])
