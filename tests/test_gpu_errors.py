"""GPU tests of the C ABI's error contract.  The reference has no error returns (its cg, spmv
and generate_matrix, numerics.F90:41-162 and SparseMatrix.F90:68-232, assume a resident, well
formed matrix); here a call made in the wrong order, or with an argument outside what the
reference could pass, returns HPCCG_ERR_STATE (-3) or HPCCG_ERR_ARG (-1) with a message in
hpccg_last_error before any kernel is launched, and the context stays usable afterwards."""
import numpy as np
import pytest

import hpccg_amd
import oracle
from hpccg_amd import learned

pytestmark = pytest.mark.gpu

STATE, ARG = "(-3)", "(-1)"


@pytest.fixture()
def fresh():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


def _raises(code, words, fn, *a, **k):
    with pytest.raises(hpccg_amd.HpccgError) as e:
        fn(*a, **k)
    msg = str(e.value)
    assert code in msg and words in msg, msg


def _still_usable(ctx):
    A = oracle.Problem(6, 5, 4)
    ctx.upload(A.rowindex, A.colindex, A.values)
    x = np.random.default_rng(0).standard_normal(A.n)
    assert np.array_equal(ctx.spmv(x), oracle.spmv(A, x))


def test_solver_calls_before_a_matrix_is_resident(fresh):
    n = 64
    b = hpccg_amd.DeviceArray.from_numpy(fresh, np.ones(n))
    x = hpccg_amd.DeviceArray.from_numpy(fresh, np.zeros(n))
    _raises(STATE, "no matrix", fresh.cg, np.ones(n), max_iter=5)
    _raises(STATE, "no matrix", fresh.cg_mixed, np.ones(n), max_iter=5)
    _raises(STATE, "no matrix", fresh.cg_dev, b, x, max_iter=5)
    _raises(STATE, "no matrix", fresh.cg_dev_noalloc, b, x, max_iter=5)
    _raises(STATE, "no matrix", fresh.cg_mixed_dev, b, x, max_iter=5)
    _raises(STATE, "no matrix", fresh.spmv_dev, b, x)
    _raises(STATE, "no matrix", fresh.fcg_learned_dev, (4, 4, 4), b, x, max_iter=5)
    assert np.array_equal(x.numpy(), np.zeros(n))  # nothing was written
    _still_usable(fresh)


def test_network_calls_before_weights(fresh):
    dims = (8, 8, 8)
    A = oracle.Problem(*dims)
    fresh.upload(A.rowindex, A.colindex, A.values)
    r = hpccg_amd.DeviceArray.from_numpy(fresh, np.ones(A.n, np.float32))
    z = hpccg_amd.DeviceArray.from_numpy(fresh, np.zeros(A.n, np.float32))
    b = hpccg_amd.DeviceArray.from_numpy(fresh, A.b)
    x = hpccg_amd.DeviceArray.from_numpy(fresh, np.zeros(A.n))
    _raises(STATE, "no network", fresh.cnn_apply_dev, dims, r, 1.0, z)
    _raises(STATE, "no network", fresh.fcg_learned_dev, dims, b, x, max_iter=5)
    _raises(STATE, "no network", fresh.cnn_get_weights)
    _still_usable(fresh)


def test_flexible_cg_rejects_a_grid_that_is_not_the_matrix(fresh):
    w = learned.load()
    assert w is not None
    fresh.cnn_set_weights(w)
    n = 16
    A = oracle.Problem(n, n, n)
    fresh.upload(A.rowindex, A.colindex, A.values)
    b = hpccg_amd.DeviceArray.from_numpy(fresh, A.b)
    x = hpccg_amd.DeviceArray.from_numpy(fresh, np.zeros(A.n))
    _raises(ARG, "grid does not match", fresh.fcg_learned_dev, (n, n, n + 1), b, x, max_iter=5)
    _raises(ARG, "bad argument", fresh.fcg_learned_dev, (n, n, n), b, x, max_iter=5, window=0)
    _raises(ARG, "bad argument", fresh.cnn_apply_dev, (n, n, n), x, 0.0, b)  # inv_scale must be > 0
    _raises(ARG, "bad argument", fresh.cnn_apply_dev, (0, n, n), x, 1.0, b)
    assert np.array_equal(x.numpy(), np.zeros(A.n))
    tol = 1e-10 * np.linalg.norm(A.b)
    res = fresh.fcg_learned_dev((n, n, n), b, x, max_iter=60, tolerance=tol)
    assert res["fcg_iters"] + res["cg_iters"] <= 60 and res["true_normr"] < tol, res


def test_batch_calls_before_and_after_setup(fresh):
    L = hpccg_amd._lib.lib()
    b = hpccg_amd.DeviceArray.from_numpy(fresh, np.ones(8))
    x = hpccg_amd.DeviceArray.from_numpy(fresh, np.zeros(8))
    assert L.hpccg_batch_solve_dev(fresh.h, b.ptr, x.ptr, 1e-8, 10, 25, -1, None, None, None, None) == -3
    assert b"no batch" in L.hpccg_last_error()
    _raises(STATE, "no batch", fresh.batch_info)
    _raises(ARG, "bad argument", fresh.batch_setup, np.zeros((0, 3), np.int32))
    _raises(ARG, "extents", fresh.batch_setup, [[8, 8, 8], [8, 0, 8]])
    info = fresh.batch_setup([[8, 8, 8], [9, 10, 8]])
    b = hpccg_amd.DeviceArray.from_numpy(fresh, np.ones(info["rows"]))
    x = hpccg_amd.DeviceArray.from_numpy(fresh, np.zeros(info["rows"]))
    nv = len(hpccg_amd.Context.batch_variants())
    _raises(ARG, "variant out of range", fresh.batch_solve_dev, b, x, variant=nv)
    _raises(ARG, "bad argument", fresh.batch_solve_dev, b, x, max_iter=-1)
    res = fresh.batch_solve_dev(b, x, rel_tol=1e-8, max_iter=100)
    assert list(res["converged"]) == [1, 1]


def test_bad_modes_and_counts(fresh):
    A = oracle.Problem(5, 4, 3)
    fresh.upload(A.rowindex, A.colindex, A.values)
    _raises(ARG, "unknown dot mode", fresh.set_dot_mode, 7)
    _raises(ARG, "max_iter", fresh.cg, A.b, max_iter=-1)
    _raises(ARG, "max_iter", fresh.cg_mixed, A.b, max_iter=-1)
    x = hpccg_amd.DeviceArray.from_numpy(fresh, np.ones(A.n))
    y = hpccg_amd.DeviceArray.from_numpy(fresh, np.zeros(A.n))
    _raises(ARG, "bad argument", fresh.time_spmv, x, y, reps=0)
    _raises(ARG, "bad argument", fresh.time_copy, y, x, reps=0)
    _raises(ARG, "level", fresh.cg_timed, A.b, max_iter=5, level=0)
    _still_usable(fresh)
