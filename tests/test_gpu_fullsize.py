"""Full-size GPU parity: the benchmark workloads at their own sizes, every answer checked by the
FP64 oracle (oracle.spmv forms the true residual b - A x on the CPU) or by the reference's own
fixtures (tests/golden/cg_solve_sines.json: the reference cg solving the same seeded systems).

  * 256^3 / 128^3 generate_matrix on the device: every CSR array equal to the oracle's
    (SparseMatrix.F90:138-175).
  * FP64 CG (tree dots) at 128^3 (seed 13) and 256^3 (seed 14): the reference's iteration count,
    its max|x - x*| to 1e-3 relative, true relative residual <= 1e-10; SEQUENTIAL dots at 128^3:
    niters, normr and max|x - x*| every bit the reference's (numerics.F90:108-147).
  * mixed-precision CG at 256^3 (seed 14): true relative residual <= 1e-10 and max|x - x*|
    within 2x of the reference FP64 solve's.
  * learned flexible CG at 128^3 and 256^3 with the shipped network: true relative residual
    <= 1e-10, no fall-back, at most a third of the reference's iterations, max|x - x*| within 2x.
  * 64 batched systems, extents U{16..96} (seed 2024), x*_s = U(-1,1) (seed 5000 + s),
    b_s = A_s x*_s in FP64: every system converged to 1e-8 in its own true residual.

Host memory: the 256^3 oracle problem holds 5.4 GB (briefly 11 GB while the device CSR is
compared); each case builds its problem once per module.
"""
import math

import numpy as np
import pytest

import hpccg_amd
import oracle
from conftest import load_golden

pytestmark = pytest.mark.gpu

SINES = {c["grid"][0]: c for c in load_golden("cg_solve_sines.json")["cases"]}


@pytest.fixture(scope="module")
def ctx():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


class _Cache:
    """One full-size problem at a time: (n, oracle problem, x*, b, device b, device x)."""
    key = None
    val = None


@pytest.fixture(scope="module")
def cache():
    yield _Cache
    _Cache.key = _Cache.val = None


def problem(ctx, cache, n):
    if cache.key != n:
        cache.key = cache.val = None
        c = SINES[n]
        A = oracle.Problem(n, n, n)
        xs = oracle.xstar_sines(n, n, n, c["seed"])
        b = oracle.spmv(A, xs)
        import hashlib
        assert hashlib.sha256(b.tobytes()).hexdigest() == c["sha_b"]  # the reference's b = A x*
        ctx.generate(n, n, n)
        cache.val = {"A": A, "xs": xs, "b": b, "c": c, "normb": math.sqrt(oracle.ddot1(b)),
                     "bd": hpccg_amd.DeviceArray.from_numpy(ctx, b), "xd": hpccg_amd.DeviceArray(ctx, A.n)}
        cache.key = n
    P = cache.val
    P["xd"].upload(np.zeros(P["A"].n))
    return P


def true_relres(P, x):
    r = P["b"] - oracle.spmv(P["A"], x)
    return math.sqrt(oracle.ddot1(r)) / P["normb"]


def xerr(P, x):
    return float(np.max(np.abs(x - P["xs"])))


@pytest.mark.parametrize("n", [128, 256])
def test_device_generate_matrix_full_size_equals_oracle(ctx, n):
    A = oracle.Problem(n, n, n)
    out = ctx.generate(n, n, n, want_host=True)
    assert A.nnz == (3 * n - 2) ** 3 == ctx.matrix_info()["nnz"]
    for k in ("rowindex", "colindex", "values", "b", "x", "xexact"):
        ref = getattr(A, k)
        assert np.array_equal(out[k], ref), k
    assert ctx.matrix_format()["ncodes"] == 27


@pytest.mark.parametrize("n", [128, 256])
def test_fp64_cg_tree_full_size_vs_reference_solve(ctx, cache, n):
    P = problem(ctx, cache, n)
    c = P["c"]
    res = ctx.cg_dev(P["bd"], P["xd"], max_iter=c["max_iter"], tolerance=c["abs_tol"], want_hist=False)
    x = P["xd"].numpy()
    assert res["status"] == 0
    assert res["niters"] == c["niters"], (res["niters"], c["niters"])
    rel = true_relres(P, x)
    assert rel <= 1e-10, rel
    assert abs(rel - c["true_relres"]) <= 0.05 * c["true_relres"], (rel, c["true_relres"])
    assert abs(xerr(P, x) - c["xerr_max"]) <= 1e-3 * c["xerr_max"], (xerr(P, x), c["xerr_max"])


def test_fp64_cg_sequential_128_bitwise_vs_reference_solve(ctx, cache):
    P = problem(ctx, cache, 128)
    c = P["c"]
    ctx.set_dot_mode(hpccg_amd.DOT_SEQUENTIAL)
    try:
        res = ctx.cg_dev(P["bd"], P["xd"], max_iter=c["max_iter"], tolerance=c["abs_tol"], want_hist=False)
    finally:
        ctx.set_dot_mode(hpccg_amd.DOT_TREE)
    x = P["xd"].numpy()
    assert (res["niters"], res["normr"]) == (c["niters"], c["normr"])
    assert xerr(P, x) == c["xerr_max"]
    assert math.sqrt(oracle.ddot1(x - P["xs"])) == c["xerr_l2"]


def test_mixed_cg_256_full_size(ctx, cache):
    P = problem(ctx, cache, 256)
    c = P["c"]
    res = ctx.cg_mixed_dev(P["bd"], P["xd"], max_iter=2000, tolerance=c["abs_tol"], replace_every=25,
                           want_hist=False)
    x = P["xd"].numpy()
    assert res["status"] == 0
    rel = true_relres(P, x)
    assert rel <= 1e-10, rel
    assert abs(res["true_normr"] / P["normb"] - rel) <= 1e-3 * rel  # it reports the true residual
    assert xerr(P, x) <= 2 * c["xerr_max"], (xerr(P, x), c["xerr_max"])
    assert res["niters"] <= 1.1 * c["niters"] + 5, (res["niters"], c["niters"])


@pytest.mark.parametrize("n", [128, 256])
def test_learned_fcg_full_size(ctx, cache, n):
    from hpccg_amd import learned
    w = learned.load()
    assert w is not None
    ctx.cnn_set_weights(w)
    P = problem(ctx, cache, n)
    c = P["c"]
    res = ctx.fcg_learned_dev((n, n, n), P["bd"], P["xd"], max_iter=2000, tolerance=c["abs_tol"],
                              replace_every=25)
    x = P["xd"].numpy()
    assert res["status"] == 0 and not res["fell_back"], res
    rel = true_relres(P, x)
    assert rel <= 1e-10, rel
    assert res["fcg_iters"] <= c["niters"] / 3, (res["fcg_iters"], c["niters"])
    assert xerr(P, x) <= 2 * c["xerr_max"], (xerr(P, x), c["xerr_max"])


def test_batched_64_systems_full_size(ctx):
    dims = oracle.batch_dims(64, 16, 96, 2024)
    assert np.array_equal(dims, hpccg_amd.batch_dims(64, 16, 96, 2024))
    info = ctx.batch_setup(dims)
    off = info["row_off"]
    assert info["count"] == 64 and info["nnz"] == sum((3 * a - 2) * (3 * b - 2) * (3 * c - 2) for a, b, c in dims)
    xs, bs, probs = [], [], []
    for s, d in enumerate(dims):
        A = oracle.Problem(*map(int, d))
        x = oracle.xstar_uniform(A.n, 5000 + s)
        xs.append(x)
        bs.append(oracle.spmv(A, x))
        probs.append(A)
    b = np.concatenate(bs)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(info["rows"]))
    res = ctx.batch_solve_dev(bd, xd, rel_tol=1e-8, max_iter=2000)
    x = xd.numpy()
    assert res["converged"].all(), np.flatnonzero(res["converged"] == 0)
    for s, A in enumerate(probs):
        xs_ = x[off[s]:off[s + 1]]
        r = bs[s] - oracle.spmv(A, xs_)
        rel = math.sqrt(oracle.ddot1(r)) / math.sqrt(oracle.ddot1(bs[s]))
        assert rel <= 1e-8, (s, dims[s], rel)
        assert res["relres"][s] == pytest.approx(rel, rel=1e-3, abs=1e-14)
        assert 20 <= res["niters"][s] <= 60, (s, res["niters"][s])
