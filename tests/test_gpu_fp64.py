"""GPU parity of the FP64 path (libhpccg_hip through its C ABI) against the pinned CPU oracle
and the reference's golden vectors.

Bar: bitwise for generate_matrix, spmv and the waxpby family in both dot modes, and for the
whole cg history in SEQUENTIAL mode.  TREE mode reorders the dots; its per-iteration residual
norms are held to 1e-12 relative against the near-exact-dot restatement (oracle.cg with
exact_dots) while ||r_k|| / ||r_0|| > 1e-10 (below that, histories are rounding noise), and
to the reference's final ||x - x*|| within 1e-13 absolute.
"""
import hashlib
import math

import numpy as np
import pytest

import hpccg_amd
import oracle
from conftest import load_golden

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ctx():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


def upload(ctx, A):
    ctx.upload(A.rowindex, A.colindex, A.values)


# ------------------------------------------------------------------------------ generate_matrix
@pytest.mark.parametrize("case", load_golden("matrix_and_kernels.json")["generate_matrix"],
                         ids=lambda c: "x".join(map(str, c["grid"])))
def test_device_generate_matrix_bitwise(ctx, case):
    out = ctx.generate(*case["grid"], want_host=True)
    assert sha(out["rowindex"]) == case["sha_rowindex"]
    assert sha(out["colindex"]) == case["sha_colindex"]
    assert sha(out["values"]) == case["sha_values"]
    assert sha(out["b"]) == case["sha_b"]
    assert sha(out["x"]) == case["sha_x"]
    assert sha(out["xexact"]) == case["sha_xexact"]
    assert ctx.matrix_info()["nnz"] == case["nnz"]


@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 7, 3), (9, 1, 1), (33, 17, 5), (130, 3, 2)])
def test_device_generate_matches_oracle_ragged(ctx, dims):
    A = oracle.Problem(*dims)
    out = ctx.generate(*dims, want_host=True)
    assert np.array_equal(out["rowindex"], A.rowindex)
    assert np.array_equal(out["colindex"], A.colindex)
    assert np.array_equal(out["values"], A.values)
    assert np.array_equal(out["b"], A.b)


# ------------------------------------------------------------------------------ kernels
@pytest.mark.parametrize("mode", [hpccg_amd.DOT_SEQUENTIAL, hpccg_amd.DOT_TREE])
@pytest.mark.parametrize("case", load_golden("matrix_and_kernels.json")["kernels"],
                         ids=lambda c: "x".join(map(str, c["grid"])))
def test_kernels_vs_reference_golden(ctx, case, mode):
    ctx.set_dot_mode(mode)
    nx, ny, nz = case["grid"]
    n = nx * ny * nz
    A = oracle.Problem(nx, ny, nz)
    upload(ctx, A)
    x = oracle.xstar_uniform(n, case["seed_x"])
    y = oracle.xstar_uniform(n, case["seed_y"])
    a, b = case["alpha"], case["beta"]
    assert sha(ctx.spmv(x)) == case["sha_spmv"]
    assert sha(ctx.waxpby(a, x, b, y)) == case["sha_waxpby"]
    assert sha(ctx.wwpby(b, x, y)) == case["sha_wwpby"]
    assert sha(ctx.wxbw(x, b, y)) == case["sha_wxbw"]
    assert sha(ctx.vcopy(x)) == case["sha_vcopy"]
    d1, d2 = ctx.ddot(x, y), ctx.ddot1(x)
    if mode == hpccg_amd.DOT_SEQUENTIAL:
        assert (d1, d2) == (case["ddot"], case["ddot1"])
    else:
        ex = oracle.ddot_exact(x, y)
        scale = float(np.sum(np.abs(x * y)))
        assert abs(d1 - ex) <= 4 * math.log2(n + 1) * 2 ** -53 * scale
        assert abs(d2 - oracle.ddot_exact(x, x)) <= 4 * math.log2(n + 1) * 2 ** -53 * d2
    ctx.set_dot_mode(hpccg_amd.DOT_TREE)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 255, 256, 257, 1000, 4097, 100003])
def test_dots_ragged_lengths(ctx, n):
    x = oracle.xstar_uniform(n, 7 + n)
    y = oracle.xstar_uniform(n, 9 + n)
    ctx.set_dot_mode(hpccg_amd.DOT_SEQUENTIAL)
    assert ctx.ddot(x, y) == oracle.ddot(x, y)
    assert ctx.ddot1(x) == oracle.ddot1(x)
    ctx.set_dot_mode(hpccg_amd.DOT_TREE)
    assert abs(ctx.ddot(x, y) - oracle.ddot_exact(x, y)) <= 1e-14 * max(1.0, float(np.sum(np.abs(x * y))))
    xi = np.round(x * 1000)
    assert ctx.ddot1(xi) == float(np.sum(xi.astype(np.int64) ** 2))  # exact in any order


@pytest.mark.parametrize("n", [0, 1, 63, 257, 100003, 2_000_003])
def test_compute_residual_exact(ctx, n):
    """max |v1 - v2| equals the oracle's bit for bit (a maximum is order-free); repeated calls do
    not carry the previous result (the reference's SAVEd accumulator is not reproduced)."""
    v1 = oracle.xstar_uniform(n, 31 + n)
    v2 = oracle.xstar_uniform(n, 37 + n)
    assert ctx.compute_residual(v1, v2) == oracle.compute_residual(v1, v2)
    if n:
        small = v2 + 1e-3 * (v1 - v2)
        assert ctx.compute_residual(v1, small) == oracle.compute_residual(v1, small)
        d1 = hpccg_amd.DeviceArray.from_numpy(ctx, v1)
        d2 = hpccg_amd.DeviceArray.from_numpy(ctx, small)
        assert ctx.compute_residual_dev(d1, d2) == oracle.compute_residual(v1, small)
        spike = v1.copy()
        spike[n // 2] += 7.5  # the maximum sits in one element
        r = ctx.compute_residual(spike, v1)
        assert r == oracle.compute_residual(spike, v1) and abs(r - 7.5) < 1e-14


def test_spmv_generic_csr_long_rows(ctx):
    """A CSR that is not a 27-point stencil (rows up to 200 nnz): the per-row fallback path."""
    rng = np.random.default_rng(3)
    n = 3000
    lens = rng.integers(0, 200, n)
    lens[::7] = 0  # empty rows
    rowindex = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) + 1 for l in lens]).astype(np.int32)
    vals = rng.standard_normal(len(cols))
    ctx.upload(rowindex, cols, vals)

    class G:  # minimal oracle.Problem stand-in for oracle.spmv
        pass
    A = G()
    A.n, A.rowindex, A.colindex, A.values = n, rowindex, cols, vals
    x = rng.standard_normal(n)
    assert np.array_equal(ctx.spmv(x), oracle.spmv(A, x))
    assert ctx.matrix_info()["max_row_nnz"] == lens.max()


def test_spmv_generic_csr_keeps_int32_columns(ctx):
    """Thousands of distinct offsets col - row: the sliced copy keeps int32 columns."""
    rng = np.random.default_rng(5)
    n = 2000
    lens = rng.integers(1, 40, n)
    rowindex = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) + 1 for l in lens]).astype(np.int32)
    vals = rng.standard_normal(len(cols))
    ctx.upload(rowindex, cols, vals)
    assert ctx.matrix_format()["ncodes"] == 0
    A = type("P", (), {})()  # minimal oracle.Problem stand-in for oracle.spmv
    A.n, A.rowindex, A.colindex, A.values = n, rowindex, cols, vals
    x = rng.standard_normal(n)
    assert np.array_equal(ctx.spmv(x), oracle.spmv(A, x))


def test_upload_of_a_large_random_csr_falls_back_quickly(ctx):
    """Millions of distinct offsets: collecting them stops after a bounded number of probes per
    row (so the upload ends; the test's own time limit is the guard), and the copy keeps int32
    columns."""
    rng = np.random.default_rng(11)
    n, k = 300_000, 30
    cols = np.sort(rng.integers(0, n, (n, k)), axis=1).astype(np.int32).ravel() + 1
    rowindex = (1 + k * np.arange(n + 1)).astype(np.int32)
    vals = rng.standard_normal(n * k)
    ctx.upload(rowindex, cols, vals)
    assert ctx.matrix_format()["ncodes"] == 0
    A = type("P", (), {})()  # minimal oracle.Problem stand-in for oracle.spmv
    A.n, A.rowindex, A.colindex, A.values = n, rowindex, cols, vals
    x = rng.standard_normal(n)
    assert np.array_equal(ctx.spmv(x), oracle.spmv(A, x))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 4099])
def test_spmv_column_codes_banded_with_empty_rows(ctx, n):
    """At most 256 distinct offsets: one-byte column codes.  A banded CSR with 40 offsets (so
    slices are not 27 wide), rows of 0..40 entries, empty rows and a ragged last slice."""
    rng = np.random.default_rng(n)
    others = np.arange(-300, 301)
    offsets = np.sort(np.concatenate([[0], rng.choice(others[others != 0], size=39, replace=False)]))
    rows = []
    for i in range(n):
        c = i + offsets
        c = c[(c >= 0) & (c < n)]
        keep = rng.random(len(c)) < (0.0 if i % 11 == 3 else 0.8)
        keep |= (c == i) & (i % 11 != 3)  # every row but the empty ones holds its diagonal
        rows.append(c[keep])
    lens = np.array([len(r) for r in rows])
    rowindex = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    cols = (np.concatenate(rows) + 1).astype(np.int32)
    vals = rng.standard_normal(len(cols))
    assert len(cols) > 0
    ctx.upload(rowindex, cols, vals)
    fmt = ctx.matrix_format()
    assert 1 <= fmt["ncodes"] <= 40
    A = type("P", (), {})()
    A.n, A.rowindex, A.colindex, A.values = n, rowindex, cols, vals
    x = rng.standard_normal(n)
    assert np.array_equal(ctx.spmv(x), oracle.spmv(A, x))


@pytest.mark.parametrize("dims", [(130, 5, 4), (256, 4, 3), (64, 3, 3)])
def test_spmv_uniform_code_slices_bitwise(ctx, dims):
    """Grids whose x-lines hold whole 64-row slices of interior rows: those slices read one
    shared code sequence.  The product must stay the reference's bit for bit."""
    A = oracle.Problem(*dims)
    ctx.upload(A.rowindex, A.colindex, A.values)
    assert ctx.matrix_format()["ncodes"] == 27
    x = np.random.default_rng(sum(dims)).standard_normal(A.n)
    assert np.array_equal(ctx.spmv(x), oracle.spmv(A, x))


def test_generated_matrix_uses_27_column_codes(ctx):
    """The 27-point matrix is stored with 27 offset codes: 9 bytes per slot instead of 12."""
    nx, ny, nz = 20, 30, 10
    out = ctx.generate(nx, ny, nz, want_host=True)
    fmt = ctx.matrix_format()
    assert fmt["ncodes"] == 27
    n = nx * ny * nz
    ns = (n + 63) // 64
    lens = np.zeros(ns * 64, np.int64)
    lens[:n] = np.diff(out["rowindex"])
    slots = 64 * int(lens.reshape(ns, 64).max(axis=1).sum())  # slice width = its longest row
    assert fmt["copy_bytes"] == slots * (8 + 1) + 192 * ns + (ns + 1) * 8


def test_upload_rejects_bad_csr(ctx):
    with pytest.raises(hpccg_amd.HpccgError):
        ctx.upload(np.array([0, 1], np.int32), np.array([1], np.int32), np.array([1.0]))  # 0-based
    with pytest.raises(hpccg_amd.HpccgError):
        ctx.upload(np.array([1, 2], np.int32), np.array([5], np.int32), np.array([1.0]))  # col out of range


# ------------------------------------------------------------------------------ cg
def _golden_problem(g):
    nx, ny, nz = g["grid"]
    A = oracle.Problem(nx, ny, nz)
    if "seed" in g:
        xs = oracle.xstar_sines(nx, ny, nz, g["seed"])
        b = oracle.spmv(A, xs)
    else:
        xs, b = A.xexact, A.b
    return A, b, xs


@pytest.mark.parametrize("fname", ["cg_hist_5x5x5_ones.json", "cg_hist_16x8x12_ones.json",
                                   "cg_hist_32x32x32_ones.json", "cg_hist_32x32x32_seed11.json",
                                   "cg_hist_128x128x128_ones.json"])
def test_cg_sequential_mode_bitwise_vs_reference(ctx, fname):
    g = load_golden(fname)
    A, b, xs = _golden_problem(g)
    upload(ctx, A)
    ctx.set_dot_mode(hpccg_amd.DOT_SEQUENTIAL)
    try:
        x, nit, normr, hist, st = ctx.cg(b, max_iter=g["kmax"])
    finally:
        ctx.set_dot_mode(hpccg_amd.DOT_TREE)
    assert st == 0 and nit == g["kmax"]
    assert hist.tolist() == g["hist"]
    assert normr == g["hist"][-2]
    assert oracle.compute_residual(x, xs) == g["xerr_max"]


@pytest.mark.parametrize("fname", ["cg_hist_32x32x32_ones.json", "cg_hist_32x32x32_seed11.json",
                                   "cg_hist_16x8x12_ones.json", "cg_hist_128x128x128_ones.json"])
def test_cg_tree_mode_within_1e12(ctx, fname):
    g = load_golden(fname)
    A, b, xs = _golden_problem(g)
    upload(ctx, A)
    x, nit, normr, hist, st = ctx.cg(b, max_iter=g["kmax"])
    assert st == 0 and nit == g["kmax"]
    ref = np.array(g["hist"])
    if A.n <= 64 ** 3:
        _, _, _, hex_, _ = oracle.cg(A, b, max_iter=g["kmax"], exact_dots=True)
    else:
        # the exact-dot history at 128^3 is a committed fixture (oracle.cg(exact_dots=True), 15 s on a CPU)
        ex = load_golden("cg_exacthist_{}_ones.json".format("x".join(map(str, g["grid"]))))
        assert ex["kmax"] == g["kmax"] and "seed" not in g
        hex_ = np.array(ex["hist"])
    live = ref / ref[0] > 1e-10
    dev = np.max(np.abs(hist[live] - hex_[live]) / hex_[live])
    assert dev <= 1e-12, dev
    # against the sequential reference the gap is the reference's own summation error (~N eps)
    dev_ref = np.max(np.abs(hist[live] - ref[live]) / ref[live])
    assert dev_ref <= max(1e-12, 4 * A.n * 2 ** -53), dev_ref
    assert abs(oracle.compute_residual(x, xs) - g["xerr_max"]) <= 1e-13


@pytest.mark.parametrize("mode", [hpccg_amd.DOT_SEQUENTIAL, hpccg_amd.DOT_TREE])
def test_cg_solve_to_tolerance(ctx, mode):
    for c in load_golden("cg_solve_sines.json")["cases"]:
        nx, ny, nz = c["grid"]
        if nx > 64:
            continue
        A = oracle.Problem(nx, ny, nz)
        xs = oracle.xstar_sines(nx, ny, nz, c["seed"])
        b = oracle.spmv(A, xs)
        upload(ctx, A)
        ctx.set_dot_mode(mode)
        x, nit, normr, hist, st = ctx.cg(b, max_iter=c["max_iter"], tolerance=c["abs_tol"])
        ctx.set_dot_mode(hpccg_amd.DOT_TREE)
        if mode == hpccg_amd.DOT_SEQUENTIAL:
            assert (nit, normr) == (c["niters"], c["normr"])
            assert float(np.max(np.abs(x - xs))) == c["xerr_max"]
        else:
            assert abs(nit - c["niters"]) <= 1
            assert normr < c["abs_tol"] * 1e3
            r = b - oracle.spmv(A, x)
            assert math.sqrt(oracle.ddot1(r)) / c["normb"] <= 2e-10


def test_cg_edge_cases(ctx):
    A = oracle.Problem(6, 5, 4)
    upload(ctx, A)
    for mode in (hpccg_amd.DOT_SEQUENTIAL, hpccg_amd.DOT_TREE):
        ctx.set_dot_mode(mode)
        # max_iter = 0: Prepare only; normr = ||r_0||, x untouched
        x, nit, normr, hist, st = ctx.cg(A.b, max_iter=0)
        assert nit == 0 and normr == math.sqrt(oracle.ddot1(A.b)) and np.all(x == 0)
        # tolerance above ||r_0||: converged at the top of iteration 1
        x, nit, normr, hist, st = ctx.cg(A.b, max_iter=50, tolerance=1e9)
        assert nit == 0 and np.all(x == 0)
        # nonzero x0 (the reference's cg starts from the x passed in)
        x0 = oracle.xstar_uniform(A.n, 5)
        xr, nr_, normr_r, hr, _ = oracle.cg(A, A.b, x0=x0, max_iter=20)
        x, nit, normr, hist, st = ctx.cg(A.b, x0=x0, max_iter=20)
        if mode == hpccg_amd.DOT_SEQUENTIAL:
            assert np.array_equal(x, xr) and hist.tolist() == hr.tolist()
        else:
            assert np.max(np.abs(x - xr)) < 1e-10
        # a start of negative zeros: Prepare skips A x as for +0 and the reference, which forms
        # A x, gets the same r = b (every row sum of A (-0) is +0.0)
        xr0, _, _, hr0, _ = oracle.cg(A, A.b, x0=np.full(A.n, -0.0), max_iter=20)
        x, nit, normr, hist, st = ctx.cg(A.b, x0=np.full(A.n, -0.0), max_iter=20)
        if mode == hpccg_amd.DOT_SEQUENTIAL:
            assert np.array_equal(x, xr0) and hist.tolist() == hr0.tolist()
        # breakdown: b = 0 gives p.Ap = 0 at iteration 1 -> NaN, reported as a warning
        x, nit, normr, hist, st = ctx.cg(np.zeros(A.n), max_iter=3)
        assert st == hpccg_amd._lib.WARN_BREAKDOWN
    ctx.set_dot_mode(hpccg_amd.DOT_TREE)


def test_cg_lagging_test_boundary(ctx):
    """Tolerance placed exactly at / just above ||r_j||: the exit iteration follows the
    reference's `normr < tolerance` at the top of iteration j+2 (device sqrt must round
    correctly for this to match)."""
    g = load_golden("cg_hist_16x8x12_ones.json")
    A = oracle.Problem(*g["grid"])
    upload(ctx, A)
    ctx.set_dot_mode(hpccg_amd.DOT_SEQUENTIAL)
    try:
        for j in (3, 10, 25):
            h = g["hist"][j]
            _, nit, normr, _, _ = ctx.cg(A.b, max_iter=150, tolerance=h)  # not < : keeps going
            _, nit_o, normr_o, _, _ = oracle.cg(A, A.b, max_iter=150, tolerance=h)
            assert (nit, normr) == (nit_o, normr_o)
            t = np.nextafter(h, np.inf)
            _, nit, normr, _, _ = ctx.cg(A.b, max_iter=150, tolerance=t)
            assert nit == j + 1 and normr == h
    finally:
        ctx.set_dot_mode(hpccg_amd.DOT_TREE)


def test_cg_device_resident_and_graph_reuse(ctx):
    A = oracle.Problem(32, 32, 32)
    upload(ctx, A)
    b = hpccg_amd.DeviceArray.from_numpy(ctx, A.b)
    x = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
    r1 = ctx.cg_dev(b, x, max_iter=150)
    x1 = x.numpy()
    x.upload(np.zeros(A.n))
    r2 = ctx.cg_dev(b, x, max_iter=150)  # replays the cached graph
    assert np.array_equal(x1, x.numpy())
    assert r1["hist"].tolist() == r2["hist"].tolist()
    assert r1["ms"] > 0
