"""Pins the CPU oracle (oracle/hpccg_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible restatement mistake (a dropped stencil term, a wrong
row order, a transposed operand, a reordered sum) fails at least one of them: golden vectors
from the reference's own Fortran build (tests/golden, bitwise), closed forms, the Kronecker
structure of the 27-point matrix, and CG's finite-termination property.
"""
import glob
import hashlib
import math
import os

import numpy as np
import pytest

import oracle
from conftest import GOLD, load_golden


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tridiag_ones(n):
    return np.eye(n) + np.eye(n, k=1) + np.eye(n, k=-1)


# ----------------------------------------------------------------------------- generate_matrix
@pytest.mark.parametrize("case", load_golden("matrix_and_kernels.json")["generate_matrix"],
                         ids=lambda c: "x".join(map(str, c["grid"])))
def test_generate_matrix_bitwise_vs_reference(case):
    A = oracle.Problem(*case["grid"])
    assert A.n == case["n"] and A.nnz == case["nnz"]
    assert sha(A.rowindex) == case["sha_rowindex"]
    assert sha(A.colindex) == case["sha_colindex"]
    assert sha(A.values) == case["sha_values"]
    assert sha(A.b) == case["sha_b"]
    assert sha(A.x) == case["sha_x"]
    assert sha(A.xexact) == case["sha_xexact"]
    # SparseMatrix.F90:111: the reference reports the nominal 27*N
    assert case["total_nnz"] == 27 * case["n"]


@pytest.mark.parametrize("dims", [(3, 4, 5), (32, 32, 32), (17, 9, 4), (1, 1, 1), (1, 5, 3)])
def test_closed_forms(dims):
    nx, ny, nz = dims
    A = oracle.Problem(*dims)
    assert A.nnz == (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
    rowlen = np.diff(A.rowindex)
    # each row length is the product of the 1-D neighbour counts; b_i = 28 - nnz_i
    c = lambda n: np.array([min(i + 1, n - 1) - max(i - 1, 0) + 1 for i in range(n)])
    expect = (c(nz)[:, None, None] * c(ny)[None, :, None] * c(nx)[None, None, :]).ravel()
    assert np.array_equal(rowlen, expect)
    assert np.array_equal(A.b, 28.0 - rowlen)
    # ||b||^2 is an exact integer sum, order independent
    assert oracle.ddot1(A.b) == float(np.sum((28 - rowlen).astype(np.int64) ** 2))
    # columns ascend strictly within each row, 1-based and in range
    for i in range(0, A.n, max(1, A.n // 97)):
        cols = A.colindex[A.rowindex[i] - 1: A.rowindex[i + 1] - 1]
        assert np.all(np.diff(cols) > 0) and cols[0] >= 1 and cols[-1] <= A.n
    # A*1 = b exactly
    assert np.array_equal(oracle.spmv(A, np.ones(A.n)), A.b)


def test_cube_sizes_table():
    # closed forms at 32^3: nnz = (3n-2)^3 = 830,584, ||b||^2 = 662,360
    A = oracle.Problem(32, 32, 32)
    assert A.nnz == 830584
    assert oracle.ddot1(A.b) == 662360.0
    assert repr(math.sqrt(662360.0)) == "813.8550239446827"


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 4), (5, 5, 5), (4, 3, 2)])
def test_dense_kronecker_identity(dims):
    """A = 28 I - T_z (x) T_y (x) T_x with T = tridiag(1,1,1), row index ix fastest."""
    nx, ny, nz = dims
    A = oracle.Problem(*dims)
    D = np.zeros((A.n, A.n))
    for i in range(A.n):
        for k in range(A.rowindex[i] - 1, A.rowindex[i + 1] - 1):
            D[i, A.colindex[k] - 1] = A.values[k]
    K = np.kron(tridiag_ones(nz), np.kron(tridiag_ones(ny), tridiag_ones(nx)))
    assert np.array_equal(D, 28.0 * np.eye(A.n) - K)


@pytest.mark.parametrize("dims,modes", [((12, 10, 8), (1, 2, 3)), ((9, 9, 9), (9, 1, 5)), ((32, 32, 32), (31, 2, 17))])
def test_spmv_kronecker_eigenpairs(dims, modes):
    A = oracle.Problem(*dims)
    vs, lam = [], 1.0
    for n, j in zip(dims, modes):
        i = np.arange(1, n + 1)
        vs.append(np.sin(np.pi * j * i / (n + 1)))
        lam *= 1.0 + 2.0 * np.cos(np.pi * j / (n + 1))
    lam = 28.0 - lam
    v = (vs[2][:, None, None] * vs[1][None, :, None] * vs[0][None, None, :]).ravel()
    r = oracle.spmv(A, v) - lam * v
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(v) * 37


# ----------------------------------------------------------------------------- vector kernels
@pytest.mark.parametrize("case", load_golden("matrix_and_kernels.json")["kernels"],
                         ids=lambda c: "x".join(map(str, c["grid"])))
def test_kernels_bitwise_vs_reference(case):
    nx, ny, nz = case["grid"]
    n = nx * ny * nz
    A = oracle.Problem(nx, ny, nz)
    x = oracle.xstar_uniform(n, case["seed_x"])
    y = oracle.xstar_uniform(n, case["seed_y"])
    a, b = case["alpha"], case["beta"]
    outs = {"spmv": oracle.spmv(A, x), "waxpby": oracle.waxpby(a, x, b, y), "wwpby": oracle.wwpby(b, x, y),
            "wxbw": oracle.wxbw(x, b, y), "vcopy": x.copy()}
    for k, v in outs.items():
        assert sha(v) == case["sha_" + k], k
        if k in case:
            assert np.array_equal(v, np.array(case[k])), k
    assert oracle.ddot(x, y) == case["ddot"]
    assert oracle.ddot1(x) == case["ddot1"]


def test_ddot_exact_vs_fsum():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(50000) * np.exp(rng.uniform(-20, 20, 50000))
    y = rng.standard_normal(50000)
    ref = math.fsum(float(a) * float(b) for a, b in zip(x.tolist(), y.tolist()))  # products rounded
    # fsum of rounded products differs from the exact dot by at most the product roundings
    bound = sum(abs(a * b) for a, b in zip(x.tolist(), y.tolist())) * 2.0 ** -53
    assert abs(oracle.ddot_exact(x, y) - ref) <= bound
    # an integer-valued dot is exact in every order
    xi = rng.integers(-1000, 1000, 10000).astype(np.float64)
    assert oracle.ddot_exact(xi, xi) == float(np.sum(xi.astype(np.int64) ** 2)) == oracle.ddot1(xi)


# ----------------------------------------------------------------------------- cg
HIST_FILES = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLD, "cg_hist_*.json")))


@pytest.mark.parametrize("fname", HIST_FILES)
def test_cg_history_bitwise_vs_reference(fname):
    g = load_golden(fname)
    nx, ny, nz = g["grid"]
    A = oracle.Problem(nx, ny, nz)
    b = A.b if "seed" not in g else oracle.spmv(A, oracle.xstar_sines(nx, ny, nz, g["seed"]))
    if "sha_b" in g:
        assert sha(b) == g["sha_b"]
    x, nit, normr, hist, st = oracle.cg(A, b, max_iter=g["kmax"])
    assert st == 0 and nit == g["kmax"]
    assert hist.tolist() == g["hist"]  # ||r_0|| .. ||r_kmax||, every bit
    xs = A.xexact if "seed" not in g else oracle.xstar_sines(nx, ny, nz, g["seed"])
    assert oracle.compute_residual(x, xs) == g["xerr_max"]
    assert normr == g["hist"][-2]  # cg returns ||r_{kmax-1}||


@pytest.mark.parametrize("case", load_golden("cg_solve_sines.json")["cases"],
                         ids=lambda c: "x".join(map(str, c["grid"])))
def test_cg_solve_to_tolerance_vs_reference(case):
    """Seeded solves to 1e-10 up to 256^3 (the 256^3 case holds ~11 GB of host memory briefly and
    takes ~50 s): niters, normr and max|x - x*| every bit equal to the reference's."""
    for c in [case]:
        nx, ny, nz = c["grid"]
        A = oracle.Problem(nx, ny, nz)
        xs = oracle.xstar_sines(nx, ny, nz, c["seed"])
        b = oracle.spmv(A, xs)
        assert sha(b) == c["sha_b"]
        x, nit, normr, _, st = oracle.cg(A, b, max_iter=c["max_iter"], tolerance=c["abs_tol"])
        assert (nit, normr) == (c["niters"], c["normr"])
        assert float(np.max(np.abs(x - xs))) == c["xerr_max"]


def test_cg_finite_termination():
    """b spanning m eigenvectors with distinct eigenvalues: CG solves it in m iterations."""
    dims = (10, 9, 8)
    A = oracle.Problem(*dims)

    def eig(modes):
        vs, lam = [], 1.0
        for n, j in zip(dims, modes):
            i = np.arange(1, n + 1)
            vs.append(np.sin(np.pi * j * i / (n + 1)))
            lam *= 1.0 + 2.0 * np.cos(np.pi * j / (n + 1))
        return (vs[2][:, None, None] * vs[1][None, :, None] * vs[0][None, None, :]).ravel(), 28.0 - lam

    for m, modes in [(1, [(1, 1, 1)]), (3, [(1, 1, 1), (2, 5, 3), (7, 2, 6)])]:
        b = sum(eig(md)[0] for md in modes)
        x, nit, normr, hist, st = oracle.cg(A, b, max_iter=m + 1)
        assert hist[m] <= 1e-11 * hist[0], (m, hist)
        assert hist[m - 1] > 1e-3 * hist[0]


def _np_cg(A, b, kmax, mistake=None):
    """numpy restatement of numerics.F90:92-147 with an optional planted mistake."""
    def spmv(v):
        y = oracle.spmv(A, v)
        if mistake == "dropped_term":  # last stencil term of every row lost
            last = A.rowindex[1:] - 2
            y = y - A.values[last] * v[A.colindex[last] - 1]
        return y
    x = np.zeros(A.n)
    r = b - spmv(x) if mistake != "sign" else b + spmv(x)
    rtrans = oracle.ddot1(r)
    hist = [np.sqrt(rtrans)]
    p = r.copy()
    for k in range(1, kmax + 1):
        if k > 1:
            old = rtrans
            rtrans = oracle.ddot1(r)
            beta = rtrans / old
            p = beta * r + p if mistake == "transposed" else r + beta * p
            hist.append(np.sqrt(rtrans))
        Ap = spmv(p)
        alpha = rtrans / oracle.ddot(p, Ap)
        r = r + (-alpha) * Ap if mistake != "sign" else r + alpha * Ap
        x = x + alpha * p
    return hist


@pytest.mark.parametrize("mistake", [None, "transposed", "dropped_term", "sign"])
def test_planted_mistakes_are_caught(mistake):
    """The bitwise history pin accepts the faithful restatement and rejects each planted mistake."""
    g = load_golden("cg_hist_16x8x12_ones.json")
    A = oracle.Problem(*g["grid"])
    hist = _np_cg(A, A.b, 30, mistake)
    ref = g["hist"][:30]
    if mistake is None:
        assert hist == ref
    else:
        dev = max(abs(a - b) / b for a, b in zip(hist, ref))
        assert dev > 1e-6, (mistake, dev)
    # a reordered (near-exact) dot also leaves the bitwise history
    if mistake is None:
        _, _, _, hist_exact, _ = oracle.cg(A, max_iter=g["kmax"], exact_dots=True)
        assert hist_exact.tolist() != g["hist"]


def test_exact_dot_history_fixture_is_the_reordered_yardstick():
    """The committed exact-dot 128^3 history (oracle.cg with exact_dots, tests/golden/make_golden.py
    --exact) starts at the closed-form ||b|| and departs from the reference's sequential-sum
    history by the sequential sum's own error: about N eps / 2 relative (1.6e-10 at 128^3), not
    more, while ||r_k|| / ||r_0|| > 1e-10; and it agrees to 5e-13 at 32^3, where N eps is small."""
    ex = load_golden("cg_exacthist_128x128x128_ones.json")
    ref = np.array(load_golden("cg_hist_128x128x128_ones.json")["hist"])
    h = np.array(ex["hist"])
    assert h[0] == ref[0] == math.sqrt(11916248.0)
    live = ref / ref[0] > 1e-10
    dev = np.abs(h[live] - ref[live]) / ref[live]
    n = 128 ** 3
    assert 0.1 * n * 2.0 ** -53 <= dev[2:].max() <= n * 2.0 ** -53, dev.max()
    g32 = load_golden("cg_hist_32x32x32_ones.json")
    A = oracle.Problem(32, 32, 32)
    _, _, _, h32, _ = oracle.cg(A, max_iter=150, exact_dots=True)
    r32 = np.array(g32["hist"])
    live = r32 / r32[0] > 1e-10
    assert np.max(np.abs(h32[live] - r32[live]) / r32[live]) <= 5e-13


# ----------------------------------------------------------------------------- seeded inputs
def test_seeded_generators():
    a = oracle.xstar_sines(12, 10, 8, 7)
    assert np.array_equal(a, oracle.xstar_sines(12, 10, 8, 7))
    assert not np.array_equal(a, oracle.xstar_sines(12, 10, 8, 8))
    assert np.all(np.abs(a) <= 8.05)
    # the 8-sine part is smooth: second differences along x are far below the field's scale
    u = oracle.xstar_uniform(100000, 3)
    assert u.min() >= -1.0 and u.max() < 1.0 and abs(u.mean()) < 0.02
    assert abs(np.var(u) - 1.0 / 3.0) < 0.01
    d = oracle.batch_dims(64, 16, 96, 2024)
    assert d.shape == (64, 3) and d.min() >= 16 and d.max() <= 96
    assert len(np.unique(d)) > 30
