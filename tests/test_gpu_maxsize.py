"""The largest grid the reference's CSR can index: its SparseMatrixT keeps rowindex, colindex
and total_nnz as default (32-bit) integers (SparseMatrix.F90:44-53), and rowindex(N+1) =
nnz + 1 must stay below 2^31.  For a cube that is 430^3: 79,507,000 rows and
(3*430 - 2)^3 = 2,136,719,872 nonzeros (2^31 - 1 = 2,147,483,647), 49 GB of device arrays.
No oracle runs at this size; the checks are closed forms: the device-built matrix's counts,
b = A*1 summing to 28 N - nnz (generate_matrix's b, SparseMatrix.F90:170-176), the
reference benchmark's 150 FP64 CG iterations from x0 = 0 reaching x* = 1, and mixed CG
meeting 1e-10 in the FP64 true residual.  The next grid up is refused before anything is
allocated."""
import math

import numpy as np
import pytest

import hpccg_amd

pytestmark = pytest.mark.gpu

N1 = 430


@pytest.fixture(scope="module")
def big():
    c = hpccg_amd.Context(0)
    c.generate(N1, N1, N1)
    yield c
    c.close()


def test_max_grid_matrix_counts(big):
    info = big.matrix_info()
    n, nnz = N1 ** 3, (3 * N1 - 2) ** 3
    assert nnz + 1 < 2 ** 31 - 1 < (3 * (N1 + 1) - 2) ** 3 + 1
    assert (info["nrow"], info["nnz"], info["max_row_nnz"]) == (n, nnz, 27)
    assert big.matrix_format()["ncodes"] == 27


def test_max_grid_fp64_cg_reaches_the_exact_solution(big):
    n = N1 ** 3
    ones = hpccg_amd.DeviceArray.from_numpy(big, np.ones(n))
    b = hpccg_amd.DeviceArray(big, n)
    big.spmv_dev(ones, b)
    del ones
    bh = b.numpy()
    assert float(bh.sum()) == 28.0 * n - (3 * N1 - 2) ** 3  # row i of A*1 is 28 - (its nonzeros)
    assert bh.min() == 1.0 and bh.max() == 20.0            # interior rows / the 8 corner rows
    x = hpccg_amd.DeviceArray.from_numpy(big, np.zeros(n))
    res = big.cg_dev(b, x, max_iter=150, tolerance=0.0)
    assert res["niters"] == 150 and res["status"] == 0
    hist = res["hist"]
    assert hist[0] == math.sqrt(float(bh @ bh))
    assert hist[150] < 1e-12 * hist[0]
    assert float(np.max(np.abs(x.numpy() - 1.0))) < 1e-12


def test_max_grid_mixed_cg_meets_the_tolerance(big):
    n = N1 ** 3
    rng = np.random.default_rng(430)
    xs = hpccg_amd.DeviceArray.from_numpy(big, rng.uniform(-1.0, 1.0, n))
    b = hpccg_amd.DeviceArray(big, n)
    big.spmv_dev(xs, b)
    normb = math.sqrt(big.ddot_dev(b, b))
    x = hpccg_amd.DeviceArray.from_numpy(big, np.zeros(n))
    res = big.cg_mixed_dev(b, x, max_iter=300, tolerance=1e-10 * normb, want_hist=False)
    assert res["status"] == 0 and res["niters"] < 300
    r = hpccg_amd.DeviceArray(big, n)
    big.spmv_dev(x, r)  # FP64 true residual b - A x, formed on the host from the FP64 SpMV
    rel = float(np.linalg.norm(b.numpy() - r.numpy())) / normb
    assert rel <= 1e-10, rel


def test_one_grid_larger_is_refused(big):
    with pytest.raises(hpccg_amd.HpccgError, match="too large"):
        big.generate(N1 + 1, N1 + 1, N1 + 1)
    assert big.matrix_info()["nrow"] == N1 ** 3  # the resident matrix is untouched
