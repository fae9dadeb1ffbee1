"""GPU tests of the matrix-free form of generate_matrix's operator (SparseMatrix.F90:138-175):
bitwise equal to the CSR SpMV on the generated matrix (both sum each row from 0.0 in the
matrix's column order, SparseMatrix.F90:225-229) and to the oracle's row loop, on cubes and
ragged grids down to a single row (nx a multiple of 4 takes the kernel that owns four rows per
thread, other grids the one-row kernel); and usable without any matrix resident."""
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


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (5, 1, 7), (4, 1, 1), (4, 3, 2), (8, 5, 3), (16, 8, 12),
                                  (37, 21, 19), (64, 64, 64)])
def test_stencil_spmv_bitwise_equals_csr_spmv(ctx, dims):
    n = int(np.prod(dims))
    ctx.generate(*dims)
    rng = np.random.default_rng(sum(dims))
    xh = rng.standard_normal(n)
    xh[::7] *= 1e-300  # tiny and subnormal-range products round the same way in both
    x = hpccg_amd.DeviceArray.from_numpy(ctx, xh)
    y1 = hpccg_amd.DeviceArray(ctx, n)
    y2 = hpccg_amd.DeviceArray(ctx, n)
    ctx.spmv_dev(x, y1)
    ctx.spmv_stencil_dev(dims, x, y2)
    a, b = y1.numpy(), y2.numpy()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    if n <= 20000:
        A = oracle.Problem(*dims)
        assert np.array_equal(b, oracle.spmv(A, xh))


def test_stencil_spmv_needs_no_matrix():
    c = hpccg_amd.Context(0)
    try:
        dims = (9, 7, 5)
        n = int(np.prod(dims))
        x = hpccg_amd.DeviceArray.from_numpy(c, np.ones(n))
        y = hpccg_amd.DeviceArray(c, n)
        c.spmv_stencil_dev(dims, x, y)  # A * 1 = 28 - (nonzeros of the row)
        A = oracle.Problem(*dims)
        assert np.array_equal(y.numpy(), 28.0 - np.diff(A.rowindex))
        with pytest.raises(hpccg_amd.HpccgError, match="bad argument"):
            c.spmv_stencil_dev((0, 7, 5), x, y)
    finally:
        c.close()
