"""CPU-only checks of the C ABI: the library loads and exports every declared entry point, and
host-side argument validation fails cleanly.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

import pytest

import hpccg_amd
from conftest import ROOT


def test_library_built_and_loads():
    assert os.path.exists(hpccg_amd.lib_path), "run __graft_entry__.build()"
    hpccg_amd.lib()


def test_every_declared_symbol_is_exported():
    names = hpccg_amd.exported_symbols()
    assert len(names) >= 30
    L = hpccg_amd.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_headers_cite_reference_or_say_none():
    """Each reference drop-in in the header cites the reference file:line it replaces."""
    txt = open(os.path.join(ROOT, "include", "hpccg_hip.h")).read()
    for proc, cite in [("hpccg_spmv", "SparseMatrix.F90:213"), ("hpccg_ddot", "vector_operations.F90:40"),
                       ("hpccg_ddot1", "vector_operations.F90:73"), ("hpccg_waxpby", "vector_operations.F90:105"),
                       ("hpccg_wwpby", "vector_operations.F90:123"), ("hpccg_wxbw", "vector_operations.F90:139"),
                       ("hpccg_vcopy", "vector_operations.F90:155"), ("hpccg_cg", "numerics.F90:41"),
                       ("hpccg_matrix_upload", "SparseMatrix.F90:68"), ("hpccg_generate_matrix", "SparseMatrix.F90:138")]:
        i = txt.index(f"int {proc}(")
        assert cite in txt[max(0, i - 700):i], proc


def test_null_handle_is_an_error_not_a_crash():
    L = hpccg_amd.lib()
    assert L.hpccg_create(None, 0) == -1
    assert b"null" in L.hpccg_last_error()
    assert L.hpccg_set_dot_mode(None, 0) == -1
    assert L.hpccg_destroy(None) == 0


def test_no_gpu_create_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(hpccg_amd.HpccgError):
        hpccg_amd.Context(0)


def test_product_never_imports_oracle():
    """The shipped package and library sources do not reference the oracle."""
    pkg = os.path.join(ROOT, "hpccg-f90_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".hpp", ".h", ".F90", ".f90", "Makefile")):
                txt = open(os.path.join(dp, f), errors="replace").read()
                assert not re.search(r"\boracle\b", txt), os.path.join(dp, f)


def test_product_input_generators_match_oracle():
    """The library's seeded-input generators (host code) equal the oracle's, bit for bit."""
    import numpy as np
    import oracle
    for dims, seed in [((12, 10, 8), 7), ((33, 5, 17), 14), ((1, 1, 1), 0)]:
        assert np.array_equal(hpccg_amd.xstar_sines(*dims, seed), oracle.xstar_sines(*dims, seed))
    assert np.array_equal(hpccg_amd.xstar_uniform(12345, 9), oracle.xstar_uniform(12345, 9))
    assert np.array_equal(hpccg_amd.batch_dims(64, 16, 96, 2024), oracle.batch_dims(64, 16, 96, 2024))
