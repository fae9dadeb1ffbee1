"""CPU checks of the benchmark's measurement arithmetic: the SpMV byte counts its roofline uses
(closed forms of generate_matrix's matrix, SparseMatrix.F90:106-185) and the reading of
rocprofv3 counter output into per-dispatch traffic.  No GPU is used."""
import csv
import importlib.util
import os

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("n1,csr,fmt", [(128, 710_858_660, 543_629_756), (256, 5_729_005_476, 4_380_640_188)])
def test_spmv_bytes_closed_forms(bench, n1, csr, fmt):
    """nnz = (3n-2)^3; CSR bytes = 12 per entry + (N+1) row pointers + x once + y; the coded copy
    stores a one-byte column code instead of the int32 column (9 bytes per entry)."""
    n, nnz = n1 ** 3, (3 * n1 - 2) ** 3
    assert bench.spmv_bytes(n, nnz) == csr
    assert bench.spmv_bytes(n, nnz, s_idx=1) == fmt
    assert csr - fmt == 3 * nnz


@pytest.mark.parametrize("n1,fmt16", [(128, 192_394_732), (256, 1_549_691_884)])
def test_fp16_spmv_format_bytes(bench, n1, fmt16):
    """The mixed SpMV's format: 2-byte values and 1-byte codes (3 per entry), the row pointers,
    FP32 x once and FP32 y."""
    n, nnz = n1 ** 3, (3 * n1 - 2) ** 3
    assert bench.spmv_bytes(n, nnz, s_val=2, s_idx=1, s_x=4, s_y=4) == fmt16
    assert fmt16 == 3 * nnz + 4 * (n + 1) + 8 * n


def _write_counter_csv(path, rows):
    cols = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(cols, r)))


def test_counter_kib_means_per_kernel(bench, tmp_path):
    """Per-dispatch means for the SpMV and the copy kernel only, of the requested counter only."""
    sub = tmp_path / "host"
    sub.mkdir()
    _write_counter_csv(sub / "run_counter_collection.csv", [
        (1, "void hpccg::k_cg_spmv<true>(hpccg::SellRef<double>, double const*)", "FETCH_SIZE", "100.0"),
        (2, "void hpccg::k_cg_spmv<true>(hpccg::SellRef<double>, double const*)", "FETCH_SIZE", "300.0"),
        (3, "hpccg::k_copy16(unsigned int __vector(4) const*)", "FETCH_SIZE", "50.0"),
        (4, "void hpccg::k_cg_spmv<false>(hpccg::SellRef<double>)", "FETCH_SIZE", "9999.0"),
        (5, "hpccg::k_generate(int, int, int)", "FETCH_SIZE", "7.0"),
        (6, "hpccg::k_copy16(unsigned int __vector(4) const*)", "WRITE_SIZE", "80.0"),
        (7, "hpccg::km_spmv(hpccg::SellRef<__half>, float const*, float*, hpccg::MixState const*, double*)",
         "FETCH_SIZE", "40.0"),
        (8, "hpccg::km_spmv_finish(double const*, int, hpccg::MixState*)", "FETCH_SIZE", "1.0"),
    ])
    got = bench.counter_kib(str(tmp_path), "FETCH_SIZE")
    assert got == {"spmv": (200.0, 2), "copy": (50.0, 1), "spmv16": (40.0, 1)}
    assert bench.counter_kib(str(tmp_path), "WRITE_SIZE") == {"copy": (80.0, 1)}


def test_traffic_measurement_can_be_switched_off(bench, monkeypatch):
    """Without the profiler (or when switched off) the benchmark reports the stored measurement."""
    monkeypatch.setenv("HPCCG_BENCH_NO_COUNTERS", "1")
    assert bench.measure_traffic(256) is None
