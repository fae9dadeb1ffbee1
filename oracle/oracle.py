"""ctypes view of liboracle.so, the CPU restatement of the reference CG path.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are
the only callers, always as the checker.  Every function mirrors one reference procedure of
RRZE-HPC/HPCCG-F90 (MPI/ variant), cited on the C side in hpccg_oracle.c.

Pinning (what checks each function, beyond retyping it):
  generate_matrix  bitwise vs the reference's own generate_matrix (oracle/_ref harness "gen"),
                   closed forms for nnz / row classes / ||b||^2, A = 28 I - T (x) T (x) T, A*1 = b
  spmv             bitwise vs the reference spmv on random x; Kronecker eigenpairs
  ddot, ddot1      bitwise vs the reference; exact integer sums
  waxpby family    bitwise vs the reference
  cg               per-iteration ||r_k|| bitwise vs the reference cg (fixtures in tests/golden)
  ddot_exact       vs math.fsum on random data
  xstar_*/dims     no reference counterpart (the reference hard-codes x* = 1); pinned by
                   closed-form properties (range, seed determinism) in tests/test_oracle.py
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "lib"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        L.oracle_generate_matrix.restype = ctypes.c_int64
        L.oracle_generate_matrix.argtypes = [I, I, I, P, P, P, P, P, P]
        L.oracle_spmv.restype = None
        L.oracle_spmv.argtypes = [I, P, P, P, P, P]
        L.oracle_ddot.restype = D
        L.oracle_ddot.argtypes = [I, P, P]
        L.oracle_ddot1.restype = D
        L.oracle_ddot1.argtypes = [I, P]
        L.oracle_ddot_exact.restype = D
        L.oracle_ddot_exact.argtypes = [I, P, P]
        L.oracle_waxpby.restype = None
        L.oracle_waxpby.argtypes = [I, D, P, D, P, P]
        L.oracle_wwpby.restype = None
        L.oracle_wwpby.argtypes = [I, D, P, P]
        L.oracle_wxbw.restype = None
        L.oracle_wxbw.argtypes = [I, P, D, P]
        L.oracle_vcopy.restype = None
        L.oracle_vcopy.argtypes = [I, P, P]
        L.oracle_cg.restype = I
        L.oracle_cg.argtypes = [I, P, P, P, P, P, I, D, ctypes.POINTER(I), ctypes.POINTER(D), P, I]
        L.oracle_compute_residual.restype = D
        L.oracle_compute_residual.argtypes = [I, P, P]
        L.oracle_xstar_sines.restype = None
        L.oracle_xstar_sines.argtypes = [I, I, I, ctypes.c_uint64, P]
        L.oracle_xstar_uniform.restype = None
        L.oracle_xstar_uniform.argtypes = [ctypes.c_int64, ctypes.c_uint64, P]
        L.oracle_batch_dims.restype = None
        L.oracle_batch_dims.argtypes = [I, I, I, ctypes.c_uint64, P]
        F = ctypes.c_float
        L.oracle_bf16.restype = F
        L.oracle_bf16.argtypes = [F]
        L.oracle_cnn_forward.restype = None
        L.oracle_cnn_forward.argtypes = [I, I, I, I, P, P, P, P, P, P, P, I]
        L.oracle_cnn_backward.restype = None
        L.oracle_cnn_backward.argtypes = [I, I, I, I, P, P, P, P, P, P, P, P, P, P, I]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data


class Problem:
    """generate_matrix output: 1-based CSR exactly as the reference stores it."""

    def __init__(self, nx: int, ny: int, nz: int):
        n = nx * ny * nz
        self.dims = (nx, ny, nz)
        self.n = n
        self.rowindex = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(27 * n, dtype=np.int32)
        val = np.zeros(27 * n, dtype=np.float64)
        self.x = np.zeros(n)
        self.b = np.zeros(n)
        self.xexact = np.zeros(n)
        nnz = lib().oracle_generate_matrix(nx, ny, nz, _p(self.rowindex), _p(col), _p(val),
                                           _p(self.x), _p(self.b), _p(self.xexact))
        self.nnz = int(nnz)
        self.colindex = col[: self.nnz].copy()
        self.values = val[: self.nnz].copy()


def spmv(A: Problem, x: np.ndarray) -> np.ndarray:
    y = np.zeros(A.n)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().oracle_spmv(A.n, _p(A.rowindex), _p(A.colindex), _p(A.values), _p(x), _p(y))
    return y


def ddot(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().oracle_ddot(len(x), _p(x), _p(y))


def ddot1(x) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_ddot1(len(x), _p(x))


def ddot_exact(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().oracle_ddot_exact(len(x), _p(x), _p(y))


def waxpby(alpha, x, beta, y) -> np.ndarray:
    w = np.zeros(len(x))
    lib().oracle_waxpby(len(x), alpha, _p(np.ascontiguousarray(x)), beta, _p(np.ascontiguousarray(y)), _p(w))
    return w


def wwpby(beta, y, w) -> np.ndarray:
    w = np.array(w, dtype=np.float64)
    lib().oracle_wwpby(len(w), beta, _p(np.ascontiguousarray(y)), _p(w))
    return w


def wxbw(x, beta, w) -> np.ndarray:
    w = np.array(w, dtype=np.float64)
    lib().oracle_wxbw(len(w), _p(np.ascontiguousarray(x)), beta, _p(w))
    return w


def cg(A: Problem, b=None, x0=None, max_iter=150, tolerance=0.0, exact_dots=False):
    """Returns (x, niters, normr, hist[0..max_iter], status)."""
    b = A.b if b is None else np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    hist = np.zeros(max_iter + 1)
    nit = ctypes.c_int(0)
    nr = ctypes.c_double(0.0)
    st = lib().oracle_cg(A.n, _p(A.rowindex), _p(A.colindex), _p(A.values), _p(b), _p(x),
                         max_iter, tolerance, ctypes.byref(nit), ctypes.byref(nr), _p(hist),
                         1 if exact_dots else 0)
    return x, nit.value, nr.value, hist, st


def compute_residual(v1, v2) -> float:
    v1 = np.ascontiguousarray(v1, dtype=np.float64)
    v2 = np.ascontiguousarray(v2, dtype=np.float64)
    return lib().oracle_compute_residual(len(v1), _p(v1), _p(v2))


def xstar_sines(nx, ny, nz, seed) -> np.ndarray:
    out = np.zeros(nx * ny * nz)
    lib().oracle_xstar_sines(nx, ny, nz, seed, _p(out))
    return out


def xstar_uniform(n, seed) -> np.ndarray:
    out = np.zeros(n)
    lib().oracle_xstar_uniform(n, seed, _p(out))
    return out


def batch_dims(count, lo, hi, seed) -> np.ndarray:
    out = np.zeros(3 * count, dtype=np.int32)
    lib().oracle_batch_dims(count, lo, hi, seed, _p(out))
    return out.reshape(count, 3)


NW1, NW2, NW3 = 16 * 27, 16 * 16 * 27, 16 * 27


def split_weights(w):
    w = np.ascontiguousarray(w, dtype=np.float32)
    return w[:NW1].copy(), w[NW1:NW1 + NW2].copy(), w[NW1 + NW2:].copy()


def cnn_forward(w, u, dims, batch=1, quant=True):
    """Returns (z, h1, h2) of the 3-layer network (hpccg_oracle.h)."""
    W1, W2, W3 = split_weights(w)
    nx, ny, nz = dims
    V = nx * ny * nz * batch
    u = np.ascontiguousarray(u, dtype=np.float32)
    h1, h2, z = np.zeros(V * 16, np.float32), np.zeros(V * 16, np.float32), np.zeros(V, np.float32)
    lib().oracle_cnn_forward(nx, ny, nz, batch, _p(W1), _p(W2), _p(W3), _p(u), _p(h1), _p(h2), _p(z),
                             1 if quant else 0)
    return z, h1, h2


def cnn_backward(w, u, dz, dims, batch=1, quant=True):
    W1, W2, W3 = split_weights(w)
    z, h1, h2 = cnn_forward(w, u, dims, batch, quant)
    nx, ny, nz = dims
    dW1, dW2, dW3 = np.zeros(NW1, np.float32), np.zeros(NW2, np.float32), np.zeros(NW3, np.float32)
    u = np.ascontiguousarray(u, dtype=np.float32)
    dz = np.ascontiguousarray(dz, dtype=np.float32)
    lib().oracle_cnn_backward(nx, ny, nz, batch, _p(W1), _p(W2), _p(W3), _p(u), _p(h1), _p(h2), _p(dz),
                              _p(dW1), _p(dW2), _p(dW3), 1 if quant else 0)
    return np.concatenate([dW1, dW2, dW3])


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def ref_binary(variant: str, name: str) -> str | None:
    """Path of a reference binary built by oracle/Makefile into oracle/_ref, or None."""
    p = os.path.join(REF_DIR, variant, name)
    return p if os.path.exists(p) else None
