"""ctypes declarations for include/hpccg_hip.h and a thin object wrapper."""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
lib_path = os.path.join(PKG_ROOT, "lib", "libhpccg_hip.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "hpccg_hip.h")

DOT_SEQUENTIAL = 0
DOT_TREE = 1
WARN_BREAKDOWN = 1
NOALLOC_MAX_ROWS = 2097152  # include/hpccg_hip.h: static workspace of hpccg_cg_dev_noalloc
NOALLOC_MAX_ITER = 4094

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
D = ctypes.c_double
PI32 = ctypes.POINTER(ctypes.c_int32)
PI64 = ctypes.POINTER(ctypes.c_int64)
PD = ctypes.POINTER(ctypes.c_double)
PINT = ctypes.POINTER(ctypes.c_int)

# name: (restype, argtypes)
_SIGS = {
    "hpccg_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_int]),
    "hpccg_destroy": (ctypes.c_int, [P]),
    "hpccg_last_error": (ctypes.c_char_p, []),
    "hpccg_set_dot_mode": (ctypes.c_int, [P, ctypes.c_int]),
    "hpccg_get_dot_mode": (ctypes.c_int, [P, PINT]),
    "hpccg_device_info": (ctypes.c_int, [P, PINT, PINT, PINT]),
    "hpccg_dev_malloc": (ctypes.c_int, [P, ctypes.POINTER(P), I64]),
    "hpccg_dev_free": (ctypes.c_int, [P, P]),
    "hpccg_memcpy_h2d": (ctypes.c_int, [P, P, P, I64]),
    "hpccg_memcpy_d2h": (ctypes.c_int, [P, P, P, I64]),
    "hpccg_memcpy_d2d": (ctypes.c_int, [P, P, P, I64]),
    "hpccg_matrix_upload": (ctypes.c_int, [P, I32, I32, P, P, P]),
    "hpccg_generate_matrix": (ctypes.c_int, [P, I32, I32, I32, P, P, P, P, P, P, PI64]),
    "hpccg_matrix_info": (ctypes.c_int, [P, PI32, PI64, PI32]),
    "hpccg_matrix_format": (ctypes.c_int, [P, PI32, PI64]),
    "hpccg_spmv": (ctypes.c_int, [P, P, P]),
    "hpccg_ddot": (ctypes.c_int, [P, I32, P, P, PD]),
    "hpccg_ddot1": (ctypes.c_int, [P, I32, P, PD]),
    "hpccg_compute_residual": (ctypes.c_int, [P, I32, P, P, PD]),
    "hpccg_compute_residual_dev": (ctypes.c_int, [P, I32, P, P, PD]),
    "hpccg_waxpby": (ctypes.c_int, [P, I32, D, P, D, P, P]),
    "hpccg_wwpby": (ctypes.c_int, [P, I32, D, P, P]),
    "hpccg_wxbw": (ctypes.c_int, [P, I32, P, D, P]),
    "hpccg_vcopy": (ctypes.c_int, [P, I32, P, P]),
    "hpccg_cg": (ctypes.c_int, [P, P, P, I32, D, PI32, PD, P]),
    "hpccg_spmv_dev": (ctypes.c_int, [P, P, P]),
    "hpccg_ddot_dev": (ctypes.c_int, [P, I32, P, P, PD]),
    "hpccg_waxpby_dev": (ctypes.c_int, [P, I32, D, P, D, P, P]),
    "hpccg_wwpby_dev": (ctypes.c_int, [P, I32, D, P, P]),
    "hpccg_wxbw_dev": (ctypes.c_int, [P, I32, P, D, P]),
    "hpccg_cg_dev": (ctypes.c_int, [P, P, P, I32, D, PI32, PD, P, PD]),
    "hpccg_cg_dev_noalloc": (ctypes.c_int, [P, P, P, I32, D, PI32, PD, P, PD]),
    "hpccg_cg_timed": (ctypes.c_int, [P, P, P, I32, D, I32, PI32, PD, P, PD, PD, P, P]),
    "hpccg_cg_mixed": (ctypes.c_int, [P, P, P, I32, D, I32, PI32, PD, P, P]),
    "hpccg_cg_mixed_dev": (ctypes.c_int, [P, P, P, I32, D, I32, PI32, PD, P, P, PD]),
    "hpccg_spmv_mixed_dev": (ctypes.c_int, [P, P, P]),
    "hpccg_time_spmv_mixed": (ctypes.c_int, [P, P, P, I32, PD]),
    "hpccg_batch_setup": (ctypes.c_int, [P, I32, P]),
    "hpccg_batch_info": (ctypes.c_int, [P, PI32, PI64, PI64, P]),
    "hpccg_batch_spmv_dev": (ctypes.c_int, [P, P, P]),
    "hpccg_batch_variants": (ctypes.c_int, [PI32, P, P]),
    "hpccg_batch_autotune": (ctypes.c_int, [P, P, P, I32, PI32, P]),
    "hpccg_batch_solve_dev": (ctypes.c_int, [P, P, P, D, I32, I32, I32, P, P, P, PD]),
    "hpccg_cnn_num_weights": (ctypes.c_int, []),
    "hpccg_cnn_init": (ctypes.c_int, [P, ctypes.c_uint64]),
    "hpccg_cnn_set_weights": (ctypes.c_int, [P, P]),
    "hpccg_cnn_get_weights": (ctypes.c_int, [P, P]),
    "hpccg_cnn_forward": (ctypes.c_int, [P, I32, I32, I32, I32, P, P, P, P]),
    "hpccg_cnn_backward": (ctypes.c_int, [P, I32, I32, I32, I32, P, P, P]),
    "hpccg_cnn_make_dataset": (ctypes.c_int, [P, I32, I32, I32, ctypes.c_uint64]),
    "hpccg_cnn_dataset_pair": (ctypes.c_int, [P, I32, P, P, P]),
    "hpccg_cnn_train": (ctypes.c_int, [P, I32, I32, ctypes.c_float, ctypes.c_uint64, P, PD]),
    "hpccg_cnn_apply_dev": (ctypes.c_int, [P, I32, I32, I32, P, ctypes.c_float, P]),
    "hpccg_fcg_learned_dev": (ctypes.c_int, [P, I32, I32, I32, P, P, I32, D, I32, I32, D, PI32, PI32, PI32, PD,
                                             P, PD]),
    "hpccg_xstar_sines": (ctypes.c_int, [I32, I32, I32, ctypes.c_uint64, P]),
    "hpccg_xstar_uniform": (ctypes.c_int, [I64, ctypes.c_uint64, P]),
    "hpccg_batch_dims": (ctypes.c_int, [I32, I32, I32, ctypes.c_uint64, P]),
    "hpccg_time_spmv": (ctypes.c_int, [P, P, P, I32, PD]),
    "hpccg_time_copy": (ctypes.c_int, [P, P, P, I64, I32, PD]),
    "hpccg_time_cnn_apply": (ctypes.c_int, [P, I32, I32, I32, P, P, I32, PD]),
    "hpccg_spmv_stencil_dev": (ctypes.c_int, [P, I32, I32, I32, P, P]),
    "hpccg_time_spmv_stencil": (ctypes.c_int, [P, I32, I32, I32, P, P, I32, PD]),
}

_lib = None


class HpccgError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load libhpccg_hip.so (raises if it was not built -- there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            raise HpccgError(f"{lib_path} is missing: run __graft_entry__.build() (make -C hpccg-f90_amd)")
        L = ctypes.CDLL(lib_path)
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Function names declared in include/hpccg_hip.h."""
    import re
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hpccg_[a-z0-9_]+)\s*\(", txt)))


def _check(status, what=""):
    if status < 0:
        msg = lib().hpccg_last_error().decode(errors="replace")
        raise HpccgError(f"{what} failed ({status}): {msg}")
    return status


def _ptr(a: np.ndarray):
    if not a.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def xstar_sines(nx, ny, nz, seed):
    out = np.zeros(nx * ny * nz)
    _check(lib().hpccg_xstar_sines(nx, ny, nz, seed, _ptr(out)), "xstar_sines")
    return out


def xstar_uniform(n, seed):
    out = np.zeros(n)
    _check(lib().hpccg_xstar_uniform(n, seed, _ptr(out)), "xstar_uniform")
    return out


def batch_dims(count, lo, hi, seed):
    out = np.zeros(3 * count, dtype=np.int32)
    _check(lib().hpccg_batch_dims(count, lo, hi, seed, _ptr(out)), "batch_dims")
    return out.reshape(count, 3)


class DeviceArray:
    """A device buffer owned by a Context (float64 unless dtype says otherwise)."""

    def __init__(self, ctx: "Context", n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = P()
        _check(lib().hpccg_dev_malloc(ctx.h, ctypes.byref(p), self.nbytes), "dev_malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(lib().hpccg_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), self.nbytes), "h2d")

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        _check(lib().hpccg_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            lib().hpccg_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class Context:
    """One device + one HIP stream + the resident matrix (hpccg_ctx)."""

    def __init__(self, device: int = 0, dot_mode: int = DOT_TREE):
        h = P()
        _check(lib().hpccg_create(ctypes.byref(h), device), "hpccg_create")
        self.h = h.value
        self.n = 0
        self.set_dot_mode(dot_mode)

    def close(self):
        if self.h:
            lib().hpccg_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup
    def set_dot_mode(self, mode: int):
        _check(lib().hpccg_set_dot_mode(self.h, mode), "set_dot_mode")

    def device_info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().hpccg_device_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"num_cu": a.value, "num_xcd": b.value, "gfx": c.value}

    def upload(self, rowindex, colindex, values):
        rowindex = np.ascontiguousarray(rowindex, dtype=np.int32)
        colindex = np.ascontiguousarray(colindex, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        n = len(rowindex) - 1
        _check(lib().hpccg_matrix_upload(self.h, n, n, _ptr(rowindex), _ptr(colindex), _ptr(values)),
               "matrix_upload")
        self.n = n

    def generate(self, nx, ny, nz, want_host=False):
        """Device-side generate_matrix; optionally returns the host copies."""
        n = nx * ny * nz
        nnz = ctypes.c_int64()
        if want_host:
            nnz_c = (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
            out = {"rowindex": np.zeros(n + 1, np.int32), "colindex": np.zeros(nnz_c, np.int32),
                   "values": np.zeros(nnz_c), "x": np.zeros(n), "b": np.zeros(n), "xexact": np.zeros(n)}
            args = [_ptr(out[k]) for k in ("rowindex", "colindex", "values", "x", "b", "xexact")]
        else:
            out, args = None, [None] * 6
        _check(lib().hpccg_generate_matrix(self.h, nx, ny, nz, *args, ctypes.byref(nnz)), "generate_matrix")
        self.n = n
        return out

    def matrix_info(self):
        n, nnz, mr = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        _check(lib().hpccg_matrix_info(self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(mr)))
        return {"nrow": n.value, "nnz": nnz.value, "max_row_nnz": mr.value}

    def matrix_format(self):
        """Device format of the uploaded matrix: `ncodes` offsets in the column-code table (0:
        int32 columns) and `copy_bytes` of the FP64 sliced copy."""
        nc, cb = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().hpccg_matrix_format(self.h, ctypes.byref(nc), ctypes.byref(cb)))
        return {"ncodes": nc.value, "copy_bytes": cb.value}

    # ---- host-array drop-ins
    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        _check(lib().hpccg_spmv(self.h, _ptr(x), _ptr(y)), "spmv")
        return y

    def ddot(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        r = ctypes.c_double()
        _check(lib().hpccg_ddot(self.h, len(x), _ptr(x), _ptr(y), ctypes.byref(r)), "ddot")
        return r.value

    def ddot1(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = ctypes.c_double()
        _check(lib().hpccg_ddot1(self.h, len(x), _ptr(x), ctypes.byref(r)), "ddot1")
        return r.value

    def compute_residual(self, v1, v2):
        v1 = np.ascontiguousarray(v1, dtype=np.float64)
        v2 = np.ascontiguousarray(v2, dtype=np.float64)
        r = ctypes.c_double()
        _check(lib().hpccg_compute_residual(self.h, len(v1), _ptr(v1), _ptr(v2), ctypes.byref(r)), "compute_residual")
        return r.value

    def compute_residual_dev(self, v1: "DeviceArray", v2: "DeviceArray"):
        r = ctypes.c_double()
        _check(lib().hpccg_compute_residual_dev(self.h, v1.n, v1.ptr, v2.ptr, ctypes.byref(r)), "compute_residual_dev")
        return r.value

    def waxpby(self, alpha, x, beta, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        w = np.zeros(len(x))
        _check(lib().hpccg_waxpby(self.h, len(x), alpha, _ptr(x), beta, _ptr(y), _ptr(w)), "waxpby")
        return w

    def wwpby(self, beta, y, w):
        y = np.ascontiguousarray(y, dtype=np.float64)
        w = np.array(w, dtype=np.float64)
        _check(lib().hpccg_wwpby(self.h, len(w), beta, _ptr(y), _ptr(w)), "wwpby")
        return w

    def wxbw(self, x, beta, w):
        x = np.ascontiguousarray(x, dtype=np.float64)
        w = np.array(w, dtype=np.float64)
        _check(lib().hpccg_wxbw(self.h, len(w), _ptr(x), beta, _ptr(w)), "wxbw")
        return w

    def vcopy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(len(x))
        _check(lib().hpccg_vcopy(self.h, len(x), _ptr(x), _ptr(y)), "vcopy")
        return y

    def cg(self, b, x0=None, max_iter=150, tolerance=0.0):
        """cg(A,b,x,max_iter,tolerance,niters,normr) -> (x, niters, normr, hist, status)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(max_iter + 1)
        nit, nr = ctypes.c_int32(), ctypes.c_double()
        st = _check(lib().hpccg_cg(self.h, _ptr(b), _ptr(x), max_iter, tolerance, ctypes.byref(nit),
                                   ctypes.byref(nr), _ptr(hist)), "cg")
        return x, nit.value, nr.value, hist, st

    def cg_timed(self, b, x0=None, max_iter=150, tolerance=0.0, level=2):
        """cg with the reference's timer regions -> dict(x, niters, normr, hist, status, t_total,
        t_prepare, t_iter[niters], t_spmv[niters]) (seconds; per-iteration arrays at level >= 2)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(max_iter + 1)
        ti, ts = np.zeros(max(max_iter, 1)), np.zeros(max(max_iter, 1))
        nit, nr, tt, tp = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        st = _check(lib().hpccg_cg_timed(self.h, _ptr(b), _ptr(x), max_iter, tolerance, level, ctypes.byref(nit),
                                         ctypes.byref(nr), _ptr(hist), ctypes.byref(tt), ctypes.byref(tp),
                                         _ptr(ti), _ptr(ts)), "cg_timed")
        k = nit.value
        return {"x": x, "niters": k, "normr": nr.value, "hist": hist, "status": st, "t_total": tt.value,
                "t_prepare": tp.value, "t_iter": ti[:k] if level >= 2 else None,
                "t_spmv": ts[:k] if level >= 2 else None}

    # ---- device-array entry points
    def cg_dev(self, b: DeviceArray, x: DeviceArray, max_iter=150, tolerance=0.0, want_hist=True):
        hist = np.zeros(max_iter + 1) if want_hist else None
        nit, nr, ms = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        st = _check(lib().hpccg_cg_dev(self.h, b.ptr, x.ptr, max_iter, tolerance, ctypes.byref(nit),
                                       ctypes.byref(nr), _ptr(hist) if want_hist else None,
                                       ctypes.byref(ms)), "cg_dev")
        return {"niters": nit.value, "normr": nr.value, "hist": hist, "status": st, "ms": ms.value}

    def cg_dev_noalloc(self, b: DeviceArray, x: DeviceArray, max_iter=150, tolerance=0.0, want_hist=True):
        """hpccg_cg_dev on the device's static workspace: allocates nothing (limits in the header)."""
        hist = np.zeros(max_iter + 1) if want_hist else None
        nit, nr, ms = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        st = _check(lib().hpccg_cg_dev_noalloc(self.h, b.ptr, x.ptr, max_iter, tolerance, ctypes.byref(nit),
                                               ctypes.byref(nr), _ptr(hist) if want_hist else None,
                                               ctypes.byref(ms)), "cg_dev_noalloc")
        return {"niters": nit.value, "normr": nr.value, "hist": hist, "status": st, "ms": ms.value}

    def cg_mixed(self, b, x0=None, max_iter=1000, tolerance=0.0, replace_every=25):
        """Mixed-precision CG on host arrays -> (x, niters, true_normr, hist_rec, hist_true, status)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        hr, ht = np.zeros(max_iter + 1), np.zeros(max_iter + 1)
        nit, tn = ctypes.c_int32(), ctypes.c_double()
        st = _check(lib().hpccg_cg_mixed(self.h, _ptr(b), _ptr(x), max_iter, tolerance, replace_every,
                                         ctypes.byref(nit), ctypes.byref(tn), _ptr(hr), _ptr(ht)), "cg_mixed")
        return x, nit.value, tn.value, hr, ht, st

    def cg_mixed_dev(self, b: DeviceArray, x: DeviceArray, max_iter=1000, tolerance=0.0, replace_every=25,
                     want_hist=True):
        hr = np.zeros(max_iter + 1) if want_hist else None
        ht = np.zeros(max_iter + 1) if want_hist else None
        nit, tn, ms = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        st = _check(lib().hpccg_cg_mixed_dev(self.h, b.ptr, x.ptr, max_iter, tolerance, replace_every,
                                             ctypes.byref(nit), ctypes.byref(tn),
                                             _ptr(hr) if want_hist else None, _ptr(ht) if want_hist else None,
                                             ctypes.byref(ms)), "cg_mixed_dev")
        return {"niters": nit.value, "true_normr": tn.value, "hist_rec": hr, "hist_true": ht, "status": st,
                "ms": ms.value}

    def spmv_mixed_dev(self, x: DeviceArray, y: DeviceArray):
        _check(lib().hpccg_spmv_mixed_dev(self.h, x.ptr, y.ptr), "spmv_mixed_dev")

    def time_spmv_mixed(self, x: DeviceArray, y: DeviceArray, reps=20):
        ms = ctypes.c_double()
        _check(lib().hpccg_time_spmv_mixed(self.h, x.ptr, y.ptr, reps, ctypes.byref(ms)), "time_spmv_mixed")
        return ms.value

    # ---- batched solves
    def batch_setup(self, dims):
        dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1))
        _check(lib().hpccg_batch_setup(self.h, len(dims) // 3, _ptr(dims)), "batch_setup")
        return self.batch_info()

    def batch_info(self):
        cnt, E, nnz = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().hpccg_batch_info(self.h, ctypes.byref(cnt), ctypes.byref(E), ctypes.byref(nnz), None))
        off = np.zeros(cnt.value + 1, dtype=np.int64)
        _check(lib().hpccg_batch_info(self.h, None, None, None, _ptr(off)))
        return {"count": cnt.value, "rows": E.value, "nnz": nnz.value, "row_off": off}

    def batch_spmv_dev(self, x: DeviceArray, y: DeviceArray):
        _check(lib().hpccg_batch_spmv_dev(self.h, x.ptr, y.ptr), "batch_spmv_dev")

    @staticmethod
    def batch_variants():
        n = ctypes.c_int32()
        lib().hpccg_batch_variants(ctypes.byref(n), None, None)
        rows, waves = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        lib().hpccg_batch_variants(ctypes.byref(n), _ptr(rows), _ptr(waves))
        return list(zip(rows.tolist(), waves.tolist()))  # (rows per wave, waves per workgroup)

    def batch_autotune(self, b: DeviceArray, x: DeviceArray, probe_iters=8):
        nv = len(self.batch_variants())
        ms = np.zeros(nv)
        best = ctypes.c_int32()
        _check(lib().hpccg_batch_autotune(self.h, b.ptr, x.ptr, probe_iters, ctypes.byref(best), _ptr(ms)),
               "batch_autotune")
        return best.value, ms

    def batch_solve_dev(self, b: DeviceArray, x: DeviceArray, rel_tol=1e-8, max_iter=200, replace_every=25,
                        variant=-1):
        S = self.batch_info()["count"]
        nit, rel, conv = np.zeros(S, np.int32), np.zeros(S), np.zeros(S, np.int32)
        ms = ctypes.c_double()
        st = _check(lib().hpccg_batch_solve_dev(self.h, b.ptr, x.ptr, rel_tol, max_iter, replace_every, variant,
                                                _ptr(nit), _ptr(rel), _ptr(conv), ctypes.byref(ms)), "batch_solve")
        return {"niters": nit, "relres": rel, "converged": conv, "ms": ms.value, "status": st}

    # ---- learned preconditioner
    @staticmethod
    def cnn_num_weights():
        return lib().hpccg_cnn_num_weights()

    def cnn_init(self, seed=1):
        _check(lib().hpccg_cnn_init(self.h, seed), "cnn_init")

    def cnn_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.cnn_num_weights()
        _check(lib().hpccg_cnn_set_weights(self.h, _ptr(w)), "cnn_set_weights")

    def cnn_get_weights(self):
        w = np.zeros(self.cnn_num_weights(), np.float32)
        _check(lib().hpccg_cnn_get_weights(self.h, _ptr(w)), "cnn_get_weights")
        return w

    def cnn_forward(self, u, dims, batch=1, want_act=False):
        nx, ny, nz = dims
        u = np.ascontiguousarray(u, dtype=np.float32)
        V = nx * ny * nz * batch
        z = np.zeros(V, np.float32)
        h1 = np.zeros(V * 16, np.float32) if want_act else None
        h2 = np.zeros(V * 16, np.float32) if want_act else None
        _check(lib().hpccg_cnn_forward(self.h, nx, ny, nz, batch, _ptr(u), _ptr(h1) if want_act else None,
                                       _ptr(h2) if want_act else None, _ptr(z)), "cnn_forward")
        return (z, h1, h2) if want_act else z

    def cnn_backward(self, u, dz, dims, batch=1):
        nx, ny, nz = dims
        u = np.ascontiguousarray(u, dtype=np.float32)
        dz = np.ascontiguousarray(dz, dtype=np.float32)
        dW = np.zeros(self.cnn_num_weights(), np.float32)
        _check(lib().hpccg_cnn_backward(self.h, nx, ny, nz, batch, _ptr(u), _ptr(dz), _ptr(dW)), "cnn_backward")
        return dW

    def cnn_make_dataset(self, n=64, npairs=20000, crop=24, seed=7):
        _check(lib().hpccg_cnn_make_dataset(self.h, n, npairs, crop, seed), "cnn_make_dataset")

    def cnn_dataset_pair(self, i, crop):
        r = np.zeros(crop ** 3, np.float32)
        e = np.zeros(crop ** 3, np.float32)
        m = np.zeros(6, np.int32)
        _check(lib().hpccg_cnn_dataset_pair(self.h, i, _ptr(r), _ptr(e), _ptr(m)), "cnn_dataset_pair")
        return r, e, m

    def cnn_train(self, steps, batch=32, lr=2e-3, seed=11):
        loss = np.zeros(steps, np.float32)
        ms = ctypes.c_double()
        _check(lib().hpccg_cnn_train(self.h, steps, batch, lr, seed, _ptr(loss), ctypes.byref(ms)), "cnn_train")
        return loss, ms.value

    def cnn_apply_dev(self, dims, r: DeviceArray, inv_scale, z: DeviceArray):
        _check(lib().hpccg_cnn_apply_dev(self.h, *dims, r.ptr, inv_scale, z.ptr), "cnn_apply_dev")

    def fcg_learned_dev(self, dims, b: DeviceArray, x: DeviceArray, max_iter=1000, tolerance=0.0, replace_every=25,
                        window=10, min_factor=2.0, want_hist=False):
        fi, ci, fb = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        tn, ms = ctypes.c_double(), ctypes.c_double()
        hist = np.zeros(max_iter + 1) if want_hist else None
        st = _check(lib().hpccg_fcg_learned_dev(self.h, *dims, b.ptr, x.ptr, max_iter, tolerance, replace_every,
                                                window, min_factor, ctypes.byref(fi), ctypes.byref(ci),
                                                ctypes.byref(fb), ctypes.byref(tn),
                                                _ptr(hist) if want_hist else None, ctypes.byref(ms)),
                    "fcg_learned_dev")
        return {"fcg_iters": fi.value, "cg_iters": ci.value, "fell_back": bool(fb.value), "true_normr": tn.value,
                "ms": ms.value, "status": st, "hist": hist}

    def spmv_dev(self, x: DeviceArray, y: DeviceArray):
        _check(lib().hpccg_spmv_dev(self.h, x.ptr, y.ptr), "spmv_dev")

    def ddot_dev(self, x: DeviceArray, y: DeviceArray):
        r = ctypes.c_double()
        _check(lib().hpccg_ddot_dev(self.h, x.n, x.ptr, y.ptr, ctypes.byref(r)), "ddot_dev")
        return r.value

    def time_spmv(self, x: DeviceArray, y: DeviceArray, reps=20):
        ms = ctypes.c_double()
        _check(lib().hpccg_time_spmv(self.h, x.ptr, y.ptr, reps, ctypes.byref(ms)), "time_spmv")
        return ms.value

    def spmv_stencil_dev(self, dims, x: DeviceArray, y: DeviceArray):
        """Matrix-free y = A x of generate_matrix's 27-point operator on the grid `dims`."""
        _check(lib().hpccg_spmv_stencil_dev(self.h, *dims, x.ptr, y.ptr), "spmv_stencil_dev")

    def time_spmv_stencil(self, dims, x: DeviceArray, y: DeviceArray, reps=20):
        ms = ctypes.c_double()
        _check(lib().hpccg_time_spmv_stencil(self.h, *dims, x.ptr, y.ptr, reps, ctypes.byref(ms)),
               "time_spmv_stencil")
        return ms.value

    def time_cnn_apply(self, dims, r: DeviceArray, z: DeviceArray, reps=20):
        """Mean time in ms of the fused network inference on one field (HIP events)."""
        ms = ctypes.c_double()
        _check(lib().hpccg_time_cnn_apply(self.h, *dims, r.ptr, z.ptr, reps, ctypes.byref(ms)), "time_cnn_apply")
        return ms.value

    def time_copy(self, dst: DeviceArray, src: DeviceArray, reps=20):
        ms = ctypes.c_double()
        _check(lib().hpccg_time_copy(self.h, dst.ptr, src.ptr, min(dst.nbytes, src.nbytes), reps,
                                     ctypes.byref(ms)), "time_copy")
        return ms.value
