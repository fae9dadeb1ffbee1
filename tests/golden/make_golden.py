#!/usr/bin/env python3
"""Regenerate the golden fixtures in tests/golden/ from the reference's own Fortran path.

Runs oracle/_ref/serial/ref_harness (the reference's unmodified MPI/src modules, serial
amdflang -O3 build, compiled by oracle/Makefile) and oracle/_ref/serial/HPCCG (the reference's
own main).  Needs the reference sources to build oracle/_ref; the fixtures it writes are data
only (inputs described by grid and seed, outputs as exact float reprs and SHA-256 digests).

    python tests/golden/make_golden.py [--big]     # --big adds 128^3 / 256^3 cases (minutes)
    python tests/golden/make_golden.py --exact     # only the exact-dot 128^3 history (~15 s)
    python tests/golden/make_golden.py --timers    # only the TIMINGLEV=3 stdout (timer values masked)

The exact-dot history (cg_exacthist_*.json) is written by the C restatement (oracle.cg with
exact_dots=True: the reference's cg loop with both dots rounded once from a double-double
accumulation), not by the reference binary: it is the yardstick for reordered (tree) dots, whose
distance from the reference's own sequential sums grows like N*eps (1.6e-10 relative at 128^3).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

GOLD = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "serial", "ref_harness")
MAIN = os.path.join(ROOT, "oracle", "_ref", "serial", "HPCCG")
WORKERS = 8


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(args, cwd):
    out = subprocess.run(args, cwd=cwd, check=True, capture_output=True, text=True).stdout
    return [ln.split()[1:] for ln in out.splitlines() if ln.strip().startswith("H ")]


def gen_case(tmp, nx, ny, nz):
    pre = os.path.join(tmp, f"g{nx}_{ny}_{nz}")
    rows = run([HARNESS, "gen", str(nx), str(ny), str(nz), pre], tmp)
    n, nnz, total_nrow, total_nnz = (int(v) for v in rows[0][1:])
    rd = lambda ext, dt: np.fromfile(pre + ext, dtype=dt)
    rec = {
        "grid": [nx, ny, nz], "n": n, "nnz": nnz, "total_nrow": total_nrow, "total_nnz": total_nnz,
        "sha_rowindex": sha(rd(".rowindex", np.int32)), "sha_colindex": sha(rd(".colindex", np.int32)),
        "sha_values": sha(rd(".values", np.float64)), "sha_b": sha(rd(".b", np.float64)),
        "sha_x": sha(rd(".x", np.float64)), "sha_xexact": sha(rd(".xexact", np.float64)),
    }
    for ext in (".rowindex", ".colindex", ".values", ".b", ".x", ".xexact"):
        os.remove(pre + ext)
    return rec


def kernels_case(tmp, nx, ny, nz, seed_x, seed_y, alpha, beta, full):
    n = nx * ny * nz
    x = oracle.xstar_uniform(n, seed_x)
    y = oracle.xstar_uniform(n, seed_y)
    fx, fy, fo = (os.path.join(tmp, f"k{nx}_{ny}_{nz}.{s}") for s in ("x", "y", "out"))
    x.tofile(fx)
    y.tofile(fy)
    rows = run([HARNESS, "kernels", str(nx), str(ny), str(nz), fx, fy, repr(alpha), repr(beta), fo], tmp)
    out = np.fromfile(fo, dtype=np.float64).reshape(5, n)
    d1, d2 = (float(v) for v in rows[0][1:])
    rec = {"grid": [nx, ny, nz], "seed_x": seed_x, "seed_y": seed_y, "alpha": alpha, "beta": beta,
           "ddot": d1, "ddot1": d2}
    names = ("spmv", "waxpby", "wwpby", "wxbw", "vcopy")
    for i, nm in enumerate(names):
        rec["sha_" + nm] = sha(out[i])
        if full:
            rec[nm] = [float(v) for v in out[i]]
    for f in (fx, fy, fo):
        os.remove(f)
    return rec


def _chunks(k0, k1, parts):
    """Split [k0,k1] into contiguous ranges of roughly equal sum(K) (work of K fresh solves)."""
    ks = list(range(k0, k1 + 1))
    total = sum(ks)
    out, cur, acc = [], [], 0
    for k in ks:
        cur.append(k)
        acc += k
        if acc >= total / parts and len(out) < parts - 1:
            out.append((cur[0], cur[-1]))
            cur, acc = [], 0
    if cur:
        out.append((cur[0], cur[-1]))
    return out


def reference_b(tmp, nx, ny, nz, xstar):
    """b = A x* with the reference's own spmv (harness 'kernels' mode, first output)."""
    n = nx * ny * nz
    fx, fy, fo = (os.path.join(tmp, f"b{nx}_{ny}_{nz}.{s}") for s in ("x", "y", "out"))
    xstar.tofile(fx)
    xstar.tofile(fy)
    run([HARNESS, "kernels", str(nx), str(ny), str(nz), fx, fy, "1.0", "1.0", fo], tmp)
    with open(fo, "rb") as f:
        b = np.fromfile(f, dtype=np.float64, count=n)
    for f in (fx, fy, fo):
        os.remove(f)
    return b


def hist_case(tmp, nx, ny, nz, kmax, seed=None):
    """||r_0..r_kmax|| of reference cg, from kmax+1 fresh solves with max_iter = K."""
    extra = []
    rec = {"grid": [nx, ny, nz], "kmax": kmax, "rhs": "A*1" if seed is None else f"sines seed {seed}"}
    if seed is not None:
        xs = oracle.xstar_sines(nx, ny, nz, seed)
        b = reference_b(tmp, nx, ny, nz, xs)
        fb = os.path.join(tmp, f"h{nx}_{ny}_{nz}_{seed}.b")
        fx = os.path.join(tmp, f"h{nx}_{ny}_{nz}_{seed}.xs")
        b.tofile(fb)
        xs.tofile(fx)
        extra = [fb, fx]
        rec["seed"] = seed
        rec["sha_b"] = sha(b)
    jobs = _chunks(1, kmax + 1, WORKERS)
    hist = {}
    xerr = None
    with cf.ThreadPoolExecutor(WORKERS) as ex:
        futs = {ex.submit(run, [HARNESS, "hist", str(nx), str(ny), str(nz), str(a), str(c)] + extra, tmp): (a, c)
                for a, c in jobs}
        futs[ex.submit(run, [HARNESS, "hist", str(nx), str(ny), str(nz), str(kmax), str(kmax)] + extra, tmp)] = None
        for f in cf.as_completed(futs):
            rows = f.result()
            for r in rows:
                if r[0] == "NORMR":
                    hist[int(r[1])] = float(r[3])
                elif r[0] == "XERR" and futs[f] is None:
                    xerr = (float(r[2]), float(r[3]))
    rec["hist"] = [hist[k] for k in range(1, kmax + 2)]  # hist[j] = ||r_j||, j = 0..kmax
    rec["xerr_max"], rec["xerr_l2"] = xerr  # after the max_iter = kmax solve
    for f in extra:
        os.remove(f)
    return rec


def solve_case(tmp, nx, ny, nz, seed, rel_tol, max_iter):
    n = nx * ny * nz
    xs = oracle.xstar_sines(nx, ny, nz, seed)
    b = reference_b(tmp, nx, ny, nz, xs)
    normb = float(np.sqrt(oracle.ddot1(b)))
    fb = os.path.join(tmp, f"s{nx}.b")
    fx = os.path.join(tmp, f"s{nx}.x")
    b.tofile(fb)
    rows = run([HARNESS, "solve", str(nx), str(ny), str(nz), str(max_iter), repr(rel_tol * normb), fb, fx], tmp)
    niters, normr = int(rows[0][1]), float(rows[0][2])
    x = np.fromfile(fx, dtype=np.float64, count=n)
    A = oracle.Problem(nx, ny, nz)
    r = b - oracle.spmv(A, x)
    rec = {"grid": [nx, ny, nz], "seed": seed, "rel_tol": rel_tol, "abs_tol": rel_tol * normb,
           "max_iter": max_iter, "normb": normb, "sha_b": sha(b), "niters": niters, "normr": normr,
           "true_relres": float(np.sqrt(oracle.ddot1(r)) / normb),
           "xerr_max": float(np.max(np.abs(x - xs))),
           "xerr_l2": float(np.sqrt(oracle.ddot1(x - xs)))}
    for f in (fb, fx):
        os.remove(f)
    return rec


def exact_hist_case(nx, ny, nz, kmax):
    """||r_0..r_kmax|| of the cg loop with near-exact dots (oracle restatement), b = A*1."""
    A = oracle.Problem(nx, ny, nz)
    x, nit, _, hist, st = oracle.cg(A, max_iter=kmax, exact_dots=True)
    assert st == 0 and nit == kmax
    return {"source": "oracle.cg(exact_dots=True): the reference cg loop (numerics.F90:92-147) with "
                      "both dots rounded once from a double-double sum",
            "grid": [nx, ny, nz], "kmax": kmax, "rhs": "A*1", "hist": [float(v) for v in hist],
            "xerr_max": float(np.max(np.abs(x - 1.0)))}


def mask_timers(lines):
    """Stdout lines with the measured numbers replaced: TIMER values by <t>, MFLOPS by <r>."""
    out = []
    for ln in lines:
        if ln.strip().startswith("TIMER:"):
            ln = re.sub(r"(TIMER:\s+\S+\s+)\S+(\s+sec)", r"\1<t>\2", ln)
        elif ln.strip().startswith("MFLOPS SPARSEMV"):
            ln = re.sub(r"(MFLOPS SPARSEMV\s+)\S+", r"\1<r>", ln)
        out.append(ln)
    return out


def write_timers():
    """stdout of the reference built with -DTIMINGLEV=3 (every timer region on) at 16^3."""
    exe = os.path.join(ROOT, "oracle", "_ref", "serial3", "HPCCG")
    out = subprocess.run([exe, "16", "16", "16"], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(GOLD, "ref_stdout_t3_16x16x16.txt"), "w") as f:
        f.write("\n".join(mask_timers(out.splitlines())) + "\n")


def write_exact():
    rec = exact_hist_case(128, 128, 128, 150)
    with open(os.path.join(GOLD, "cg_exacthist_128x128x128_ones.json"), "w") as f:
        json.dump(rec, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true")
    ap.add_argument("--exact", action="store_true", help="only write the exact-dot 128^3 history")
    ap.add_argument("--timers", action="store_true", help="only write the TIMINGLEV=3 stdout fixture")
    args = ap.parse_args()
    if args.exact:
        write_exact()
        print("exact-dot history written to", GOLD)
        return
    if args.timers:
        write_timers()
        print("TIMINGLEV=3 stdout written to", GOLD)
        return
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref not built: run `make -C oracle ref` (needs the reference sources)")
    tmp = tempfile.mkdtemp(prefix="hpccg_golden_")

    gen = [gen_case(tmp, *g) for g in [(2, 2, 2), (7, 5, 6), (16, 8, 12), (32, 32, 32)]]
    kern = [kernels_case(tmp, 6, 5, 4, 101, 102, 0.7, -1.3, True),
            kernels_case(tmp, 16, 8, 12, 103, 104, -2.5, 0.3, False),
            kernels_case(tmp, 32, 32, 32, 105, 106, 1.0 / 3.0, -7.0 / 9.0, False)]
    with open(os.path.join(GOLD, "matrix_and_kernels.json"), "w") as f:
        json.dump({"source": "reference MPI/src via oracle/_ref/serial/ref_harness",
                   "generate_matrix": gen, "kernels": kern}, f, indent=1)

    hists = [hist_case(tmp, 32, 32, 32, 150), hist_case(tmp, 16, 8, 12, 150),
             hist_case(tmp, 5, 5, 5, 40), hist_case(tmp, 32, 32, 32, 150, seed=11)]
    if args.big:
        hists.append(hist_case(tmp, 128, 128, 128, 150))
    for h in hists:
        g = "x".join(str(v) for v in h["grid"])
        tag = "ones" if "seed" not in h else f"seed{h['seed']}"
        with open(os.path.join(GOLD, f"cg_hist_{g}_{tag}.json"), "w") as f:
            json.dump({"source": "reference cg via oracle/_ref/serial/ref_harness hist", **h}, f, indent=1)

    solves = [solve_case(tmp, 32, 32, 32, 11, 1e-10, 2000), solve_case(tmp, 64, 64, 64, 12, 1e-10, 2000)]
    if args.big:
        write_exact()
        solves += [solve_case(tmp, 128, 128, 128, 13, 1e-10, 2000), solve_case(tmp, 256, 256, 256, 14, 1e-10, 2000)]
    with open(os.path.join(GOLD, "cg_solve_sines.json"), "w") as f:
        json.dump({"source": "reference cg via oracle/_ref/serial/ref_harness solve, b = reference spmv(x*)",
                   "cases": solves}, f, indent=1)

    write_timers()
    out = subprocess.run([MAIN, "32", "32", "32"], check=True, capture_output=True, text=True).stdout
    lines = [ln for ln in out.splitlines() if "TIMER" not in ln]
    with open(os.path.join(GOLD, "ref_stdout_32x32x32.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    os.rmdir(tmp)
    print("golden fixtures written to", GOLD)


if __name__ == "__main__":
    main()
