#!/usr/bin/env python3
"""Seconds per k-point of lapw1's hot path (HAMILT + HNS + SECLR4) on one MI355X,
plain FP64 and with the split-precision HNS update (INT8 digit slices).

Workload: the 64-atom rocksalt supercell (2x2x2, Mg/O, RMT 2.0/1.6, RKmax 7,
real/inversion-symmetric, one k-point at Gamma, N = 5951 = 5695 plane waves +
256 local orbitals, lowest ~20% of the spectrum = 1184 eigenpairs), built from
the seeded WIEN2k input deck whose reference outputs are stored in
tests/golden/c2_rs64.npz.  One step is one k-point: HAMILT, the diagonal swap,
HNS and SECLR4 through the library's C-ABI with the per-run tables and the
basis already resident on the device (uploaded before the timed region).

The CPU baseline is the reference lapw1 itself (compiled from its sources by
oracle/Makefile.ref into oracle/_ref/), run on this host on the same deck for
the same k-point, timed by its own phase clocks; it runs before the GPU is
touched.

The k-point is timed twice: with every product in FP64 ("fp64") and with the
HNS rank-K update in split precision (--split-slices INT8 digits per element,
FP64 recombination).  `value` is the split-precision time when its
eigenvalues stay within 1e-8 Ry of the reference with identical NE/NBELW
(checked in the run), else the FP64 time; both are in the line.  Prints one
JSON line.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import resource
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(REPO, "tests", "golden", "c2_rs64.npz")
REF_EXE = os.path.join(REPO, "oracle", "_ref", "lapw1_real")

METRIC = "s per k-point (H/S setup+eigensolve), split-prec vs FP64 at ΔE≤1e-8 Ry"
PEAK_FP64_TFLOPS = 78.6
PEAK_INT8_TOPS = 5000.0
DE_BAR = 1e-8  # Ry: split precision must stay this close to the reference eigenvalues
PEAK_HBM_GBS = 8000.0


def load_package():
    name = "hack_wein2k_amd"
    spec = importlib.util.spec_from_file_location(name, os.path.join(REPO, "hack-wein2k_amd", "__init__.py"),
                                                  submodule_search_locations=[os.path.join(REPO, "hack-wein2k_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_workload(L):
    raw = np.load(FIXTURE)
    run = L.RunData({k[4:]: raw[k] for k in raw.files if k.startswith("run/")}, bool(int(raw["meta/complex"])))
    kd = {k[3:]: raw[k] for k in raw.files if k.startswith("k0/")}
    basis = L.Basis(int(kd["NV"]), int(kd["NLO"]), kd["KZZ"], kd["XK"], kd["YK"], kd["ZK"], kd["RK"])
    return run, basis, kd


def cpu_baseline(kd, threads: int):
    """The reference lapw1 on the same deck (same seeded inputs, same window)."""
    if not os.path.exists(REF_EXE):
        return None
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    import w2k_inputs as W
    case = W.CASES["c2_rs64"]()
    case.window = (float(kd["EL"]), float(kd["EU"]))
    with tempfile.TemporaryDirectory(prefix="w2k_ref_") as d:
        W.write_case(case, d)
        env = dict(os.environ, OPENBLAS_NUM_THREADS=str(threads), OMP_NUM_THREADS=str(threads))

        def unlimited_stack():
            resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

        t0 = time.time()
        r = subprocess.run([REF_EXE, "nodump"], cwd=d, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           preexec_fn=unlimited_stack, timeout=900)
        wall = time.time() - t0
        tf = os.path.join(d, "ref_times.txt")
        if r.returncode != 0 or not os.path.exists(tf):
            sys.stderr.write(r.stdout.decode(errors="replace")[-2000:])
            return None
        t = {k: float(v) for k, v in (l.split() for l in open(tf))}
    per_k = t["hamilt_wall_s"] + t["hns_wall_s"] + t["seclr4_wall_s"]
    return {
        "value": round(per_k, 3),
        "unit": "s/k-point",
        "cores": threads,
        "kind": "reference",
        "sample": (f"the whole workload: one k-point of the 64-atom rocksalt deck through the reference lapw1 "
                   f"(flang build of SRC_lapw1); HAMILT {t['hamilt_wall_s']:.2f} s and HNS {t['hns_wall_s']:.2f} s "
                   f"on 1 thread (serial in the reference), SECLR4 {t['seclr4_wall_s']:.2f} s with OpenBLAS on "
                   f"{threads} threads; process wall {wall:.1f} s"),
    }


def _progress(msg: str):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def copy_calibration(L):
    """A plain device copy of 512 MiB (3 launches): the reference point that
    scales the profiler's FETCH_SIZE / WRITE_SIZE counters to HBM bytes."""
    import ctypes
    lib = L.load_library()
    lib.w2k_test_copy.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    n = 64 << 20
    ms = ctypes.c_double(0.0)
    rc = lib.w2k_test_copy(n, 3, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"w2k_test_copy failed ({rc})")
    return {"kernel": "k_copy", "bytes_read": 8 * n, "bytes_written": 8 * n, "ms": round(ms.value, 4),
            "GB/s": round(16.0 * n / (ms.value * 1e-3) / 1e9, 1)}


TRAFFIC_FILE = os.path.join(REPO, "profiles", "c2_rs64_hbm_counters.json")


def measured_traffic(kname: str):
    """HBM bytes per launch of the roofline kernel from the committed counter
    passes (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each scaled by the
    same counter's reading on the calibration copy), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        rec = json.load(f).get("kernels", {}).get(kname)
    return None if rec is None else rec.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="skip the extra k-point with per-launch SYMV events (for profiler passes)")
    ap.add_argument("--split-slices", type=int, default=4,
                    help="INT8 digits per element for the split-precision pass (0: FP64 only)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-k measurement (64 k-points of the 8-atom cell)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 512 MiB device copy (calibrates the profiler's byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(k-points shard across GPUs with no exchange; run one process per GPU)", file=sys.stderr)
        sys.exit(2)

    pkg = load_package()
    L = pkg.lapw1
    run, basis, kd = load_workload(L)
    el, eu, nume = float(kd["EL"]), float(kd["EU"]), int(kd["NUME"])

    cpu = None if args.no_cpu_baseline else cpu_baseline(kd, args.cpu_threads)

    copy = None
    if args.copy_calibration:
        copy = copy_calibration(L)
        _progress(f"copy calibration: {copy['GB/s']} GB/s")
    ctx = L.Lapw1(run)
    ctx.set_basis(basis)
    names = L.Lapw1.TIMING_NAMES

    def timed(slices: int):
        """W warmup and K timed k-points with the HNS rank-K update in FP64
        (slices = 0) or in split precision (INT8 digit slices)."""
        ctx.set_split_precision(slices)
        for _ in range(args.warmup):
            ctx.calkpt(el, eu, nume)
        ctx.sync()
        acc = np.zeros(len(names))
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ne, nb = ctx.calkpt(el, eu, nume)
            acc += ctx.timings()
        ctx.sync()
        t1 = time.perf_counter()
        ev, _ = ctx.get_eig()
        tag = "FP64" if slices == 0 else f"split precision ({slices} digits)"
        _progress(f"{args.warmup} warmup + {args.steps} timed k-point(s) done, {tag}")
        return {"ms": (t1 - t0) * 1e3 / args.steps, "dev": acc / args.steps, "ne": ne, "nb": nb, "ev": ev}

    fp = timed(0)
    sp = timed(args.split_slices) if args.split_slices > 0 else None
    ctx.set_split_precision(0)
    tk = None
    if not args.no_kernel_timing:
        # one more k-point, outside the timed region, with per-launch events
        # around the tridiagonalisation SYMV (its kernel-level roofline)
        ctx.set_kernel_timing(True)
        ctx.calkpt(el, eu, nume)
        tk = dict(zip(names, ctx.timings()))
        ctx.set_kernel_timing(False)
    ctx.close()

    def check(r):
        ok = r["ne"] == int(kd["NE"]) and r["nb"] == int(kd["NBELW"])
        return ok, (float(np.max(np.abs(r["ev"] - kd["EIGVAL"]))) if ok else float("nan"))

    ok, dmax = check(fp)
    ne, nb, dev, ms_fp64 = fp["ne"], fp["nb"], fp["dev"], fp["ms"]
    split = None
    ms_step = ms_fp64
    if sp is not None:
        sok, sd = check(sp)
        de_fp = float(np.max(np.abs(sp["ev"] - fp["ev"]))) if sok else float("nan")
        usable = sok and sd <= DE_BAR
        split = {"slices": args.split_slices, "ms_per_step": round(sp["ms"], 2),
                 "max_abs_eig_dev_vs_reference_ry": sd, "max_abs_eig_dev_vs_fp64_ry": de_fp,
                 "ne": sp["ne"], "nbelw": sp["nb"], "meets_de_bar": usable,
                 "hns_rank_k_ms": round(float(sp["dev"][names.index("hns_gemm")]), 3)}
        if usable:
            ms_step = sp["ms"]

    nv, nlo, n = basis.nv, basis.nlo, basis.n
    nat, lt = run.nat, run.ltop
    sum_lindex = _sum_lindex(run)
    f_hamilt = 8.0 * nv * nv * nat * lt
    f_hns = 4.0 * n * n * sum_lindex
    t = dict(zip(names, dev))
    kernels = {
        "k_hamilt_pw": ("mfma", f_hamilt, t["hamilt_pw_kernel"], 1),
        "k_dgemm (HNS rank-K)": ("mfma", f_hns, t["hns_gemm"], 1),
    }
    if tk is not None and tk["tri_symv_launches"] > 0:
        # algorithmic bytes of the SYMV at step i: the lower triangle
        # (diagonal included) of the trailing order m = n - i - 1, FP64
        nsym = int(tk["tri_symv_launches"])
        sym_bytes = sum(8.0 * m * (m + 1) / 2 for m in range(n - 1, n - 1 - nsym, -1))
        kernels["k_tri_symv"] = ("hbm", sym_bytes, tk["tri_symv_total"], nsym)
    kname, (bound, work, kms, nl) = max(kernels.items(), key=lambda kv: kv[1][2])
    if bound == "hbm":
        achieved, peak, unit = work / (kms * 1e-3) / 1e9, PEAK_HBM_GBS, "GB/s"
    else:
        achieved, peak, unit = work / (kms * 1e-3) / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s"
    others = {}
    for kn, (bd, wk, ms_, nl_) in kernels.items():
        if kn == kname or ms_ <= 0:
            continue
        if bd == "hbm":
            others[kn] = {"ms": round(ms_, 3), "GB/s": round(wk / (ms_ * 1e-3) / 1e9, 1)}
        else:
            others[kn] = {"ms": round(ms_, 3), "TFLOP/s": round(wk / (ms_ * 1e-3) / 1e12, 2)}
    if split is not None and split["hns_rank_k_ms"] > 0:
        # the same FP64-equivalent work; the INT8 rate counts the S(S+1)/2
        # digit-pair products actually issued
        sl, tms = split["slices"], split["hns_rank_k_ms"]
        others["k_split_gemm (HNS rank-K, INT8 digits)"] = {
            "ms": tms, "FP64-equivalent TFLOP/s": round(f_hns / (tms * 1e-3) / 1e12, 2),
            "INT8 TOP/s": round(f_hns * sl * (sl + 1) / 2 / (tms * 1e-3) / 1e12, 1), "INT8 peak TOP/s": PEAK_INT8_TOPS}
    use_split = split is not None and split["meets_de_bar"]
    out = {
        "metric": METRIC,
        "value": round(ms_step / 1e3, 4),
        "unit": "s/k-point",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not use_split else "f64 (HNS rank-K: int8 digit slices, FP64 recombination)",
        "data": "synthetic: seeded WIEN2k input deck (tests/golden/w2k_inputs.py)",
        "config": {"workload": f"64-atom rocksalt 2x2x2 (Mg/O), real, Gamma, RKmax 7, N={n} (NV {nv} + NLO {nlo}), "
                               f"NE={ne}", "nat": nat, "lt": lt, "sum_lindex": sum_lindex},
        "roofline": {"kernel": kname, "bound": bound, "achieved": round(achieved, 2), "peak": peak,
                     "unit": unit, "frac": round(achieved / peak, 4), "traffic": measured_traffic(kname),
                     "work_per_launch": work / nl, "ms_per_launch": round(kms / nl, 4), "launches": nl,
                     "other_kernels": others},
        "cpu_baseline": cpu,
        "phases_ms": {k: round(v, 2) for k, v in t.items() if k != "tri_launch_mode"},
        "tri_launch_mode": {0: "direct", 1: "hip graph replay", 2: "hip graph capture"}.get(int(t.get("tri_launch_mode", 0)), "?"),
        "check": {"ne": ne, "nbelw": nb, "ne_ref": int(kd["NE"]), "max_abs_eig_dev_ry": dmax},
        "fp64": {"value": round(ms_fp64 / 1e3, 4), "ms_per_step": round(ms_fp64, 2), "max_abs_eig_dev_ry": dmax},
        "split_precision": split,
        "value_is": "split precision" if use_split else "fp64",
    }
    if copy is not None:
        out["copy_calibration"] = copy
    if not args.no_batched:
        out["batched"] = batched_kpoints(L)
    print(json.dumps(out))


BATCH_FIXTURE = os.path.join(REPO, "tests", "golden", "c4_ortho8.npz")


def batched_kpoints(L):
    """64 k-points of the 8-atom orthorhombic complex cell (N 1936-1978):
    a serial loop of single k-point calls against one batched call, both over
    every k-point with its own window; eigenvalues checked against the
    reference (1e-10 Ry) and the batch against the loop (bit-identical)."""
    raw = np.load(BATCH_FIXTURE)
    run = L.RunData({k[4:]: raw[k] for k in raw.files if k.startswith("run/")}, bool(int(raw["meta/complex"])))
    nk = int(raw["meta/nk"])
    kds = [{k[len(f"k{q}/"):]: raw[k] for k in raw.files if k.startswith(f"k{q}/")} for q in range(nk)]
    bases = [L.Basis(int(kd["NV"]), int(kd["NLO"]), kd["KZZ"], kd["XK"], kd["YK"], kd["ZK"], kd["RK"]) for kd in kds]
    el = [float(kd["EL"]) for kd in kds]
    eu = [float(kd["EU"]) for kd in kds]
    nume = max(int(kd["NUME"]) for kd in kds)
    ctx = L.Lapw1(run)
    # serial loop, largest basis first (one context, one k-point per call)
    order = sorted(range(nk), key=lambda q: -bases[q].n)
    ctx.set_basis(bases[order[0]])
    ctx.calkpt(el[order[0]], eu[order[0]], nume)  # warm-up (workspaces)
    ctx.sync()
    ser = [None] * nk
    t0 = time.perf_counter()
    for q in order:
        ctx.set_basis(bases[q])
        ne, nb = ctx.calkpt(el[q], eu[q], nume)
        ev, z = ctx.get_eig()
        ser[q] = (ne, nb, ev, z)
    t_serial = time.perf_counter() - t0
    _progress(f"batched-k cell: serial loop {t_serial:.3f} s")
    big = order[:8]  # warm-up (worker streams sized by the largest bases)
    ctx.calkpt_batched([bases[q] for q in big], [el[q] for q in big], [eu[q] for q in big], nume)
    t0 = time.perf_counter()
    bat = ctx.calkpt_batched(bases, el, eu, nume)
    t_batch = time.perf_counter() - t0
    ctx.close()
    _progress(f"batched-k cell: one batched call {t_batch:.3f} s")
    dmax = 0.0
    same = True
    for q in range(nk):
        ne, nb, ev, z = bat[q]
        ok = ne == int(kds[q]["NE"]) and nb == int(kds[q]["NBELW"])
        dmax = max(dmax, float(np.max(np.abs(ev - kds[q]["EIGVAL"]))) if ok else float("inf"))
        same = same and ser[q][0] == ne and np.array_equal(ser[q][2], ev) and np.array_equal(ser[q][3], z)
    ns = [b.n for b in bases]
    return {"workload": f"8-atom orthorhombic cell, 2 species, complex, {nk} k-points, N {min(ns)}-{max(ns)}, "
                        f"NE {min(int(k['NE']) for k in kds)}-{max(int(k['NE']) for k in kds)}",
            "serial_s_per_kpoint": round(t_serial / nk, 4), "batched_s_per_kpoint": round(t_batch / nk, 4),
            "speedup": round(t_serial / t_batch, 2), "max_abs_eig_dev_ry": dmax, "batched_equals_serial": same,
            "cpu_reference_s_per_kpoint_8_threads_build_host": round(float(raw["meta/ref_seconds"]) / nk, 3)}


def _sum_lindex(run) -> int:
    """sum over all equivalent atoms of the number of distinct (L0,M0) of HNS."""
    lqns = run.arr["lqns"]
    lqind = run.arr["lqind"]
    mult = run.arr["mult"]
    tot = 0
    for a in range(run.nat):
        seen = set()
        for q in range(int(lqind[a])):
            seen.add((int(lqns[2, q, a]), int(lqns[5, q, a])))
        tot += len(seen) * int(mult[a])
    return tot


if __name__ == "__main__":
    main()
