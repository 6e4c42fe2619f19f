"""Writes the golden fixtures under tests/golden/ by running the reference lapw1.

Run in the build container only (needs /root/reference and the binaries that
`make -f oracle/Makefile.ref` builds into oracle/_ref/):

    python tests/golden/make_fixtures.py [case ...]

For every case it writes the seeded WIEN2k input deck (w2k_inputs.py), runs the
reference lapw1 through the dump driver (oracle/ref_dump.F90), and stores in
<case>.npz:
  run-level module state at CALKPT entry (cropped to the entries the hot path
  reads: LQNS/GFAC to LQIND, VNS* to the LM indices in use, the WARP box to
  its non-zero extent -- WARPIN returns 0 outside the box, warpin.f:31-37);
  per k-point: the basis (KZZ, XK..RK), HS/HSDIAG after HAMILT and after HNS,
  the AL/BL tables, and NE, NBELW, EIGVAL, Z after SECLR4.
The batched-k cell (c4_ortho8, 64 k-points) stores per k-point only the basis,
window, NE, NBELW and EIGVAL.
Large cells store per k-point only the basis, eigenvalues, a seeded sample of
H/S entries, per-column sums and norms (the matrices are hundreds of MB).

The MgO example shipped with the reference (SRC_elast/example-MgO/elast/eos)
is run as well; its published eigenvalues (eos.output1_1) are stored next to
the reference-built ones so tests can pin the build against published output.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import w2k_inputs as W  # noqa: E402
from w2kdump import read_dump  # noqa: E402

REF_DIR = os.environ.get("W2K_REF_DIR", os.path.join(REPO, "oracle", "_ref"))
MGO_DIR = "/root/reference/SRC_elast/example-MgO/elast/eos"

RUN_KEYS_SCALAR = ["NAT", "NDF", "LTOP", "LNSMAX", "INS", "NLO", "MULTMAX", "LMAXP",
                   "LOMAXP", "NLOATP", "NSLMAXP", "LMMXP", "NGAUP", "VI"]
RUN_KEYS_ARR = ["MULT", "POS", "RMT", "V", "P", "PE", "DP", "DPE", "PEI", "E", "LAPW",
                "LOOR", "ILO", "ALO", "BLO", "CLO", "ELO", "PLO", "DPLO", "PELO", "DPELO",
                "PEILO", "PI12LO", "PE12LO", "ROTLOC", "ROTIJ", "LQIND"]


def _unlimited_stack():
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_reference(workdir: str, complex_: bool, threads: int = 8, dump: bool = True, mode: str = "") -> float:
    exe = os.path.join(REF_DIR, "lapw1_complex" if complex_ else "lapw1_real")
    env = dict(os.environ, OPENBLAS_NUM_THREADS=str(threads))
    t0 = time.time()
    args = [exe] + ([mode] if mode else ([] if dump else ["nodump"]))
    # the reference keeps automatic arrays on the stack that grow with the
    # number of atoms (the 100-atom cell overflows the default 8 MiB)
    r = subprocess.run(args, cwd=workdir, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       preexec_fn=_unlimited_stack)
    dt = time.time() - t0
    if r.returncode != 0 or not os.path.exists(os.path.join(workdir, "ref_times.txt")):
        sys.stdout.write(r.stdout.decode(errors="replace")[-3000:])
        raise RuntimeError(f"reference lapw1 failed in {workdir}")
    return dt


def crop_run(run: dict) -> dict:
    out = {k: np.asarray(run[k]) for k in RUN_KEYS_SCALAR + RUN_KEYS_ARR}
    nat = int(run["NAT"])
    lqind = np.atleast_1d(run["LQIND"]).astype(int)
    nq = max(int(lqind.max()), 1)
    out["LQNS"] = run["LQNS"][:, :nq, :]
    out["GFAC"] = run["GFAC"][:nq, :]
    lmx_used = int(run["LQNS"][3, :nq, :].max()) if lqind.max() > 0 else 1
    lmx_used = max(lmx_used, 1)
    for k in ("VNS1", "VNS2", "VNS3", "VNS1LO", "VNS2LO", "VNS3LO"):
        out[k] = run[k][:, :lmx_used, :, :]
    # radial mesh, spherical potential (r V, Hartree), LM lists, REL
    if "JRJ" in run:
        jri = int(np.max(np.atleast_1d(run["JRJ"])))
        lmmax = np.atleast_1d(run["LMMAX"]).astype(int)
        nl = max(int(lmmax.max()), 1)
        out["JRJ"], out["DH"], out["RO"] = run["JRJ"], run["DH"], run["RO"]
        out["VR"] = run["VR"][:jri, :]
        out["LMMAX"] = run["LMMAX"]
        out["LM"] = run["LM"][:, :nl, :]
        out["REL"] = run["REL"]
    for k in ("BR2", "PIA", "ORTHO", "CXZ", "RKM"):
        if k in run:
            out[k] = run[k]
    if int(run["INS"]) == 0 and "NKK" in run:
        for k in ("IORD", "IMAT", "TAU", "NKK", "JA", "JB", "JC", "POTK"):
            out[k] = run[k]
    if int(run["INS"]) == 0:
        w = run["WARP"]
        k1, k2, k3 = int(run["WARP1"]), int(run["WARP2"]), int(run["WARP3"])
        nz = np.argwhere(w != 0)
        if len(nz) == 0:
            c1 = c2 = c3 = 0
        else:
            c = np.abs(nz - np.array([k1, k2, k3])).max(axis=0)
            c1, c2, c3 = (int(x) for x in c)
        out["WARP1"], out["WARP2"], out["WARP3"] = np.asarray(c1), np.asarray(c2), np.asarray(c3)
        out["WARP"] = w[k1 - c1:k1 + c1 + 1, k2 - c2:k2 + c2 + 1, k3 - c3:k3 + c3 + 1]
    return out


def kdata(kd: dict, full: bool, rng: np.random.Generator | None = None) -> dict:
    n = int(kd["HSROWS"])
    out = {k: np.asarray(kd[k]) for k in ("NV", "NLO", "HSROWS", "SX", "SY", "SZ", "EL", "EU", "NUME",
                                          "NE", "NBELW")}
    out["KZZ"] = kd["KZZ"][:, :n]
    for k in ("XK", "YK", "ZK", "RK"):
        out[k] = kd[k][:n]
    ne = int(kd["NE"])
    out["EIGVAL"] = np.atleast_1d(kd["EIGVAL"])[:ne]
    if "HS_HAMILT" not in kd:
        return out  # basis-only dump: no matrices
    if full:
        for k in ("HS_HAMILT", "HSDIAG_HAMILT", "HS_HNS", "HSDIAG_HNS", "AL", "BL"):
            out[k] = kd[k]
        out["Z"] = kd["Z"][:, :ne]
    else:
        for stage in ("HAMILT", "HNS"):
            hs = kd[f"HS_{stage}"]
            m = 4096
            ii = rng.integers(0, n, size=m)
            jj = rng.integers(0, n, size=m)
            out[f"SAMPLE_I_{stage}"] = ii.astype(np.int32)
            out[f"SAMPLE_J_{stage}"] = jj.astype(np.int32)
            out[f"SAMPLE_V_{stage}"] = hs[ii, jj]
            out[f"COLSUM_{stage}"] = hs.sum(axis=0)
            out[f"ROWSUM_{stage}"] = hs.sum(axis=1)
            out[f"HSDIAG_{stage}"] = kd[f"HSDIAG_{stage}"]
            out[f"MAXABS_{stage}"] = np.asarray(np.abs(hs).max())
        # eigenvectors: first/last columns only
        out["Z_FIRST"] = kd["Z"][:, 0]
        out["Z_LAST"] = kd["Z"][:, ne - 1]
    return out


def save_case(name: str, workdir: str, nk_full: int, complex_: bool, seconds: float, extra: dict | None = None):
    run = read_dump(os.path.join(workdir, "run.bin"))
    payload = {f"run/{k}": v for k, v in crop_run(run).items()}
    vlm = os.path.join(workdir, "vlm_read.npy")
    if os.path.exists(vlm):
        payload["run/VLM"] = np.load(vlm)
    ks = sorted(f for f in os.listdir(workdir) if re.match(r"k\d{4}\.bin$", f))
    rng = np.random.Generator(np.random.PCG64(W.SEED_BASE + 999))
    for i, kf in enumerate(ks):
        kd = read_dump(os.path.join(workdir, kf))
        for k, v in kdata(kd, i < nk_full, rng).items():
            payload[f"k{i}/{k}"] = v
    if name in ("c1_fcc", "lo_cplx"):
        # the reference's vector file (TAPEWF, unit 10) of every k-point
        with open(os.path.join(workdir, "case.vector"), "rb") as f:
            payload["vector/case_vector"] = np.frombuffer(f.read(), np.uint8)
    payload["meta/nk"] = np.asarray(len(ks))
    payload["meta/complex"] = np.asarray(int(complex_))
    payload["meta/ref_seconds"] = np.asarray(seconds)
    times = dict(l.split() for l in open(os.path.join(workdir, "ref_times.txt")))
    for k, v in times.items():
        payload[f"meta/{k}"] = np.asarray(float(v))
    if extra:
        payload.update(extra)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **payload)
    print(f"{name}: {len(ks)} k-points, N={[int(read_dump(os.path.join(workdir, k))['HSROWS']) for k in ks[:1]]}, "
          f"{os.path.getsize(path)/1024:.0f} KiB, reference {seconds:.2f} s")


def parse_published_eigs(path: str, nblocks: int):
    """Eigenvalue blocks ('EIGENVALUES ARE:') of a published lapw1 output."""
    txt = open(path).read().split("\n")
    blocks = []
    i = 0
    while i < len(txt) and len(blocks) < nblocks:
        if "EIGENVALUES ARE:" in txt[i]:
            vals = []
            j = i + 1
            while j < len(txt) and txt[j].strip() and re.fullmatch(r"[\s0-9.+\-Ee]+", txt[j]):
                vals += [float(x) for x in txt[j].split()]
                j += 1
            blocks.append(np.array(vals))
            i = j
        else:
            i += 1
    return blocks


def make_mgo(tmp: str):
    d = os.path.join(tmp, "mgo")
    os.makedirs(d)
    sub_struct = [("ATOMS:", "ATOMS "), ("ATOM=", "ATOM ")]
    sub_pot = [("ATOM NUMBER =", "ATOM NUMBER  "), ("ATOMNUMBER  =", "ATOMNUMBER   "),
               ("NUMBER OF LM=", "NUMBER OF LM "), ("LM(R) FOR L=", "LM(R) FOR L ")]
    # 100atomfix_struct / 100atomfix_clm edits (100atomfix_struct:9-14, 100atomfix_clm:8-15)
    for ext, subs in (("struct", sub_struct), ("vsp", sub_pot), ("vns", sub_pot)):
        s = open(os.path.join(MGO_DIR, f"eos.{ext}")).read()
        for a, b in subs:
            s = s.replace(a, b)
        open(os.path.join(d, f"case.{ext}"), "w").write(s)
    shutil.copy(os.path.join(MGO_DIR, "eos.in1"), os.path.join(d, "case.in1"))
    # first 12 k-points of the published list (output1_1 holds k 1..55)
    kl = open(os.path.join(MGO_DIR, "eos.klist")).read().split("\n")
    open(os.path.join(d, "case.klist"), "w").write("\n".join(kl[:12] + ["END"]) + "\n")
    open(os.path.join(d, "lapw1.def"), "w").write(W.DEF_TEMPLATE.format(c="case"))
    sec = run_reference(d, False)
    pub = parse_published_eigs(os.path.join(MGO_DIR, "eos.output1_1"), 12)
    extra = {f"published/k{i}": b for i, b in enumerate(pub)}
    save_case("mgo", d, 1, False, sec, extra)


def make_c3(name: str, workdir: str, seconds: float, frac: float = 0.2):
    """Fixture of the 100-atom complex cell from one wide-window run: the
    window is narrowed to the lowest `frac` of the spectrum after the fact."""
    run = read_dump(os.path.join(workdir, "run.bin"))
    payload = {f"run/{k}": v for k, v in crop_run(run).items()}
    kd = read_dump(os.path.join(workdir, "k0001.bin"))
    n = int(kd["HSROWS"])
    ev = np.atleast_1d(kd["EIGVAL"])[: int(kd["NE"])]
    target = int(round(frac * n))
    assert len(ev) > target, (len(ev), target)
    kd["EU"] = np.asarray(0.5 * (ev[target - 1] + ev[target]))
    kd["NE"] = np.asarray(target)
    rng = np.random.Generator(np.random.PCG64(W.SEED_BASE + 999))
    out = kdata(kd, False, rng)
    out["Z_LAST"] = kd["Z"][:, target - 1]
    out["Z_MID"] = kd["Z"][:, target // 2]
    out["Z_MID_INDEX"] = np.asarray(target // 2)
    for k, v in out.items():
        payload[f"k0/{k}"] = v
    payload["meta/nk"] = np.asarray(1)
    payload["meta/complex"] = np.asarray(1)
    payload["meta/ref_seconds"] = np.asarray(seconds)
    times = dict(l.split() for l in open(os.path.join(workdir, "ref_times.txt")))
    for k, v in times.items():
        payload[f"meta/{k}"] = np.asarray(float(v))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **payload)
    print(f"{name}: N={n}, NE={target}, EU={float(kd['EU']):.6f}, {os.path.getsize(path)/1024:.0f} KiB, "
          f"reference {seconds:.1f} s")


def add_lattice(name: str, tmp: str):
    """Add COORS's lattice (BR2, PIA, ORTHO, CXZ, RKM) to an existing fixture
    from a COORS-only pass of the reference over the same deck ('coors'
    dump: run state and bases, no HAMILT/HNS/SECLR4), after checking that the
    pass reproduces the stored bases."""
    case = W.CASES[name]()
    d = os.path.join(tmp, name + "_coors")
    W.write_case(case, d)
    run_reference(d, case.complex, mode="coors")
    run = read_dump(os.path.join(d, "run.bin"))
    path = os.path.join(HERE, f"{name}.npz")
    z = np.load(path, allow_pickle=False)
    payload = {k: z[k] for k in z.files}
    for i in range(int(payload["meta/nk"])):
        kd = read_dump(os.path.join(d, f"k{i + 1:04d}.bin"))
        n = int(kd["HSROWS"])
        assert np.array_equal(kd["KZZ"][:, :n], payload[f"k{i}/KZZ"]), f"{name} k{i}: bases differ"
    for k in ("BR2", "PIA", "ORTHO", "CXZ", "RKM"):
        payload[f"run/{k}"] = np.asarray(run[k])
    np.savez_compressed(path, **payload)
    print(f"{name}: lattice added ({os.path.getsize(path)/1024:.0f} KiB)")


def main(argv):
    names = argv or ["mgo", "c1_fcc", "lo_mult2", "lo_cplx", "rs8"]
    # W2K_FIXTURE_WORKDIR: run the reference there and keep its outputs (the
    # 100-atom cell takes about an hour)
    keep = os.environ.get("W2K_FIXTURE_WORKDIR")
    tmp = keep or tempfile.mkdtemp(prefix="w2kfix_")
    os.makedirs(tmp, exist_ok=True)
    try:
        for name in names:
            if name.startswith("lattice:"):
                add_lattice(name.split(":", 1)[1], tmp)
                continue
            if name == "mgo":
                make_mgo(tmp)
                continue
            case = W.CASES[name]()
            d = os.path.join(tmp, name)
            if name == "c4_ortho8":
                # many k-points: the run state, bases and eigenvalues only
                # ('basis' dump); window from the first k-point so that the
                # lowest ~10% of its spectrum is computed
                first = W.CASES[name](nk=1)
                W.write_case(first, d)
                run_reference(d, True, mode="basis")
                kd = read_dump(os.path.join(d, "k0001.bin"))
                ev = np.atleast_1d(kd["EIGVAL"])
                target = int(round(0.1 * int(kd["HSROWS"])))
                case.window = (case.window[0], float(0.5 * (ev[target - 1] + ev[target])))
                shutil.rmtree(d)
                W.write_case(case, d)
                sec = run_reference(d, True, mode="basis")
                save_case(name, d, 0, True, sec)
                continue
            if name == "c3_lowsym100":
                # one pass with a wide window capped at NUME (4000, the lowest
                # eigenvalues: the tridiagonal matrix does not split), then
                # the window (EL, EU'] of the lowest ~20% of the spectrum is
                # cut from it: same matrices, same eigenvalues (bisection to
                # ABSTOL = 0 does not depend on the window)
                done = os.path.join(d, "ref_times.txt")
                if keep and os.path.exists(done) and os.path.exists(os.path.join(d, "k0001.bin")):
                    sec = float(dict(l.split() for l in open(done))["total_wall_s"])  # kept run
                else:
                    W.write_case(case, d)
                    sec = run_reference(d, True)
                make_c3(name, d, sec)
                continue
            if name == "c2_rs64":
                # window chosen so that the lowest ~20% of the spectrum is
                # computed ("all occupied eigenpairs"): a first pass with a
                # wide window gives the spectrum, the second uses (EL, EU]
                W.write_case(case, d)
                run_reference(d, case.complex, dump=True)
                kd = read_dump(os.path.join(d, "k0001.bin"))
                target = int(round(0.2 * int(kd["HSROWS"])))
                ev = np.atleast_1d(kd["EIGVAL"])
                assert len(ev) > target, (len(ev), target)
                case.window = (case.window[0], float(0.5 * (ev[target - 1] + ev[target])))
                shutil.rmtree(d)
            W.write_case(case, d)
            sec = run_reference(d, case.complex)
            nk_full = {"c2_rs64": 0, "c5_cubic16": 0, "c1_fcc": 2}.get(name, 1)
            save_case(name, d, nk_full, case.complex, sec)
    finally:
        if not keep:
            shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main(sys.argv[1:])
