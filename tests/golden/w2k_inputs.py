"""Seeded synthetic WIEN2k_07 input decks (struct, in1, klist, vsp, vns, def).

Test infrastructure: writes the files the reference lapw1 reads, in the fixed
formats of its readers, so that the reference (built under oracle/_ref) can
produce golden vectors on cells whose size matches the configurations of the
lapw1 hot path.  Formats follow:

  struct   inilpw.f:227-317 with FORMAT 5000-5060 (inilpw.f:533-541) and
           latgen.f:311-315 (IORD, 3x(3I2,F11.8) per operation); the header
           spelling is the one the readers accept after 100atomfix_struct.
  in1      inilpw.f:324-351 (MODUS, RKmax/lmax/lnsmax), atpar.f:134-208
           (per-atom global E, per-l lines FORMAT 5000 = (1X,I1,2F10.5,A4,i2);
           a repeated l adds a local orbital), inilpw.f:399 (FORMAT 5090).
  klist    FORMAT 5101 (A10,4I10,3F5.2,A3), terminated by 'END'.
  vsp      atpar.f:92-101 (FORMAT 5070/5010/5020/5060/5040/5050): r*V00 in Ry.
  vns      atpar.f:352-375 (VLM(r) per LM, Ry) followed by the interstitial
           star list read by rdswar.f:19-29.
  def      the lapw1 unit map written by x_lapw:311-323.

Potential recipe (all in Ry): r*V00(r) = -2Z[(1-w)exp(-kr) + w exp(-k'r)]
+ r*sum_k c_k exp(-a_k r^2); V_LM(r) (L>=1) of amplitude ~1e-2 Ry shaped as
(r/R)^L; interstitial potential V(r) with 100 seeded stars V(G_s) = A u_s (G_1/|G_s|)^2,
real when the cell has inversion and with V(-G) = conj V(G) otherwise,
written as its warped form (V Theta)(Q) for all |Q| <= 2 Kmax.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

SEED_BASE = 20261202
NPT = 781
R0 = 1.0e-4


@dataclasses.dataclass
class AtomType:
    name: str
    z: float
    rmt: float
    positions: List[Tuple[float, float, float]]  # fractional, MULT = len(positions)
    iatnr: int = -1  # <0: non-cubic site (atpar.f:423-441)
    # local orbitals: list of (l, E) -> extra 'l' lines in in1 (atpar.f:181-188)
    los: List[Tuple[int, float]] = dataclasses.field(default_factory=list)
    apw_lo: List[int] = dataclasses.field(default_factory=list)  # l values run as APW+lo
    vlm: List[Tuple[int, int]] = dataclasses.field(default_factory=list)  # (L, M) of V_LM


@dataclasses.dataclass
class Case:
    name: str
    title: str
    lattice: str  # 'F', 'P', 'B'
    alat: Tuple[float, float, float]
    atoms: List[AtomType]
    ops: List[Tuple[np.ndarray, np.ndarray]]  # (3x3 int acting on fractional coords, tau)
    rkmax: float = 7.0
    lmax: int = 10
    lnsmax: int = 4
    eglob: float = 0.30
    kpoints: List[Tuple[int, int, int, int]] = dataclasses.field(default_factory=lambda: [(0, 0, 0, 1)])
    window: Tuple[float, float] = (-9.0, 5.0)  # (EL, EU] in Ry
    nstars: int = 100
    seed: int = SEED_BASE
    complex: bool = False
    alpha: Tuple[float, float, float] = (90.0, 90.0, 90.0)
    vns_amp: float = 1.0e-2
    potk_amp: float = 0.05
    # False: an empty case.vns, so that ATPAR's read of it ends at once and
    # INS = IOSTAT /= 0 (atpar.f:105): no non-spherical potential and no
    # warped interstitial box, HAMILT takes its USTPHX branch (hamilt.F:348-364)
    vns: bool = True


def ops_identity() -> List[Tuple[np.ndarray, np.ndarray]]:
    return [(np.eye(3, dtype=int), np.zeros(3))]


def ops_inversion() -> List[Tuple[np.ndarray, np.ndarray]]:
    return [(np.eye(3, dtype=int), np.zeros(3)), (-np.eye(3, dtype=int), np.zeros(3))]


def ops_oh() -> List[Tuple[np.ndarray, np.ndarray]]:
    """The 48 signed permutation matrices (point group m-3m), no translations."""
    import itertools
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), dtype=int)
            for i, (p, s) in enumerate(zip(perm, signs)):
                m[i, p] = s
            out.append((m, np.zeros(3)))
    # identity first, as in WIEN2k struct files
    out.sort(key=lambda t: (0 if (t[0] == np.eye(3, dtype=int)).all() else 1))
    return out


# ----------------------------------------------------------------------------
# lattice helpers.  Reciprocal vectors are integer triples in the conventional
# axes (the KZZ convention of coors.f for P, F and B cells: all-even/all-odd
# for F, h+k+l even for B), |G| = 2 pi sqrt((h/a)^2 + (k/b)^2 + (l/c)^2).
def rvfac(case: "Case") -> float:
    return {"F": 4.0, "B": 2.0}.get(case.lattice[0], 1.0)


def cell_volume(case: "Case") -> float:
    a, b, c = case.alat
    return a * b * c / rvfac(case)


def gvectors(case: "Case", gmax: float) -> Tuple[np.ndarray, np.ndarray]:
    a, b, c = case.alat
    n1, n2, n3 = (int(gmax * x / (2 * math.pi)) + 1 for x in (a, b, c))
    g = np.array(np.meshgrid(np.arange(-n1, n1 + 1), np.arange(-n2, n2 + 1), np.arange(-n3, n3 + 1),
                             indexing="ij")).reshape(3, -1).T
    if case.lattice[0] == "F":
        par = g % 2
        g = g[(par[:, 0] == par[:, 1]) & (par[:, 1] == par[:, 2])]
    elif case.lattice[0] == "B":
        g = g[g.sum(axis=1) % 2 == 0]
    glen = 2 * math.pi * np.sqrt((g[:, 0] / a) ** 2 + (g[:, 1] / b) ** 2 + (g[:, 2] / c) ** 2)
    keep = glen <= gmax + 1e-12
    return g[keep], glen[keep]


def step_function(case: "Case", q: np.ndarray) -> np.ndarray:
    """Theta(Q) exactly as HAMILT builds the interstitial overlap
    (hamilt.F:289-331): 1 - sum_a V_a MULT_a at Q=0, otherwise
    sum_a 3 V_a (x cos x - sin x)/x^3 sum_eq exp(2 pi i Q.tau), x = |Q| R_a."""
    a, b, c = case.alat
    om = cell_volume(case)
    qlen = 2 * math.pi * np.sqrt((q[:, 0] / a) ** 2 + (q[:, 1] / b) ** 2 + (q[:, 2] / c) ** 2)
    th = np.zeros(len(q), np.complex128)
    zero = qlen < 1e-12
    for at in case.atoms:
        va = 4 * math.pi * at.rmt ** 3 / (3 * om)
        x = np.where(zero, 1.0, qlen * at.rmt)
        besr = (x * np.cos(x) - np.sin(x)) / x ** 3
        ph = np.zeros(len(q), np.complex128)
        for p in at.positions:
            ph += np.exp(2j * math.pi * (q @ np.asarray(p)))
        th += np.where(zero, -va * len(at.positions), 3 * va * besr * ph)
    th[zero] += 1.0
    return th


def kmax(case: "Case") -> float:
    return case.rkmax / min(at.rmt for at in case.atoms)


def _fmt_e(x: float) -> str:
    """A 19-character E-format field accepted by Fortran Ew.d input."""
    return f"{x:19.12E}"


def write_struct(case: Case, path: str) -> None:
    lines = [case.title.ljust(80)[:80]]
    nat = len(case.atoms)
    lines.append(f"{case.lattice:<4}LATTICE,NONEQUIV.ATOMS {nat:3d}")
    lines.append("MODE OF CALC=RELA unit=bohr")
    a, b, c = case.alat
    al, be, ga = case.alpha
    lines.append(f"{a:10.6f}{b:10.6f}{c:10.6f}{al:10.6f}{be:10.6f}{ga:10.6f}")
    for it, at in enumerate(case.atoms, start=1):
        iat = -it if at.iatnr < 0 else it
        x, y, z = at.positions[0]
        lines.append(f"ATOM{iat:4d}: X={x:10.8f} Y={y:10.8f} Z={z:10.8f}")
        lines.append(f"          MULT={len(at.positions):2d}          ISPLIT= 8")
        for (x, y, z) in at.positions[1:]:
            lines.append(f"    {iat:4d}: X={x:10.8f} Y={y:10.8f} Z={z:10.8f}")
        lines.append(f"{at.name:<10} NPT={NPT:5d}  R0={R0:10.8f} RMT={at.rmt:10.4f}   Z:{at.z:10.5f}")
        lines.append("LOCAL ROT MATRIX:    1.0000000 0.0000000 0.0000000")
        lines.append("                     0.0000000 1.0000000 0.0000000")
        lines.append("                     0.0000000 0.0000000 1.0000000")
    lines.append(f"{len(case.ops):4d}      NUMBER OF SYMMETRY OPERATIONS")
    for k, (m, tau) in enumerate(case.ops, start=1):
        for j in range(3):
            lines.append(f"{m[j,0]:2d}{m[j,1]:2d}{m[j,2]:2d}{tau[j]:11.8f}")
        lines.append(f"{k:8d}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_in1(case: Case, path: str) -> None:
    lines = ["WFFIL        (WFPRI, SUPWF)"]
    lines.append(f"{case.rkmax:6.2f}{case.lmax:10d}{case.lnsmax:5d} (R-MT*K-MAX; MAX L IN WF, V-NMT")
    for at in case.atoms:
        llines = []
        done = set()
        # exceptions for l = 0..3: LAPW/APW+lo at the global energy, then LOs
        for l in range(4):
            has_lo = [e for (ll, e) in at.los if ll == l]
            if not has_lo and l not in at.apw_lo:
                continue
            iapw = 1 if l in at.apw_lo else 0
            llines.append(f" {l:1d}{case.eglob:10.5f}{0.0:10.5f}CONT{iapw:2d}")
            for e in has_lo:
                llines.append(f" {l:1d}{e:10.5f}{0.0:10.5f}CONT 0")
            done.add(l)
        lines.append(f"{case.eglob:6.2f}{len(llines):5d}  0   (GLOBAL E-PARAMETER WITH n OTHER CHOICES, global APW/LAPW)")
        lines.extend(llines)
    lines.append(f"K-VECTORS FROM UNIT:4{case.window[0]:10.1f}{case.window[1]:10.1f}   emin/emax window")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_klist(case: Case, path: str) -> None:
    lines = []
    for i, (ix, iy, iz, idv) in enumerate(case.kpoints, start=1):
        name = f"{i:10d}"
        if i == 1:
            lines.append(f"{name}{ix:10d}{iy:10d}{iz:10d}{idv:10d}{1.0:5.2f}{case.window[0]:5.1f}{case.window[1]:5.1f}")
        else:
            lines.append(f"{name}{ix:10d}{iy:10d}{iz:10d}{idv:10d}{1.0:5.2f}")
    lines.append("END")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def radial_mesh(rmt: float) -> np.ndarray:
    dx = math.log(rmt / R0) / (NPT - 1)
    return R0 * np.exp(dx * np.arange(NPT))


def write_vsp(case: Case, path: str, rng: np.random.Generator) -> None:
    out = ["     TOTAL SPHERICAL POTENTIAL IN MT SPHERES       1.ITERATION", "     NORM:  V*R", ""]
    species_cache = {}
    for it, at in enumerate(case.atoms, start=1):
        key = (at.name, at.z, at.rmt)
        if key not in species_cache:
            c = rng.uniform(-0.5, 0.5, size=4)
            a = rng.uniform(0.5, 4.0, size=4)
            species_cache[key] = (c, a)
        c, a = species_cache[key]
        r = radial_mesh(at.rmt)
        w = 0.15
        rv = -2.0 * at.z * ((1 - w) * np.exp(-1.6 * r) + w * np.exp(-0.45 * r))
        rv += r * (c[None, :] * np.exp(-a[None, :] * r[:, None] ** 2)).sum(axis=1)
        out.append(f"   ATOMNUMBER  {it:3d}")
        out.append("   NUMBER OF LM  1")
        out.append("")
        out.append("")
        out.append("   VLM(R) FOR L  0   M= 0")
        out.append("")
        for j in range(0, NPT, 4):
            out.append("   " + "".join(_fmt_e(v) for v in rv[j:j + 4]))
        out.extend([""] * 6)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def _stars(case: Case, g: np.ndarray, glen: np.ndarray):
    """Group reciprocal vectors into stars under the point operations of the
    struct (reciprocal action m^{-T}); returns (representative, member list)."""
    rops = [np.rint(np.linalg.inv(m).T).astype(int) for (m, _) in case.ops]
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], np.round(glen, 9)))
    seen = set()
    out = []
    for idx in order:
        t = tuple(int(x) for x in g[idx])
        if t in seen:
            continue
        members = sorted({tuple(int(x) for x in (r @ g[idx])) for r in rops})
        seen.update(members)
        out.append((t, members, float(glen[idx])))
    return out


def warped_potential(case: Case, rng: np.random.Generator):
    """Stars of the warped interstitial potential (V Theta)(Q) for every Q with
    |Q| <= 2 Kmax (all differences G_J - G_I of the basis), written as lapw0
    does (the interstitial section read by rdswar.f).  V(r) is smooth: its
    first `nstars` stars carry seeded values V(G) = A u_s (G_1/|G|)^2 (V(0) =
    v0), and (V Theta)(Q) = sum_G' V(G') Theta(Q - G') is formed with the same
    analytic Theta as HAMILT so that the interstitial operator stays
    consistent with the overlap (no ghost states from the potential)."""
    has_inv = any((m == -np.eye(3, dtype=int)).all() for (m, _) in case.ops)
    gv, glv = gvectors(case, 12.0 * 2 * math.pi / min(case.alat))
    vstars = _stars(case, gv, glv)[: case.nstars]
    g1 = vstars[1][2] if len(vstars) > 1 else 1.0
    vg = {}
    for (t, members, gl) in vstars:
        if gl < 1e-12:
            val = complex(-0.30, 0.0)
        else:
            neg = tuple(-x for x in t)
            if neg in vg and neg not in members:
                val = vg[neg].conjugate()
            else:
                u = rng.uniform(-1.0, 1.0)
                ph = 0.0 if has_inv else rng.uniform(-math.pi, math.pi)
                val = case.potk_amp * u * (g1 / gl) ** 2 * complex(math.cos(ph), math.sin(ph))
        for mbr in members:
            vg[mbr] = val
    qmax = 2.0 * kmax(case) * 1.02 + 1e-9
    q, qlen = gvectors(case, qmax)
    gp = np.array(list(vg.keys()), int)
    vp = np.array(list(vg.values()), np.complex128)
    # Theta on the box of all differences Q - G', then the convolution by gathers
    lo = q.min(axis=0) - gp.max(axis=0)
    hi = q.max(axis=0) - gp.min(axis=0)
    dims = hi - lo + 1
    box = np.array(np.meshgrid(*(np.arange(lo[i], hi[i] + 1) for i in range(3)), indexing="ij")).reshape(3, -1).T
    th = step_function(case, box).reshape(tuple(dims))
    warp = np.zeros(len(q), np.complex128)
    for j in range(len(gp)):
        d = q - gp[j] - lo
        warp += vp[j] * th[d[:, 0], d[:, 1], d[:, 2]]
    if has_inv:
        warp = warp.real.astype(np.complex128)
    wmap = {tuple(int(x) for x in qq): w for qq, w in zip(q, warp)}
    out = []
    for (t, members, gl) in _stars(case, q, qlen):
        out.append((t, wmap[t]))
    return out


def write_vns(case: Case, path: str, rng: np.random.Generator) -> list:
    """Writes case.vns; returns per atom the V_LM(r) arrays as the reference
    reads them back (the written text parsed), in file order."""
    out = ["     TOTAL NONSPHERICAL POTENTIAL OF MT-SPHERES", "     NORM  OF VLM(R) = VLM(R)", ""]
    vlm_read = []
    for it, at in enumerate(case.atoms, start=1):
        r = radial_mesh(at.rmt)
        out.append(f"   ATOMNUMBER  {it:3d}")
        out.append(f"   NUMBER OF LM{len(at.vlm):3d}")
        out.append("")
        out.append("")
        vlm_read.append([])
        for (L, M) in at.vlm:
            amp = case.vns_amp * rng.uniform(0.5, 1.5) * (1 if rng.uniform() < 0.5 else -1)
            x = r / at.rmt
            v = amp * x ** L * np.exp(-2.0 * (x - 0.7) ** 2)
            out.append(f"   VLM(R) FOR L{L:3d}   M={M:2d}")
            out.append("")
            txt = [_fmt_e(val) for val in v]
            vlm_read[-1].append(np.array([float(t.replace("D", "E")) for t in txt]))
            for j in range(0, NPT, 4):
                out.append("   " + "".join(txt[j:j + 4]))
            out.append("")
            out.append("")
        out.extend([""] * 4)
    out.append("     TOTAL POTENTIAL IN INTERSTITIAL")
    out.append("")
    stars = warped_potential(case, rng)
    out.append(f" NUMBER OF PW{len(stars):6d}")
    for (t, v) in stars:
        out.append(f"   {t[0]:5d}{t[1]:5d}{t[2]:5d}{_fmt_e(v.real)}{_fmt_e(v.imag)}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return vlm_read


DEF_TEMPLATE = """ 4,'{c}.klist',          'unknown','formatted',0
 5,'{c}.in1',   'old',    'formatted',0
 6,'{c}.output1','unknown','formatted',0
10,'{c}.vector', 'unknown','unformatted',9000
11,'{c}.energy', 'unknown','formatted',0
18,'{c}.vsp',       'old',    'formatted',0
19,'{c}.vns',       'unknown','formatted',0
20,'{c}.struct',         'old',    'formatted',0
21,'{c}.scf1',   'unknown','formatted',0
55,'{c}.vec',            'unknown','formatted',0
71,'{c}.nsh','unknown','formatted',0
"""


def write_case(case: Case, d: str) -> None:
    os.makedirs(d, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(case.seed))
    c = "case"
    write_struct(case, os.path.join(d, f"{c}.struct"))
    write_in1(case, os.path.join(d, f"{c}.in1"))
    write_klist(case, os.path.join(d, f"{c}.klist"))
    write_vsp(case, os.path.join(d, f"{c}.vsp"), rng)
    if case.vns:
        vlm = write_vns(case, os.path.join(d, f"{c}.vns"), rng)
        # V_LM(r) as the reference reads it (NPT, max LM count, atom types),
        # for the radial-integral tests; the reference does not read this file
        nl = max([len(a) for a in vlm] + [1])
        arr = np.zeros((NPT, nl, len(vlm)))
        for it, a in enumerate(vlm):
            for k, v in enumerate(a):
                arr[:, k, it] = v
        np.save(os.path.join(d, "vlm_read.npy"), arr)
    else:
        open(os.path.join(d, f"{c}.vns"), "w").close()
    with open(os.path.join(d, "lapw1.def"), "w") as f:
        f.write(DEF_TEMPLATE.format(c=c))


# ----------------------------------------------------------------------------
# the cells
def _vlm_set(lnsmax: int, rng: np.random.Generator, count: int) -> List[Tuple[int, int]]:
    pool = [(L, M) for L in range(1, 2 * lnsmax + 1) for M in range(0, L + 1)]
    idx = rng.choice(len(pool), size=min(count, len(pool)), replace=False)
    return sorted(pool[i] for i in idx)


def case_c1_fcc(nk: int = 10) -> Case:
    """Monatomic fcc a=7.60 bohr, Z=13, RMT 2.50, RKmax 7, lmax 10, lnsmax 4,
    10 seeded k-points on a 1/8 grid, real (inversion symmetric); the window
    (-9, 4] Ry holds the lowest 20 to 23 eigenvalues of every k-point."""
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 0))
    ks = [(0, 0, 0, 8)]
    while len(ks) < nk:
        k = tuple(int(x) for x in rng.integers(0, 5, size=3))
        if (k + (8,)) not in ks:
            ks.append(k + (8,))
    at = AtomType("Al", 13.0, 2.50, [(0.0, 0.0, 0.0)], iatnr=-1,
                  vlm=[(2, 0), (4, 0), (4, 4), (6, 0), (6, 4)])
    return Case("c1_fcc", "fcc Al-like cell, seeded potential", "F", (7.60, 7.60, 7.60), [at],
                ops_inversion(), kpoints=ks, window=(-9.0, 4.0), seed=SEED_BASE + 0)


def case_lo_mult2(complex_: bool = False) -> Case:
    """Two atom types, one with MULT=2 (inversion partners), local orbitals
    (two on one l), APW+lo on one l; orthorhombic P cell."""
    seed = SEED_BASE + 11 + (1 if complex_ else 0)
    a1 = AtomType("Ti", 22.0, 2.0, [(0.0, 0.0, 0.0)], los=[(1, -3.0)], apw_lo=[0],
                  vlm=[(2, 0), (2, 2), (4, 0), (4, 2), (4, 4)])
    a2 = AtomType("O", 8.0, 1.6, [(0.25, 0.3, 0.4), (0.75, 0.7, 0.6)],
                  los=[(0, -2.0)], apw_lo=[0], vlm=[(1, 0), (2, 1), (3, 2), (4, 3)])
    ops = ops_inversion()
    if complex_:
        a2.positions = [(0.2, 0.3, 0.35), (0.7, 0.6, 0.7)]
        ops = ops_identity()
        a2.positions = [(0.25, 0.3, 0.4)]
        a3 = AtomType("N", 7.0, 1.5, [(0.62, 0.55, 0.7)], los=[(1, -0.4)], vlm=[(1, 1), (2, 0), (3, 3)])
        atoms = [a1, a2, a3]
    else:
        atoms = [a1, a2]
    return Case("lo_cplx" if complex_ else "lo_mult2", "LO/APW+lo/MULT test cell", "P", (6.1, 6.6, 7.2), atoms,
                ops, rkmax=5.5, lmax=8, kpoints=[(1, 2, 3, 8), (0, 0, 0, 1)] if complex_ else [(0, 0, 0, 1), (1, 0, 2, 6)],
                window=(-4.0, 4.0), seed=seed, complex=complex_)


def case_ins_lo(complex_: bool = False) -> Case:
    """The LO/APW+lo/MULT cell without a non-spherical potential file: INS /= 0,
    so HAMILT builds the interstitial part with USTPHX and no warped box."""
    c = case_lo_mult2(complex_)
    c.name = "ins_lo_cplx" if complex_ else "ins_lo"
    c.title = "LO/APW+lo/MULT test cell, no vns (INS /= 0)"
    c.vns = False
    return c


def rocksalt_positions(n: int) -> Tuple[list, list]:
    """Cation/anion fractional positions of an n x n x n supercell of the
    8-atom conventional rocksalt cell."""
    fcc = [(0, 0, 0), (0, .5, .5), (.5, 0, .5), (.5, .5, 0)]
    cat, an = [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for (x, y, z) in fcc:
                    cat.append(((x + i) / n, (y + j) / n, (z + k) / n))
                    an.append((((x + .5) % 1 + i) / n, (y + j) / n, (z + k) / n))
    return cat, an


def case_rocksalt(n: int, rkmax: float, name: str, seed: int) -> Case:
    """Rocksalt n^3 supercell, Z=12/8, RMT 2.0/1.6, every atom its own type
    (MULT=1), APW+lo for l=0,1 on every atom (4 local orbitals per atom);
    real (inversion)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cat, an = rocksalt_positions(n)
    atoms = []
    for p in cat:
        atoms.append(AtomType("Mg", 12.0, 2.0, [p], apw_lo=[0, 1], vlm=[(2, 0), (4, 0), (4, 4)]))
    for p in an:
        atoms.append(AtomType("O", 8.0, 1.6, [p], apw_lo=[0, 1], vlm=[(2, 0), (4, 0), (4, 4)]))
    a = 7.95 * n
    return Case(name, f"rocksalt {n}x{n}x{n} supercell ({8*n**3} atoms)", "P", (a, a, a), atoms,
                ops_inversion(), rkmax=rkmax, kpoints=[(0, 0, 0, 1)], window=(-9.0, 0.5), seed=seed)


def case_c4_ortho8(nk: int = 64) -> Case:
    """Batched-k cell: 8-atom orthorhombic P cell, two species (4 x Z=12,
    RMT 2.0 and 4 x Z=8, RMT 1.6) at seeded random positions (no inversion,
    so the complex build; nearest distances above the RMT sums), RKmax 7,
    lmax 10, lnsmax 4, no local orbitals, 64 seeded general k-points; basis
    sizes around 1900-2000."""
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 3))
    alat = (10.4, 11.2, 11.85)
    species = [("Mg", 12.0, 2.0)] * 4 + [("O", 8.0, 1.6)] * 4
    pos: List[Tuple[float, float, float]] = []
    for _, _, r in species:
        while True:
            f = tuple(float(x) for x in rng.uniform(0.0, 1.0, size=3))
            ok = True
            for (g, (_, _, r2)) in zip(pos, species):
                d = np.array(f) - np.array(g)
                d -= np.round(d)
                if np.linalg.norm(d * np.array(alat)) < r + r2 + 0.3:
                    ok = False
                    break
            if ok:
                pos.append(f)
                break
    vlms = [[(1, 0), (2, 0), (2, 2), (3, 1), (4, 0)], [(1, 1), (2, -1), (3, 0), (4, 2)]]
    atoms = [AtomType(nm, z, r, [p], vlm=vlms[i % 2]) for i, ((nm, z, r), p) in enumerate(zip(species, pos))]
    ks = []
    while len(ks) < nk:
        k = tuple(int(x) for x in rng.integers(0, 8, size=3)) + (8,)
        if k not in ks:
            ks.append(k)
    return Case("c4_ortho8", "8-atom orthorhombic cell, 2 species, seeded positions", "P", alat, atoms,
                ops_identity(), rkmax=7.0, kpoints=ks, window=(-9.0, 12.0), seed=SEED_BASE + 3, complex=True)


def _random_positions(rng: np.random.Generator, alat, radii, margin: float) -> List[Tuple[float, float, float]]:
    """Seeded fractional positions, one per radius, with every periodic
    distance at least the sum of the two radii plus `margin` (random
    sequential addition; the cells here pack to about a quarter)."""
    pos: List[Tuple[float, float, float]] = []
    A = np.asarray(alat, float)
    for r in radii:
        for _ in range(100000):
            f = rng.uniform(0.0, 1.0, size=3)
            if pos:
                d = np.asarray(pos) - f
                d -= np.round(d)
                dist = np.linalg.norm(d * A, axis=1)
                if np.any(dist < np.asarray(radii[:len(pos)]) + r + margin):
                    continue
            pos.append(tuple(float(x) for x in f))
            break
        else:
            raise RuntimeError("no room for another sphere")
    return pos


def case_c3_lowsym100(rkmax: float = 6.5, name: str = "c3_lowsym100") -> Case:
    """Split-precision cell: 100 atoms of 3 species (34 x Z=22 RMT 2.0,
    33 x Z=12 RMT 1.8, 33 x Z=8 RMT 1.6) at seeded random positions in an
    orthorhombic P cell with no symmetry but the identity (so the complex
    build), every atom its own type, APW+lo on l = 0, 1 (4 local orbitals per
    atom), RKmax 6.5, lmax 10, lnsmax 4, one general k-point; N about 12400."""
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 2))
    alat = (21.2, 22.0, 22.8)
    species = [("Ti", 22.0, 2.0)] * 34 + [("Mg", 12.0, 1.8)] * 33 + [("O", 8.0, 1.6)] * 33
    order = rng.permutation(len(species))
    species = [species[i] for i in order]
    pos = _random_positions(rng, alat, [s[2] for s in species], 0.05)
    vlms = [[(1, 0), (2, 0), (2, 2), (3, 1), (4, 0)], [(1, 1), (2, -1), (3, 0), (4, 2)],
            [(2, 1), (3, -2), (4, 4)]]
    atoms = [AtomType(nm, z, r, [p], apw_lo=[0, 1], vlm=vlms[i % 3])
             for i, ((nm, z, r), p) in enumerate(zip(species, pos))]
    return Case(name, "100-atom cell, 3 species, seeded positions, no symmetry", "P", alat, atoms,
                ops_identity(), rkmax=rkmax, kpoints=[(1, 3, 2, 8)], window=(-9.0, 12.0), seed=SEED_BASE + 2,
                complex=True)


def case_c5_cubic16(nk: int = 8) -> Case:
    """Adjoint cell: a CsCl-type 2x2x2 supercell (16 atoms, Z=12 RMT 2.0 and
    Z=8 RMT 1.6) in a cubic P cell of 12.8 bohr, every atom its own type,
    inversion (real build); RKmax 7 (N about 3000 with 64 local orbitals);
    V_LM with L <= 4; 200 seeded V(G) stars; 8 seeded k-points on a 1/4
    grid."""
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 5))
    atoms = []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                atoms.append(AtomType("Mg", 12.0, 2.0, [(i / 2, j / 2, k / 2)], apw_lo=[0, 1],
                                      vlm=[(2, 0), (2, 2), (4, 0), (4, 4)]))
    for i in range(2):
        for j in range(2):
            for k in range(2):
                atoms.append(AtomType("O", 8.0, 1.6, [((i + .5) / 2, (j + .5) / 2, (k + .5) / 2)], apw_lo=[0, 1],
                                      vlm=[(2, 0), (4, 0), (4, 2)]))
    ks = []
    while len(ks) < nk:
        q = tuple(int(x) for x in rng.integers(0, 4, size=3)) + (4,)
        if q not in ks:
            ks.append(q)
    a = 12.8
    return Case("c5_cubic16", "CsCl-type 2x2x2 supercell (16 atoms), adjoint cell", "P", (a, a, a), atoms,
                ops_inversion(), rkmax=7.0, kpoints=ks, window=(-9.0, 0.5), nstars=200, seed=SEED_BASE + 5)


def _with_window(c: Case, w) -> Case:
    c.window = w
    return c


CASES = {
    "c1_fcc": case_c1_fcc,
    "lo_mult2": lambda: case_lo_mult2(False),
    "lo_cplx": lambda: case_lo_mult2(True),
    "rs8": lambda: case_rocksalt(1, 5.0, "rs8", SEED_BASE + 21),
    "c2_rs64": lambda: _with_window(case_rocksalt(2, 7.0, "c2_rs64", SEED_BASE + 1), (-9.0, 12.0)),
    "c4_ortho8": case_c4_ortho8,
    "c3_lowsym100": case_c3_lowsym100,
    "c5_cubic16": case_c5_cubic16,
    "ins_lo": lambda: case_ins_lo(False),
    "ins_lo_cplx": lambda: case_ins_lo(True),
}
