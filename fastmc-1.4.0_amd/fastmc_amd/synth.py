"""Seeded synthetic workloads in the reference's input formats.

Frameworks are rigid, frozen lattices: an fcc-like lattice in the cell, a
random subset of its sites kept to reach the requested atom count, Gaussian
jitter, 4 element types with LJ epsilon ~ U[0.02, 0.3] kcal/mol and sigma ~
U[2.5, 3.8] A, partial charges ~ N(0, 0.4) e shifted to net zero.  Guests are
rigid CO2, N2 (massless centre charge), 5-site CH4 and 4-site water.  All
pairs use Lorentz-Berthelot mixing written as ``lj`` VDW lines.

Workloads:
  c1  cubic 25.8 A, 424 atoms, CO2, 298 K, 0.15 bar, kmax 8
  c2  2x2x2 of c1 (3392 atoms, 51.6 A), CO2/N2 0.15/0.85 bar, kmax 12
  c3  256 chains: 32 frameworks (300-6000 atoms, orthorhombic 20-60 A sides,
      replicated to >= 2 rcut) x 8 CO2 pressures 0.01-10 bar, kmax 8-14
  c4  serving requests: 300-10000 atoms, 1-3 of CO2/N2/CH4, 0.01-30 bar,
      273-348 K
  c5  3x3x3 of c1-like cells (~11.4k atoms, 77.4 A), CO2/CH4/H2O at 10 bar,
      kmax 16
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import inputs

KCAL = 418.4
RCUT = 12.5

# guest site data: name, mass, charge, LJ eps (kcal/mol), sigma (A)
GUEST_SITES = {
    "Cx": (12.0107, 0.0537, 2.80),   # CO2 carbon
    "Ox": (15.9994, 0.1570, 3.05),   # CO2 oxygen
    "Nn": (14.0067, 0.0715, 3.31),   # N2 nitrogen
    "Nc": (0.0, 0.0, 0.0),           # N2 massless centre
    "Cm": (12.0107, 0.0660, 3.50),   # CH4 carbon
    "Hm": (1.00794, 0.0300, 2.50),   # CH4 hydrogen
    "Ow": (15.9994, 0.1550, 3.1536),  # water oxygen
    "Hw": (1.00794, 0.0, 0.0),
    "Mw": (0.0, 0.0, 0.0),
    "Ar": (39.948, 0.2381, 3.405),   # argon, one uncharged LJ site
    "Ce": (12.0107, 0.0660, 3.50),   # ethane carbon
    "He": (1.00794, 0.0300, 2.50),   # ethane hydrogen
    "Co": (12.0107, 0.0526, 3.385),  # carbon monoxide carbon
    "Oo": (15.9994, 0.1610, 2.955),  # carbon monoxide oxygen
}


def guest_template(name: str):
    """Rigid guest geometry: list of (site, mass, charge, x, y, z)."""
    if name == "co2":
        d = 1.16
        return [("Cx", 12.0107, 0.70, 0.0, 0.0, 0.0),
                ("Ox", 15.9994, -0.35, 0.0, 0.0, d),
                ("Ox", 15.9994, -0.35, 0.0, 0.0, -d)]
    if name == "n2":
        h = 0.55
        return [("Nn", 14.0067, -0.482, 0.0, 0.0, h),
                ("Nc", 0.0, 0.964, 0.0, 0.0, 0.0),
                ("Nn", 14.0067, -0.482, 0.0, 0.0, -h)]
    if name == "ch4":
        b = 1.09 / math.sqrt(3.0)
        return [("Cm", 12.0107, -0.24, 0.0, 0.0, 0.0),
                ("Hm", 1.00794, 0.06, b, b, b),
                ("Hm", 1.00794, 0.06, -b, -b, b),
                ("Hm", 1.00794, 0.06, -b, b, -b),
                ("Hm", 1.00794, 0.06, b, -b, -b)]
    if name == "h2o":
        roh, ang = 0.9572, math.radians(104.52)
        hx = roh * math.sin(ang / 2)
        hz = roh * math.cos(ang / 2)
        return [("Ow", 15.9994, 0.0, 0.0, 0.0, 0.0),
                ("Hw", 1.00794, 0.52, hx, 0.0, hz),
                ("Hw", 1.00794, 0.52, -hx, 0.0, hz),
                ("Mw", 0.0, -1.04, 0.0, 0.0, 0.15)]
    if name == "ar":
        return [("Ar", 39.948, 0.0, 0.0, 0.0, 0.0)]
    if name == "co":
        # two-site carbon monoxide, C-O 1.128 A, point charges -/+0.2 e
        d = 1.128 / 2
        return [("Co", 12.0107, -0.2, 0.0, 0.0, d), ("Oo", 15.9994, 0.2, 0.0, 0.0, -d)]
    if name == "c2h6":
        # rigid staggered ethane, all 8 atoms: C-C 1.54 A, C-H 1.09 A,
        # tetrahedral angles, q_C = -0.18 e, q_H = +0.06 e
        cc, ch = 1.54, 1.09
        th = math.acos(-1.0 / 3.0)
        rho, dz = ch * math.sin(th), ch * math.cos(th)
        sites = [("Ce", 12.0107, -0.18, 0.0, 0.0, cc / 2), ("Ce", 12.0107, -0.18, 0.0, 0.0, -cc / 2)]
        for k in range(3):
            a = 2.0 * math.pi * k / 3.0
            sites.append(("He", 1.00794, 0.06, rho * math.cos(a), rho * math.sin(a), cc / 2 - dz))
        for k in range(3):
            a = 2.0 * math.pi * k / 3.0 + math.pi / 3.0
            sites.append(("He", 1.00794, 0.06, rho * math.cos(a), rho * math.sin(a), -cc / 2 + dz))
        return sites
    raise ValueError(name)


GUEST_TITLE = {"co2": "CO2 rigid 3-site", "n2": "N2 3-site with centre charge",
               "ch4": "CH4 rigid 5-site", "h2o": "H2O rigid 4-site",
               "ar": "Ar single LJ site", "c2h6": "C2H6 rigid 8-site",
               "co": "CO rigid 2-site"}


@dataclasses.dataclass
class Framework:
    cell: np.ndarray            # (9,) DL_POLY cell, orthorhombic here
    names: List[str]
    mass: List[float]
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    q: np.ndarray
    elements: List[Tuple[str, float, float, float]]  # name, mass, eps, sigma

    @property
    def n(self):
        return len(self.x)

    def replicate(self, na, nb, nc) -> "Framework":
        a, b, c = self.cell[0], self.cell[4], self.cell[8]
        xs, ys, zs, qs, ns, ms = [], [], [], [], [], []
        for i in range(na):
            for j in range(nb):
                for k in range(nc):
                    xs.append(self.x + i * a)
                    ys.append(self.y + j * b)
                    zs.append(self.z + k * c)
                    qs.append(self.q)
                    ns += self.names
                    ms += self.mass
        cell = np.array([na * a, 0, 0, 0, nb * b, 0, 0, 0, nc * c], float)
        return Framework(cell, ns, ms, np.concatenate(xs), np.concatenate(ys),
                         np.concatenate(zs), np.concatenate(qs), self.elements)


def make_framework(rng: np.random.Generator, sides: Sequence[float],
                   natoms: int, jitter: float = 0.3, prefix: str = "E") -> Framework:
    """Jittered fcc-like lattice with natoms sites in an orthorhombic cell."""
    sides = np.asarray(sides, float)
    vol = float(np.prod(sides))
    # fcc cells per side so that 4 * n_a n_b n_c >= natoms
    a0 = (4.0 * vol / natoms) ** (1.0 / 3.0)
    nc = np.maximum(1, np.round(sides / a0)).astype(int)
    while 4 * int(np.prod(nc)) < natoms:
        nc[np.argmin(sides / nc)] += 1
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = []
    for i in range(nc[0]):
        for j in range(nc[1]):
            for k in range(nc[2]):
                for bvec in basis:
                    pts.append(((i + bvec[0]) / nc[0], (j + bvec[1]) / nc[1],
                                (k + bvec[2]) / nc[2]))
    pts = np.array(pts)
    keep = np.sort(rng.choice(len(pts), size=natoms, replace=False))
    frac = pts[keep]
    xyz = frac * sides + rng.normal(0.0, jitter, size=(natoms, 3))
    xyz = np.mod(xyz, sides)
    elements = []
    masses = [65.38, 14.0067, 15.9994, 12.0107]
    for e in range(4):
        eps = float(rng.uniform(0.02, 0.3))
        sig = float(rng.uniform(2.5, 3.8))
        elements.append((f"{prefix}{e + 1}", masses[e], eps, sig))
    kind = rng.integers(0, 4, size=natoms)
    q = rng.normal(0.0, 0.4, size=natoms)
    q = q - q.mean()
    names = [elements[k][0] for k in kind]
    mass = [elements[k][1] for k in kind]
    cell = np.array([sides[0], 0, 0, 0, sides[1], 0, 0, 0, sides[2]], float)
    return Framework(cell, names, mass, xyz[:, 0].copy(), xyz[:, 1].copy(),
                     xyz[:, 2].copy(), q, elements)


def lb_pairs(elements, guests: Sequence[str], fw_names: Sequence[str] = ()):
    """Unique type list in the order readfield assigns it (first appearance
    over the FIELD molecules: guest sites, then framework atoms,
    readinputs.f:186-206) and Lorentz-Berthelot lj lines for every pair."""
    types = []
    params = {}
    for g in guests:
        for (n, m, q, x, y, z) in guest_template(g):
            if n not in params:
                types.append(n)
                params[n] = GUEST_SITES[n][1:]
    el = {e[0]: (e[2], e[3]) for e in elements}
    for n in list(fw_names) + [e[0] for e in elements]:
        if n not in params:
            types.append(n)
            params[n] = el[n]
    lines = []
    for i, a in enumerate(types):
        for b in types[i:]:
            ea, sa = params[a]
            eb, sb = params[b]
            eps = math.sqrt(ea * eb)
            sig = 0.5 * (sa + sb) if (sa > 0 and sb > 0) else 0.0
            lines.append((a, b, "lj", [eps, sig]))
    return types, lines


@dataclasses.dataclass
class Workload:
    """One chain's system + thermodynamic state."""
    name: str
    fw: Framework
    guests: List[str]
    press_bar: List[float]
    temp: float
    kmax: Tuple[int, int, int]
    alpha: float = 0.26
    rcut: float = RCUT
    delr: float = 1.0
    seed: Tuple[int, int, int, int] = (12, 34, 56, 78)

    def write(self, directory: str, steps: int = 10000, eqsteps: int = 10000):
        os.makedirs(directory, exist_ok=True)
        types, vdw = lb_pairs(self.fw.elements, self.guests, self.fw.names)
        gl = []
        for g in self.guests:
            gl.append(dict(name=GUEST_TITLE[g], sites=guest_template(g), nummols=0))
        # framework as one molecule type with every atom listed
        fw_sites = [(self.fw.names[i], self.fw.mass[i], float(self.fw.q[i]))
                    for i in range(self.fw.n)]
        inputs.write_field(os.path.join(directory, "FIELD"), self.name, gl,
                           [dict(name="Framework", sites=fw_sites, nummols=1)],
                           vdw)
        atoms = [(self.fw.names[i], self.fw.x[i], self.fw.y[i], self.fw.z[i])
                 for i in range(self.fw.n)]
        inputs.write_config(os.path.join(directory, "CONFIG"), self.name, 2,
                            self.fw.cell, atoms)
        inputs.write_control(os.path.join(directory, "CONTROL"), self.name,
                             self.temp, self.press_bar, rcut=self.rcut,
                             delr=self.delr,
                             ewald=(self.alpha, *self.kmax), steps=steps,
                             eqsteps=eqsteps)

    def system_arrays(self):
        """Arrays for the fmc_* / orc_* builder calls, in FIELD order."""
        types, vdw = lb_pairs(self.fw.elements, self.guests, self.fw.names)
        tindex = {t: i + 1 for i, t in enumerate(types)}
        guests = []
        for g in self.guests:
            tpl = guest_template(g)
            m = np.array([s[1] for s in tpl])
            x = np.array([s[3] for s in tpl]); y = np.array([s[4] for s in tpl])
            z = np.array([s[5] for s in tpl])
            if m.sum() > 0:
                x = x - (x * m).sum() / m.sum()
                y = y - (y * m).sum() / m.sum()
                z = z - (z * m).sum() / m.sum()
            guests.append(dict(x=x, y=y, z=z, q=np.array([s[2] for s in tpl]),
                               type=np.array([tindex[s[0]] for s in tpl]),
                               mass=m))
        fw = dict(x=self.fw.x, y=self.fw.y, z=self.fw.z, q=self.fw.q,
                  type=np.array([tindex[n] for n in self.fw.names]),
                  frozen=np.ones(self.fw.n, dtype=np.int32))
        vdwt = [(tindex[a], tindex[b], 2, [p[0] * KCAL, p[1], 0.0, 0.0, 0.0])
                for (a, b, kw, p) in vdw]
        return dict(imcon=2 if not _cubic(self.fw.cell) else 1,
                    cell=self.fw.cell, alpha=self.alpha, kmax=self.kmax,
                    rcut=self.rcut, delr=self.delr, ntpatm=len(types),
                    vdw=vdwt, fw=fw, guests=guests)


def _cubic(cell):
    return abs(cell[0] - cell[4]) < 1e-12 and abs(cell[0] - cell[8]) < 1e-12


def c1(seed: int = 1) -> Workload:
    rng = np.random.default_rng(seed)
    fw = make_framework(rng, (25.8, 25.8, 25.8), 424)
    return Workload("c1 plain cubic 25.8 A, 424 atoms, CO2", fw, ["co2"], [0.15],
                    298.0, (8, 8, 8), seed=(12, 34, 56, 78))


def c2(seed: int = 1) -> Workload:
    base = c1(seed).fw
    fw = base.replicate(2, 2, 2)
    return Workload("c2 2x2x2 supercell 51.6 A, 3392 atoms, CO2/N2", fw,
                    ["co2", "n2"], [0.15, 0.85], 298.0, (12, 12, 12),
                    seed=(12, 34, 56, 78))


def _kmax_for(side: float) -> int:
    return int(min(14, max(8, round(side * 8.0 / 25.8))))


def c3(seed: int = 3, n_frameworks: int = 32, n_press: int = 8) -> List[Workload]:
    rng = np.random.default_rng(seed)
    out = []
    press = np.logspace(-2, 1, n_press)
    for f in range(n_frameworks):
        n = int(round(math.exp(rng.uniform(math.log(300), math.log(6000)))))
        side0 = (n / 0.0247) ** (1.0 / 3.0)
        sides = np.clip(side0 * rng.uniform(0.8, 1.25, size=3), 20.0, 60.0)
        fw = make_framework(rng, sides, n, prefix=f"F{f % 10}")
        reps = [max(1, math.ceil(2 * RCUT / s)) for s in sides]
        if reps != [1, 1, 1]:
            fw = fw.replicate(*reps)
        km = tuple(_kmax_for(fw.cell[4 * d]) for d in range(3))
        for p in range(n_press):
            s = (int(rng.integers(1, 178)), int(rng.integers(1, 178)),
                 int(rng.integers(1, 178)), int(rng.integers(0, 169)))
            out.append(Workload(f"c3 fw{f} P={press[p]:.4g} bar", fw, ["co2"],
                                [float(press[p])], 298.0, km, seed=s))
    return out


def c4_request(rng: np.random.Generator, idx: int) -> Workload:
    n = int(round(math.exp(rng.uniform(math.log(300), math.log(10000)))))
    side0 = (n / 0.0247) ** (1.0 / 3.0)
    sides = np.clip(side0 * rng.uniform(0.85, 1.18, size=3), 15.0, 80.0)
    fw = make_framework(rng, sides, n, prefix="R")
    reps = [max(1, math.ceil(2 * RCUT / s)) for s in sides]
    if reps != [1, 1, 1]:
        fw = fw.replicate(*reps)
    ng = int(rng.integers(1, 4))
    species = list(rng.choice(["co2", "n2", "ch4"], size=ng, replace=False))
    ptot = float(math.exp(rng.uniform(math.log(0.01), math.log(30.0))))
    frac = rng.dirichlet(np.ones(ng))
    km = tuple(_kmax_for(fw.cell[4 * d]) for d in range(3))
    s = (int(rng.integers(1, 178)), int(rng.integers(1, 178)),
         int(rng.integers(1, 178)), int(rng.integers(0, 169)))
    return Workload(f"c4 request {idx}", fw, species,
                    [float(ptot * f) for f in frac],
                    float(rng.uniform(273.0, 348.0)), km, seed=s)


def c5(seed: int = 5) -> Workload:
    rng = np.random.default_rng(seed)
    fw = make_framework(rng, (25.8, 25.8, 25.8), 424).replicate(3, 3, 3)
    return Workload("c5 3x3x3 supercell 77.4 A, CO2/CH4/H2O", fw,
                    ["co2", "ch4", "h2o"], [10.0 / 3] * 3, 298.0, (16, 16, 16),
                    seed=(12, 34, 56, 78))
