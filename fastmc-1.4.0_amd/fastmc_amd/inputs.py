"""Readers and writers for the reference's CONTROL / FIELD / CONFIG files.

The subset understood here is the one the trial-move path needs; keyword
matching and number extraction follow the reference's parser:

* records are matched by lower-case substring (``findstring``,
  parse_module.f:449-491) in the order of the reference's if-chains
  (readinputs.f:53-290 for FIELD, :594-842 for CONTROL, utility_pack.f:
  310-414 for the CONTROL pre-scan);
* numbers are pulled left to right with the arithmetic of ``dblstr`` /
  ``intstr`` (parse_module.f:100-288), which builds decimals digit by digit,
  so parsed values equal the reference's to the last bit;
* CONFIG coordinates are read as three 20-character fields
  (readinputs.f:374-376).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Tuple

ENGUNIT = {"kcal": 418.4, "ev": 9648.530821, "kj": 1.0e2, "internal": 1.0}
BOLTZ = 1.3806503e-23

VDW_KEYS = {"12-6": 1, "lj": 2, "nm": 3, "buck": 4, "bhm": 5, "hbnd": 6,
            "snm": 7, "hcnm": 7, "mors": 8, "wca": 9}


class Record:
    """A record consumed left to right like the reference's work string."""

    def __init__(self, text: str):
        # records are fixed-length and blank padded in the reference
        self.s = text + " " * 16

    def dblstr(self) -> float:
        # parse_module.f:171-288
        s = self.s
        n = len(s)
        val, sn, ksn, ten, one = 0.0, 1.0, "+", 10.0, 1.0
        iexp, start, ldot, final = 0, False, False, False
        lst = 0
        consumed_extra = 0
        while lst < n and not final:
            ch = s[lst]
            lst += 1
            flag = False
            if "0" <= ch <= "9":
                val = ten * val + one * float(ord(ch) - 48)
                flag = True
                start = True
            if ch == ".":
                flag = True
                ten = 1.0
                ldot = True
                start = True
            if flag and ksn == "-":
                sn = -1.0
            if ldot:
                one = one / 10.0
            ksn = ch
            if ksn == "D":
                ksn = "d"
            if ksn == "E":
                ksn = "e"
            if start:
                if ksn in ("d", "e"):
                    rest = Record(s[lst:])
                    iexp, idum = rest._intstr_pos()
                    consumed_extra = idum
                    final = True
                if not flag:
                    final = True
        val = sn * val * (10.0 ** iexp)
        lst = lst + consumed_extra
        self.s = s[max(lst - 1, 0):] if lst <= n else ""
        return val

    def _intstr_pos(self) -> Tuple[int, int]:
        # parse_module.f:100-169; returns (value, chars consumed)
        s = self.s
        n = len(s)
        isn, lst, ksn, val = 1, 0, "+", 0
        count, final = False, False
        while lst < n and not final:
            ch = s[lst]
            lst += 1
            flag = False
            if "0" <= ch <= "9":
                val = 10 * val + (ord(ch) - 48)
                count = True
                flag = True
            if count and not flag:
                final = True
            if flag and ksn == "-":
                isn = -1
            ksn = ch
        return isn * val, lst

    def intstr(self) -> int:
        v, lst = self._intstr_pos()
        self.s = self.s[max(lst - 1, 0):] if lst <= len(self.s) else ""
        return v

    def getword(self, width: int = 8) -> str:
        # parse_module.f getword: next blank-delimited word
        s = self.s.lstrip(" \t")
        parts = s.split(None, 1)
        if not parts:
            self.s = ""
            return ""
        w = parts[0]
        self.s = parts[1] if len(parts) > 1 else ""
        return w[:width]


def _clean(line: str) -> str:
    return line.rstrip("\n").replace("\t", " ")


# ---------------------------------------------------------------------------
# FIELD
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Molecule:
    name: str
    guest: bool
    nummols: int
    names: List[str]
    mass: List[float]
    charge: List[float]
    frozen: List[int]
    types: List[int]               # 1-based unique atom type (ltpsit)
    gx: List[float] = dataclasses.field(default_factory=list)  # COM-centred
    gy: List[float] = dataclasses.field(default_factory=list)
    gz: List[float] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class Field:
    title: str
    engunit: float
    molecules: List[Molecule]
    unqatm: List[str]
    vdw: List[Tuple[int, int, int, List[float]]]   # (a, b, key, prm*engunit)

    @property
    def guests(self) -> List[Molecule]:
        return [m for m in self.molecules if m.guest]

    @property
    def frameworks(self) -> List[Molecule]:
        return [m for m in self.molecules if not m.guest]


def read_field(path: str) -> Field:
    """readfield (readinputs.f:11-307) + define_van_der_waals
    (vdw_terms.f:1-187) for the subset on the path."""
    with open(path) as f:
        lines = [_clean(l) for l in f]
    it = iter(lines)
    title = next(it).strip()
    engunit = 1.0
    molecules: List[Molecule] = []
    unqatm: List[str] = []
    vdw = []
    for raw in it:
        rec = raw.lower().strip()
        if rec == "" or rec.startswith("#"):
            continue
        if "units" in rec:
            rest = rec[5:].strip()
            if not rest:
                engunit = 1.0
            elif "ev" in rec:
                engunit = ENGUNIT["ev"]
            elif "kcal" in rec:
                engunit = ENGUNIT["kcal"]
            elif "kj" in rec:
                engunit = ENGUNIT["kj"]
            elif "k" in rec:
                engunit = BOLTZ
        elif "vdw" in rec:
            ntpvdw = Record(rec).intstr()
            for _ in range(ntpvdw):
                r = Record(next(it))
                a1 = r.getword()
                a2 = r.getword()
                r.s = r.s.lower()
                kw = r.getword(4)
                key = None
                for k, v in VDW_KEYS.items():
                    if kw.startswith(k[:4]) and (len(kw) >= len(k[:4])):
                        key = v
                        break
                if kw == "tab":
                    key = 0
                prm = [r.dblstr() for _ in range(5)]
                if key is None:
                    raise ValueError(f"unknown vdw potential {kw!r}")
                ka = unqatm.index(a1) + 1 if a1 in unqatm else 0
                kb = unqatm.index(a2) + 1 if a2 in unqatm else 0
                prm[0] *= engunit
                if key == 1:
                    prm[1] *= engunit
                elif key == 4:
                    prm[2] *= engunit
                elif key == 5:
                    prm[3] *= engunit
                    prm[4] *= engunit
                elif key == 6:
                    prm[1] *= engunit
                vdw.append((ka, kb, key, prm))
        elif "molecu" in rec:
            ntpmls = Record(rec).intstr()
            for _ in range(ntpmls):
                nameline = next(it)
                guest = "&guest" in nameline.lower()
                mol = Molecule(name=nameline.strip(), guest=guest, nummols=0,
                               names=[], mass=[], charge=[], frozen=[],
                               types=[])
                while True:
                    r2 = next(it).lower().strip()
                    if "nummol" in r2:
                        mol.nummols = Record(r2).intstr()
                    elif "atoms" in r2:
                        numsit = Record(r2).intstr()
                        ksite = 0
                        while ksite < numsit:
                            r = Record(next(it))
                            atom = r.getword()
                            w = r.dblstr()
                            q = r.dblstr()
                            if guest:
                                nrept = 1
                                mol.gx.append(r.dblstr())
                                mol.gy.append(r.dblstr())
                                mol.gz.append(r.dblstr())
                                ifrz = 0
                            else:
                                nrept = r.intstr()
                                ifrz = r.intstr()
                            if nrept == 0:
                                nrept = 1
                            if atom not in unqatm:
                                unqatm.append(atom)
                            t = unqatm.index(atom) + 1
                            for _r in range(nrept):
                                mol.names.append(atom)
                                mol.mass.append(w)
                                mol.charge.append(q)
                                mol.frozen.append(ifrz)
                                mol.types.append(t)
                                ksite += 1
                    elif "finish" in r2:
                        break
                if guest:
                    # readinputs.f:226-253: centre template on its COM
                    m = cx = cy = cz = 0.0
                    for i in range(len(mol.names)):
                        cx += mol.gx[i] * mol.mass[i]
                        cy += mol.gy[i] * mol.mass[i]
                        cz += mol.gz[i] * mol.mass[i]
                        m += mol.mass[i]
                    cx, cy, cz = cx / m, cy / m, cz / m
                    mol.gx = [v - cx for v in mol.gx]
                    mol.gy = [v - cy for v in mol.gy]
                    mol.gz = [v - cz for v in mol.gz]
                molecules.append(mol)
        elif "close" in rec:
            break
    return Field(title, engunit, molecules, unqatm, vdw)


# ---------------------------------------------------------------------------
# CONFIG
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Config:
    title: str
    levcfg: int
    imcon: int
    cell: List[float]
    coords: Dict[int, List[Tuple[float, float, float]]]   # molecule index


def _f20(line: str) -> Tuple[float, float, float]:
    vals = []
    for k in range(3):
        fld = line[20 * k:20 * (k + 1)].strip().replace("d", "e").replace("D", "e")
        vals.append(float(fld) if fld else 0.0)
    return vals[0], vals[1], vals[2]


def read_config(path: str, field: Field) -> Config:
    """readconfig (readinputs.f:309-513) for levcfg 0-2."""
    with open(path) as f:
        lines = [_clean(l) for l in f]
    title = lines[0].strip()
    r = Record(lines[1])
    levcfg = r.intstr()
    imcon = r.intstr()
    pos = 2
    cell = [0.0] * 9
    if imcon > 0:
        for row in range(3):
            rr = Record(lines[pos])
            for c in range(3):
                cell[3 * row + c] = rr.dblstr()
            pos += 1
    coords: Dict[int, List[Tuple[float, float, float]]] = {}
    for k, mol in enumerate(field.molecules):
        lst = []
        for _m in range(mol.nummols):
            for s in range(len(mol.names)):
                name = lines[pos].strip()[:8].strip()
                if name != mol.names[s]:
                    raise ValueError(
                        f"unidentified atom label {name!r} (expected "
                        f"{mol.names[s]!r}) at CONFIG line {pos + 1}")
                lst.append(_f20(lines[pos + 1]))
                pos += 2 + levcfg
        coords[k] = lst
    return Config(title, levcfg, imcon, cell, coords)


# ---------------------------------------------------------------------------
# CONTROL
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class GuestControl:
    press_pa: float = 0.0
    freq: List[float] = dataclasses.field(default_factory=lambda: [0.0] * 9)
    freq_set: bool = False
    delr: Optional[float] = None
    disp_ratio: Optional[float] = None
    tran_ratio: Optional[float] = None


@dataclasses.dataclass
class Control:
    title: str = ""
    temp: float = 0.0
    steps: int = 1
    eqsteps: int = 0
    rcut: float = 0.0
    delr: float = 1.0
    rvdw: float = 0.0
    ewald_precision: Optional[float] = None
    ewald_sum: Optional[Tuple[float, int, int, int]] = None
    numguests: int = 1000
    nhis: int = 0
    nwind: int = 100000
    maxguest: int = 3000
    maxatm: int = 15000
    guests: Dict[int, GuestControl] = dataclasses.field(default_factory=dict)
    gfreq: List[Optional[float]] = dataclasses.field(default_factory=lambda: [None] * 9)
    disp_ratio: float = 0.5
    tran_ratio: float = 0.5
    rota_ratio: float = 0.5
    restart: bool = False


_MOVE_KEYS = ["ins", "del", "dis", "jum", "fle", "swa", "tra", "rot", "mul"]


def read_control(path: str) -> Control:
    """initscan (utility_pack.f:310-414) + readcontrol (readinputs.f:515-914)
    for the subset on the path."""
    with open(path) as f:
        lines = [_clean(l) for l in f]
    c = Control(title=lines[0].strip())
    it = iter(lines[1:])
    for raw in it:
        rec = raw.lower().strip()
        if rec == "" or rec.startswith("#"):
            continue
        if "restart" in rec:
            c.restart = True
        elif "temp" in rec:
            c.temp = Record(rec).dblstr()
        elif "&guest" in rec:
            idg = Record(rec).intstr()
            g = c.guests.setdefault(idg, GuestControl())
            for raw2 in it:
                r2 = raw2.lower().strip()
                if r2 == "" or r2.startswith("#"):
                    continue
                if "fract" in r2:
                    pass
                elif "press" in r2:
                    g.press_pa = Record(r2).dblstr() * 1.0e5
                elif "move" in r2:
                    g.freq_set = True
                    for i, k in enumerate(_MOVE_KEYS):
                        if k in r2:
                            g.freq[i] = Record(r2).dblstr()
                            break
                elif "delr" in r2:
                    g.delr = Record(r2).dblstr()
                elif "accep" in r2:
                    if "dis" in r2:
                        g.disp_ratio = Record(r2).dblstr()
                    elif "tra" in r2:
                        g.tran_ratio = Record(r2).dblstr()
                elif "&end" in r2:
                    break
        elif "steps" in rec:
            c.steps = Record(rec).intstr()
        elif "numg" in rec:
            c.numguests = Record(rec).intstr()
        elif "hist" in rec:
            c.nhis = Record(rec).intstr()
        elif "equilibr" in rec:
            c.eqsteps = Record(rec).intstr()
        elif "cutoff" in rec or "cut" in rec:
            c.rcut = Record(rec).dblstr()
        elif "delr" in rec:
            c.delr = Record(rec).dblstr()
        elif "ewald" in rec:
            if "precision" in rec:
                c.ewald_precision = Record(rec).dblstr()
            else:
                r = Record(rec)
                c.ewald_sum = (r.dblstr(), r.intstr(), r.intstr(), r.intstr())
        elif "rvdw" in rec:
            c.rvdw = Record(rec).dblstr()
        elif "averaging window" in rec:
            c.nwind = Record(rec).intstr()
        elif "max guest atoms" in rec:
            c.maxguest = Record(rec).intstr()
        elif "max framework atoms" in rec:
            c.maxatm = Record(rec).intstr()
        elif "move" in rec:
            for i, k in enumerate(_MOVE_KEYS):
                if k in rec:
                    c.gfreq[i] = Record(rec).dblstr()
                    break
        elif "accep" in rec:
            if "dis" in rec:
                c.disp_ratio = Record(rec).dblstr()
            elif "tra" in rec:
                c.tran_ratio = Record(rec).dblstr()
            elif "rot" in rec:
                c.rota_ratio = Record(rec).dblstr()
        elif "finish" in rec:
            break
    # guests without their own move lines take the general ones
    for g in c.guests.values():
        if not g.freq_set:
            g.freq = [v if v is not None else 0.0 for v in c.gfreq]
        if g.delr is None:
            g.delr = c.delr
        if g.disp_ratio is None:
            g.disp_ratio = c.disp_ratio
        if g.tran_ratio is None:
            g.tran_ratio = c.tran_ratio
    return c


# ---------------------------------------------------------------------------
# writers (used by the synthetic-workload generator)
# ---------------------------------------------------------------------------

def write_field(path: str, title: str, guests, frameworks, vdw_lines,
                units: str = "kcal") -> None:
    """guests: list of dict(name, sites=[(name, mass, q, x, y, z)], nummols);
    frameworks: list of dict(name, sites=[(name, mass, q)], nummols)."""
    out = [title, f"UNITS   {units}", f"molecular types {len(guests) + len(frameworks)}"]
    for g in guests:
        out.append(f"&guest {g['name']}")
        out.append(f"NUMMOLS {g.get('nummols', 0)}")
        out.append(f"ATOMS {len(g['sites'])}")
        for (n, m, q, x, y, z) in g["sites"]:
            out.append(f"{n:<8s} {m:14.6f} {q:12.6f} {x:12.6f} {y:12.6f} {z:12.6f}")
        out.append("rigid 1")
        out.append("finish")
    for fw in frameworks:
        out.append(fw["name"])
        out.append(f"NUMMOLS {fw.get('nummols', 1)}")
        out.append(f"ATOMS {len(fw['sites'])}")
        for (n, m, q) in fw["sites"]:
            out.append(f"{n:<8s} {m:14.6f} {q:18.14f}      1      1")
        out.append("finish")
    out.append(f"VDW {len(vdw_lines)}")
    for (a, b, kw, p) in vdw_lines:
        out.append(f"{a:<6s} {b:<6s} {kw} " + " ".join(f"{v:.6f}" for v in p))
    out.append("close")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_config(path: str, title: str, imcon: int, cell, atoms) -> None:
    """atoms: list of (name, x, y, z) in FIELD molecule order."""
    out = [title, f"{0:10d}{imcon:10d}{len(atoms):10d}"]
    for row in range(3):
        out.append("".join(f"{cell[3 * row + c]:20.12f}" for c in range(3)))
    for i, (n, x, y, z) in enumerate(atoms):
        out.append(f"{n:<8s}{i + 1:10d}")
        out.append(f"{x:20.12f}{y:20.12f}{z:20.12f}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_control(path: str, title: str, temp: float, guests_bar,
                  rcut: float = 12.5, delr: float = 1.0, ewald=None,
                  precision: Optional[float] = None, steps: int = 10000,
                  eqsteps: int = 10000, moves=(1.0, 1.0, 1.0),
                  extra: Optional[List[str]] = None) -> None:
    out = [title, f"temperature  {temp:.6f}", "",
           f"steps    {steps}", f"equilibration    {eqsteps}", "",
           f"cutoff          {rcut:.6f} angstrom", f"delr            {delr} angstrom"]
    if ewald is not None:
        a, k1, k2, k3 = ewald
        out.append(f"ewald sum {a} {k1} {k2} {k3}")
    else:
        out.append(f"ewald precision  {precision or 1e-6:.0e}".replace("e-0", "d-"))
    out.append("")
    for i, pbar in enumerate(guests_bar):
        out += [f"&guest {i + 1}", f"  pressure (bar) {pbar:.6f}", "&end"]
    out.append("")
    out += [f"move insert {moves[0]}", f"move delete {moves[1]}",
            f"move displace {moves[2]}"]
    if extra:
        out += extra
    out.append("finish")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


# ---------------------------------------------------------------------------
# REVIVE / REVCON (restart files, utility_pack.f:1772-1862)
# ---------------------------------------------------------------------------

def read_fortran_records(path: str) -> List[bytes]:
    """Records of a Fortran sequential unformatted file (4-byte little-endian
    length markers before and after each record)."""
    import struct
    raw = open(path, "rb").read()
    recs, p = [], 0
    while p < len(raw):
        if p + 4 > len(raw):
            raise ValueError(f"{path}: truncated record marker at byte {p}")
        n = struct.unpack_from("<i", raw, p)[0]
        if n < 0 or p + 8 + n > len(raw):
            raise ValueError(f"{path}: bad record length {n} at byte {p}")
        tail = struct.unpack_from("<i", raw, p + 4 + n)[0]
        if tail != n:
            raise ValueError(f"{path}: record markers {n} and {tail} differ at byte {p}")
        recs.append(raw[p + 4:p + 4 + n])
        p += 8 + n
    return recs


def write_fortran_records(path: str, recs: List[bytes]) -> None:
    import struct
    with open(path, "wb") as f:
        for r in recs:
            m = struct.pack("<i", len(r))
            f.write(m + r + m)


@dataclasses.dataclass
class Revive:
    """The restart state the reference writes every 1000 moves (``revive``,
    utility_pack.f:1800-1814) and reads back (``revread``, :1835-1862)."""
    production: int                 # the LOGICAL as stored (non-zero = true)
    nummols: List[int]              # per FIELD molecule type
    chainstats: List[float]         # 1 + 14 per guest type (gcmc.f:216-230)
    natms: int                      # guest + framework atoms
    coords: List[Tuple["object", "object", "object"]]   # per molecule type,
                                    # sites molecule-major (framwkxxx(mol, :))
    energy: List[float]             # per guest type (kcal/mol)

    @property
    def is_production(self) -> bool:
        return self.production != 0


def read_revive(path: str, field: Field) -> Revive:
    """revread (utility_pack.f:1835-1862): the record sizes are checked
    against the FIELD's molecule types."""
    import numpy as np
    recs = read_fortran_records(path)
    nt = len(field.molecules)
    ng = len(field.guests)
    if len(recs) != 5 + 3 * nt:
        raise ValueError(f"{path}: {len(recs)} records, expected {5 + 3 * nt} "
                         f"for {nt} molecule types")
    production = int(np.frombuffer(recs[0], "<i4")[0])
    nummols = [int(v) for v in np.frombuffer(recs[1], "<i4")]
    chainstats = [float(v) for v in np.frombuffer(recs[2], "<f8")]
    natms = int(np.frombuffer(recs[3], "<i4")[0])
    if len(nummols) != nt or len(chainstats) != 14 * ng + 1:
        raise ValueError(f"{path}: {len(nummols)} molecule counts and "
                         f"{len(chainstats)} statistics do not fit the FIELD")
    coords = []
    total = 0
    for k, mol in enumerate(field.molecules):
        n = nummols[k] * len(mol.names)
        xyz = tuple(np.frombuffer(recs[4 + 3 * k + c], "<f8").copy() for c in range(3))
        if any(len(a) != n for a in xyz):
            raise ValueError(f"{path}: molecule type {k + 1} holds "
                             f"{len(xyz[0])} sites, expected {n}")
        coords.append(xyz)
        total += n
    if total != natms:
        raise ValueError(f"{path}: {total} sites listed, natms = {natms}")
    energy = [float(v) for v in np.frombuffer(recs[-1], "<f8")]
    if len(energy) != ng:
        raise ValueError(f"{path}: {len(energy)} energies for {ng} guest types")
    return Revive(production, nummols, chainstats, natms, coords, energy)


def write_revive(path: str, rev: Revive) -> None:
    """revive (utility_pack.f:1800-1814)."""
    import numpy as np
    recs = [np.array([rev.production], "<i4").tobytes(),
            np.array(rev.nummols, "<i4").tobytes(),
            np.array(rev.chainstats, "<f8").tobytes(),
            np.array([rev.natms], "<i4").tobytes()]
    for xyz in rev.coords:
        recs += [np.asarray(a, "<f8").tobytes() for a in xyz]
    recs.append(np.array(rev.energy, "<f8").tobytes())
    write_fortran_records(path, recs)


def apply_revive(field: Field, config: Config, rev: Revive) -> Tuple[Field, Config]:
    """Copies of field and config holding the REVIVE's molecule counts and
    coordinates, as revread leaves the module arrays for the restart
    (gcmc.f:361-386)."""
    import copy
    f = copy.deepcopy(field)
    c = copy.deepcopy(config)
    for k, mol in enumerate(f.molecules):
        mol.nummols = rev.nummols[k]
        x, y, z = rev.coords[k]
        c.coords[k] = [(float(a), float(b), float(d)) for a, b, d in zip(x, y, z)]
    return f, c


def fortran_g(x: float, w: int = 20, d: int = 10) -> str:
    """Fortran Gw.d output editing: F(w-4).(d-e) and four blanks when the
    value rounded to d digits has decimal exponent 0 <= e <= d, else Ew.d."""
    if x == 0.0:
        return f"{0.0:.{d - 1}f}".rjust(w - 4) + "    "
    m = f"{abs(x):.{d - 1}e}"
    e = int(m.split("e")[1]) + 1
    if 0 <= e <= d:
        f = f"{x:.{d - e}f}" + ("." if e == d else "")
        return f.rjust(w - 4) + "    "
    return fortran_e(x, w, d)


def fortran_e(x: float, w: int = 20, d: int = 10) -> str:
    """Fortran Ew.d output editing: 0.ddd...E+ee, or 0.ddd...+eee when the
    exponent needs three digits."""
    if x == 0.0:
        return ("0." + "0" * d + "E+00").rjust(w)
    m, ex = f"{abs(x):.{d - 1}e}".split("e")
    e = int(ex) + 1
    tail = f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"
    s = ("-" if x < 0 else "") + "0." + m.replace(".", "") + tail
    return s.rjust(w)


def revcon_atoms(field: Field, coords) -> List[Tuple[str, float, float, float]]:
    """(name, x, y, z) in the order ``condense`` builds the atom arrays
    (utility_pack.f:3337-3380): guest types first, then the frameworks, each
    in FIELD order.  coords maps the FIELD molecule index to its sites
    (Config.coords, or the (x, y, z) arrays of a Revive)."""
    out = []
    order = ([k for k, m in enumerate(field.molecules) if m.guest] +
             [k for k, m in enumerate(field.molecules) if not m.guest])
    for k in order:
        mol = field.molecules[k]
        c = coords[k]
        if isinstance(c, tuple) and len(c) == 3 and not isinstance(c[0], tuple):
            c = list(zip(*c))
        ns = len(mol.names)
        for i, (x, y, z) in enumerate(c):
            out.append((mol.names[i % ns], float(x), float(y), float(z)))
    return out


def write_revcon(path: str, title: str, imcon: int, cell, atoms,
                 eng: float = 0.0, levcfg: int = 0) -> None:
    """revive's REVCON (utility_pack.f:1788-1798): a CONFIG-format file,
    '(80a1)' title, '(2i10,e20.10)' levcfg imcon energy, '(3f20.12)' cell,
    then per atom '(a8,i10)' name index and '(3g20.10)' x y z.  The file can
    be read back as a CONFIG."""
    out = [title[:80].ljust(80), f"{levcfg:10d}{imcon:10d}" + fortran_e(eng)]
    if imcon > 0:
        for row in range(3):
            out.append("".join(f"{cell[3 * row + c]:20.12f}" for c in range(3)))
    for i, (n, x, y, z) in enumerate(atoms):
        out.append(f"{n[:8]:<8s}{i + 1:10d}")
        out.append(fortran_g(x) + fortran_g(y) + fortran_g(z))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
