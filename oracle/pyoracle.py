"""ctypes wrapper of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker.  It loads
``oracle/_build/liboracle.so`` (``make -C oracle``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
ABI_LIB_PATH = os.path.join(HERE, "_build", "libfastmc_oracle_abi.so")

_lib = None

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(DP)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(IP)


class MoveRec(C.Structure):
    _fields_ = [("kind", C.c_int), ("iguest", C.c_int), ("accepted", C.c_int),
                ("nmol_after", C.c_int), ("estep", C.c_double),
                ("e1", C.c_double), ("e2", C.c_double), ("vdw", C.c_double),
                ("e3", C.c_double), ("delrc", C.c_double)]


class ChainParams(C.Structure):
    _fields_ = [("temp", C.c_double), ("ntpguest", C.c_int),
                ("press", C.c_double * 8), ("freq", (C.c_double * 9) * 8),
                ("delr", C.c_double * 8), ("disp_ratio", C.c_double * 8),
                ("tran_ratio", C.c_double * 8), ("rota_ratio", C.c_double),
                ("seed", C.c_int * 4), ("engunit", C.c_double)]


class DuniState(C.Structure):
    _fields_ = [("u", C.c_float * 97), ("c", C.c_float), ("cd", C.c_float),
                ("cm", C.c_float), ("ir", C.c_int), ("jr", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        d, i, v, p = C.c_double, C.c_int, None, C.c_void_p
        sig = {
            "orc_dcell": (v, [DP, DP]),
            "orc_invert": (v, [DP, DP, DP]),
            "orc_images": (v, [i, i, DP, DP, DP, DP]),
            "orc_volume": (d, [i, DP]),
            "orc_ewald_precision": (v, [d, d, i, DP, DP, IP]),
            "orc_mxgrid": (i, [d]),
            "orc_erfcgen": (d, [d, d, i, DP]),
            "orc_forgen": (d, [i, IP, DP, d, i, DP]),
            "orc_kvec_count": (i, [DP, IP]),
            "orc_ewald1": (d, [i, DP, d, IP, d, i, DP, DP, DP, DP, DP, DP, IP, DP]),
            "orc_ewald1_guest": (d, [i, DP, d, IP, d, i, DP, DP, DP, DP, i, i,
                                     DP, DP, DP, DP]),
            "orc_ewald2": (d, [d, i, IP, DP, DP, DP, d, d, d]),
            "orc_srfrce": (d, [i, i, IP, DP, IP, IP, IP, DP, i, d, d]),
            "orc_ewald3": (d, [d, i, DP, DP, DP, DP, d, d]),
            "orc_ewald3_guest": (d, [i, DP, i, DP, DP, DP, DP, d, d]),
            "orc_lrcorrect": (d, [i, i, IP, IP, IP, IP, DP, d, d, DP]),
            "orc_gstlrcorrect": (d, [i, i, IP, IP, DP, d]),
            "orc_guestlistgen": (v, [i, DP, d, d, i, DP, DP, DP, IP, i, DP, DP,
                                     DP, IP, IP]),
            "orc_duni_init": (v, [p, i, i, i, i]),
            "orc_duni_next": (d, [p]),
            "orc_duni_seed_from_clock": (v, [i, i, IP]),
            "orc_energy_eval": (i, [d, d, d, d, i, d, d, i]),
            "orc_rotation": (v, [DP, DP, DP, i, d, d, d, d]),
            "orc_random_ins": (v, [DP, i, DP, DP, DP, d, d, d, d, d, d, d, DP,
                                   DP, DP]),
            "orc_random_disp": (v, [DP, DP, i, DP, DP, DP, DP, d, d, d, d, d,
                                    d, d, d, d]),
            "orc_sys_create": (p, []),
            "orc_sys_destroy": (v, [p]),
            "orc_sys_set_cell": (i, [p, i, DP]),
            "orc_sys_set_ewald": (i, [p, d, i, i, i, d]),
            "orc_sys_set_cutoffs": (i, [p, d, d, d]),
            "orc_sys_set_vdw": (i, [p, i, i, IP, IP, IP, DP]),
            "orc_sys_add_framework": (i, [p, i, DP, DP, DP, DP, IP, IP]),
            "orc_sys_add_guest_type": (i, [p, i, DP, DP, DP, DP, IP, DP]),
            "orc_sys_add_guest_molecule": (i, [p, i, DP, DP, DP]),
            "orc_sys_finalize": (i, [p]),
            "orc_sys_nk": (i, [p]),
            "orc_sys_nmol": (i, [p, i]),
            "orc_sys_nsites": (i, [p, i]),
            "orc_sys_volume": (d, [p]),
            "orc_sys_ewald3en": (d, [p, i]),
            "orc_sys_get_molecule": (v, [p, i, i, DP, DP, DP]),
            "orc_sys_get_sk": (v, [p, DP, DP]),
            "orc_sys_get_table": (i, [p, i, i, DP]),
            "orc_sys_grid_info": (v, [p, IP, DP, DP]),
            "orc_sys_recip_full": (d, [p, i]),
            "orc_sys_trial": (i, [p, i, i, i, DP, DP, DP, DP]),
            "orc_sys_commit": (i, [p]),
            "orc_sys_tail_delta": (d, [p, i, i, d]),
            "orc_sys_energy_full": (v, [p, DP]),
            "orc_chain_create": (p, [p, C.POINTER(ChainParams)]),
            "orc_chain_destroy": (v, [p]),
            "orc_chain_run": (i, [p, C.c_long, C.POINTER(MoveRec)]),
            "orc_chain_energy": (d, [p, i]),
            "orc_chain_adapt_state": (v, [p, i, DP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# convenience wrappers
# ---------------------------------------------------------------------------

class Duni:
    def __init__(self, i, j, k, l):
        self.st = DuniState()
        lib().orc_duni_init(C.byref(self.st), i, j, k, l)

    def __call__(self) -> float:
        return lib().orc_duni_next(C.byref(self.st))


class OracleSystem:
    """orc_sys built from explicit arrays (same builder sequence as the
    product's fmc_* set-up calls)."""

    def __init__(self, imcon, cell, alpha, kmax, rcut, delr, ntpatm, vdw,
                 fw, guests, rvdw=0.0, epsq=1.0, guest_mols=None):
        L = lib()
        self.L = L
        self.h = L.orc_sys_create()
        self._keep = []
        c, cp = _d(cell)
        L.orc_sys_set_cell(self.h, imcon, cp)
        L.orc_sys_set_ewald(self.h, alpha, int(kmax[0]), int(kmax[1]),
                            int(kmax[2]), epsq)
        L.orc_sys_set_cutoffs(self.h, rcut, delr, rvdw)
        ta, tap = _i([v[0] for v in vdw] or [0])
        tb, tbp = _i([v[1] for v in vdw] or [0])
        ky, kyp = _i([v[2] for v in vdw] or [0])
        pr, prp = _d(np.array([v[3][:5] for v in vdw]).reshape(-1) if vdw else [0.0] * 5)
        L.orc_sys_set_vdw(self.h, ntpatm, len(vdw), tap, tbp, kyp, prp)
        x, xp = _d(fw["x"]); y, yp = _d(fw["y"]); z, zp = _d(fw["z"])
        q, qp = _d(fw["q"]); t, tp = _i(fw["type"])
        fr, frp = _i(fw.get("frozen", np.ones(len(fw["x"]))))
        L.orc_sys_add_framework(self.h, len(x), xp, yp, zp, qp, tp, frp)
        self.nsites = []
        for g in guests:
            gx, gxp = _d(g["x"]); gy, gyp = _d(g["y"]); gz, gzp = _d(g["z"])
            gq, gqp = _d(g["q"]); gt, gtp = _i(g["type"]); gm, gmp = _d(g["mass"])
            L.orc_sys_add_guest_type(self.h, len(gx), gxp, gyp, gzp, gqp, gtp, gmp)
            self.nsites.append(len(gx))
        if guest_mols:
            for (gtype, mx, my, mz) in guest_mols:
                a, ap = _d(mx); b, bp = _d(my); cc, ccp = _d(mz)
                L.orc_sys_add_guest_molecule(self.h, gtype, ap, bp, ccp)
        L.orc_sys_finalize(self.h)

    def __del__(self):
        try:
            if self.h:
                self.L.orc_sys_destroy(self.h)
        except Exception:
            pass

    @property
    def nk(self):
        return self.L.orc_sys_nk(self.h)

    @property
    def volume(self):
        return self.L.orc_sys_volume(self.h)

    def nmol(self, g):
        return self.L.orc_sys_nmol(self.h, g)

    def ewald3en(self, g):
        return self.L.orc_sys_ewald3en(self.h, g)

    def recip_full(self, all_atoms=1):
        return self.L.orc_sys_recip_full(self.h, all_atoms)

    def energy_full(self):
        out = np.zeros(6)
        self.L.orc_sys_energy_full(self.h, out.ctypes.data_as(DP))
        return out

    def trial(self, kind, gtype, imol=-1, x=None, y=None, z=None):
        ns = self.nsites[gtype]
        x = np.zeros(ns) if x is None else x
        y = np.zeros(ns) if y is None else y
        z = np.zeros(ns) if z is None else z
        a, ap = _d(x); b, bp = _d(y); c, cp = _d(z)
        out = np.zeros(3)
        r = self.L.orc_sys_trial(self.h, kind, gtype, imol, ap, bp, cp,
                                 out.ctypes.data_as(DP))
        if r != 0:
            raise RuntimeError(f"orc_sys_trial failed {r}")
        return out

    def commit(self):
        if self.L.orc_sys_commit(self.h) != 0:
            raise RuntimeError("orc_sys_commit failed")

    def tail_delta(self, gtype, sign, elrc):
        return self.L.orc_sys_tail_delta(self.h, gtype, sign, elrc)

    def molecule(self, g, m):
        ns = self.nsites[g]
        x, y, z = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        self.L.orc_sys_get_molecule(self.h, g, m, x.ctypes.data_as(DP),
                                    y.ctypes.data_as(DP), z.ctypes.data_as(DP))
        return x, y, z

    def sk(self):
        re = np.zeros(self.nk); im = np.zeros(self.nk)
        self.L.orc_sys_get_sk(self.h, re.ctypes.data_as(DP), im.ctypes.data_as(DP))
        return re, im

    def table(self, which, t=0):
        mg = C.c_int(); d1 = C.c_double(); d2 = C.c_double()
        self.L.orc_sys_grid_info(self.h, C.byref(mg), C.byref(d1), C.byref(d2))
        out = np.zeros(mg.value + 1)
        self.L.orc_sys_get_table(self.h, which, t, out.ctypes.data_as(DP))
        return out

    def grid_info(self):
        mg = C.c_int(); d1 = C.c_double(); d2 = C.c_double()
        self.L.orc_sys_grid_info(self.h, C.byref(mg), C.byref(d1), C.byref(d2))
        return mg.value, d1.value, d2.value


class OracleChain:
    def __init__(self, system: OracleSystem, temp, press_pa, seed,
                 freq=None, delr=1.0, engunit=418.4, ratios=(0.5, 0.5, 0.5)):
        self.sys = system
        p = ChainParams()
        p.temp = temp
        p.ntpguest = len(press_pa)
        for g, pr in enumerate(press_pa):
            p.press[g] = pr
            p.delr[g] = delr
            p.disp_ratio[g] = ratios[0]
            p.tran_ratio[g] = ratios[1]
            if freq is not None:
                for k in range(9):
                    p.freq[g][k] = freq[g][k] if k < len(freq[g]) else 0.0
        p.rota_ratio = ratios[2]
        for k in range(4):
            p.seed[k] = seed[k]
        p.engunit = engunit
        self.p = p
        self.h = lib().orc_chain_create(system.h, C.byref(p))

    def run(self, n):
        recs = (MoveRec * n)()
        r = lib().orc_chain_run(self.h, n, recs)
        if r != 0:
            raise RuntimeError(f"orc_chain_run failed {r}")
        return recs

    def energy(self, g):
        return lib().orc_chain_energy(self.h, g)

    def __del__(self):
        try:
            lib().orc_chain_destroy(self.h)
        except Exception:
            pass


def system_from_inputs(field, config, control, guest_mols_from_config=True):
    """Build an OracleSystem from parsed reference input files."""
    imcon = config.imcon
    cell = config.cell
    rcut = control.rcut
    if control.ewald_sum is not None:
        alpha, k1, k2, k3 = control.ewald_sum
        kmax = (k1, k2, k3)
    else:
        a = C.c_double()
        km = (C.c_int * 3)()
        cc, ccp = _d(cell)
        lib().orc_ewald_precision(control.ewald_precision, rcut, imcon, ccp,
                                  C.byref(a), km)
        alpha, kmax = a.value, (km[0], km[1], km[2])
    fx, fy, fz, fq, ft, ff = [], [], [], [], [], []
    guests, gmols = [], []
    for k, mol in enumerate(field.molecules):
        if mol.guest:
            gi = len(guests)
            guests.append(dict(x=mol.gx, y=mol.gy, z=mol.gz, q=mol.charge,
                               type=mol.types, mass=mol.mass))
            if guest_mols_from_config:
                crd = config.coords.get(k, [])
                ns = len(mol.names)
                for m in range(mol.nummols):
                    pts = crd[m * ns:(m + 1) * ns]
                    gmols.append((gi, [p[0] for p in pts], [p[1] for p in pts],
                                  [p[2] for p in pts]))
        else:
            crd = config.coords[k]
            ns = len(mol.names)
            for m in range(mol.nummols):
                for s in range(ns):
                    x, y, z = crd[m * ns + s]
                    fx.append(x); fy.append(y); fz.append(z)
                    fq.append(mol.charge[s]); ft.append(mol.types[s])
                    ff.append(mol.frozen[s])
    fw = dict(x=fx, y=fy, z=fz, q=fq, type=ft, frozen=ff)
    return OracleSystem(imcon, cell, alpha, kmax, rcut, control.delr,
                        len(field.unqatm), field.vdw, fw, guests,
                        rvdw=control.rvdw, guest_mols=gmols), alpha, kmax
