"""ctypes binding of include/fastmc_amd.h and a thin Python mirror of the
single-chain context (the reference's trial-move interface: trial energies
per move, commit on acceptance)."""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libfastmc_amd.so")

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

ERRORS = {-1: "FMC_ERR_ARG", -2: "FMC_ERR_STATE", -3: "FMC_ERR_UNSUPPORTED",
          -4: "FMC_ERR_CAPACITY", -5: "FMC_ERR_HIP", -6: "FMC_ERR_NODEVICE",
          -7: "FMC_ERR_CELL"}

INSERT, DELETE, DISPLACE = 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1


class FmcError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: {ERRORS.get(code, code)}")
        self.code = code


class ChainParams(C.Structure):
    _fields_ = [("temp", C.c_double), ("ntpguest", C.c_int),
                ("seed", C.c_int * 4), ("press", C.c_double * 4),
                ("freq", (C.c_double * 9) * 4), ("delr", C.c_double * 4),
                ("disp_ratio", C.c_double * 4), ("tran_ratio", C.c_double * 4),
                ("rota_ratio", C.c_double), ("engunit", C.c_double),
                ("eq_steps", C.c_longlong), ("max_moves", C.c_longlong),
                ("recompute_every", C.c_longlong), ("window", C.c_longlong)]


class ChainStats(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("prod_steps", C.c_longlong),
                ("accepted", C.c_longlong * 6), ("attempted", C.c_longlong * 6),
                ("nmol", C.c_int * 4), ("energy", C.c_double * 4),
                ("sum_n", C.c_double * 4), ("sum_n2", C.c_double * 4),
                ("sum_e", C.c_double * 4), ("delrdisp", C.c_double * 4),
                ("rotangle", C.c_double), ("tran_delr", C.c_double * 4),
                ("rota_rotangle", C.c_double), ("e_acc", C.c_double),
                ("e_full", C.c_double), ("drift", C.c_double),
                ("sk_dev", C.c_double), ("recomputes", C.c_longlong),
                ("sum_en", C.c_double * 4), ("sum_e2", C.c_double * 4),
                ("win_len", C.c_longlong), ("win_closed", C.c_int),
                ("win", ((C.c_double * 4) * 5) * 5),
                ("sel_window", C.c_longlong), ("sel_count", C.c_longlong),
                ("sel_base", C.c_double * 4), ("sel_mean", C.c_double * 6),
                ("sel_second", C.c_double * 6), ("sel_var", C.c_double * 6),
                ("tile_threads", C.c_int), ("tile_kpt", C.c_int),
                ("tile_kgroup", C.c_int), ("tile_kunroll", C.c_int)]


_lib = None

_SIGS = {
    "fmc_ctx_create": [C.POINTER(C.c_void_p)],
    "fmc_ctx_destroy": [C.c_void_p],
    "fmc_set_cell": [C.c_void_p, C.c_int, DP],
    "fmc_set_ewald": [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double],
    "fmc_set_ewald_precision": [C.c_void_p, C.c_double, C.c_double],
    "fmc_set_cutoffs": [C.c_void_p, C.c_double, C.c_double, C.c_double],
    "fmc_set_vdw": [C.c_void_p, C.c_int, C.c_int, IP, IP, IP, DP],
    "fmc_add_framework": [C.c_void_p, C.c_int, DP, DP, DP, DP, IP, IP],
    "fmc_add_guest_type": [C.c_void_p, C.c_int, DP, DP, DP, DP, IP, DP, IP],
    "fmc_add_guest_molecule": [C.c_void_p, C.c_int, DP, DP, DP],
    "fmc_set_capacity": [C.c_void_p, C.c_int],
    "fmc_ctx_finalize": [C.c_void_p],
    "fmc_get_info": [C.c_void_p, IP, DP, IP, DP, DP],
    "fmc_guest_ewald3": [C.c_void_p, C.c_int, DP],
    "fmc_guest_count": [C.c_void_p, C.c_int, IP],
    "fmc_get_molecule": [C.c_void_p, C.c_int, C.c_int, DP, DP, DP],
    "fmc_get_table": [C.c_void_p, C.c_int, C.c_int, DP],
    "fmc_get_sk": [C.c_void_p, DP, DP],
    "fmc_recip_full": [C.c_void_p, C.c_int, DP],
    "fmc_trial": [C.c_void_p, C.c_int, C.c_int, C.c_int, DP, DP, DP, DP],
    "fmc_commit": [C.c_void_p],
    "fmc_tail_delta": [C.c_void_p, C.c_int, C.c_int, C.c_double, DP],
    "fmc_energy_full": [C.c_void_p, DP],
    "fmc_energy_full_noalloc": [C.c_void_p, DP],
    "fmc_ctx_timing": [C.c_void_p, C.c_int, DP],
    "fmc_batch_create": [C.POINTER(C.c_void_p), C.c_int, C.c_int],
    "fmc_batch_destroy": [C.c_void_p],
    "fmc_batch_add_chain": [C.c_void_p, C.c_void_p, C.POINTER(ChainParams), IP],
    "fmc_batch_remove_chain": [C.c_void_p, C.c_int],
    "fmc_batch_checkpoint": [C.c_void_p, C.c_int, C.c_void_p, C.c_longlong,
                             C.POINTER(C.c_longlong)],
    "fmc_batch_restore_chain": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, IP],
    "fmc_batch_get_molecules": [C.c_void_p, C.c_int, C.c_int, DP, DP, DP, C.c_int, IP],
    "fmc_batch_run": [C.c_void_p, C.c_longlong],
    "fmc_batch_launch": [C.c_void_p, C.c_longlong],
    "fmc_batch_sync": [C.c_void_p],
    "fmc_batch_launch_budget": [C.c_void_p, C.c_longlong, C.c_double],
    "fmc_batch_query": [C.c_void_p, IP],
    "fmc_batch_stats": [C.c_void_p, C.c_int, C.POINTER(ChainStats)],
    "fmc_batch_enable_log": [C.c_void_p, C.c_int, C.c_longlong],
    "fmc_batch_get_log": [C.c_void_p, C.c_int, C.c_longlong, DP],
    "fmc_batch_resync": [C.c_void_p, DP],
    "fmc_batch_recompute": [C.c_void_p, DP, DP],
    "fmc_batch_launch_cycles": [C.c_void_p, C.c_longlong, C.c_longlong],
    "fmc_batch_poll": [C.c_void_p, IP, C.c_int, IP],
    "fmc_batch_timing": [C.c_void_p, C.c_int, DP, C.POINTER(C.c_longlong)],
    "fmc_batch_set_tile": [C.c_void_p, C.c_int],
    "fmc_batch_get_tile": [C.c_void_p, IP],
    "fmc_batch_set_tile_shape": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int],
    "fmc_batch_get_tile_shape": [C.c_void_p, IP, IP, IP, IP, IP],
    "fmc_chain_footprint": [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)],
    "fmc_version": [IP, IP],
    "fmc_device_ok": [],
}

EXPORTED = sorted(_SIGS)


def lib():
    """Load libfastmc_amd.so; raise if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FmcError(-6, f"HIP library not built: {LIB_PATH}; run "
                           "__graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        _lib = L
    return _lib


def check(code, what):
    if code != 0:
        raise FmcError(code, what)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(DP)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(IP)


class Context:
    """One chain's system on the GPU (the drop-in single-chain path)."""

    def __init__(self, imcon, cell, alpha, kmax, rcut, delr, ntpatm, vdw, fw,
                 guests, rvdw=0.0, epsq=1.0, guest_mols=None,
                 max_guest_atoms=3000, precision_eps=None):
        L = lib()
        self.L = L
        h = C.c_void_p()
        check(L.fmc_ctx_create(C.byref(h)), "fmc_ctx_create")
        self.h = h
        c, cp = _d(cell)
        check(L.fmc_set_cell(h, imcon, cp), "fmc_set_cell")
        check(L.fmc_set_cutoffs(h, rcut, delr, rvdw), "fmc_set_cutoffs")
        if precision_eps is not None:
            check(L.fmc_set_ewald_precision(h, precision_eps, epsq),
                  "fmc_set_ewald_precision")
        else:
            check(L.fmc_set_ewald(h, alpha, int(kmax[0]), int(kmax[1]),
                                  int(kmax[2]), epsq), "fmc_set_ewald")
        ta, tap = _i([v[0] for v in vdw] or [0])
        tb, tbp = _i([v[1] for v in vdw] or [0])
        ky, kyp = _i([v[2] for v in vdw] or [0])
        pr, prp = _d(np.array([list(v[3][:5]) + [0.0] * (5 - len(v[3][:5]))
                               for v in vdw]).reshape(-1) if vdw else [0.0] * 5)
        check(L.fmc_set_vdw(h, ntpatm, len(vdw), tap, tbp, kyp, prp), "fmc_set_vdw")
        x, xp = _d(fw["x"]); y, yp = _d(fw["y"]); z, zp = _d(fw["z"])
        q, qp = _d(fw["q"]); t, tp = _i(fw["type"])
        fr, frp = _i(fw.get("frozen", np.ones(len(fw["x"]))))
        check(L.fmc_add_framework(h, len(x), xp, yp, zp, qp, tp, frp),
              "fmc_add_framework")
        self.nsites = []
        for g in guests:
            gx, gxp = _d(g["x"]); gy, gyp = _d(g["y"]); gz, gzp = _d(g["z"])
            gq, gqp = _d(g["q"]); gt, gtp = _i(g["type"]); gm, gmp = _d(g["mass"])
            out = C.c_int()
            check(L.fmc_add_guest_type(h, len(gx), gxp, gyp, gzp, gqp, gtp, gmp,
                                       C.byref(out)), "fmc_add_guest_type")
            self.nsites.append(len(gx))
        for (gtype, mx, my, mz) in (guest_mols or []):
            a, ap = _d(mx); b, bp = _d(my); cc, ccp = _d(mz)
            check(L.fmc_add_guest_molecule(h, gtype, ap, bp, ccp),
                  "fmc_add_guest_molecule")
        check(L.fmc_set_capacity(h, max_guest_atoms), "fmc_set_capacity")
        check(L.fmc_ctx_finalize(h), "fmc_ctx_finalize")

    @classmethod
    def from_arrays(cls, A, **kw):
        return cls(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                   A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"], **kw)

    def close(self):
        if getattr(self, "h", None):
            self.L.fmc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        nk = C.c_int(); v = C.c_double(); mg = C.c_int()
        d1 = C.c_double(); d2 = C.c_double()
        check(self.L.fmc_get_info(self.h, C.byref(nk), C.byref(v), C.byref(mg),
                                  C.byref(d1), C.byref(d2)), "fmc_get_info")
        return dict(nk=nk.value, volm=v.value, mxgrid=mg.value,
                    drewd=d1.value, dlrpot=d2.value)

    def ewald3en(self, g):
        v = C.c_double()
        check(self.L.fmc_guest_ewald3(self.h, g, C.byref(v)), "fmc_guest_ewald3")
        return v.value

    def nmol(self, g):
        v = C.c_int()
        check(self.L.fmc_guest_count(self.h, g, C.byref(v)), "fmc_guest_count")
        return v.value

    def molecule(self, g, m):
        ns = self.nsites[g]
        x, y, z = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        check(self.L.fmc_get_molecule(self.h, g, m, x.ctypes.data_as(DP),
                                      y.ctypes.data_as(DP), z.ctypes.data_as(DP)),
              "fmc_get_molecule")
        return x, y, z

    def table(self, which, t=0):
        out = np.zeros(self.info()["mxgrid"] + 1)
        check(self.L.fmc_get_table(self.h, which, t, out.ctypes.data_as(DP)),
              "fmc_get_table")
        return out

    def sk(self):
        nk = self.info()["nk"]
        re, im = np.zeros(nk), np.zeros(nk)
        check(self.L.fmc_get_sk(self.h, re.ctypes.data_as(DP), im.ctypes.data_as(DP)),
              "fmc_get_sk")
        return re, im

    def recip_full(self, all_atoms=1):
        e = C.c_double()
        check(self.L.fmc_recip_full(self.h, all_atoms, C.byref(e)), "fmc_recip_full")
        return e.value

    def trial(self, kind, gtype, imol=-1, x=None, y=None, z=None):
        # an unknown guest type is rejected by the C ABI (FMC_ERR_ARG)
        ns = self.nsites[gtype] if 0 <= gtype < len(self.nsites) else 8
        x = np.zeros(ns) if x is None else x
        y = np.zeros(ns) if y is None else y
        z = np.zeros(ns) if z is None else z
        a, ap = _d(x); b, bp = _d(y); c, cp = _d(z)
        out = np.zeros(3)
        check(self.L.fmc_trial(self.h, kind, gtype, imol, ap, bp, cp,
                               out.ctypes.data_as(DP)), "fmc_trial")
        return out

    def commit(self):
        check(self.L.fmc_commit(self.h), "fmc_commit")

    def tail_delta(self, gtype, sign, elrc):
        v = C.c_double()
        check(self.L.fmc_tail_delta(self.h, gtype, sign, elrc, C.byref(v)),
              "fmc_tail_delta")
        return v.value

    def energy_full(self):
        out = np.zeros(6)
        check(self.L.fmc_energy_full(self.h, out.ctypes.data_as(DP)), "fmc_energy_full")
        return out

    def energy_full_noalloc(self):
        """energy_full on the library's static device scratch (no device
        allocation); the device state must exist already."""
        out = np.zeros(6)
        check(self.L.fmc_energy_full_noalloc(self.h, out.ctypes.data_as(DP)),
              "fmc_energy_full_noalloc")
        return out


def context_from_inputs(field, config, control, **kw):
    """Context built from parsed reference input files (FIELD order)."""
    fx, fy, fz, fq, ft, ff = [], [], [], [], [], []
    guests, gmols = [], []
    for k, mol in enumerate(field.molecules):
        ns = len(mol.names)
        crd = config.coords.get(k, [])
        if mol.guest:
            gi = len(guests)
            guests.append(dict(x=mol.gx, y=mol.gy, z=mol.gz, q=mol.charge,
                               type=mol.types, mass=mol.mass))
            for m in range(mol.nummols):
                pts = crd[m * ns:(m + 1) * ns]
                gmols.append((gi, [p[0] for p in pts], [p[1] for p in pts],
                              [p[2] for p in pts]))
        else:
            for m in range(mol.nummols):
                for s in range(ns):
                    x, y, z = crd[m * ns + s]
                    fx.append(x); fy.append(y); fz.append(z)
                    fq.append(mol.charge[s]); ft.append(mol.types[s])
                    ff.append(mol.frozen[s])
    fw = dict(x=fx, y=fy, z=fz, q=fq, type=ft, frozen=ff)
    if control.ewald_sum is not None:
        alpha, k1, k2, k3 = control.ewald_sum
        return Context(config.imcon, config.cell, alpha, (k1, k2, k3),
                       control.rcut, control.delr, len(field.unqatm),
                       field.vdw, fw, guests, rvdw=control.rvdw,
                       guest_mols=gmols, max_guest_atoms=control.maxguest, **kw)
    return Context(config.imcon, config.cell, 0.0, (0, 0, 0), control.rcut,
                   control.delr, len(field.unqatm), field.vdw, fw, guests,
                   rvdw=control.rvdw, guest_mols=gmols,
                   max_guest_atoms=control.maxguest,
                   precision_eps=control.ewald_precision, **kw)
