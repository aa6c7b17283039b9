"""Pins the oracle's trial positions (random_ins, utility_pack.f:2545-2650;
random_disp + random_rot, :2679-2763 and :2929-3000, with the quaternion
matrix of rotation, :2887-2926) against a construction the mathematics
fixes: a rotation by the angle theta about the axis u = normalize(2 r2 - 1,
2 r3 - 1, 2 r4 - 1) is Rodrigues' rotation, which scipy builds from the
rotation vector theta * u.  theta is 2 pi r1 for an insertion and r1 *
rotangle for a displacement, whose rotation acts about the centre of mass
after the translation and the wrap of the centre of mass into the cell.

The template is methane (five sites, not coplanar), so the rotation is
determined completely; det R = +1 is checked on the matrix the oracle's
output implies.  The same checks run against oracle builds with a planted
mistake -- the matrix transposed, a reflection, the axis components
swapped -- and must reject each of them."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import ORACLE
from fastmc_amd import synth

CELL = np.array([25.0, 0.0, 0.0, 3.0, 24.0, 0.0, 2.0, 4.0, 23.0])   # a, b, c rows


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _template():
    g = synth.Workload("t", synth.c1().fw, ["ch4"], [1.0], 298.0, (8, 8, 8))
    t = g.system_arrays()["guests"][0]
    return [np.ascontiguousarray(t[k], float) for k in ("x", "y", "z", "mass")]


def _axis(r2, r3, r4):
    b = np.array([2 * r2 - 1, 2 * r3 - 1, 2 * r4 - 1])
    return b / np.linalg.norm(b)


def _bind(L):
    d, i, DP = C.c_double, C.c_int, C.POINTER(C.c_double)
    L.orc_random_ins.restype = None
    L.orc_random_ins.argtypes = [DP, i, DP, DP, DP, d, d, d, d, d, d, d, DP, DP, DP]
    L.orc_random_disp.restype = None
    L.orc_random_disp.argtypes = [DP, DP, i, DP, DP, DP, DP, d, d, d, d, d, d, d, d, d]
    L.orc_invert.restype = None
    L.orc_invert.argtypes = [DP, DP, C.POINTER(C.c_double)]
    return L


def _implied_matrix(src, dst):
    """The linear map taking the centred rows of src onto those of dst."""
    a = src - src.mean(axis=0)
    b = dst - dst.mean(axis=0)
    m, *_ = np.linalg.lstsq(a, b, rcond=None)
    return m.T


def _rotation_errors(L, n=200, seed=11):
    """Largest deviation of the oracle's insertion and displacement
    positions from the Rodrigues construction, and the extreme values of
    det R of the matrices they imply."""
    tx, ty, tz, m = _template()
    ns = len(tx)
    T = np.vstack([tx, ty, tz]).T
    cell = CELL.copy()
    A = cell.reshape(3, 3).T          # columns a, b, c
    rcell = np.zeros(9)
    det = C.c_double()
    L.orc_invert(_dp(cell), _dp(rcell), C.byref(det))
    rng = np.random.default_rng(seed)
    worst, dets = 0.0, []
    nx, ny, nz = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    for _ in range(n):
        ra, rb, rc, r1, r2, r3, r4 = rng.uniform(0.0, 1.0, 7)
        L.orc_random_ins(_dp(cell), ns, _dp(tx), _dp(ty), _dp(tz), ra, rb, rc, r1, r2,
                         r3, r4, _dp(nx), _dp(ny), _dp(nz))
        got = np.vstack([nx, ny, nz]).T.copy()
        R = Rotation.from_rotvec(2 * math.pi * r1 * _axis(r2, r3, r4)).as_matrix()
        want = T @ R.T + A @ np.array([ra, rb, rc])
        worst = max(worst, float(np.abs(got - want).max()))
        dets.append(np.linalg.det(_implied_matrix(T, got)))
        # a displacement of that molecule
        delr, rotangle = 1.3, math.pi / 3
        ra, rb, rc, r1, r2, r3, r4 = rng.uniform(0.0, 1.0, 7)
        before = got.copy()
        L.orc_random_disp(_dp(cell), _dp(rcell), ns, _dp(m), _dp(nx), _dp(ny), _dp(nz),
                          delr, rotangle, ra, rb, rc, r1, r2, r3, r4)
        got = np.vstack([nx, ny, nz]).T.copy()
        moved = before + (2 * np.array([ra, rb, rc]) - 1) * (0.5 * delr)
        com = (moved * m[:, None]).sum(axis=0) / m.sum()
        f = np.linalg.solve(A, com)
        wrapped = A @ (f - np.floor(f))
        R = Rotation.from_rotvec(r1 * rotangle * _axis(r2, r3, r4)).as_matrix()
        want = (moved - com) @ R.T + wrapped
        worst = max(worst, float(np.abs(got - want).max()))
        dets.append(np.linalg.det(_implied_matrix(before - before.mean(0), got)))
    return worst, min(dets), max(dets)


def test_trial_positions_are_rodrigues_rotations(oracle):
    worst, dmin, dmax = _rotation_errors(_bind(oracle.lib()))
    print(f"\nworst deviation {worst:.2e} A; det R in [{dmin:.15f}, {dmax:.15f}]")
    assert worst < 1e-12
    assert abs(dmin - 1.0) < 1e-12 and abs(dmax - 1.0) < 1e-12


# planted mistakes: each edits the quaternion matrix of orc_rotation (the
# restatement of utility_pack.f:2906-2914) or the axis of orc_quat
MUTANTS = {
    "transposed": [("Q[1] = 2.0 * q2 * q3 - 2.0 * q1 * q4;", "Q[1] = 2.0 * q2 * q3 + 2.0 * q1 * q4;"),
                   ("Q[3] = 2.0 * q2 * q3 + 2.0 * q1 * q4;", "Q[3] = 2.0 * q2 * q3 - 2.0 * q1 * q4;"),
                   ("Q[2] = 2.0 * q2 * q4 + 2.0 * q1 * q3;", "Q[2] = 2.0 * q2 * q4 - 2.0 * q1 * q3;"),
                   ("Q[6] = 2.0 * q2 * q4 - 2.0 * q1 * q3;", "Q[6] = 2.0 * q2 * q4 + 2.0 * q1 * q3;"),
                   ("Q[5] = 2.0 * q3 * q4 - 2.0 * q1 * q2;", "Q[5] = 2.0 * q3 * q4 + 2.0 * q1 * q2;"),
                   ("Q[7] = 2.0 * q3 * q4 + 2.0 * q1 * q2;", "Q[7] = 2.0 * q3 * q4 - 2.0 * q1 * q2;")],
    "reflection": [("Q[8] = 2.0 * q4 * q4 + 2.0 * q1 * q1 - 1.0;",
                    "Q[8] = -(2.0 * q4 * q4 + 2.0 * q1 * q1 - 1.0);"),
                   ("Q[6] = 2.0 * q2 * q4 - 2.0 * q1 * q3;", "Q[6] = -(2.0 * q2 * q4 - 2.0 * q1 * q3);"),
                   ("Q[7] = 2.0 * q3 * q4 + 2.0 * q1 * q2;", "Q[7] = -(2.0 * q3 * q4 + 2.0 * q1 * q2);")],
    "axis_swapped": [("double u1 = beta1 / norm, u2 = beta2 / norm, u3 = beta3 / norm;",
                      "double u1 = beta2 / norm, u2 = beta1 / norm, u3 = beta3 / norm;")],
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_rotation_mistakes_are_rejected(name, tmp_path):
    src = open(os.path.join(ORACLE, "fastmc_oracle.c")).read()
    for a, b in MUTANTS[name]:
        assert src.count(a) == 1, a
        src = src.replace(a, b)
    c = tmp_path / "mutant.c"
    c.write_text(src)
    so = tmp_path / "libmutant.so"
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off",
                           "-I", ORACLE, "-o", str(so), str(c), "-lm"])
    worst, dmin, dmax = _rotation_errors(_bind(C.CDLL(str(so))), n=40)
    print(f"\n{name}: worst deviation {worst:.2e} A, det R in [{dmin:.3f}, {dmax:.3f}]")
    assert worst > 1e-3
