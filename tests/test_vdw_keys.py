"""The nine vdW forms (vdw_terms.f:189-396 tables, 555-650 tails) pinned
by identities between the forms, so that a dropped term, a wrong sign or a
swapped parameter in any of them fails:

* 12-6 with A = 4 eps sigma^12, B = 4 eps sigma^6, and n-m with n = 12,
  m = 6, r0 = 2^(1/6) sigma, are the Lennard-Jones form;
* n-m with n = 12, m = 10 is the 12-10 hydrogen bond with A = 5 E0 r0^12,
  B = 6 E0 r0^10;
* the shifted n-m form with its cutoff far away is the n-m form, and with a
  cutoff inside the table it and its slope vanish there;
* Buckingham without dispersion is a pure exponential, without repulsion
  the 12-6 dispersion; Born-Huggins-Meyer reduces to Buckingham, and its
  r^-8 term times r^8 is constant;
* Morse is -E0 at r0, -3/4 E0 at r0 + ln2/k and 0 at r0 - ln2/k;
* WCA is the Lennard-Jones form shifted by eps and by delta, cut at its
  minimum.

The tails follow the same identities, with the r^-8 and n-m tails against
numerical integrals of r^2 V(r), and no tail for the shifted n-m, Morse and
WCA forms (the reference's lrcorrect has no branch for them).  The
product's host tables and tails equal the oracle's on a system that uses
every form."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

from keyed import key_params, keyed_arrays

EPS, SIG = 0.2 * 418.4, 3.1
RV = 12.5


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _table(oracle, key, prm, rvdw=RV):
    L = oracle.lib()
    mg = L.orc_mxgrid(rvdw)
    vvv = np.zeros(2 * (mg + 1))
    keys = np.array([key], dtype=np.int32)
    p = np.zeros(5)
    p[:len(prm)] = prm
    dl = L.orc_forgen(1, _ip(keys), _dp(p), rvdw, mg, _dp(vvv))
    r = np.arange(mg + 1) * dl
    return r[1:], vvv[1:mg + 1], dl


def _same(a, b, scale, tol=1e-12):
    assert np.max(np.abs(a - b)) <= tol * scale, np.max(np.abs(a - b)) / scale


def test_12_6_and_n_m_are_lennard_jones(oracle):
    r, lj, _ = _table(oracle, 2, [EPS, SIG])
    keep = r > 0.5 * SIG          # away from the r^-12 wall's huge values
    scale = np.abs(lj[keep]).max()
    _, v1, _ = _table(oracle, 1, [4 * EPS * SIG ** 12, 4 * EPS * SIG ** 6])
    _same(v1[keep], lj[keep], scale)
    _, v3, _ = _table(oracle, 3, [EPS, 12.0, 6.0, 2 ** (1 / 6) * SIG])
    _same(v3[keep], lj[keep], scale)
    # the order of n and m does not matter (vdw_terms.f takes max and min)
    _, v3b, _ = _table(oracle, 3, [EPS, 6.0, 12.0, 2 ** (1 / 6) * SIG])
    _same(v3b[keep], lj[keep], scale)


def test_n_m_with_12_10_is_the_hydrogen_bond(oracle):
    e0, r0 = 0.3 * 418.4, 2.9
    r, v3, _ = _table(oracle, 3, [e0, 12.0, 10.0, r0])
    _, v6, _ = _table(oracle, 6, [5 * e0 * r0 ** 12, 6 * e0 * r0 ** 10])
    keep = r > 0.5 * r0
    _same(v6[keep], v3[keep], np.abs(v3[keep]).max())


@pytest.mark.parametrize("n,m", [(12.0, 6.0), (9.0, 6.0)])
def test_shifted_n_m_far_cutoff_is_n_m(oracle, n, m):
    r0 = 3.4
    r, v3, _ = _table(oracle, 3, [EPS, n, m, r0])
    _, v7, _ = _table(oracle, 7, [EPS, n, m, r0, 1e4 * r0])
    keep = r > 0.5 * r0
    _same(v7[keep], v3[keep], np.abs(v3[keep]).max(), tol=1e-11)


def test_shifted_n_m_vanishes_with_its_slope_at_the_cutoff(oracle):
    rc = 10.0
    r, v7, dl = _table(oracle, 7, [EPS, 12.0, 6.0, 3.4, rc])
    assert np.all(v7[r > rc] == 0.0)
    # value and slope zero at rc: a quartic fitted to V(rc - x) over the
    # last 0.3 A has no constant or linear part (the unshifted form has
    # V(rc) ~ -0.26 and V'(rc) ~ 0.15 here, some 1e3 c2 and 1e1 c2)
    sel = (r <= rc) & (r > rc - 0.3)
    x = rc - r[sel]
    c4, c3, c2, c1, c0 = np.polyfit(x, v7[sel], 4)
    assert abs(c0) < 1e-4 * abs(c2) and abs(c1) < 1e-2 * abs(c2), (c0, c1, c2)
    assert c2 < 0.0


def test_buckingham_and_born_huggins_meyer(oracle):
    A, rho, Cd = 1.0e6, 0.27, 2.0e4
    r, v4, _ = _table(oracle, 4, [A, rho, 0.0])
    keep = r > 0.5
    # no dispersion: ln V is the line ln A - r / rho
    assert np.max(np.abs(np.log(v4[keep]) + r[keep] / rho - math.log(A))) < 1e-12
    # no repulsion: the 12-6 form's dispersion
    _, v4d, _ = _table(oracle, 4, [0.0, rho, Cd])
    _, v1d, _ = _table(oracle, 1, [0.0, Cd])
    _same(v4d[keep], v1d[keep], np.abs(v1d[keep]).max())
    # BHM A exp(B (s - r)) is Buckingham with A e^{B s}, rho = 1/B
    B, s = 1.0 / rho, 2.2
    _, v5, _ = _table(oracle, 5, [A, B, s, 0.0, 0.0])
    _, v4s, _ = _table(oracle, 4, [A * math.exp(B * s), rho, 0.0])
    _same(v5[keep], v4s[keep], np.abs(v4s[keep]).max())
    _, v5c, _ = _table(oracle, 5, [0.0, B, s, Cd, 0.0])
    _same(v5c[keep], v4d[keep], np.abs(v4d[keep]).max())
    D = 3.0e5
    _, v5d, _ = _table(oracle, 5, [0.0, B, s, 0.0, D])
    assert np.max(np.abs(v5d[keep] * r[keep] ** 8 + D)) < 1e-12 * D


def test_morse_points(oracle):
    _, _, dl = _table(oracle, 2, [EPS, SIG])
    i0, j = 350, 40
    r0, k = i0 * dl, math.log(2.0) / (j * dl)
    r, v8, _ = _table(oracle, 8, [EPS, r0, k])
    # table entry i (1-based) sits at r = i dl, so v8[i - 1] is V(i dl)
    assert abs(v8[i0 - 1] + EPS) < 1e-12 * EPS
    assert abs(v8[i0 + j - 1] + 0.75 * EPS) < 1e-12 * EPS
    assert abs(v8[i0 - j - 1]) < 1e-12 * EPS
    assert np.all(v8[r > 2 * r0] < 0.0) and np.all(np.diff(v8[i0:]) > 0.0)


@pytest.mark.parametrize("shift", [0, 25])
def test_wca_is_shifted_lennard_jones(oracle, shift):
    r, lj, dl = _table(oracle, 2, [EPS, SIG])
    delta = shift * dl
    _, v9, _ = _table(oracle, 9, [EPS, SIG, delta])
    rr = r - delta
    inside = (rr > 0.5 * SIG) & (rr < 2 ** (1 / 6) * SIG - 1e-9)
    idx = np.nonzero(inside)[0]
    want = lj[idx - shift] + EPS
    _same(v9[idx], want, np.abs(want).max())
    assert np.all(v9[rr > 2 ** (1 / 6) * SIG + 1e-9] == 0.0)
    if shift:
        assert np.all(v9[rr <= 0.0] == 0.0)


def _tail(oracle, key, prm, rc=RV):
    """eadd of one self pair (steadd, vdw_terms.f:555-650)."""
    L = oracle.lib()
    p = np.zeros(10)
    p[:len(prm)] = prm
    numtyp = np.array([10], dtype=np.int32)
    numfrz = np.array([0], dtype=np.int32)
    lst = np.array([1], dtype=np.int32)
    ltp = np.array([key, -1], dtype=np.int32)
    st = np.zeros(1)
    L.orc_lrcorrect(2, 1, _ip(numtyp), _ip(numfrz), _ip(lst), _ip(ltp), _dp(p), rc,
                    20000.0, _dp(st))
    return st[0]


def test_tails_by_identity_and_integral(oracle):
    rm = 2 ** (1 / 6) * SIG
    lj = _tail(oracle, 2, [EPS, SIG])
    assert _tail(oracle, 1, [4 * EPS * SIG ** 12, 4 * EPS * SIG ** 6]) == pytest.approx(lj, rel=1e-13)
    assert _tail(oracle, 3, [EPS, 12.0, 6.0, rm]) == pytest.approx(lj, rel=1e-13)
    e0, r0 = 0.3 * 418.4, 2.9
    assert _tail(oracle, 6, [5 * e0 * r0 ** 12, 6 * e0 * r0 ** 10]) == pytest.approx(
        _tail(oracle, 3, [e0, 12.0, 10.0, r0]), rel=1e-13)
    Cd, D = 2.0e4, 3.0e5
    t1 = _tail(oracle, 1, [0.0, Cd])
    assert _tail(oracle, 4, [1.0e6, 0.27, Cd]) == pytest.approx(t1, rel=1e-15)
    assert _tail(oracle, 5, [1.0e6, 3.7, 2.2, Cd, 0.0]) == pytest.approx(t1, rel=1e-15)
    i8 = integrate.quad(lambda r: -D / r ** 6, RV, np.inf, epsabs=0, epsrel=1e-13)[0]
    assert _tail(oracle, 5, [0.0, 3.7, 2.2, 0.0, D]) == pytest.approx(i8, rel=1e-11)
    n, m = 9.0, 6.0
    v = lambda r: r * r * EPS / (n - m) * (m * (r0 / r) ** n - n * (r0 / r) ** m)
    inm = integrate.quad(v, RV, np.inf, epsabs=0, epsrel=1e-13)[0]
    assert _tail(oracle, 3, [EPS, n, m, r0]) == pytest.approx(inm, rel=1e-11)
    for key, prm in ((7, [EPS, 12.0, 6.0, rm, 0.0]), (8, [EPS, rm, 2.0]), (9, [EPS, SIG, 0.0])):
        assert _tail(oracle, key, prm) == 0.0


def test_host_tables_and_tails_match_oracle_every_key(oracle):
    """The product's tables (fmc_ctx_finalize) on a system using all nine
    forms equal the oracle's, and so do the tail changes of inserts and
    deletes."""
    import fastmc_amd
    from fastmc_amd import synth
    rng = np.random.default_rng(21)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("keys", fw, ["co2", "n2"], [4.0, 8.0], 298.0, (8, 8, 8))
    A, used = keyed_arrays(w.system_arrays())
    assert used == list(range(1, 10))
    ctx = fastmc_amd.Context.from_arrays(A)
    s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                            A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    for t, (_, _, key, prm) in enumerate(A["vdw"]):
        v_p, v_o = ctx.table(1, t), s.table(1, t)
        assert np.allclose(v_p, v_o, rtol=1e-13, atol=1e-300), (t, key)
        assert np.any(v_p != 0.0) == (prm[0] != 0.0), (t, key)
    for g in range(len(A["guests"])):
        for sign in (1, -1):
            assert ctx.tail_delta(g, sign, 0.25) == pytest.approx(
                s.tail_delta(g, sign, 0.25), rel=1e-13, abs=1e-12)


def test_key_params_wells():
    """The keyed systems' parameter choice: each form has its well near
    2^(1/6) sigma with a depth of order eps (the repulsive forms aside)."""
    for key in (1, 2, 3, 6, 7, 8):
        p = key_params(key, EPS, SIG)
        assert len(p) == 5
    p4 = key_params(4, EPS, SIG)
    assert p4[2] < 1e-5 * EPS * (2 ** (1 / 6) * SIG) ** 6
