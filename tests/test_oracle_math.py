"""Pins the CPU oracle against what the mathematics fixes: closed forms,
brute force on small inputs and invariants.  Each check is chosen so that a
plausible slip (wrong table offset, dropped term, sign, transposed cell
index) fails it."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, special

from fastmc_amd import synth

R4PIE0 = 138935.4835


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


@pytest.fixture(scope="module")
def small(oracle):
    w = synth.c1(seed=7)
    A = w.system_arrays()
    s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"],
                            A["rcut"], A["delr"], A["ntpatm"], A["vdw"],
                            A["fw"], A["guests"])
    return w, A, s


def test_erfc_table_matches_erfc(oracle):
    L = oracle.lib()
    alpha, rcut = 0.26, 12.5
    mg = L.orc_mxgrid(rcut)
    assert mg == 1254  # utility_pack.f:725 with rvdw 12.5
    erc = np.zeros(mg + 1)
    drewd = L.orc_erfcgen(alpha, rcut, mg, _dp(erc))
    assert drewd == rcut / (mg - 4)
    r = np.arange(1, mg + 1) * drewd
    # A&S 7.1.26: |erfc error| <= 1.5e-7; table entry i is at r = i*drewd
    err = np.abs(erc[1:] * r - special.erfc(alpha * r))
    assert err.max() < 1.5e-7
    assert erc[0] == erc[1]


def test_lj_table_closed_form(oracle):
    L = oracle.lib()
    mg = 1254
    eps, sig = 0.2 * 418.4, 3.1
    vvv = np.zeros(2 * (mg + 1))
    keys = np.array([2], dtype=np.int32)
    prm = np.array([eps, sig, 0, 0, 0], float)
    dl = L.orc_forgen(1, _ip(keys), _dp(prm), 12.5, mg, _dp(vvv))
    v = vvv[:mg + 1]
    i0 = int(round(sig / dl))
    assert abs(i0 * dl - sig) < 1e-12 and abs(v[i0]) < 1e-9 * eps
    imin = int(round(2 ** (1 / 6) * sig / dl))
    assert abs(v[imin] + eps) < 1e-3 * eps
    assert np.all(vvv[mg + 1:] == 0.0)  # undefined-pair zero table


def test_interpolation_exact_on_quadratic(oracle):
    """The 3-point scheme of ewald2/srfrce reproduces any quadratic in r
    exactly, so a shifted index or wrong p would show."""
    L = oracle.lib()
    mg, drewd, rcut = 1254, 0.01, 12.5
    grid = np.arange(mg + 1) * drewd
    f = lambda r: 3.0 - 0.7 * r + 0.05 * r * r
    tab = f(grid)
    rng = np.random.default_rng(1)
    r = rng.uniform(0.02, 12.4, size=400)
    rsq = r * r
    chg = np.array([1.0, 0.5])
    il = np.ones(400, dtype=np.int32)
    got = L.orc_ewald2(2.0, 400, _ip(il), _dp(rsq), _dp(chg), _dp(tab), drewd,
                       rcut, 1.0)
    want = (f(r) * 2.0 * R4PIE0 * 0.5).sum()
    assert abs(got - want) < 1e-9 * abs(want)
    # srfrce on the same quadratic table for a single pair type
    ltype = np.array([1, 1], dtype=np.int32)
    lst = np.array([1], dtype=np.int32)
    ltp = np.array([2, -1], dtype=np.int32)
    got2 = L.orc_srfrce(1, 400, _ip(il), _dp(rsq), _ip(ltype), _ip(lst),
                        _ip(ltp), _dp(tab), mg, rcut, drewd)
    assert abs(got2 - f(r).sum()) < 1e-9 * abs(f(r).sum())
    # cut is strict r^2 < rcut^2
    one = np.array([rcut * rcut]); il1 = np.array([1], dtype=np.int32)
    assert L.orc_ewald2(1.0, 1, _ip(il1), _dp(one), _dp(chg), _dp(tab),
                        drewd, rcut, 1.0) == 0.0


def test_images_minimum_image(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    cell = np.array([20.0, 0, 0, 0, 23.0, 0, 0, 0, 27.0])
    d = rng.uniform(-60, 60, size=(500, 3))
    x, y, z = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    L.orc_images(2, 500, _dp(cell), _dp(x), _dp(y), _dp(z))
    for k, L_ in enumerate((20.0, 23.0, 27.0)):
        v = (x, y, z)[k]
        assert np.all(np.abs(v) <= L_ / 2 + 1e-12)
        n = (d[:, k] - v) / L_
        assert np.allclose(n, np.round(n), atol=1e-9)
    # triclinic: result is a lattice translate with fractional coords in
    # [-1/2, 1/2]
    cell3 = np.array([20.0, 0, 0, 3.0, 22.0, 0, -4.0, 2.0, 25.0])
    x, y, z = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    L.orc_images(3, 500, _dp(cell3), _dp(x), _dp(y), _dp(z))
    H = cell3.reshape(3, 3).T  # columns a, b, c
    fr_in = np.linalg.solve(H, d.T)
    fr_out = np.linalg.solve(H, np.vstack([x, y, z]))
    assert np.all(np.abs(fr_out) <= 0.5 + 1e-12)
    assert np.allclose(fr_in - fr_out, np.round(fr_in - fr_out), atol=1e-9)


def _ref_widths(H):
    """ewald.f:79-84 takes dcell of rcell, i.e. the perpendicular widths of
    the parallelepiped spanned by the COLUMNS of H^-1 (for non-orthogonal
    cells this differs from the reciprocal cell, whose vectors are rows)."""
    B = np.linalg.inv(H)
    a, b, c = B[:, 0], B[:, 1], B[:, 2]
    v = abs(a @ np.cross(b, c))
    return (v / np.linalg.norm(np.cross(b, c)), v / np.linalg.norm(np.cross(c, a)),
            v / np.linalg.norm(np.cross(a, b)))


def _brute_recip(cell, kmax, alpha, x, y, z, q):
    """Direct Ewald reciprocal sum over the full k sphere (both halves) with
    explicit exp(i k.r), no recurrences."""
    H = cell.reshape(3, 3).T
    G = 2 * math.pi * np.linalg.inv(H)  # rows are reciprocal vectors
    widths = _ref_widths(H)
    rcp = min(kmax[0] * widths[0], kmax[1] * widths[1], kmax[2] * widths[2])
    rcp = rcp * 1.05 * 2 * math.pi
    e = 0.0
    pos = np.vstack([x, y, z])
    for l in range(-kmax[0], kmax[0] + 1):
        for m in range(-kmax[1], kmax[1] + 1):
            for n in range(-kmax[2], kmax[2] + 1):
                if l == 0 and m == 0 and n == 0:
                    continue
                k = l * G[0] + m * G[1] + n * G[2]
                k2 = k @ k
                if k2 > rcp * rcp:
                    continue
                s = (q * np.exp(1j * (k @ pos))).sum()
                e += math.exp(-k2 / (4 * alpha * alpha)) / k2 * abs(s) ** 2
    vol = abs(np.linalg.det(H))
    return (2 * math.pi / vol) * R4PIE0 * e - R4PIE0 * alpha / math.sqrt(math.pi) * (q * q).sum()


@pytest.mark.parametrize("imcon,cell", [
    (1, [14.0, 0, 0, 0, 14.0, 0, 0, 0, 14.0]),
    (3, [15.0, 0, 0, 2.5, 14.0, 0, -3.0, 1.5, 16.0]),
])
def test_ewald1_vs_brute_force(oracle, imcon, cell):
    L = oracle.lib()
    cell = np.array(cell, float)
    rng = np.random.default_rng(3)
    n = 17
    fr = rng.uniform(0, 1, size=(n, 3))
    H = cell.reshape(3, 3).T
    pos = (H @ fr.T)
    x, y, z = pos[0].copy(), pos[1].copy(), pos[2].copy()
    q = rng.normal(0, 0.5, n)
    q -= q.mean()
    kmax = np.array([5, 4, 6], dtype=np.int32)
    alpha = 0.3
    nk = L.orc_kvec_count(_dp(cell), _ip(kmax))
    ckc, cks = np.zeros(nk), np.zeros(nk)
    newld = C.c_int()
    e = L.orc_ewald1(imcon, _dp(cell), alpha, _ip(kmax), 1.0, n, _dp(x),
                     _dp(y), _dp(z), _dp(q), _dp(ckc), _dp(cks),
                     C.byref(newld), None)
    assert newld.value == nk
    want = _brute_recip(cell, kmax, alpha, x, y, z, q)
    assert abs(e - want) < 1e-10 * abs(want)


def test_kvec_count_brute_force(oracle):
    L = oracle.lib()
    for cell, km in [([25.8, 0, 0, 0, 25.8, 0, 0, 0, 25.8], (8, 8, 8)),
                     ([51.6, 0, 0, 0, 51.6, 0, 0, 0, 51.6], (12, 12, 12)),
                     ([35.6552, 0, 0, 0, 29.0805, 0, -12.425126850515, 0,
                       25.59422285182], (9, 8, 8)),
                     ([15.0, 0, 0, 2.5, 14.0, 0, -3.0, 1.5, 16.0], (5, 4, 6))]:
        cell = np.array(cell, float)
        kmax = np.array(km, dtype=np.int32)
        H = cell.reshape(3, 3).T
        G = 2 * math.pi * np.linalg.inv(H)
        widths = _ref_widths(H)
        rcp = min(km[0] * widths[0], km[1] * widths[1], km[2] * widths[2])
        rcp2 = (rcp * 1.05 * 2 * math.pi) ** 2
        cnt = 0
        for l in range(0, km[0] + 1):
            for m in range(-km[1], km[1] + 1):
                for n in range(-km[2], km[2] + 1):
                    # half space: l>0, or l==0 and (m>0 or m==0 and n>0)
                    if l == 0 and (m < 0 or (m == 0 and n <= 0)):
                        continue
                    k = l * G[0] + m * G[1] + n * G[2]
                    if k @ k <= rcp2:
                        cnt += 1
        assert L.orc_kvec_count(_dp(cell), _ip(kmax)) == cnt
    # the k-vector counts of the C1 (25.8 A, kmax 8) and C2 (51.6 A, kmax 12)
    # workloads
    assert L.orc_kvec_count(_dp(np.array([25.8, 0, 0, 0, 25.8, 0, 0, 0, 25.8])),
                            _ip(np.array([8, 8, 8], dtype=np.int32))) == 1258
    assert L.orc_kvec_count(_dp(np.array([51.6, 0, 0, 0, 51.6, 0, 0, 0, 51.6])),
                            _ip(np.array([12, 12, 12], dtype=np.int32))) == 4192


def test_guest_update_equals_full_difference(oracle, small):
    """ewald1_guest insert/delete/displace == difference of two full
    ewald1 sums (R1 vs R2), and insert-then-delete restores S(k)."""
    w, A, s = small
    L = oracle.lib()
    g = A["guests"][0]
    rng = np.random.default_rng(4)
    cell = np.asarray(A["cell"], float)
    fw = A["fw"]
    kmax = np.array(A["kmax"], dtype=np.int32)
    e_fw = s.recip_full(1)
    x0, y0, z0 = [np.asarray(g[k]) + rng.uniform(0, 25.8) for k in "xyz"]
    e1 = s.trial(1, 0, -1, x0, y0, z0)[0]
    X = np.concatenate([x0, fw["x"]]); Y = np.concatenate([y0, fw["y"]])
    Z = np.concatenate([z0, fw["z"]]); Q = np.concatenate([g["q"], fw["q"]])
    nk = s.nk
    ckc, cks = np.zeros(nk), np.zeros(nk)
    e_all = L.orc_ewald1(A["imcon"], _dp(cell), A["alpha"], _ip(kmax), 1.0,
                         len(X), _dp(X), _dp(Y), _dp(Z), _dp(Q), _dp(ckc),
                         _dp(cks), None, None)
    assert abs(e1 - (e_all - e_fw)) < 1e-9 * abs(e_fw)
    s.commit()
    re, im = s.sk()
    assert np.allclose(re, ckc, atol=1e-9) and np.allclose(im, cks, atol=1e-9)
    # displace: new position vs full recompute
    x1, y1, z1 = x0 + 1.3, y0 - 0.4, z0 + 2.2
    e_d = s.trial(3, 0, 0, x1, y1, z1)[0]
    X2 = np.concatenate([x1, fw["x"]]); Y2 = np.concatenate([y1, fw["y"]])
    Z2 = np.concatenate([z1, fw["z"]])
    e_all2 = L.orc_ewald1(A["imcon"], _dp(cell), A["alpha"], _ip(kmax), 1.0,
                          len(X), _dp(X2), _dp(Y2), _dp(Z2), _dp(Q), _dp(ckc),
                          _dp(cks), None, None)
    assert abs(e_d - (e_all2 - e_all)) < 1e-9 * abs(e_fw)
    # delete restores the framework structure factor
    e_del = s.trial(2, 0, 0)[0]
    assert abs(e_del - (e_fw - e_all)) < 1e-9 * abs(e_fw)
    s.commit()
    s.recip_full(1)


def test_tail_correction_closed_form(oracle):
    """lrcorrect eadd equals the integral 2 pi/V N_i N_j int_rc^inf r^2 V(r)
    dr for LJ."""
    L = oracle.lib()
    eps, sig, rc, vol = 0.15 * 418.4, 3.3, 12.5, 20000.0
    f = lambda r: r * r * 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
    integral = integrate.quad(f, rc, np.inf, epsabs=1e-14, epsrel=1e-12)[0]
    numtyp = np.array([7, 300], dtype=np.int32)
    numfrz = np.array([0, 300], dtype=np.int32)
    lst = np.array([1, 1, 1], dtype=np.int32)  # all pairs use potential 1
    ltp = np.array([2, -1], dtype=np.int32)
    prm = np.array([eps, sig, 0, 0, 0, 0, 0, 0, 0, 0], float)
    st = np.zeros(3)
    elrc = L.orc_lrcorrect(2, 2, _ip(numtyp), _ip(numfrz), _ip(lst), _ip(ltp),
                           _dp(prm), rc, vol, _dp(st))
    want = 2 * math.pi / vol * (7 * 7 + 2 * 7 * 300 + (300 * 300 - 300 * 300)) * integral
    assert abs(elrc - want) < 1e-9 * abs(want)
    again = L.orc_gstlrcorrect(2, 2, _ip(numtyp), _ip(numfrz), _dp(st), vol)
    assert abs(again - elrc) < 1e-12 * abs(elrc)


def test_ewald3_closed_form(oracle):
    L = oracle.lib()
    alpha = 0.26
    for r in (1.16, 2.32, 0.03):  # 0.03 uses the small-r series branch
        dx = np.array([r]); dy = np.zeros(1); dz = np.zeros(1)
        q2 = np.array([-0.35])
        e = L.orc_ewald3(0.7, 1, _dp(q2), _dp(dx), _dp(dy), _dp(dz), alpha, 1.0)
        want = -0.7 * -0.35 * R4PIE0 * special.erf(alpha * r) / r
        assert abs(e - want) < 2e-7 * R4PIE0 * 0.7 * 0.35 / r


def test_guestlistgen_brute_force(oracle, small):
    w, A, s = small
    L = oracle.lib()
    fw = A["fw"]
    n = len(fw["x"])
    cell = np.asarray(A["cell"], float)
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 25.8, size=(3, 3))
    nx, ny, nz = p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()
    ind = np.array([-1, -1, -1], dtype=np.int32)
    ent = np.zeros(3, dtype=np.int32)
    lst = np.zeros(3 * n, dtype=np.int32)
    X, Y, Z = [np.ascontiguousarray(fw[k], dtype=float) for k in "xyz"]
    L.orc_guestlistgen(A["imcon"], _dp(cell), 12.5, 1.0, 3, _dp(nx), _dp(ny),
                       _dp(nz), _ip(ind), n, _dp(X), _dp(Y), _dp(Z), _ip(ent),
                       _ip(lst))
    for i in range(3):
        d = np.vstack([nx[i] - X, ny[i] - Y, nz[i] - Z])
        d -= 25.8 * np.round(d / 25.8)
        want = set(np.nonzero((d * d).sum(0) <= 13.5 ** 2)[0])
        got = set(lst[i * n:i * n + ent[i]])
        assert got == want


def test_energy_eval_closed_forms(oracle):
    L = oracle.lib()
    beta = 1 / (1.9872041e-3 * 298.0)
    vol, p, T = 17173.512e-30, 1.5e4, 298.0
    kb = 1.3806503e-23
    e = 0.7
    for rande in (0.01, 0.2, 0.5, 0.9):
        pins = vol * p / (4 * kb * T) * math.exp(-beta * e)
        assert L.orc_energy_eval(e, rande, vol, p, 3, T, beta, 1) == (rande < pins)
        pdel = 3 * kb * T / (vol * p) * math.exp(-beta * e)
        assert L.orc_energy_eval(e, rande, vol, p, 3, T, beta, 2) == (rande < pdel)
        assert L.orc_energy_eval(e, rande, vol, p, 3, T, beta, 3) == (rande < math.exp(-beta * e))
    assert L.orc_energy_eval(0.5, 0.999, vol, 0.0, 3, T, beta, 2) == 1  # P<1e-7
    assert L.orc_energy_eval(-1000.0, 0.999, vol, p, 3, T, beta, 1) == 1  # e^700 cap


def test_proposals_rigid_and_in_cell(oracle, small):
    w, A, s = small
    L = oracle.lib()
    g = A["guests"][0]
    cell = np.asarray(A["cell"], float)
    rcell = np.zeros(9); det = C.c_double()
    L.orc_invert(_dp(cell), _dp(rcell), C.byref(det))
    gx, gy, gz = [np.ascontiguousarray(g[k], float) for k in "xyz"]
    ns = len(gx)
    nx, ny, nz = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    rng = np.random.default_rng(6)
    d0 = np.linalg.norm(np.vstack([gx - gx[0], gy - gy[0], gz - gz[0]]), axis=0)
    for _ in range(50):
        r = rng.uniform(0, 1, 7)
        L.orc_random_ins(_dp(cell), ns, _dp(gx), _dp(gy), _dp(gz), *r,
                         _dp(nx), _dp(ny), _dp(nz))
        d = np.linalg.norm(np.vstack([nx - nx[0], ny - ny[0], nz - nz[0]]), axis=0)
        assert np.allclose(d, d0, atol=1e-12)
        m = np.asarray(g["mass"])
        com = np.array([(nx * m).sum(), (ny * m).sum(), (nz * m).sum()]) / m.sum()
        assert np.allclose(com, r[:3] * 25.8, atol=1e-10)
        r2 = rng.uniform(0, 1, 7)
        L.orc_random_disp(_dp(cell), _dp(rcell), ns, _dp(m), _dp(nx), _dp(ny),
                          _dp(nz), 1.0, math.pi / 3, *r2)
        d = np.linalg.norm(np.vstack([nx - nx[0], ny - ny[0], nz - nz[0]]), axis=0)
        assert np.allclose(d, d0, atol=1e-12)
        com2 = np.array([(nx * m).sum(), (ny * m).sum(), (nz * m).sum()]) / m.sum()
        assert np.all(com2 >= -1e-12) and np.all(com2 < 25.8 + 1e-12)


def test_chain_energy_bookkeeping(oracle):
    """The chain's running sum of accepted estep equals a from-scratch
    single_point energy of the final configuration (signs of the delete /
    displace assembly, ewald3 and tail bookkeeping)."""
    import numpy as np_
    rng = np_.random.default_rng(11)
    fw = synth.make_framework(rng, (25.8, 25.8, 25.8), 200)
    w = synth.Workload("porous test", fw, ["co2"], [10.0], 298.0, (8, 8, 8))
    A = w.system_arrays()
    s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"],
                            A["rcut"], A["delr"], A["ntpatm"], A["vdw"],
                            A["fw"], A["guests"])
    ch = oracle.OracleChain(s, 298.0, [10.0e5], (12, 34, 56, 78))
    recs = ch.run(1500)
    assert s.nmol(0) >= 3
    kinds = {(r.kind, r.accepted) for r in recs}
    assert (1, 1) in kinds and (3, 1) in kinds
    e = s.energy_full()
    sp = (e[1] + e[2] + e[3] + e[4] + e[5]) / 418.4
    assert abs(ch.energy(0) - sp) < 1e-8 * max(1.0, abs(sp))


def test_reciprocal_change_rounding_on_the_large_cell(oracle):
    """How far FP64 can pin a displacement's reciprocal energy change on the
    3392-atom 51.6 A cell at kmax 12 (Nk 4192).  The exact change, summed
    in 64-bit-mantissa extended precision, is compared with (a) the
    oracle, which follows the reference's term-by-term form a_k (|S_new|^2 -
    |S_old|^2) (ewald.f:311-319), and (b) the cancellation-free form a_k (2
    S.dS + |dS|^2) in FP64 that the HIP kernels use.  The reference's form
    cancels |S|^2 terms of the framework's structure factor, so its
    rounding error reaches ~1e-11 of the component; the cancellation-free
    form stays near 1e-14.  A GPU-vs-oracle comparison on this cell is
    therefore bounded by the oracle's own rounding, relative to the
    components rather than to a small estep."""
    from fastmc_amd import synth
    w = synth.c2()
    A = w.system_arrays()
    s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                            A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    L = float(A["cell"][0])
    km = 12
    # the half space and spherical cut of ewald.f:81-84,141-213 on a cubic cell
    lmn = [(l, m, n) for l in range(0, km + 1) for m in range(-km, km + 1)
           for n in range(-km, km + 1)
           if (l > 0 or m > 0 or (m == 0 and n > 0)) and l * l + m * m + n * n <= (1.05 * km) ** 2]
    lmn = np.array(lmn, dtype=np.longdouble)
    assert len(lmn) == s.nk == 4192
    ld = np.longdouble
    twopi = 2 * ld(np.pi)
    k2 = (twopi / ld(L)) ** 2 * (lmn ** 2).sum(axis=1)
    alpha = ld(A["alpha"])
    ak = np.exp(-k2 / (4 * alpha * alpha)) / k2

    def sfac(x, y, z, q):
        f = np.stack([np.asarray(x, ld), np.asarray(y, ld), np.asarray(z, ld)], axis=1) / ld(L)
        re = np.zeros(len(lmn), ld)
        im = np.zeros(len(lmn), ld)
        for a in range(0, len(f), 256):
            ph = twopi * (f[a:a + 256] @ lmn.T)
            qq = np.asarray(q[a:a + 256], ld)[:, None]
            re += (qq * np.cos(ph)).sum(axis=0)
            im += (qq * np.sin(ph)).sum(axis=0)
        return re, im

    fw = A["fw"]
    Sre, Sim = sfac(fw["x"], fw["y"], fw["z"], fw["q"])
    g = A["guests"][0]
    q = np.asarray(g["q"])
    factor = ld(2.0) * (twopi / ld(L) ** 3) * ld(138935.4835)
    rng = np.random.default_rng(5)
    worst_ref = worst_cf = 0.0
    for t in range(6):
        p0 = rng.uniform(2.0, L - 2.0, 3)
        old = [np.asarray(g[c]) + p0[i] for i, c in enumerate("xyz")]
        new = [o + rng.uniform(-0.3, 0.3) for o in old]
        s.trial(1, 0, -1, *old)
        s.commit()
        e_orc = s.trial(3, 0, 0, *new)[0]
        s.trial(2, 0, 0)
        s.commit()
        so_re, so_im = sfac(*old, q)
        sn_re, sn_im = sfac(*new, q)
        o_re, o_im = Sre + so_re, Sim + so_im
        n_re, n_im = o_re - so_re + sn_re, o_im - so_im + sn_im
        exact = float(factor * (ak * ((n_re ** 2 + n_im ** 2) - (o_re ** 2 + o_im ** 2))).sum())
        # the cancellation-free form in FP64
        d_re = (sn_re - so_re).astype(np.float64)
        d_im = (sn_im - so_im).astype(np.float64)
        O_re, O_im = o_re.astype(np.float64), o_im.astype(np.float64)
        cf = float(factor) * float(np.sum(ak.astype(np.float64) *
                                          (2.0 * (O_re * d_re + O_im * d_im) + d_re * d_re + d_im * d_im)))
        scale = max(abs(exact), 1.0)
        worst_ref = max(worst_ref, abs(e_orc - exact) / scale)
        worst_cf = max(worst_cf, abs(cf - exact) / scale)
        print(f"\ntrial {t}: exact {exact:.12f}, oracle - exact {e_orc - exact:.2e}, "
              f"cancellation-free - exact {cf - exact:.2e} (internal units)")
    assert worst_cf < 1e-12
    assert worst_ref < 1e-9
