"""Pins the CPU oracle against the known answers the reference's own files
hold (examples/restart/1_...: OUTPUT:113, branch01/REVIVE) and against the
published RANMAR sequence that the reference's duni generator implements."""
import numpy as np

from conftest import revive_config


def test_kat_framework_recip_plus_self(oracle, restart_example):
    ex = restart_example
    s, alpha, kmax = oracle.system_from_inputs(ex["field"], ex["config"],
                                               ex["control"])
    kat = ex["kat"]
    # initscan's "ewald precision" derivation (utility_pack.f:380-392)
    assert abs(alpha - kat["alpha"]) < 1e-7
    assert tuple(kmax) == tuple(kat["kmax"])
    assert s.nk == kat["nk"]
    e = s.recip_full(0) / ex["field"].engunit
    # printed with 6 decimals: agreement within the print rounding
    assert abs(e - kat["framework_recip_plus_self_kcal"]) <= 5e-7
    # single_point's full energy of the guest-free framework is the same
    full = s.energy_full()
    assert abs(full[0] / ex["field"].engunit - e) < 1e-9
    assert np.all(full[1:] == 0.0)


def test_kat_revive_configuration_energy(oracle, restart_example):
    """The REVIVE energy is the reference's running sum of accepted estep
    over 226M moves from an empty framework; the single_point energy of the
    REVIVE configuration must reproduce it (R1-R8 end to end)."""
    ex = restart_example
    field, config, e_revive = revive_config(ex)
    s, _, _ = oracle.system_from_inputs(field, config, ex["control"])
    assert s.nmol(0) == 90
    e = s.energy_full()
    spenergy = (e[1] + e[2] + e[3] + e[4] + e[5]) / field.engunit
    assert abs(spenergy - e_revive) < 1e-7, (spenergy, e_revive)
    # every component is non-trivial (a dropped term would show)
    assert all(abs(v) > 1.0 for v in e[1:6])


def test_duni_ranmar_known_answer(oracle):
    """Marsaglia-Zaman-Tsang universal generator test: seeds ij=1802,
    kl=9373; after 20000 draws the next six values times 2^24 are
    published integers."""
    ij, kl = 1802, 9373
    i = (ij // 177) % 177 + 2
    j = ij % 177 + 2
    k = (kl // 169) % 178 + 1
    l = kl % 169
    r = oracle.Duni(i, j, k, l)
    for _ in range(20000):
        r()
    got = [r() * 4096.0 * 4096.0 for _ in range(6)]
    assert got == [6533892.0, 14220222.0, 7275067.0, 6172232.0, 8354498.0,
                   10633180.0]


def test_duni_range_and_clock_seed(oracle):
    import ctypes as C
    out = (C.c_int * 4)()
    oracle.lib().orc_duni_seed_from_clock(59, 999, out)
    i, j, k, l = out[0], out[1], out[2], out[3]
    assert 1 <= i <= 177 and 1 <= j <= 177 and 1 <= k <= 177 and 0 <= l <= 167
    r = oracle.Duni(i, j, k, l)
    v = np.array([r() for _ in range(5000)])
    assert v.min() >= 0.0 and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.02
    # values are multiples of 2^-24 (float32 arithmetic, gcmc.f:2719)
    assert np.all(np.mod(v * 16777216.0, 1.0) == 0.0)
