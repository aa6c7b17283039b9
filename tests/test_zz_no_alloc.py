"""fmc_energy_full_noalloc: single_point with every partial sum in the
library's static device scratch instead of device memory allocated per
call.  It runs the same kernels in the same order as fmc_energy_full, so
the two agree bit for bit; two calls in a row, the second on a smaller
system, each match; the oracle's single_point pins the values.  Calls on a
context without device state, and systems whose partial sums exceed the
scratch, are refused."""
import numpy as np
import pytest

from conftest import revive_config

pytestmark = pytest.mark.gpu


def _close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1.0))


def test_noalloc_equals_energy_full_on_two_calls(oracle, restart_example):
    import fastmc_amd as fmc
    from fastmc_amd import synth
    fmc.lib()
    ex = restart_example
    # first call: the reference's restart example, 1584 framework atoms and
    # 90 CO2 (Nk 870)
    field, config, e_rev = revive_config(ex)
    big = fmc.context_from_inputs(field, config, ex["control"])
    e_alloc = big.energy_full()
    e_none = big.energy_full_noalloc()
    print(f"\nrestart example: {e_none}")
    assert np.array_equal(e_alloc, e_none)
    sp = (e_none[1] + e_none[2] + e_none[3] + e_none[4] + e_none[5]) / field.engunit
    assert abs(sp - e_rev) < 1e-7     # known answer 2
    # second call, smaller: the 424-atom cubic framework with three CO2
    A = synth.c1().system_arrays()
    small = fmc.Context.from_arrays(A)
    orc = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                              A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    g = A["guests"][0]
    for shift in (3.1, 9.7, 17.3):
        xyz = [np.asarray(g[k]) + shift + 0.37 * i for i, k in enumerate("xyz")]
        small.trial(1, 0, -1, *xyz)
        small.commit()
        orc.trial(1, 0, -1, *xyz)
        orc.commit()
    e_none2 = small.energy_full_noalloc()
    e_alloc2 = small.energy_full()
    print(f"c1 with 3 CO2: {e_none2}")
    assert np.array_equal(e_alloc2, e_none2)
    assert _close(e_none2, orc.energy_full(), 1e-10)
    # and the first context again, after the smaller call left its sums in
    # the scratch
    assert np.array_equal(big.energy_full_noalloc(), e_alloc)


def test_noalloc_refuses_without_device_state_and_beyond_the_scratch():
    import fastmc_amd as fmc
    from fastmc_amd import synth
    from fastmc_amd.capi import FmcError
    fmc.lib()
    fresh = fmc.Context.from_arrays(synth.c1().system_arrays())
    with pytest.raises(FmcError) as e:
        fresh.energy_full_noalloc()
    assert e.value.code == -2
    # 27136 atoms and kmax 16: 27 chunks x Nk recip partial sums exceed 4 MiB
    base = synth.c1()
    w = synth.Workload("c1 4x4x4", base.fw.replicate(4, 4, 4), ["co2"], [0.15], 298.0,
                       (16, 16, 16))
    huge = fmc.Context.from_arrays(w.system_arrays())
    assert huge.info()["nk"] * 27 * 16 > 4 << 20
    huge.recip_full(0)
    with pytest.raises(FmcError) as e:
        huge.energy_full_noalloc()
    assert e.value.code == -4
