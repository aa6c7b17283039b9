"""Host side of the product (no GPU): the tables, k-vector count, ewald3
exclusion constants and tail corrections that libfastmc_amd builds in
fmc_ctx_finalize, against the oracle on the same inputs, and the errors
the set-up reports.  Device state is built only by the first single-chain
call, so all of this runs on a machine without a GPU."""
import numpy as np
import pytest


def _pair(oracle, A):
    import fastmc_amd
    ctx = fastmc_amd.Context.from_arrays(A)
    s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                            A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    return ctx, s


def _systems():
    from fastmc_amd import synth
    return {"c1": synth.c1().system_arrays(), "c2": synth.c2().system_arrays(),
            "c5": synth.c5().system_arrays()}


@pytest.mark.parametrize("name", ["c1", "c2", "c5"])
def test_tables_and_constants_match_oracle(oracle, name):
    A = _systems()[name]
    ctx, s = _pair(oracle, A)
    pi = ctx.info()
    mg, drewd, dlrpot = s.grid_info()
    assert pi["nk"] == s.nk
    assert pi["mxgrid"] == mg
    assert pi["drewd"] == pytest.approx(drewd, rel=1e-15)
    assert pi["dlrpot"] == pytest.approx(dlrpot, rel=1e-15)
    assert pi["volm"] == pytest.approx(s.volume, rel=1e-15)
    e_p, e_o = ctx.table(0), s.table(0)
    assert np.allclose(e_p, e_o, rtol=1e-14, atol=0.0)
    for t in range(len(A["vdw"]) + 1):
        v_p, v_o = ctx.table(1, t), s.table(1, t)
        assert np.allclose(v_p, v_o, rtol=1e-13, atol=1e-300), t
    for g in range(len(A["guests"])):
        assert ctx.ewald3en(g) == pytest.approx(s.ewald3en(g), rel=1e-13)
        for sign in (1, -1):
            assert ctx.tail_delta(g, sign, 0.25) == pytest.approx(
                s.tail_delta(g, sign, 0.25), rel=1e-13, abs=1e-12)


def test_restart_example_host_side(oracle, restart_example):
    """The reference example's inputs through the product's readers: alpha
    and kmax from 'ewald precision 1d-6' and Nk as the reference ran them."""
    import fastmc_amd
    ex = restart_example
    ctx = fastmc_amd.context_from_inputs(ex["field"], ex["config"], ex["control"])
    assert ctx.info()["nk"] == ex["kat"]["nk"]
    s, _, _ = oracle.system_from_inputs(ex["field"], ex["config"], ex["control"])
    assert ctx.info()["nk"] == s.nk


def test_setup_errors_without_gpu(oracle):
    import fastmc_amd
    from fastmc_amd import synth
    from fastmc_amd.capi import FmcError
    A = synth.c1().system_arrays()
    ctx = fastmc_amd.Context.from_arrays(A)
    # the first device call reports the missing device, loudly
    with pytest.raises(FmcError) as ei:
        ctx.trial(1, 0, -1, [0.0] * 3, [0.0] * 3, [1.16, 0.0, -1.16])
    assert ei.value.code in (-6, -5)
    # a cutoff beyond half the cell width is the reference's error 95
    B = dict(A)
    B["rcut"] = 14.0
    with pytest.raises(FmcError) as ei:
        fastmc_amd.Context.from_arrays(B)
    assert ei.value.code == -7
    # parallelepiped and slab conventions outside 1-3 are not on this path
    B = dict(A)
    B["imcon"] = 4
    with pytest.raises(FmcError) as ei:
        fastmc_amd.Context.from_arrays(B)
    assert ei.value.code == -3
