"""Large-cell stress case: 3x3x3 supercell of the 424-atom generator cell
(11448 atoms, 77.4 A, kmax 16, Nk ~10^4: S(k) in HBM, the streaming form),
ternary CO2 / CH4 (5 sites) / H2O (4 sites, massless M site) at 10 bar
total and 298 K.  FP64 on the GPU reproduces the oracle's trajectory; the
mixed-precision run is compared with it on the same states (estep error),
and both report their energy drift against an FP64 full recompute."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_c5_fp64_parity_and_mixed_error(oracle):
    import fastmc_amd as fmc
    from fastmc_amd import synth
    from fastmc_amd.batch import Batch
    w = synth.c5()
    A = w.system_arrays()
    press = [p * 1e5 for p in w.press_bar]
    n = 600
    s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                            A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    ch = oracle.OracleChain(s, w.temp, press, w.seed, delr=A["delr"])
    recs = ch.run(n)
    ref = np.array([[r.kind, r.iguest, r.accepted, r.nmol_after, r.estep, r.e1, r.e2,
                     r.vdw] for r in recs])
    ctx = fmc.Context.from_arrays(A)
    assert ctx.info()["nk"] > 9000
    logs, stats = {}, {}
    for prec in (0, 1):
        b = Batch(1, prec)
        cid = b.add_chain(ctx, w.temp, press, w.seed, delr=A["delr"])
        b.enable_log(cid, n)
        b.run(n)
        assert b.tile() & 0x10000          # streaming form
        logs[prec] = b.get_log(cid, n)
        skdev, drift = b.recompute()
        stats[prec] = b.stats(cid)
        b.close()
    # FP64: the oracle's trajectory
    a = logs[0]
    assert np.array_equal(a[:, :4], ref[:, :4])
    # each component within 1e-10 of max(its magnitude, 100 internal units,
    # i.e. an absolute 1e-8 = 2.4e-11 kcal/mol): at kmax 16 the reciprocal
    # sum's ~10^4 terms reach ~5e4 internal units and cancel, so FP64
    # rounding in a different summation order alone moves it by ~1e-9;
    # estep within 1e-10 of the summed component magnitudes
    comp = np.abs(a[:, 5:8] - ref[:, 5:8]) / np.maximum(np.abs(ref[:, 5:8]), 100.0)
    scale = np.abs(ref[:, 5:8]).sum(axis=1) / 418.4
    rel = np.abs(a[:, 4] - ref[:, 4]) / np.maximum(np.maximum(np.abs(ref[:, 4]), scale), 1e-3)
    i = int(np.argmax(rel))
    print(f"\nC5 FP64: worst component rel {comp.max():.2e}, worst estep rel {rel.max():.2e} "
          f"(move {i}: estep {ref[i, 4]:.6e}, components {ref[i, 5:8]})")
    assert comp.max() < 1e-10 and rel.max() < 1e-10
    assert abs(stats[0]["drift"]) < 1e-8 * max(1.0, abs(stats[0]["e_full"]))
    # mixed vs FP64 on the common prefix
    m = logs[1]
    same = np.all(m[:, :4] == a[:, :4], axis=1)
    k = int(np.argmin(same)) if not same.all() else n
    kT = 1.9872041e-3 * w.temp
    sel = np.abs(a[:k, 4]) < 50 * kT
    err = np.abs(m[:k, 4] - a[:k, 4])[sel]
    print(f"\nC5: {int(ref[:, 2].sum())} accepted; mixed identical for {k} moves; "
          f"estep err max {err.max():.2e} kcal/mol ({err.max() / kT:.2e} kT), median "
          f"{np.median(err):.2e}; drift FP64 {stats[0]['drift']:.2e}, mixed "
          f"{stats[1]['drift']:.2e} kcal/mol; loadings FP64 {stats[0]['nmol']} "
          f"mixed {stats[1]['nmol']}")
    assert k > 100
    assert err.max() < 1e-3 * kT
    assert abs(stats[1]["drift"]) < 1e-2


def test_c5_loading_fp32_bulk_vs_fp64():
    """Loading deltas: 64 chains per precision, 12000 moves from the empty
    framework; per-species mean loadings over the second half agree within
    four combined standard errors (chains are independent samples)."""
    import fastmc_amd as fmc
    from fastmc_amd import synth
    from fastmc_amd.batch import Batch
    w = synth.c5()
    A = w.system_arrays()
    press = [p * 1e5 for p in w.press_bar]
    ctx = fmc.Context.from_arrays(A)
    nch, n = 64, 12000
    res = {}
    for prec in (0, 1):
        b = Batch(nch, prec)
        for i in range(nch):
            b.add_chain(ctx, w.temp, press, (i % 177 + 1, 17, 29, (5 * i) % 168),
                        delr=A["delr"], eq_steps=n // 2)
        for _ in range(4):
            b.run(n // 4)
        _, drift = b.recompute()
        per = np.array([[b.stats(c)["sum_n"][g] / b.stats(c)["prod_steps"]
                         for g in range(3)] for c in range(nch)])
        res[prec] = (per.mean(0), per.std(0, ddof=1) / np.sqrt(nch), drift)
        b.close()
    (m0, e0, d0), (m1, e1, d1) = res[0], res[1]
    print(f"\nC5 loadings (CO2, CH4, H2O): FP64 {np.round(m0, 3)} +/- {np.round(e0, 3)}; "
          f"mixed {np.round(m1, 3)} +/- {np.round(e1, 3)}; max |drift| FP64 {d0:.2e}, "
          f"mixed {d1:.2e} kcal/mol")
    assert np.all(np.abs(m0 - m1) <= 4 * np.sqrt(e0 ** 2 + e1 ** 2) + 1e-9)
