"""Known answer 3: the reference's example run (CO2 in the triclinic
framework of examples/restart, 303.15 K, 0.994808 bar) reports, over
226,407,200 moves (OUTPUT:138-161):
  <N> = 91.811778 and <E> = -756.809150 kcal/mol, each with the population
  standard deviation of its 5 block means (4.229880 and 38.860);
  insertion ratio 0.000690166 and deletion ratio 0.000688845 (accepted /
  attempted over the whole run, gcmc.f:2234-2240).
Chains started from the run's own final configuration (REVIVE, 90 CO2)
with independent duni seeds run the same move loop.  Every chain begins in
the same state and makes only a few hundred accepted insertions, so the
spread of the chain means measures the noise of this sample, not an
equilibrium error; the bars are the reference's own block deviations:
  * <N> and <E> within the reference's block deviation;
  * the insertion and deletion ratios within sigma_N / <N> of the
    reference's (a ratio scales as 1/N at fixed chemical potential, so the
    reference's loading uncertainty is its uncertainty) plus three spreads
    of the chain ratios.
Both precisions."""
import numpy as np
import pytest

from conftest import revive_config

pytestmark = pytest.mark.gpu

N_CHAINS = 256
EQ, PROD = 100000, 900000


def _run(fmc, ex, prec):
    from fastmc_amd.batch import Batch
    field, config, _ = revive_config(ex)
    ctx = fmc.context_from_inputs(field, config, ex["control"])
    b = Batch(N_CHAINS, prec)
    for i in range(N_CHAINS):
        seed = (i % 177 + 1, (7 * i) % 177 + 1, (13 * i + 5) % 177 + 1, (11 * i) % 168)
        b.add_chain(ctx, ex["control"].temp, [0.994808e5], seed, delr=1.0,
                    eq_steps=EQ)
    step = 50000
    for _ in range((EQ + PROD) // step):
        b.run(step)
        print(".", end="", flush=True)
    st = [b.stats(c) for c in range(N_CHAINS)]
    b.close()
    for s in st:
        assert s["prod_steps"] == PROD
    return st


@pytest.mark.parametrize("prec", [1, 0])
def test_loading_energy_and_acceptance_match_reference(prec, restart_example):
    import fastmc_amd as fmc
    kat = restart_example["kat"]
    st = _run(fmc, restart_example, prec)
    prod = sum(s["prod_steps"] for s in st)
    n_mean = sum(s["sum_n"][0] for s in st) / prod
    e_mean = sum(s["sum_e"][0] for s in st) / prod
    n_chain = np.array([s["sum_n"][0] / s["prod_steps"] for s in st])
    out = {}
    for slot, name in ((1, "insertion"), (2, "deletion"), (3, "displacement")):
        acc = np.array([s["accepted"][slot] for s in st], float)
        att = np.array([s["attempted"][slot] for s in st], float)
        r = acc.sum() / att.sum()
        spread = (acc / att).std(ddof=1) / np.sqrt(len(st))
        out[name] = (r, spread, acc.sum())
    print(f"\nprecision {prec}: <N> = {n_mean:.3f} (spread of chain means {n_chain.std(ddof=1) / 16:.3f}),"
          f" reference {kat['loading_mean']} +/- {kat['loading_sigma_5block']}; <E> = {e_mean:.3f},"
          f" reference {kat['energy_mean']} +/- {kat['energy_sigma_5block']}")
    for name, (r, sp, na) in out.items():
        ref = kat.get(f"{name}_ratio")
        print(f"  {name} ratio {r:.9f} +/- {sp:.9f} ({int(na)} accepted), reference {ref}")
    assert abs(n_mean - kat["loading_mean"]) < kat["loading_sigma_5block"]
    assert abs(e_mean - kat["energy_mean"]) < kat["energy_sigma_5block"]
    rel_n = kat["loading_sigma_5block"] / kat["loading_mean"]
    for name in ("insertion", "deletion"):
        r, sp, _ = out[name]
        ref = kat[f"{name}_ratio"]
        assert abs(r - ref) < rel_n * ref + 3 * sp, name
