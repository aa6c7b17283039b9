"""Known answer 3: the reference's example run (CO2 in the triclinic
framework of examples/restart, 303.15 K, 0.994808 bar) reports
<N> = 91.811778 with a 5-block standard error of 4.229880
(OUTPUT:151-159).  Chains started from the run's own final configuration
(REVIVE, 90 CO2) with independent duni seeds must average within that
error, in both precisions."""
import numpy as np
import pytest

from conftest import revive_config

pytestmark = pytest.mark.gpu

N_CHAINS = 256
EQ, PROD = 60000, 240000


def _mean_loading(fmc, ex, prec):
    from fastmc_amd.batch import Batch
    field, config, _ = revive_config(ex)
    ctx = fmc.context_from_inputs(field, config, ex["control"])
    b = Batch(N_CHAINS, prec)
    for i in range(N_CHAINS):
        seed = (i % 177 + 1, (7 * i) % 177 + 1, (13 * i + 5) % 177 + 1, (11 * i) % 168)
        b.add_chain(ctx, ex["control"].temp, [0.994808e5], seed, delr=1.0,
                    eq_steps=EQ)
    step = 30000
    for _ in range((EQ + PROD) // step):
        b.run(step)
        print(".", end="", flush=True)
    sn = sp = 0.0
    per = []
    for c in range(N_CHAINS):
        st = b.stats(c)
        assert st["prod_steps"] == PROD
        sn += st["sum_n"][0]
        sp += st["prod_steps"]
        per.append(st["sum_n"][0] / st["prod_steps"])
    b.close()
    per = np.array(per)
    return sn / sp, per.std(ddof=1) / np.sqrt(len(per))


@pytest.mark.parametrize("prec", [1, 0])
def test_loading_within_reference_error(prec, restart_example):
    import fastmc_amd as fmc
    kat = restart_example["kat"]
    mean, se = _mean_loading(fmc, restart_example, prec)
    print(f"\nprecision {prec}: <N> = {mean:.3f} +/- {se:.3f} (chains), reference "
          f"{kat['loading_mean']} +/- {kat['loading_sigma_5block']}")
    assert abs(mean - kat["loading_mean"]) < kat["loading_sigma_5block"]
