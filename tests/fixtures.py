"""Loaders for the committed reference fixtures (tests/golden/)."""
import functools
import importlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@functools.lru_cache(maxsize=None)
def mini():
    """(arrays, Population, params, hz_params) of the 20k-person fixture."""
    csa = importlib.import_module("covid-sim_amd")
    z = np.load(os.path.join(GOLDEN, "mini_fixture.npz"))
    arrays = {k: z[k] for k in z.files}
    popa = {k: v for k, v in arrays.items()
            if not k.startswith(("s0_", "s160_", "hz_", "g_"))}
    pop = csa.population.Population(popa)
    params = csa.population.load_params(os.path.join(GOLDEN, "mini_params.bin"))
    hzp = csa.population.load_params(os.path.join(GOLDEN, "mini_hz_params.bin"))
    return arrays, pop, params, hzp


def hz_population():
    """The fixture population with the hazard grid's person flags."""
    csa = importlib.import_module("covid-sim_amd")
    arrays, pop, params, hzp = mini()
    popa = dict(pop.a)
    popa["p_flags"] = arrays["hz_p_flags"]
    return csa.population.Population(popa)
