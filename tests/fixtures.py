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


@functools.lru_cache(maxsize=None)
def c1():
    """(Population, params) of the 200k-person C1 region, as the reference's
    set-up builds it from covid-sim_amd/inputs/c1 (/c:1)."""
    csa = importlib.import_module("covid-sim_amd")
    z = np.load(os.path.join(GOLDEN, "c1_population.npz"))
    pop = csa.population.Population({k: z[k] for k in z.files})
    params = csa.population.load_params(os.path.join(os.path.dirname(HERE), "covid-sim_amd", "inputs", "c1", "params.bin"))
    return pop, params


@functools.lru_cache(maxsize=None)
def mini_care():
    """(Population, params) of the 20k-person region with care homes: mini's
    population with the reference's care-home resident flags, and the
    parameters the reference derives with the care-home keys
    (tests/golden/make_fixtures.py CARE_KEYS)."""
    csa = importlib.import_module("covid-sim_amd")
    arrays, pop, params, _ = mini()
    popa = dict(pop.a)
    popa["p_flags"] = np.load(os.path.join(GOLDEN, "mini_care_flags.npz"))["p_flags"]
    params = csa.population.load_params(os.path.join(GOLDEN, "mini_care_params.bin"))
    return csa.population.Population(popa), params


@functools.lru_cache(maxsize=None)
def mini_int():
    """(Population, params) of the 20k-person region with the intervention
    subset switched on (tests/golden/make_fixtures.py INT_PRE, INT_PARAMS)."""
    csa = importlib.import_module("covid-sim_amd")
    arrays, pop, params, _ = mini()
    return pop, csa.population.load_params(os.path.join(GOLDEN, "mini_int_params.bin"))
