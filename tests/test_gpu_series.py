"""The library's samples by age group and by administrative unit equal the
oracle's (whose sums are pinned against the totals and the state in
test_series_by_age_unit.py), with step graphs on and off, with and without
the intervention subset, and across a series continued over two calls."""
import importlib

import numpy as np
import pytest

from fixtures import mini, mini_int
from oracle_lib import OracleEngine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,graphs", [("plain", True), ("plain", False), ("interventions", True)])
def test_series_by_age_and_unit_equal_oracle(which, graphs):
    eng_mod = importlib.import_module("covid-sim_amd.engine")
    arrays, pop, params, _ = mini()
    if which == "interventions":
        pop, params = mini_int()
    na = pop.n_adunits
    seeds = np.array([31, 32, 33], np.uint64)
    g = eng_mod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    g.set_step_graphs(graphs)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, 60 * tspd, tspd)   # a series continued over two calls
        eng.run_steps(60.0, 60 * tspd, tspd)
    assert np.array_equal(g.timeseries(120), o.timeseries(120))
    assert np.array_equal(g.timeseries_age(120), o.timeseries_age(120))
    assert np.array_equal(g.timeseries_adunit(120, na), o.timeseries_adunit(120, na))
    for r in range(3):
        assert np.array_equal(g.adunit_counters(r, na), o.adunit_counters(r, na))
    assert g.timeseries_age(120)[:, :, 0].sum() > 1000
