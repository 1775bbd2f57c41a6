"""Regions of uneven size advanced together (cs_batch_*): every output step,
each region's step graph is launched on its own stream before any wait.
Each region's realisations
must come out exactly as cs_run_steps gives them on that region alone, and
as the oracle gives them; an engine whose work lists overflow inside the
batch must be grown and brought level without changing any result."""
import importlib

import numpy as np
import pytest

from fixtures import c1, mini, mini_int
from oracle_lib import OracleEngine

DAYS = 60


def regions():
    """(population, params, seeds, per-realisation set-up) of three regions:
    the 20k-person region, the 200k-person C1 region, and the 20k region
    with interventions, per-realisation LocalBeta and trigger thresholds"""
    _, mpop, mpar, _ = mini()
    cpop, cpar = c1()
    ipop, ipar = mini_int()
    return [
        (mpop, mpar, np.array([11, 12], np.uint64), {}),
        (cpop, cpar, np.array([21, 22, 23], np.uint64), {}),
        (ipop, ipar, np.array([31, 32], np.uint64),
         {"beta": np.array([0.9, 1.15]) * ipar.local_beta, "thr": np.array([8.0, 60.0])}),
    ]


def make(engmod, pop, params, seeds, extra, capacity=None):
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    if capacity:
        g.set_queue_capacity(*capacity)
    if "beta" in extra:
        g.set_realisation_params(local_beta=extra["beta"])
    if "thr" in extra:
        g.set_realisation_triggers(extra["thr"])
    g.reset(10)
    return g


def results(eng, pop, n, csa):
    """series, then per realisation its state and its counters (sumRad2, a
    double sum in no fixed order, aside)"""
    out = [eng.timeseries(DAYS)]
    for r in range(n):
        st = csa.population.State(pop)
        eng.store_state(r, st)
        out.append(st)
        c = eng.counters(r)
        out.append(tuple(repr(np.array(v).tolist()) if hasattr(v, "__len__") else v
                         for v in (getattr(c, f) for f, _ in c._fields_ if f != "sumRad2")))
    return out


def same(a, b):
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1:], b[1:]):
        if isinstance(x, tuple):
            assert x == y
        else:
            assert x.compare(y) == []


@pytest.mark.gpu
@pytest.mark.parametrize("overflow", [False, True])
def test_batch_equals_each_region_alone(overflow):
    engmod = importlib.import_module("covid-sim_amd.engine")
    csa = importlib.import_module("covid-sim_amd")
    regs = regions()
    # with overflow: the C1 engine starts with tiny lists, so its replays
    # inside the batch halt and it is grown and brought level
    caps = [None, (1024, 64) if overflow else None, None]
    eng = [make(engmod, p, q, s, x, c) for (p, q, s, x), c in zip(regs, caps)]
    b = engmod.Batch(eng)
    tspd = int(regs[0][1].time_steps_per_day)
    b.run_steps(0.0, DAYS * tspd, tspd)
    if overflow:
        assert eng[1].queue_capacity()[0] > 1024
    for k, (pop, params, seeds, extra) in enumerate(regs):
        alone = make(engmod, pop, params, seeds, extra)
        alone.run_steps(0.0, DAYS * tspd, tspd)
        same(results(eng[k], pop, len(seeds), csa), results(alone, pop, len(seeds), csa))
        # each region's epidemic grew
        assert (eng[k].timeseries(DAYS)[:, :, 6].sum(axis=1) > 50).all(), k
    # the batch's realisations are the oracle's too (the small regions)
    for k in (0, 2):
        pop, params, seeds, extra = regs[k]
        o = OracleEngine(params, pop)
        o.alloc(seeds)
        if "beta" in extra:
            o.set_realisation_params(local_beta=extra["beta"])
        if "thr" in extra:
            o.set_realisation_triggers(extra["thr"])
        o.reset(10)
        o.run_steps(0.0, DAYS * tspd, tspd)
        assert np.array_equal(eng[k].timeseries(DAYS), o.timeseries(DAYS)), k
    b.close()


@pytest.mark.gpu
def test_batch_continues_a_series():
    """Two batch calls of 30 days give the series of one 60-day call."""
    engmod = importlib.import_module("covid-sim_amd.engine")
    regs = regions()[:2]
    tspd = int(regs[0][1].time_steps_per_day)
    one = [make(engmod, p, q, s, x) for p, q, s, x in regs]
    two = [make(engmod, p, q, s, x) for p, q, s, x in regs]
    engmod.Batch(one).run_steps(0.0, DAYS * tspd, tspd)
    b = engmod.Batch(two)
    b.run_steps(0.0, 30 * tspd, tspd)
    b.run_steps(30.0, 30 * tspd, tspd)
    for x, y in zip(one, two):
        assert np.array_equal(x.timeseries(DAYS), y.timeseries(DAYS))
