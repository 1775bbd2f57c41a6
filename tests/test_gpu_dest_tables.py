"""Checkpointed destination tables against the full ones and the oracle.

With checkpointed tables (cs_set_dest_tables(stride > 0)) the library keeps
no NC x NC table: max_trans is evaluated where it is read and each cum_trans
row only at every stride-th cell.  The reference's destination rule
(src/Sweep.cpp:757-762, tables of src/CovidSim.cpp:2752-2802 and
src/Kernels.cpp:48-63) must come out bit for bit: tables expanded from the
checkpoints equal the full tables, every destination equals the full
tables' for the same uniforms (at random and at the exact cumulative
boundaries), and whole trajectories equal the oracle's (which keeps full
tables).
"""
import importlib

import numpy as np
import pytest

from fixtures import c1, mini
from oracle_lib import OracleEngine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engmod():
    return importlib.import_module("covid-sim_amd.engine")


def _engine(engmod, params, pop, seeds, stride, index=True):
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.set_dest_tables(stride)
    g.set_dest_index(index)
    g.alloc(np.asarray(seeds, np.uint64))
    assert g.dest_tables() == stride
    return g


def _tables_equal(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.ascontiguousarray(x).view(np.uint32), np.ascontiguousarray(y).view(np.uint32))


@pytest.mark.parametrize("stride", [1, 3, 7, 16, 49, 64])
def test_tables_expand_to_the_full_ones(engmod, stride):
    """C1 region (49 cells), after the reset and after 40 days (S0 changed)."""
    pop, params = c1()
    seeds = [5, 6]
    full = _engine(engmod, params, pop, seeds, 0)
    ck = _engine(engmod, params, pop, seeds, stride)
    tspd = int(params.time_steps_per_day)
    for eng in (full, ck):
        eng.reset(10)
    for r in range(2):
        _tables_equal(full.cell_tables(r, pop.n_cells), ck.cell_tables(r, pop.n_cells))
    for eng in (full, ck):
        eng.run_steps(0.0, 40 * tspd, tspd)
        eng.update_probs()
    for r in range(2):
        _tables_equal(full.cell_tables(r, pop.n_cells), ck.cell_tables(r, pop.n_cells))


def _edges_and_buckets(rng, prefixes):
    """random uniforms, each prefix value and the doubles either side of it,
    and the bucket edges k / 1024 and the doubles either side of them"""
    edges = prefixes.astype(np.float64)
    edges = edges[edges < 1.0]
    k = np.arange(1, 1024) / 1024.0
    u = np.concatenate([rng.random(4000), edges, np.nextafter(edges, 0.0), np.nextafter(edges, 1.0),
                        k, np.nextafter(k, 0.0), np.nextafter(k, 1.0)])
    return u[(u > 0.0) & (u < 1.0)]


@pytest.mark.parametrize("stride,index", [(1, True), (5, True), (16, True), (5, False), (16, False)])
def test_destinations_equal_at_the_boundaries(engmod, stride, index):
    """Every source cell of C1 after 40 days: random uniforms, every
    normalised prefix value itself and the doubles just below and above it
    (where a rounding slip in a redone block would show), and the bucket
    edges of the brackets; checkpoints with and without the rows' bucket
    index."""
    pop, params = c1()
    full = _engine(engmod, params, pop, [9], 0)
    ck = _engine(engmod, params, pop, [9], stride, index)
    tspd = int(params.time_steps_per_day)
    for eng in (full, ck):
        eng.reset(10)
        eng.run_steps(0.0, 40 * tspd, tspd)
        eng.update_probs()
    _, ct, _, _, _ = full.cell_tables(0, pop.n_cells)
    rng = np.random.default_rng(8)
    for c in range(pop.n_cells):
        u = _edges_and_buckets(rng, ct[c])
        a, b = full.select_dest(0, c, u), ck.select_dest(0, c, u)
        assert np.array_equal(a, b), c
        # the inverse-CDF definition: the first cell whose prefix reaches u
        assert np.array_equal(b, np.argmax(ct[c][None, :].astype(np.float64) >= u[:, None], axis=1))


@pytest.mark.parametrize("graphs", [False, True])
def test_trajectory_with_checkpoints_matches_oracle(engmod, csa, graphs):
    """The 20k-person fixture (16 cells) with checkpoints every 5 cells (a
    partial last block): 3 realisations through 120 days, series, state and
    counters equal to the oracle's full-table run."""
    arrays, pop, params, _ = mini()
    seeds = [11, 12, 13]
    g = _engine(engmod, params, pop, seeds, 5)
    g.set_step_graphs(graphs)
    o = OracleEngine(params, pop)
    o.alloc(np.asarray(seeds, np.uint64))
    steps = int(120 * params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, steps, int(params.time_steps_per_day))
    assert np.array_equal(g.timeseries(120), o.timeseries(120))
    for r in range(3):
        a, b = csa.population.State(pop), csa.population.State(pop)
        g.store_state(r, a)
        o.store_state(r, b)
        assert a.compare(b) == []
    wg, wo = g.work_counts(), o.work_counts()
    for k in ("n_sp_draw", "n_sp_susc", "n_new"):
        assert wg[k] == wo[k], k


def test_many_cell_region_equals_full_tables(engmod, csa):
    """A generated region of 1,156 cells (300k persons): checkpoints every
    128 cells (10 blocks a row, the last partial) against full tables,
    2 realisations through 80 days with the S-drop table rebuilds; the
    series and every person's state equal."""
    popgen = importlib.import_module("covid-sim_amd.popgen")
    arrays, geom = popgen.generate(300000, side_deg=2.55, seed=5)
    pop = csa.population.Population(arrays)
    assert pop.n_cells == 1156
    _, params = c1()
    params = type(params).from_buffer_copy(params)
    popgen.apply_geometry(params, geom)
    seeds = [21, 22]
    full = _engine(engmod, params, pop, seeds, 0)
    ck = _engine(engmod, params, pop, seeds, 128)
    tspd = int(params.time_steps_per_day)
    for eng in (full, ck):
        eng.reset(10)
        eng.run_steps(0.0, 80 * tspd, tspd)
    ta, tb = full.timeseries(80), ck.timeseries(80)
    assert np.array_equal(ta, tb)
    assert ta[:, -1, 1:5].sum() > 0.2 * pop.n_people  # the epidemic spread over the region
    for r in range(2):
        a, b = csa.population.State(pop), csa.population.State(pop)
        full.store_state(r, a)
        ck.store_state(r, b)
        assert a.compare(b) == []


def test_many_cell_destinations(engmod, csa):
    """The 1,156-cell region after 30 days: for 60 source cells, random
    uniforms, prefix values and bucket edges give the same destination from
    full tables (InvCDF brackets), checkpoints every 16 and every 128 cells
    with the rows' bucket index, and checkpoints without it; each is the first
    cell whose prefix reaches u."""
    popgen = importlib.import_module("covid-sim_amd.popgen")
    arrays, geom = popgen.generate(300000, side_deg=2.55, seed=5)
    pop = csa.population.Population(arrays)
    _, params = c1()
    params = type(params).from_buffer_copy(params)
    popgen.apply_geometry(params, geom)
    engs = [_engine(engmod, params, pop, [3], 0), _engine(engmod, params, pop, [3], 16),
            _engine(engmod, params, pop, [3], 128), _engine(engmod, params, pop, [3], 16, False)]
    tspd = int(params.time_steps_per_day)
    for eng in engs:
        eng.reset(10)
        eng.run_steps(0.0, 30 * tspd, tspd)
        eng.update_probs()
    _, ct, _, _, _ = engs[0].cell_tables(0, pop.n_cells)
    rng = np.random.default_rng(4)
    for c in rng.choice(pop.n_cells, 60, replace=False):
        u = _edges_and_buckets(rng, ct[c])
        want = np.argmax(ct[c][None, :].astype(np.float64) >= u[:, None], axis=1)
        for eng in engs:
            assert np.array_equal(eng.select_dest(0, int(c), u), want), c


def test_automatic_choice(engmod):
    """-1 picks full tables for a small region."""
    pop, params = c1()
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    assert g.dest_tables() == -1
    g.alloc(np.array([1], np.uint64))
    assert g.dest_tables() == 0
