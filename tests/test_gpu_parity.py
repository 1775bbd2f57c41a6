"""The HIP library against the oracle and the reference's golden vectors.

Bar: bit-exact for tables, indices, random streams, samplers and whole
trajectories (the GPU and the oracle share the counter-based random streams
and the deterministic event/list rules); hazards within 1e-6 relative of
the reference (in practice exact: both evaluate the reference's double
products in its order).  sumRad2 is the one float accumulated with atomics
in arbitrary order: compared to 1e-12 relative.
"""
import importlib

import numpy as np
import pytest

from fixtures import c1, hz_population, mini, mini_int
from oracle_lib import OracleEngine

pytestmark = pytest.mark.gpu

HAZARD_RTOL = 1e-6


@pytest.fixture(scope="module")
def engmod():
    return importlib.import_module("covid-sim_amd.engine")


@pytest.fixture(scope="module")
def pair(engmod):
    """GPU engine and oracle on the fixture population, one realisation."""
    arrays, pop, params, _ = mini()
    seeds = np.array([777], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    return g, o


def test_philox_known_answers(pair):
    """Philox4x32-10 known-answer vectors (Salmon et al., Random123 kat_vectors)."""
    g, o = pair
    ck = np.array([0, 0, 0, 0, 0, 0,
                   0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff,
                   0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0], np.uint32)
    want = np.array([0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8,
                     0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd,
                     0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1], np.uint32)
    assert np.array_equal(g.philox(ck), want)
    assert np.array_equal(o.philox(ck), want)


@pytest.mark.parametrize("which,a,b", [(0, 0, 0), (1, 30, 0.05), (1, 400, 0.3), (1, 5000, 0.97), (1, 12, 0.5),
                                       (2, 0.7, 0), (2, 9.99, 0), (2, 10.0, 0), (2, 250.0, 0), (3, 0, 0),
                                       (4, 1.0, 1.0), (4, 4.0, 4.0), (4, 2.0, 0.5)])
def test_samplers_bit_exact(pair, which, a, b):
    g, o = pair
    n = 3000
    x = g.sample(which, 0x1234567890ABCDEF, 77, 5, a, b, n)
    y = o.sample(which, 0x1234567890ABCDEF, 77, 5, a, b, n)
    assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_kernel_lookup_vs_reference(pair):
    g, o = pair
    a = mini()[0]
    out = g.kernel_num(a["g_kernel_r2"])
    assert np.array_equal(out.view(np.uint64), a["g_kernel_num"].view(np.uint64))


def test_dist2_vs_reference(pair):
    g, o = pair
    a = mini()[0]
    out = g.dist2(a["g_dist_a"], a["g_dist_b"])
    assert np.array_equal(out, a["g_dist2"])


def test_cell_tables_vs_reference(pair, csa):
    g, o = pair
    arrays, pop, params, _ = mini()
    st = csa.population.State.from_dump(pop, arrays, "s0_")
    g.load_state(0, st)
    mt, ct, inv, tp, s0 = g.cell_tables(0, pop.n_cells)
    assert np.array_equal(mt.view(np.uint32), arrays["s0_max_trans"].view(np.uint32))
    assert np.array_equal(ct.view(np.uint32), arrays["s0_cum_trans"].view(np.uint32))
    assert np.array_equal(tp.view(np.uint32), arrays["s0_tot_prob"].view(np.uint32))
    assert np.array_equal(inv, arrays["s0_invcdf"])
    assert np.array_equal(s0, arrays["s0_cell_S0"])


def test_choose_vs_reference(pair):
    g, o = pair
    a = mini()[0]
    tab, u, mean, val = a["g_choose_table"], a["g_choose_u"], a["g_choose_mean"], a["g_choose_value"]
    for t in np.unique(tab):
        sel = tab == t
        assert np.array_equal(g.choose(int(t), u[sel], mean[sel]), val[sel])


def test_hazards_vs_reference(engmod, csa):
    arrays, pop, params, hzp = mini()
    hzp.interventions_active = 1
    hpop = hz_population()
    g = engmod.Engine(0)
    g.set_params(hzp)
    g.set_population(hpop)
    g.alloc(np.array([1], np.uint64))
    st = csa.population.State.from_dump(hpop, arrays, "hz_")
    g.load_state(0, st)
    person, other, pt, ts, ref = arrays["g_hz_person"], arrays["g_hz_other"], arrays["g_hz_pt"], arrays["g_hz_ts"], arrays["g_hz_out"]
    for kind in range(8):
        sec = pt if kind in (1, 5) else other
        out = np.zeros(len(person))
        for t in np.unique(ts):
            sel = np.nonzero(ts == t)[0]
            out[sel] = g.hazards(0, kind, int(t), person[sel], sec[sel])
        rel = np.abs(out - ref[:, kind]) / np.maximum(np.abs(ref[:, kind]), 1e-300)
        assert rel.max() <= HAZARD_RTOL, "kind %d max rel %g" % (kind, rel.max())


def test_indexing_vs_oracle(pair, csa):
    """Infector binary search and destination-cell selection for random u."""
    g, o = pair
    arrays, pop, params, _ = mini()
    st = csa.population.State.from_dump(pop, arrays, "s160_")
    g.load_state(0, st)
    o.load_state(0, st)
    rng = np.random.default_rng(3)
    u = rng.random(20000)
    u = np.clip(u, 1e-15, 1 - 1e-15)
    for n in (1, 2, 3, 7, 64, 1000):
        ci = np.cumsum(rng.random(n)).astype(np.float32)
        ci /= ci[-1]
        ci[-1] = 1.0
        ci[rng.random(n) < 0.2] *= -1
        assert np.array_equal(g.select_infector(ci, u), o.select_infector(ci, u))
    for c in range(pop.n_cells):
        got = g.select_dest(0, c, u)
        assert np.array_equal(got, o.select_dest(0, c, u))
        _, ct, _, _, _ = g.cell_tables(0, pop.n_cells)
        # inverse-CDF definition: first l with cum[l] >= u
        want = np.argmax(ct[c][None, :] >= u[:, None], axis=1)
        assert np.array_equal(got, want)


def _compare_counters(a, b):
    for f, _ in a._fields_:
        va, vb = getattr(a, f), getattr(b, f)
        if f == "sumRad2":
            assert abs(va - vb) <= 1e-12 * max(abs(va), 1.0), f
        elif hasattr(va, "__len__"):
            assert np.array_equal(np.array(va), np.array(vb)), f
        else:
            assert va == vb, f


def test_one_step_matches_oracle(pair, csa):
    """One InfectSweep + IncubRecoverySweep from the reference's own state at
    day 40 of its run: every person, cell and place array equal."""
    g, o = pair
    arrays, pop, params, _ = mini()
    st = csa.population.State.from_dump(pop, arrays, "s160_")
    for eng in (g, o):
        eng.load_state(0, st)
        eng.counters(0, reset=True)
    t = 160 / params.time_steps_per_day
    for eng in (g, o):
        eng.infect_sweep(t)
    a, b = csa.population.State(pop), csa.population.State(pop)
    g.store_state(0, a)
    o.store_state(0, b)
    assert a.compare(b) == []
    for eng in (g, o):
        eng.incub_recovery_sweep(t)
    g.store_state(0, a)
    o.store_state(0, b)
    assert a.compare(b) == []
    ca, cb = g.counters(0), o.counters(0)
    _compare_counters(ca, cb)
    assert ca.cumI > 20  # the step really infected people


def test_trajectory_matches_oracle(engmod, csa):
    """From the realisation reset (seeding included) through 120 days, in one
    batched launch of 3 realisations: daily series, final state and
    counters equal to the oracle's."""
    arrays, pop, params, _ = mini()
    seeds = np.array([11, 12, 13], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    steps = int(120 * params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, steps, int(params.time_steps_per_day))
    tg, to = g.timeseries(120), o.timeseries(120)
    assert tg.shape == (3, 120, 21)
    assert np.array_equal(tg, to)
    for r in range(3):
        a, b = csa.population.State(pop), csa.population.State(pop)
        g.store_state(r, a)
        o.store_state(r, b)
        assert a.compare(b) == []
        _compare_counters(g.counters(r), o.counters(r))
    # the epidemic took off in every realisation
    assert (tg[:, -1, 3] + tg[:, -1, 4] > pop.n_people // 2).all()
    # work counts equal, except the cum_trans reads of the destination
    # search: the oracle scans from the InvCDF start as the reference does,
    # the library binary-searches the bucket's bracket (same cells)
    wg, wo = g.work_counts(), o.work_counts()
    for k in wg:
        if k not in ("n_cells", "n_sp_scan"):
            assert wg[k] == wo[k], k


def test_invariants_on_gpu(engmod, csa):
    """Population conservation and list consistency after a GPU run."""
    arrays, pop, params, _ = mini()
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(np.array([99], np.uint64))
    g.reset(10)
    g.run_steps(0.0, 200, 4)
    ts = g.timeseries(50)
    assert (ts[0, :, :5].sum(axis=1) == pop.n_people).all()
    st = csa.population.State(pop)
    g.store_state(0, st)
    a = st.a
    cs_ = pop.a["c_member_start"]
    for c in range(pop.n_cells):
        seg = a["cell_perm"][cs_[c]:cs_[c + 1]]
        assert np.array_equal(np.sort(seg), np.sort(pop.a["c_members"][cs_[c]:cs_[c + 1]]))
        assert np.array_equal(a["listpos"][seg], np.arange(len(seg)))
        S, L, I = a["cell_S"][c], a["cell_L"][c], a["cell_I"][c]
        inf = a["inf"][seg]
        assert (inf[:S] == 0).all()
        assert (inf[S:S + L] == 1).all()
        assert np.isin(inf[S + L:S + L + I], [-1, 2, -2]).all()
        assert np.isin(inf[S + L + I:], [3, -3, 5, -5]).all()


@pytest.mark.parametrize("graphs,events,draw_chunks", [(False, 0, 0), (True, 512, 4), (False, 512, 4)])
def test_launch_modes_and_overflow_recovery(engmod, csa, graphs, events, draw_chunks):
    """Replayed step graphs or direct launches, and event-queue / draw-list
    overflows (tiny capacities force the halt, grow and re-run path in the
    middle of a replay and of a direct step): the same 60-day trajectory as
    the oracle.  Work counts are left out: a re-run step counts twice."""
    arrays, pop, params, _ = mini()
    seeds = np.array([21, 22], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    g.set_step_graphs(graphs)
    if events:
        g.set_queue_capacity(events, draw_chunks)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, 60 * tspd + 2, tspd)  # 2 steps past the last sample: the direct tail
    assert np.array_equal(g.timeseries(61), o.timeseries(61))
    if events:
        ev, dr = g.queue_capacity()
        assert ev > events and dr > draw_chunks  # the lists grew
    for r in range(2):
        a, b = csa.population.State(pop), csa.population.State(pop)
        g.store_state(r, a)
        o.store_state(r, b)
        assert a.compare(b) == []


@pytest.mark.parametrize("cap,graphs", [(0, True), (0, False), (100, True)])
def test_list_moves_through_device_memory(engmod, csa, cap, graphs):
    """Cells with more list moves in a step than the move capacity take the
    device-memory bitmap path (k_list_move_big); capacity 0 sends every
    moving cell there, 100 mixes it with the LDS block path (65-100 moves)
    and the wave path (up to 64): the 90-day trajectory and the final
    states stay the oracle's."""
    arrays, pop, params, _ = mini()
    seeds = np.array([51, 52], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.set_move_capacity(cap)
    g.alloc(seeds)
    g.set_step_graphs(graphs)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, 90 * tspd, tspd)
    assert np.array_equal(g.timeseries(90), o.timeseries(90))
    assert g.timeseries(90)[:, -1, 3].min() > 1000
    for r in range(2):
        a, b = csa.population.State(pop), csa.population.State(pop)
        g.store_state(r, a)
        o.store_state(r, b)
        assert a.compare(b) == []
    with pytest.raises(Exception):
        g.set_move_capacity(4097)


@pytest.mark.parametrize("graphs", [True, False])
def test_latent_due_list_overflow(engmod, csa, graphs):
    """Latent due lists far too small for the epidemic: the steps whose list
    overflows scan every latent person instead, and the 60-day trajectory
    stays the oracle's."""
    arrays, pop, params, _ = mini()
    seeds = np.array([41, 42], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    g.set_step_graphs(graphs)
    g.set_due_capacity(8)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, 60 * tspd, tspd)
    assert np.array_equal(g.timeseries(60), o.timeseries(60))
    assert g.timeseries(60)[:, -1, 3].min() > 100  # the epidemic grew past the lists
    for r in range(2):
        a, b = csa.population.State(pop), csa.population.State(pop)
        g.store_state(r, a)
        o.store_state(r, b)
        assert a.compare(b) == []


def test_autotune_keeps_results(engmod, csa):
    """The autotuner times the event phase on the current state without
    changing it; launch shapes and capacities do not change results."""
    arrays, pop, params, _ = mini()
    seeds = np.array([31, 32], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, 30 * tspd, tspd)
    tune = g.autotune(30.0, reps=2)
    assert tune["grid_waves"] > 0 and tune["grid_cells"] > 0 and tune["event_phase_ms"] > 0
    for eng in (g, o):
        eng.run_steps(30.0, 30 * tspd, tspd)
    assert np.array_equal(g.timeseries(60), o.timeseries(60))


def test_household_beyond_the_size_table_is_rejected(engmod, csa):
    """A household holds at most CS_MAX_HOUSEHOLD_SIZE persons (the
    reference sets nh = nhr from the size table, src/SetupModel.cpp:1287-1288):
    a larger one is refused by the host checks, and the engine then takes a
    valid population and runs it."""
    arrays, pop, params, _ = mini()
    popa = dict(pop.a)
    nh = popa["h_nh"].copy()
    nh[0] = csa.abi.CS_MAX_HOUSEHOLD_SIZE + 1
    popa["h_nh"] = nh
    g = engmod.Engine(0)
    g.set_params(params)
    with pytest.raises(engmod.CSError, match="household 0 out of range"):
        g.set_population(csa.population.Population(popa))
    g.set_population(pop)
    g.alloc(np.array([5], np.uint64))
    o = OracleEngine(params, pop)
    o.alloc(np.array([5], np.uint64))
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, 10 * tspd, tspd)
    assert np.array_equal(g.timeseries(10), o.timeseries(10))


def test_per_realisation_params_match_oracle(engmod, csa):
    """Distinct LocalBeta and seed counts per realisation
    (cs_set_realisation_params, the parameter sweep of a batched run): the
    60-day series, states and counters equal the oracle's with the same
    overrides, and the overrides really change the epidemics."""
    arrays, pop, params, _ = mini()
    seeds = np.array([51, 52, 53, 54], np.uint64)
    lb = np.array([0.5, 1.0, 1.6, 2.2]) * params.local_beta
    ns = np.array([1, 10, 40, 5], np.int32)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.set_realisation_params(local_beta=lb, n_seed=ns)
        eng.reset(10)
        eng.run_steps(0.0, 60 * tspd, tspd)
    tg = g.timeseries(60)
    assert np.array_equal(tg, o.timeseries(60))
    for r in range(len(seeds)):
        a, b = csa.population.State(pop), csa.population.State(pop)
        g.store_state(r, a)
        o.store_state(r, b)
        assert a.compare(b) == []
        _compare_counters(g.counters(r), o.counters(r))
    # the seed counts took effect (the first sample, after day 1, has at
    # least the seeds infected) and beta orders the growth
    assert (pop.n_people - tg[:, 0, 0] >= ns).all() and pop.n_people - tg[2, 0, 0] >= 40
    assert tg[3, -1, 0] < tg[0, -1, 0]


@pytest.mark.parametrize("cap", [None, 0])
def test_block_list_moves_match_oracle(engmod, csa, cap):
    """Cells with more than 64 list moves in one step take the block path of
    the batched move (a wave takes smaller ones); with move capacity 0 every
    moving cell takes the device-memory bitmap path instead, here over list
    segments of thousands of persons (many bitmap words per thread).  On the
    200k-person C1 region with its LocalBeta raised fourfold (so the peak is
    sharp), step by step through days 34-37: every state equal to the
    oracle's, and the per-cell move counts of each kind (S->L, L->I,
    I->R/D), derived from the cell totals, exceed 64 in some cell."""
    pop, params = c1()
    seeds = np.array([61, 62], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    if cap is not None:
        g.set_move_capacity(cap)
    g.alloc(seeds)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    day0 = 34
    for eng in (g, o):
        eng.set_realisation_params(local_beta=np.full(2, 4.0 * params.local_beta))
        eng.reset(10)
        eng.run_steps(0.0, day0 * tspd, tspd)

    def cells(eng, r):
        st = csa.population.State(pop)
        eng.store_state(r, st)
        a = st.a
        return st, np.stack([a["cell_S"], a["cell_L"], a["cell_I"], a["cell_R"] + a["cell_D"]]).astype(np.int64)

    most = np.zeros(3, np.int64)
    prev = [cells(o, r)[1] for r in range(2)]
    for k in range(12):
        t = day0 + k / tspd
        for eng in (g, o):
            eng.run_steps(t, 1, 1)
        for r in range(2):
            a, ca = cells(g, r)
            b, cb = cells(o, r)
            assert a.compare(b) == [], (k, r)
            s_l = prev[r][0] - cb[0]
            l_i = prev[r][1] + s_l - cb[1]
            i_r = cb[3] - prev[r][3]
            most = np.maximum(most, [s_l.max(), l_i.max(), i_r.max()])
            prev[r] = cb
    assert (most > 64).all(), most


@pytest.mark.parametrize("graphs", [True, False])
def test_interventions_match_oracle(engmod, csa, graphs):
    """The intervention subset (global ICU triggers, an alert that raises
    case detection, case isolation, household quarantine, school and
    university closure with absenteeism and carers) over 120 days of three
    realisations, with replayed step graphs or direct launches: series,
    states, counters and the intervention state (alert, trigger, start
    times, closed microcells) equal to the oracle's."""
    pop, params = mini_int()
    seeds = np.array([71, 72, 73], np.uint64)
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    g.set_step_graphs(graphs)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.reset(10)
        eng.run_steps(0.0, 120 * tspd, tspd)
    assert np.array_equal(g.timeseries(120), o.timeseries(120))
    for r in range(3):
        a, b = csa.population.State(pop), csa.population.State(pop)
        g.store_state(r, a)
        o.store_state(r, b)
        assert a.compare(b) == [], r
        ca, cb = g.counters(r), o.counters(r)
        _compare_counters(ca, cb)
        sa, sb = g.intervention_state(r), o.intervention_state(r)
        for f, _ in sa._fields_:
            assert getattr(sa, f) == getattr(sb, f), (r, f)
        # every intervention ran: the alert, all three start times, and
        # quarantined households, closure absences and isolations
        assert sa.alert == 1 and sa.place_close_start < 1e9 and sa.case_isolation_start < 1e9
        assert ca.cumHQ > 0 and ca.cumAPCS > 0
        assert (a.a["isolation_start_time"] != 65534).sum() > 0


@pytest.mark.parametrize("graphs", [True, False])
def test_per_realisation_triggers_match_oracle(engmod, csa, graphs):
    """Per-realisation trigger thresholds (the same value for the three
    thresholds of realisation r: the run's own, a higher one, one never
    reached) over 120 days: series, counters and intervention state equal to
    the oracle's, and the start times ordered by threshold."""
    pop, params = mini_int()
    seeds = np.array([81, 81, 81], np.uint64)
    thr = np.array([5.0, 40.0, 1e9])
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    g.set_step_graphs(graphs)
    o = OracleEngine(params, pop)
    o.alloc(seeds)
    tspd = int(params.time_steps_per_day)
    for eng in (g, o):
        eng.set_realisation_triggers(thr)
        eng.reset(10)
        eng.run_steps(0.0, 120 * tspd, tspd)
    assert np.array_equal(g.timeseries(120), o.timeseries(120))
    starts = []
    for r in range(3):
        _compare_counters(g.counters(r), o.counters(r))
        sa, sb = g.intervention_state(r), o.intervention_state(r)
        for f, _ in sa._fields_:
            assert getattr(sa, f) == getattr(sb, f), (r, f)
        starts.append(sa.case_isolation_start)
    # one seed, so the runs agree until the lowest threshold is crossed
    assert starts[0] < starts[1] < 1e9 and starts[2] >= 1e10, starts


def test_interventions_refused_outside_the_subset(engmod, csa):
    """Parameter sets outside the intervention subset are refused with
    CS_EUNSUPPORTED (never approximated), and a single sweep refuses an
    intervention set (its triggers and TreatSweep run between sweeps)."""
    pop, params = mini_int()
    g = engmod.Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(np.array([5], np.uint64))
    g.reset(10)
    with pytest.raises(engmod.CSError, match="only in cs_run_steps"):
        g.infect_sweep(0.0)
    for field, value, msg in [("do_global_triggers", 0, "global triggers"), ("soc_dist_time_start_base", 5.0, "social distancing"),
                              ("place_close_cell_inc_stop_thresh", 3, "stop threshold"), ("do_per_capita_triggers", 1, "per-capita")]:
        bad = csa.abi.cs_params()
        C_ = __import__("ctypes")
        C_.memmove(C_.addressof(bad), C_.addressof(params), C_.sizeof(bad))
        setattr(bad, field, value)
        g.set_params(bad)
        with pytest.raises(engmod.CSError, match=msg):
            g.reset(10)
    g.set_params(params)
    g.reset(10)
    g.run_steps(0.0, 8, 4)
