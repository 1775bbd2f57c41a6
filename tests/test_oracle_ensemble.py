"""The oracle's and the GPU's stochastic trajectories against the
reference's, as ensembles.

The random streams differ by design (counter-based streams replace the
reference's per-thread generators), so single trajectories cannot match;
the distribution of trajectories must.  Each reference ensemble is 200
realisations of the unmodified reference (tests/golden/make_ensemble.py) on
one of four workloads, all with the population the reference's own set-up
builds (/c:1):

  mini       the 20k-person fixture region, 120 days
  mini_care  the same region with care homes (workplaces whose members aged
             55 and over are residents, every care-home scaling away from 1;
             tests/golden/make_fixtures.py CARE_KEYS), 120 days
  mini_int   the same region with the intervention subset: global triggers
             on ICU cases, case isolation, household quarantine, school and
             university closure, an alert that raises case detection
             (tests/golden/make_fixtures.py INT_PRE, INT_PARAMS), 120 days
  c1         the 200k-person C1 region, 200 days

Compared: attack rate, peak day, peak daily incidence, cumulative cases and
deaths, each with a two-sample Kolmogorov-Smirnov test (p > 0.01) and
agreement of the two ensembles' 95% confidence intervals for the mean (the
intervals overlap); the daily infections and cases of every day with a
per-day KS test, Bonferroni-corrected over the days (family-wise level 0.01;
120 uncorrected tests at 0.01 would reject a true null in 70% of suites); and
the final infections by route; with interventions, the start times of
place closure, case isolation and household quarantine.  A 1000-versus-1000 run of the mini
comparison (not part of the suite) gave |z| < 1.8 for every statistic.

On the CPU the oracle's 200 realisations are compared (mini, mini_care,
mini_int).  On
the GPU the same 200 seeds, advanced together by one batched run, must give
the oracle's 200 trajectories bit for bit and pass the same tests.  For c1,
where 200 oracle realisations of 200k persons would take minutes, the GPU's
200 are compared with the reference's, and its first two realisations with
the oracle's bit for bit.
"""
import functools
import importlib
import os

import numpy as np
import pytest
from scipy import stats

from fixtures import GOLDEN, c1, mini, mini_care, mini_int
from oracle_lib import OracleEngine

N_RUNS = 200
ALPHA = 0.01
SEEDS = np.arange(1, N_RUNS + 1, dtype=np.uint64) * np.uint64(2654435761)


def workload(name):
    if name == "mini":
        _, pop, params, _ = mini()
    elif name == "mini_care":
        pop, params = mini_care()
    elif name == "mini_int":
        pop, params = mini_int()
    else:
        pop, params = c1()
    z = np.load(os.path.join(GOLDEN, "%s_ensemble.npz" % name))
    workload.starts = z["starts"] if "starts" in z.files else None
    return pop, params, z["daily"], z["itype"].astype(float)


@functools.lru_cache(maxsize=None)
def oracle_ensemble(name):
    pop, params, ref, _ = workload(name)
    days = ref.shape[1]
    o = OracleEngine(params, pop)
    o.alloc(SEEDS)
    o.reset(10)
    o.run_steps(0.0, days * int(params.time_steps_per_day), int(params.time_steps_per_day))
    itype = np.array([list(o.counters(r).cumItype) for r in range(N_RUNS)], float)
    oracle_ensemble.starts[name] = start_times(o, N_RUNS)
    return o.timeseries(days), itype


oracle_ensemble.starts = {}


def start_times(eng, n):
    """[run, 3]: place closure, case isolation, household quarantine start
    (days; 1e10 while unset)"""
    out = np.zeros((n, 3))
    for r in range(n):
        s = eng.intervention_state(r)
        out[r] = (s.place_close_start, s.case_isolation_start, s.hquarantine_start)
    return out


def check_start_times(ref_steps, got_days, tspd=4):
    """The intervention start times: set in the same share of runs, and the
    same distribution where set (KS p > 0.01, overlapping 95% CIs)"""
    for k in range(3):
        a = ref_steps[:, k][ref_steps[:, k] >= 0] / float(tspd)
        b = got_days[:, k][got_days[:, k] < 1e9]
        assert abs(len(a) - len(b)) <= 3 * np.sqrt(max(len(a), 1)) + 2, (k, len(a), len(b))
        if len(a) > 10 and len(b) > 10:
            assert stats.ks_2samp(a, b).pvalue > ALPHA, (k, a.mean(), b.mean())
            assert cis_overlap(a, b), (k, a.mean(), b.mean())


SUMMARIES = ["attack rate", "peak day", "peak incidence", "cumulative cases", "deaths"]


def summary(daily, n):
    attack = (daily[:, -1, 3] + daily[:, -1, 4]) / n
    inc = daily[:, :, 5]
    return attack, inc.argmax(axis=1), inc.max(axis=1), daily[:, :, 6].sum(axis=1), daily[:, -1, 4]


def welch_ci_contains_zero(a, b, level=0.95):
    d = a.mean() - b.mean()
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    z = stats.norm.ppf(0.5 + level / 2)
    return abs(d) <= z * se + 1e-12


def cis_overlap(a, b, level=0.95):
    """The two samples' confidence intervals for their means overlap."""
    z = stats.norm.ppf(0.5 + level / 2)
    ha = z * a.std(ddof=1) / np.sqrt(len(a))
    hb = z * b.std(ddof=1) / np.sqrt(len(b))
    return abs(a.mean() - b.mean()) <= ha + hb + 1e-12


def check_against_reference(ref, ref_it, got, got_it, n):
    """Every ensemble comparison of the module docstring."""
    assert (got[:, :, :5].sum(axis=2) == n).all()
    assert (ref[:, :, :5].sum(axis=2) == n).all()
    for idx, name in enumerate(SUMMARIES):
        a, b = summary(ref, n)[idx].astype(float), summary(got, n)[idx].astype(float)
        assert stats.ks_2samp(a, b).pvalue > ALPHA, (name, a.mean(), b.mean())
        assert cis_overlap(a, b), (name, a.mean(), b.mean())
    days = ref.shape[1]
    for field in (5, 6):
        pv = np.array([stats.ks_2samp(ref[:, d, field], got[:, d, field]).pvalue for d in range(days)])
        assert pv.min() > ALPHA / days, (field, pv.argmin(), pv.min())
    # final infections by route: household, within- and between-group place
    # per place type, spatial (StateT.cumItype); each route's mean within a
    # Welch 99.9% interval of the reference's (0 is unused, 15 = seeds,
    # cleared in the reference before the first sample)
    for q in range(1, 15):
        if ref_it[:, q].sum() == 0 and got_it[:, q].sum() == 0:
            continue
        assert welch_ci_contains_zero(ref_it[:, q], got_it[:, q], level=0.999), (q, ref_it[:, q].mean(), got_it[:, q].mean())


@pytest.mark.parametrize("name", ["mini", "mini_care", "mini_int"])
def test_oracle_ensemble_matches_reference(name):
    pop, params, ref, ref_it = workload(name)
    got, got_it = oracle_ensemble(name)
    check_against_reference(ref, ref_it, got, got_it, pop.n_people)
    if name == "mini_int":
        check_start_times(workload.starts, oracle_ensemble.starts[name])


def test_interventions_change_the_epidemic():
    """The intervention fixture is not inert: in the reference's own
    ensemble every run triggers all three interventions, and they lower the
    attack rate against the same region without them."""
    _, _, ref, _ = workload("mini")
    _, _, iref, _ = workload("mini_int")
    starts = workload.starts
    assert (starts >= 0).all()
    assert summary(iref, 20000)[0].mean() < summary(ref, 20000)[0].mean() - 0.05


def test_care_homes_change_the_epidemic():
    """The care-home fixture is not inert: its reference ensemble differs
    from the plain region's in the infections within care homes (route of
    place type 3) and in attack rate."""
    _, _, ref, ref_it = workload("mini")
    _, _, cref, cref_it = workload("mini_care")
    assert stats.ks_2samp(ref_it[:, 2 + 3], cref_it[:, 2 + 3]).pvalue < 1e-6
    assert stats.ks_2samp(summary(ref, 20000)[0], summary(cref, 20000)[0]).pvalue < 1e-3


def gpu_run(pop, params, seeds, days):
    g = importlib.import_module("covid-sim_amd.engine").Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(seeds)
    g.reset(10)
    tspd = int(params.time_steps_per_day)
    g.run_steps(0.0, days * tspd, tspd)
    itype = np.array([list(g.counters(r).cumItype) for r in range(len(seeds))], float)
    gpu_run.starts = start_times(g, len(seeds))
    return g.timeseries(days), itype


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["mini", "mini_care", "mini_int"])
def test_gpu_ensemble_is_the_oracle_ensemble_and_matches_the_reference(name):
    """The GPU's 200 realisations, advanced together by the batched run,
    reproduce the oracle's 200 bit for bit, and pass the ensemble tests
    against the reference's 200."""
    pop, params, ref, ref_it = workload(name)
    orc, orc_it = oracle_ensemble(name)
    gpu, gpu_it = gpu_run(pop, params, SEEDS, ref.shape[1])
    assert gpu.shape == orc.shape and np.array_equal(gpu, orc)
    assert np.array_equal(gpu_it, orc_it)
    check_against_reference(ref, ref_it, gpu, gpu_it, pop.n_people)
    if name == "mini_int":
        assert np.array_equal(gpu_run.starts, oracle_ensemble.starts[name])
        check_start_times(workload.starts, gpu_run.starts)


@pytest.mark.gpu
def test_gpu_c1_ensemble_matches_the_reference():
    """C1 (200k persons, 200 days): the GPU's 200 realisations against the
    reference's 200, and its first two realisations equal to the oracle's."""
    pop, params, ref, ref_it = workload("c1")
    days = ref.shape[1]
    gpu, gpu_it = gpu_run(pop, params, SEEDS, days)
    check_against_reference(ref, ref_it, gpu, gpu_it, pop.n_people)
    o = OracleEngine(params, pop)
    o.alloc(SEEDS[:2])
    o.reset(10)
    tspd = int(params.time_steps_per_day)
    o.run_steps(0.0, days * tspd, tspd)
    assert np.array_equal(gpu[:2], o.timeseries(days))
