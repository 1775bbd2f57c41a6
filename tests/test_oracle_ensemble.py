"""The oracle's stochastic trajectories against the reference's, as ensembles.

The random streams differ by design (counter-based streams replace the
reference's per-thread generators), so single trajectories cannot match;
the distribution of trajectories must.  The reference ensemble
(tests/golden/mini_ensemble.npz) is 200 realisations of the unmodified
reference on the 20k-person fixture region (tests/golden/make_ensemble.py);
the oracle runs 200 realisations on the same population and parameters.

Compared: attack rate, peak day, peak daily incidence, cumulative cases and
deaths, each with a two-sample Kolmogorov-Smirnov test (p > 0.01) and
agreement of the two ensembles' 95% confidence intervals for the mean (the
intervals overlap); and the daily infections and cases of every day with a
per-day KS test,
Bonferroni-corrected over the days (family-wise level 0.01; 120
uncorrected tests at 0.01 would reject a true null in 70% of suites).
A 1000-versus-1000 run of the same comparison (not part of the suite)
gave |z| < 1.8 for every statistic.

On the GPU the same 200 seeds, advanced together by one batched run, must
give the oracle's 200 trajectories bit for bit and pass the same tests
against the reference's 200.
"""
import ctypes as C
import functools
import os

import numpy as np
import pytest
from scipy import stats

from fixtures import GOLDEN, mini
from oracle_lib import OracleEngine

N_RUNS = 200
ALPHA = 0.01


@functools.lru_cache(maxsize=None)
def ensembles():
    ref = np.load(os.path.join(GOLDEN, "mini_ensemble.npz"))["daily"]
    arrays, pop, params, _ = mini()
    days = ref.shape[1]
    o = OracleEngine(params, pop)
    o.alloc(np.arange(1, N_RUNS + 1, dtype=np.uint64) * np.uint64(2654435761))
    o.reset(10)
    o.run_steps(0.0, days * int(params.time_steps_per_day), int(params.time_steps_per_day))
    ensembles.itype = np.array([list(o.counters(r).cumItype) for r in range(N_RUNS)], float)
    return ref, o.timeseries(days), pop.n_people


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


def test_population_conserved():
    ref, orc, n = ensembles()
    assert (orc[:, :, :5].sum(axis=2) == n).all()
    assert (ref[:, :, :5].sum(axis=2) == n).all()


@pytest.mark.parametrize("name", SUMMARIES)
def test_summary_statistics(name):
    ref, orc, n = ensembles()
    idx = SUMMARIES.index(name)
    a, b = summary(ref, n)[idx].astype(float), summary(orc, n)[idx].astype(float)
    assert stats.ks_2samp(a, b).pvalue > ALPHA, (a.mean(), b.mean())
    assert cis_overlap(a, b), (a.mean(), b.mean())


@pytest.mark.parametrize("field,name", [(5, "infections"), (6, "cases")])
def test_daily_incidence_per_day_ks(field, name):
    ref, orc, n = ensembles()
    days = ref.shape[1]
    pv = np.array([stats.ks_2samp(ref[:, d, field], orc[:, d, field]).pvalue for d in range(days)])
    assert pv.min() > ALPHA / days, (name, pv.argmin(), pv.min())


def test_infection_routes():
    """Final infections by route (household, within-group and between-group
    place per place type, spatial; StateT.cumItype) agree in mean: each
    route's oracle mean lies within a Welch 99.9% interval of the
    reference's."""
    ref_it = np.load(os.path.join(GOLDEN, "mini_ensemble.npz"))["itype"].astype(float)
    ensembles()
    orc_it = ensembles.itype
    for q in range(1, 15):  # 0 unused, 15 = seeds (cleared in the reference before the first sample)
        if ref_it[:, q].sum() == 0 and orc_it[:, q].sum() == 0:
            continue
        assert welch_ci_contains_zero(ref_it[:, q], orc_it[:, q], level=0.999), (q, ref_it[:, q].mean(), orc_it[:, q].mean())


@pytest.mark.gpu
def test_gpu_ensemble_is_the_oracle_ensemble_and_matches_the_reference():
    """The GPU's 200 realisations, advanced together by the batched run,
    reproduce the oracle's 200 bit for bit, and pass the same ensemble tests
    against the reference's 200 (summaries by KS and overlapping 95% CIs,
    daily incidence by per-day KS)."""
    import importlib
    ref, orc, n = ensembles()
    arrays, pop, params, _ = mini()
    g = importlib.import_module("covid-sim_amd.engine").Engine(0)
    g.set_params(params)
    g.set_population(pop)
    g.alloc(np.arange(1, N_RUNS + 1, dtype=np.uint64) * np.uint64(2654435761))
    g.reset(10)
    days = ref.shape[1]
    tspd = int(params.time_steps_per_day)
    g.run_steps(0.0, days * tspd, tspd)
    gpu = g.timeseries(days)
    assert gpu.shape == orc.shape and np.array_equal(gpu, orc)
    for idx, name in enumerate(SUMMARIES):
        a, b = summary(ref, n)[idx].astype(float), summary(gpu, n)[idx].astype(float)
        assert stats.ks_2samp(a, b).pvalue > ALPHA and cis_overlap(a, b), (name, a.mean(), b.mean())
    for field in (5, 6):
        pv = np.array([stats.ks_2samp(ref[:, d, field], gpu[:, d, field]).pvalue for d in range(days)])
        assert pv.min() > ALPHA / days, (field, pv.argmin(), pv.min())
