"""Calibration (covid-sim_amd/calibrate.py).

CPU: the likelihood is CalcLikelihood's negative binomial term
(src/CovidSim.cpp:3778-3784), checked against scipy's negative binomial
pmf with the same parameters; the pattern search finds the maximum of a
smooth test function; the crest refinement recovers the maximum of a
sheared quadratic ridge from starts anywhere in its basin, and starts across
the basin return the same fit on rough ridges; both calibration drivers run
end to end on a stand-in simulator and agree on R0.
GPU: on the 20k-person fixture region, an emulator trained on simulator
ensembles guides calibration to a fit inside plain calibration's 95%
likelihood region, with fewer simulator launches -- what the emulator gains
over the method without it (bench.py --calibration measures the speed-up
and the agreement of R0 on the 1M-person region).
"""
import importlib

import numpy as np
import pytest
from scipy import stats

cal = importlib.import_module("covid-sim_amd.calibrate")


def test_nb_loglik_is_negative_binomial():
    rng = np.random.default_rng(0)
    data = rng.integers(0, 60, 365).astype(float)
    model = rng.uniform(0, 50, 365)
    k = 10.0
    knb = 1 + model / k
    pnb = k / (1 + k)
    want = stats.nbinom.logpmf(data, knb, 1 - pnb).sum()
    assert abs(cal.nb_loglik(data, model, k) - want) < 1e-8 * abs(want)


def test_pattern_search_finds_maximum():
    f = lambda th: -((th[:, 0] - 2.13) ** 2) * 40 - (th[:, 1] + 0.21) ** 2 * 9
    x, fx, n = cal.pattern_search(f, (2.4, 0.0), (0.4, 0.4))
    assert abs(x[0] - 2.13) < 0.01 and abs(x[1] + 0.21) < 0.02


def ridge(th, r0=2.16, ls=0.17, slope=-3.25):
    """A sheared quadratic ridge like the calibration likelihood's: steep
    across the crest line ls + slope (R0 - r0), flat along it."""
    u = th[:, 0] - r0
    v = th[:, 1] - (ls + slope * u)
    return -567.0 - 1300.0 * u * u - 5000.0 * v * v


@pytest.mark.parametrize("start", [(2.16, 0.17, -3.25), (2.08, 0.45, -2.5), (2.26, -0.15, -4.0), (2.2, 0.2, -3.0)])
def test_crest_profile_finds_the_maximum_of_a_ridge(start):
    """On a sheared quadratic ridge the column parabolas and the profile
    parabola are exact, so the fit is the maximum itself, from any start in
    the basin."""
    x, fx, k, n = cal.crest_profile(ridge, *start)
    assert abs(x[0] - 2.16) < 1e-9 and abs(x[1] - 0.17) < 1e-9, (x, k, n)
    assert k <= 5


@pytest.mark.parametrize("amp", [1.0, 2.0, 4.0])
def test_crest_profile_agrees_across_starts_on_rough_ridges(amp):
    """With roughness of `amp` log-likelihood units (a 16-seed mean's is
    about 2), starts across the basin settle on the same lattice window and
    return the same fit -- what lets two calibrations agree; at twice that
    roughness neighbouring windows can both be fixed points, and the fits
    still agree within 0.01.  Every fit is near the smooth ridge's maximum."""
    rough = lambda th: ridge(th) + amp * np.sin(97.0 * th[:, 0]) * np.cos(41.0 * th[:, 1])
    fits = np.array([cal.crest_profile(rough, *st)[0] for st in
                     [(2.161, 0.17, -3.1), (2.19, 0.07, -3.4), (2.12, 0.3, -3.0), (2.2, 0.05, -3.75), (2.14, 0.21, -2.75)]])
    if amp <= 2.0:
        assert (fits == fits[0]).all(), fits
    assert np.ptp(fits[:, 0]) < 0.01, fits
    assert np.abs(fits[:, 0] - 2.16).max() < 0.01, fits


def test_crest_from_grid():
    r0 = np.linspace(1.8, 3.0, 13)
    ls = np.linspace(-0.7, 0.7, 8)
    th = np.array([(a, b) for a in r0 for b in ls])
    c, l, slope = cal.crest_from_grid(r0, ls, ridge(th).reshape(len(r0), len(ls)))
    assert abs(c - 2.2) < 1e-9  # the grid row nearest the maximum
    assert abs(l - (0.17 - 3.25 * 0.04)) < 0.05 and abs(slope + 3.25) < 0.5


class CurveSim:
    """Stand-in simulator: a deterministic death curve whose timing trades
    R0 against the seeding scale (the likelihood ridge of the real one)."""

    launches = 0
    realisations = 0

    def run(self, theta, seeds):
        theta = np.atleast_2d(theta)
        self.launches += -(-len(theta) * len(seeds) // 64)
        self.realisations += len(theta) * len(seeds)
        t = np.arange(cal.DAYS)[None, :]
        r = (theta[:, :1] - 1.0) / 5.0                  # growth rate per day
        peak = 40.0 + (np.log(1e4) - theta[:, 1:] * np.log(10.0) * 1.0) / r
        width = 1.2 / r
        deaths = 6000.0 * (theta[:, :1] - 1.0) / 1.15 * np.exp(-0.5 * ((t - peak) / width) ** 2) / (width * 2.5066)
        return deaths * 0.5, deaths


class ExactEmulator:
    def forward(self, x):
        th = np.column_stack([x[:, 0] * 0.35 + 2.4, x[:, 1] * 0.4])
        inc, d = CurveSim().run(th, [0])
        return cal.from_curves(inc, d)


def test_calibrations_agree_on_a_stand_in_simulator():
    """Both drivers run end to end on the stand-in: the plain grid and the
    emulator's dense grid start the same crest refinement, which returns
    the same fit; the emulator-guided one uses far fewer launches."""
    sim = CurveSim()
    target = np.round(sim.run(np.array([[2.15, 0.25]]), [0])[1][0])
    seeds = np.arange(16)
    plain = cal.calibrate_plain(sim, target, seeds)
    emul = cal.calibrate_emulated(sim, ExactEmulator(), target, seeds)
    assert abs(plain["theta"][0] - 2.15) < 0.02 and abs(emul["theta"][0] - 2.15) < 0.02, (plain, emul)
    assert abs(plain["theta"][0] - emul["theta"][0]) < 0.02
    assert plain["realisations"] >= 5 * emul["realisations"], (plain["realisations"], emul["realisations"])


def test_encoding_round_trip():
    inc = np.array([[0.0, 1.0, 50.0] + [3.0] * 362])
    d = np.array([[0.0, 0.0, 2.0] + [1.0] * 362])
    i2, d2 = cal.to_curves(cal.from_curves(inc, d))
    assert np.allclose(i2, inc, rtol=1e-5) and np.allclose(d2, d, atol=1e-4)


@pytest.mark.gpu
def test_emulator_guided_calibration_uses_fewer_launches():
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from fixtures import mini
    csa = importlib.import_module("covid-sim_amd")
    arrays, pop, params, _ = mini()

    class MiniSim(cal.Simulator):
        def __init__(self):
            eng_mod = importlib.import_module("covid-sim_amd.engine")
            self.pop, self.params = pop, params
            self.eng = eng_mod.Engine(0)
            self.eng.set_params(params)
            self.eng.set_population(pop)
            c = self.eng.transmission_coeffs(77)
            self.r0_hh, self.r0_pl, self.sp_r0 = c.r0_household, c.r0_places, c.spatial_r0
            self.n_real, self.launches, self.realisations, self.sim_seconds = 64, 0, 0, 0.0

    sim = MiniSim()
    rng = np.random.default_rng(1)
    th, y = cal.training_set(sim, 256, np.arange(1000, 1004, dtype=np.uint64), rng)
    emu, losses = cal.train_emulator(cal.encode(th), y, epochs=600, batch=64)
    assert losses[-1] < 0.1 * losses[0]
    _, d = sim.run(np.array([[2.2, 0.1]]), np.arange(5000, 5016, dtype=np.uint64))
    target = np.round(d[0])
    # 32 seeds per candidate: the deaths of a 20k-person region leave a flat
    # ridge of likelihood along R0 x seeding scale, rough at 16 seeds
    seeds = np.arange(100, 132, dtype=np.uint64)
    plain = cal.calibrate_plain(sim, target, seeds)
    emul = cal.calibrate_emulated(sim, emu, target, seeds)
    # the two fits are compared by the simulator's likelihood at each (same
    # seeds): the emulator-guided fit lies in the plain fit's 95% likelihood
    # region for two parameters (within chi2(2, 0.95) / 2 = 3.0 units)
    f = cal.nb_loglik(target, sim.run(np.array([plain["theta"], emul["theta"]]), seeds)[1])
    assert f[1] >= f[0] - 3.0, (plain, emul, f)
    assert emul["launches"] < plain["launches"], (plain, emul)
