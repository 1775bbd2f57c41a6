"""Calibration (covid-sim_amd/calibrate.py).

CPU: the likelihood is CalcLikelihood's negative binomial term
(src/CovidSim.cpp:3778-3784), checked against scipy's negative binomial
pmf with the same parameters; the pattern search finds the maximum of a
smooth test function; the crest refinement recovers the maximum and slope
of a sheared quadratic ridge from starts anywhere in its basin, and returns
identical fits from starts that snap to the same design.
GPU: on the 20k-person fixture region, an emulator trained on simulator
ensembles guides calibration to a simulator-verified likelihood within 2
log-likelihood units of plain calibration's with fewer simulator launches
-- what the emulator gains over the method without it.
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


def test_crest_fit_recovers_maximum_and_slope():
    rng = np.random.default_rng(3)
    z = rng.uniform(-0.12, 0.12, (15, 2))
    c0 = np.array([2.1, 0.4])
    pts = c0 + z
    got = cal.crest_fit(z, ridge(pts))
    assert got is not None
    (u, v), slope = got
    assert abs(c0[0] + u - 2.16) < 1e-9 and abs(c0[1] + v - (0.17 - 3.25 * (c0[0] + u - 2.16))) < 1e-9
    # in unsheared coordinates the crest slope is the ridge's own
    assert abs(slope + 3.25) < 1e-9
    # a convex fit is refused
    assert cal.crest_fit(z, -ridge(pts)) is None


@pytest.mark.parametrize("start", [(2.16, 0.17, -3.25), (2.08, 0.45, -2.5), (2.26, -0.15, -4.0), (2.2, 0.2, -3.0)])
def test_refine_converges_on_the_crest_from_any_start_in_the_basin(start):
    x, fx, k = cal.refine(ridge, *start)
    assert abs(x[0] - 2.16) < 1e-6 and abs(x[1] - 0.17) < 1e-6, (x, k)
    assert k <= 3


def test_refine_is_deterministic_on_its_lattice():
    """Starts in the same basin end on the same snapped design, so they
    evaluate the same final candidates and return the same fit, even on a
    rough objective."""
    rough = lambda th: ridge(th) + 2.0 * np.sin(97.0 * th[:, 0]) * np.cos(41.0 * th[:, 1])
    a = cal.refine(rough, 2.161, 0.17, -3.1)
    b = cal.refine(rough, 2.170, 0.19, -3.4)
    assert np.array_equal(a[0], b[0]), (a, b)
    assert abs(a[0][0] - 2.16) < 0.01


def test_crest_from_grid():
    r0 = np.linspace(1.8, 3.0, 13)
    ls = np.linspace(-0.7, 0.7, 8)
    th = np.array([(a, b) for a in r0 for b in ls])
    c, l, slope = cal.crest_from_grid(r0, ls, ridge(th).reshape(len(r0), len(ls)))
    assert abs(c - 2.2) < 1e-9  # the grid row nearest the maximum
    assert abs(l - (0.17 - 3.25 * 0.04)) < 0.05 and abs(slope + 3.25) < 0.5


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
            self.n_real, self.launches, self.sim_seconds = 64, 0, 0.0

    sim = MiniSim()
    rng = np.random.default_rng(1)
    th, y = cal.training_set(sim, 128, np.arange(1000, 1004, dtype=np.uint64), rng)
    emu, losses = cal.train_emulator(cal.encode(th), y, epochs=200, batch=64)
    assert losses[-1] < 0.1 * losses[0]
    _, d = sim.run(np.array([[2.2, 0.1]]), np.arange(5000, 5016, dtype=np.uint64))
    target = np.round(d[0])
    seeds = np.arange(100, 116, dtype=np.uint64)
    plain = cal.calibrate_plain(sim, target, seeds)
    emul = cal.calibrate_emulated(sim, emu, target, seeds)
    # deaths of a 20k-person region leave a ridge of near-equal likelihood
    # along R0 x seeding scale, so the two searches are compared by the
    # simulator-verified likelihood they reach, not by position on the ridge
    assert emul["loglik"] >= plain["loglik"] - 2.0, (plain, emul)
    assert emul["launches"] < plain["launches"], (plain, emul)
