"""Calibration of R0 and the seeding scale to a death curve, with and
without the emulator.

The reference calibrates by re-running the simulator: RunModel's
re-run-to-trigger loop rescales P.SeedingScaling (src/CovidSim.cpp:3485-3689)
and CalcLikelihood scores a run against daily deaths with a negative
binomial likelihood (src/CovidSim.cpp:3691-3852).  Here every candidate
(R0, seeding scale) is scored with that likelihood on the mean daily deaths
of a GPU ensemble (16 seeds, common random numbers across candidates), and
two searches are compared:

  plain      a coarse grid over the whole (R0, log10 seeding scale) box,
             four candidates per 64-realisation launch, then the crest
             refinement (crest_profile) started from the grid's crest;
  emulator   the trained emulator scores a dense grid in one batched forward
             pass; the simulator verifies and refines the crest it proposes
             with the same refinement.

The likelihood has a long ridge (more seeds with a lower R0 give a similar
death curve), so a local axis search from a default start stops early on
the ridge (calibrate_pattern shows it); the plain calibration needs its
global stage, and the emulator replaces exactly that stage.

Per-realisation parameters: LocalBeta from R0 through the reproduction
numbers InitTransmissionCoeffs gives for the population
(src/SetupModel.cpp:832-838: LocalBeta = (R0 - R0household - R0places) /
spatial R0), and the seed count (int)(10 * scale + 0.5)
(src/CovidSim.cpp:1213).

Emulator inputs are 10 numbers: the normalised R0 and log10 seeding scale,
then 8 inputs for intervention fractions and trigger thresholds that the
no-intervention simulator does not read (held at 0).  Outputs are 730
numbers: log1p of mean daily infections and of mean daily deaths over 365
days, scaled by 1/8.
"""
import math
import time

import numpy as np
from scipy.special import gammaln

N_IN, N_HIDDEN, N_OUT = 10, 512, 730
DAYS = 365
R0_RANGE = (1.8, 3.0)
LOGS_RANGE = (-0.7, 0.7)   # log10 seeding scale
OUT_SCALE = 8.0


NEG_BIN_K = 1.0  # CalcLikelihood's NegBinK (read from its data file there); 1: no over-dispersion


def nb_loglik(data, model, k=NEG_BIN_K):
    """CalcLikelihood's negative binomial term for one column of death counts
    (src/CovidSim.cpp:3778-3784), summed over days."""
    data = np.asarray(data, np.float64)
    model = np.asarray(model, np.float64)
    knb = 1.0 + model / k
    pnb = k / (1.0 + k)
    return np.sum(gammaln(data + knb) - gammaln(data + 1) - gammaln(knb) + knb * math.log(1.0 - pnb) + data * math.log(pnb), axis=-1)


def encode(theta):
    """(R0, log10 scale) rows -> emulator inputs [n, 10]."""
    theta = np.atleast_2d(np.asarray(theta, np.float64))
    x = np.zeros((len(theta), N_IN), np.float32)
    x[:, 0] = (theta[:, 0] - 2.4) / 0.35
    x[:, 1] = theta[:, 1] / 0.4
    return x


def to_curves(y):
    y = np.maximum(np.asarray(y, np.float64) * OUT_SCALE, 0.0)
    c = np.expm1(y)
    return c[:, :DAYS], c[:, DAYS:]


def from_curves(inc, deaths):
    return (np.log1p(np.concatenate([inc, deaths], axis=1)) / OUT_SCALE).astype(np.float32)


class Simulator:
    """Batched ensembles of one population; 64 realisations per launch."""

    progress = None  # optional callable(str), called every 8 launches

    def __init__(self, n_people=1_000_000, pop_seed=2024, n_real=64, itc_seed=77):
        import importlib
        import os
        csa = importlib.import_module("covid-sim_amd")
        eng_mod = importlib.import_module("covid-sim_amd.engine")
        popgen = importlib.import_module("covid-sim_amd.popgen")
        here = os.path.dirname(os.path.abspath(__file__))
        arrays, geom = popgen.generate(n_people, seed=pop_seed)
        self.pop = csa.population.Population(arrays)
        self.params = csa.population.load_params(os.path.join(here, "inputs", "c1", "params.bin"))
        popgen.apply_geometry(self.params, geom)
        self.eng = eng_mod.Engine(0)
        self.eng.set_params(self.params)
        self.eng.set_population(self.pop)
        c = self.eng.transmission_coeffs(itc_seed)
        self.r0_hh, self.r0_pl, self.sp_r0 = c.r0_household, c.r0_places, c.spatial_r0
        self.n_real = n_real
        self.launches = 0
        self.realisations = 0
        self.sim_seconds = 0.0

    def beta(self, r0):
        return np.maximum(0.0, (np.asarray(r0, np.float64) - self.r0_hh - self.r0_pl) / self.sp_r0)

    def run(self, theta, seeds):
        """Mean daily infections and deaths [n_theta, DAYS] over `seeds`
        (the same seeds for every candidate)."""
        theta = np.atleast_2d(np.asarray(theta, np.float64))
        seeds = np.asarray(seeds, np.uint64)
        jobs = [(i, s) for i in range(len(theta)) for s in seeds]
        inc = np.zeros((len(theta), DAYS))
        dth = np.zeros((len(theta), DAYS))
        tspd = int(round(self.params.time_steps_per_day))
        for b0 in range(0, len(jobs), self.n_real):
            batch = jobs[b0:b0 + self.n_real]  # the last launch holds only what is left
            t0 = time.perf_counter()
            self.eng.alloc(np.array([s for _, s in batch], np.uint64))
            th = theta[[i for i, _ in batch]]
            nseed = np.maximum(1, (10.0 * 10.0 ** th[:, 1] + 0.5).astype(np.int32))
            self.eng.set_realisation_params(self.beta(th[:, 0]), nseed)
            self.eng.reset(10)
            self.eng.run_steps(0.0, DAYS * tspd, tspd)
            ts = self.eng.timeseries(DAYS)
            self.sim_seconds += time.perf_counter() - t0
            self.launches += 1
            self.realisations += len(batch)
            if self.progress and self.launches % 8 == 0:
                self.progress("%d simulator launches, %.1f s" % (self.launches, self.sim_seconds))
            for k, (i, _) in enumerate(batch):
                inc[i] += ts[k, :, 5]
                dth[i] += ts[k, :, 7]
        return inc / len(seeds), dth / len(seeds)


def pattern_search(score, start, step, min_step=0.01, max_evals=400, fstart=None):
    """Maximise score over (R0, log10 scale): evaluate the four axis
    neighbours of the incumbent together (one simulator launch), move to the
    best improvement, else halve both steps; stop when the R0 step falls
    below min_step.  fstart: the score of start when already known."""
    box = np.array([R0_RANGE, LOGS_RANGE])
    x = np.clip(np.asarray(start, np.float64), box[:, 0], box[:, 1])
    fx = score(x[None])[0] if fstart is None else fstart
    step = np.asarray(step, np.float64)
    evals = 1
    while step[0] >= min_step and evals < max_evals:
        cand = np.array([x + [step[0], 0], x - [step[0], 0], x + [0, step[1]], x - [0, step[1]]])
        cand = np.clip(cand, box[:, 0], box[:, 1])
        f = score(cand)
        evals += len(cand)
        j = int(np.argmax(f))
        if f[j] > fx:
            x, fx = cand[j], f[j]
        else:
            step = step / 2
    return x, fx, evals


# Local refinement shared by both calibrations.  The log-likelihood has a
# long diagonal ridge in (R0, log10 scale): more seeds with a lower R0 give a
# similar death curve.  Along the crest it is flat (within ~3 units over
# +-0.05 in R0) and rough at the 1-2 unit level (16-seed means, integer seed
# counts); it is not quadratic over a design wide enough to average the
# roughness out.  So the refinement fits the two directions apart: across the
# crest a parabola through three points of each R0 column gives the column's
# maximum, and along it a parabola through the column maxima gives R0.
# Every candidate lies on one lattice (R0 step 0.025, log10 scale step
# 0.05), and the window of columns re-centres on the lattice column nearest
# its fitted maximum until it stays: with common random numbers the
# simulator's likelihood is a deterministic function, so two calibrations
# that start in the same basin evaluate the same final window and return the
# same fit.
LAT_R0, LAT_LS = 0.025, 0.05
PROFILE_HALF = 2   # columns on each side of the window's centre


def crest_profile(score, r0, ls, slope, half=PROFILE_HALF, max_windows=8):
    """The crest refinement from a crest estimate: ls + slope (R0 - r0).
    Returns (R0, log10 scale) of the fitted maximum, the best evaluated
    log-likelihood, the windows fitted and the candidates evaluated."""
    cache = {}

    def ev(keys):  # lattice indices (i, j): R0 = i LAT_R0, log10 scale = j LAT_LS
        new = [k for k in dict.fromkeys(keys) if k not in cache]
        if new:
            f = score(np.array([(i * LAT_R0, j * LAT_LS) for i, j in new]))
            cache.update(zip(new, (float(x) for x in f)))

    ci = int(round(r0 / LAT_R0))
    seen = set()
    est = (r0, ls)
    for k in range(1, max_windows + 1):
        cols = list(range(ci - half, ci + half + 1))
        js = {i: int(round((ls + slope * (i * LAT_R0 - r0)) / LAT_LS)) for i in cols}
        for _ in range(6):  # bracket every column's maximum on the lattice, all columns per batch
            ev([(i, js[i] + d) for i in cols for d in (-1, 0, 1)])
            moved = False
            for i in cols:
                y = [cache[(i, js[i] + d)] for d in (-1, 0, 1)]
                if not (y[1] >= y[0] and y[1] >= y[2]):
                    js[i] += 1 if y[2] > y[1] else -1
                    moved = True
            if not moved:
                break
        prof = []
        for i in cols:
            y0, y1, y2 = (cache[(i, js[i] + d)] for d in (-1, 0, 1))
            den = y0 - 2 * y1 + y2
            off = 0.5 * (y0 - y2) / den if den < 0 else 0.0
            prof.append((i * LAT_R0, (js[i] + off) * LAT_LS, y1 - 0.25 * (y0 - y2) * off))
        prof = np.array(prof)
        p = np.polyfit(prof[:, 0] - ci * LAT_R0, prof[:, 2], 2)
        u = -p[1] / (2 * p[0]) if p[0] < 0 else prof[int(np.argmax(prof[:, 2])), 0] - ci * LAT_R0
        u = float(np.clip(u, -half * LAT_R0, half * LAT_R0))
        r0 = ci * LAT_R0 + u
        q = np.polyfit(prof[:, 0], prof[:, 1], 1)  # the crest through the columns' maxima
        ls, slope = float(np.polyval(q, r0)), float(q[0])
        est = (r0, ls)
        ni = int(round(r0 / LAT_R0))
        if ni == ci or ni in seen:
            break
        seen.add(ci)
        ci = ni
    return np.array(est), max(cache.values()), k, len(cache)


def crest_from_grid(r0, ls, ll, around=None, halfwidth=0.1):
    """The crest of a log-likelihood grid ll[len(r0), len(ls)]: for each R0,
    the maximising log10 scale (parabola through the best three); returns the
    best R0 (or `around`), its crest log10 scale and the crest slope fitted
    over R0 within halfwidth."""
    cr = []
    for i in range(len(r0)):
        j = int(np.clip(np.argmax(ll[i]), 1, len(ls) - 2))
        p = np.polyfit(ls[j - 1:j + 2], ll[i, j - 1:j + 2], 2)
        x = -p[1] / (2 * p[0]) if p[0] < 0 else ls[j]
        x = float(np.clip(x, ls[j - 1], ls[j + 1]))
        cr.append((x, float(np.polyval(p, x)) if p[0] < 0 else float(ll[i, j])))
    cr = np.array(cr)
    i0 = int(np.argmax(cr[:, 1])) if around is None else int(np.argmin(np.abs(r0 - around)))
    sel = np.abs(r0 - r0[i0]) <= halfwidth + 1e-9
    slope = np.polyfit(r0[sel], cr[sel, 0], 1)[0] if sel.sum() >= 2 else 0.0
    return float(r0[i0]), float(cr[i0, 0]), float(slope)


def _score(sim, target_deaths, seeds):
    return lambda th: nb_loglik(target_deaths, sim.run(th, seeds)[1])


def calibrate_plain(sim, target_deaths, seeds, grid=(13, 8)):
    """Simulator only: a coarse grid over the whole box (R0 step 0.1, log10
    scale step 0.2), then the crest refinement from the grid's crest."""
    t0 = time.perf_counter()
    l0, r_0 = sim.launches, sim.realisations
    score = _score(sim, target_deaths, seeds)
    r0 = np.linspace(*R0_RANGE, grid[0])
    ls = np.linspace(*LOGS_RANGE, grid[1])
    f = score(np.array([(a, b) for a in r0 for b in ls])).reshape(len(r0), len(ls))
    c0, l0s, slope = crest_from_grid(r0, ls, f, halfwidth=0.1)
    x, fx, k, n = crest_profile(score, c0, l0s, slope)
    return {"theta": x.tolist(), "loglik": fx, "start": [c0, l0s, slope], "candidates": f.size + n,
            "windows": k, "launches": sim.launches - l0, "realisations": sim.realisations - r_0,
            "seconds": time.perf_counter() - t0}


def calibrate_pattern(sim, target_deaths, seeds, start=(2.4, 0.0), step=(0.4, 0.4)):
    """Simulator only, local: the axis pattern search from a default start
    (stops on the likelihood ridge; kept to show why a global stage is needed)."""
    t0 = time.perf_counter()
    l0 = sim.launches
    x, fx, n = pattern_search(_score(sim, target_deaths, seeds), start, step)
    return {"theta": x.tolist(), "loglik": float(fx), "candidates": n, "launches": sim.launches - l0,
            "seconds": time.perf_counter() - t0}


def calibrate_emulated(sim, emu, target_deaths, seeds, grid=(121, 57)):
    """The emulator scores a dense grid (R0 step 0.01) in one batched forward
    pass and proposes the crest around its best candidate; the simulator
    verifies and refines the proposal with the same crest refinement as the
    plain calibration (more designs when the proposal is off: the plain
    simulator as fallback)."""
    t0 = time.perf_counter()
    l0, r_0 = sim.launches, sim.realisations
    r0 = np.linspace(*R0_RANGE, grid[0])
    ls = np.linspace(*LOGS_RANGE, grid[1])
    th = np.array([(a, b) for a in r0 for b in ls])
    _, d_pred = to_curves(emu.forward(encode(th)))
    ll = nb_loglik(target_deaths, d_pred).reshape(len(r0), len(ls))
    c0, l0s, slope = crest_from_grid(r0, ls, ll, halfwidth=0.1)
    t_emul = time.perf_counter() - t0
    x, fx, k, n = crest_profile(_score(sim, target_deaths, seeds), c0, l0s, slope)
    return {"theta": x.tolist(), "loglik": fx, "proposal": [c0, l0s, slope], "candidates": n,
            "windows": k, "launches": sim.launches - l0, "realisations": sim.realisations - r_0,
            "seconds": time.perf_counter() - t0, "emulator_seconds": t_emul}


def training_set(sim, n_theta, seeds, rng):
    """Latin-hypercube candidates and their ensemble mean curves."""
    u = (rng.permuted(np.tile(np.arange(n_theta), (2, 1)), axis=1).T + rng.random((n_theta, 2))) / n_theta
    th = np.column_stack([R0_RANGE[0] + u[:, 0] * (R0_RANGE[1] - R0_RANGE[0]),
                          LOGS_RANGE[0] + u[:, 1] * (LOGS_RANGE[1] - LOGS_RANGE[0])])
    inc, dth = sim.run(th, seeds)
    return th, from_curves(inc, dth)


def train_emulator(x, y, epochs=1500, batch=128, lr=1e-3, seed=5):
    """Minibatch Adam on the GPU; the learning rate drops tenfold for the
    last third of the epochs."""
    import importlib
    emu = importlib.import_module("covid-sim_amd.emulator").Emulator(N_IN, N_HIDDEN, N_OUT, seed=seed)
    emu.set_data(x, y)
    e1 = (2 * epochs) // 3
    losses = np.concatenate([emu.train(e1, batch, lr, seed=seed), emu.train(epochs - e1, batch, lr / 10, seed=seed + 1)])
    return emu, losses


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def emulator_error(emu, th, y, y_other=None):
    """Relative L2 error of the emulator's curves against the simulator's
    ensemble means, for infections and deaths (held-out candidates).  With
    y_other, a second simulator estimate from independent seeds, also the
    same error between the two simulator estimates: the noise floor of a
    mean over that many seeds."""
    pi, pd = to_curves(emu.forward(encode(th)))
    ti, td = to_curves(y)
    out = {"infections": rel_l2(pi, ti), "deaths": rel_l2(pd, td)}
    if y_other is not None:
        oi, od = to_curves(y_other)
        out["simulator_noise_infections"] = rel_l2(oi, ti)
        out["simulator_noise_deaths"] = rel_l2(od, td)
    return out
