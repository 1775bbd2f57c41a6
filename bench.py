"""Benchmark: simulated person-days/s of a batched ensemble on one MI355X.

Workload (the ensemble the headline metric is quoted on, at the size the
target names): 64 realisations of one 5,000,000-person synthetic region
advance together, 300 days at 4 timesteps/day, from a reset with 10 random
seed infections each.  A "step" is one complete ensemble run (reset + 1200
timesteps of InfectSweep + IncubRecoverySweep + the S-drop UpdateProbs
check); `value` = realisations x persons x days / wall seconds of the timed
steps, with the population and tables already resident in HBM.

Parameters are the 200k-person region's derived parameter set
(covid-sim_amd/inputs/c1/params.bin); the population comes from
covid-sim_amd/popgen.py at the same density and the region's geometry.

The timed steps replay one captured HIP graph per output step (--no-graphs:
direct launches).  The reported roofline is for the kernel with the largest
device time, from HIP events recorded around each of its launches on the
library's stream during one more, profiled pass of the same workload with
direct launches; its algorithmic bytes per launch come from the event
counts the library keeps and the per-event byte costs below.

The CPU baseline runs the reference itself (oracle/_ref/covidsim_ref, built
from the reference sources) on its 200k-person workload with all host
threads, for a bounded number of realisations.

--config c3 prints the line of the uneven-regions ensemble instead: 64
realisations over five generated regions of 1M, 2M, 5M, 10M and 20M persons
(13/13/13/13/12 realisations), each realisation with its own R0 (LocalBeta
from one InitTransmissionCoeffs pass per region) and intervention trigger
threshold (covid-sim_amd/inputs/c3/realisations.json), the intervention
subset on (covid-sim_amd/inputs/c3/params.bin), 300 days; the regions advance
as one batch (every output step, each region's step graph launched on its
own stream before any wait, so the regions run side by side); `value` = the
sum over realisations of the region's persons x days / wall seconds.

--calibration prints a second kind of line instead: the wall time of
fitting R0 and the seeding scale to a death curve with the emulator
(covid-sim_amd/calibrate.py) against plain simulator-only calibration on
the same GPU, both verified by the simulator, with the emulator's training
cost and held-out error.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# minimum bytes touched per event (the per-event costs of the step's
# algorithmic-bytes formula); fixed, never tuned to the measurement
BYTES = {"n_inf": 64, "n_hh_inf": 16, "n_hh_member": 40, "n_hh_scan": 40, "n_place_link": 32, "n_place_cand": 44,
         "n_sp_draw": 48, "n_sp_susc": 40, "n_new": 40, "n_active": 16, "n_transition": 30,
         "n_cell_steps": 32, "n_sp_scan": 4, "n_sp_probe": 4}
# the events each kernel's bytes are counted from.  The household kernel is
# charged 40 B only per member its scans actually read (n_hh_scan): a
# household already found to hold no susceptible member is skipped, and its
# infector then reads no household record either, so neither the skipped
# members (n_hh_member counts them) nor a per-household 16 B is charged
KERNEL_UNITS = {
    "infect_households": ["n_inf", "n_hh_scan"],
    "infect_places": ["n_place_link", "n_place_cand"],
    "cell_spatial_sum": ["n_cell_steps"],
    "spatial_draws": ["n_sp_draw", "n_sp_susc", "n_sp_scan", "n_sp_probe"],
    "resolve": ["n_new"],
    "latent_scan": ["n_active"],
    "infectious_scan": ["n_transition"],
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(max_seconds=25.0):
    """Time the reference's day loop on its 200k-person workload."""
    exe = os.path.join(ROOT, "oracle", "_ref", "covidsim_ref")
    inp = os.path.join(ROOT, "covid-sim_amd", "inputs", "c1")
    if not os.path.exists(exe):
        return None
    threads = len(os.sched_getaffinity(0))
    threads = max(1, min(threads, 96))  # MAX_NUM_THREADS (src/Constants.h)
    n_people, days = 200000, 200

    def run(nr):
        tmp = tempfile.mkdtemp()
        cmd = [exe, "/c:%d" % threads, "/NR:%d" % nr,
               "/PP:" + os.path.join(inp, "pre-params.txt"), "/P:" + os.path.join(inp, "input-noint-params.txt"),
               "/O:" + os.path.join(tmp, "out"), "/D:" + os.path.join(inp, "density.txt"),
               "/A:" + os.path.join(inp, "admin-params.txt"), "98798150", "729101", "17389101", "4797132"]
        p = subprocess.Popen(cmd, stderr=subprocess.PIPE, stdout=subprocess.DEVNULL, cwd=tmp)
        marks = []  # wall time of every realisation's t=0 line and End of run
        buf = b""
        while True:
            chunk = p.stderr.read1(65536) if hasattr(p.stderr, "read1") else p.stderr.read(4096)
            if not chunk:
                break
            now = time.perf_counter()
            buf += chunk
            # the progress lines use \r; scan complete markers only
            while True:
                i0 = buf.find(b"t=0 ")
                i1 = buf.find(b"End of run")
                cands = [(i, k) for i, k in ((i0, "start"), (i1, "end")) if i >= 0]
                if not cands:
                    buf = buf[-32:]
                    break
                i, k = min(cands)
                marks.append((k, now))
                buf = buf[i + 4:]
        p.wait()
        loop = 0.0
        start = None
        for k, tt in marks:
            if k == "start" and start is None:
                start = tt
            elif k == "end" and start is not None:
                loop += tt - start
                start = None
        return loop, p.returncode

    loop1, rc = run(1)
    if rc != 0 or loop1 <= 0:
        return None
    nr = int(max(1, min(400, (max_seconds - loop1) / max(loop1, 1e-3))))
    loop, rc = run(nr)
    if rc != 0 or loop <= 0:
        nr, loop = 1, loop1
    return {"value": nr * n_people * days / loop, "unit": "person-days/s", "cores": threads, "kind": "reference",
            "sample": "reference covidsim_ref, its 200k-person 200-day workload (4 steps/day, no interventions), "
                      "%d realisation(s), day-loop wall %.2f s from stderr markers, /c:%d" % (nr, loop, threads),
            "realisations": nr}


def gpu_same_workload(nr, csa, eng_mod):
    """The GPU on the CPU baseline's own workload: the 200k-person region as
    the reference's set-up builds it from covid-sim_amd/inputs/c1
    (tests/golden/c1_population.npz, built with /c:1), its parameters, 200
    days at 4 steps/day, as many realisations as the baseline ran, one
    launch; person-days/s of the second of two runs (the first captures the
    step graphs)."""
    z = np.load(os.path.join(ROOT, "tests", "golden", "c1_population.npz"))
    pop = csa.population.Population({k: z[k] for k in z.files})
    params = csa.population.load_params(os.path.join(ROOT, "covid-sim_amd", "inputs", "c1", "params.bin"))
    eng = eng_mod.Engine(0)
    eng.set_params(params)
    eng.set_population(pop)
    eng.alloc(np.arange(1, nr + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
    tspd = int(round(params.time_steps_per_day))
    dt = 0.0
    for _ in range(2):
        eng.sync()
        ta = time.perf_counter()
        eng.reset(10)
        eng.run_steps(0.0, 200 * tspd, tspd)
        eng.sync()
        dt = time.perf_counter() - ta
    eng.close()
    return nr * pop.n_people * 200 / dt, dt


def copy_probe(nbytes):
    """One device-to-device hipMemcpy of nbytes (the HIP runtime's copy
    kernel, __amd_rocclr_copyBuffer): a transfer of known size for scaling
    the profiler's FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)):
            raise RuntimeError("copy probe: hipMalloc of %d bytes failed" % nbytes)
    rc = hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) or hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # 3: device to device
    rc = rc or hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError("copy probe: HIP error %d" % rc)
    return nbytes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--people", type=int, default=5_000_000)
    ap.add_argument("--realisations", type=int, default=64)
    ap.add_argument("--days", type=int, default=300)
    ap.add_argument("--warmup-days", type=int, default=60, help="days simulated by each warmup step (the autotuner "
                                                                "then tunes on that state)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--chunk-days", type=int, default=50, help="days per run_steps call inside a step (progress lines)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel directly in the timed steps too "
                                                            "(for runs under rocprofv3 tracing)")
    ap.add_argument("--calibration", action="store_true", help="measure emulator-guided vs plain calibration instead")
    ap.add_argument("--profile-kinds", action="store_true", help="with --config large: time every launch by kind "
                    "(HIP events, direct launches)")
    ap.add_argument("--icu-trigger", type=float, default=0.0, help="with --config large: the uneven-regions "
                    "intervention subset (case isolation, household quarantine, place closure) with global ICU "
                    "triggers at this many cumulative ICU cases; 0 (default) no interventions")
    ap.add_argument("--dest-stride", type=int, default=-1, help="destination tables (headline and large), 0 full, "
                    "> 0 checkpoints every so many cells, -1 (default) chosen by the table memory")
    ap.add_argument("--cells", type=int, nargs=2, default=[668, 336], metavar=("WIDE", "HIGH"),
                    help="with --config large: the region's cells (0.075 degrees each)")
    ap.add_argument("--config", choices=["headline", "c3", "large"], default="headline",
                    help="c3: the 64-realisation ensemble over five regions of uneven size (see above)")
    ap.add_argument("--summarise-pmc", metavar="DIR", help="print per-kernel sums of a rocprofv3 --pmc output "
                                                            "directory (counter_collection.csv) and exit")
    ap.add_argument("--summarise-trace", metavar="DIR", help="print a per-kernel table of a rocprofv3 --kernel-trace "
                                                              "output directory (kernel_trace.csv) and exit")
    ap.add_argument("--pmc-last", type=int, default=0, help="with --summarise-pmc or --summarise-trace: also the "
                                                             "means over each kernel's last N dispatches")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_SUMMARY", "WRITE_SUMMARY"),
                    help="write profiles/traffic.json from two --summarise-pmc outputs of this workload (with "
                         "--pmc-last and --copy-probe-mb as in the profiled runs) and exit")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="before the run, copy this many MiB device-to-device (a copy of known size for "
                         "scaling the profiler's byte counters)")
    ap.add_argument("--cal-people", type=int, default=1_000_000)
    ap.add_argument("--cal-train", type=int, default=512, help="training candidates (8 seeds each)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the library has no multi-GPU path (an ensemble would shard "
              "by realisation, one batch of realisations per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.summarise_pmc:
        return summarise_pmc(args.summarise_pmc, args.pmc_last)
    if args.summarise_trace:
        return summarise_trace(args.summarise_trace, args.pmc_last)
    if args.write_traffic:
        return write_traffic(args.write_traffic[0], args.write_traffic[1], args.pmc_last,
                             os.path.join(ROOT, "profiles", "traffic.json"),
                             {"realisations": args.realisations, "persons": args.people, "days": args.days,
                              "copy_probe_bytes": args.copy_probe_mb << 20})
    if args.calibration:
        return calibration(args)
    if args.config == "c3":
        return uneven_regions(args)
    if args.config == "large":
        return large_region(args)
    import importlib
    csa = importlib.import_module("covid-sim_amd")
    eng_mod = importlib.import_module("covid-sim_amd.engine")
    popgen = importlib.import_module("covid-sim_amd.popgen")

    t0 = time.perf_counter()
    arrays, geom = popgen.generate(args.people, seed=2024)
    pop = csa.population.Population(arrays)
    params = csa.population.load_params(os.path.join(ROOT, "covid-sim_amd", "inputs", "c1", "params.bin"))
    popgen.apply_geometry(params, geom)
    log("population %d persons, %d cells (%.1f s)" % (pop.n_people, pop.n_cells, time.perf_counter() - t0))
    eng = eng_mod.Engine(0)
    eng.set_params(params)
    eng.set_population(pop)
    seeds = np.arange(1, args.realisations + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    if args.dest_stride >= 0:
        eng.set_dest_tables(args.dest_stride)
    eng.alloc(seeds)
    if args.no_graphs:
        eng.set_step_graphs(False)
    tspd = int(round(params.time_steps_per_day))
    steps = args.days * tspd
    log("set-up done (%.1f s); warmup %d x %d days" % (time.perf_counter() - t0, args.warmup, args.warmup_days))
    for _ in range(args.warmup):
        eng.reset(10)
        eng.run_steps(0.0, args.warmup_days * tspd, tspd)
    if args.warmup and not args.no_autotune:
        tune = eng.autotune(float(args.warmup_days))
        log("autotune at day %d: %s" % (args.warmup_days, json.dumps(tune)))
    eng.sync()
    if not args.copy_probe_mb:  # counter runs keep torch out of the profiled process
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.synchronize()
        except Exception:
            pass
    if args.copy_probe_mb:
        log("copy probe: %d bytes" % copy_probe(args.copy_probe_mb << 20))
    chunk = max(1, args.chunk_days) * tspd

    def one_step(k, label):
        eng.sync()
        ta = time.perf_counter()
        eng.reset(10)
        for s0 in range(0, steps, chunk):
            n = min(chunk, steps - s0)
            eng.run_steps(s0 / tspd, n, tspd)  # one series across the calls
            eng.sync()
            log("  %s %d: day %d (%.1f s)" % (label, k, (s0 + n) // tspd, time.perf_counter() - ta))
        dt = time.perf_counter() - ta
        log("%s %d: %.2f s" % (label, k, dt))
        return dt

    # the timed steps: replayed step graphs, nothing recorded per launch
    times = [one_step(k, "step") for k in range(args.steps)]
    # one more pass of the same workload with a HIP event pair around every
    # launch (direct launches): the per-kernel times and work counts of the
    # roofline; its kernels are the ones the timed steps replayed
    eng.work_counts(reset=True)
    eng.set_profiling(True)
    t_prof = one_step(0, "profiled step")
    eng.set_profiling(False)
    prof = eng.profile()
    wc = eng.work_counts()
    wc["n_cell_steps"] = args.realisations * pop.n_cells * prof["cell_spatial_sum"][1]  # cells visited per launch
    ts = eng.timeseries(args.days)
    total = float(sum(times))
    person_days = float(args.realisations) * args.people * args.days * args.steps
    value = person_days / total
    # the roofline of every kernel with counted work, the dominant one first
    def roof(k):
        ms, launches = prof[k]
        b = sum(BYTES[u] * wc[u] for u in KERNEL_UNITS[k]) / max(launches, 1)
        a = b / (ms / max(launches, 1) * 1e-3) / 1e9
        tr = measured_traffic(k, args)
        return {"kernel": k, "ms_total": round(ms, 1), "launches": launches, "bytes_per_launch": round(b),
                "ms_per_launch": round(ms / max(launches, 1), 4), "achieved_gbs": round(a, 2),
                "frac": round(a / HBM_PEAK_GBS, 5), "traffic_bytes_per_launch": tr,
                "traffic_over_algorithmic": round(tr / b, 2) if tr and b else None}
    ranked = sorted((k for k in prof if k in KERNEL_UNITS), key=lambda k: -prof[k][0])
    top = [roof(k) for k in ranked[:5]]
    dom = ranked[0]
    ms, launches = prof[dom]
    alg_bytes, achieved = top[0]["bytes_per_launch"], top[0]["achieved_gbs"]
    attack = float(np.mean((ts[:, -1, 3] + ts[:, -1, 4]) / args.people)) if ts.size else 0.0
    log("kernel ms: " + ", ".join("%s %.0f/%d" % (k, v[0], v[1]) for k, v in prof.items()))
    log("work: " + json.dumps(wc))
    log("mean attack rate %.3f" % attack)
    total_alg = sum(BYTES[u] * wc[u] for u in BYTES)
    traffic = top[0]["traffic_bytes_per_launch"]
    base = None if args.no_cpu_baseline else cpu_baseline()
    if base:  # the GPU on the baseline's own sample: the like-for-like ratio
        g, dt = gpu_same_workload(base["realisations"], csa, eng_mod)
        base["gpu_same_workload"] = {"value": g, "unit": "person-days/s", "seconds": round(dt, 3),
                                     "ratio_to_cpu": round(g / base["value"], 2),
                                     "workload": "the reference's own sample: the 200k-person region its set-up "
                                                 "builds, 200 days, %d realisations in one launch" % base["realisations"]}
    out = {
        "metric": "person-days/s, 64-run ensemble",
        "value": value,
        "unit": "person-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched ensemble: %d realisations x %d-person synthetic region, %d days, 4 steps/day, "
                               "10 seeds, no interventions" % (args.realisations, args.people, args.days),
                   "realisations": args.realisations, "persons": args.people, "days": args.days, "cells": pop.n_cells,
                   "mean_attack_rate": round(attack, 4), "launches": "direct launches" if args.no_graphs else "replayed step graphs",
                   "dest_tables": "full" if eng.dest_tables() == 0 else "checkpoints every %d cells" % eng.dest_tables(),
                   "all_kernel_algorithmic_gbs": round(total_alg / t_prof / 1e9, 2)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "bytes_per_launch": round(alg_bytes), "ms_per_launch": round(ms / max(launches, 1), 4),
                     "measured": "HIP events around every launch of a profiled pass of the same workload "
                                 "(%.2f s, direct launches)" % t_prof,
                     "top_kernels": top},
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


def uneven_regions(args):
    """The uneven-regions ensemble (--config c3) as one batch."""
    import importlib
    csa = importlib.import_module("covid-sim_amd")
    eng_mod = importlib.import_module("covid-sim_amd.engine")
    popgen = importlib.import_module("covid-sim_amd.popgen")
    inp = os.path.join(ROOT, "covid-sim_amd", "inputs", "c3")
    with open(os.path.join(inp, "realisations.json")) as f:
        spec = json.load(f)
    regions = spec["regions"]
    reg = np.array(spec["realisation_region"])
    r0 = np.array(spec["r0"], np.float64)
    thr = np.array(spec["trigger_threshold"], np.float64)
    seeds_all = np.arange(1, len(reg) + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    t0 = time.perf_counter()
    engines, sizes, ncells, betas = [], [], [], []
    for k, n in enumerate(regions):
        arrays, geom = popgen.generate(n, seed=2024 + k)
        pop = csa.population.Population(arrays)
        params = csa.population.load_params(os.path.join(inp, "params.bin"))
        popgen.apply_geometry(params, geom)
        eng = eng_mod.Engine(0)
        eng.set_params(params)
        eng.set_population(pop)
        c = eng.transmission_coeffs(77 + k)  # this region's R0 components (its LocalBeta for params.r0 too)
        mine = np.where(reg == k)[0]
        beta = np.maximum(0.0, (r0[mine] - c.r0_household - c.r0_places) / c.spatial_r0)
        eng.alloc(seeds_all[mine])
        eng.set_realisation_params(local_beta=beta)
        eng.set_realisation_triggers(thr[mine])
        if args.no_graphs:
            eng.set_step_graphs(False)
        engines.append(eng)
        sizes.append(len(mine) * n)
        ncells.append(pop.n_cells)
        betas.append([float(x) for x in beta])
        log("region %d: %d persons, %d cells, %d realisations, R0 %.2f-%.2f (%.1f s)"
            % (k, n, pop.n_cells, len(mine), r0[mine].min(), r0[mine].max(), time.perf_counter() - t0))
    batch = eng_mod.Batch(engines)
    tspd = int(round(params.time_steps_per_day))
    steps = args.days * tspd
    for _ in range(args.warmup):
        for e in engines:
            e.reset(10)
        batch.run_steps(0.0, args.warmup_days * tspd, tspd)
    if args.warmup and not args.no_autotune:
        for k, e in enumerate(engines):
            log("autotune region %d: %s" % (k, json.dumps(e.autotune(float(args.warmup_days)))))
    for e in engines:
        e.sync()
    chunk = max(1, args.chunk_days) * tspd

    def one_step(k, label):
        for e in engines:
            e.sync()
        ta = time.perf_counter()
        for e in engines:
            e.reset(10)
        for s0 in range(0, steps, chunk):
            n = min(chunk, steps - s0)
            batch.run_steps(s0 / tspd, n, tspd)
            for e in engines:
                e.sync()
            log("  %s %d: day %d (%.1f s)" % (label, k, (s0 + n) // tspd, time.perf_counter() - ta))
        dt = time.perf_counter() - ta
        log("%s %d: %.2f s" % (label, k, dt))
        return dt

    times = [one_step(k, "step") for k in range(args.steps)]
    # the profiled pass: every launch timed, so the engines run one after
    # another with direct launches; per-kernel times and work summed
    for e in engines:
        e.work_counts(reset=True)
        e.set_profiling(True)
    t_prof = one_step(0, "profiled step")
    prof, wc = {}, {}
    for e, nc, (k, n) in zip(engines, ncells, enumerate(regions)):
        e.set_profiling(False)
        pr = e.profile()
        for name, (ms, launches) in pr.items():
            a = prof.setdefault(name, [0.0, 0])
            a[0] += ms
            a[1] += launches
        w = e.work_counts()
        w["n_cell_steps"] = int((reg == k).sum()) * nc * pr["cell_spatial_sum"][1]  # cells visited per launch
        for u, v in w.items():
            wc[u] = wc.get(u, 0) + v
    total = float(sum(times))
    person_days = float(sum(sizes)) * args.days * args.steps
    value = person_days / total
    ranked = sorted((k for k in prof if k in KERNEL_UNITS), key=lambda k: -prof[k][0])
    top = []
    for name in ranked[:5]:
        ms, launches = prof[name]
        b = sum(BYTES[u] * wc.get(u, 0) for u in KERNEL_UNITS[name]) / max(launches, 1)
        a = b / (ms / max(launches, 1) * 1e-3) / 1e9
        top.append({"kernel": name, "ms_total": round(ms, 1), "launches": launches, "bytes_per_launch": round(b),
                    "ms_per_launch": round(ms / max(launches, 1), 4), "achieved_gbs": round(a, 2),
                    "frac": round(a / HBM_PEAK_GBS, 5)})
    attack = []
    for e, n in zip(engines, regions):
        ts = e.timeseries(args.days)
        attack.append(round(float(np.mean((ts[:, -1, 3] + ts[:, -1, 4]) / n)), 4))
    starts = []
    for e in engines:
        for r in range(e.n_real):
            s = e.intervention_state(r)
            starts.append(s.place_close_start < 1e9)
    log("kernel ms: " + ", ".join("%s %.0f/%d" % (k, v[0], v[1]) for k, v in prof.items()))
    log("mean attack rate by region: %s; interventions started in %d of %d realisations" % (attack, sum(starts), len(starts)))
    dom = top[0]
    out = {
        "metric": "person-days/s, 64-run ensemble",
        "value": value,
        "unit": "person-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "uneven-regions ensemble: 64 realisations over 5 synthetic regions (1M/2M/5M/10M/20M "
                               "persons, 13/13/13/13/12 realisations), per-realisation R0 ~ U[1.8, 3.0] and ICU "
                               "trigger threshold ~ U[20, 500], case isolation, household quarantine and place "
                               "closure on, %d days, 4 steps/day, 10 seeds; the regions' step graphs replayed side by side"
                               % args.days,
                   "realisations": int(len(reg)), "regions": regions, "cells": ncells, "days": args.days,
                   "person_realisations": int(sum(sizes)), "mean_attack_rate_by_region": attack,
                   "interventions_started": int(sum(starts)),
                   "launches": "direct launches" if args.no_graphs else "replayed step graphs, one stream per region"},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["frac"], "traffic": None,
                     "bytes_per_launch": dom["bytes_per_launch"], "ms_per_launch": dom["ms_per_launch"],
                     "measured": "HIP events around every launch of a profiled pass of the same workload "
                                 "(%.2f s, engines one after another, direct launches)" % t_prof,
                     "top_kernels": top},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def large_region(args):
    """One large region (--config large): its population generated on the
    device (cs_popgen_*: SetupPopulation, AssignHouseholdAges,
    AssignPeopleToPlaces, StratifyPlaces) from a synthetic density, then one
    realisation simulated with the destination tables the region's size
    calls for (checkpointed beyond NC x NC that fit)."""
    import importlib
    csa = importlib.import_module("covid-sim_amd")
    eng_mod = importlib.import_module("covid-sim_amd.engine")
    sm = importlib.import_module("covid-sim_amd.setup_model")
    inp = os.path.join(ROOT, "covid-sim_amd", "inputs", "c1")
    base_pg = sm.load_popgen_params(os.path.join(inp, "popgen_params.bin"))
    base_kp = csa.population.load_params(os.path.join(inp, "params.bin"))
    if args.icu_trigger > 0:  # the intervention subset of the uneven-regions parameters
        base_kp = csa.population.load_params(os.path.join(ROOT, "covid-sim_amd", "inputs", "c3", "params.bin"))
    ncw, nch = args.cells
    pg, kp, dens, country, adunit = sm.synthetic_region(args.people, ncw, nch, 2025, base_pg, base_kp)
    log("region: %d persons, %d x %d cells, %d microcells" % (args.people, ncw, nch, pg.mcw * pg.mch))
    gen = sm.PopGen(0)
    t0 = time.perf_counter()
    sizes = gen.run(pg, kp, dens, country, adunit, 77)
    t_gen = time.perf_counter() - t0
    log("generated on the device in %.2f s (%.3g persons/s): %d households, %d populated cells, places %s, "
        "unplaced %s" % (t_gen, args.people / t_gen, sizes.n_households, sizes.n_cells, list(sizes.n_places)[:4],
                         list(sizes.n_unplaced)[:4]))
    t1 = time.perf_counter()
    arrays = gen.fetch(pg)
    gen.close()
    pop = csa.population.Population(arrays)
    eng = eng_mod.Engine(0)
    eng.set_params(kp)
    eng.set_population(pop)
    c = eng.transmission_coeffs(5)
    if args.dest_stride >= 0:
        eng.set_dest_tables(args.dest_stride)
    eng.alloc(np.array([0x9E3779B97F4A7C15], np.uint64))
    if args.icu_trigger > 0:
        eng.set_realisation_triggers(np.array([args.icu_trigger], np.float64))
    stride = eng.dest_tables()
    t_setup = time.perf_counter() - t1
    log("engine set-up %.1f s (host copy and checks); R0 %.2f (household %.2f, places %.2f), LocalBeta %.4g; "
        "destination tables: %s" % (t_setup, c.r0, c.r0_household, c.r0_places, c.local_beta,
                                    "full" if stride == 0 else "checkpoints every %d cells" % stride))
    tspd = int(round(kp.time_steps_per_day))
    if args.no_graphs:
        eng.set_step_graphs(False)
    if args.profile_kinds:
        eng.set_profiling(True)  # HIP events around every launch (direct launches)
    eng.reset(10)
    eng.sync()
    ta = time.perf_counter()
    chunk = max(1, args.chunk_days) * tspd
    steps = args.days * tspd
    for s0 in range(0, steps, chunk):
        n = min(chunk, steps - s0)
        eng.run_steps(s0 / tspd, n, tspd)
        eng.sync()
        log("  day %d (%.1f s)" % ((s0 + n) // tspd, time.perf_counter() - ta))
    dt = time.perf_counter() - ta
    ts = eng.timeseries(args.days)
    attack = float((ts[0, -1, 3] + ts[0, -1, 4]) / args.people)
    itv = None
    if args.icu_trigger > 0:
        st = eng.intervention_state(0)
        itv = {"icu_trigger": args.icu_trigger, "case_isolation_prop": kp.case_isolation_prop,
               "household_quarantine_prop": kp.hquarantine_prop_house_compliant,
               "alert": int(st.alert),
               "case_isolation_start_day": st.case_isolation_start if st.case_isolation_start < 1e9 else None,
               "household_quarantine_start_day": st.hquarantine_start if st.hquarantine_start < 1e9 else None,
               "place_closure_start_day": st.place_close_start if st.place_close_start < 1e9 else None}
        log("interventions (ICU trigger %g): %s" % (args.icu_trigger, itv))
    kinds = None
    if args.profile_kinds:
        kinds = {k: [round(ms, 1), n] for k, (ms, n) in eng.profile().items() if n}
        log("kernel ms by kind: " + ", ".join("%s %.0f/%d" % (k, v[0], v[1]) for k, v in kinds.items()))
    out = {
        "metric": "person-days/s, one large region",
        "value": args.people * args.days / dt,
        "unit": "person-days/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "one region of %d persons on %d x %d cells (%d populated), population generated on "
                               "the device, 1 realisation, %d days, 4 steps/day, 10 seeds, %s"
                               % (args.people, ncw, nch, sizes.n_cells, args.days,
                                  "case isolation 70%%, household quarantine 50%%, school and university closure "
                                  "when cumulative ICU cases pass %g" % args.icu_trigger if args.icu_trigger > 0
                                  else "no interventions"),
                   "persons": args.people, "cells": int(sizes.n_cells), "days": args.days,
                   "generation_seconds": round(t_gen, 3), "generation_persons_per_s": args.people / t_gen,
                   "engine_setup_seconds": round(t_setup, 1),
                   "dest_tables": "full" if stride == 0 else "checkpoints every %d cells" % stride,
                   "r0": round(c.r0, 3), "attack_rate": round(attack, 4),
                   "launches": "direct launches, every launch timed" if args.profile_kinds else
                               ("direct launches" if args.no_graphs else "replayed step graphs"),
                   "kernel_ms_by_kind": kinds, "interventions": itv},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


# the kernels behind each timed kind whose traffic is summed (the profiler
# measures kernels; the library times kinds)
KIND_KERNELS = {"infect_places": ["infect_places", "place_contacts"],
                "resolve": ["resolve_claim", "resolve_apply", "resolve_reset"]}


def measured_traffic(kernel, args):
    """HBM bytes per launch of the timed kind `kernel` from the committed
    profiler summary (profiles/traffic.json: rocprofv3 FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the same counters' reading for a copy
    of known size), summed over the kind's kernels the summary holds, when
    it was measured on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    w = t.get("workload", {})
    if (w.get("realisations"), w.get("persons"), w.get("days")) != (args.realisations, args.people, args.days):
        return None
    got = [t["kernels"][k]["hbm_bytes_per_launch"] for k in KIND_KERNELS.get(kernel, [kernel]) if k in t.get("kernels", {})]
    return sum(got) if got else None


def write_traffic(fetch_summary, write_summary, last, out, workload):
    """profiles/traffic.json from two --summarise-pmc outputs (FETCH_SIZE and
    WRITE_SIZE, KiB per dispatch): each counter is scaled by bytes moved /
    counter reading of the largest __amd_rocclr_copyBuffer dispatch (the
    --copy-probe-mb copy), and a kernel's HBM bytes per launch are the scaled
    means over its last `last` dispatches."""
    def parse(path, counter):
        rows = {}
        with open(path) as f:
            for line in f:
                parts = [x.strip() for x in line.split("|")]
                name = short_name(parts[0])
                m = re.search(counter + r" total \S+ mean \S+ max (\S+)(?: mean_last\d+ (\S+))?", line)
                if m:
                    rows[name] = (float(m.group(1)), float(m.group(2)) if m.group(2) else None)
        return rows
    fr, wr = parse(fetch_summary, "FETCH_SIZE"), parse(write_summary, "WRITE_SIZE")
    probe = workload.pop("copy_probe_bytes")
    f_fetch = probe / (fr["__amd_rocclr_copyBuffer"][0] * 1024.0)
    f_write = probe / (wr["__amd_rocclr_copyBuffer"][0] * 1024.0)
    kernels = {}
    for name in fr:
        if not name.startswith("cs::k_") or name not in wr or fr[name][1] is None:
            continue
        fb, wb = fr[name][1] * 1024.0 * f_fetch, wr[name][1] * 1024.0 * f_write
        kernels[name[len("cs::k_"):]] = {"hbm_bytes_per_launch": round(fb + wb), "fetch_bytes": round(fb),
                                         "write_bytes": round(wb), "fetch_kib_raw": fr[name][1],
                                         "write_kib_raw": wr[name][1], "dispatches_averaged": last}
    doc = {"workload": workload, "copy_probe_bytes": probe,
           "scale": {"fetch": f_fetch, "write": f_write,
                     "how": "bytes of one device-to-device copy / the counter's KiB x 1024 for that copy"},
           "kernels": kernels}
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print(json.dumps(doc, indent=1))


def summarise_pmc(d, last=0):
    """Per-kernel totals and per-dispatch means of every counter in a
    rocprofv3 --pmc CSV output directory; with last > 0 also the means over
    each kernel's last `last` dispatches (the profiled pass of a bench run)."""
    import csv
    import glob
    vals = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                k, c = row["Kernel_Name"], row["Counter_Name"]
                key = (int(row["Dispatch_Id"]), c)
                vals.setdefault(k, {})
                vals[k][key] = vals[k].get(key, 0.0) + float(row["Counter_Value"])
    for k in sorted(vals, key=lambda k: -sum(vals[k].values())):
        per = {}
        for (did, c), v in vals[k].items():
            per.setdefault(c, {})[did] = v
        parts = []
        for c in sorted(per):
            ids = sorted(per[c])
            tot = sum(per[c].values())
            txt = "%s total %.6g mean %.6g max %.6g" % (c, tot, tot / len(ids), max(per[c].values()))
            if last and len(ids) >= last:
                txt += " mean_last%d %.6g" % (last, sum(per[c][i] for i in ids[-last:]) / last)
            parts.append(txt)
        print("%s | dispatches %d | %s" % (k[:80], len(per[min(per)]), " | ".join(parts)))


def short_name(k):
    """A kernel's name without its argument list ("(anonymous namespace)::"
    kept out of the cut) and without the return type and template arguments
    the profiler prints for a template instance ("void cs::k_x<false>")."""
    k = k.replace("(anonymous namespace)::", "")
    k = k.split("(")[0]
    if k.startswith("void "):
        k = k[len("void "):]
    return k.split("<")[0]


def summarise_trace(d, last=0):
    """Per-kernel dispatch count, total and mean duration from a rocprofv3
    --kernel-trace CSV output directory (kernel_trace.csv), as a markdown
    table; with last > 0 also the mean over each kernel's last `last`
    dispatches (for a kernel launched once per timestep: the profiled pass of
    a bench run), with the LDS and scratch sizes the trace records."""
    import csv
    import glob
    runs = {}
    for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                k = row["Kernel_Name"]
                dur = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
                runs.setdefault(k, []).append((int(row["Dispatch_Id"]), dur, row["LDS_Block_Size"], row["Scratch_Size"]))
    tot_all = sum(x[1] for v in runs.values() for x in v)
    head = "| kernel | dispatches | total ms | share | mean us |"
    if last:
        head += " mean us, last %d | LDS B | scratch B/lane |" % last
    print(head)
    print("|" + "---|" * (head.count("|") - 1))
    for k in sorted(runs, key=lambda k: -sum(x[1] for x in runs[k])):
        v = sorted(runs[k])
        tot = sum(x[1] for x in v)
        line = "| %s | %d | %.1f | %.2f%% | %.1f |" % (short_name(k), len(v), tot / 1e6, 100.0 * tot / tot_all,
                                                      tot / len(v) / 1e3)
        if last:
            tail = v[-last:]
            line += " %s | %s | %s |" % ("%.1f" % (sum(x[1] for x in tail) / len(tail) / 1e3)
                                       if len(v) >= last else "-", v[-1][2], v[-1][3])
        print(line)


def calibration(args):
    import importlib
    cal = importlib.import_module("covid-sim_amd.calibrate")
    t0 = time.perf_counter()
    sim = cal.Simulator(n_people=args.cal_people)
    log("calibration population %d persons (%.1f s)" % (sim.pop.n_people, time.perf_counter() - t0))
    sim.progress = log
    rng = np.random.default_rng(12)
    train_seeds = np.arange(1000, 1008, dtype=np.uint64)
    t1 = time.perf_counter()
    th_tr, y_tr = cal.training_set(sim, args.cal_train, train_seeds, rng)
    th_te, y_te = cal.training_set(sim, 64, train_seeds + np.uint64(500), rng)
    t_data = time.perf_counter() - t1
    inc2, d2 = sim.run(th_te, train_seeds + np.uint64(900))  # independent seeds: the noise floor
    y_te2 = cal.from_curves(inc2, d2)
    log("training data: %d + 64 candidates x 8 seeds (+ 64 x 8 for the noise floor), %d launches, %.1f s"
        % (args.cal_train, sim.launches, t_data))
    t1 = time.perf_counter()
    emu, losses = cal.train_emulator(cal.encode(th_tr), y_tr)
    t_train = time.perf_counter() - t1
    err = cal.emulator_error(emu, th_te, y_te, y_te2)
    log("emulator: loss %.4g -> %.4g in %.1f s; held-out relative L2 error %s" % (losses[0], losses[-1], t_train, json.dumps(err)))
    hidden = np.array([2.15, 0.25])
    _, d = sim.run(hidden[None], np.arange(5000, 5016, dtype=np.uint64))
    target = np.round(d[0])
    seeds = np.arange(100, 116, dtype=np.uint64)
    plain = cal.calibrate_plain(sim, target, seeds)
    log("plain: %s" % json.dumps(plain))
    emul = cal.calibrate_emulated(sim, emu, target, seeds)
    log("emulator-guided: %s" % json.dumps(emul))
    local = cal.calibrate_pattern(sim, target, seeds)
    log("local pattern search alone: %s" % json.dumps(local))
    out = {
        "metric": "calibration wall-time, emulator-guided vs plain (speed-up)",
        "value": plain["seconds"] / emul["seconds"],
        "unit": "x",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "bf16 emulator, f64 simulator",
        "data": "synthetic",
        "config": {"workload": "fit R0 and seeding scale to a 365-day death curve of a %d-person synthetic region "
                               "(hidden R0 %.2f, log10 scale %.2f); candidates verified with 16-seed GPU ensembles"
                               % (sim.pop.n_people, hidden[0], hidden[1]),
                   "emulator": "10-512-512-730 GELU MLP, BF16 MFMA kernels, FP32 accumulation and Adam",
                   "training_candidates": args.cal_train, "training_seeds": 8},
        "plain": plain, "emulator_guided": emul, "local_pattern_search": local,
        "hidden_theta": hidden.tolist(),
        "fitted_r0_difference": abs(plain["theta"][0] - emul["theta"][0]),
        "realisation_ratio": plain["realisations"] / max(emul["realisations"], 1),
        "training_data_seconds": t_data, "training_seconds": t_train,
        # the whole cost: one calibration that also pays for its emulator's
        # training data and training, and the number of calibrations on one
        # region and parameter family after which the emulator has paid off
        "speedup_with_training": plain["seconds"] / (emul["seconds"] + t_data + t_train),
        "break_even_calibrations": (t_data + t_train) / max(plain["seconds"] - emul["seconds"], 1e-9),
        "emulator_heldout_rel_l2": err,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
