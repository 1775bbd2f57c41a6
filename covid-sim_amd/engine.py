"""The HIP library (libcovidsim_hip.so) behind the C ABI, from Python.

Every method is a thin call of one cs_* entry point of
include/covidsim_hip.h; errors raise CSError with the library's message.
There is no fallback: without the built library, or without a GPU, the
engine cannot be created.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcovidsim_hip.so")

_lib = None


class CSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("covidsim_hip error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("%s is not built (run __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        abi.declare(L, "cs_")
        P = C.POINTER
        L.cs_create.restype = C.c_int
        L.cs_create.argtypes = [P(C.c_void_p), C.c_int]
        L.cs_set_params.argtypes = [C.c_void_p, P(abi.cs_params), P(C.c_double), P(C.c_double)]
        L.cs_set_params.restype = C.c_int
        L.cs_set_population.argtypes = [C.c_void_p, P(abi.cs_population)]
        L.cs_set_population.restype = C.c_int
        L.cs_select_infector.argtypes = [C.c_void_p, P(C.c_float), C.c_int, P(C.c_double), P(C.c_int32), C.c_int]
        L.cs_select_infector.restype = C.c_int
        L.cs_select_dest_cell.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), P(C.c_int32), C.c_int]
        L.cs_select_dest_cell.restype = C.c_int
        L.cs_philox.argtypes = [C.c_void_p, P(C.c_uint32), P(C.c_uint32), C.c_int]
        L.cs_philox.restype = C.c_int
        L.cs_sample.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                P(C.c_double), C.c_int]
        L.cs_sample.restype = C.c_int
        L.cs_init_transmission_coeffs.argtypes = [C.c_void_p, P(C.c_double), C.c_uint64, P(abi.cs_transmission_coeffs)]
        L.cs_init_transmission_coeffs.restype = C.c_int
        L.cs_version.restype = C.c_char_p
        L.cs_last_error.restype = C.c_char_p
        L.cs_last_error.argtypes = [C.c_void_p]
        L.cs_batch_create.argtypes = [P(C.c_void_p), P(C.c_void_p), C.c_int]
        L.cs_batch_create.restype = C.c_int
        L.cs_batch_destroy.argtypes = [C.c_void_p]
        L.cs_batch_destroy.restype = None
        L.cs_batch_last_error.argtypes = [C.c_void_p]
        L.cs_batch_last_error.restype = C.c_char_p
        L.cs_batch_run_steps.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
        L.cs_batch_run_steps.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Engine:
    """One engine (one device, a population, a set of realisations)."""

    prefix = "cs_"

    def __init__(self, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.cs_create(C.byref(self.h), int(device))
        self._check(rc)
        self._keep = []

    def _check(self, rc):
        if rc:
            raise CSError(rc, self.L.cs_last_error(self.h).decode())

    def call(self, name, *args):
        self._check(getattr(self.L, self.prefix + name)(self.h, *args))

    def close(self):
        if self.h:
            self.L.cs_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up ----
    def set_params(self, params, kl=None, kh=None):
        self._params = params
        self._check(self.L.cs_set_params(self.h, C.byref(params),
                                         _p(kl, C.c_double) if kl is not None else None,
                                         _p(kh, C.c_double) if kh is not None else None))

    def set_population(self, pop):
        self._pop = pop
        self._check(self.L.cs_set_population(self.h, C.byref(pop.view())))

    def alloc(self, seeds):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        self.n_real = len(seeds)
        self.call("alloc_realisations", len(seeds), _p(seeds, C.c_uint64))

    def set_realisation_params(self, local_beta=None, n_seed=None):
        lb = None if local_beta is None else np.ascontiguousarray(local_beta, np.float64)
        ns = None if n_seed is None else np.ascontiguousarray(n_seed, np.int32)
        self.call("set_realisation_params", None if lb is None else _p(lb, C.c_double),
                  None if ns is None else _p(ns, C.c_int32))

    def set_realisation_triggers(self, threshold=None):
        th = None if threshold is None else np.ascontiguousarray(threshold, np.float64)
        self.call("set_realisation_triggers", None if th is None else _p(th, C.c_double))

    def reset(self, n_seed):
        self.call("reset_realisations", int(n_seed))

    def load_state(self, r, st):
        self.call("load_state", int(r), C.byref(st.view()))

    def store_state(self, r, st):
        self.call("store_state", int(r), C.byref(st.view()))

    # ---- the path ----
    def infect_sweep(self, t):
        self.call("infect_sweep", float(t))

    def incub_recovery_sweep(self, t):
        self.call("incub_recovery_sweep", float(t))

    def update_probs(self):
        self.call("update_probs")

    def run_steps(self, t0, n_steps, steps_per_sample):
        self.call("run_steps", float(t0), int(n_steps), int(steps_per_sample))

    def set_queue_capacity(self, events=0, draw_chunks=0):
        self.call("set_queue_capacity", int(events), int(draw_chunks))

    def set_due_capacity(self, per_step):
        self.call("set_due_capacity", int(per_step))

    def queue_capacity(self):
        a, b = C.c_uint32(), C.c_uint32()
        self.call("get_queue_capacity", C.byref(a), C.byref(b))
        return a.value, b.value

    def set_move_capacity(self, per_cell):
        """cells with more list moves in a step than per_cell move through
        device memory instead of LDS (0: every cell; results are the same)"""
        self.call("set_move_capacity", int(per_cell))

    def set_launch_shape(self, grid_waves=0, grid_cells=0):
        self.call("set_launch_shape", int(grid_waves), int(grid_cells))

    def autotune(self, t, reps=3):
        """Launch shapes and queue capacity for the current state (cs_autotune)."""
        tu = abi.cs_tuning()
        self.call("autotune", float(t), int(reps), C.byref(tu))
        return {f: getattr(tu, f) for f, _ in tu._fields_}

    def set_dest_tables(self, stride):
        """Destination tables at the next alloc: 0 full, > 0 checkpoint stride, -1 automatic."""
        self.call("set_dest_tables", int(stride))

    def set_dest_index(self, enable):
        """Checkpointed tables' per-row bucket index at the next alloc (default on)."""
        self.call("set_dest_index", int(bool(enable)))

    def dest_tables(self):
        s = C.c_int()
        self.call("get_dest_tables", C.byref(s))
        return s.value

    def set_step_graphs(self, on):
        self.call("set_step_graphs", int(bool(on)))

    def sync(self):
        self.call("sync")

    KERNEL_KINDS = ["infect_households", "spatial_draws", "resolve", "list_move", "latent_scan",
                    "infectious_scan", "update_probs", "cell_spatial_sum", "infect_places", "other"]

    def set_profiling(self, on):
        self.call("set_profiling", int(bool(on)))

    def profile(self, reset=False):
        n = len(self.KERNEL_KINDS)
        ms = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        self.call("get_profile", _p(ms, C.c_double), _p(cnt, C.c_int64), int(bool(reset)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.KERNEL_KINDS)}

    # ---- results ----
    def counters(self, r, reset=False):
        c = abi.cs_counters()
        self.call("get_counters", int(r), C.byref(c), int(bool(reset)))
        return c

    def intervention_state(self, r):
        """cs_get_intervention_state of realisation r"""
        st = abi.cs_intervention_state()
        self._check(self.L.cs_get_intervention_state(self.h, int(r), C.byref(st)))
        return st

    def timeseries(self, max_samples):
        """[n_real, n_samples, CS_TS_FIELDS] int32: per sample S, L, I, R, D,
        incidence of infections, cases and deaths, incidence of each severity
        kind (abi.CS_TS_INC_SEV + CS_SK_*) and severity prevalence
        (abi.CS_TS_PREV_SEV + CS_SK_MILD .. CS_SK_CRITRECOV)."""
        out = np.zeros((self.n_real, max_samples, abi.CS_TS_FIELDS), np.int32)
        n = C.c_int()
        self.call("get_timeseries", _p(out, C.c_int32), int(max_samples), C.byref(n))
        return out[:, :n.value]

    def timeseries_age(self, max_samples):
        """[n_real, n_samples, CS_TSA_FIELDS, CS_NUM_AGE_GROUPS] int32: per
        sample and age group the incidence of infections, cases, deaths and
        each severity kind, and severity prevalence."""
        out = np.zeros((self.n_real, max_samples, abi.CS_TSA_FIELDS, abi.CS_NUM_AGE_GROUPS), np.int32)
        n = C.c_int()
        self.call("get_timeseries_age", _p(out, C.c_int32), int(max_samples), C.byref(n))
        return out[:, :n.value]

    def timeseries_adunit(self, max_samples, n_adunits):
        """[n_real, n_samples, CS_TSU_FIELDS, n_adunits] int32: per sample
        and administrative unit the incidence of infections, cases,
        detected cases and deaths."""
        out = np.zeros((self.n_real, max_samples, abi.CS_TSU_FIELDS, n_adunits), np.int32)
        n = C.c_int()
        self.call("get_timeseries_adunit", _p(out, C.c_int32), int(max_samples), C.byref(n))
        return out[:, :n.value]

    def adunit_counters(self, r, n_adunits, reset=False):
        """[n_adunits, CS_TSU_FIELDS] int64: cumI, cumC, cumDC, cumD by unit."""
        out = np.zeros((n_adunits, abi.CS_TSU_FIELDS), np.int64)
        self.call("get_adunit_counters", int(r), _p(out, C.c_int64), int(bool(reset)))
        return out

    def work_counts(self, reset=False):
        w = abi.cs_work_counts()
        self.call("get_work_counts", C.byref(w), int(bool(reset)))
        return {f: getattr(w, f) for f, _ in w._fields_}

    # ---- components ----
    def transmission_coeffs(self, seed, prop_age_group=None):
        """InitTransmissionCoeffs with per-person draws from the stream `seed`;
        sets the engine's local_beta unless the parameters fix it."""
        out = abi.cs_transmission_coeffs()
        pa = None if prop_age_group is None else np.ascontiguousarray(prop_age_group, np.float64)
        self.call("init_transmission_coeffs", None if pa is None else _p(pa, C.c_double), int(seed), C.byref(out))
        return out

    def transmission_coeffs_from_draws(self, infness, rec, infstat=None, prop_age_group=None):
        out = abi.cs_transmission_coeffs()
        infness = np.ascontiguousarray(infness, np.float32)
        rec = np.ascontiguousarray(rec, np.uint16)
        st = None if infstat is None else np.ascontiguousarray(infstat, np.int8)
        pa = None if prop_age_group is None else np.ascontiguousarray(prop_age_group, np.float64)
        self.call("transmission_coeffs_from_draws", None if pa is None else _p(pa, C.c_double), _p(infness, C.c_float),
                  _p(rec, C.c_uint16), None if st is None else _p(st, C.c_int8), C.byref(out))
        return out

    def kernel_num(self, r2):
        r2 = np.ascontiguousarray(r2, np.float64)
        out = np.zeros_like(r2)
        self.call("kernel_num", _p(r2, C.c_double), _p(out, C.c_double), len(r2))
        return out

    def dist2(self, a, b):
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        out = np.zeros(len(a))
        self.call("dist2_persons", _p(a, C.c_int32), _p(b, C.c_int32), _p(out, C.c_double), len(a))
        return out

    def cell_tables(self, r, nc):
        mt = np.zeros((nc, nc), np.float32)
        ct = np.zeros((nc, nc), np.float32)
        inv = np.zeros((nc, 1025), np.int32)
        tp = np.zeros(nc, np.float32)
        s0 = np.zeros(nc, np.int32)
        f = C.c_float
        self.call("get_cell_tables", int(r), _p(mt, f), _p(ct, f), _p(inv, C.c_int32), _p(tp, f), _p(s0, C.c_int32))
        return mt, ct, inv, tp, s0

    def hazards(self, r, kind, ts, person, other):
        person = np.ascontiguousarray(person, np.int32)
        other = np.ascontiguousarray(other, np.int32)
        out = np.zeros(len(person))
        self.call("eval_hazards", int(r), int(kind), int(ts), _p(person, C.c_int32), _p(other, C.c_int32),
                  _p(out, C.c_double), len(person))
        return out

    def choose(self, table, u, mean):
        u = np.ascontiguousarray(u, np.float64)
        mean = np.ascontiguousarray(mean, np.float64)
        out = np.zeros(len(u), np.uint16)
        self.call("icdf_choose", int(table), _p(u, C.c_double), _p(mean, C.c_double), _p(out, C.c_uint16), len(u))
        return out

    def select_infector(self, cell_inf, u):
        ci = np.ascontiguousarray(cell_inf, np.float32)
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros(len(u), np.int32)
        self._check(self.L.cs_select_infector(self.h, _p(ci, C.c_float), len(ci), _p(u, C.c_double),
                                              _p(out, C.c_int32), len(u)))
        return out

    def select_dest(self, r, cell, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros(len(u), np.int32)
        self._check(self.L.cs_select_dest_cell(self.h, int(r), int(cell), _p(u, C.c_double), _p(out, C.c_int32), len(u)))
        return out

    def philox(self, ck):
        ck = np.ascontiguousarray(ck, np.uint32)
        n = len(ck) // 6
        out = np.zeros(4 * n, np.uint32)
        self._check(self.L.cs_philox(self.h, _p(ck, C.c_uint32), _p(out, C.c_uint32), n))
        return out

    def sample(self, which, key, c0, c1, a, b, n):
        out = np.zeros(n)
        self._check(self.L.cs_sample(self.h, int(which), int(key), int(c0), int(c1), float(a), float(b),
                                     _p(out, C.c_double), int(n)))
        return out


class Batch:
    """Engines of one device (a region each) advanced together by one graph
    launch per output step (cs_batch_*).  The engines stay the caller's and
    keep their own results."""

    def __init__(self, engines):
        self.L = lib()
        self.engines = list(engines)
        hs = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        self.h = C.c_void_p()
        rc = self.L.cs_batch_create(C.byref(self.h), hs, len(self.engines))
        if rc:
            raise CSError(rc, "cs_batch_create")

    def run_steps(self, t0, n_steps, steps_per_sample):
        rc = self.L.cs_batch_run_steps(self.h, float(t0), int(n_steps), int(steps_per_sample))
        if rc:
            raise CSError(rc, self.L.cs_batch_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.cs_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

