"""ctypes wrapper of the CPU oracle (oracle/_build/libcs_oracle.so).

Test infrastructure: the oracle is the checker for the GPU library and is
never used by the package itself.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "libcs_oracle.so")


def load():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    import importlib
    abi = importlib.import_module("covid-sim_amd.abi")
    lib = C.CDLL(LIB)
    abi.declare(lib, "or_")
    P = C.POINTER
    lib.or_create.restype = C.c_int
    lib.or_create.argtypes = [P(C.c_void_p), P(abi.cs_params), P(C.c_double), P(C.c_double), P(abi.cs_population)]
    lib.or_select_infector.argtypes = [P(C.c_float), C.c_int, P(C.c_double), P(C.c_int32), C.c_int]
    lib.or_select_dest_cell.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), P(C.c_int32), C.c_int]
    lib.or_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), C.c_int]
    lib.or_philox.restype = None
    lib.or_sample.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                              C.c_int, P(C.c_double)]
    lib.or_sample.restype = C.c_int
    lib.or_init_transmission_coeffs.argtypes = [C.c_void_p, P(C.c_double), C.c_uint64, P(abi.cs_transmission_coeffs),
                                                P(C.c_float), P(C.c_float), P(C.c_uint16)]
    lib.or_init_transmission_coeffs.restype = C.c_int
    lib.or_household_ages.argtypes = [P(abi.cs_popgen_params), C.c_uint64, P(C.c_int32), C.c_int, P(C.c_uint8)]
    lib.or_household_ages.restype = C.c_int
    lib.or_nearest_places.argtypes = [P(abi.cs_params), P(abi.cs_popgen_params), C.c_int, C.c_uint64, P(C.c_int32),
                                      P(C.c_float), P(C.c_int32), C.c_int, P(C.c_int64), P(C.c_float), P(C.c_int32),
                                      P(C.c_int32)]
    lib.or_nearest_places.restype = C.c_int
    lib.or_log.argtypes = [C.c_double]
    lib.or_log.restype = C.c_double
    lib.or_exp.argtypes = [C.c_double]
    lib.or_exp.restype = C.c_double
    return lib, abi


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Oracle:
    def __init__(self, params, pop, kl=None, kh=None):
        self.lib, self.abi = load()
        self.params, self.pop = params, pop
        self.h = C.c_void_p()
        self._kl, self._kh = kl, kh
        rc = self.lib.or_create(C.byref(self.h), C.byref(params),
                                dptr(kl) if kl is not None else None, dptr(kh) if kh is not None else None,
                                C.byref(pop.view()))
        self.check(rc)

    def check(self, rc):
        if rc:
            raise RuntimeError("oracle error %d: %s" % (rc, self.lib.or_last_error(self.h)))

    def __del__(self):
        try:
            self.lib.or_destroy(self.h)
        except Exception:
            pass

    def call(self, name, *args):
        self.check(getattr(self.lib, "or_" + name)(self.h, *args))


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleEngine(Oracle):
    """The oracle behind the same method names as the package's Engine."""

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

    def infect_sweep(self, t):
        self.call("infect_sweep", float(t))

    def incub_recovery_sweep(self, t):
        self.call("incub_recovery_sweep", float(t))

    def update_probs(self):
        self.call("update_probs")

    def run_steps(self, t0, n_steps, sps):
        self.call("run_steps", float(t0), int(n_steps), int(sps))

    def counters(self, r, reset=False):
        c = self.abi.cs_counters()
        self.call("get_counters", int(r), C.byref(c), int(bool(reset)))
        return c

    def intervention_state(self, r):
        s = self.abi.cs_intervention_state()
        self.call("get_intervention_state", int(r), C.byref(s))
        return s

    def timeseries(self, max_samples):
        out = np.zeros((self.n_real, max_samples, self.abi.CS_TS_FIELDS), np.int32)
        n = C.c_int()
        self.call("get_timeseries", _p(out, C.c_int32), int(max_samples), C.byref(n))
        return out[:, :n.value]

    def timeseries_age(self, max_samples):
        """[n_real, n_samples, CS_TSA_FIELDS, CS_NUM_AGE_GROUPS] int32: per
        sample and age group the incidence of infections, cases, deaths and
        each severity kind, and severity prevalence."""
        out = np.zeros((self.n_real, max_samples, self.abi.CS_TSA_FIELDS, self.abi.CS_NUM_AGE_GROUPS), np.int32)
        n = C.c_int()
        self.call("get_timeseries_age", _p(out, C.c_int32), int(max_samples), C.byref(n))
        return out[:, :n.value]

    def timeseries_adunit(self, max_samples, n_adunits):
        """[n_real, n_samples, CS_TSU_FIELDS, n_adunits] int32: per sample
        and administrative unit the incidence of infections, cases,
        detected cases and deaths."""
        out = np.zeros((self.n_real, max_samples, self.abi.CS_TSU_FIELDS, n_adunits), np.int32)
        n = C.c_int()
        self.call("get_timeseries_adunit", _p(out, C.c_int32), int(max_samples), C.byref(n))
        return out[:, :n.value]

    def adunit_counters(self, r, n_adunits, reset=False):
        """[n_adunits, CS_TSU_FIELDS] int64: cumI, cumC, cumDC, cumD by unit."""
        out = np.zeros((n_adunits, self.abi.CS_TSU_FIELDS), np.int64)
        self.call("get_adunit_counters", int(r), _p(out, C.c_int64), int(bool(reset)))
        return out

    def work_counts(self, reset=False):
        w = self.abi.cs_work_counts()
        self.call("get_work_counts", C.byref(w), int(bool(reset)))
        return {f: getattr(w, f) for f, _ in w._fields_}

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
        self.lib.or_select_infector(_p(ci, C.c_float), len(ci), _p(u, C.c_double), _p(out, C.c_int32), len(u))
        return out

    def select_dest(self, r, cell, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros(len(u), np.int32)
        self.check(self.lib.or_select_dest_cell(self.h, int(r), int(cell), _p(u, C.c_double), _p(out, C.c_int32), len(u)))
        return out

    def transmission_coeffs(self, seed, prop_age_group=None, draws=False):
        """InitTransmissionCoeffs with the stream `seed`; with draws=True also
        the per-person (susc, infectiousness, infectious steps)."""
        out = self.abi.cs_transmission_coeffs()
        n = self.pop.n_people
        su, inf, rec = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.uint16)
        pa = None if prop_age_group is None else np.ascontiguousarray(prop_age_group, np.float64)
        self.check(self.lib.or_init_transmission_coeffs(self.h, None if pa is None else _p(pa, C.c_double), int(seed),
                                                        C.byref(out), _p(su, C.c_float), _p(inf, C.c_float),
                                                        _p(rec, C.c_uint16)))
        return (out, su, inf, rec) if draws else out

    def transmission_coeffs_from_draws(self, infness, rec, infstat=None, prop_age_group=None):
        out = self.abi.cs_transmission_coeffs()
        infness = np.ascontiguousarray(infness, np.float32)
        rec = np.ascontiguousarray(rec, np.uint16)
        st = None if infstat is None else np.ascontiguousarray(infstat, np.int8)
        pa = None if prop_age_group is None else np.ascontiguousarray(prop_age_group, np.float64)
        self.call("transmission_coeffs_from_draws", None if pa is None else _p(pa, C.c_double), _p(infness, C.c_float),
                  _p(rec, C.c_uint16), None if st is None else _p(st, C.c_int8), C.byref(out))
        return out

    def philox(self, ck):
        ck = np.ascontiguousarray(ck, np.uint32)
        n = len(ck) // 6
        out = np.zeros(4 * n, np.uint32)
        self.lib.or_philox(_p(ck, C.c_uint32), _p(out, C.c_uint32), n)
        return out

    def sample(self, which, key, c0, c1, a, b, n):
        """One draw per stream {key, c0, c1, i}, as the GPU's cs_sample."""
        out = np.zeros(n)
        one = np.zeros(1)
        for i in range(n):
            self.lib.or_sample(int(which), int(key), int(c0), int(c1), i, float(a), float(b), 1, _p(one, C.c_double))
            out[i] = one[0]
        return out


def household_ages(pg, seed, sizes):
    """AssignHouseholdAges of the oracle for the given household sizes."""
    lib, _ = load()
    sizes = np.ascontiguousarray(sizes, np.int32)
    out = np.zeros(int(sizes.sum()), np.uint8)
    rc = lib.or_household_ages(C.byref(pg), C.c_uint64(seed), iptr(sizes), len(sizes),
                               out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc:
        raise RuntimeError("or_household_ages: %d" % rc)
    return out


def nearest_places(kp, pg, tp, seed, persons, hloc, pmcell, pl_off, pl_loc, mc_country):
    """The oracle's nearest-places choice for the given persons."""
    lib, _ = load()
    persons = np.ascontiguousarray(persons, np.int32)
    hloc = np.ascontiguousarray(hloc, np.float32)
    pmcell = np.ascontiguousarray(pmcell, np.int32)
    pl_off = np.ascontiguousarray(pl_off, np.int64)
    pl_loc = np.ascontiguousarray(pl_loc, np.float32)
    mc_country = np.ascontiguousarray(mc_country, np.int32)
    out = np.zeros(len(persons), np.int32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    rc = lib.or_nearest_places(C.byref(kp), C.byref(pg), int(tp), C.c_uint64(seed), iptr(persons), fp(hloc), iptr(pmcell),
                               len(persons), pl_off.ctypes.data_as(C.POINTER(C.c_int64)), fp(pl_loc), iptr(mc_country),
                               iptr(out))
    if rc:
        raise RuntimeError("or_nearest_places: %d" % rc)
    return out
