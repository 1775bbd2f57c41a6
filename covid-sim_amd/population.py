"""Host-side containers for the C ABI's population and state views.

A population is a dict of numpy arrays with the names the reference-export
format uses (one .npy file per array, written by a run of CovidSim's own
set-up code or by the synthetic generator in this package):

  p_hh p_mcell p_pcell p_age p_flags p_place_links p_place_group_links
  h_first h_nh h_nhr h_loc  c_index c_n c_member_start c_members  m_adunit
  pl{t}_n pl{t}_mcell pl{t}_member_start pl{t}_members pl{t}_groups_start
  pl{t}_group_start pl{t}_group_size   meta  kernel_lookup kernel_hi_res

`meta` = [n_people, n_households, n_cells, n_cells_total, n_mcells,
n_adunits, n_place_types, nmcl].  params.bin holds a raw cs_params.
"""
import ctypes as C
import os

import numpy as np

from . import abi

NPT_MAX = abi.CS_MAX_PLACE_TYPES

_PERSON_STATE = [
    # (cs_state field, dump name, dtype)
    ("inf", "inf", np.int8), ("severity_current", "sev_cur", np.uint8),
    ("severity_final", "sev_final", np.uint8), ("bits", "bits", np.uint8),
    ("listpos", "listpos", np.int32), ("infector", "infector", np.int32),
    ("infect_type", "infect_type", np.int16), ("infectiousness", "infectiousness", np.float32),
    ("susc", "susc", np.float32), ("prob_care", "prob_care", np.float32),
    ("infection_time", "t_infection", np.uint16), ("latent_time", "t_latent", np.uint16),
    ("recovery_or_death_time", "t_recovery", np.uint16), ("sari_time", "t_sari", np.uint16),
    ("critical_time", "t_critical", np.uint16), ("stepdown_time", "t_stepdown", np.uint16),
    ("detected_time", "t_detected", np.uint16), ("absent_start_time", "t_absent_start", np.uint16),
    ("absent_stop_time", "t_absent_stop", np.uint16), ("treat_start_time", "t_treat_start", np.uint16),
    ("treat_stop_time", "t_treat_stop", np.uint16), ("vacc_start_time", "t_vacc_start", np.uint16),
    ("isolation_start_time", "t_isolation_start", np.uint16), ("quar_start_time", "t_quar_start", np.uint16),
    ("quar_comply", "quar_comply", np.uint8), ("dct_traced", "dct_traced", np.uint8),
]
_CELL_STATE = [("cell_S", "cell_S"), ("cell_L", "cell_L"), ("cell_I", "cell_I"), ("cell_R", "cell_R"),
               ("cell_D", "cell_D"), ("cell_cumTC", "cell_cumTC")]


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


_CT = {np.int8: C.c_int8, np.uint8: C.c_uint8, np.int16: C.c_int16, np.uint16: C.c_uint16,
       np.int32: C.c_int32, np.float32: C.c_float, np.float64: C.c_double}


def load_params(path):
    raw = open(path, "rb").read()
    if len(raw) != C.sizeof(abi.cs_params):
        raise ValueError("%s: %d bytes, cs_params is %d" % (path, len(raw), C.sizeof(abi.cs_params)))
    p = abi.cs_params()
    C.memmove(C.addressof(p), raw, len(raw))
    return p


def save_params(p, path):
    with open(path, "wb") as f:
        f.write(bytes(p))


class Population:
    """Population arrays plus a cs_population view that points into them."""

    def __init__(self, arrays):
        self.a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        meta = self.a["meta"]
        self.n_people, self.n_households, self.n_cells, self.n_cells_total, self.n_mcells, \
            self.n_adunits, self.n_place_types = [int(x) for x in meta[:7]]
        self.nmcl = int(meta[7]) if len(meta) > 7 else 0
        self.n_places = [len(self.a["pl%d_n" % t]) if t < self.n_place_types else 0 for t in range(NPT_MAX)]
        self._view = None

    @classmethod
    def from_dir(cls, d):
        arrays = {}
        for fn in os.listdir(d):
            if fn.endswith(".npy"):
                name = fn[:-4]
                if name.startswith(("p_", "h_", "c_", "m_", "pl", "meta", "kernel_")) and "close" not in name:
                    arrays[name] = np.load(os.path.join(d, fn))
        return cls(arrays)

    def view(self):
        if self._view is not None:
            return self._view
        a = self.a
        v = abi.cs_population()
        v.n_people, v.n_households, v.n_cells = self.n_people, self.n_households, self.n_cells
        v.n_cells_total, v.n_mcells, v.n_adunits = self.n_cells_total, self.n_mcells, self.n_adunits
        for f, name, t in [("hh", "p_hh", C.c_int32), ("mcell", "p_mcell", C.c_int32), ("pcell", "p_pcell", C.c_int32),
                           ("age", "p_age", C.c_uint8), ("flags", "p_flags", C.c_uint8),
                           ("place_links", "p_place_links", C.c_int32),
                           ("place_group_links", "p_place_group_links", C.c_uint16),
                           ("hh_first", "h_first", C.c_int32), ("hh_nh", "h_nh", C.c_uint16),
                           ("hh_nhr", "h_nhr", C.c_uint16), ("hh_loc", "h_loc", C.c_float),
                           ("cell_index", "c_index", C.c_int32), ("cell_n", "c_n", C.c_int32),
                           ("cell_member_start", "c_member_start", C.c_int32),
                           ("cell_members", "c_members", C.c_int32), ("mcell_adunit", "m_adunit", C.c_int32)]:
            setattr(v, f, _ptr(a[name], t))
        for t in range(NPT_MAX):
            if t < self.n_place_types:
                v.n_places[t] = self.n_places[t]
                for f, suf in [("place_n", "n"), ("place_mcell", "mcell"), ("place_member_start", "member_start"),
                               ("place_members", "members"), ("place_groups_start", "groups_start"),
                               ("place_group_start", "group_start"), ("place_group_size", "group_size")]:
                    arr = a["pl%d_%s" % (t, suf)]
                    if arr.size == 0:
                        arr = a["pl%d_%s" % (t, suf)] = np.zeros(1, np.int32)
                    getattr(v, f)[t] = _ptr(arr, C.c_int32)
        self._view = v
        return v


class State:
    """One realisation's dynamic state with a cs_state view."""

    def __init__(self, pop, arrays=None):
        self.pop = pop
        N, NC, NM = pop.n_people, pop.n_cells, pop.n_mcells
        self.a = {}
        for f, name, dt in _PERSON_STATE:
            self.a[f] = np.zeros(N, dt)
        for f, name in _CELL_STATE:
            self.a[f] = np.zeros(NC, np.int32)
        self.a["cell_perm"] = np.zeros(N, np.int32)
        for t in range(pop.n_place_types):
            self.a["pcs%d" % t] = np.zeros(pop.n_places[t], np.uint16)
            self.a["pce%d" % t] = np.zeros(pop.n_places[t], np.uint16)
        self.a["mcell_socdist"] = np.zeros(NM, np.uint8)
        self.a["mcell_moverest"] = np.zeros(NM, np.uint8)
        if arrays:
            for k, v in arrays.items():
                self.a[k][...] = v
        self._view = None

    @classmethod
    def from_dump(cls, pop, src, prefix):
        """State named `prefix*` from a dump directory or a mapping of arrays."""
        if isinstance(src, str):
            get = lambda n: np.load(os.path.join(src, n + ".npy"))  # noqa: E731
        else:
            get = lambda n: src[n]  # noqa: E731
        s = cls(pop)
        for f, name, dt in _PERSON_STATE:
            s.a[f][...] = get(prefix + name).astype(dt)
        for f, name in _CELL_STATE:
            s.a[f][...] = get(prefix + name)
        s.a["cell_perm"][...] = get(prefix + "cell_perm")
        for t in range(pop.n_place_types):
            s.a["pcs%d" % t][...] = get("%spl%d_close_start" % (prefix, t))
            s.a["pce%d" % t][...] = get("%spl%d_close_end" % (prefix, t))
        s.a["mcell_socdist"][...] = get(prefix + "m_socdist")
        s.a["mcell_moverest"][...] = get(prefix + "m_moverest")
        return s

    def view(self):
        if self._view is not None:
            return self._view
        v = abi.cs_state()
        for f, name, dt in _PERSON_STATE:
            setattr(v, f, _ptr(self.a[f], _CT[dt]))
        for f, name in _CELL_STATE:
            setattr(v, f, _ptr(self.a[f], C.c_int32))
        v.cell_perm = _ptr(self.a["cell_perm"], C.c_int32)
        for t in range(self.pop.n_place_types):
            arr_s, arr_e = self.a["pcs%d" % t], self.a["pce%d" % t]
            if arr_s.size == 0:
                arr_s = self.a["pcs%d" % t] = np.zeros(1, np.uint16)
                arr_e = self.a["pce%d" % t] = np.zeros(1, np.uint16)
            v.place_close_start[t] = _ptr(arr_s, C.c_uint16)
            v.place_close_end[t] = _ptr(arr_e, C.c_uint16)
        v.mcell_socdist = _ptr(self.a["mcell_socdist"], C.c_uint8)
        v.mcell_moverest = _ptr(self.a["mcell_moverest"], C.c_uint8)
        self._view = v
        return v

    def compare(self, other, fields=None):
        """Names of the arrays that differ from other's."""
        bad = []
        for k, v in self.a.items():
            if fields and k not in fields:
                continue
            if not np.array_equal(v, other.a[k]):
                bad.append(k)
        return bad
