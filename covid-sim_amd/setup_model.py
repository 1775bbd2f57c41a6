"""Population set-up on the GPU, with the reference's set-up file formats.

  read_density(path, pg)        the text density file as SetupModel reads it
                                (src/SetupModel.cpp:73-103): x y pop country
                                admin-code per line
  microcells(points, pg, codes) per-microcell mean density, country and
                                administrative unit, as SetupPopulation
                                aggregates the points (src/SetupModel.cpp:882-1008)
  write_density_bin(...)        the binary density file (/M:, 0xf0f0f0f0,
                                src/SetupModel.cpp:944-987)
  write_network / read_network  the network file (/S:, /L:,
                                src/SetupModel.cpp:2603-2693)
  read_admin_units(path)        the units an admin file defines (code,
                                country, name), as ReadParams keeps them
                                (src/ReadParams.cpp:1814-1833)
  generate(...)                 a whole population from the device generator
                                (cs_popgen_* of include/covidsim_hip.h), as
                                the arrays of population.Population
"""
import ctypes as C
import os

import numpy as np

from . import abi

NETWORK_FILE_VERSION = 1  # src/SetupModel.h:96
DENSITY_MAGIC = 0xF0F0F0F0
BINFILE = np.dtype([("x", "<f8"), ("y", "<f8"), ("pop", "<f8"), ("cnt", "<i4"), ("ad", "<i4")])  # src/SetupModel.h:98-102


def load_popgen_params(path):
    raw = open(path, "rb").read()
    g = abi.cs_popgen_params()
    if len(raw) != C.sizeof(g):
        raise ValueError("%s: %d bytes, cs_popgen_params is %d" % (path, len(raw), C.sizeof(g)))
    C.memmove(C.addressof(g), raw, len(raw))
    return g


def read_admin_units(path):
    """[(code, country, name)] of the units an admin file defines: the
    "Codes and country/province names for admin units" triples whose country
    is among "List of names of countries to include" and whose code is not
    zero, in file order (src/ReadParams.cpp:1814-1833).  A file without
    country names defines no named units: []."""
    blocks, key = {}, None
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("[") and line.endswith("]"):
                key = line[1:-1]
                blocks[key] = []
            elif line and key is not None:
                blocks[key] += line.split()
    trip = blocks.get("Codes and country/province names for admin units", [])
    countries = blocks.get("List of names of countries to include", [])
    nc = abs(int(blocks.get("Number of countries to include", ["0"])[0]))
    countries = countries[:nc]
    units = []
    for i in range(len(trip) // 3):
        code, cnt, name = trip[3 * i:3 * i + 3]
        if cnt in countries and int(code) != 0:
            units.append((int(code), cnt, name))
    return units


def read_density(path, pg):
    """Density points as BinFile records, longitudes west of the cut line
    moved by 360 degrees (src/SetupModel.cpp:73-103)."""
    rows = np.loadtxt(path, ndmin=2)
    pts = np.zeros(len(rows), BINFILE)
    x = rows[:, 0]
    pts["x"] = np.where(x >= pg.longitude_cut_line, x, x + 360)
    pts["y"] = rows[:, 1]
    pts["pop"] = rows[:, 2]
    pts["cnt"] = rows[:, 3].astype(np.int32)
    pts["ad"] = rows[:, 4].astype(np.int32) if (rows.shape[1] > 4 and pg.do_ad_units) else 0
    return pts


def microcells(points, pg, codes=None):
    """(mean density, country, administrative unit index, kept) per
    microcell, and the sums SetupPopulation keeps before averaging.

    codes maps an administrative code to its unit index (the units the
    admin file defines, with boundaries); None numbers the codes in order
    of first appearance (src/SetupModel.cpp:889-922).  Points outside the
    box or of a unit not in codes are dropped, as the reference drops them.
    """
    nm = pg.mcw * pg.mch
    dens = np.zeros(nm)
    num = np.zeros(nm, np.int64)
    country = np.ones(nm, np.int32)
    adunit = np.full(nm, -1, np.int32)
    x0, y0 = pg.bbox_x0, pg.bbox_y0
    x1 = x0 + pg.mcw * pg.mc_width
    y1 = y0 + pg.mch * pg.mc_height
    seen = {} if codes is None else dict(codes)
    kept = np.zeros(len(points), bool)
    for r, p in enumerate(points):
        code = int(p["ad"])
        if pg.do_ad_units:
            key = (code % pg.adunit_level1_mask) // pg.adunit_level1_divisor
            if codes is None:
                seen.setdefault(key, len(seen))
                unit = seen[key]
            else:
                if code not in seen:
                    continue
                unit = seen[code]
        else:
            unit = 0
        x, y = float(p["x"]), float(p["y"])
        if not (x0 <= x < x1 and y0 <= y < y1):
            continue
        j = int(np.floor((x - x0) / pg.mc_width + 0.1))
        k = int(np.floor((y - y0) / pg.mc_height + 0.1))
        l = j * pg.mch + k
        if l >= nm:
            continue
        dens[l] += float(p["pop"])
        country[l] = int(p["cnt"])
        num[l] += 1
        adunit[l] = unit
        kept[r] = True
    sums = dens.copy()
    mean = np.where(num > 0, dens / np.maximum(num, 1), 0.0)
    return mean, country, adunit, sums, kept


def write_density_bin(path, pg, sums, country, adunit, codes_of_unit):
    """The binary density file the reference writes from a text density
    (/M:): one record per microcell with a unit, at its centre, with the
    summed population (src/SetupModel.cpp:955-987)."""
    l = np.nonzero(adunit >= 0)[0]
    rec = np.zeros(len(l), BINFILE)
    rec["x"] = (pg.mc_width * ((l // pg.mch).astype(np.float64) + 0.5)) + pg.bbox_x0
    rec["y"] = (pg.mc_height * ((l % pg.mch).astype(np.float64) + 0.5)) + pg.bbox_y0
    rec["ad"] = [codes_of_unit[int(u)] if pg.do_ad_units else 0 for u in adunit[l]]
    rec["pop"] = sums[l]
    rec["cnt"] = country[l]
    with open(path, "wb") as f:
        f.write(np.array([DENSITY_MAGIC, len(l)], "<u4").tobytes())
        f.write(rec.tobytes())


def read_density_bin(path):
    raw = open(path, "rb").read()
    magic, n = np.frombuffer(raw[:8], "<u4")
    if magic != DENSITY_MAGIC:
        raise ValueError("%s: not a binary density file" % path)
    return np.frombuffer(raw[8:8 + int(n) * BINFILE.itemsize], BINFILE).copy()


def write_network(path, place_links, n_place_types, seed1, seed2):
    """The network file (/S:): version, place types, population, the two
    set-up seeds, then each person's links to places of the first
    n_place_types types (src/SetupModel.cpp:2661-2693)."""
    links = np.ascontiguousarray(np.asarray(place_links)[:, :n_place_types], "<i4")
    with open(path, "wb") as f:
        f.write(np.array([NETWORK_FILE_VERSION, n_place_types, len(links), seed1, seed2], "<i4").tobytes())
        f.write(links.tobytes())


def read_network(path):
    """(n_place_types, pop_size, seed1, seed2, links [pop][types]) of a
    network file (src/SetupModel.cpp:2603-2659)."""
    raw = open(path, "rb").read()
    ver, npt, n, s1, s2 = np.frombuffer(raw[:20], "<i4")
    if ver != NETWORK_FILE_VERSION:
        raise ValueError("%s: network file version %d" % (path, ver))
    links = np.frombuffer(raw[20:20 + 4 * int(npt) * int(n)], "<i4").reshape(int(n), int(npt)).copy()
    return int(npt), int(n), int(s1), int(s2), links


CELL_DEG = 0.075  # [Grid size] of the configs
NMCL = 9


def synthetic_region(n_people, ncw, nch, seed, base_pg, base_kp, lat0=25.0, sigma=1.5):
    """Inputs of the generator for a synthetic region of ncw x nch cells of
    0.075 degrees (9 x 9 microcells each) with its bottom edge at latitude
    lat0: a lognormal(3, sigma) mean density per microcell, one country and
    one administrative unit, the base parameter sets' other fields.
    Returns (cs_popgen_params, cs_params, density, country, adunit)."""
    from . import popgen
    pg = type(base_pg).from_buffer_copy(base_pg)
    kp = type(base_kp).from_buffer_copy(base_kp)
    pg.pop_size = int(n_people)
    pg.ncw, pg.nch, pg.nmcl = int(ncw), int(nch), NMCL
    pg.mcw, pg.mch = int(ncw) * NMCL, int(nch) * NMCL
    pg.mc_width = pg.mc_height = CELL_DEG / NMCL
    pg.bbox_x0, pg.bbox_y0 = 0.0, float(lat0)
    W, H = ncw * CELL_DEG, nch * CELL_DEG
    longest = max(popgen.dist2_utm(0, 0, W, H, lat0), popgen.dist2_utm(0, 0, W, 0, lat0), popgen.dist2_utm(0, 0, 0, H, lat0))
    popgen.apply_geometry(kp, {"nch": int(nch), "cell_width": CELL_DEG, "cell_height": CELL_DEG, "degrees_width": W,
                               "degrees_height": H, "bbox_bottom_left_y": float(lat0), "longest_dist2": longest})
    nm = pg.mcw * pg.mch
    rng = np.random.default_rng(seed)
    density = rng.lognormal(3.0, sigma, nm)
    return pg, kp, density, np.ones(nm, np.int32), np.zeros(nm, np.int32)


class PopGen:
    """The device generator (cs_popgen_*)."""

    def __init__(self, device=0):
        from .engine import lib
        self.L = lib()
        P = C.POINTER
        L = self.L
        L.cs_popgen_create.argtypes = [P(C.c_void_p), C.c_int]
        L.cs_popgen_destroy.argtypes = [C.c_void_p]
        L.cs_popgen_destroy.restype = None
        L.cs_popgen_last_error.argtypes = [C.c_void_p]
        L.cs_popgen_last_error.restype = C.c_char_p
        L.cs_popgen_run.argtypes = [C.c_void_p, P(abi.cs_popgen_params), P(abi.cs_params), P(C.c_double), P(C.c_int32),
                                    P(C.c_int32), C.c_uint64]
        L.cs_popgen_get_sizes.argtypes = [C.c_void_p, P(abi.cs_popgen_sizes)]
        L.cs_popgen_fetch.argtypes = [C.c_void_p, P(abi.cs_popgen_out)]
        L.cs_popgen_household_ages.argtypes = [C.c_void_p, P(abi.cs_popgen_params), C.c_uint64, P(C.c_int32), C.c_int,
                                               P(C.c_uint8)]
        self.h = C.c_void_p()
        rc = L.cs_popgen_create(C.byref(self.h), int(device))
        if rc:
            raise RuntimeError("cs_popgen_create: %d" % rc)

    def _check(self, rc):
        if rc:
            from .engine import CSError
            raise CSError(rc, self.L.cs_popgen_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.cs_popgen_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def household_ages(self, pg, seed, sizes):
        sizes = np.ascontiguousarray(sizes, np.int32)
        out = np.zeros(int(sizes.sum()), np.uint8)
        self._check(self.L.cs_popgen_household_ages(self.h, C.byref(pg), C.c_uint64(seed),
                                                    sizes.ctypes.data_as(C.POINTER(C.c_int32)), len(sizes),
                                                    out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def run(self, pg, kp, mcell_density, mcell_country, mcell_adunit, seed):
        d = np.ascontiguousarray(mcell_density, np.float64)
        c = np.ascontiguousarray(mcell_country, np.int32)
        a = np.ascontiguousarray(mcell_adunit, np.int32)
        self._check(self.L.cs_popgen_run(self.h, C.byref(pg), C.byref(kp), d.ctypes.data_as(C.POINTER(C.c_double)),
                                         c.ctypes.data_as(C.POINTER(C.c_int32)), a.ctypes.data_as(C.POINTER(C.c_int32)),
                                         C.c_uint64(seed)))
        self._adunit = a
        return self.sizes()

    def sizes(self):
        s = abi.cs_popgen_sizes()
        self._check(self.L.cs_popgen_get_sizes(self.h, C.byref(s)))
        return s

    def fetch(self, pg):
        """The generated population as population.Population arrays (plus
        pl{t}_loc, the place locations)."""
        s = self.sizes()
        N, NH, NC, NM = s.n_people, s.n_households, s.n_cells, s.n_mcells
        MPT = abi.CS_MAX_PLACE_TYPES
        a = {"p_hh": np.zeros(N, np.int32), "p_mcell": np.zeros(N, np.int32), "p_pcell": np.zeros(N, np.int32),
             "p_age": np.zeros(N, np.uint8), "p_flags": np.zeros(N, np.uint8),
             "p_place_links": np.zeros((N, MPT), np.int32), "p_place_group_links": np.zeros((N, MPT), np.uint16),
             "h_first": np.zeros(NH, np.int32), "h_nh": np.zeros(NH, np.uint16), "h_loc": np.zeros((NH, 2), np.float32),
             "c_index": np.zeros(NC, np.int32), "c_n": np.zeros(NC, np.int32),
             "c_member_start": np.zeros(NC + 1, np.int32), "c_members": np.zeros(N, np.int32),
             "m_adunit": np.zeros(NM, np.int32)}
        o = abi.cs_popgen_out()
        P = C.POINTER
        for f, k, t in [("hh", "p_hh", C.c_int32), ("mcell", "p_mcell", C.c_int32), ("pcell", "p_pcell", C.c_int32),
                        ("age", "p_age", C.c_uint8), ("place_links", "p_place_links", C.c_int32),
                        ("place_group_links", "p_place_group_links", C.c_uint16), ("hh_first", "h_first", C.c_int32),
                        ("hh_nh", "h_nh", C.c_uint16), ("hh_loc", "h_loc", C.c_float),
                        ("cell_index", "c_index", C.c_int32), ("cell_n", "c_n", C.c_int32),
                        ("cell_member_start", "c_member_start", C.c_int32), ("cell_members", "c_members", C.c_int32),
                        ("mcell_adunit", "m_adunit", C.c_int32)]:
            setattr(o, f, a[k].ctypes.data_as(P(t)))
        npt = pg.num_place_types
        for t in range(npt):
            npl, nm, ng = s.n_places[t], s.n_members[t], s.n_groups[t]
            pre = "pl%d_" % t
            a[pre + "n"] = np.zeros(npl, np.int32)
            a[pre + "mcell"] = np.zeros(npl, np.int32)
            a[pre + "loc"] = np.zeros((npl, 2), np.float32)
            a[pre + "member_start"] = np.zeros(npl + 1, np.int32)
            a[pre + "members"] = np.zeros(nm, np.int32)
            a[pre + "groups_start"] = np.zeros(npl + 1, np.int32)
            a[pre + "group_start"] = np.zeros(ng, np.int32)
            a[pre + "group_size"] = np.zeros(ng, np.int32)
            for f, k, ty in [("place_n", "n", C.c_int32), ("place_mcell", "mcell", C.c_int32),
                             ("place_member_start", "member_start", C.c_int32), ("place_members", "members", C.c_int32),
                             ("place_groups_start", "groups_start", C.c_int32),
                             ("place_group_start", "group_start", C.c_int32),
                             ("place_group_size", "group_size", C.c_int32), ("place_loc", "loc", C.c_float)]:
                getattr(o, f)[t] = a[pre + k].ctypes.data_as(P(ty))
        self._check(self.L.cs_popgen_fetch(self.h, C.byref(o)))
        a["h_nhr"] = a["h_nh"].copy()
        n_adunits = int(max(a["m_adunit"].max() + 1, 1))
        a["meta"] = np.array([N, NH, NC, s.n_cells_total, NM, n_adunits, npt, pg.nmcl], np.int32)
        return a


def generate(pg, kp, mcell_density, mcell_country, mcell_adunit, seed, device=0):
    """A population from the device generator: (arrays, sizes)."""
    g = PopGen(device)
    try:
        sizes = g.run(pg, kp, mcell_density, mcell_country, mcell_adunit, seed)
        return g.fetch(pg), sizes
    finally:
        g.close()
