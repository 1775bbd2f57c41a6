"""Synthetic populations in the cs_population layout, at any size.

A vectorised generator with the structure CovidSim's SetupModel produces
(src/SetupModel.cpp:848-2601): a density grid at 1/120 degree (one point per
microcell, 9x9 microcells per 0.075-degree cell), households drawn from a
size distribution and laid in contiguous person ranges inside one
microcell, ages from a 17-band pyramid, and four place types (primary and
secondary schools, universities, workplaces with power-law sizes) filled
with geographically close members and split into groups of the type's
group-size parameter.  It is not the reference's generator (which also
models household age structure and place choice by kernel); it reproduces
its scale and shape so the infection step sees realistic work.

generate() returns (population arrays, geometry) where geometry holds the
cs_params fields that depend on the region: nch, cell sizes, region size,
bounding-box latitude and the kernel table's delta (the longest dist2UTM
across the region divided by the table size, src/SetupModel.cpp:202-209).
"""
import math

import numpy as np

# UK age pyramid of tests/uk-input/admin-params.txt ([Age distribution of population])
AGE_PYRAMID = np.array([0.057810002, 0.060683584, 0.05827917, 0.054298865, 0.060021798, 0.066052894,
                        0.069334292, 0.067586774, 0.06346124, 0.063284341, 0.068269448, 0.066860984,
                        0.057523138, 0.04981529, 0.049914383, 0.035974234, 0.050829564])
HOUSEHOLD_PMF = np.array([0.30, 0.34, 0.15, 0.13, 0.05, 0.03])
CELL_DEG = 0.075
NMCL = 9
MC_DEG = CELL_DEG / NMCL

# place types: (min age, max age, proportion attending, mean size)
PLACE_TYPES = [(5, 10, 1.0, 200.0), (11, 17, 0.98, 800.0), (18, 21, 0.4, 3300.0), (18, 64, 0.75, 14.28)]
# workplace sizes: heavy-tailed, mean 14.3, capped at 5000 (the reference's
# 200k-person workload realises mean 14.2, maximum 173)
WORK_LOG_MU, WORK_LOG_SIGMA, WORK_MAX = 2.15, 1.0, 5000


def _sinx_tables():
    sinx = np.array([math.sin(math.pi * i / 180) for i in range(361)])
    cosx = np.array([math.cos(math.pi * i / 180) for i in range(361)])
    a = np.array([math.asin(math.sqrt(i / 1000.0)) for i in range(1001)])
    return sinx, cosx, a * a


def dist2_utm(x1, y1, x2, y2, lat0):
    """dist2UTM (src/Dist.cpp:23-48) for the kernel-table range."""
    sinx, cosx, asin2sqx = _sinx_tables()
    x = abs(x1 - x2) / 2
    y = abs(y1 - y2) / 2
    xi, yi = math.floor(x), math.floor(y)
    x -= xi
    y -= yi
    x = (1 - x) * sinx[int(xi)] + x * sinx[int(xi) + 1]
    y = (1 - y) * sinx[int(yi)] + y * sinx[int(yi) + 1]
    yt = abs(y1 + lat0)
    yi = math.floor(yt)
    cy1 = yt - yi
    cy1 = (1 - cy1) * cosx[int(yi)] + cy1 * cosx[int(yi) + 1]
    yt = abs(y2 + lat0)
    yi = math.floor(yt)
    cy2 = yt - yi
    cy2 = (1 - cy2) * cosx[int(yi)] + cy2 * cosx[int(yi) + 1]
    x = abs(1000 * (y * y + x * x * cy1 * cy2))
    xi = math.floor(x)
    x -= xi
    y = (1 - x) * asin2sqx[int(xi)] + x * asin2sqx[int(xi) + 1]
    er = ((1852 * 60 * 360) / 3.1415926535) / 2
    return 4 * er * er * y


def _fill(order_key, members, sizes, rng):
    """Assign members (sorted by order_key) to consecutive places of the given
    sizes; returns place index per member (or -1) and realised sizes."""
    m = members[np.argsort(order_key[members], kind="stable")]
    cap = np.cumsum(sizes)
    n = min(len(m), int(cap[-1]))
    place = np.searchsorted(cap, np.arange(n), side="right")
    out = np.full(len(members), -1, np.int64)
    idx = np.empty(len(m), np.int64)
    idx[:n] = place
    idx[n:] = -1
    back = np.empty(len(m), np.int64)
    back[np.argsort(order_key[members], kind="stable")] = np.arange(len(m))
    out = idx[back]
    real = np.bincount(place, minlength=len(sizes))
    return out, real


def generate(n_people, side_deg=None, seed=1, lat0=51.6, group_size=(25, 25, 100, 10)):
    rng = np.random.default_rng(seed)
    if side_deg is None:  # the 200k / 0.5-degree density of the reference workload
        side_deg = 0.5 * math.sqrt(n_people / 200000.0)
    npts = max(1, int(round(side_deg * 120)))
    ncw = 4 * int(math.ceil(npts / NMCL / 4))
    nch = ncw
    tot_mc_high = nch * NMCL
    # density: lognormal(3, 1.5) per microcell, multinomial allocation
    w = rng.lognormal(3.0, 1.5, size=npts * npts)
    cnt = rng.multinomial(n_people, w / w.sum())
    mx = np.repeat(np.arange(npts), npts)
    my = np.tile(np.arange(npts), npts)
    mcid = mx * tot_mc_high + my
    order = np.argsort(mcid, kind="stable")
    mcid, cnt = mcid[order], cnt[order]
    # persons in microcell order; households are runs of consecutive persons
    person_mc = np.repeat(mcid, cnt)
    sizes = rng.choice(np.arange(1, 7), size=n_people, p=HOUSEHOLD_PMF)
    cs = np.cumsum(sizes)
    nhh = int(np.searchsorted(cs, n_people)) + 1
    sizes = sizes[:nhh]
    sizes[-1] -= cs[nhh - 1] - n_people
    hfirst = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    hmc = person_mc[hfirst]
    hh = np.repeat(np.arange(nhh, dtype=np.int32), sizes)
    mcell = hmc[hh].astype(np.int32)  # a household lives in its first member's microcell
    hx = (hmc // tot_mc_high) * MC_DEG + rng.random(nhh) * MC_DEG
    hy = (hmc % tot_mc_high) * MC_DEG + rng.random(nhh) * MC_DEG
    hloc = np.stack([hx, hy], 1).astype(np.float32)
    cx = (mcell // tot_mc_high) // NMCL
    cy = (mcell % tot_mc_high) // NMCL
    pcell = (cx * nch + cy).astype(np.int32)
    ages = (rng.choice(17, size=n_people, p=AGE_PYRAMID / AGE_PYRAMID.sum()) * 5 + rng.integers(0, 5, n_people)).astype(np.uint8)
    # geographic order key for place filling: cell, then a random jitter
    gkey = pcell.astype(np.int64) * 1000000 + rng.integers(0, 1000000, n_people)
    plinks = np.full((n_people, 5), -1, np.int32)
    pglinks = np.zeros((n_people, 5), np.uint16)
    arrays = {}
    for t, (a0, a1, prop, mean) in enumerate(PLACE_TYPES):
        elig = np.nonzero((ages >= a0) & (ages <= a1) & (rng.random(n_people) < prop))[0]
        if t == 3:
            nplace = int(len(elig) / 14.28 * 1.02) + 1
            s = np.minimum(WORK_MAX, np.ceil(rng.lognormal(WORK_LOG_MU, WORK_LOG_SIGMA, nplace))).astype(np.int64)
        else:
            nplace = max(1, int(round(len(elig) / mean)))
            s = np.maximum(1, rng.poisson(len(elig) / nplace, nplace))
        place, real = _fill(gkey, elig, s, rng)
        keep = real > 0
        remap = np.cumsum(keep) - 1
        ok = place >= 0
        members_p = elig[ok]
        pl = remap[place[ok]]
        real = real[keep]
        np_ = len(real)
        order_m = np.lexsort((rng.random(len(pl)), pl))  # group members: place, then shuffled
        mem = members_p[order_m].astype(np.int32)
        mstart = np.concatenate([[0], np.cumsum(real)]).astype(np.int32)
        gs = group_size[t]
        ng = np.maximum(1, np.ceil(real / gs).astype(np.int64))
        gstart_list, gsize_list = [], []
        within = np.arange(len(mem)) - np.repeat(mstart[:-1], real)
        per = np.ceil(real / ng).astype(np.int64)
        gidx = within // np.repeat(per, real)
        for k_ in range(np_):
            n_ = real[k_]
            p_ = per[k_]
            starts = np.arange(0, n_, p_)
            gstart_list.append(starts)
            gsize_list.append(np.minimum(p_, n_ - starts))
        gstart = np.concatenate(gstart_list).astype(np.int32)
        gsize = np.concatenate(gsize_list).astype(np.int32)
        gs0 = np.concatenate([[0], np.cumsum([len(x) for x in gstart_list])]).astype(np.int32)
        plinks[mem, t] = np.repeat(np.arange(np_, dtype=np.int32), real)
        pglinks[mem, t] = gidx.astype(np.uint16)
        pmc = mcell[mem[mstart[:-1]]]
        arrays.update({"pl%d_n" % t: real.astype(np.int32), "pl%d_mcell" % t: pmc.astype(np.int32),
                       "pl%d_member_start" % t: mstart, "pl%d_members" % t: mem,
                       "pl%d_groups_start" % t: gs0, "pl%d_group_start" % t: gstart,
                       "pl%d_group_size" % t: gsize})
    # populated cells in ascending Cells order, members in person order
    ncell_tot = ncw * nch
    counts = np.bincount(pcell, minlength=ncell_tot)
    cidx = np.nonzero(counts)[0].astype(np.int32)
    corder = np.argsort(pcell, kind="stable").astype(np.int32)
    cn = counts[cidx].astype(np.int32)
    cstart = np.concatenate([[0], np.cumsum(cn)]).astype(np.int32)
    nm = ncell_tot * NMCL * NMCL
    arrays.update({
        "p_hh": hh, "p_mcell": mcell, "p_pcell": pcell, "p_age": ages,
        "p_flags": np.zeros(n_people, np.uint8), "p_place_links": plinks, "p_place_group_links": pglinks,
        "h_first": hfirst, "h_nh": sizes.astype(np.uint16), "h_nhr": sizes.astype(np.uint16), "h_loc": hloc,
        "c_index": cidx, "c_n": cn, "c_member_start": cstart, "c_members": corder,
        "m_adunit": np.zeros(nm, np.int32),
        "meta": np.array([n_people, nhh, len(cidx), ncell_tot, nm, 1, 4, NMCL], np.int32),
    })
    W = ncw * CELL_DEG
    H = nch * CELL_DEG
    longest = max(dist2_utm(0, 0, W, H, lat0), dist2_utm(0, 0, W, 0, lat0), dist2_utm(0, 0, 0, H, lat0))
    geom = {"nch": nch, "cell_width": CELL_DEG, "cell_height": CELL_DEG, "degrees_width": W,
            "degrees_height": H, "bbox_bottom_left_y": lat0, "longest_dist2": longest}
    return arrays, geom


def apply_geometry(params, geom):
    """Set the region-dependent cs_params fields for a generated population."""
    params.nch = geom["nch"]
    params.cell_width = geom["cell_width"]
    params.cell_height = geom["cell_height"]
    params.degrees_width = geom["degrees_width"]
    params.degrees_height = geom["degrees_height"]
    params.bbox_bottom_left_y = geom["bbox_bottom_left_y"]
    params.kernel_delta = geom["longest_dist2"] / params.kernel_size
    return params
