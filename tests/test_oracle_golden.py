"""The CPU oracle against vectors the reference itself produced.

The fixtures (tests/golden/mini_*) were written by the golden harness: the
unmodified CovidSim objects on the 20k-person fixture region, with the two
sweeps observed (oracle/golden_harness.cpp).  These tests pin the oracle's
restatement of each deterministic part of the path before the oracle is
trusted as the GPU's checker.
"""
import ctypes as C

import numpy as np
import pytest

from fixtures import hz_population, mini
from oracle_lib import Oracle, dptr, iptr, load


@pytest.fixture(scope="module")
def orc():
    arrays, pop, params, hzp = mini()
    o = Oracle(params, pop)
    o.call("alloc_realisations", 1, (C.c_uint64 * 1)(12345))
    return o


def test_kernel_lookup_bit_exact(orc):
    """KernelLookup::num (src/Kernels.cpp:108-129) on tables built from the
    kernel formula (src/Kernels.cpp:17-45): bit-exact at 9.5k radii,
    including r2 = 0, r2 = max and the hi-res/low-res boundary."""
    a = mini()[0]
    r2 = a["g_kernel_r2"].astype(np.float64)
    ref = a["g_kernel_num"]
    out = np.zeros_like(r2)
    orc.call("kernel_num", dptr(r2), dptr(out), len(r2))
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))


def test_kernel_out_of_range_is_an_error():
    """r2 beyond the table is fatal in the reference (src/Kernels.cpp:110-115)."""
    arrays, pop, params, hzp = mini()
    o = Oracle(params, pop)
    r2 = np.array([params.kernel_delta * params.kernel_size * 1.01])
    out = np.zeros(1)
    with pytest.raises(RuntimeError, match="r too large"):
        o.call("kernel_num", dptr(r2), dptr(out), 1)


def test_dist2_bit_exact(orc):
    """dist2 between persons (src/Dist.cpp:49-60), incl. a == b."""
    a = mini()[0]
    pa, pb = a["g_dist_a"].astype(np.int32), a["g_dist_b"].astype(np.int32)
    out = np.zeros(len(pa))
    orc.call("dist2_persons", iptr(pa), iptr(pb), dptr(out), len(pa))
    assert np.array_equal(out, a["g_dist2"])
    assert np.all(out[pa == pb] == 0)


def test_cell_tables_bit_exact(orc, csa):
    """max_trans (src/Kernels.cpp:48-63) and UpdateProbs' S0, cum_trans,
    tot_prob, InvCDF (src/CovidSim.cpp:2752-2802) against the reference's
    tables at the start of the run."""
    arrays, pop, params, hzp = mini()
    st = csa.population.State.from_dump(pop, arrays, "s0_")
    orc.call("load_state", 0, C.byref(st.view()))
    NC = pop.n_cells
    mt = np.zeros((NC, NC), np.float32)
    ct = np.zeros((NC, NC), np.float32)
    inv = np.zeros((NC, 1025), np.int32)
    tp = np.zeros(NC, np.float32)
    s0 = np.zeros(NC, np.int32)
    fp = lambda x: x.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    orc.call("get_cell_tables", 0, fp(mt), fp(ct), iptr(inv), fp(tp), iptr(s0))
    assert np.array_equal(mt.view(np.uint32), arrays["s0_max_trans"].view(np.uint32))
    assert np.array_equal(s0, arrays["s0_cell_S0"])
    assert np.array_equal(ct.view(np.uint32), arrays["s0_cum_trans"].view(np.uint32))
    assert np.array_equal(tp.view(np.uint32), arrays["s0_tot_prob"].view(np.uint32))
    assert np.array_equal(inv, arrays["s0_invcdf"])
    # InvCDF's definition, by brute force: the first m with cum[m]*1024 >= k
    for j in range(NC):
        for k in range(0, 1025, 7):
            m = int(np.argmax(ct[j] * np.float32(1024) >= np.float32(k)))
            assert inv[j, k] == m


def test_icdf_choose_bit_exact(orc):
    """InverseCdf::choose (src/InverseCdf.cpp:60-74) for the uniform the
    reference drew, over all 13 delay tables."""
    a = mini()[0]
    tab, u, mean, val = a["g_choose_table"], a["g_choose_u"], a["g_choose_mean"], a["g_choose_value"]
    for t in np.unique(tab):
        sel = tab == t
        uu, mm = u[sel].copy(), mean[sel].copy()
        out = np.zeros(sel.sum(), np.uint16)
        orc.call("icdf_choose", int(t), dptr(uu), dptr(mm), out.ctypes.data_as(C.POINTER(C.c_uint16)), len(uu))
        assert np.array_equal(out, val[sel]), "table %d" % t


def test_hazards_match_reference(csa):
    """The eight hazards of src/CalcInfSusc.cpp on 6000 synthetic person
    states with every intervention factor exercised (perturbed parameters).
    The reference computes in double; the oracle restates the same products
    in the same order, so the match is exact."""
    arrays, pop, params, hzp = mini()
    hzp.interventions_active = 1
    hpop = hz_population()
    o = Oracle(hzp, hpop)
    o.call("alloc_realisations", 1, (C.c_uint64 * 1)(1))
    st = csa.population.State.from_dump(hpop, arrays, "hz_")
    o.call("load_state", 0, C.byref(st.view()))
    person, other, pt, ts = arrays["g_hz_person"], arrays["g_hz_other"], arrays["g_hz_pt"], arrays["g_hz_ts"]
    ref = arrays["g_hz_out"]
    for kind in range(8):
        sec = pt if kind in (1, 5) else other
        out = np.zeros(len(person))
        for t in np.unique(ts):
            sel = np.nonzero(ts == t)[0]
            pp, oo = person[sel].astype(np.int32), sec[sel].astype(np.int32)
            tmp = np.zeros(len(sel))
            o.call("eval_hazards", 0, kind, int(t), iptr(pp), iptr(oo), dptr(tmp), len(sel))
            out[sel] = tmp
        rel = np.abs(out - ref[:, kind]) / np.maximum(np.abs(ref[:, kind]), 1e-300)
        assert rel.max() == 0.0, "hazard kind %d: max rel err %g" % (kind, rel.max())
    # the grid really exercises the factors: values vary and include non-trivial products
    assert len(np.unique(np.round(ref[:, 0], 12))) > 1000


def test_log_exp_accuracy():
    """The oracle's fixed log/exp polynomials (shared with the device) agree
    with libm to within 2 ulp over the ranges the samplers use."""
    lib, _ = load()
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(1e-300, 1, 20000), rng.uniform(1, 1e6, 20000), np.exp(rng.uniform(-700, 700, 20000)),
                         [1.0, 2.0, 0.5, 1e-308, 5e-324]])
    got = np.array([lib.or_log(float(x)) for x in xs])
    ref = np.log(xs)
    ulp = np.abs(got - ref) / np.spacing(np.abs(ref) + 1e-300)
    assert np.nanmax(ulp[ref != 0]) <= 2.0
    assert lib.or_log(1.0) == 0.0
    ys = np.concatenate([rng.uniform(-745, 709, 40000), rng.uniform(-1, 1, 20000), [0.0, -0.0]])
    got = np.array([lib.or_exp(float(y)) for y in ys])
    ref = np.exp(ys)
    ok = ref > 1e-300
    ulp = np.abs(got[ok] - ref[ok]) / np.spacing(ref[ok])
    assert ulp.max() <= 2.0
