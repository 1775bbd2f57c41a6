"""cs_run_steps continues one series across calls: a run split into chunks
at sample boundaries gives the same samples and end state as one call
(oracle on the CPU; the GPU, chunked, against the oracle in one call).
The series' severity fields are pinned by invariants on the oracle."""
import numpy as np
import pytest

from fixtures import mini
from oracle_lib import OracleEngine


def run(engine_cls_or_obj, chunks, days=60):
    arrays, pop, params, _ = mini()
    e = engine_cls_or_obj(params, pop)
    e.alloc(np.array([11, 12], np.uint64))
    e.reset(10)
    tspd = int(params.time_steps_per_day)
    d0 = 0
    for c in chunks:
        e.run_steps(float(d0), c * tspd, tspd)
        d0 += c
    assert d0 == days
    return e.timeseries(days + 5), e.counters(0), e.counters(1)


def test_oracle_chunked_series_equals_single_call():
    a, ca0, ca1 = run(OracleEngine, [60])
    b, cb0, cb1 = run(OracleEngine, [25, 5, 30])
    assert a.shape == b.shape == (2, 60, 21)
    assert np.array_equal(a, b)
    assert ca0.cumI == cb0.cumI and ca1.cumI == cb1.cumI


def test_oracle_series_severity_fields():
    """The severity fields of the series agree with what fixes them
    independently: the incidence columns sum to the cumulative counters,
    every death comes from one of the three death kinds, and the last
    prevalence equals a count of the stored persons by current severity."""
    import importlib
    csa = importlib.import_module("covid-sim_amd")
    abi = importlib.import_module("covid-sim_amd.abi")
    arrays, pop, params, _ = mini()
    e = OracleEngine(params, pop)
    e.alloc(np.array([11, 12], np.uint64))
    e.reset(10)
    tspd = int(params.time_steps_per_day)
    e.run_steps(0.0, 90 * tspd, tspd)
    ts = e.timeseries(90)
    inc, prev = abi.CS_TS_INC_SEV, abi.CS_TS_PREV_SEV
    deaths = slice(inc + abi.CS_SK_DEATH_ILI, inc + abi.CS_SK_DEATH_CRITICAL + 1)
    for r in range(2):
        c = e.counters(r)
        s = ts[r]
        assert np.array_equal(s[:, inc:inc + abi.CS_NUM_SEV_KINDS].sum(0), np.array(c.sev_cum[:abi.CS_NUM_SEV_KINDS]))
        assert np.array_equal(s[:, abi.CS_TS_INC_DEATH], s[:, deaths].sum(1))
        assert s[:, abi.CS_TS_INC_DEATH].sum() == s[-1, abi.CS_TS_D] > 0
        st = csa.population.State(pop)
        e.store_state(r, st)
        sev = st.a["severity_current"]
        count = [np.count_nonzero(sev == k) for k in (abi.CS_SEV_MILD, abi.CS_SEV_ILI, abi.CS_SEV_SARI,
                                                      abi.CS_SEV_CRITICAL, abi.CS_SEV_STEPDOWN)]
        assert list(s[-1, prev:prev + 5]) == count, (list(s[-1, prev:prev + 5]), count)


def gpu_engine(params, pop):
    import importlib
    g = importlib.import_module("covid-sim_amd.engine").Engine(0)
    g.set_params(params)
    g.set_population(pop)
    return g


@pytest.mark.gpu
def test_gpu_chunked_series_equals_oracle():
    a, ca0, ca1 = run(OracleEngine, [60])
    b, cb0, cb1 = run(gpu_engine, [25, 5, 30])
    assert np.array_equal(a, b)
    assert ca0.cumI == cb0.cumI and ca1.cumI == cb1.cumI
