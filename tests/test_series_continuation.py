"""cs_run_steps continues one series across calls: a run split into chunks
at sample boundaries gives the same samples and end state as one call
(oracle on the CPU; the GPU, chunked, against the oracle in one call)."""
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
    assert a.shape == b.shape == (2, 60, 8)
    assert np.array_equal(a, b)
    assert ca0.cumI == cb0.cumI and ca1.cumI == cb1.cumI


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
