"""CPU-side checks: the C ABI library's exports, and the oracle's samplers
against exact distributions (the GPU samplers are bit-compared with these
in the GPU tests, so pinning the oracle's pins both)."""
import ctypes as C
import os
import re

import numpy as np
import pytest
from scipy import stats

from oracle_lib import load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    """libcovidsim_hip.so loads without a GPU and exports what the header declares."""
    path = os.path.join(ROOT, "covid-sim_amd", "libcovidsim_hip.so")
    if not os.path.exists(path):
        pytest.skip("library not built")
    lib = C.CDLL(path)
    text = open(os.path.join(ROOT, "include", "covidsim_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", text)))
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def _draws(which, a, b, n, key=99, c0=7, c1=3):
    lib, _ = load()
    out = np.zeros(1)
    vals = np.zeros(n)
    for i in range(n):
        lib.or_sample(which, key, c0, c1, i, float(a), float(b), 1, out.ctypes.data_as(C.POINTER(C.c_double)))
        vals[i] = out[0]
    return vals


def _chi2_pvalue(x, pmf, support):
    obs = np.array([(x == k).sum() for k in support], float)
    exp = pmf * len(x)
    keep = exp >= 5
    o = np.append(obs[keep], obs[~keep].sum() + (len(x) - obs.sum()))
    e = np.append(exp[keep], exp[~keep].sum() + len(x) * (1 - pmf.sum()))
    m = e > 0
    return stats.chisquare(o[m], e[m] * o[m].sum() / e[m].sum()).pvalue


@pytest.mark.parametrize("n,p", [(20, 0.1), (200, 0.04), (25, 0.6), (3000, 0.02), (500, 0.3), (5000, 0.97)])
def test_binomial_matches_pmf(n, p):
    """Both branches (inversion below n*min(p,1-p)=10, BTRS above) and the
    p > 1/2 reflection follow Binomial(n,p) (cf. ignbin_mt, src/Rand.cpp:404-662)."""
    x = _draws(1, n, p, 20000)
    support = np.arange(0, n + 1)
    assert _chi2_pvalue(x, stats.binom.pmf(support, n, p), support) > 1e-3
    assert abs(x.mean() - n * p) < 5 * np.sqrt(n * p * (1 - p) / len(x))


@pytest.mark.parametrize("mu", [0.3, 3.0, 9.9, 10.0, 37.0, 400.0])
def test_poisson_matches_pmf(mu):
    """Inversion below 10 and PTRS from 10 follow Poisson(mu) (cf. ignpoi_mt)."""
    x = _draws(2, mu, 0, 20000)
    support = np.arange(0, int(mu + 12 * np.sqrt(mu) + 20))
    assert _chi2_pvalue(x, stats.poisson.pmf(support, mu), support) > 1e-3


def test_normal_and_gamma_match():
    x = _draws(3, 0, 0, 20000)
    assert stats.kstest(x, "norm").pvalue > 1e-3
    for rate, shape in [(1.0, 1.0), (4.0, 4.0), (2.0, 0.5)]:
        g = _draws(4, rate, shape, 20000)
        assert stats.kstest(g, stats.gamma(a=shape, scale=1 / rate).cdf).pvalue > 1e-3


@pytest.mark.parametrize("k,n", [(1, 7), (2, 2), (5, 40), (32, 33), (33, 200), (150, 200), (200, 200)])
def test_sample_without_replacement_uniform(k, n):
    """Floyd (k <= 32), selection sampling (k > 32) and the k == n case all
    give distinct indices, each included with probability k/n
    (SampleWithoutReplacement, src/Rand.cpp:1109-1243)."""
    lib, _ = load()
    reps = 4000
    counts = np.zeros(n)
    out = np.zeros(n)
    for i in range(reps):
        m = lib.or_sample(5, 1234, 55, i, 0, float(k), float(n), n, out.ctypes.data_as(C.POINTER(C.c_double)))
        assert m == k
        idx = out[:k].astype(int)
        assert len(set(idx)) == k and idx.min() >= 0 and idx.max() < n
        counts[idx] += 1
    if k < n:
        assert stats.chisquare(counts).pvalue > 1e-3
        assert abs(counts.mean() / reps - k / n) < 1e-9


def test_philox_known_answers_oracle():
    lib, _ = load()
    ck = np.array([0, 0, 0, 0, 0, 0, 0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0], np.uint32)
    out = np.zeros(8, np.uint32)
    lib.or_philox(ck.ctypes.data_as(C.POINTER(C.c_uint32)), out.ctypes.data_as(C.POINTER(C.c_uint32)), 2)
    assert list(out) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8, 0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]
