"""Ensemble summaries of a batched run, in the layout of CovidSim's summary
outputs (SaveSummaryResults, src/CovidSim.cpp:1960-2400).

For every sample, each field's mean over the realisations, c * sum(x), and
its variance, c * sum(x^2) - (c * sum(x))^2 with c = 1 / realisations (the
TSMean / TSVar arithmetic of src/CovidSim.cpp:1976-1999), in tab-separated
columns named as the reference names them; variances carry the reference's
"_v" suffix (its .severity.xls convention).  Only the fields the device
series holds are written (cs_get_timeseries: totals, incidence of
infections, cases and deaths, severity incidence and prevalence); the
reference's intervention, contact-tracing and radius columns have no source
on this path.
"""
import numpy as np

from . import abi

# (reference column name, series field)
FIELDS = [
    ("S", abi.CS_TS_S), ("L", abi.CS_TS_L), ("I", abi.CS_TS_I), ("R", abi.CS_TS_R), ("D", abi.CS_TS_D),
    ("incI", abi.CS_TS_INC_I), ("incC", abi.CS_TS_INC_C), ("incDeath", abi.CS_TS_INC_DEATH),
    ("incMild", abi.CS_TS_INC_SEV + abi.CS_SK_MILD), ("incILI", abi.CS_TS_INC_SEV + abi.CS_SK_ILI),
    ("incSARI", abi.CS_TS_INC_SEV + abi.CS_SK_SARI), ("incCritical", abi.CS_TS_INC_SEV + abi.CS_SK_CRITICAL),
    ("incCritRecov", abi.CS_TS_INC_SEV + abi.CS_SK_CRITRECOV),
    ("incDeath_ILI", abi.CS_TS_INC_SEV + abi.CS_SK_DEATH_ILI),
    ("incDeath_SARI", abi.CS_TS_INC_SEV + abi.CS_SK_DEATH_SARI),
    ("incDeath_Critical", abi.CS_TS_INC_SEV + abi.CS_SK_DEATH_CRITICAL),
    ("Mild", abi.CS_TS_PREV_SEV + abi.CS_SK_MILD), ("ILI", abi.CS_TS_PREV_SEV + abi.CS_SK_ILI),
    ("SARI", abi.CS_TS_PREV_SEV + abi.CS_SK_SARI), ("Critical", abi.CS_TS_PREV_SEV + abi.CS_SK_CRITICAL),
    ("CritRecov", abi.CS_TS_PREV_SEV + abi.CS_SK_CRITRECOV),
]


def moments(series):
    """series [n_real, n_samples, CS_TS_FIELDS] -> (mean, variance), each
    [n_samples, len(FIELDS)], with the reference's arithmetic."""
    x = np.asarray(series, np.float64)[:, :, [f for _, f in FIELDS]]
    c = 1.0 / x.shape[0]
    mean = c * x.sum(axis=0)
    var = c * (x * x).sum(axis=0) - mean * mean
    return mean, var


def write(path, series, sample_days=1.0, t0=0.0):
    """Write the ensemble summary of `series` to `path` (one row per sample;
    t = t0 + (sample + 1) * sample_days, the time each sample is taken)."""
    mean, var = moments(series)
    names = [n for n, _ in FIELDS]
    with open(path, "w") as f:
        f.write("t\t" + "\t".join(names) + "\t" + "\t".join(n + "_v" for n in names) + "\n")
        for i in range(mean.shape[0]):
            row = [t0 + (i + 1) * sample_days] + list(mean[i]) + list(var[i])
            f.write("\t".join("%.10f" % v for v in row) + "\n")


def read(path):
    """(header names, values [n_samples, columns]) of a summary file."""
    with open(path) as f:
        names = f.readline().rstrip("\n").split("\t")
    return names, np.loadtxt(path, skiprows=1, ndmin=2)
