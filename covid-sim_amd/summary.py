"""Ensemble summaries of a batched run, in the layout of CovidSim's summary
outputs (SaveSummaryResults, src/CovidSim.cpp:1960-2400): write (totals and
severity), write_age (.age.xls) and write_adunit (.adunit.xls).

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


def _mean(x):
    """c * sum over realisations (TSMean arithmetic), c = 1 / realisations"""
    x = np.asarray(x, np.float64)
    return (1.0 / x.shape[0]) * x.sum(axis=0)


def _rows(series, start, sample_days, t0):
    """(times, per-sample means) with the optional sample taken at t0 first"""
    m = _mean(series)
    if start is None:
        return [t0 + (i + 1) * sample_days for i in range(m.shape[0])], m
    return ([t0 + i * sample_days for i in range(m.shape[0] + 1)],
            np.concatenate([_mean(start)[None], m], axis=0))


def write_age(path, series_age, ages, sample_days=1.0, t0=0.0, start=None):
    """The ensemble's age file (.age.xls of SaveSummaryResults,
    src/CovidSim.cpp:2035-2063): per sample the mean number of infections,
    cases and deaths of each 5-year age group since the realisation began,
    then a `dist` row with the population's count in each group (AgeDist).
    The reference's columns are cumulative although its fields are named
    incIa, incCa, incDa: RecordSample copies StateT's cumIa, cumCa, cumDa
    (src/CovidSim.cpp:3392-3400), which only InitModel zeroes (:928, :954).
    series_age is [n_real, n_samples, CS_TSA_FIELDS, CS_NUM_AGE_GROUPS]
    (Engine.timeseries_age, per-sample incidence), accumulated here per
    realisation; ages the persons' ages in years.  start, shaped
    [n_real, CS_TSA_FIELDS, CS_NUM_AGE_GROUPS], holds the counts at t0 (the
    seed infections: Engine.counters' cumIa and cumCa after the reset): given,
    the accumulation starts from it and it is the first row, as in the
    reference's file; without it the counts start from zero."""
    g = abi.CS_NUM_AGE_GROUPS
    x = np.asarray(series_age, np.int64)
    flds = [abi.CS_TSA_INC_I, abi.CS_TSA_INC_C, abi.CS_TSA_INC_DEATH]
    base = np.zeros((x.shape[0], len(flds), g), np.int64) if start is None else np.asarray(start, np.int64)[:, flds]
    cum = base[:, None] + np.cumsum(x[:, :, flds], axis=1)  # [n_real, n_samples, 3, groups]
    if start is not None:
        cum = np.concatenate([base[:, None], cum], axis=1)
        ts = [t0 + i * sample_days for i in range(cum.shape[1])]
    else:
        ts = [t0 + (i + 1) * sample_days for i in range(cum.shape[1])]
    m = _mean(cum)
    cols = []
    for tag in ("I", "C", "D"):
        cols += ["%s%d-%d" % (tag, 5 * i, 5 * (i + 1)) for i in range(g)]
    dist = np.bincount(np.minimum(np.asarray(ages, np.int64) // 5, g - 1), minlength=g)
    with open(path, "w") as f:
        f.write("t\t" + "\t".join(cols) + "\n")
        for i, t in enumerate(ts):
            f.write("\t".join("%.10f" % v for v in [t] + list(m[i].reshape(-1))) + "\n")
        f.write("dist" + "".join("\t%.10f" % v for v in dist) + "\n")


def write_adunit(path, series_adunit, names, unit_n, pop_by_adunit=None, sample_days=1.0, t0=0.0,
                 start=None):
    """The ensemble's administrative-unit file (.adunit.xls of
    SaveSummaryResults, src/CovidSim.cpp:2065-2090): a header of columns
    I_<name>, C_<name>, DC_<name>, T_<name> followed by each unit's person
    count (unit_n, AdUnits[i].n) and its PopByAdunit[i][1] entry
    (pop_by_adunit, zero when not given, as it stays unless a regional
    demography file is read, src/SetupModel.cpp:1018-1045); then per sample the mean
    incidence of infections, cases and detected cases of each unit and the
    mean cumulative treated count, which is zero: treatment is not part of
    this path.  series_adunit is [n_real, n_samples, CS_TSU_FIELDS, n_adunits]
    (Engine.timeseries_adunit); start, [n_real, CS_TSU_FIELDS, n_adunits],
    the sample taken at t0 (Engine.adunit_counters after the reset), written
    first when given."""
    na = len(names)
    ts, m = _rows(series_adunit, start, sample_days, t0)
    if pop_by_adunit is None:
        pop_by_adunit = np.zeros(na)
    cols = []
    for tag in ("I", "C", "DC", "T"):
        cols += ["%s_%s" % (tag, n) for n in names]
    with open(path, "w") as f:
        f.write("t\t" + "\t".join(cols) + "".join("\t%d" % int(v) for v in unit_n[:na])
                + "".join("\t%.10f" % v for v in pop_by_adunit[:na]) + "\n")
        for i, t in enumerate(ts):
            row = [t]
            for fld in (abi.CS_TSU_INC_I, abi.CS_TSU_INC_C, abi.CS_TSU_INC_DC):
                row += list(m[i, fld, :na])
            row += [0.0] * na
            f.write("\t".join("%.10f" % v for v in row) + "\n")
