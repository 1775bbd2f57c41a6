"""Write the parameter and density files of the 200k-person single-region
workload (and of a small 20k-person fixture region) in the reference's own
input formats.

The parameter files start from the reference's UK integration-test inputs
(tests/uk-input/{pre,admin,input-noint}-params.txt, read from REF) and
override the keys that define the workload:
  * 0.5 x 0.5 degree region, 1/120 degree density grid, lognormal(3, 1.5)
    cell populations (numpy default_rng(12345)) scaled to the target size;
  * household sizes 0.30/0.34/0.15/0.13/0.05/0.03;
  * school means 200/800, workplace power law on [1, 5000] with power 1.6;
  * power-law kernel a = 4 km, b = 3; R0 = 2.4; 10 seeds at random places;
  * 200 days at 4 steps/day; no interventions; in-run calibration and
    bitmaps off.
The density format is `x y pop country admin` per line
(src/SetupModel.cpp:73-103).

Run once in a container that has the reference mounted; the outputs are
committed, so nothing at test or bench time reads the reference.
"""
import os
import re
import sys

import numpy as np

REF = os.environ.get("REF", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "..", "..", "covid-sim_amd", "inputs")


def override(text, values):
    """Replace the value block after each `[Key]` header named in values."""
    lines = text.split("\n")
    out, i, seen = [], 0, set()
    while i < len(lines):
        line = lines[i]
        m = re.match(r"^\[(.*)\]\s*$", line)
        if m and m.group(1) in values:
            key = m.group(1)
            out.append(line)
            out.append(str(values[key]))
            seen.add(key)
            i += 1
            # skip the old value lines up to the next blank line or header
            while i < len(lines) and lines[i].strip() and not lines[i].startswith("["):
                i += 1
            continue
        out.append(line)
        i += 1
    for key, val in values.items():
        if key not in seen:
            out += ["", "[%s]" % key, str(val)]
    return "\n".join(out) + "\n"


def density(path, lon0, lat0, n_side, total, seed):
    rng = np.random.default_rng(seed)
    pop = rng.lognormal(mean=3.0, sigma=1.5, size=n_side * n_side)
    pop *= total / pop.sum()
    with open(path, "w") as f:
        k = 0
        for i in range(n_side):
            for j in range(n_side):
                x = lon0 + (i + 0.5) / 120.0
                y = lat0 + (j + 0.5) / 120.0
                f.write("%.6f\t%.6f\t%.6f\t54\t541600\n" % (x, y, pop[k]))
                k += 1


def make(name, n_side, total, days):
    d = os.path.join(OUT, name)
    os.makedirs(d, exist_ok=True)
    src = os.path.join(REF, "tests", "uk-input")
    pre = open(os.path.join(src, "pre-params.txt")).read()
    adm = open(os.path.join(src, "admin-params.txt")).read()
    noint = open(os.path.join(src, "input-noint-params.txt")).read()
    pre = override(pre, {
        "Output bitmap": 0,
        "Sampling time": days,
        "Initial number of infecteds": 10,
        "Reproduction number": 2.4,
        "Day of year trigger is reached": -1,
        "Alert trigger starts after interventions": 0,
        "Use global triggers for interventions": 0,
        "Use ICU case triggers for interventions": 0,
    })
    adm = override(adm, {
        "Initial number of infecteds": 10,
        "Population size": total,
        "Fix population size at specified value": 1,
        "Household size distribution": "0.30\t0.34\t0.15\t0.13\t0.05\t0.03\t0\t0\t0\t0",
        "Mean size of place types": "200\t800\t3300\t14.28",
        "Power of place size distribution": "0\t0\t0\t1.6",
        "Offset of place size distribution": "0\t0\t0\t1",
        "Maximum of place size distribution": "0\t0\t0\t5000",
        "Kernel type": 2,
        "Kernel scale": 4000,
        "Kernel Shape": 3,
    })
    open(os.path.join(d, "pre-params.txt"), "w").write(pre)
    open(os.path.join(d, "admin-params.txt"), "w").write(adm)
    open(os.path.join(d, "input-noint-params.txt"), "w").write(noint)
    density(os.path.join(d, "density.txt"), -3.5, 51.6, n_side, total, 12345)
    print("wrote", d)


if __name__ == "__main__":
    make("c1", 60, 200000, 200)
    # small fixture region: 0.25 x 0.25 degrees, 20k people, 120 days
    make("mini", 30, 20000, 120)
    sys.exit(0)
