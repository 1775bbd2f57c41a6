"""ctypes view of include/covidsim_hip.h.

The structs are generated from the header text at import time, so the C
header is the single definition of the ABI.  Only the constructs the header
uses are understood: `#define NAME value`, enums with explicit or implicit
values, and `typedef struct { ... } name;` bodies whose members are scalars,
pointers, and fixed-size (possibly 2-D) arrays of the integer/float types.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "covidsim_hip.h")

_TYPES = {
    "double": C.c_double, "float": C.c_float, "int": C.c_int,
    "int8_t": C.c_int8, "uint8_t": C.c_uint8, "int16_t": C.c_int16, "uint16_t": C.c_uint16,
    "int32_t": C.c_int32, "uint32_t": C.c_uint32, "int64_t": C.c_int64, "uint64_t": C.c_uint64,
}


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _parse(text):
    text = _strip_comments(text)
    consts = {}
    for m in re.finditer(r"#define\s+(CS_\w+)\s+\(?(-?\d+)\)?", text):
        consts[m.group(1)] = int(m.group(2))
    for m in re.finditer(r"enum\s*\{([^}]*)\}", text):
        val = 0
        for item in m.group(1).split(","):
            item = item.strip()
            if not item:
                continue
            if "=" in item:
                name, v = [x.strip() for x in item.split("=")]
                val = int(v)
            else:
                name = item
            consts[name] = val
            val += 1

    def ev(expr):
        expr = expr.strip()
        for k in sorted(consts, key=len, reverse=True):
            expr = re.sub(r"\b%s\b" % k, str(consts[k]), expr)
        return int(eval(expr, {"__builtins__": {}}))

    structs = {}
    for m in re.finditer(r"typedef\s+struct\s+(\w+)\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        body, name = m.group(2), m.group(3)
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            decl = decl.replace("const ", "")
            tm = re.match(r"(\w+)\s*(.*)", decl)
            base, rest = tm.group(1), tm.group(2)
            for part in re.split(r",(?![^\[]*\])", rest):
                part = part.strip()
                ptr = part.count("*")
                part = part.replace("*", "").strip()
                dims = [ev(d) for d in re.findall(r"\[([^\]]+)\]", part)]
                fname = re.sub(r"\[.*", "", part).strip()
                ct = _TYPES[base]
                if ptr:
                    ct = C.POINTER(ct)
                for d in reversed(dims):
                    ct = ct * d
                fields.append((fname, ct))
        structs[name] = fields
    return consts, structs


with open(HEADER) as _f:
    CONSTS, _STRUCTS = _parse(_f.read())


def _make(name):
    return type(name, (C.Structure,), {"_fields_": _STRUCTS[name]})


cs_params = _make("cs_params")
cs_population = _make("cs_population")
cs_state = _make("cs_state")
cs_counters = _make("cs_counters")
cs_work_counts = _make("cs_work_counts")
cs_transmission_coeffs = _make("cs_transmission_coeffs")
cs_tuning = _make("cs_tuning")
cs_intervention_state = _make("cs_intervention_state")
cs_popgen_params = _make("cs_popgen_params")
cs_popgen_sizes = _make("cs_popgen_sizes")
cs_popgen_out = _make("cs_popgen_out")

globals().update({k: v for k, v in CONSTS.items()})


def declare(lib, prefix):
    """Set argtypes/restype for the cs_* (or or_*) entry points of lib."""
    P, I, D, F = C.POINTER, C.c_int, C.c_double, C.c_float
    vp = C.c_void_p
    sig = {
        "create": None,
        "destroy": (None, [vp]),
        "last_error": (C.c_char_p, [vp]),
        "alloc_realisations": (I, [vp, I, P(C.c_uint64)]),
        "reset_realisations": (I, [vp, I]),
        "set_realisation_params": (I, [vp, P(D), P(C.c_int32)]),
        "set_realisation_triggers": (I, [vp, P(D)]),
        "load_state": (I, [vp, I, P(cs_state)]),
        "store_state": (I, [vp, I, P(cs_state)]),
        "infect_sweep": (I, [vp, D]),
        "incub_recovery_sweep": (I, [vp, D]),
        "update_probs": (I, [vp]),
        "run_steps": (I, [vp, D, I, I]),
        "set_step_graphs": (I, [vp, I]),
        "set_dest_tables": (I, [vp, I]),
        "get_dest_tables": (I, [vp, P(I)]),
        "set_launch_shape": (I, [vp, C.c_uint32, C.c_uint32]),
        "set_move_capacity": (I, [vp, C.c_int]),
        "set_dest_index": (I, [vp, C.c_int]),
        "autotune": (I, [vp, D, I, P(cs_tuning)]),
        "set_queue_capacity": (I, [vp, C.c_uint32, C.c_uint32]),
        "get_queue_capacity": (I, [vp, P(C.c_uint32), P(C.c_uint32)]),
        "get_counters": (I, [vp, I, P(cs_counters), I]),
        "get_intervention_state": (I, [vp, I, P(cs_intervention_state)]),
        "get_timeseries": (I, [vp, P(C.c_int32), I, P(I)]),
        "get_timeseries_age": (I, [vp, P(C.c_int32), I, P(I)]),
        "get_timeseries_adunit": (I, [vp, P(C.c_int32), I, P(I)]),
        "get_adunit_counters": (I, [vp, I, P(C.c_int64), I]),
        "get_work_counts": (I, [vp, P(cs_work_counts), I]),
        "kernel_num": (I, [vp, P(D), P(D), I]),
        "dist2_persons": (I, [vp, P(C.c_int32), P(C.c_int32), P(D), I]),
        "get_cell_tables": (I, [vp, I, P(F), P(F), P(C.c_int32), P(F), P(C.c_int32)]),
        "eval_hazards": (I, [vp, I, I, C.c_uint16, P(C.c_int32), P(C.c_int32), P(D), I]),
        "icdf_choose": (I, [vp, I, P(D), P(D), P(C.c_uint16), I]),
        "transmission_coeffs_from_draws": (I, [vp, P(D), P(F), P(C.c_uint16), P(C.c_int8), P(cs_transmission_coeffs)]),
        "set_profiling": (I, [vp, I]),
        "get_profile": (I, [vp, P(D), P(C.c_int64), I]),
        "sync": (I, [vp]),
    }
    for k, v in sig.items():
        if v is None:
            continue
        fn = getattr(lib, prefix + k, None)
        if fn is None:
            continue
        fn.restype, fn.argtypes = v
    return lib
