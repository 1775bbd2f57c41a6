"""ctypes wrapper of the CPU oracle (oracle/_build/libcs_oracle.so).

Test infrastructure: the oracle is the checker for the GPU library and is
never used by the package itself.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "libcs_oracle.so")


def load():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    import importlib
    abi = importlib.import_module("covid-sim_amd.abi")
    lib = C.CDLL(LIB)
    abi.declare(lib, "or_")
    P = C.POINTER
    lib.or_create.restype = C.c_int
    lib.or_create.argtypes = [P(C.c_void_p), P(abi.cs_params), P(C.c_double), P(C.c_double), P(abi.cs_population)]
    lib.or_select_infector.argtypes = [P(C.c_float), C.c_int, P(C.c_double), P(C.c_int32), C.c_int]
    lib.or_select_dest_cell.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), P(C.c_int32), C.c_int]
    lib.or_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), C.c_int]
    lib.or_philox.restype = None
    lib.or_sample.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                              C.c_int, P(C.c_double)]
    lib.or_sample.restype = C.c_int
    lib.or_log.argtypes = [C.c_double]
    lib.or_log.restype = C.c_double
    lib.or_exp.argtypes = [C.c_double]
    lib.or_exp.restype = C.c_double
    return lib, abi


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Oracle:
    def __init__(self, params, pop, kl=None, kh=None):
        self.lib, self.abi = load()
        self.params, self.pop = params, pop
        self.h = C.c_void_p()
        self._kl, self._kh = kl, kh
        rc = self.lib.or_create(C.byref(self.h), C.byref(params),
                                dptr(kl) if kl is not None else None, dptr(kh) if kh is not None else None,
                                C.byref(pop.view()))
        self.check(rc)

    def check(self, rc):
        if rc:
            raise RuntimeError("oracle error %d: %s" % (rc, self.lib.or_last_error(self.h)))

    def __del__(self):
        try:
            self.lib.or_destroy(self.h)
        except Exception:
            pass

    def call(self, name, *args):
        self.check(getattr(self.lib, "or_" + name)(self.h, *args))
