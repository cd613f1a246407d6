"""ctypes binding of the C oracle (test infrastructure only).

Loads oracle/_build/libprocell_oracle.so (built by oracle/build.py) and
exposes the primitive generators and the full simulation as numpy-returning
functions.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PROCELL_ORACLE_LIB selects another build of the same source (the sanitizer
# build of oracle/build.py).
LIB_PATH = os.environ.get("PROCELL_ORACLE_LIB") or os.path.join(_HERE, "_build", "libprocell_oracle.so")

STAT_NAMES = ("divisions", "recorded", "dropped", "lost", "rejected", "cells")

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class _Problem(ctypes.Structure):
    _fields_ = [
        ("values", _dp), ("counts", _u64p), ("n_rows", ctypes.c_uint64),
        ("proportion", _dp), ("mean", _dp), ("sd", _dp), ("n_types", ctypes.c_uint32),
        ("t_max", ctypes.c_double), ("phi", ctypes.c_double), ("seed", ctypes.c_uint64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.oracle_block.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, _u32p]
        L.oracle_u01_open_closed.restype = ctypes.c_double
        L.oracle_u01_open_closed.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_u01_closed_open.restype = ctypes.c_double
        L.oracle_u01_closed_open.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_log.restype = ctypes.c_double
        L.oracle_log.argtypes = [ctypes.c_double]
        L.oracle_sqrt.restype = ctypes.c_double
        L.oracle_sqrt.argtypes = [ctypes.c_double]
        L.oracle_sincos_turns.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _dp, _dp]
        L.oracle_normal_pair.argtypes = [_u32p, _dp, _dp]
        L.oracle_resolve_phi.restype = ctypes.c_double
        L.oracle_resolve_phi.argtypes = [ctypes.POINTER(_Problem)]
        L.oracle_max_generation.restype = ctypes.c_int
        L.oracle_max_generation.argtypes = [ctypes.POINTER(_Problem)]
        L.oracle_simulate.restype = ctypes.c_int
        L.oracle_simulate.argtypes = [ctypes.POINTER(_Problem), _i64p, _i64p, _u64p, ctypes.c_int]
        L.oracle_relaxed.restype = ctypes.c_int
        L.oracle_relaxed.argtypes = [ctypes.POINTER(_Problem), ctypes.c_double, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_run_files.restype = ctypes.c_int
        L.oracle_run_files.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_int]
        _lib = L
    return _lib


def philox(ctr, key):
    out = (ctypes.c_uint32 * 4)()
    lib().oracle_philox4x32_10((ctypes.c_uint32 * 4)(*ctr), (ctypes.c_uint32 * 2)(*key), out)
    return list(out)


def block(seed, cell, node, attempt):
    out = (ctypes.c_uint32 * 4)()
    lib().oracle_block(seed, cell, node, attempt, out)
    return list(out)


def u01_open_closed(hi, lo):
    return lib().oracle_u01_open_closed(hi, lo)


def u01_closed_open(hi, lo):
    return lib().oracle_u01_closed_open(hi, lo)


def log(x):
    return lib().oracle_log(x)


def sqrt(x):
    return lib().oracle_sqrt(x)


def sincos_turns(hi, lo):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().oracle_sincos_turns(hi, lo, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def normal_pair(words):
    z0, z1 = ctypes.c_double(), ctypes.c_double()
    lib().oracle_normal_pair((ctypes.c_uint32 * 4)(*words), ctypes.byref(z0), ctypes.byref(z1))
    return z0.value, z1.value


class Problem:
    """Holds numpy buffers alive for the duration of a call."""

    def __init__(self, values, counts, proportion, mean, sd, t_max, phi, seed):
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.counts = np.ascontiguousarray(counts, dtype=np.uint64)
        self.proportion = np.ascontiguousarray(proportion, dtype=np.float64)
        self.mean = np.ascontiguousarray(mean, dtype=np.float64)
        self.sd = np.ascontiguousarray(sd, dtype=np.float64)
        self.c = _Problem(
            self.values.ctypes.data_as(_dp), self.counts.ctypes.data_as(_u64p), len(self.values),
            self.proportion.ctypes.data_as(_dp), self.mean.ctypes.data_as(_dp), self.sd.ctypes.data_as(_dp),
            len(self.proportion), float(t_max), float(phi), int(seed) & (2**64 - 1))

    @classmethod
    def from_workload(cls, w, seed):
        return cls(w.values, w.counts, w.proportion, w.mean, w.sd, w.t_max, w.phi, seed)


def resolve_phi(p: Problem) -> float:
    return lib().oracle_resolve_phi(ctypes.byref(p.c))


def max_generation(p: Problem) -> int:
    return lib().oracle_max_generation(ctypes.byref(p.c))


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle error {code}")
        self.code = code


def simulate(p: Problem, n_threads: int = 0, want_drops: bool = False):
    """Returns (tally[n_nonzero_rows, kdim, n_types] int64, drops or None,
    stats dict)."""
    kmax = max_generation(p)
    if kmax < 0:
        raise OracleError(kmax)
    n_nz = int(np.count_nonzero(p.counts))
    T = len(p.proportion)
    tally = np.zeros((n_nz, kmax + 1, max(T, 1)), dtype=np.int64)
    drops = np.zeros((n_nz, kmax + 1), dtype=np.int64) if want_drops else None
    stats = np.zeros(len(STAT_NAMES), dtype=np.uint64)
    rc = lib().oracle_simulate(ctypes.byref(p.c), tally.ctypes.data_as(_i64p),
                               drops.ctypes.data_as(_i64p) if want_drops else None,
                               stats.ctypes.data_as(_u64p), int(n_threads))
    if rc != 0:
        raise OracleError(rc)
    return tally, drops, {k: int(v) for k, v in zip(STAT_NAMES, stats)}


def relaxed(p: Problem, beta: float, g=None, n_threads: int = 0):
    """Relaxed estimator: returns (h[n_nonzero_rows, kdim, n_types] float64,
    grad[3, n_types] = dL/d(mean, sd, proportion) for g = dL/dh, or None)."""
    kmax = max_generation(p)
    if kmax < 0:
        raise OracleError(kmax)
    n_nz = int(np.count_nonzero(p.counts))
    T = len(p.proportion)
    h = np.zeros((n_nz, kmax + 1, max(T, 1)), dtype=np.float64)
    grad = np.zeros((3, max(T, 1)), dtype=np.float64)
    gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
    if gg is not None and gg.shape != h.shape:
        raise ValueError("g must have the shape of h")
    rc = lib().oracle_relaxed(ctypes.byref(p.c), float(beta), gg.ctypes.data_as(_dp) if gg is not None else None,
                              h.ctypes.data_as(_dp), grad.ctypes.data_as(_dp), int(n_threads))
    if rc != 0:
        raise OracleError(rc)
    return h, (grad[:, :T] if g is not None else None)


def run_files(histogram, types, t_max, phi, track_ratio, seed, output, n_threads=0):
    return lib().oracle_run_files(histogram.encode(), types.encode(), float(t_max), float(phi),
                                  int(bool(track_ratio)), int(seed), output.encode(), int(n_threads))
