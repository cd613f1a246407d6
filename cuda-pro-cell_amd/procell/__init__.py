"""Python binding of libprocell (ctypes over the C ABI in include/procell.h).

The simulation itself runs in the gfx950 HIP kernels of libprocell.so; this
module only marshals arrays.  There is no fallback: if the library is not
built, or no gfx950 device is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
BUILD_DIR = os.path.join(PKG_ROOT, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libprocell.so")
CLI_PATH = os.path.join(BUILD_DIR, "procell")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "procell.h")

OK = 0
E_INVALID = -1
E_PROPORTIONS = -2
E_ZERO_MEAN = -3
E_NONFINITE = -4
E_PHI = -5
E_DEPTH = -6
E_TOO_MANY_CELLS = -7
E_CAPACITY = -8
E_HIP = -9
E_NO_DEVICE = -10
E_IO = -11
E_OUT_OF_MEMORY = -12
E_SCRATCH = -13

FLAG_TIMING = 1

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class Row(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double), ("count", ctypes.c_uint64)]


class CellType(ctypes.Structure):
    _fields_ = [("proportion", ctypes.c_double), ("mean", ctypes.c_double), ("sd", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("segments", ctypes.c_uint32), ("idle_lanes", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class Params(ctypes.Structure):
    _fields_ = [("rows", ctypes.POINTER(Row)), ("n_rows", ctypes.c_uint64),
                ("types", ctypes.POINTER(CellType)), ("n_types", ctypes.c_uint32),
                ("t_max", ctypes.c_double), ("phi", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("track_ratio", ctypes.c_int32), ("device", ctypes.c_int32),
                ("tuning", Tuning), ("stream", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("divisions", ctypes.c_uint64), ("recorded", ctypes.c_uint64),
                ("dropped", ctypes.c_uint64), ("lost", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("cells", ctypes.c_uint64),
                ("pool", ctypes.c_uint64), ("ms_tree", ctypes.c_double), ("ms_run", ctypes.c_double),
                ("workgroups", ctypes.c_uint32), ("block_size", ctypes.c_uint32),
                ("lds_bytes", ctypes.c_uint32), ("window_rows", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ResultStruct(ctypes.Structure):
    _fields_ = [("capacity", ctypes.c_uint64), ("value", _dp), ("count", _i64p),
                ("type_count", _i64p), ("n_out", ctypes.c_uint64), ("raw_tally", _i64p),
                ("raw_tally_len", ctypes.c_uint64), ("stats", Stats)]


class AdamStruct(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_double), ("beta1", ctypes.c_double), ("beta2", ctypes.c_double),
                ("eps", ctypes.c_double)]


class ProcellError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        msg = strerror(code) if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")
        self.code = code


_lib = None

EXPORTS = (
    "procell_simulate", "procell_simulate_noalloc", "procell_output_capacity", "procell_tally_shape", "procell_resolve_phi",
    "procell_strerror", "procell_abi_version", "procell_plan_create", "procell_plan_run",
    "procell_plan_result", "procell_plan_destroy", "procell_plan_stream", "procell_plan_kernel_times",
    "procell_read_histogram",
    "procell_read_types", "procell_write_histogram", "procell_free", "procell_debug_draws",
    "procell_batch_create", "procell_batch_run", "procell_batch_result", "procell_batch_autotune",
    "procell_plan_autotune",
    "procell_relaxed_create", "procell_relaxed_shape", "procell_relaxed_forward", "procell_relaxed_backward",
    "procell_relaxed_set_stride", "procell_relaxed_expansions", "procell_relaxed_set_cuts", "procell_relaxed_destroy",
    "procell_relaxed_jacobian", "procell_relaxed_create_batch", "procell_relaxed_set_shape",
    "procell_calib_create", "procell_calib_step", "procell_calib_theta", "procell_calib_loss",
    "procell_calib_stream", "procell_calib_destroy",
)


def lib():
    """Load libprocell.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with cuda-pro-cell_amd/build.py")
    # torch wheels carry their own HIP runtime under the same SONAME
    # (libamdhip64.so.7).  Two runtimes in one process cannot both open the
    # GPU: whichever initialises second finds no device.  Loading torch first
    # makes the dynamic loader bind libprocell to torch's copy, so that
    # tensors, streams and this library share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(Params)
    R = ctypes.POINTER(ResultStruct)
    L.procell_simulate.argtypes = [P, R]
    L.procell_simulate_noalloc.argtypes = [P, R]
    L.procell_output_capacity.argtypes = [P, _u64p]
    L.procell_tally_shape.argtypes = [P, _u64p, _u32p]
    L.procell_resolve_phi.argtypes = [P, _dp]
    L.procell_strerror.restype = ctypes.c_char_p
    L.procell_strerror.argtypes = [ctypes.c_int]
    L.procell_plan_create.argtypes = [P, ctypes.POINTER(ctypes.c_void_p)]
    L.procell_plan_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.procell_plan_result.argtypes = [ctypes.c_void_p, R]
    L.procell_plan_destroy.argtypes = [ctypes.c_void_p]
    L.procell_plan_stream.restype = ctypes.c_void_p
    L.procell_plan_stream.argtypes = [ctypes.c_void_p]
    L.procell_plan_kernel_times.argtypes = [ctypes.c_void_p, _dp, ctypes.c_uint32, _u32p]
    L.procell_read_histogram.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(Row)), _u64p]
    L.procell_read_types.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(CellType)), _u32p]
    L.procell_write_histogram.argtypes = [ctypes.c_char_p, R, ctypes.c_uint32, ctypes.c_int32]
    L.procell_free.argtypes = [ctypes.c_void_p]
    L.procell_debug_draws.argtypes = [ctypes.c_int32, ctypes.c_uint64, _u64p, _u64p, _u32p,
                                      ctypes.c_uint64, _u32p, _dp]
    if hasattr(L, "procell_batch_create"):  # (libraries built before the batch API lack it)
        L.procell_batch_create.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
        L.procell_batch_run.argtypes = [ctypes.c_void_p, _u64p]
        L.procell_batch_result.argtypes = [ctypes.c_void_p, ctypes.c_uint32, R]
        L.procell_batch_autotune.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint32, ctypes.POINTER(Tuning)]
        L.procell_plan_autotune.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint32, ctypes.POINTER(Tuning)]
    if hasattr(L, "procell_relaxed_create"):
        L.procell_relaxed_create.argtypes = [P, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]
        L.procell_relaxed_shape.argtypes = [ctypes.c_void_p, _u64p, _u32p, _u32p]
        L.procell_relaxed_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.procell_relaxed_backward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]
        L.procell_relaxed_expansions.argtypes = [ctypes.c_void_p, _u64p]
        L.procell_relaxed_set_cuts.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.procell_relaxed_set_stride.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
        L.procell_relaxed_destroy.argtypes = [ctypes.c_void_p]
    if hasattr(L, "procell_relaxed_jacobian"):
        L.procell_relaxed_jacobian.argtypes = [ctypes.c_void_p] * 5
        L.procell_relaxed_create_batch.argtypes = [P, ctypes.c_uint32, ctypes.c_double,
                                                   ctypes.POINTER(ctypes.c_void_p)]
        L.procell_relaxed_set_shape.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _u64p, _u64p, _u32p, _u32p]
    if hasattr(L, "procell_calib_create"):
        L.procell_calib_create.argtypes = [P, ctypes.c_uint32, ctypes.c_double, R, ctypes.POINTER(AdamStruct),
                                           ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
        L.procell_calib_step.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.procell_calib_theta.argtypes = [ctypes.c_void_p, _dp, ctypes.c_uint64]
        L.procell_calib_loss.argtypes = [ctypes.c_void_p, _dp, ctypes.c_uint32, _u64p, _u64p]
        L.procell_calib_stream.restype = ctypes.c_void_p
        L.procell_calib_stream.argtypes = [ctypes.c_void_p]
        L.procell_calib_destroy.argtypes = [ctypes.c_void_p]
    _lib = L
    return L


def strerror(code: int) -> str:
    return lib().procell_strerror(int(code)).decode()


def _check(rc: int, what: str = ""):
    if rc != OK:
        raise ProcellError(rc, what)


class Problem:
    """Inputs of one simulation, kept alive for the C structs that point at them."""

    def __init__(self, values, counts, proportion, mean, sd, t_max, phi=0.0, seed=0,
                 track_ratio=True, device=0, block_size=0, blocks_per_cu=0, segments=0, idle_lanes=0, flags=0):
        values = np.ascontiguousarray(values, dtype=np.float64)
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        self.rows = (Row * max(len(values), 1))()
        for i, (v, c) in enumerate(zip(values, counts)):
            self.rows[i].value = float(v)
            self.rows[i].count = int(c)
        n_types = len(proportion)
        self.types = (CellType * max(n_types, 1))()
        for j in range(n_types):
            self.types[j].proportion = float(proportion[j])
            self.types[j].mean = float(mean[j])
            self.types[j].sd = float(sd[j])
        self.n_types = n_types
        self.values, self.counts = values, counts
        self.proportion = np.array(proportion, dtype=np.float64)
        self.mean = np.array(mean, dtype=np.float64)
        self.sd = np.array(sd, dtype=np.float64)
        self.params = Params()
        self.params.rows = ctypes.cast(self.rows, ctypes.POINTER(Row))
        self.params.n_rows = len(values)
        self.params.types = ctypes.cast(self.types, ctypes.POINTER(CellType))
        self.params.n_types = n_types
        self.params.t_max = float(t_max)
        self.params.phi = float(phi)
        self.params.seed = int(seed) & (2**64 - 1)
        self.params.track_ratio = int(bool(track_ratio))
        self.params.device = int(device)
        self.params.tuning = Tuning(block_size, blocks_per_cu, segments, idle_lanes, flags)
        self.params.stream = None

    @classmethod
    def from_workload(cls, w, seed=0, **kw):
        kw.setdefault("track_ratio", w.track_ratio)
        return cls(w.values, w.counts, w.proportion, w.mean, w.sd, w.t_max, w.phi, seed, **kw)

    def output_capacity(self) -> int:
        cap = ctypes.c_uint64()
        _check(lib().procell_output_capacity(ctypes.byref(self.params), ctypes.byref(cap)), "output_capacity")
        return cap.value

    def tally_shape(self):
        rows = ctypes.c_uint64()
        kdim = ctypes.c_uint32()
        _check(lib().procell_tally_shape(ctypes.byref(self.params), ctypes.byref(rows), ctypes.byref(kdim)),
               "tally_shape")
        return rows.value, kdim.value, max(self.n_types, 1)

    def resolve_phi(self) -> float:
        phi = ctypes.c_double()
        _check(lib().procell_resolve_phi(ctypes.byref(self.params), ctypes.byref(phi)), "resolve_phi")
        return phi.value


@dataclass
class Result:
    value: np.ndarray        # float64 [n_out]
    count: np.ndarray        # int64 [n_out]
    type_count: np.ndarray   # int64 [n_out, n_types] (file order)
    raw_tally: np.ndarray    # int64 [nonzero_rows, kdim, n_types]
    stats: dict


class _ResultBuffers:
    def __init__(self, prob: Problem, want_raw=True):
        cap = prob.output_capacity()
        T = max(prob.n_types, 1)
        self.value = np.zeros(max(cap, 1), dtype=np.float64)
        self.count = np.zeros(max(cap, 1), dtype=np.int64)
        self.type_count = np.zeros((max(cap, 1), T), dtype=np.int64)
        shape = prob.tally_shape()
        self.raw = np.zeros(shape, dtype=np.int64) if want_raw else None
        self.r = ResultStruct()
        self.r.capacity = cap
        self.r.value = self.value.ctypes.data_as(_dp)
        self.r.count = self.count.ctypes.data_as(_i64p)
        self.r.type_count = self.type_count.ctypes.data_as(_i64p)
        if want_raw and self.raw.size:
            self.r.raw_tally = self.raw.ctypes.data_as(_i64p)
            self.r.raw_tally_len = self.raw.size
        self.T = prob.n_types

    def result(self) -> Result:
        n = self.r.n_out
        return Result(self.value[:n].copy(), self.count[:n].copy(),
                      self.type_count[:n, : max(self.T, 0)].copy(),
                      self.raw.copy() if self.raw is not None else None,
                      self.r.stats.as_dict())


def simulate(prob: Problem, want_raw: bool = True) -> Result:
    """procell_simulate(): one end-to-end run on the GPU."""
    buf = _ResultBuffers(prob, want_raw)
    _check(lib().procell_simulate(ctypes.byref(prob.params), ctypes.byref(buf.r)), "simulate")
    return buf.result()


def simulate_noalloc(prob: Problem, want_raw: bool = True) -> Result:
    """procell_simulate_noalloc(): procell_simulate with every device buffer in the
    library's static per-GPU scratch array (no device allocation)."""
    buf = _ResultBuffers(prob, want_raw)
    _check(lib().procell_simulate_noalloc(ctypes.byref(prob.params), ctypes.byref(buf.r)), "simulate_noalloc")
    return buf.result()


class Plan:
    """procell_plan_*: inputs resident on the device, runs enqueued on one stream."""

    def __init__(self, prob: Problem, stream=None):
        self.prob = prob
        if stream is not None:
            prob.params.stream = ctypes.c_void_p(int(stream))
        self.handle = ctypes.c_void_p()
        try:
            _check(lib().procell_plan_create(ctypes.byref(prob.params), ctypes.byref(self.handle)), "plan_create")
        finally:
            prob.params.stream = None
        self._buf = None

    def run(self, seed: int):
        _check(lib().procell_plan_run(self.handle, int(seed) & (2**64 - 1)), "plan_run")

    def result(self, want_raw: bool = True) -> Result:
        buf = _ResultBuffers(self.prob, want_raw)
        _check(lib().procell_plan_result(self.handle, ctypes.byref(buf.r)), "plan_result")
        return buf.result()

    def kernel_times(self, max_runs: int = 256):
        """Device ms of the simulation kernel for each run since the last call
        (needs flags=FLAG_TIMING)."""
        ms = np.zeros(max_runs, dtype=np.float64)
        n = ctypes.c_uint32()
        _check(lib().procell_plan_kernel_times(self.handle, ms.ctypes.data_as(_dp), max_runs, ctypes.byref(n)),
               "plan_kernel_times")
        return ms[: n.value].copy()

    def autotune(self, seeds, reps: int = 2) -> dict:
        """Times every candidate layout (resident workgroups per CU x segments
        per workgroup) and idle-lane threshold on this plan's inputs (seeds:
        one per set), keeps the fastest and returns its knobs."""
        s = np.ascontiguousarray(np.atleast_1d(np.asarray(seeds, dtype=np.uint64)))
        n_sets = len(getattr(self, "problems", [None]))
        if s.size != n_sets:
            raise ValueError(f"{s.size} seeds for {n_sets} sets")
        t = Tuning()
        _check(lib().procell_plan_autotune(self.handle, s.ctypes.data_as(_u64p), int(reps), ctypes.byref(t)),
               "plan_autotune")
        return {"blocks_per_cu": t.blocks_per_cu, "segments": t.segments, "idle_lanes": t.idle_lanes,
                "block_size": t.block_size}

    @property
    def stream(self) -> int:
        return lib().procell_plan_stream(self.handle) or 0

    def close(self):
        if self.handle:
            lib().procell_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch(Plan):
    """procell_batch_*: many parameter sets run by one launch of each kernel
    (a replayed hipGraph); set i's result equals Plan(problems[i]) run with
    seeds[i]."""

    def __init__(self, problems, stream=None):
        self.problems = list(problems)
        self.prob = self.problems[0] if self.problems else None
        arr = (Params * max(len(self.problems), 1))(*[p.params for p in self.problems])
        if stream is not None and self.problems:
            arr[0].stream = ctypes.c_void_p(int(stream))
        self.handle = ctypes.c_void_p()
        _check(lib().procell_batch_create(arr, len(self.problems), ctypes.byref(self.handle)), "batch_create")
        self._buf = None

    def _seeds(self, seeds):
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        if s.shape != (len(self.problems),):
            raise ValueError("one seed per parameter set")
        return s

    def run(self, seeds):
        s = self._seeds(seeds)
        _check(lib().procell_batch_run(self.handle, s.ctypes.data_as(_u64p)), "batch_run")

    def result(self, i: int = 0, want_raw: bool = True) -> Result:
        buf = _ResultBuffers(self.problems[i], want_raw)
        _check(lib().procell_batch_result(self.handle, int(i), ctypes.byref(buf.r)), "batch_result")
        return buf.result()

    def autotune(self, seeds, reps: int = 2) -> dict:
        """procell_plan_autotune on these sets (one seed each)."""
        return Plan.autotune(self, seeds, reps)


class Relaxed:
    """procell_relaxed_*: the relaxed estimator of one histogram (t_max, phi,
    seed and quiescent types fixed by `prob`) at temperature `beta`.

    forward / backward take device tensors (float64, contiguous, on the
    plan's device): theta [3, n_types] = mean, sd, proportion; h and g
    [n_rows, kdim, n_types]; grad [3, n_types].  They run on `stream`
    (torch's current stream by default), in order with the caller's work."""

    def __init__(self, prob: Problem, beta: float):
        self.prob = prob
        self.beta = float(beta)
        self.handle = ctypes.c_void_p()
        _check(lib().procell_relaxed_create(ctypes.byref(prob.params), self.beta, ctypes.byref(self.handle)),
               "relaxed_create")
        n, k, t = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().procell_relaxed_shape(self.handle, ctypes.byref(n), ctypes.byref(k), ctypes.byref(t)),
               "relaxed_shape")
        self.shape = (int(n.value), int(k.value), int(t.value))
        self.stride = (1, 0)

    @staticmethod
    def _ptr(x, shape, what):
        import torch

        if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous():
            raise ValueError(f"{what}: a contiguous float64 CUDA tensor is required")
        if tuple(x.shape) != tuple(shape):
            raise ValueError(f"{what}: shape {tuple(x.shape)}, expected {tuple(shape)}")
        return ctypes.c_void_p(x.data_ptr())

    @staticmethod
    def _stream(stream):
        if stream is None:
            import torch

            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        return ctypes.c_void_p(int(stream))

    def forward(self, theta, h, stream=None):
        T = self.shape[2]
        _check(lib().procell_relaxed_forward(self.handle, self._ptr(theta, (3, T), "theta"),
                                             self._ptr(h, self.shape, "h"), self._stream(stream)),
               "relaxed_forward")

    def backward(self, theta, g, grad, stream=None):
        T = self.shape[2]
        _check(lib().procell_relaxed_backward(self.handle, self._ptr(theta, (3, T), "theta"),
                                              self._ptr(g, self.shape, "g"), self._ptr(grad, (3, T), "grad"),
                                              self._stream(stream)),
               "relaxed_backward")

    def jacobian(self, theta, h, jac, stream=None):
        """h and jac [2, n_rows, kdim, n_types] = dh/dmean, dh/dsd of each
        entry's own type, in one pass (dh/dproportion_j = h / proportion_j)."""
        T = self.shape[2]
        _check(lib().procell_relaxed_jacobian(self.handle, self._ptr(theta, (3, T), "theta"),
                                              self._ptr(h, self.shape, "h"),
                                              self._ptr(jac, (2,) + tuple(self.shape), "jac"),
                                              self._stream(stream)),
               "relaxed_jacobian")

    def set_stride(self, stride: int, phase: int = 0):
        """Later passes walk only the initial cells c with (c + phase) a
        multiple of `stride` (a power of two); 1 walks every cell."""
        _check(lib().procell_relaxed_set_stride(self.handle, int(stride), int(phase)), "relaxed_set_stride")
        self.stride = (int(stride), int(phase) % int(stride))

    def set_cuts(self, cuts, stream=None):
        """Enqueues a copy of the last call's walks cut short, per set, into
        `cuts` (an int64 CUDA tensor with one entry per set); no wait."""
        import torch

        n = len(getattr(self, "shapes", [None]))
        if cuts.dtype != torch.int64 or not cuts.is_cuda or not cuts.is_contiguous() or cuts.numel() != n:
            raise ValueError(f"cuts: a contiguous int64 CUDA tensor of {n} entries is required")
        _check(lib().procell_relaxed_set_cuts(self.handle, ctypes.c_void_p(cuts.data_ptr()), self._stream(stream)),
               "relaxed_set_cuts")

    def expansions(self) -> int:
        """Expansions of the last call (roots drawn plus nodes expanded); waits."""
        n = ctypes.c_uint64()
        _check(lib().procell_relaxed_expansions(self.handle, ctypes.byref(n)), "relaxed_expansions")
        return int(n.value)

    def close(self):
        if self.handle:
            lib().procell_relaxed_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RelaxedBatch(Relaxed):
    """procell_relaxed_create_batch: one relaxed plan over many problems.
    theta, h, g, grad and jac are flat device tensors, the sets' arrays
    concatenated in order: theta [3 * types_total] (set i's [3, n_types] at
    theta_base[i]), h and g [h_total] (set i's [n_rows, kdim, n_types] at
    h_base[i]), grad [3 * types_total], jac [2, h_total]."""

    def __init__(self, problems, beta: float):
        self.problems = list(problems)
        self.beta = float(beta)
        n = len(self.problems)
        arr = (Params * max(n, 1))(*[p.params for p in self.problems])
        self.handle = ctypes.c_void_p()
        _check(lib().procell_relaxed_create_batch(arr, n, self.beta, ctypes.byref(self.handle)), "relaxed_create_batch")
        self.shapes, self.h_base, self.theta_base = [], [], []
        tb = 0
        for i in range(n):
            hb, r, k, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
            _check(lib().procell_relaxed_set_shape(self.handle, i, ctypes.byref(hb), ctypes.byref(r),
                                                   ctypes.byref(k), ctypes.byref(t)), "relaxed_set_shape")
            self.shapes.append((int(r.value), int(k.value), int(t.value)))
            self.h_base.append(int(hb.value))
            self.theta_base.append(tb)
            tb += 3 * int(t.value)
        self.types_total = tb // 3
        last = self.shapes[-1]
        self.h_total = self.h_base[-1] + last[0] * last[1] * last[2]
        self.shape = (self.h_total,)
        self.stride = (1, 0)

    def forward(self, theta, h, stream=None):
        _check(lib().procell_relaxed_forward(self.handle, self._ptr(theta, (3 * self.types_total,), "theta"),
                                             self._ptr(h, (self.h_total,), "h"), self._stream(stream)),
               "relaxed_forward")

    def backward(self, theta, g, grad, stream=None):
        n3 = 3 * self.types_total
        _check(lib().procell_relaxed_backward(self.handle, self._ptr(theta, (n3,), "theta"),
                                              self._ptr(g, (self.h_total,), "g"), self._ptr(grad, (n3,), "grad"),
                                              self._stream(stream)),
               "relaxed_backward")

    def jacobian(self, theta, h, jac, stream=None):
        _check(lib().procell_relaxed_jacobian(self.handle, self._ptr(theta, (3 * self.types_total,), "theta"),
                                              self._ptr(h, (self.h_total,), "h"),
                                              self._ptr(jac, (2, self.h_total), "jac"), self._stream(stream)),
               "relaxed_jacobian")


class Calibration:
    """procell_calib_*: batched calibration of many parameter sets, each to
    its own target histogram, with the relaxed estimator at temperature
    `beta` and an on-device Adam step replayed from one hipGraph.

    problems: their histograms, t_max, phi, relaxed-stream seeds and the
    starting parameters; targets: per set (value, count) arrays of a target
    histogram in the plain method's output form (e.g. a Result); cell_stride:
    a power of two k walks the initial cells of phase (step mod k) only."""

    def __init__(self, problems, beta, targets, lr=0.02, beta1=0.9, beta2=0.999, eps=1e-8, stream=None,
                 cell_stride=1):
        self.problems = list(problems)
        n = len(self.problems)
        if len(targets) != n:
            raise ValueError("one target per parameter set")
        arr = (Params * max(n, 1))(*[p.params for p in self.problems])
        if stream is not None and n:
            arr[0].stream = ctypes.c_void_p(int(stream))
        self._keep = []
        tr = (ResultStruct * max(n, 1))()
        for i, tg in enumerate(targets):
            value, count = (tg.value, tg.count) if hasattr(tg, "value") else tg
            v = np.ascontiguousarray(value, dtype=np.float64)
            c = np.ascontiguousarray(count, dtype=np.int64)
            self._keep += [v, c]
            tr[i].capacity = len(v)
            tr[i].n_out = len(v)
            tr[i].value = v.ctypes.data_as(_dp)
            tr[i].count = c.ctypes.data_as(_i64p)
        adam = AdamStruct(lr, beta1, beta2, eps)
        self.handle = ctypes.c_void_p()
        _check(lib().procell_calib_create(arr, n, float(beta), tr, ctypes.byref(adam), int(cell_stride),
                                          ctypes.byref(self.handle)), "calib_create")
        self.n_types = [max(1, p.n_types) for p in self.problems]

    def step(self, n: int = 1):
        _check(lib().procell_calib_step(self.handle, int(n)), "calib_step")

    def theta(self):
        """Current parameters: one [3, n_types] array (mean, sd, proportion) per set."""
        out = np.zeros(3 * sum(self.n_types), dtype=np.float64)
        _check(lib().procell_calib_theta(self.handle, out.ctypes.data_as(_dp), len(out)), "calib_theta")
        res, at = [], 0
        for T in self.n_types:
            res.append(out[at: at + 3 * T].reshape(3, T).copy())
            at += 3 * T
        return res

    def loss(self):
        """(loss of each set at the last step's parameters, steps taken)."""
        out = np.zeros(max(len(self.problems), 1), dtype=np.float64)
        steps, exp = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().procell_calib_loss(self.handle, out.ctypes.data_as(_dp), len(out), ctypes.byref(steps),
                                        ctypes.byref(exp)), "calib_loss")
        self.last_expansions = int(exp.value)
        return out[: len(self.problems)], int(steps.value)

    @property
    def stream(self) -> int:
        return lib().procell_calib_stream(self.handle) or 0

    def close(self):
        if self.handle:
            lib().procell_calib_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def debug_draws(seed, cells, nodes, attempts, device=0):
    """Philox blocks and Box-Muller pairs computed by a device kernel."""
    cells = np.ascontiguousarray(cells, dtype=np.uint64)
    nodes = np.ascontiguousarray(nodes, dtype=np.uint64)
    attempts = np.ascontiguousarray(attempts, dtype=np.uint32)
    n = len(cells)
    words = np.zeros((n, 4), dtype=np.uint32)
    z = np.zeros((n, 2), dtype=np.float64)
    _check(lib().procell_debug_draws(device, int(seed), cells.ctypes.data_as(_u64p), nodes.ctypes.data_as(_u64p),
                                     attempts.ctypes.data_as(_u32p), n, words.ctypes.data_as(_u32p),
                                     z.ctypes.data_as(_dp)), "debug_draws")
    return words, z


def read_histogram(path: str):
    rows = ctypes.POINTER(Row)()
    n = ctypes.c_uint64()
    _check(lib().procell_read_histogram(path.encode(), ctypes.byref(rows), ctypes.byref(n)), "read_histogram")
    try:
        values = np.array([rows[i].value for i in range(n.value)], dtype=np.float64)
        counts = np.array([rows[i].count for i in range(n.value)], dtype=np.uint64)
    finally:
        lib().procell_free(rows)
    return values, counts


def read_types(path: str):
    types = ctypes.POINTER(CellType)()
    n = ctypes.c_uint32()
    _check(lib().procell_read_types(path.encode(), ctypes.byref(types), ctypes.byref(n)), "read_types")
    try:
        out = np.array([[types[i].proportion, types[i].mean, types[i].sd] for i in range(n.value)],
                       dtype=np.float64).reshape(-1, 3)
    finally:
        lib().procell_free(types)
    return out[:, 0].copy(), out[:, 1].copy(), out[:, 2].copy()


def write_histogram(path, value, count, type_count=None, track_ratio=False):
    """procell_write_histogram() on host arrays (no device needed)."""
    value = np.ascontiguousarray(value, dtype=np.float64)
    count = np.ascontiguousarray(count, dtype=np.int64)
    r = ResultStruct()
    r.n_out = r.capacity = len(value)
    r.value = value.ctypes.data_as(_dp)
    r.count = count.ctypes.data_as(_i64p)
    T = 0
    if type_count is not None:
        type_count = np.ascontiguousarray(type_count, dtype=np.int64)
        T = type_count.shape[1] if type_count.ndim == 2 else 0
        r.type_count = type_count.ctypes.data_as(_i64p)
    _check(lib().procell_write_histogram(path.encode() if path else None, ctypes.byref(r), T,
                                         int(bool(track_ratio))), "write_histogram")
