"""MI355X-native restarted-PDHG LP engine (Niryukti-compatible surface).

The compute path is the native library ``niryukti_amd/lib/libniryukti.so``:
hand-written gfx950 HIP kernels behind C-ABIs declared in ``include/``.  This
module is a thin ctypes binding over those C-ABIs; it never computes anything
itself and raises if the native library is missing.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = [
    "load_library", "LP", "Model", "solve_lp", "Instance", "Backend", "Batch", "NativeSession", "Server",
    "device_count", "device_name", "ABI_VERSION", "CLI_PATH",
]

_PKG = Path(__file__).resolve().parent
LIBRARY_PATH = Path(os.environ.get("NIRYUKTI_LIBRARY", _PKG / "lib" / "libniryukti.so"))
ABI_VERSION = 1
CLI_PATH = _PKG / "bin" / "niryukti"

_f64p = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class _LP(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("ptr", _i64p), ("index", _i64p), ("value", _f64p),
                ("c", _f64p), ("lb", _f64p), ("ub", _f64p), ("rl", _f64p), ("ru", _f64p),
                ("offset", C.c_double), ("sense", C.c_double)]


class _Options(C.Structure):
    _fields_ = [("device", C.c_char_p), ("method", C.c_char_p), ("scaling", C.c_char_p),
                ("primal_weight", C.c_char_p), ("matrix_precision", C.c_char_p), ("gpu_indices", C.c_char_p),
                ("tol", C.c_double), ("time_limit", C.c_double), ("iteration_limit", C.c_int64),
                ("check_every", C.c_int), ("scaling_passes", C.c_int), ("threads", C.c_int),
                ("presolve", C.c_int), ("restart", C.c_int), ("adaptive", C.c_int), ("graphs", C.c_int),
                ("verbose", C.c_int), ("device_checks", C.c_int), ("device_prepare", C.c_int),
                ("power_iterations", C.c_int), ("polishing", C.c_int), ("initial_x", _f64p),
                ("initial_y", _f64p), ("checkpoint_path", C.c_char_p), ("resume_path", C.c_char_p),
                ("checkpoint_nodes", C.c_int64)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_char * 32), ("message", C.c_char * 256), ("objective", C.c_double),
                ("primal_residual", C.c_double), ("dual_residual", C.c_double), ("relative_gap", C.c_double),
                ("kkt_error", C.c_double), ("complementarity", C.c_double), ("dual_bound", C.c_double),
                ("primal_absolute", C.c_double), ("dual_absolute", C.c_double), ("finite", C.c_int),
                ("iterations", C.c_int64), ("restarts", C.c_int64), ("rejected_steps", C.c_int64),
                ("weight_updates", C.c_int64), ("host_candidate_checks", C.c_int64),
                ("device_candidate_checks", C.c_int64),
                ("removed_rows", C.c_int64), ("removed_columns", C.c_int64), ("seconds", C.c_double),
                ("preprocess_seconds", C.c_double), ("transfer_seconds", C.c_double),
                ("iteration_seconds", C.c_double), ("verification_seconds", C.c_double),
                ("index_bits", C.c_int), ("device_prepare", C.c_int), ("operator_norm_estimate", C.c_double),
                ("polishing_iterations", C.c_int64), ("polishing_attempts", C.c_int64),
                ("initial_step", C.c_double), ("initial_weight", C.c_double), ("first_restart_weight", C.c_double),
                ("first_restart_iteration", C.c_int64), ("fp32_iterations", C.c_int64), ("x", _f64p), ("y", _f64p)]


def _struct_dtype(cls) -> np.dtype:
    """numpy view of a ctypes structure of scalars, char arrays and pointers."""
    kinds = {C.c_double: np.float64, C.c_int: np.int32, C.c_int64: np.int64}
    names, formats, offsets = [], [], []
    for name, ctype in cls._fields_:
        if ctype in kinds:
            fmt = kinds[ctype]
        elif issubclass(ctype, C.Array) and ctype._type_ is C.c_char:
            fmt = f"S{ctype._length_}"
        else:  # pointers
            fmt = np.uint64
        names.append(name)
        formats.append(fmt)
        offsets.append(getattr(cls, name).offset)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": C.sizeof(cls)})


_RESULT_DTYPE = _struct_dtype(_Result)


class _Accuracy(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("objective", "primal", "dual", "gap", "kkt", "complementarity",
                                          "lower_bound", "primal_absolute", "dual_absolute")] + [("finite", C.c_int)]


class _PdhgLP(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("ptr", _i64p), ("index", _i64p), ("value", _f64p),
                ("c", _f64p), ("lb", _f64p), ("ub", _f64p), ("rl", _f64p), ("ru", _f64p)]


class _PdhgConfig(C.Structure):
    _fields_ = [("adaptive", C.c_int), ("halpern", C.c_int), ("reflection", C.c_double),
                ("residual_restarts", C.c_int), ("graphs", C.c_int), ("index_bits", C.c_int),
                ("fp32_values", C.c_int), ("fp32_arithmetic", C.c_int)]


class _PdhgOriginal(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("ptr", _i64p), ("index", _i64p), ("value", _f64p),
                ("c", _f64p), ("lb", _f64p), ("ub", _f64p), ("rl", _f64p), ("ru", _f64p), ("offset", C.c_double),
                ("cols_map", _i64p), ("lower_rows", _i64p), ("upper_rows", _i64p), ("fixed", _f64p),
                ("column_scale", _f64p), ("row_scale", _f64p), ("objective_scale", C.c_double)]


class _PdhgAccuracy(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("objective", "primal", "dual", "gap", "kkt")] + \
               [("finite", C.c_int), ("farkas_possible", C.c_int)]


class _PdhgStatus(C.Structure):
    _fields_ = [("accepted_total", C.c_int64), ("rejected_total", C.c_int64), ("restart_requested", C.c_int),
                ("index_bits", C.c_int), ("factor", C.c_double), ("fp32_phase", C.c_int), ("pad", C.c_int)]


class _Scaled(C.Structure):
    _fields_ = [("value", _f64p), ("c", _f64p), ("lb", _f64p), ("ub", _f64p), ("rl", _f64p), ("ru", _f64p),
                ("row_scale", _f64p), ("column_scale", _f64p), ("objective_scale", C.c_double),
                ("t_ptr", _i64p), ("t_index", _i64p), ("t_value", _f64p), ("device_seconds", C.c_double),
                ("seconds", C.c_double)]


class _BatchStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "queued", "active", "finished", "arena_bytes",
                                         "arena_used_bytes", "launches", "host_certifications",
                                         "host_farkas_checks")] + \
        [(k, C.c_double) for k in ("device_seconds", "advance_seconds", "check_seconds", "advance_bytes")] + \
        [("attempts", C.c_int64), ("class_seconds", C.c_double * 5), ("class_bytes", C.c_double * 5),
         ("class_launches", C.c_int64 * 5), ("handoffs", C.c_int64), ("handoffs_running", C.c_int64),
         ("class_threads", C.c_int32 * 3), ("class_rows", C.c_int32 * 3), ("class_tuned", C.c_int32 * 3),
         ("pad", C.c_int32), ("shape_seconds", C.c_double * 4), ("shape_bytes", C.c_double * 4),
         ("shape_launches", C.c_int64 * 4), ("certify_wait_seconds", C.c_double), ("settle_seconds", C.c_double),
         ("admit_seconds", C.c_double)]


_lib = None


def load_library() -> C.CDLL:
    """Loads the native engine; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.exists():
        raise RuntimeError(f"native engine library not found at {LIBRARY_PATH}; build it with "
                           "`python -m niryukti_amd.build` (the engine has no non-native fallback)")
    lib = C.CDLL(str(LIBRARY_PATH))
    lib.niryukti_default_options.argtypes = [C.POINTER(_Options)]
    lib.niryukti_solve_lp.argtypes = [C.POINTER(_LP), C.POINTER(_Options), C.POINTER(_Result)]
    lib.niryukti_solve_lp.restype = C.c_int
    lib.niryukti_engine_last_error.restype = C.c_char_p
    lib.niryukti_instance_create.argtypes = [C.c_int, C.c_uint64, C.c_int64, C.c_int64]
    lib.niryukti_instance_create.restype = C.c_void_p
    lib.niryukti_instance_destroy.argtypes = [C.c_void_p]
    lib.niryukti_instance_view.argtypes = [C.c_void_p, C.POINTER(_LP), C.POINTER(C.c_double),
                                           C.POINTER(_f64p), C.POINTER(_f64p)]
    lib.npdhg_last_error.restype = C.c_char_p
    lib.npdhg_device_count.restype = C.c_int
    lib.npdhg_device_name.argtypes = [C.c_int, C.c_char_p, C.c_int]
    lib.npdhg_create.argtypes = [C.POINTER(_PdhgLP), _f64p, _f64p, C.POINTER(_PdhgConfig), C.POINTER(C.c_void_p)]
    lib.npdhg_destroy.argtypes = [C.c_void_p]
    lib.npdhg_advance.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int)]
    lib.npdhg_status_get.argtypes = [C.c_void_p, C.POINTER(_PdhgStatus)]
    lib.npdhg_candidates.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
    lib.npdhg_reset.argtypes = [C.c_void_p, _f64p, _f64p]
    lib.npdhg_reset_candidate.argtypes = [C.c_void_p, C.c_int]
    lib.npdhg_set_fp32_phase.argtypes = [C.c_void_p, C.c_int]
    lib.npdhg_spmv.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p]
    lib.niryukti_prepare_lp.argtypes = [C.POINTER(_LP), C.POINTER(_Options), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), _i64p, _i64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                        _f64p, _f64p, _i64p, _i64p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.niryukti_prepare_lp.restype = C.c_int
    lib.niryukti_verify_lp.argtypes = [C.POINTER(_LP), _f64p, _f64p, C.c_int, C.POINTER(_Accuracy)]
    lib.niryukti_verify_lp.restype = C.c_int
    lib.niryukti_power_norm.argtypes = [C.POINTER(_LP), C.c_int]
    lib.niryukti_power_norm.restype = C.c_double
    lib.npdhg_scale_lp.argtypes = [C.POINTER(_PdhgLP), C.c_int, C.POINTER(_Scaled)]
    lib.npdhg_scale_lp.restype = C.c_int
    lib.npdhg_scale_last_error.restype = C.c_char_p
    lib.npdhg_scale_lp_noalloc.argtypes = [C.POINTER(_PdhgLP), C.c_int, C.POINTER(_Scaled), C.c_void_p]
    lib.npdhg_scale_lp_noalloc.restype = C.c_int
    lib.niryukti_batch_create.argtypes = [C.POINTER(_Options), C.c_int64, C.c_int64, C.POINTER(C.c_void_p)]
    lib.niryukti_batch_destroy.argtypes = [C.c_void_p]
    lib.niryukti_batch_submit.argtypes = [C.c_void_p, C.POINTER(_LP), _f64p, _f64p, C.POINTER(C.c_int64)]
    lib.niryukti_batch_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
    lib.niryukti_batch_run.argtypes = [C.c_void_p, C.c_int]
    lib.niryukti_batch_admit.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    lib.niryukti_batch_result.argtypes = [C.c_void_p, C.c_int64, C.POINTER(_Result)]
    lib.niryukti_batch_stats_get.argtypes = [C.c_void_p, C.POINTER(_BatchStats)]
    lib.niryukti_batch_collect.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(_Result), C.c_int64, C.c_int,
                                           C.POINTER(C.c_int64)]
    lib.niryukti_batch_last_error.restype = C.c_char_p
    lib.niryukti_batch_finished.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]
    lib.niryukti_batch_release.argtypes = [C.c_void_p, C.c_int64]
    lib.niryukti_batch_set_handoff.argtypes = [C.c_void_p, C.c_int64, C.c_int]
    lib.niryukti_model_load.argtypes = [C.c_char_p, C.c_char_p]
    lib.niryukti_model_load.restype = C.c_void_p
    lib.niryukti_model_destroy.argtypes = [C.c_void_p]
    lib.niryukti_model_view.argtypes = [C.c_void_p, C.POINTER(_LP), C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
    lib.niryukti_model_write.argtypes = [C.c_void_p, C.c_char_p]
    lib.niryukti_model_write.restype = C.c_int
    lib.niryukti_result_json.argtypes = [C.c_void_p, C.POINTER(_Result), C.c_char_p]
    lib.niryukti_result_json.restype = C.c_char_p
    _lib = lib
    return lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray, kind=_f64p):
    return a.ctypes.data_as(kind)


@dataclass
class LP:
    """An LP in the reference's canonical form: min c^T x + offset, rl <= A x <= ru, lb <= x <= ub."""

    ptr: np.ndarray
    index: np.ndarray
    value: np.ndarray
    c: np.ndarray
    lb: np.ndarray
    ub: np.ndarray
    rl: np.ndarray
    ru: np.ndarray
    offset: float = 0.0
    sense: float = 1.0

    def __post_init__(self):
        self.ptr = np.ascontiguousarray(self.ptr, dtype=np.int64).reshape(-1)
        self.index = np.ascontiguousarray(self.index, dtype=np.int64).reshape(-1)
        for name in ("value", "c", "lb", "ub", "rl", "ru"):
            setattr(self, name, _f64(getattr(self, name)).reshape(-1))
        # The native side reads exactly these extents from the raw pointers.
        m, n = len(self.rl), len(self.c)
        if len(self.ptr) != m + 1:
            raise ValueError(f"ptr has {len(self.ptr)} entries; rows + 1 = {m + 1} expected")
        if self.ptr[0] != 0 or np.any(np.diff(self.ptr) < 0):
            raise ValueError("ptr must start at 0 and be nondecreasing")
        nnz = int(self.ptr[-1])
        if len(self.index) != nnz or len(self.value) != nnz:
            raise ValueError(f"index and value must have ptr[-1] = {nnz} entries "
                             f"(got {len(self.index)} and {len(self.value)})")
        if len(self.lb) != n or len(self.ub) != n:
            raise ValueError(f"lb and ub must have cols = {n} entries (got {len(self.lb)} and {len(self.ub)})")
        if len(self.ru) != m:
            raise ValueError(f"ru must have rows = {m} entries (got {len(self.ru)})")

    @property
    def rows(self) -> int:
        return len(self.rl)

    @property
    def cols(self) -> int:
        return len(self.c)

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def _struct(self) -> _LP:
        return _LP(self.rows, self.cols, _ptr(self.ptr, _i64p), _ptr(self.index, _i64p), _ptr(self.value),
                   _ptr(self.c), _ptr(self.lb), _ptr(self.ub), _ptr(self.rl), _ptr(self.ru),
                   float(self.offset), float(self.sense))

    @classmethod
    def from_dense(cls, A, c, lb, ub, rl, ru, offset=0.0, sense=1.0) -> "LP":
        A = np.asarray(A, dtype=np.float64)
        rows, cols = np.nonzero(A)
        ptr = np.zeros(A.shape[0] + 1, dtype=np.int64)
        np.add.at(ptr, rows + 1, 1)
        return cls(np.cumsum(ptr), cols, A[rows, cols], c, lb, ub, rl, ru, offset, sense)


def _warm_array(val, size: int, what: str) -> np.ndarray:
    """A warm-start vector of exactly `size` entries (the native side copies
    that many doubles from the pointer)."""
    arr = _f64(val).reshape(-1)
    if arr.size != size:
        raise ValueError(f"{what} has {arr.size} entries; {size} expected")
    return arr


def _options_struct(lib, options: dict, keep: list, lp: "LP | None" = None) -> _Options:
    opts = _Options()
    lib.niryukti_default_options(C.byref(opts))
    for key, val in options.items():
        if key in ("initial_x", "initial_y"):
            if lp is None:
                raise ValueError(f"{key} needs the LP it warm-starts")
            arr = _warm_array(val, lp.cols if key == "initial_x" else lp.rows, key)
            keep.append(arr)
            setattr(opts, key, _ptr(arr))
        elif key in _OPTION_STRINGS:
            b = str(val).encode()
            keep.append(b)
            setattr(opts, key, b)
        elif key in ("presolve", "restart", "adaptive", "graphs", "verbose"):
            setattr(opts, key, int(bool(val)))
        elif hasattr(opts, key):
            setattr(opts, key, val)
        else:
            raise ValueError(f"unknown option {key!r}")
    return opts


def _result_dict(res: _Result, x, y) -> dict:
    out = {name: getattr(res, name) for name, _ in _Result._fields_ if name not in ("status", "message", "x", "y")}
    out["status"] = res.status.decode()
    out["message"] = res.message.decode()
    out["x"] = x
    out["y"] = y
    return out


_OPTION_STRINGS = ("device", "method", "scaling", "primal_weight", "matrix_precision", "gpu_indices",
                   "checkpoint_path", "resume_path")


def solve_lp(lp: LP, **options) -> dict:
    """Solves `lp` with the GPU engine; keyword options mirror the reference's Options."""
    lib = load_library()
    keep = []
    opts = _options_struct(lib, options, keep, lp)
    x = np.zeros(lp.cols)
    y = np.zeros(lp.rows)
    res = _Result()
    res.x = _ptr(x)
    res.y = _ptr(y)
    s = lp._struct()
    rc = lib.niryukti_solve_lp(C.byref(s), C.byref(opts), C.byref(res))
    if rc != 0:
        raise RuntimeError(lib.niryukti_engine_last_error().decode())
    return _result_dict(res, x, y)


def prepare_lp(lp: LP, **options) -> dict:
    """The engine's host prepare step (presolve + scaling) on its own; returns
    the prepared arrays, scales and maps, or {"failure": message}."""
    lib = load_library()
    keep = []
    opts = _options_struct(lib, options, keep, lp)
    m, n, nnz = lp.rows, lp.cols, lp.nnz
    out = {k: np.zeros(max(size, 1)) for k, size in (("value", nnz), ("c", n), ("lb", n), ("ub", n), ("rl", m),
                                                      ("ru", m), ("column_scale", n), ("row_scale", m))}
    out.update(ptr=np.zeros(m + 1, np.int64), index=np.zeros(max(nnz, 1), np.int64), col_map=np.zeros(max(n, 1), np.int64),
               row_map=np.zeros(max(m, 1), np.int64))
    rows, cols, osc, off = C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
    held = lp._struct()
    rc = lib.niryukti_prepare_lp(C.byref(held), C.byref(opts), C.byref(rows), C.byref(cols), _ptr(out["ptr"], _i64p),
                                 _ptr(out["index"], _i64p), _ptr(out["value"]), _ptr(out["c"]), _ptr(out["lb"]),
                                 _ptr(out["ub"]), _ptr(out["rl"]), _ptr(out["ru"]), _ptr(out["column_scale"]),
                                 _ptr(out["row_scale"]), _ptr(out["col_map"], _i64p), _ptr(out["row_map"], _i64p),
                                 C.byref(osc), C.byref(off))
    if rc < 0:
        raise RuntimeError(lib.niryukti_engine_last_error().decode())
    if rc > 0:
        return {"failure": lib.niryukti_engine_last_error().decode()}
    r, c = rows.value, cols.value
    k = int(out["ptr"][r])
    return {"failure": "", "rows": r, "cols": c, "ptr": out["ptr"][:r + 1], "index": out["index"][:k],
            "value": out["value"][:k], "c": out["c"][:c], "lb": out["lb"][:c], "ub": out["ub"][:c],
            "rl": out["rl"][:r], "ru": out["ru"][:r], "column_scale": out["column_scale"][:c],
            "row_scale": out["row_scale"][:r], "col_map": out["col_map"][:c], "row_map": out["row_map"][:r],
            "objective_scale": osc.value, "offset": off.value}


def verify_lp(lp: LP, x, y, safe_bound: bool = True) -> dict:
    """Original-space accuracy of (x, y) by the engine's certifying verifier
    (the reference's vantage::verify)."""
    lib = load_library()
    x, y = _warm_array(x, lp.cols, "x"), _warm_array(y, lp.rows, "y")
    s = lp._struct()
    a = _Accuracy()
    if lib.niryukti_verify_lp(C.byref(s), _ptr(x), _ptr(y), int(bool(safe_bound)), C.byref(a)) != 0:
        raise RuntimeError(lib.niryukti_engine_last_error().decode())
    return {k: getattr(a, k) for k, _ in _Accuracy._fields_}


def power_norm(lp: LP, iterations: int) -> float:
    """Power-iteration estimate of ||A||_2 (the reference's --power-iterations)."""
    lib = load_library()
    s = lp._struct()
    return float(lib.niryukti_power_norm(C.byref(s), int(iterations)))


def scale_lp(lp: LP, passes: int = 5) -> dict:
    """Ruiz passes, objective scale and the explicit transpose of a presolved
    LP on the device (include/niryukti_prepare.h).  Returns the scaled arrays,
    the factors and A^T (t_ptr, t_index, t_value)."""
    lib = load_library()
    m, n, nnz = lp.rows, lp.cols, lp.nnz
    s = _PdhgLP(m, n, _ptr(lp.ptr, _i64p), _ptr(lp.index, _i64p), _ptr(lp.value), _ptr(lp.c), _ptr(lp.lb),
                _ptr(lp.ub), _ptr(lp.rl), _ptr(lp.ru))
    out = {k: np.zeros(size) for k, size in (("value", nnz), ("c", n), ("lb", n), ("ub", n), ("rl", m), ("ru", m),
                                             ("row_scale", m), ("column_scale", n), ("t_value", nnz))}
    out["t_ptr"] = np.zeros(n + 1, np.int64)
    out["t_index"] = np.zeros(nnz, np.int64)
    r = _Scaled(*(_ptr(out[k]) for k in ("value", "c", "lb", "ub", "rl", "ru", "row_scale", "column_scale")), 1.0,
                _ptr(out["t_ptr"], _i64p), _ptr(out["t_index"], _i64p), _ptr(out["t_value"]), 0.0, 0.0)
    if lib.npdhg_scale_lp(C.byref(s), int(passes), C.byref(r)) != 0:
        raise RuntimeError(lib.npdhg_scale_last_error().decode())
    out.update(objective_scale=r.objective_scale, device_seconds=r.device_seconds, seconds=r.seconds)
    return out


class Model:
    """A model read by the native readers: a .json / .mps path, or native JSON
    (text or dict).  `lp` holds the canonical arrays; `result_json` renders a
    result with the reference's document keys."""

    def __init__(self, path=None, *, json_model=None):
        lib = load_library()
        if (path is None) == (json_model is None):
            raise ValueError("give exactly one of path and json_model")
        text = None
        if json_model is not None:
            text = json_model if isinstance(json_model, str) else json.dumps(json_model, allow_nan=False)
        h = lib.niryukti_model_load(os.fsencode(str(path)) if path is not None else None,
                                    text.encode() if text is not None else None)
        if not h:
            raise ValueError(lib.niryukti_engine_last_error().decode())
        self._lib = lib
        self._h = C.c_void_p(h)
        s = _LP()
        name, fp = C.c_char_p(), C.c_char_p()
        lib.niryukti_model_view(self._h, C.byref(s), C.byref(name), C.byref(fp))
        m, n = s.rows, s.cols
        nnz = s.ptr[m]

        def view(p, count, dtype):
            return np.ctypeslib.as_array(p, shape=(count,)).copy() if count else np.zeros(0, dtype=dtype)

        self.lp = LP(view(s.ptr, m + 1, np.int64), view(s.index, nnz, np.int64), view(s.value, nnz, np.float64),
                     view(s.c, n, np.float64), view(s.lb, n, np.float64), view(s.ub, n, np.float64),
                     view(s.rl, m, np.float64), view(s.ru, m, np.float64), s.offset, s.sense)
        self.name = name.value.decode()
        self.fingerprint = fp.value.decode()

    def write(self, path) -> None:
        """Writes .mps or native .json by extension (the reference's `convert`)."""
        if self._lib.niryukti_model_write(self._h, os.fsencode(str(path))) != 0:
            raise ValueError(self._lib.niryukti_engine_last_error().decode())

    def result_json(self, result: dict, backend: str = "hip") -> dict:
        res = _Result()
        for key, _ in _Result._fields_:
            if key in ("x", "y"):
                continue
            if key in ("status", "message"):
                setattr(res, key, str(result.get(key, "")).encode()[:250])
            elif key in result:
                setattr(res, key, result[key])
        x = _f64(result["x"]) if result.get("x") is not None else None
        y = _f64(result["y"]) if result.get("y") is not None else None
        res.x = _ptr(x) if x is not None and len(x) == self.lp.cols else None
        res.y = _ptr(y) if y is not None and len(y) == self.lp.rows else None
        raw = self._lib.niryukti_result_json(self._h, C.byref(res), backend.encode())
        if raw is None:
            raise RuntimeError(self._lib.niryukti_engine_last_error().decode())
        return json.loads(raw)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.niryukti_model_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Instance:
    """Seeded synthetic LP with a planted optimum, generated natively."""

    def __init__(self, config: int, seed: int = 42, scale: int = 1, index: int = 0):
        lib = load_library()
        h = lib.niryukti_instance_create(int(config), int(seed), int(scale), int(index))
        if not h:
            raise RuntimeError(lib.niryukti_engine_last_error().decode())
        self._h = C.c_void_p(h)
        s = _LP()
        obj = C.c_double()
        px, py = _f64p(), _f64p()
        lib.niryukti_instance_view(self._h, C.byref(s), C.byref(obj), C.byref(px), C.byref(py))
        m, n = s.rows, s.cols
        nnz = s.ptr[m] if m >= 0 else 0

        def view(p, count, dtype):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            return np.ctypeslib.as_array(p, shape=(count,)).copy()

        self.lp = LP(view(s.ptr, m + 1, np.int64), view(s.index, nnz, np.int64), view(s.value, nnz, np.float64),
                     view(s.c, n, np.float64), view(s.lb, n, np.float64), view(s.ub, n, np.float64),
                     view(s.rl, m, np.float64), view(s.ru, m, np.float64))
        self.planted_objective = obj.value
        self.planted_x = view(px, n, np.float64)
        self.planted_y = view(py, m, np.float64)
        lib.niryukti_instance_destroy(self._h)
        self._h = None


class Backend:
    """Direct handle on the HIP iteration backend (the npdhg_* seam)."""

    def __init__(self, lp: LP, x0, y0, adaptive=True, halpern=False, reflection=0.0, residual_restarts=False,
                 graphs=True, index_bits=0, fp32_values=False, fp32_arithmetic=False):
        self._lib = load_library()
        self.lp = lp
        self._s = _PdhgLP(lp.rows, lp.cols, _ptr(lp.ptr, _i64p), _ptr(lp.index, _i64p), _ptr(lp.value),
                          _ptr(lp.c), _ptr(lp.lb), _ptr(lp.ub), _ptr(lp.rl), _ptr(lp.ru))
        cfg = _PdhgConfig(int(adaptive), int(halpern), float(reflection), int(residual_restarts), int(graphs),
                          int(index_bits), int(fp32_values), int(fp32_arithmetic))
        x0, y0 = _f64(x0), _f64(y0)
        h = C.c_void_p()
        self._check(self._lib.npdhg_create(C.byref(self._s), _ptr(x0), _ptr(y0), C.byref(cfg), C.byref(h)))
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._lib.npdhg_last_error().decode())

    def advance(self, count: int, tau: float, sigma: float) -> int:
        acc = C.c_int()
        self._check(self._lib.npdhg_advance(self._h, int(count), float(tau), float(sigma), C.byref(acc)))
        return acc.value

    def status(self) -> dict:
        s = _PdhgStatus()
        self._check(self._lib.npdhg_status_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _PdhgStatus._fields_}

    def candidates(self):
        x, y = np.zeros(self.lp.cols), np.zeros(self.lp.rows)
        xa, ya = np.zeros(self.lp.cols), np.zeros(self.lp.rows)
        self._check(self._lib.npdhg_candidates(self._h, _ptr(x), _ptr(y), _ptr(xa), _ptr(ya)))
        return x, y, xa, ya

    def reset(self, x, y):
        x, y = _f64(x), _f64(y)
        self._check(self._lib.npdhg_reset(self._h, _ptr(x), _ptr(y)))

    def set_fp32_phase(self, on: bool):
        """FP32 arithmetic phase on / off (backends created with fp32_arithmetic)."""
        self._check(self._lib.npdhg_set_fp32_phase(self._h, int(bool(on))))

    def reset_candidate(self, average: bool):
        self._check(self._lib.npdhg_reset_candidate(self._h, int(bool(average))))

    def attach_original(self, lp0: LP, cols_map, lower_rows, upper_rows, fixed, column_scale, row_scale,
                        objective_scale: float):
        """Device-side checks against `lp0` with the prepare step's maps."""
        keep = [np.ascontiguousarray(cols_map, dtype=np.int64), np.ascontiguousarray(lower_rows, dtype=np.int64),
                np.ascontiguousarray(upper_rows, dtype=np.int64), _f64(fixed), _f64(column_scale), _f64(row_scale)]
        o = _PdhgOriginal(lp0.rows, lp0.cols, _ptr(lp0.ptr, _i64p), _ptr(lp0.index, _i64p), _ptr(lp0.value),
                          _ptr(lp0.c), _ptr(lp0.lb), _ptr(lp0.ub), _ptr(lp0.rl), _ptr(lp0.ru), float(lp0.offset),
                          _ptr(keep[0], _i64p), _ptr(keep[1], _i64p), _ptr(keep[2], _i64p), _ptr(keep[3]),
                          _ptr(keep[4]), _ptr(keep[5]), float(objective_scale))
        self._check(self._lib.npdhg_attach_original(self._h, C.byref(o)))
        self._lp0 = lp0

    def evaluate(self):
        cur, avg = _PdhgAccuracy(), _PdhgAccuracy()
        self._check(self._lib.npdhg_evaluate(self._h, C.byref(cur), C.byref(avg)))
        return tuple({k: getattr(a, k) for k, _ in _PdhgAccuracy._fields_} for a in (cur, avg))

    def keep_best(self, average: bool):
        self._check(self._lib.npdhg_keep_best(self._h, int(bool(average))))

    def original_point(self, which: int):
        x, y = np.zeros(self._lp0.cols), np.zeros(self._lp0.rows)
        self._check(self._lib.npdhg_original_point(self._h, int(which), _ptr(x), _ptr(y)))
        return x, y

    def farkas_displacement(self, average: bool):
        possible = C.c_int()
        ray = np.zeros(self._lp0.rows)
        self._check(self._lib.npdhg_farkas_displacement(self._h, int(bool(average)), C.byref(possible), _ptr(ray)))
        return bool(possible.value), ray

    def epoch_displacement(self, average: bool):
        dx, dy = C.c_double(), C.c_double()
        self._check(self._lib.npdhg_epoch_displacement(self._h, int(bool(average)), C.byref(dx), C.byref(dy)))
        return dx.value, dy.value

    def set_epoch(self, average: bool):
        self._check(self._lib.npdhg_set_epoch(self._h, int(bool(average))))

    def spmv(self, x, transpose=False):
        x = _f64(x)
        out = np.zeros(self.lp.cols if transpose else self.lp.rows)
        self._check(self._lib.npdhg_spmv(self._h, int(bool(transpose)), _ptr(x), _ptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.npdhg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Batch:
    """Batched / served engine (include/niryukti_batch.h): many LPs resident in one
    HBM arena, one workgroup per LP, continuous admission between launches."""

    def __init__(self, arena_bytes: int = 8 << 30, max_slots: int = 4096, **options):
        self._lib = load_library()
        self._keep = []
        opts = _options_struct(self._lib, options, self._keep)
        h = C.c_void_p()
        self._check(self._lib.niryukti_batch_create(C.byref(opts), int(arena_bytes), int(max_slots), C.byref(h)))
        self._h = h
        self._shapes = {}

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(self._lib.niryukti_batch_last_error().decode())
        return rc

    def submit(self, lp: LP, x0=None, y0=None) -> int:
        """Queues `lp` (copied) for prepare and admission; returns its ticket.
        x0 / y0: optional original-space warm start of cols / rows entries."""
        s = lp._struct()
        wx = _warm_array(x0, lp.cols, "x0") if x0 is not None else None
        wy = _warm_array(y0, lp.rows, "y0") if y0 is not None else None
        px = _ptr(wx) if wx is not None else None
        py = _ptr(wy) if wy is not None else None
        t = C.c_int64()
        self._check(self._lib.niryukti_batch_submit(self._h, C.byref(s), px, py, C.byref(t)))
        del wx, wy
        self._shapes[t.value] = (lp.cols, lp.rows)
        return t.value

    def step(self, checks_per_launch: int = 1) -> int:
        running = C.c_int64()
        self._check(self._lib.niryukti_batch_step(self._h, int(checks_per_launch), C.byref(running)))
        return running.value

    def admit(self) -> int:
        n = C.c_int64()
        self._check(self._lib.niryukti_batch_admit(self._h, C.byref(n)))
        return n.value

    def run(self, checks_per_launch: int = 5) -> None:
        self._check(self._lib.niryukti_batch_run(self._h, int(checks_per_launch)))

    def result(self, ticket: int):
        # an unknown or released ticket gets the engine's own error
        n, m = self._shapes.get(ticket, (0, 0))
        x, y = np.zeros(n), np.zeros(m)
        res = _Result()
        res.x, res.y = _ptr(x), _ptr(y)
        done = self._check(self._lib.niryukti_batch_result(self._h, int(ticket), C.byref(res)))
        return _result_dict(res, x, y) if done else None

    def set_handoff(self, iterations: int, workers: int = 4) -> None:
        """LPs still running after `iterations` iterations continue on the
        grid-wide engine (0 disables); see include/niryukti_batch.h."""
        self._check(self._lib.niryukti_batch_set_handoff(self._h, int(iterations), int(workers)))

    def finished(self, capacity: int = 4096) -> list:
        """Tickets finished since the last call, in finishing order."""
        buf = (C.c_int64 * capacity)()
        n = C.c_int64()
        self._check(self._lib.niryukti_batch_finished(self._h, buf, int(capacity), C.byref(n)))
        return list(buf[:n.value])

    def collect(self, capacity: int = 8192, release: bool = True):
        """Finished tickets and their results (every field but x and y) in one
        call: (tickets, numpy structured array with the fields of
        niryukti_result, status and message as bytes); with `release` each
        LP's host copy is freed."""
        if getattr(self, "_collect_cap", 0) < capacity:
            self._collect_tickets = (C.c_int64 * capacity)()
            self._collect_results = (_Result * capacity)()
            self._collect_cap = capacity
        n = C.c_int64()
        self._check(self._lib.niryukti_batch_collect(self._h, self._collect_tickets, self._collect_results,
                                                     int(capacity), 1 if release else 0, C.byref(n)))
        k = n.value
        tickets = np.frombuffer(self._collect_tickets, dtype=np.int64, count=k).copy()
        rows = np.frombuffer(self._collect_results, dtype=_RESULT_DTYPE, count=k).copy()
        if release:
            for t in tickets.tolist():
                self._shapes.pop(t, None)
        return tickets, rows

    def release(self, ticket: int) -> None:
        """Frees the host copy of a finished LP; the ticket is invalid afterwards."""
        self._check(self._lib.niryukti_batch_release(self._h, int(ticket)))
        self._shapes.pop(ticket, None)

    def stats(self) -> dict:
        st = _BatchStats()
        self._check(self._lib.niryukti_batch_stats_get(self._h, C.byref(st)))
        out = {}
        for k, _ in _BatchStats._fields_:
            v = getattr(st, k)
            out[k] = list(v) if k.startswith(("class_", "shape_")) else v
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.niryukti_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def device_count() -> int:
    return load_library().npdhg_device_count()


def device_name(index: int = 0) -> str:
    buf = C.create_string_buffer(256)
    load_library().npdhg_device_name(index, buf, 256)
    return buf.value.decode()


from .session import NativeSession  # noqa: E402  (uses load_library above)
from .serve import Server  # noqa: E402
