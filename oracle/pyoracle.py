"""ctypes binding of the CPU oracle (oracle/pdhg_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
benchmark's CPU-baseline leg, never by the engine.  Build the library with
`make -f oracle/oracle.mk`.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
LIB = ROOT / "lib" / "liboracle_pdhg.so"
CLI = ROOT / "lib" / "pdhg_oracle"
REF_DRIVER = ROOT / "_ref" / "ref_driver"

_f64p = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class _LP(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("ptr", _i64p), ("index", _i64p), ("value", _f64p),
                ("c", _f64p), ("lb", _f64p), ("ub", _f64p), ("rl", _f64p), ("ru", _f64p), ("offset", C.c_double)]


class _Options(C.Structure):
    _fields_ = [("tol", C.c_double), ("time_limit", C.c_double), ("iteration_limit", C.c_int64),
                ("check_every", C.c_int), ("scaling_passes", C.c_int), ("threads", C.c_int),
                ("presolve", C.c_int), ("restart", C.c_int), ("adaptive", C.c_int), ("method", C.c_char_p),
                ("scaling", C.c_char_p), ("primal_weight", C.c_char_p), ("initial_x", _f64p), ("initial_y", _f64p)]


class _Accuracy(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("objective", "primal", "dual", "gap", "kkt", "lower_bound",
                                          "complementarity", "primal_absolute", "dual_absolute")] + [("finite", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_char * 32), ("accuracy", _Accuracy)] + \
        [(k, C.c_int64) for k in ("iterations", "restarts", "rejected_steps", "weight_updates",
                                  "host_candidate_checks", "removed_rows", "removed_columns")] + \
        [(k, C.c_double) for k in ("seconds", "iteration_seconds", "preprocess_seconds", "verification_seconds")] + \
        [(k, C.c_double) for k in ("initial_step", "initial_weight", "first_restart_weight")] + \
        [("first_restart_iteration", C.c_int64), ("x", _f64p), ("y", _f64p)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle library missing at {LIB}; run `make -f oracle/oracle.mk`")
        l = C.CDLL(str(LIB))
        l.oracle_default_options.argtypes = [C.POINTER(_Options)]
        l.oracle_solve.argtypes = [C.POINTER(_LP), C.POINTER(_Options), C.POINTER(_Result), C.c_char_p, C.c_int]
        l.oracle_verify.argtypes = [C.POINTER(_LP), _f64p, _f64p, C.c_int, C.POINTER(_Accuracy), C.c_char_p, C.c_int]
        l.oracle_infeasibility_bound.argtypes = [C.POINTER(_LP), _f64p]
        l.oracle_infeasibility_bound.restype = C.c_double
        l.oracle_prepare.argtypes = [C.POINTER(_LP), C.POINTER(_Options), _i64p, _i64p, _f64p, _f64p, _f64p,
                                     _i64p, _i64p, C.c_char_p, C.c_int]
        l.oracle_prepared_model.argtypes = [C.POINTER(_LP), C.POINTER(_Options), _i64p, _i64p, _i64p, _i64p,
                                            _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, C.POINTER(C.c_double),
                                            C.c_char_p, C.c_int]
        l.oracle_backend_create.argtypes = [C.POINTER(_LP), _f64p, _f64p, C.c_int, C.c_int, C.c_double, C.c_int,
                                            C.c_char_p, C.c_int]
        l.oracle_backend_create.restype = C.c_void_p
        l.oracle_backend_advance.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_char_p, C.c_int]
        l.oracle_backend_candidates.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
        l.oracle_backend_reset.argtypes = [C.c_void_p, _f64p, _f64p]
        l.oracle_backend_rejected.argtypes = [C.c_void_p]
        l.oracle_backend_rejected.restype = C.c_int64
        l.oracle_backend_restart_requested.argtypes = [C.c_void_p]
        l.oracle_backend_destroy.argtypes = [C.c_void_p]
        l.oracle_pid_create.restype = C.c_void_p
        l.oracle_pid_update.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        l.oracle_pid_update.restype = C.c_double
        l.oracle_pid_destroy.argtypes = [C.c_void_p]
        _lib = l
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _Held:
    """Keeps numpy arrays alive while a C struct points into them."""

    def __init__(self, lp):
        self.ptr = np.ascontiguousarray(lp.ptr, dtype=np.int64)
        self.index = np.ascontiguousarray(lp.index, dtype=np.int64)
        self.arrays = {k: _f(getattr(lp, k)) for k in ("value", "c", "lb", "ub", "rl", "ru")}
        a = self.arrays
        self.s = _LP(len(a["rl"]), len(a["c"]), self.ptr.ctypes.data_as(_i64p), self.index.ctypes.data_as(_i64p),
                     *(a[k].ctypes.data_as(_f64p) for k in ("value", "c", "lb", "ub", "rl", "ru")),
                     float(getattr(lp, "offset", 0.0)))


def solve(lp, **kw) -> dict:
    l = lib()
    h = _Held(lp)
    o = _Options()
    l.oracle_default_options(C.byref(o))
    keep = []
    for k, v in kw.items():
        if k in ("method", "scaling", "primal_weight"):
            b = str(v).encode()
            keep.append(b)
            setattr(o, k, b)
        elif k in ("initial_x", "initial_y"):
            arr = _f(v)
            keep.append(arr)
            setattr(o, k, arr.ctypes.data_as(_f64p))
        elif k in ("presolve", "restart", "adaptive"):
            setattr(o, k, int(bool(v)))
        else:
            setattr(o, k, v)
    x, y = np.zeros(h.s.cols), np.zeros(h.s.rows)
    r = _Result()
    r.x, r.y = x.ctypes.data_as(_f64p), y.ctypes.data_as(_f64p)
    err = C.create_string_buffer(512)
    if l.oracle_solve(C.byref(h.s), C.byref(o), C.byref(r), err, 512) != 0:
        raise RuntimeError(err.value.decode())
    out = {k: getattr(r, k) for k, _ in _Result._fields_ if k not in ("status", "accuracy", "x", "y")}
    out.update({k: getattr(r.accuracy, k) for k, _ in _Accuracy._fields_})
    out["status"] = r.status.decode()
    out["x"], out["y"] = x, y
    return out


def verify(lp, x, y, safe=True) -> dict:
    l = lib()
    h = _Held(lp)
    x, y = _f(x), _f(y)
    a = _Accuracy()
    err = C.create_string_buffer(512)
    if l.oracle_verify(C.byref(h.s), x.ctypes.data_as(_f64p), y.ctypes.data_as(_f64p), int(safe), C.byref(a), err, 512):
        raise RuntimeError(err.value.decode())
    return {k: getattr(a, k) for k, _ in _Accuracy._fields_}


def infeasibility_bound(lp, ray) -> float:
    h = _Held(lp)
    ray = _f(ray)
    return lib().oracle_infeasibility_bound(C.byref(h.s), ray.ctypes.data_as(_f64p))


def prepare(lp, **kw) -> dict:
    l = lib()
    h = _Held(lp)
    o = _Options()
    l.oracle_default_options(C.byref(o))
    keep = []
    for k, v in kw.items():
        if k == "scaling":
            b = str(v).encode()
            keep.append(b)
            o.scaling = b
        elif k == "presolve":
            o.presolve = int(bool(v))
        else:
            setattr(o, k, v)
    rows, cols = C.c_int64(), C.c_int64()
    cs, rs = np.zeros(h.s.cols), np.zeros(h.s.rows)
    cm, rm = np.zeros(h.s.cols, np.int64), np.zeros(h.s.rows, np.int64)
    osc = C.c_double()
    fail = C.create_string_buffer(256)
    rc = l.oracle_prepare(C.byref(h.s), C.byref(o), C.byref(rows), C.byref(cols), cs.ctypes.data_as(_f64p),
                          rs.ctypes.data_as(_f64p), C.byref(osc), cm.ctypes.data_as(_i64p), rm.ctypes.data_as(_i64p),
                          fail, 256)
    if rc < 0:
        raise RuntimeError(fail.value.decode())
    return {"failure": fail.value.decode(), "rows": rows.value, "cols": cols.value,
            "column_scale": cs[:cols.value], "row_scale": rs[:rows.value], "objective_scale": osc.value,
            "col_map": cm[:cols.value], "row_map": rm[:rows.value]}


def prepared_model(lp, **kw) -> dict:
    """The prepared LP itself (canonical arrays after presolve and scaling)."""
    l = lib()
    h = _Held(lp)
    o = _Options()
    l.oracle_default_options(C.byref(o))
    keep = []
    for k, v in kw.items():
        if k == "scaling":
            b = str(v).encode()
            keep.append(b)
            o.scaling = b
        elif k == "presolve":
            o.presolve = int(bool(v))
        else:
            setattr(o, k, v)
    m, n, nnz = h.s.rows, h.s.cols, int(np.asarray(lp.ptr)[-1])
    rows, cols, offset = C.c_int64(), C.c_int64(), C.c_double()
    ptr, index = np.zeros(m + 1, np.int64), np.zeros(max(nnz, 1), np.int64)
    value = np.zeros(max(nnz, 1))
    c, lb, ub, rl, ru = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
    fail = C.create_string_buffer(256)
    f = lambda a: a.ctypes.data_as(_f64p)
    rc = l.oracle_prepared_model(C.byref(h.s), C.byref(o), C.byref(rows), C.byref(cols),
                                 ptr.ctypes.data_as(_i64p), index.ctypes.data_as(_i64p), f(value), f(c), f(lb), f(ub),
                                 f(rl), f(ru), C.byref(offset), fail, 256)
    if rc < 0:
        raise RuntimeError(fail.value.decode())
    if rc > 0:
        return {"failure": fail.value.decode()}
    mr, nc = rows.value, cols.value
    k = int(ptr[mr])
    return {"failure": "", "rows": mr, "cols": nc, "ptr": ptr[:mr + 1].copy(), "index": index[:k].copy(),
            "value": value[:k].copy(), "c": c[:nc].copy(), "lb": lb[:nc].copy(), "ub": ub[:nc].copy(),
            "rl": rl[:mr].copy(), "ru": ru[:mr].copy(), "offset": offset.value}


class Backend:
    """The oracle's CPU iteration backend (src/cpu.cpp restated)."""

    def __init__(self, lp, x0, y0, adaptive=True, halpern=False, reflection=0.0, residual_restarts=False):
        self._l = lib()
        self._h_lp = _Held(lp)
        x0, y0 = _f(x0), _f(y0)
        err = C.create_string_buffer(512)
        self._h = self._l.oracle_backend_create(C.byref(self._h_lp.s), x0.ctypes.data_as(_f64p),
                                                y0.ctypes.data_as(_f64p), int(adaptive), int(halpern),
                                                float(reflection), int(residual_restarts), err, 512)
        if not self._h:
            raise RuntimeError(err.value.decode())
        self.rows, self.cols = self._h_lp.s.rows, self._h_lp.s.cols

    def advance(self, count, tau, sigma) -> int:
        err = C.create_string_buffer(512)
        r = self._l.oracle_backend_advance(self._h, int(count), float(tau), float(sigma), err, 512)
        if r < 0:
            raise RuntimeError(err.value.decode())
        return r

    def candidates(self):
        x, y, xa, ya = np.zeros(self.cols), np.zeros(self.rows), np.zeros(self.cols), np.zeros(self.rows)
        self._l.oracle_backend_candidates(self._h, *(a.ctypes.data_as(_f64p) for a in (x, y, xa, ya)))
        return x, y, xa, ya

    def reset(self, x, y):
        x, y = _f(x), _f(y)
        self._l.oracle_backend_reset(self._h, x.ctypes.data_as(_f64p), y.ctypes.data_as(_f64p))

    def rejected(self) -> int:
        return self._l.oracle_backend_rejected(self._h)

    def restart_requested(self) -> bool:
        return bool(self._l.oracle_backend_restart_requested(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            self._l.oracle_backend_destroy(self._h)
            self._h = None


class Pid:
    """The oracle's PID primal-weight controller (src/research.cpp:52-70)."""

    def __init__(self):
        self._l = lib()
        self._h = self._l.oracle_pid_create()

    def update(self, weight, dx, dy) -> float:
        return self._l.oracle_pid_update(self._h, float(weight), float(dx), float(dy))

    def __del__(self):
        if getattr(self, "_h", None):
            self._l.oracle_pid_destroy(self._h)
            self._h = None
