"""Serving loop: continuous batching of LP solves on one MI355X.

The reference serves `POST /v1/solve {"model": <native JSON>, "options": {...}}`
(python/niryukti/service.py:54-69) with one synchronous solve per worker and
answers 429 when every worker is busy.  Here requests are queued instead and
admitted as device memory frees up, by the native server of
include/niryukti_serve.h (a C++ scheduler thread and grid workers; this module
only binds it):

* LPs below `large_nonzeros` nonzeros and `batch_max_size` rows + columns go
  to one batched engine: each is a workgroup of the batched kernel, prepared
  on the engine's host workers, admitted into the HBM arena between launches,
  certified on the host and retired while the others keep running.
* Larger LPs run on the grid-wide single-LP engine, a few at a time on worker
  threads, each on its own HIP stream, admitted against the rest of the HBM
  budget.

Both paths are deterministic per LP (fixed-order reductions, no cross-LP
state), so a result is the same alone and in any mix.  `solve_request`
keeps the reference's request and response shape without the HTTP layer.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time

from . import LP, Model, _f64, _options_struct, _ptr, _Result, _result_dict, _warm_array, load_library

_REQUEST_KEYS = {"model", "options"}
_OPTION_KEYS = {"device", "method", "tol", "time_limit", "threads", "iterations"}


def grid_footprint(lp: LP) -> int:
    """HBM bytes of one single-LP solve: the reference's estimate
    (src/cuda.cu:1294-1295) for the scaled LP, twice over for the resident
    original model used by the device-side checks."""
    return int(2 * (48 * lp.nnz + 240 * (lp.rows + lp.cols + 2)))


def parse_request(request, tol: float = 1e-6, time_limit: float = 60.0, iteration_limit: int = 100_000):
    """Validates a `/v1/solve` body as the reference does
    (python/niryukti/service.py:54-69) and reads its model.  Returns the model
    and the solve options that differ from the server's defaults; raises
    ValueError for anything the reference answers with 400."""
    if not isinstance(request, dict) or set(request) - _REQUEST_KEYS:
        raise ValueError("Unknown solve fields")
    if not isinstance(request.get("model"), dict):
        raise ValueError("Provide an inline native JSON model")
    options = request.get("options", {})
    if not isinstance(options, dict) or set(options) - _OPTION_KEYS:
        raise ValueError("Unsupported solve options")
    limit = float(options.get("time_limit", time_limit))
    if not 0 < limit <= 3600:
        raise ValueError("Time limit exceeds API resource policy")
    if not 1 <= int(options.get("threads", 1)) <= 64:
        raise ValueError("Threads must be 1..64")
    model = Model(json_model=request["model"])
    extra = {}
    if "tol" in options and float(options["tol"]) != tol:
        extra["tol"] = float(options["tol"])
    if limit != time_limit:
        extra["time_limit"] = limit
    if "iterations" in options and int(options["iterations"]) != iteration_limit:
        extra["iteration_limit"] = int(options["iterations"])
    if options.get("method", "pdhg") not in ("pdhg", "auto"):
        extra["method"] = str(options["method"])
    if options.get("device", "gpu") not in ("gpu", "hip", "auto"):
        extra["device"] = str(options["device"])
    return model, extra


class _Limits(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("batch_fraction", C.c_double), ("large_nonzeros", C.c_int64),
                ("grid_workers", C.c_int), ("checks_per_launch", C.c_int), ("max_slots", C.c_int64),
                ("batch_max_size", C.c_int64)]


class _Stats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "finished", "batch_routed", "grid_routed", "grid_running",
                                         "grid_bytes_in_use", "grid_budget_bytes", "batch_rounds")]


def _bind(lib):
    if getattr(lib, "_serve_bound", False):
        return lib
    lib.niryukti_server_default_limits.argtypes = [C.POINTER(_Limits)]
    lib.niryukti_server_create.argtypes = [C.c_void_p, C.POINTER(_Limits), C.POINTER(C.c_void_p)]
    lib.niryukti_server_destroy.argtypes = [C.c_void_p]
    lib.niryukti_server_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.POINTER(C.c_int64)]
    lib.niryukti_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.POINTER(_Result),
                                         C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.niryukti_server_release.argtypes = [C.c_void_p, C.c_int64]
    lib.niryukti_server_stats_get.argtypes = [C.c_void_p, C.POINTER(_Stats)]
    lib.niryukti_server_last_error.restype = C.c_char_p
    lib._serve_bound = True
    return lib


class Request:
    """One submitted LP; `wait()` returns its result dict (or raises its error)."""

    def __init__(self, server: "Server", rid: int, lp: LP, x0, y0, route: str):
        self._server, self.id = server, rid
        self.lp, self.x0, self.y0, self.route = lp, x0, y0, route
        self.arrival = time.perf_counter()
        self.finish = None
        self.result = None
        self._latency = math.nan
        self._lock = threading.Lock()

    @property
    def latency(self) -> float:
        """Seconds from submit to completion, measured by the native server."""
        return self._latency

    def _collect(self, timeout) -> bool:
        with self._lock:
            return self._collect_locked(timeout)

    def _collect_locked(self, timeout) -> bool:
        if self.finish is not None:
            return True
        x, y = _f64([0.0] * self.lp.cols), _f64([0.0] * self.lp.rows)
        res = _Result()
        res.x, res.y = _ptr(x), _ptr(y)
        latency, route = C.c_double(), C.c_int()
        rc = self._server._lib.niryukti_server_wait(self._server._h, self.id, -1.0 if timeout is None else timeout,
                                                    C.byref(res), C.byref(latency), C.byref(route))
        if rc == 0:
            return False
        if rc < 0:
            message = self._server._lib.niryukti_server_last_error().decode()
            if message == "Unknown request":
                raise RuntimeError(message)
            self.result = RuntimeError(message)
        else:
            self.result = _result_dict(res, x, y)
        self._latency = latency.value
        self.route = "grid" if route.value == 1 else "batch"
        self.finish = self.arrival + self._latency
        self._server._release(self)
        return True

    def done(self) -> bool:
        return self._collect(0.0)

    def wait(self, timeout=None) -> dict:
        if not self._collect(timeout):
            raise TimeoutError("request still running")
        if isinstance(self.result, BaseException):
            raise self.result
        return self.result


class Server:
    """Continuous-batching LP server on the current HIP device (the native
    server of include/niryukti_serve.h)."""

    def __init__(self, hbm_budget: int = 64 << 30, batch_fraction: float = 0.75, large_nonzeros: int = 200_000,
                 grid_workers: int = 2, checks_per_launch: int = 1, max_slots: int = 4096, tol: float = 1e-6,
                 time_limit: float = 60.0, iteration_limit: int = 100_000, batch_max_size: int = 8000):
        if not 0 < batch_fraction < 1 or hbm_budget <= 0 or grid_workers < 1:
            raise ValueError("invalid server limits")
        self._lib = _bind(load_library())
        self.tol, self.time_limit, self.iteration_limit = tol, time_limit, iteration_limit
        self.large_nonzeros = large_nonzeros
        self.batch_max_size = batch_max_size
        self._keep = []
        opts = _options_struct(self._lib, dict(tol=tol, time_limit=time_limit, iteration_limit=iteration_limit),
                               self._keep)
        limits = _Limits(int(hbm_budget), float(batch_fraction), int(large_nonzeros), int(grid_workers),
                         int(checks_per_launch), int(max_slots), int(batch_max_size))
        h = C.c_void_p()
        if self._lib.niryukti_server_create(C.byref(opts), C.byref(limits), C.byref(h)) != 0:
            raise RuntimeError(self._lib.niryukti_server_last_error().decode())
        self._h = h
        self._open = {}  # requests whose results have not been collected yet
        self._open_lock = threading.Lock()

    # ------------------------------------------------------------ submit --
    def submit(self, lp: LP, x0=None, y0=None, *, route: str | None = None, **options) -> Request:
        """Queues one LP (optionally warm-started from an original-space x0/y0).
        Options other than the server's own tol / time_limit / iteration_limit
        send the LP to the grid engine, which takes per-solve options."""
        if not getattr(self, "_h", None):
            raise RuntimeError("server is closed")
        if route not in (None, "batch", "grid"):
            raise ValueError("route must be 'batch' or 'grid'")
        wx = _warm_array(x0, lp.cols, "x0") if x0 is not None else None
        wy = _warm_array(y0, lp.rows, "y0") if y0 is not None else None
        keep = []
        opts = None
        if options:
            merged = dict(tol=self.tol, time_limit=self.time_limit, iteration_limit=self.iteration_limit)
            merged.update(options)
            opts = _options_struct(self._lib, merged, keep, lp)
        s = lp._struct()
        rid = C.c_int64()
        code = {None: -1, "batch": 0, "grid": 1}[route]
        rc = self._lib.niryukti_server_submit(self._h, C.byref(s), _ptr(wx) if wx is not None else None,
                                              _ptr(wy) if wy is not None else None, code,
                                              C.byref(opts) if opts is not None else None, C.byref(rid))
        if rc != 0:
            raise ValueError(self._lib.niryukti_server_last_error().decode())
        large = lp.nnz >= self.large_nonzeros or lp.rows + lp.cols > self.batch_max_size
        guess = "grid" if (route == "grid" or options or (route is None and large)) else "batch"
        req = Request(self, rid.value, lp, x0, y0, guess)
        with self._open_lock:
            self._open[req.id] = req
        return req

    def _release(self, req: Request):
        self._lib.niryukti_server_release(self._h, int(req.id))
        with self._open_lock:
            self._open.pop(req.id, None)

    def stats(self) -> dict:
        st = _Stats()
        if self._lib.niryukti_server_stats_get(self._h, C.byref(st)) != 0:
            raise RuntimeError(self._lib.niryukti_server_last_error().decode())
        return {k: getattr(st, k) for k, _ in _Stats._fields_}

    # --------------------------------------------------------- /v1/solve --
    def solve_request(self, request) -> tuple:
        """The reference's `/v1/solve` contract: returns (HTTP status, body).
        400 for malformed input; otherwise 200 with the result document,
        whatever the solver status.  Never 429: requests queue."""
        try:
            model, extra = parse_request(request, self.tol, self.time_limit, self.iteration_limit)
        except (ValueError, TypeError, OverflowError, KeyError) as error:
            return 400, {"error": str(error)[:2048]}
        try:
            req = self.submit(model.lp, **extra)
            result = req.wait()
            body = model.result_json(result, backend="hip-batch" if req.route == "batch" else "hip")
        except (ValueError, RuntimeError) as error:
            return 400, {"error": str(error)[:2048]}
        body["serving"] = {"route": req.route, "latency_seconds": req.latency}
        return 200, body

    def close(self):
        """Finishes every submitted request (their results stay readable from
        their Request objects), then stops the server."""
        if getattr(self, "_h", None):
            with self._open_lock:
                pending = list(self._open.values())
            for req in pending:
                req._collect(None)
            self._lib.niryukti_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()
