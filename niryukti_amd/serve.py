"""Serving loop: continuous batching of LP solves on one MI355X.

The reference serves `POST /v1/solve {"model": <native JSON>, "options": {...}}`
(python/niryukti/service.py:54-69) with one synchronous solve per worker and
answers 429 when every worker is busy.  Here requests are queued instead and
admitted as device memory frees up:

* LPs below `large_nonzeros` go to one `Batch` (include/niryukti_batch.h): each
  is a workgroup of the batched kernel, admitted into the HBM arena between
  launches, certified on the host and retired while the others keep running.
* Larger LPs run on the grid-wide single-LP engine (`solve_lp`), a few at a
  time on worker threads, each on its own HIP stream, admitted against the
  rest of the HBM budget.

Both paths are deterministic per LP (fixed-order reductions, no cross-LP
state), so a result is the same alone and in any mix.  `solve_request`
keeps the reference's request and response shape without the HTTP layer.
"""
from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor

from . import LP, Batch, Model, solve_lp

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


class Request:
    """One submitted LP; `wait()` returns its result dict."""

    def __init__(self, lp: LP, x0, y0, route: str):
        self.lp, self.x0, self.y0, self.route = lp, x0, y0, route
        self.arrival = time.perf_counter()
        self.finish = None
        self.result = None
        self._done = threading.Event()

    def _complete(self, result: dict):
        self.result = result
        self.finish = time.perf_counter()
        self._done.set()

    @property
    def latency(self) -> float:
        return (self.finish - self.arrival) if self.finish is not None else math.nan

    def done(self) -> bool:
        return self._done.is_set()

    def wait(self, timeout=None) -> dict:
        if not self._done.wait(timeout):
            raise TimeoutError("request still running")
        if isinstance(self.result, BaseException):
            raise self.result
        return self.result


class Server:
    """Continuous-batching LP server on the current HIP device."""

    def __init__(self, hbm_budget: int = 64 << 30, batch_fraction: float = 0.75, large_nonzeros: int = 200_000,
                 grid_workers: int = 2, checks_per_launch: int = 1, max_slots: int = 4096, tol: float = 1e-6,
                 time_limit: float = 60.0, iteration_limit: int = 100_000):
        if not 0 < batch_fraction < 1 or hbm_budget <= 0 or grid_workers < 1:
            raise ValueError("invalid server limits")
        self.tol, self.time_limit, self.iteration_limit = tol, time_limit, iteration_limit
        self.large_nonzeros = large_nonzeros
        self.checks_per_launch = checks_per_launch
        self._batch = Batch(arena_bytes=int(hbm_budget * batch_fraction), max_slots=max_slots, tol=tol,
                            time_limit=time_limit, iteration_limit=iteration_limit)
        self._grid_budget = int(hbm_budget * (1 - batch_fraction))
        self._grid_used = 0
        self._grid_running = 0
        self._grid_cv = threading.Condition()
        self._pool = ThreadPoolExecutor(grid_workers, thread_name_prefix="niryukti-grid")
        self._cv = threading.Condition()
        self._incoming: list[Request] = []
        self._tickets: dict[int, Request] = {}
        self._closing = False
        self._error = None
        self._thread = threading.Thread(target=self._loop, name="niryukti-scheduler", daemon=True)
        self._thread.start()

    # ------------------------------------------------------------ submit --
    def submit(self, lp: LP, x0=None, y0=None, *, route: str | None = None, **options) -> Request:
        """Queues one LP (optionally warm-started from an original-space x0/y0).
        Options other than the server's own tol / time_limit / iteration_limit
        send the LP to the grid engine, which takes per-solve options."""
        if self._closing:
            raise RuntimeError("server is closed")
        if self._error is not None:
            raise RuntimeError(f"scheduler failed: {self._error}")
        if route is None:
            route = "grid" if lp.nnz >= self.large_nonzeros or options else "batch"
        if route not in ("batch", "grid"):
            raise ValueError("route must be 'batch' or 'grid'")
        req = Request(lp, x0, y0, route)
        if route == "grid":
            self._pool.submit(self._grid_solve, req, options)
        else:
            with self._cv:
                self._incoming.append(req)
                self._cv.notify()
        return req

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

    # -------------------------------------------------------------- loop --
    def _loop(self):
        incoming: list[Request] = []
        try:
            while True:
                with self._cv:
                    while not self._incoming and not self._tickets and not self._closing:
                        self._cv.wait()
                    if self._closing and not self._incoming and not self._tickets:
                        return
                    incoming, self._incoming = self._incoming, []
                while incoming:
                    req = incoming.pop(0)
                    try:
                        ticket = self._batch.submit(req.lp, req.x0, req.y0)
                    except (ValueError, RuntimeError) as error:
                        # this request alone is rejected (e.g. a nonfinite warm start)
                        req._complete(error)
                        continue
                    self._tickets[ticket] = req
                if self._tickets:
                    self._batch.step(self.checks_per_launch)
                    for t in self._batch.finished():
                        req = self._tickets.pop(t)
                        req._complete(self._batch.result(t))
                        self._batch.release(t)
        except BaseException as error:  # surface to every waiter
            self._error = error
            with self._cv:
                pending = list(self._tickets.values()) + incoming + self._incoming
                self._tickets.clear()
                self._incoming = []
            for req in pending:
                req._complete(error)

    def _grid_solve(self, req: Request, options: dict):
        need = grid_footprint(req.lp)
        with self._grid_cv:
            # Admit when it fits, or alone when it never could.
            while self._grid_running and self._grid_used + need > self._grid_budget:
                self._grid_cv.wait()
            self._grid_used += need
            self._grid_running += 1
        try:
            opts = dict(tol=self.tol, time_limit=self.time_limit, iteration_limit=self.iteration_limit)
            opts.update(options)
            if req.x0 is not None:
                opts["initial_x"] = req.x0
            if req.y0 is not None:
                opts["initial_y"] = req.y0
            req._complete(solve_lp(req.lp, **opts))
        except BaseException as error:
            req._complete(error)
        finally:
            with self._grid_cv:
                self._grid_used -= need
                self._grid_running -= 1
                self._grid_cv.notify_all()

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify()
        self._thread.join()
        self._pool.shutdown(wait=True)
        self._batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
