"""Host-side checks of the Python binding and the serving loop (no GPU).

The native side copies exactly rows/cols/nnz elements from every pointer it
is given, so the binding must refuse arrays of any other length before they
cross the C-ABI; and one bad request must not take the serving loop down.
"""
import numpy as np
import pytest


def small_lp(engine):
    return engine.LP.from_dense([[1.0, 2.0], [0.0, 1.0]], [1, 1], [0, 0], [10, 10], [4, 0], [np.inf, 3])


def test_lp_extents_are_checked(engine):
    lp = small_lp(engine)
    with pytest.raises(ValueError, match="ptr"):
        engine.LP(lp.ptr[:-1], lp.index, lp.value, lp.c, lp.lb, lp.ub, lp.rl, lp.ru)
    with pytest.raises(ValueError, match="index and value"):
        engine.LP(lp.ptr, lp.index[:-1], lp.value, lp.c, lp.lb, lp.ub, lp.rl, lp.ru)
    with pytest.raises(ValueError, match="index and value"):
        engine.LP(lp.ptr, lp.index, lp.value[:-1], lp.c, lp.lb, lp.ub, lp.rl, lp.ru)
    with pytest.raises(ValueError, match="lb and ub"):
        engine.LP(lp.ptr, lp.index, lp.value, lp.c, lp.lb[:1], lp.ub, lp.rl, lp.ru)
    with pytest.raises(ValueError, match="lb and ub"):
        engine.LP(lp.ptr, lp.index, lp.value, lp.c, lp.lb, np.append(lp.ub, 1), lp.rl, lp.ru)
    with pytest.raises(ValueError, match="ru"):
        engine.LP(lp.ptr, lp.index, lp.value, lp.c, lp.lb, lp.ub, lp.rl, lp.ru[:1])
    with pytest.raises(ValueError, match="nondecreasing"):
        engine.LP(np.array([0, 2, 1]), lp.index[:1], lp.value[:1], lp.c, lp.lb, lp.ub, lp.rl, lp.ru)


def test_warm_start_lengths_are_checked(engine):
    lp = small_lp(engine)
    with pytest.raises(ValueError, match="initial_x"):
        engine.solve_lp(lp, initial_x=[0.0])
    with pytest.raises(ValueError, match="initial_y"):
        engine.solve_lp(lp, initial_y=[0.0, 0.0, 0.0])
    with pytest.raises(ValueError, match="initial_x"):
        engine.prepare_lp(lp, initial_x=np.zeros(3))
    from niryukti_amd import _warm_array

    assert _warm_array([1, 2], 2, "x0").dtype == np.float64
    with pytest.raises(ValueError, match="x0"):
        _warm_array([1, 2, 3], 2, "x0")


class FakeBatch:
    """Stands in for the native Batch: refuses a nonfinite warm start at
    submit (as niryukti_batch_submit does) and finishes every other LP at
    the next step."""

    def __init__(self, **_):
        self.next, self.pending, self.done = 0, [], {}

    def submit(self, lp, x0=None, y0=None):
        for v in (x0, y0):
            if v is not None and not np.all(np.isfinite(v)):
                raise RuntimeError("Nonfinite warm start")
        self.next += 1
        self.pending.append(self.next)
        return self.next

    def step(self, checks):
        for t in self.pending:
            self.done[t] = {"status": "OPTIMAL", "ticket": t}
        self.pending = []
        return 0

    def finished(self):
        out = sorted(self.done)
        return out

    def result(self, t):
        return self.done[t]

    def release(self, t):
        del self.done[t]

    def close(self):
        pass


def test_serving_loop_survives_a_bad_request(engine, monkeypatch):
    from niryukti_amd import serve

    monkeypatch.setattr(serve, "Batch", FakeBatch)
    lp = small_lp(engine)
    with serve.Server(hbm_budget=1 << 30) as server:
        good = server.submit(lp)
        bad = server.submit(lp, [np.nan, 0.0], [0.0, 0.0])
        after = server.submit(lp)
        assert good.wait(10)["status"] == "OPTIMAL"
        with pytest.raises(RuntimeError, match="Nonfinite"):
            bad.wait(10)
        assert after.wait(10)["status"] == "OPTIMAL"
        # the scheduler is still alive for later requests
        assert server.submit(lp).wait(10)["status"] == "OPTIMAL"


def test_serving_loop_fatal_error_completes_every_waiter(engine, monkeypatch):
    from niryukti_amd import serve

    class Broken(FakeBatch):
        def step(self, checks):
            raise MemoryError("device lost")

    monkeypatch.setattr(serve, "Batch", Broken)
    lp = small_lp(engine)
    server = serve.Server(hbm_budget=1 << 30)
    reqs = [server.submit(lp) for _ in range(3)]
    for r in reqs:
        with pytest.raises(MemoryError):
            r.wait(10)
    with pytest.raises(RuntimeError, match="scheduler failed"):
        server.submit(lp)
    server.close()
