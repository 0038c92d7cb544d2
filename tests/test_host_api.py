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


def test_serving_loop_survives_a_bad_request(engine):
    """The native server refuses a request with a nonfinite warm start at
    submit and keeps serving: a good request submitted beside it still
    completes.  Without a GPU (this test's usual setting) that completion is
    an error ("No HIP device is visible") on the batched route and status
    UNSUPPORTED on the grid route; either way every call returns."""
    from niryukti_amd import serve

    lp = small_lp(engine)
    with serve.Server(hbm_budget=1 << 30) as server:
        good = server.submit(lp)
        with pytest.raises(ValueError, match="Nonfinite"):
            server.submit(lp, [np.nan, 0.0], [0.0, 0.0])
        with pytest.raises(ValueError, match="x0"):
            server.submit(lp, [0.0], [0.0, 0.0])
        grid = server.submit(lp, route="grid")
        after = server.submit(lp)
        for req in (good, after):
            try:
                assert req.wait(60)["status"] == "OPTIMAL"
            except RuntimeError as error:  # no device: the batched engine cannot start
                assert engine.device_count() == 0 and "device" in str(error).lower()
        status = grid.wait(60)["status"]
        assert status == ("UNSUPPORTED" if engine.device_count() == 0 else "OPTIMAL")
        assert server.stats()["finished"] == 3
        # and the server still takes requests
        assert server.submit(lp, route="grid").wait(60)["status"] == status


def test_serving_refuses_malformed_requests(engine):
    from niryukti_amd import serve

    lp = small_lp(engine)
    with serve.Server(hbm_budget=1 << 30) as server:
        with pytest.raises(ValueError, match="route"):
            server.submit(lp, route="elsewhere")
        with pytest.raises(ValueError, match="grid route"):
            server._lib.niryukti_server_last_error  # noqa: B018 (bound)
            server.submit(lp, route="batch", method="halpern")
        assert server.solve_request({"model": {}, "extra": 1})[0] == 400
        assert server.solve_request({"model": "x"})[0] == 400
    with pytest.raises(ValueError):
        serve.Server(batch_fraction=1.5)
