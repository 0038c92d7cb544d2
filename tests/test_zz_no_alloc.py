"""The no-allocation device prepare (npdhg_scale_lp_noalloc) against npdhg_scale_lp.

The entry point takes device arrays in and out and keeps its work arrays in a
static device array, so it must give bit-identical results to the allocating
entry point on consecutive calls (the second smaller than the first, so that
stale scratch from the first call would show), and refuse an LP that needs
more scratch than the array holds.  Device memory for the inputs and outputs
comes from torch (plumbing only).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _p(t, kind):
    return C.cast(C.c_void_p(t.data_ptr()), kind)


def scale_noalloc(engine, lp, passes):
    lib = engine.load_library()
    m, n, nnz = lp.rows, lp.cols, lp.nnz
    src = {k: _dev(getattr(lp, k)) for k in ("ptr", "index", "value", "c", "lb", "ub", "rl", "ru")}
    f64 = dict(dtype=torch.float64, device="cuda")
    out = {k: torch.full((max(size, 1),), np.nan, **f64) for k, size in
           (("value", nnz), ("c", n), ("lb", n), ("ub", n), ("rl", m), ("ru", m), ("row_scale", m),
            ("column_scale", n), ("t_value", nnz))}
    out["t_ptr"] = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    out["t_index"] = torch.full((max(nnz, 1),), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    s = engine._PdhgLP(m, n, _p(src["ptr"], engine._i64p), _p(src["index"], engine._i64p),
                       *(_p(src[k], engine._f64p) for k in ("value", "c", "lb", "ub", "rl", "ru")))
    r = engine._Scaled(*(_p(out[k], engine._f64p) for k in ("value", "c", "lb", "ub", "rl", "ru", "row_scale",
                                                              "column_scale")), 0.0,
                       _p(out["t_ptr"], engine._i64p), _p(out["t_index"], engine._i64p),
                       _p(out["t_value"], engine._f64p), -1.0, -1.0)
    rc = lib.npdhg_scale_lp_noalloc(C.byref(s), int(passes), C.byref(r), None)
    if rc != 0:
        raise RuntimeError(lib.npdhg_scale_last_error().decode())
    got = {k: v.cpu().numpy() for k, v in out.items()}
    sizes = {"value": nnz, "c": n, "lb": n, "ub": n, "rl": m, "ru": m, "row_scale": m, "column_scale": n,
             "t_value": nnz, "t_index": nnz, "t_ptr": n + 1}
    got = {k: v[:sizes[k]] for k, v in got.items()}
    got["objective_scale"] = r.objective_scale
    return got


KEYS = ("value", "c", "lb", "ub", "rl", "ru", "row_scale", "column_scale", "t_ptr", "t_index", "t_value")


def test_noalloc_prepare_matches_allocating_entry_point(engine):
    cases = [((1, 42, 400, 0), 5), ((0, 7, 4, 0), 5), ((2, 42, 1, 3), 5), ((2, 42, 1, 3), 0)]
    for case, passes in cases:  # consecutive calls, each smaller than the one before
        lp = engine.Instance(*case).lp
        print("no-alloc prepare", case, passes, lp.rows, lp.cols, lp.nnz, flush=True)
        want = engine.scale_lp(lp, passes)
        got = scale_noalloc(engine, lp, passes)
        for key in KEYS:
            assert np.array_equal(got[key], want[key]), (case, key)
        assert got["objective_scale"] == want["objective_scale"]


def test_noalloc_prepare_refuses_what_it_cannot_hold(engine):
    lib = engine.load_library()
    # offsets that claim 10M nonzeros: 240 MB of work arrays, more than the
    # 128 MiB scratch; refused after reading the two end offsets only
    ptr = _dev(np.array([0, 10_000_000], dtype=np.int64))
    small = _dev(np.zeros(4))
    idx = _dev(np.zeros(4, dtype=np.int64))
    torch.cuda.synchronize()
    f = lambda t: _p(t, engine._f64p)
    s = engine._PdhgLP(1, 2, _p(ptr, engine._i64p), _p(idx, engine._i64p), f(small), f(small), f(small), f(small),
                       f(small), f(small))
    out_d = torch.zeros(8, dtype=torch.float64, device="cuda")
    out_i = torch.zeros(8, dtype=torch.int64, device="cuda")
    r = engine._Scaled(*([f(out_d)] * 8), 0.0, _p(out_i, engine._i64p), _p(out_i, engine._i64p), f(out_d), 0.0, 0.0)
    assert lib.npdhg_scale_lp_noalloc(C.byref(s), 5, C.byref(r), None) == -1
    assert "scratch" in lib.npdhg_scale_last_error().decode()
    # a column index outside the matrix is caught by the device validation
    bad_ptr = _dev(np.array([0, 2], dtype=np.int64))
    bad_idx = _dev(np.array([0, 5], dtype=np.int64))
    vals = _dev(np.array([1.0, 2.0]))
    torch.cuda.synchronize()
    s = engine._PdhgLP(1, 2, _p(bad_ptr, engine._i64p), _p(bad_idx, engine._i64p), f(vals), f(small), f(small),
                       f(small), f(small), f(small))
    assert lib.npdhg_scale_lp_noalloc(C.byref(s), 5, C.byref(r), None) == -1
    assert "out of range" in lib.npdhg_scale_last_error().decode()
    # and a valid call afterwards still matches the allocating entry point
    lp = engine.Instance(2, 42, 1, 5).lp
    want, got = engine.scale_lp(lp, 5), scale_noalloc(engine, lp, 5)
    for key in KEYS:
        assert np.array_equal(got[key], want[key]), key
