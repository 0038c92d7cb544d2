"""GPU parity of the HIP iteration backend against the CPU oracle backend.

Mirrors the reference's differential CPU/GPU tests (tests/test_research.cpp:34-64
one adaptive step, :864-891 anchored trajectories with a reset at step 14,
:892-907 index widths x precisions) at the same 1e-11 relative tolerance for
FP64 values, through the npdhg_* C-ABI.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def near(a, b, tol=1e-11):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.all(np.abs(a - b) <= tol * (1 + np.abs(b))), (a, b)


def model(engine, n, rows, entries, lb=-10.0, ub=10.0):
    A = np.zeros((rows, n))
    for i, j, v in entries:
        A[i, j] += v
    return dict(A=A, c=np.zeros(n), lb=np.full(n, lb), ub=np.full(n, ub), rl=np.full(rows, -np.inf),
                ru=np.full(rows, np.inf))


def to_lp(engine, d):
    return engine.LP.from_dense(d["A"], d["c"], d["lb"], d["ub"], d["rl"], d["ru"])


def test_device_present(engine):
    assert engine.device_count() >= 1
    print(engine.device_name(0))


def test_one_adaptive_step(engine, oracle):
    # tests/test_research.cpp:34-64: the first trial (limit 1/2) is rejected.
    d = model(engine, 1, 1, [(0, 0, 1.0)])
    d["c"][:] = -1
    d["rl"][:] = d["ru"][:] = 3
    lp = to_lp(engine, d)
    cpu = oracle.Backend(lp, [0.0], [0.0], adaptive=True)
    assert cpu.advance(1, 2, 2) == 1
    assert cpu.rejected() > 0
    x, y, ax, ay = cpu.candidates()
    assert 0 < x[0] < 2
    assert (x[0] ** 2 + y[0] ** 2) / x[0] >= 2 * abs(x[0] * y[0])
    for graphs in (False, True):
        gpu = engine.Backend(lp, [0.0], [0.0], adaptive=True, graphs=graphs)
        assert gpu.advance(1, 2, 2) == 1
        assert gpu.status()["rejected_total"] == cpu.rejected()
        gx, gy, gax, gay = gpu.candidates()
        near(gx, x), near(gy, y), near(gax, ax), near(gay, ay)
        gpu.close()


@pytest.mark.parametrize("reflection", [0.0, 1.0])
@pytest.mark.parametrize("graphs", [False, True])
def test_anchored_trajectory(engine, oracle, reflection, graphs):
    # tests/test_research.cpp:864-891
    d = model(engine, 2, 2, [(0, 0, 1), (0, 1, 2), (1, 0, -1), (1, 1, 1)])
    d["c"][:] = [1, -2]
    d["rl"][:] = [-1, -2]
    d["ru"][:] = [3, 2]
    lp = to_lp(engine, d)
    cpu = oracle.Backend(lp, [0, 0], [0, 0], adaptive=False, halpern=True, reflection=reflection)
    gpu = engine.Backend(lp, [0, 0], [0, 0], adaptive=False, halpern=True, reflection=reflection, graphs=graphs,
                         index_bits=32)
    for k in range(30):
        assert cpu.advance(1, .1, .1) == gpu.advance(1, .1, .1)
        x, y, ax, ay = cpu.candidates()
        gx, gy, gax, gay = gpu.candidates()
        near(gx, x), near(gy, y), near(gax, ax), near(gay, ay)
        if k == 14:
            cpu.reset(x, y)
            gpu.reset(gx, gy)
    gpu.close()


@pytest.mark.parametrize("index_bits", [32, 64])
@pytest.mark.parametrize("fp32", [False, True])
def test_fixed_step_chunk(engine, oracle, index_bits, fp32):
    # tests/test_research.cpp:892-907 (the 2x2 matrix is exact in FP32)
    d = model(engine, 2, 2, [(0, 0, 1), (0, 1, 2), (1, 0, -1), (1, 1, 1)])
    d["c"][:] = [1, -2]
    d["rl"][:] = [-1, -2]
    d["ru"][:] = [3, 2]
    lp = to_lp(engine, d)
    cpu = oracle.Backend(lp, [0, 0], [0, 0], adaptive=False)
    gpu = engine.Backend(lp, [0, 0], [0, 0], adaptive=False, index_bits=index_bits, fp32_values=fp32)
    assert cpu.advance(30, .1, .1) == 30
    assert gpu.advance(30, .1, .1) == 30
    x, y, _, _ = cpu.candidates()
    gx, gy, _, _ = gpu.candidates()
    near(gx, x), near(gy, y)
    gpu.close()


def random_csr(rng, rows, cols, long_rows=()):
    ptr = [0]
    idx, val = [], []
    for i in range(rows):
        k = int(rng.integers(0, 9)) if i not in long_rows else min(cols, 5000)
        cs = np.sort(rng.choice(cols, size=k, replace=False))
        idx.extend(cs.tolist())
        val.extend(rng.standard_normal(k).tolist())
        ptr.append(len(idx))
    return np.array(ptr), np.array(idx), np.array(val)


def test_spmv_matches_dense(engine):
    rng = np.random.default_rng(7)
    rows, cols = 700, 6000
    ptr, idx, val = random_csr(rng, rows, cols, long_rows=(3, 400))
    A = np.zeros((rows, cols))
    for i in range(rows):
        A[i, idx[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
    lp = engine.LP(ptr, idx, val, np.zeros(cols), -np.ones(cols), np.ones(cols), -np.ones(rows), np.ones(rows))
    b = engine.Backend(lp, np.zeros(cols), np.zeros(rows))
    x = rng.standard_normal(cols)
    y = rng.standard_normal(rows)
    near(b.spmv(x), A @ x, 1e-12)
    near(b.spmv(y, transpose=True), A.T @ y, 1e-12)
    b.close()


def dense_to_csr(A):
    ptr, idx, val = [0], [], []
    for row in A:
        nz = np.flatnonzero(row)
        idx.extend(nz.tolist())
        val.extend(row[nz].tolist())
        ptr.append(len(idx))
    return np.array(ptr), np.array(idx), np.array(val)


def test_spmv_long_rows_and_columns(engine):
    """Rows and columns far longer than a lane group walks (split into several
    2,048-nonzero segments, one workgroup each) against a dense product, and
    bitwise repeatable across calls."""
    rng = np.random.default_rng(11)
    rows, cols = 2500, 12000
    A = np.zeros((rows, cols))
    for i in range(rows):
        A[i, rng.choice(cols, 6, replace=False)] = rng.standard_normal(6)
    A[5, :] = rng.standard_normal(cols)                               # 6 segments
    A[17, rng.choice(cols, 3000, replace=False)] = rng.standard_normal(3000)  # 2 segments
    A[2499, rng.choice(cols, 300, replace=False)] = 1.0               # 1 segment
    A[:, 42] = rng.standard_normal(rows)                              # long column
    A[rng.random(rows) < 0.3, 7] = -2.0
    ptr, idx, val = dense_to_csr(A)
    lp = engine.LP(ptr, idx, val, np.zeros(cols), -np.ones(cols), np.ones(cols), -np.ones(rows), np.ones(rows))
    for bits in (32, 64):
        b = engine.Backend(lp, np.zeros(cols), np.zeros(rows), index_bits=bits)
        x = rng.standard_normal(cols)
        y = rng.standard_normal(rows)
        ax, aty = b.spmv(x), b.spmv(y, transpose=True)
        near(ax, A @ x, 1e-12)
        near(aty, A.T @ y, 1e-12)
        assert np.array_equal(ax, b.spmv(x)) and np.array_equal(aty, b.spmv(y, transpose=True))
        b.close()


def test_xcd_chunked_products_track_oracle(engine, oracle, monkeypatch):
    """Config 1 at half size with column chunks forced down to 2 MB vectors
    (by default only vectors from 12 MB are chunked): A dx and A^T y run as
    XCD column chunks plus fixed-order partial sums.  Three adaptive chunks
    must track the oracle, and the chunked and direct paths must agree to
    rounding."""
    inst = engine.Instance(1, 11, 2, 0)
    lp = inst.lp
    x0, y0 = np.zeros(lp.cols), np.zeros(lp.rows)
    cpu = oracle.Backend(lp, x0, y0, adaptive=True)
    monkeypatch.setenv("NIRYUKTI_XCD_CHUNK_MIN_MB", "2")
    gpu = engine.Backend(lp, x0, y0, adaptive=True)
    monkeypatch.setenv("NIRYUKTI_XCD_CHUNKS", "0")
    direct = engine.Backend(lp, x0, y0, adaptive=True)
    for _ in range(3):
        assert cpu.advance(10, 0.05, 0.05) == gpu.advance(10, 0.05, 0.05) == direct.advance(10, 0.05, 0.05) == 10
        x, y, xa, ya = cpu.candidates()
        for b in (gpu, direct):
            gx, gy, gxa, gya = b.candidates()
            assert np.abs(gx - x).max() <= 1e-9 * (1 + np.abs(x).max())
            assert np.abs(gy - y).max() <= 1e-9 * (1 + np.abs(y).max())
            assert np.abs(gxa - xa).max() <= 1e-9 * (1 + np.abs(x).max())
            assert np.abs(gya - ya).max() <= 1e-9 * (1 + np.abs(y).max())
    assert gpu.status()["rejected_total"] == cpu.rejected() == direct.status()["rejected_total"]
    gpu.close()
    direct.close()


@pytest.mark.parametrize("case", [(0, 4, 0), (2, 1, 0), (2, 1, 1), (4, 200, 0), (1, 400, 0), (1, 20, 0),
                                  (1, 1, 0), (4, 1, 0)])
def test_adaptive_chunks_track_oracle(engine, oracle, case):
    """Three chunks of 10 adaptive steps on a generated (unscaled) LP, up to
    the full-size instances of configs 1 (1,000,000 x 2,000,000, 12.5M
    nonzeros: hub rows as long-row segments, A dx in XCD column chunks) and 4
    (1,282,000 x 2,560,000, 29.7M nonzeros: the linking rows as segments)."""
    config, scale, index = case
    inst = engine.Instance(config, 11, scale, index)
    lp = inst.lp
    x0, y0 = np.zeros(lp.cols), np.zeros(lp.rows)
    cpu = oracle.Backend(lp, x0, y0, adaptive=True)
    gpu = engine.Backend(lp, x0, y0, adaptive=True)
    for _ in range(3):
        assert cpu.advance(10, 0.05, 0.05) == gpu.advance(10, 0.05, 0.05) == 10
        x, y, xa, ya = cpu.candidates()
        gx, gy, gxa, gya = gpu.candidates()
        scale_x = 1 + np.abs(x).max()
        scale_y = 1 + np.abs(y).max()
        assert np.abs(gx - x).max() <= 1e-9 * scale_x
        assert np.abs(gy - y).max() <= 1e-9 * scale_y
        assert np.abs(gxa - xa).max() <= 1e-9 * scale_x
        assert np.abs(gya - ya).max() <= 1e-9 * scale_y
    assert gpu.status()["rejected_total"] == cpu.rejected()
    gpu.close()
