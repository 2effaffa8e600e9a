"""Plan / execute: a plan's executes equal the one-call entry points bit for
bit (FP64 against the oracle as well; FP32+refine within 1e-12 of dgbsv), on
every route (LDS, split, window, uneven queue, uneven stage), and executes
captured in a graph and replayed with new right-hand sides equal eager
executes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C, I = 0, 1
F64, F32R = 0, 1


def _np(t):
    return t.detach().cpu().numpy()


def _say(*args):
    print("[case]", *args, flush=True)


def _slow_decay_rows(n, seed):
    """c = 2.05, b = d = -1, a = e = 0: block multipliers that stay above the
    truncation threshold past a window's halo, so the batch is not local."""
    rng = np.random.default_rng(seed)
    a = np.zeros(n); e = np.zeros(n)
    b = -np.ones(n); d = -np.ones(n); c = np.full(n, 2.05)
    b[0] = 0.0; d[-1] = 0.0
    return [a, b, c, d, e, rng.uniform(-1, 1, n)]


MIXED = [1, 2, 3, 5, 64, 100, 1000, 2047, 2048, 2049, 4097, 5000, 40000]

UNIFORM = [
    # name, precision, generator, n, m, layout, route, slow-decay system
    ("f64-lds", F64, 0, 700, 33, C, 0, False),
    ("f64-split", F64, 1, 5000, 8, I, 1, False),
    ("f32r-lds", F32R, 0, 1000, 50, I, 0, False),
    ("f32r-window-n2pow20", F32R, 1, 1 << 20, 4, I, 2, False),
    ("f32r-window-contiguous", F32R, 0, 40000, 3, C, 2, False),
    ("f32r-window-n1500", F32R, 0, 1500, 6000, I, 2, False),
    ("f32r-split-slow", F32R, 0, 6000, 3, I, 1, True),
]


def _uniform_batch(gpu, kind, n, m, layout, slow):
    import torch
    dev = gpu.generate(kind, 60 + n % 97, n, m, layout)
    if slow:
        rows = _slow_decay_rows(n, 5)
        for t, v in zip(dev, rows):
            (t.view(n, m)[:, 1] if layout == I else t.view(m, n)[1, :]).copy_(torch.from_numpy(v))
    return dev


def _one_call(gpu, prec, dev, n, m, layout, y):
    import torch
    sw = torch.full((m,), -1, dtype=torch.int32, device="cuda")
    if prec == F64:
        gpu.solve(*dev[:5], y, n, m, layout)
    else:
        gpu.solve_refined(*dev[:5], y, n, m, layout, max_sweeps=8, tol=1e-12, sweeps=sw)
    torch.cuda.synchronize()
    return _np(y), _np(sw)


@pytest.mark.parametrize("case", UNIFORM, ids=[c[0] for c in UNIFORM])
def test_plan_uniform_equals_one_call(gpu, oracle_lib, case):
    import torch
    name, prec, kind, n, m, layout, route, slow = case
    _say("plan", name, n, m, layout)
    dev = _uniform_batch(gpu, kind, n, m, layout, slow)
    ins = [_np(t).copy() for t in dev]
    plan = gpu.Plan(*dev[:5], n, m, layout, precision=prec)
    assert plan.route == route, (name, plan.route)
    for rep in range(2):
        rhs = ins[5] if rep == 0 else np.sin(np.arange(n * m) * 0.3 + rep)
        y1 = torch.from_numpy(rhs.copy()).cuda()
        x1, s1 = _one_call(gpu, prec, dev, n, m, layout, y1)
        y2 = torch.from_numpy(rhs.copy()).cuda()
        sw = torch.full((m,), -1, dtype=torch.int32, device="cuda")
        plan.execute(y2, sweeps=sw)
        torch.cuda.synchronize()
        x2 = _np(y2)
        assert np.array_equal(x1, x2), (name, rep)
        if prec == F64:
            assert np.array_equal(x2, oracle_lib.pcr_batch(*ins[:5], rhs, n, m, layout)), name
        else:
            assert np.array_equal(_np(sw), s1), (name, rep)
            ref = oracle_lib.gbsv_batch(*ins[:5], rhs, n, m, layout)
            assert oracle_lib.normwise_relerr(x2, ref, n, m, layout) <= 1e-12, name
    plan.destroy()


@pytest.mark.parametrize("prec,slow", [(F64, False), (F32R, False), (F32R, True)])
def test_plan_uneven_equals_one_call(gpu, oracle_lib, prec, slow):
    import torch
    off, nn, total = gpu.uneven_layout(MIXED)
    _say("plan-uneven", prec, slow)
    dev = gpu.generate_uneven(0, 17, off, nn)
    if slow:
        rows = _slow_decay_rows(40000, 6)
        sl = slice(int(off[-1]), int(off[-1]) + 40000)
        for t, v in zip(dev, rows):
            t[sl].copy_(torch.from_numpy(v))
    ins = [_np(t).copy() for t in dev]
    plan = gpu.Plan(*dev[:5], precision=prec, offsets=off, sizes=nn)
    assert plan.route == (3 if prec == F64 else (4 if slow else 2)), plan.route
    for rep in range(2):
        rhs = ins[5] if rep == 0 else np.cos(np.arange(total) * 0.7 + rep)
        y1 = torch.from_numpy(rhs.copy()).cuda()
        s1 = torch.full((len(nn),), -1, dtype=torch.int32, device="cuda")
        if prec == F64:
            gpu.solve_uneven(*dev[:5], y1, off, nn)
        else:
            gpu.solve_uneven_refined(*dev[:5], y1, off, nn, max_sweeps=8, tol=1e-12, sweeps=s1)
        y2 = torch.from_numpy(rhs.copy()).cuda()
        s2 = torch.full((len(nn),), -1, dtype=torch.int32, device="cuda")
        plan.execute(y2, sweeps=s2)
        torch.cuda.synchronize()
        x1, x2 = _np(y1), _np(y2)
        assert np.array_equal(x1, x2), (prec, slow, rep)
        if prec == F64:
            assert np.array_equal(x2, oracle_lib.pcr_uneven(*ins[:5], rhs, off, nn))
        else:
            assert np.array_equal(_np(s1), _np(s2))
            ref = oracle_lib.gbsv_uneven(*ins[:5], rhs, off, nn)
            assert oracle_lib.uneven_relerr(x2, ref, off, nn).max() <= 1e-12


GRAPHED = [
    ("uneven-f64", F64, None),
    ("uneven-f32r", F32R, None),
    ("f32r-window-n2pow20", F32R, (1 << 20, 4, I, 1)),
    ("f32r-split-slow", F32R, (6000, 3, I, 0)),
    ("f64-split", F64, (5000, 8, I, 1)),
]


@pytest.mark.parametrize("case", GRAPHED, ids=[c[0] for c in GRAPHED])
def test_plan_execute_under_graph_capture(gpu, oracle_lib, case):
    """One execute captured in a graph, replayed three times with a new
    right-hand side copied in before each replay (eager work between the
    replays): every replay equals an eager execute of the same right-hand side."""
    import torch
    name, prec, shape = case
    _say("plan-graph", name)
    if shape is None:
        off, nn, total = gpu.uneven_layout(MIXED)
        dev = gpu.generate_uneven(0, 23, off, nn)
        plan = gpu.Plan(*dev[:5], precision=prec, offsets=off, sizes=nn)
        m = len(nn)
    else:
        n, m, layout, kind = shape
        dev = _uniform_batch(gpu, kind, n, m, layout, name.endswith("slow"))
        total = n * m
        plan = gpu.Plan(*dev[:5], n, m, layout, precision=prec)
    s = torch.cuda.Stream()
    y_cap = torch.zeros(total, dtype=torch.float64, device="cuda")
    sw_cap = torch.full((m,), -1, dtype=torch.int32, device="cuda")
    with torch.cuda.stream(s):
        plan.execute(y_cap, sweeps=sw_cap, stream=s)   # warm-up, outside the capture
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.execute(y_cap, sweeps=sw_cap, stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        rhs = np.sin(np.arange(total) * (0.1 + 0.2 * rep)) + rep
        y_eager = torch.from_numpy(rhs.copy()).cuda()
        sw_eager = torch.full((m,), -1, dtype=torch.int32, device="cuda")
        plan.execute(y_eager, sweeps=sw_eager)
        torch.cuda.synchronize()
        y_cap.copy_(torch.from_numpy(rhs))
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(_np(y_cap), _np(y_eager)), (name, rep)
        if prec == F32R:
            assert np.array_equal(_np(sw_cap), _np(sw_eager)), (name, rep)
    plan.destroy()
