"""Factor once, solve many: the factored solves against the full solves and
the oracles on the same seeded inputs.

FP64: pcrp_solve_factored_f64 must equal the oracle's FP64 PCR-Penta
(kernel.cu:32-154 restated) bit for bit (signed zeros compare equal), for
every right-hand side solved against one factor. FP32+refine: the cached
factor gives the same x and sweep counts as pcrp_solve_f32r_f64 (same
operations) and stays within 1e-12 of dgbsv, normwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C, I = 0, 1


def _np(t):
    return t.detach().cpu().numpy()


def _say(*args):
    print("[case]", *args, flush=True)


def _dev(arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a).copy()).cuda() for a in arrs]


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 63, 64, 65, 100, 128, 255, 256, 257, 511, 512, 513,
                               777, 1000, 1023, 1024])
def test_factored_f64_bit_exact(gpu, oracle_lib, n):
    """Per-system factors, both layouts, two right-hand sides per factor."""
    import torch
    m = 19
    for kind, layout in ((0, I), (1, C)):
        _say("factored-f64", n, m, kind, layout)
        arrs = oracle_lib.gen_batch(kind, 400 + n, n, m, layout)
        dev = _dev(arrs)
        F = gpu.factor(*dev[:5], n, m, layout)
        for rep in range(2):
            rhs = arrs[5] if rep == 0 else np.sin(np.arange(n * m) * 0.37 + n)
            y = torch.from_numpy(rhs.copy()).cuda()
            gpu.solve_factored(F, y, n, m, layout)
            torch.cuda.synchronize()
            ref = oracle_lib.pcr_batch(*arrs[:5], rhs, n, m, layout)
            assert np.array_equal(_np(y), ref), (n, kind, layout, rep)
        for t, h in zip(dev[:5], arrs[:5]):   # a..e untouched
            assert np.array_equal(_np(t), h)


def test_factored_f64_equals_full_solve_zero_pivots(gpu, oracle_lib):
    """Exact zero pivots (kernel.cu:122,129: the reused multiplier) are
    recorded in the factor: the factored solve equals the full solve."""
    import torch
    n, m = 16, 4
    _say("factored-zero-pivot", n, m)
    arrs = [a.copy() for a in oracle_lib.gen_batch(0, 2, n, m, C)]
    j0 = n
    arrs[0][j0:j0 + n] = 0.0
    arrs[4][j0:j0 + n] = 0.0
    arrs[2][j0 + 5] = 0.0
    ref = oracle_lib.pcr_batch(*arrs, n, m, C)
    dev = _dev(arrs)
    F = gpu.factor(*dev[:5], n, m, C)
    gpu.solve_factored(F, dev[5], n, m, C)
    torch.cuda.synchronize()
    assert np.array_equal(np.nan_to_num(_np(dev[5]), nan=7.0), np.nan_to_num(ref, nan=7.0))


@pytest.mark.parametrize("n,m", [(512, 1024), (512, 600), (1024, 37), (300, 2048), (5, 7)])
@pytest.mark.parametrize("layout", [I, C])
def test_factored_f64_shared_matrix(gpu, oracle_lib, n, m, layout):
    """One matrix for every right-hand side (the Heston v-lines' case): equal
    to the oracle on the matrix repeated m times."""
    import torch
    _say("factored-shared", n, m, layout)
    one = oracle_lib.gen_batch(0, 70 + n, n, 1, C)
    rhs = oracle_lib.gen_batch(1, 71 + n, n, m, layout)[5]
    if layout == C:
        rep = [np.tile(v, m) for v in one[:5]]
    else:
        rep = [np.repeat(v, m) for v in one[:5]]
    ref = oracle_lib.pcr_batch(*rep, rhs, n, m, layout)
    dev = _dev(one[:5])
    F = gpu.factor(*dev, n, m, layout, shared=True)
    assert F.numel() == gpu.factor_size(n, m, gpu.PREC_F64, True)
    y = torch.from_numpy(rhs.copy()).cuda()
    gpu.solve_factored(F, y, n, m, layout, shared=True)
    torch.cuda.synchronize()
    assert np.array_equal(_np(y), ref), (n, m, layout)


@pytest.mark.parametrize("n", [1, 2, 3, 33, 100, 256, 512, 700, 1024, 1500, 2048])
def test_factored_f32r_matches_full_refinement(gpu, oracle_lib, n):
    """The cached FP32 factor: same x and sweeps as the full FP32+refine solve
    (contiguous layout, one system per workgroup in both), every system within
    1e-12 of dgbsv, and a second right-hand side against the same factor."""
    import torch
    m = 24
    for kind in (0, 1):
        _say("factored-f32r", n, m, kind)
        arrs = oracle_lib.gen_batch(kind, 500 + n, n, m, C)
        dev = _dev(arrs)
        sw0 = torch.full((m,), -1, dtype=torch.int32, device="cuda")
        y0 = dev[5].clone()
        gpu.solve_refined(*dev[:5], y0, n, m, C, max_sweeps=8, tol=1e-12, sweeps=sw0)
        F = gpu.factor(*dev[:5], n, m, C, precision=gpu.PREC_F32_REFINE)
        for rep in range(2):
            rhs = arrs[5] if rep == 0 else np.cos(np.arange(n * m) * 0.11 + n)
            y = torch.from_numpy(rhs.copy()).cuda()
            sw = torch.full((m,), -1, dtype=torch.int32, device="cuda")
            gpu.solve_factored_refined(F, *dev[:5], y, n, m, C, max_sweeps=8, tol=1e-12, sweeps=sw)
            torch.cuda.synchronize()
            x = _np(y)
            if rep == 0:
                assert np.array_equal(x, _np(y0)) and np.array_equal(_np(sw), _np(sw0)), (n, kind)
            ref = oracle_lib.gbsv_batch(*arrs[:5], rhs, n, m, C)
            err = oracle_lib.normwise_relerr(x, ref, n, m, C)
            assert err <= 1e-12, (n, kind, rep, err)
            assert _np(sw).min() >= 1 and _np(sw).max() <= 3


@pytest.mark.parametrize("n,m,layout", [(512, 1024, I), (1024, 512, C), (200, 66, I), (2048, 8, C)])
def test_factored_f32r_shared_matrix(gpu, oracle_lib, n, m, layout):
    """One shared matrix, m right-hand sides (interleaved pairs load 16 bytes
    at a time): every system within 1e-12 of dgbsv."""
    import torch
    _say("factored-f32r-shared", n, m, layout)
    one = oracle_lib.gen_batch(1, 90 + n, n, 1, C)
    rhs = oracle_lib.gen_batch(0, 91 + n, n, m, layout)[5]
    rep = [np.tile(v, m) for v in one[:5]] if layout == C else [np.repeat(v, m) for v in one[:5]]
    ref = oracle_lib.gbsv_batch(*rep, rhs, n, m, layout)
    dev = _dev(one[:5])
    F = gpu.factor(*dev, n, m, layout, precision=gpu.PREC_F32_REFINE, shared=True)
    y = torch.from_numpy(rhs.copy()).cuda()
    sw = torch.full((m,), -1, dtype=torch.int32, device="cuda")
    gpu.solve_factored_refined(F, *dev, y, n, m, layout, shared=True, max_sweeps=8, tol=1e-12,
                               sweeps=sw)
    torch.cuda.synchronize()
    err = oracle_lib.normwise_relerr(_np(y), ref, n, m, layout)
    print("shared f32r", n, m, layout, "err", err, "sweeps", np.bincount(_np(sw)), flush=True)
    assert err <= 1e-12
    assert _np(sw).min() >= 1 and _np(sw).max() <= 3


@pytest.mark.parametrize("n", [1, 2, 3, 64, 65, 100, 128, 129, 200, 256, 257, 333, 511, 512])
def test_factored_f32r_shared_grouped_equals_single(gpu, oracle_lib, n):
    """Interleaved right-hand sides of one shared matrix, m even and 16-byte
    aligned (pairs of systems per workgroup, 16-byte loads): the same x and
    sweep counts as one system per workgroup (the same y at an 8-byte offset,
    which that kernel takes), bit for bit: with the default stopping rule,
    with tol = 0 (systems run max_sweeps and are flagged), and with a NaN in
    one system."""
    import torch
    m = 300   # more systems than CUs: the grouped loader's condition
    one = oracle_lib.gen_batch(1, 30 + n, n, 1, C)
    dev = _dev(one[:5])
    F = gpu.factor(*dev, n, m, I, precision=gpu.PREC_F32_REFINE, shared=True)
    for case, (tol, ms) in enumerate(((1e-12, 8), (0.0, 3), (1e-12, 8))):
        _say("f32r-shared-grouped", n, m, case)
        rhs = oracle_lib.gen_batch(0, 31 + n + case, n, m, I)[5].copy()
        if case == 2:
            rhs[(n // 2) * m + 5] = np.nan
        out = []
        for off in (0, 1):
            buf = torch.zeros(n * m + 2, dtype=torch.float64, device="cuda")
            y = buf[off:off + n * m]
            y.copy_(torch.from_numpy(rhs))
            assert (y.data_ptr() % 16 == 0) == (off == 0)
            sw = torch.full((m,), -1, dtype=torch.int32, device="cuda")
            gpu.solve_factored_refined(F, *dev, y, n, m, I, shared=True, max_sweeps=ms, tol=tol,
                                       sweeps=sw)
            torch.cuda.synchronize()
            out.append((_np(y).copy(), _np(sw).copy()))
        (x0, s0), (x1, s1) = out
        assert np.array_equal(np.nan_to_num(x0, nan=7.0), np.nan_to_num(x1, nan=7.0)), (n, case)
        assert np.array_equal(s0, s1), (n, case)
        if case == 1:   # tol = 0: only an exactly zero correction stops a system early
            assert (s0 <= ms + 1).all() and (n < 64 or (s0 == ms + 1).any())
        if case == 2:
            assert s0[5] == 9 and (np.delete(s0, 5) <= 3).all()
        if case == 0:
            rep = [np.repeat(v, m) for v in one[:5]]
            ref = oracle_lib.gbsv_batch(*rep, rhs, n, m, I)
            assert oracle_lib.normwise_relerr(x0, ref, n, m, I) <= 1e-12
