"""GPU parity: the HIP kernels against the CPU oracle on the same seeded inputs.

Plain FP64 must equal the oracle's FP64 PCR-Penta (kernel.cu:32-154 restated)
bit for bit (signed zeros compare equal); the FP32+refine form must reach
normwise relative error <= 1e-12 against the FP64 banded-LU (dgbsv) oracle.
Each test prints its case before touching the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C, I = 0, 1  # layouts


def _np(t):
    return t.detach().cpu().numpy()


def _say(*args):
    print("[case]", *args, flush=True)


def _run_f64(gpu, arrs_dev, n, m, layout):
    a, b, c, d, e, y = arrs_dev
    gpu.solve(a, b, c, d, e, y, n, m, layout)
    import torch
    torch.cuda.synchronize()
    return _np(y)


def test_generator_matches_oracle(gpu, oracle_lib):
    for kind, seed, n, m, layout in ((0, 0, 512, 256, I), (1, 5, 1000, 3, C), (0, 1, 37, 129, I)):
        _say("generate", kind, seed, n, m, layout)
        dev = gpu.generate(kind, seed, n, m, layout)
        host = oracle_lib.gen_batch(kind, seed, n, m, layout)
        for u, v in zip(dev, host):
            assert np.array_equal(_np(u), v)


@pytest.mark.parametrize("layout", [I, C])
def test_dominant_m256_n512_bit_exact(gpu, oracle_lib, layout):
    """m=256 systems of n=512, seed 0, strictly dominant."""
    n, m = 512, 256
    _say("m256-n512", layout)
    dev = gpu.generate(0, 0, n, m, layout)
    ins = [_np(t).copy() for t in dev]
    x = _run_f64(gpu, dev, n, m, layout)
    ref = oracle_lib.pcr_batch(*ins, n, m, layout)
    assert np.array_equal(x, ref)
    # a..e untouched
    for t, h in zip(dev[:5], ins[:5]):
        assert np.array_equal(_np(t), h)
    # and within FP64 PCR noise of the banded-LU solution
    xg = oracle_lib.gbsv_batch(*ins, n, m, layout)
    assert oracle_lib.normwise_relerr(x, xg, n, m, layout) < 1e-10


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 31, 33, 63, 64, 65, 100, 127, 128, 129,
                               255, 256, 257, 511, 512, 513, 777, 1000, 1023, 1024])
def test_lds_sizes_bit_exact(gpu, oracle_lib, n):
    m = 37
    for kind, layout in ((0, I), (1, C)):
        _say("lds", n, m, kind, layout)
        dev = gpu.generate(kind, 100 + n, n, m, layout)
        ins = [_np(t).copy() for t in dev]
        x = _run_f64(gpu, dev, n, m, layout)
        ref = oracle_lib.pcr_batch(*ins, n, m, layout)
        assert np.array_equal(x, ref), (n, kind, layout)


def test_reference_t2_system(gpu, oracle_lib):
    """The reference's own input (kernel.cu:5-26), incl. its out-of-range b_0 = d_{n-1} = 1."""
    for n in (8, 64, 256, 512, 1000, 1024):
        m = 3
        _say("t2", n)
        dev = gpu.init_t2(n, m, I)
        host = oracle_lib.init_t2(n, m, I)
        for u, v in zip(dev, host):
            assert np.array_equal(_np(u), v)
        x = _run_f64(gpu, dev, n, m, I)
        assert np.array_equal(x, oracle_lib.pcr_batch(*host, n, m, I))


def test_layouts_agree_bit_for_bit(gpu, oracle_lib):
    n, m = 300, 20
    _say("layouts", n, m)
    xi = _run_f64(gpu, gpu.generate(0, 3, n, m, I), n, m, I).reshape(n, m).T
    xc = _run_f64(gpu, gpu.generate(0, 3, n, m, C), n, m, C).reshape(m, n)
    assert np.array_equal(xi, xc)


def test_zero_pivot_rule(gpu, oracle_lib):
    """Exact zero pivots (c1 == 0 / c2 == 0, kernel.cu:122,129) take the
    multiplier the reference would reuse; results match the oracle."""
    n, m = 16, 4
    _say("zero-pivot", n, m)
    import torch
    arrs = oracle_lib.gen_batch(0, 2, n, m, C)
    # make the first pair equation's pivot exactly zero in system 1: c'_j = c_j + al*d_{j-1} + ar*b_{j+1}
    for k in range(6):
        arrs[k] = arrs[k].copy()
    j0 = 1 * n
    arrs[0][j0:j0 + n] = 0.0
    arrs[4][j0:j0 + n] = 0.0     # tridiagonal system: E(j).C = c_j when a=e=0
    arrs[2][j0 + 5] = 0.0        # zero main diagonal entry -> zero pivot at step 0
    ref = oracle_lib.pcr_batch(*arrs, n, m, C)
    try:
        for v in gpu.variants():   # every form, incl. the in-place form's fallback step
            gpu.tune_set(n, v["id"])
            dev = [torch.from_numpy(a).cuda() for a in arrs]
            x = _run_f64(gpu, dev, n, m, C)
            assert np.array_equal(np.nan_to_num(x, nan=7.0), np.nan_to_num(ref, nan=7.0)), v
    finally:
        gpu.tune_set(n, -1)


def test_every_lds_variant_bit_exact(gpu, oracle_lib):
    """Every compiled LDS variant (shared/redundant pair equations, NT threads
    x RPT rows, register-resident large strides) gives the oracle's bits."""
    import torch
    vs = gpu.variants()
    assert len(vs) >= 10
    try:
        for n, layout in ((1, I), (5, C), (64, I), (100, I), (255, C), (256, I), (300, I),
                          (512, I), (513, C), (777, I), (1024, I), (1024, C)):
            m = 21
            _say("variants", n, layout)
            dev = gpu.generate(0, 900 + n, n, m, layout)
            ins = [_np(t).copy() for t in dev]
            ref = oracle_lib.pcr_batch(*ins, n, m, layout)
            for v in vs:
                if v["max_n"] < n:
                    continue
                gpu.tune_set(n, v["id"])
                assert gpu.tune_get(n) == v["id"]
                y = torch.from_numpy(ins[5]).cuda()
                gpu.solve(*dev[:5], y, n, m, layout)
                torch.cuda.synchronize()
                assert np.array_equal(_np(y), ref), (n, layout, v)
    finally:
        for n in (1, 5, 64, 100, 255, 256, 300, 512, 513, 777, 1024):
            gpu.tune_set(n, -1)


def test_spaced_lds_variants_bit_exact(gpu, oracle_lib):
    """The spaced variants (a thread's rows two strides apart, each pair
    equation serving two of them, rows re-laid out in LDS every step) give the
    oracle's bits at every size from 1 to 70 and around every power of two up
    to 1024 (chains of every length, ragged last blocks, rows past n in the
    last blocks), in both layouts and for even and odd batches (the paired
    kernel and its one-system fallback)."""
    import torch
    sv = [v for v in gpu.variants() if v["kind"].endswith("spaced")]
    assert len(sv) >= 2
    sizes = sorted(set(list(range(1, 71)) + [95, 96, 97, 127, 128, 129, 191, 255, 256, 257, 383,
                                               511, 512, 513, 640, 767, 768, 769, 1000, 1023, 1024]))
    try:
        for n in sizes:
            for m, layout in ((10, I), (9, C)):
                _say("spaced", n, m, layout)
                dev = gpu.generate(0, 500 + n, n, m, layout)
                ins = [_np(t).copy() for t in dev]
                ref = oracle_lib.pcr_batch(*ins, n, m, layout)
                for v in sv:
                    if v["max_n"] < n:
                        continue
                    gpu.tune_set(n, v["id"])
                    y = torch.from_numpy(ins[5]).cuda()
                    gpu.solve(*dev[:5], y, n, m, layout)
                    torch.cuda.synchronize()
                    assert np.array_equal(_np(y), ref), (n, m, layout, v)
    finally:
        for n in sizes:
            gpu.tune_set(n, -1)


def test_paired_lds_variants_bit_exact(gpu, oracle_lib):
    """The paired variants (two neighbouring interleaved systems per workgroup,
    16-byte loads and stores) give the oracle's bits: even m on interleaved
    batches (the paired kernel), and odd m, the contiguous layout and a
    misaligned right-hand side (the one-system kernel they fall back to)."""
    import torch
    pv = [v for v in gpu.variants() if v["kind"].startswith("paired-")]
    assert pv
    sizes = (1, 2, 3, 100, 255, 256, 257, 511, 512, 513, 777, 1023, 1024)
    try:
        for n in sizes:
            for m, layout, off in ((22, I, 0), (64, I, 0), (21, I, 0), (22, C, 0), (22, I, 1)):
                _say("paired", n, m, layout, off)
                dev = gpu.generate(0, 300 + n + m, n, m, layout)
                ins = [_np(t).copy() for t in dev]
                ref = oracle_lib.pcr_batch(*ins, n, m, layout)
                for v in pv:
                    if v["max_n"] < n:
                        continue
                    gpu.tune_set(n, v["id"])
                    buf = torch.zeros(n * m + 2, dtype=torch.float64, device="cuda")
                    y = buf[off:off + n * m]
                    y.copy_(torch.from_numpy(ins[5]))
                    gpu.solve(*dev[:5], y, n, m, layout)
                    torch.cuda.synchronize()
                    assert np.array_equal(_np(y), ref), (n, m, layout, off, v)
    finally:
        for n in sizes:
            gpu.tune_set(n, -1)


def test_autotune_picks_a_fitting_variant(gpu, oracle_lib, tmp_path):
    import torch
    n, m = 300, 64
    _say("autotune", n, m)
    dev = gpu.generate(0, 1, n, m, I)
    y0 = _np(dev[5]).copy()
    best, ms, times = gpu.autotune(*dev, n, m, I, reps=2)
    assert best >= 0 and ms > 0 and times[best] == ms
    assert gpu.variants()[best]["max_n"] >= n
    assert gpu.tune_get(n) == best
    assert np.array_equal(_np(dev[5]), y0)   # autotune leaves y untouched
    path = str(tmp_path / "tune.txt")
    gpu.tune_save(path)
    gpu.tune_set(n, -1)
    gpu.tune_load(path)
    assert gpu.tune_get(n) == best
    gpu.tune_set(n, -1)
    torch.cuda.synchronize()


# ---------------- FP32 PCR + FP64 refinement ----------------

def _refined(gpu, dev, n, m, layout, max_sweeps=8, tol=1e-12):
    import torch
    sweeps = torch.full((m,), -1, dtype=torch.int32, device="cuda")
    a, b, c, d, e, y = dev
    gpu.solve_refined(a, b, c, d, e, y, n, m, layout, max_sweeps=max_sweeps, tol=tol, sweeps=sweeps)
    torch.cuda.synchronize()
    return _np(y), _np(sweeps)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 33, 100, 255, 256, 512, 700, 1000, 1024, 1500, 2048,
                               2049, 3000, 4096, 5000, 8191, 20000])
def test_refined_meets_1e12_vs_dgbsv(gpu, oracle_lib, n):
    """FP32 PCR + FP64 refinement: normwise relative error <= 1e-12 against the
    FP64 banded-LU (dgbsv) oracle, both generators, both layouts."""
    m = 24
    for kind, layout in ((0, I), (1, C)):
        _say("refined", n, m, kind, layout)
        dev = gpu.generate(kind, 300 + n, n, m, layout)
        ins = [_np(t).copy() for t in dev]
        x, sw = _refined(gpu, dev, n, m, layout)
        ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
        err = oracle_lib.normwise_relerr(x, ref, n, m, layout)
        assert err <= 1e-12, (n, kind, layout, err, sw)
        assert sw.min() >= 0 and sw.max() <= 3, sw   # the 2x2-block inner solve: ~3e-7 per solve
        for t, h in zip(dev[:5], ins[:5]):      # a..e untouched
            assert np.array_equal(_np(t), h)


def test_refined_reports_nonconvergence(gpu, oracle_lib):
    """max_sweeps = 0 stops after the bare FP32 solve and flags it (sweeps = 1)."""
    n, m = 512, 8
    _say("refined-flag", n, m)
    dev = gpu.generate(0, 5, n, m, I)
    ins = [_np(t).copy() for t in dev]
    x, sw = _refined(gpu, dev, n, m, I, max_sweeps=0, tol=1e-15)
    assert (sw == 1).all(), sw
    ref = oracle_lib.gbsv_batch(*ins, n, m, I)
    err = oracle_lib.normwise_relerr(x, ref, n, m, I)
    assert 1e-9 < err < 1e-1, err   # a plain FP32 PCR solve


def test_refined_n2048_slice(gpu, oracle_lib):
    """FP32+refine at n = 2048 on a slice (m = 256) of the m = 16384 batch."""
    n, m = 2048, 256
    _say("refined-n2048", n, m)
    dev = gpu.generate(0, 0, n, m, I)
    ins = [_np(t).copy() for t in dev]
    x, sw = _refined(gpu, dev, n, m, I)
    ref = oracle_lib.gbsv_batch(*ins, n, m, I)
    assert oracle_lib.normwise_relerr(x, ref, n, m, I) <= 1e-12
    print("sweeps histogram", np.bincount(sw), flush=True)


@pytest.mark.parametrize("layout", [I, C])
def test_refined_large_batch_window_route(gpu, oracle_lib, layout):
    """n in (1024, 2048] with at least 2^23 unknowns runs the window path
    (the one-workgroup kernel below that): 1e-12 vs dgbsv, 2 corrections."""
    n = 1100
    m = (1 << 23) // n + 1
    _say("refined-window-route", n, m, layout)
    dev = gpu.generate(0, 77, n, m, layout)
    ins = [_np(t).copy() for t in dev]
    x, sw = _refined(gpu, dev, n, m, layout)
    ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
    err = oracle_lib.normwise_relerr(x, ref, n, m, layout)
    assert err <= 1e-12, (err, np.bincount(sw))
    assert sw.min() >= 1 and sw.max() <= 3, np.bincount(sw)


@pytest.mark.parametrize("n", [64, 1000, 3000])
def test_refined_flags_nan_system(gpu, oracle_lib, n):
    """A system with a NaN coefficient is flagged (sweeps = max_sweeps + 1)
    and does not disturb the others (one-workgroup and split paths)."""
    import torch
    m = 4
    _say("refined-nan", n, m)
    dev = gpu.generate(0, 17, n, m, I)
    ins = [_np(t).copy() for t in dev]
    dev[2].view(n, m)[n // 2, 1] = float("nan")      # c of system 1
    x, sw = _refined(gpu, dev, n, m, I, max_sweeps=4)
    assert sw[1] == 5, sw
    others = [0, 2, 3]
    assert all(1 <= sw[j] <= 3 for j in others), sw
    ref = oracle_lib.gbsv_batch(*ins, n, m, I)
    xs, rs = x.reshape(n, m)[:, others], ref.reshape(n, m)[:, others]
    assert np.abs(xs - rs).max() / np.abs(rs).max() <= 1e-12


@pytest.mark.parametrize("n", [257, 300, 512, 600, 1000, 1024])
def test_refined_grouped_load_matches_single(gpu, oracle_lib, n):
    """Interleaved batches of an even number of systems, more than the GPU
    has CUs, load pairs of neighbouring systems per workgroup with 16-byte
    accesses (257 <= n <= 1024); the same systems in the contiguous layout go
    one per workgroup. The arithmetic of a system does not depend on that,
    so x and the sweep counts agree bit for bit, also when one system of a
    pair is NaN and stops with a flag while its neighbour keeps iterating."""
    import torch
    group = 2
    m = torch.cuda.get_device_properties(0).multi_processor_count + 2
    _say("refined-grouped", n, m)
    dev = gpu.generate(1, 40 + n, n, m, I)
    dev[2].view(n, m)[n // 3, group + 1] = float("nan")   # c of system group + 1
    host = [_np(t).copy() for t in dev]
    xi, swi = _refined(gpu, dev, n, m, I, max_sweeps=4)
    cont = [torch.from_numpy(np.ascontiguousarray(h.reshape(n, m).T).reshape(-1)).cuda()
            for h in host]
    xc, swc = _refined(gpu, cont, n, m, C, max_sweeps=4)
    xc_i = xc.reshape(m, n).T.reshape(-1)
    assert np.array_equal(swi, swc), (swi, swc)
    assert swi[group + 1] == 5 and all(1 <= swi[j] <= 3 for j in range(m) if j != group + 1), swi
    keep = [j for j in range(m) if j != group + 1]
    xs = xi.reshape(n, m)[:, keep]
    assert np.array_equal(xs.view(np.int64), xc_i.reshape(n, m)[:, keep].view(np.int64))
    ref = oracle_lib.gbsv_batch(*host, n, m, I)
    rs = ref.reshape(n, m)[:, keep]
    assert np.abs(xs - rs).max() / np.abs(rs).max() <= 1e-12


def test_refined_split_reports_nonconvergence(gpu, oracle_lib):
    """Split refinement path: max_sweeps = 0 gives the bare FP32 split solve,
    flagged with sweeps = 1, both layouts."""
    n, m = 4096, 5
    for layout in (I, C):
        _say("refined-split-flag", n, m, layout)
        dev = gpu.generate(0, 9, n, m, layout)
        ins = [_np(t).copy() for t in dev]
        x, sw = _refined(gpu, dev, n, m, layout, max_sweeps=0, tol=1e-15)
        assert (sw == 1).all(), sw
        ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
        err = oracle_lib.normwise_relerr(x, ref, n, m, layout)
        assert 1e-9 < err < 1e-2, err


@pytest.mark.parametrize("n", [2049, 2050, 2051, 4097, 65535, 131073])
def test_refined_split_boundary_sizes(gpu, oracle_lib, n):
    """Split refinement path at odd sizes and just past the HBM-level
    boundaries (block rows 1025, 2049, ...): 1e-12 against dgbsv within three
    corrections, both layouts, both generators."""
    m = 3
    for kind, layout in ((0, I), (1, C)):
        _say("refined-split-size", n, m, kind, layout)
        dev = gpu.generate(kind, 500 + n, n, m, layout)
        ins = [_np(t).copy() for t in dev]
        x, sw = _refined(gpu, dev, n, m, layout)
        ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
        err = oracle_lib.normwise_relerr(x, ref, n, m, layout)
        assert err <= 1e-12, (n, kind, layout, err, sw)
        assert sw.min() >= 1 and sw.max() <= 3, sw


def test_refined_split_n2pow20_weak_sample(gpu, oracle_lib):
    """n = 2^20, weakly dominant, 2 systems, FP32 + refinement:
    within 1e-12 of dgbsv."""
    n, m = 1 << 20, 2
    _say("refined-split-n2^20", n, m)
    dev = gpu.generate(1, 0, n, m, I)
    ins = [_np(t).copy() for t in dev]
    x, sw = _refined(gpu, dev, n, m, I)
    xg = oracle_lib.gbsv_batch(*ins, n, m, I)
    err = oracle_lib.normwise_relerr(x, xg, n, m, I)
    print("n=2^20 refined vs dgbsv", err, "sweeps", sw, flush=True)
    assert err <= 1e-12 and sw.max() <= 3


# ---------------- HBM split path (n above the LDS threshold) ----------------

@pytest.mark.parametrize("n", [1025, 1500, 2048, 2049, 3000, 4096, 5000, 8191, 16384])
def test_split_path_bit_exact(gpu, oracle_lib, n):
    """n > 1024: fused HBM passes + LDS stage give the unsplit oracle's bits
    (kernel.cu:158-320 + 665-791 are the same arithmetic as kernel.cu:32-154)."""
    m = 6
    for layout in (I, C):
        _say("split", n, m, layout)
        dev = gpu.generate(0, 40 + n, n, m, layout)
        ins = [_np(t).copy() for t in dev]
        x = _run_f64(gpu, dev, n, m, layout)
        ref = oracle_lib.pcr_batch(*ins, n, m, layout)
        assert np.array_equal(x, ref), (n, layout)
        for t, h in zip(dev[:5], ins[:5]):
            assert np.array_equal(_np(t), h)


def test_split_tunables_bit_exact(gpu, oracle_lib):
    """Deep splits on small systems: every LDS threshold, steps-per-pass and
    tile shape gives the same bits."""
    import torch
    saved = gpu.split_tune_get()
    n_shapes = saved[3]
    try:
        for n, m in ((40, 5), (301, 3), (1000, 4), (2500, 2)):
            dev = gpu.generate(1, 7 + n, n, m, I)
            ins = [_np(t).copy() for t in dev]
            ref = oracle_lib.pcr_batch(*ins, n, m, I)
            for lds_max in (4, 16, 64, 256):
                for k in (1, 2, 3, 5):
                    for shape in range(n_shapes):
                        try:
                            gpu.split_tune_set(lds_max, k, shape)
                        except gpu.PCRPError:
                            continue   # halo too wide for this tile shape
                        _say("split-tune", n, m, lds_max, k, shape)
                        y = torch.from_numpy(ins[5]).cuda()
                        gpu.solve(*dev[:5], y, n, m, I)
                        torch.cuda.synchronize()
                        assert np.array_equal(_np(y), ref), (n, lds_max, k, shape)
    finally:
        gpu.split_tune_set(*saved[:3])


def test_split_spaced_pass_shapes_bit_exact(gpu, oracle_lib):
    """The spaced pass shapes at the production LDS threshold: one pass round
    (n = 2048), a ragged split (n = 5000) and two rounds (n = 2^16), both
    layouts, several steps per pass. Bit-identical to the oracle."""
    import torch
    saved = gpu.split_tune_get()
    spaced_shapes = (8, 9, 10, 11)
    assert saved[3] >= 12
    try:
        for n, m in ((2048, 6), (5000, 5), (1 << 16, 2)):
            for layout in (I, C):
                dev = gpu.generate(0, 40 + n, n, m, layout)
                ins = [_np(t).copy() for t in dev]
                ref = oracle_lib.pcr_batch(*ins, n, m, layout)
                for shape in spaced_shapes:
                    for k in (3, 5):
                        try:
                            gpu.split_tune_set(1024, k, shape)
                        except gpu.PCRPError:
                            continue   # halo too wide for this tile shape
                        _say("split-spaced", n, m, layout, k, shape)
                        y = torch.from_numpy(ins[5]).cuda()
                        gpu.solve(*dev[:5], y, n, m, layout)
                        torch.cuda.synchronize()
                        assert np.array_equal(_np(y), ref), (n, m, layout, k, shape)
    finally:
        gpu.split_tune_set(*saved[:3])


def test_split_n2pow20_weak_sample(gpu, oracle_lib):
    """n = 2^20 systems (weakly dominant generator), 2 systems:
    bit-identical to the oracle and within FP64 PCR noise of dgbsv."""
    n, m = 1 << 20, 2
    _say("split-n2^20", n, m)
    dev = gpu.generate(1, 0, n, m, I)
    ins = [_np(t).copy() for t in dev]
    x = _run_f64(gpu, dev, n, m, I)
    ref = oracle_lib.pcr_batch(*ins, n, m, I)
    assert np.array_equal(x, ref)
    xg = oracle_lib.gbsv_batch(*ins, n, m, I)
    print("n=2^20 pcr64 vs dgbsv", oracle_lib.normwise_relerr(x, xg, n, m, I), flush=True)


# ---------------- uneven batches (persistent work queue) ----------------

def _uneven_case(gpu, oracle_lib, sizes, kind, seed):
    import torch
    off, n, total = gpu.uneven_layout(sizes)
    dev = gpu.generate_uneven(kind, seed, off, n)
    ins = [_np(t).copy() for t in dev]
    gpu.solve_uneven(*dev, off, n)
    torch.cuda.synchronize()
    x = _np(dev[5])
    ref = oracle_lib.pcr_uneven(*ins, off, n)
    return x, ref, ins, dev


def test_uneven_generator_matches_batch_generator(gpu, oracle_lib):
    sizes = [5, 64, 300]
    off, n, _ = gpu.uneven_layout(sizes)
    dev = gpu.generate_uneven(0, 3, off, n)
    for j, (o, nj) in enumerate(zip(off, n)):
        ref = oracle_lib.gen_batch(0, 3, int(nj), j + 1, oracle_lib.CONTIGUOUS)
        sysj = oracle_lib.system(ref, j, int(nj), j + 1, oracle_lib.CONTIGUOUS)
        for u, v in zip(dev, sysj):
            assert np.array_equal(_np(u)[o:o + nj], v)


@pytest.mark.parametrize("kind,n,m", [(1, 1 << 20, 32), (0, 2048, 16384), (0, 1024, 16384)])
def test_f64_full_benchmark_shapes_bit_exact(gpu, oracle_lib, kind, n, m):
    """Plain FP64 at the full benchmark shapes (n = 2^20, m = 32 weakly
    dominant; n = 2048, m = 16384, split path; n = 1024, m = 16384, the LDS
    kernel's own benchmark batch; interleaved): every system equal to the
    oracle's PCR-Penta bit for bit."""
    _say("f64-full", kind, n, m)
    dev = gpu.generate(kind, 0, n, m, I)
    ins = [_np(t).copy() for t in dev]
    x = _run_f64(gpu, dev, n, m, I)
    assert np.array_equal(x, oracle_lib.pcr_batch(*ins, n, m, I))


def test_uneven_f64_full_batch_bit_exact(gpu, oracle_lib):
    """The whole uneven benchmark batch (10000 systems, n log-uniform in
    [64, 131072], 1.67e8 unknowns) equals the oracle bit for bit."""
    sizes = gpu.uneven_log_uniform_sizes()
    _say("uneven-f64-full", len(sizes), int(sizes.sum()))
    x, ref, ins, dev = _uneven_case(gpu, oracle_lib, sizes, 0, 1)
    assert np.array_equal(x, ref)


def test_uneven_bit_exact_mixed_sizes(gpu, oracle_lib):
    rng = np.random.default_rng(11)
    sizes = [1, 2, 3, 5, 64, 100, 1000, 1024, 1025, 2048, 3000, 5000, 16385]
    sizes += list(rng.integers(1, 4000, 60))
    _say("uneven-mixed", len(sizes))
    x, ref, ins, dev = _uneven_case(gpu, oracle_lib, sizes, 0, 21)
    assert np.array_equal(x, ref)
    for t, h in zip(dev[:5], ins[:5]):
        assert np.array_equal(_np(t), h)


def test_uneven_deep_splits_bit_exact(gpu, oracle_lib):
    """Small LDS threshold: several pass rounds with different k per system."""
    saved = gpu.split_tune_get()
    try:
        n_shapes = saved[3]
        combos = [(16, 2, 0), (8, 3, 2), (64, 1, 1), (32, 4, 0)]
        combos += [(16, 2 + (sh % 3), sh) for sh in range(n_shapes)]
        for lds_max, k, shape in combos:
            gpu.split_tune_set(lds_max, k, shape)
            rng = np.random.default_rng(lds_max + k)
            sizes = list(rng.integers(1, 3000, 40)) + [4096, 777]
            _say("uneven-deep", lds_max, k, shape)
            x, ref, _, _ = _uneven_case(gpu, oracle_lib, sizes, 1, 5 + k)
            assert np.array_equal(x, ref), (lds_max, k, shape)
    finally:
        gpu.split_tune_set(*saved[:3])


def test_uneven_grouped_residue_layout_bit_exact(gpu, oracle_lib):
    """Deep splits at the production thresholds: systems around powers of two
    up to 2^16 (S up to 7 at lds_max 512), so the pass buffers hold residue
    groups of 2 and 4 with short last residues; several steps per pass and
    both tile families (4 sub-systems per tile, 8 per tile), with the
    redundant and the spaced step forms (shapes 8-11; 128 x 8 tiles at k = 5
    give 32 threads per sub-system, the most k allows). Bit-identical to the
    oracle in every setting."""
    saved = gpu.split_tune_get()
    sizes = [1025, 1027, 2047, 2049, 3071, 4097, 8191, 8193, 12289, 16385, 32767, 40001, 65537]
    sizes += [5, 700, 1024]
    try:
        for lds_max, k, shape in ((1024, 3, 0), (1024, 5, 2), (512, 3, 4), (512, 5, 6),
                                  (256, 4, 0), (1024, 4, 8), (512, 3, 9), (1024, 5, 10),
                                  (256, 6, 11)):
            gpu.split_tune_set(lds_max, k, shape)
            _say("uneven-grouped", lds_max, k, shape)
            x, ref, _, _ = _uneven_case(gpu, oracle_lib, sizes, 0, 60 + k)
            assert np.array_equal(x, ref), (lds_max, k, shape)
    finally:
        gpu.split_tune_set(*saved[:3])


def test_uneven_queue_step_forms_bit_exact(gpu, oracle_lib):
    """The work queue pinned to the spaced queue kernel (spaced steps for jobs
    above 736 rows, redundant below, chosen per job) and to the redundant
    one: whole systems and residue sub-systems on both sides of the switch,
    bit-identical to the oracle."""
    vs = {v["kind"] + f"-{v['nt']}x{v['rpt']}": v["id"] for v in gpu.variants()}
    rng = np.random.default_rng(77)
    sizes = [1, 2, 700, 736, 737, 800, 1023, 1024, 1473, 1475, 2047, 2049, 2943, 5889]
    sizes += list(rng.integers(1, 3000, 40))
    try:
        for name in ("spaced-256x4", "redundant-256x4"):
            gpu.tune_set(1024, vs[name])
            _say("uneven-queue", name, len(sizes))
            x, ref, _, _ = _uneven_case(gpu, oracle_lib, sizes, 0, 91)
            assert np.array_equal(x, ref), name
    finally:
        gpu.tune_set(1024, -1)


def test_uneven_log_uniform_sample(gpu, oracle_lib):
    """First 400 systems of the uneven size draw (n log-uniform in [64, 131072], seed 1)."""
    sizes = gpu.uneven_log_uniform_sizes()[:400]
    _say("uneven-c4", len(sizes), int(sizes.sum()))
    x, ref, _, _ = _uneven_case(gpu, oracle_lib, sizes, 0, 1)
    assert np.array_equal(x, ref)


def _uneven_refined_case(gpu, oracle_lib, sizes, kind, seed, max_sweeps=8):
    import torch
    off, n, total = gpu.uneven_layout(sizes)
    dev = gpu.generate_uneven(kind, seed, off, n)
    ins = [_np(t).copy() for t in dev]
    sweeps = torch.full((len(n),), -1, dtype=torch.int32, device="cuda")
    gpu.solve_uneven_refined(*dev, off, n, max_sweeps=max_sweeps, tol=1e-12, sweeps=sweeps)
    torch.cuda.synchronize()
    x = _np(dev[5])
    errs = []
    for o, nj in zip(off, n):
        sl = slice(int(o), int(o) + int(nj))
        ref = oracle_lib.gbsv(*[t[sl] for t in ins])
        errs.append(np.abs(x[sl] - ref).max() / np.abs(ref).max())
    for t, h in zip(dev[:5], ins[:5]):      # a..e untouched
        assert np.array_equal(_np(t), h)
    return np.array(errs), _np(sweeps)


def test_uneven_refined_mixed_sizes(gpu, oracle_lib):
    """Uneven FP32 + refinement: every system within 1e-12 of dgbsv (normwise),
    small systems, odd sizes, and sizes that need one or two pass rounds."""
    rng = np.random.default_rng(12)
    sizes = [1, 2, 3, 5, 64, 100, 1000, 2047, 2048, 2049, 2051, 4097, 5000, 40000, 70001]
    sizes += list(rng.integers(1, 6000, 50))
    for kind in (0, 1):
        _say("uneven-refined", len(sizes), kind)
        errs, sw = _uneven_refined_case(gpu, oracle_lib, sizes, kind, 31 + kind)
        assert errs.max() <= 1e-12, (kind, errs.max(), int(errs.argmax()))
        assert sw.min() >= 1 and sw.max() <= 3, sw


def test_uneven_refined_flags_nan_system(gpu, oracle_lib):
    """Uneven batch (window path, odd and even offsets): a system with a NaN
    coefficient reports max_sweeps + 1, the others meet 1e-12 vs dgbsv."""
    import torch
    sizes = [301, 5000, 2100, 777]
    off, n, total = gpu.uneven_layout(sizes)
    _say("uneven-refined-nan", sizes)
    dev = gpu.generate_uneven(0, 23, off, n)
    ins = [_np(t).copy() for t in dev]
    dev[2][int(off[1]) + 2500] = float("nan")          # c of system 1
    sweeps = torch.full((len(n),), -1, dtype=torch.int32, device="cuda")
    gpu.solve_uneven_refined(*dev, off, n, max_sweeps=4, tol=1e-12, sweeps=sweeps)
    torch.cuda.synchronize()
    sw, x = _np(sweeps), _np(dev[5])
    assert sw[1] == 5, sw
    for j in (0, 2, 3):
        assert 1 <= sw[j] <= 3, sw
        s = slice(int(off[j]), int(off[j]) + int(n[j]))
        ref = oracle_lib.gbsv(*[t[s] for t in ins])
        assert np.abs(x[s] - ref).max() / np.abs(ref).max() <= 1e-12, j


def test_uneven_refined_tiny_windows_on_resident_grid(gpu, oracle_lib):
    """Thousands of 1- and 2-row systems among long ones, with a tolerance so
    strict that sweeps 3+ run on the resident grid (a workgroup walks several
    windows, and a window of one block row runs no level): every system within
    1e-12 of dgbsv, and results and sweep counts equal those of solving the
    system alone (the per-window max reduction must not mix windows)."""
    import torch
    sizes = []
    for k in range(3000):
        sizes.append(1 + (k % 2))
        if k % 300 == 0:
            sizes.append(5000 + k)
    off, n, total = gpu.uneven_layout(sizes)
    _say("uneven-refined-tiny", len(sizes), total)
    dev = gpu.generate_uneven(0, 77, off, n)
    ins = [_np(t).copy() for t in dev]
    # tol below what two corrections reach (~1e-13 relative corrections): systems
    # stop at sweep 3 or later, or run out
    tol, ms = 1e-20, 6
    sweeps = torch.full((len(n),), -1, dtype=torch.int32, device="cuda")
    gpu.solve_uneven_refined(*dev, off, n, max_sweeps=ms, tol=tol, sweeps=sweeps)
    torch.cuda.synchronize()
    sw, x = _np(sweeps), _np(dev[5])
    print("tiny-window batch sweeps", np.bincount(sw), flush=True)
    assert sw.max() >= 3, sw   # sweeps from the third on run on the resident grid
    for j in range(len(n)):
        sl = slice(int(off[j]), int(off[j]) + int(n[j]))
        ref = oracle_lib.gbsv(*[t[sl] for t in ins])
        assert np.abs(x[sl] - ref).max() / np.abs(ref).max() <= 1e-12, j
    long_ids = [j for j in range(len(n)) if n[j] > 2]
    check = long_ids + list(range(0, len(n), 97))
    for j in check:
        sl = slice(int(off[j]), int(off[j]) + int(n[j]))
        # same element-offset parity as in the batch (odd offsets pair rows from row -1)
        par = int(off[j]) & 1
        one = []
        for t in ins:
            buf = np.zeros(par + int(n[j]))
            buf[par:] = t[sl]
            one.append(torch.from_numpy(buf).cuda())
        s1 = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        gpu.solve_uneven_refined(*one, [par], [int(n[j])], max_sweeps=ms, tol=tol, sweeps=s1)
        torch.cuda.synchronize()
        assert int(s1.item()) == int(sw[j]), (j, int(n[j]), int(s1.item()), int(sw[j]))
        assert np.array_equal(_np(one[5])[par:], x[sl]), (j, int(n[j]))


def test_uneven_refined_flags_nonconvergence(gpu, oracle_lib):
    """max_sweeps = 0 returns the bare FP32 solve, flagged (sweeps = 1)."""
    sizes = [10, 3000, 5000]
    _say("uneven-refined-flag", sizes)
    errs, sw = _uneven_refined_case(gpu, oracle_lib, sizes, 0, 2, max_sweeps=0)
    assert (sw == 1).all(), sw
    assert errs.max() < 1e-2 and errs.max() > 1e-9, errs


def test_uneven_refined_log_uniform_sample(gpu, oracle_lib):
    """First 300 systems of the uneven size draw, FP32 + refinement."""
    sizes = gpu.uneven_log_uniform_sizes()[:300]
    _say("uneven-refined-c4", len(sizes), int(sizes.sum()))
    errs, sw = _uneven_refined_case(gpu, oracle_lib, sizes, 0, 1)
    print("uneven sample refined: max err", errs.max(), "sweeps", np.bincount(sw), flush=True)
    assert errs.max() <= 1e-12 and sw.max() <= 3


def test_uneven_refined_full_batch(gpu, oracle_lib):
    """The whole uneven benchmark batch (10000 systems, n log-uniform in
    [64, 131072]), FP32 + refinement: every system within 1e-12 of dgbsv."""
    sizes = gpu.uneven_log_uniform_sizes()
    _say("uneven-refined-full", len(sizes), int(sizes.sum()))
    errs, sw = _uneven_refined_case(gpu, oracle_lib, sizes, 0, 1)
    print("full uneven refined: max err", errs.max(), "sweeps", np.bincount(sw), flush=True)
    assert errs.max() <= 1e-12 and sw.min() >= 1 and sw.max() <= 3


def _slow_decay_rows(n, seed):
    """Rows of c = 2.05, b = d = -1, a = e = 0 (a shifted 1-D Laplacian,
    condition ~80): its block multipliers fall below 2^-20 only after 6
    reduction levels, past the reach of a long system's window, so a batch
    holding it is solved by the pass/stage path."""
    rng = np.random.default_rng(seed)
    a = np.zeros(n); e = np.zeros(n)
    b = -np.ones(n); d = -np.ones(n); c = np.full(n, 2.05)
    b[0] = 0.0; d[-1] = 0.0
    return [a, b, c, d, e, rng.uniform(-1, 1, n)]


@pytest.mark.parametrize("layout", [I, C])
def test_refined_split_falls_back_on_slow_decay(gpu, oracle_lib, layout):
    """A uniform batch with one system whose couplings outlast the window halo
    (the others from the dominant generator) still meets 1e-12 vs dgbsv."""
    import torch
    n, m = 6000, 3
    _say("refined-split-fallback", n, m, layout)
    dev = gpu.generate(0, 41, n, m, layout)
    slow = _slow_decay_rows(n, 5)
    for t, v in zip(dev, slow):
        (t.view(n, m)[:, 1] if layout == I else t.view(m, n)[1, :]).copy_(torch.from_numpy(v))
    ins = [_np(t).copy() for t in dev]
    x, sw = _refined(gpu, dev, n, m, layout)
    ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
    err = oracle_lib.normwise_relerr(x, ref, n, m, layout)
    assert err <= 1e-12, (err, sw)
    assert sw.max() <= 8, sw


def test_uneven_refined_falls_back_on_slow_decay(gpu, oracle_lib):
    """Uneven batch with one long slow-decay system among dominant ones."""
    import torch
    sizes = [300, 5000, 2100, 9000]
    off, n, total = gpu.uneven_layout(sizes)
    _say("uneven-refined-fallback", sizes)
    dev = gpu.generate_uneven(0, 8, off, n)
    slow = _slow_decay_rows(9000, 6)
    sl = slice(int(off[3]), int(off[3]) + 9000)
    for t, v in zip(dev, slow):
        t[sl].copy_(torch.from_numpy(v))
    ins = [_np(t).copy() for t in dev]
    sweeps = torch.full((len(n),), -1, dtype=torch.int32, device="cuda")
    gpu.solve_uneven_refined(*dev, off, n, max_sweeps=8, tol=1e-12, sweeps=sweeps)
    torch.cuda.synchronize()
    x = _np(dev[5])
    for o, nj in zip(off, n):
        s = slice(int(o), int(o) + int(nj))
        ref = oracle_lib.gbsv(*[t[s] for t in ins])
        assert np.abs(x[s] - ref).max() / np.abs(ref).max() <= 1e-12, (int(nj), _np(sweeps))


@pytest.mark.parametrize("kind,n,m", [(1, 1 << 20, 32), (0, 2048, 16384)])
def test_refined_full_benchmark_shapes(gpu, oracle_lib, kind, n, m):
    """FP32 refinement at the full benchmark shapes (n = 2^20, m = 32, weakly
    dominant; n = 2048, m = 16384, strictly dominant; interleaved, window
    path): every system within 1e-12 of dgbsv in two or three corrections."""
    _say("refined-full", kind, n, m)
    dev = gpu.generate(kind, 0, n, m, I)
    ins = [_np(t).copy() for t in dev]
    x, sw = _refined(gpu, dev, n, m, I)
    ref = oracle_lib.gbsv_batch(*ins, n, m, I)
    err = oracle_lib.normwise_relerr(x, ref, n, m, I)
    print("full-shape refined vs dgbsv", err, np.bincount(sw), flush=True)
    assert err <= 1e-12 and sw.min() >= 1 and sw.max() <= 3


def test_refined_split_under_graph_capture(gpu, oracle_lib):
    """A long-system FP32 refinement captured in a graph takes the pass/stage
    path (the window path's host read-back cannot be captured): three
    replays, with eager copies and synchronisations between them, each within
    1e-12 of dgbsv and of the eager (window path) result."""
    import torch
    n, m = 4096, 3
    _say("refined-graph", n, m)
    dev = gpu.generate(0, 61, n, m, I)
    ins = [_np(t).copy() for t in dev]
    ref = oracle_lib.gbsv_batch(*ins, n, m, I)
    y0 = dev[5].clone()
    ws, nb = gpu._workspace(n, m, gpu.PREC_F32_REFINE, dev[5].device)
    x_eager, _ = _refined(gpu, dev, n, m, I)
    assert oracle_lib.normwise_relerr(x_eager, ref, n, m, I) <= 1e-12
    y = y0.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        gpu.solve_refined(*dev[:5], y, n, m, I, max_sweeps=8, tol=1e-12, stream=s, ws=ws)
    torch.cuda.synchronize()
    for k in range(3):
        y.copy_(y0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        x = _np(y)
        assert oracle_lib.normwise_relerr(x, ref, n, m, I) <= 1e-12, k
        assert oracle_lib.normwise_relerr(x, x_eager, n, m, I) <= 1e-12, k


@pytest.mark.parametrize("n", [2049, 2700, 2701, 2702, 3601, 3603, 4500, 2892, 2893, 2894,
                               3856, 3857])
def test_refined_window_edges(gpu, oracle_lib, n):
    """Long systems cut into windows (interior block rows balanced over the
    system's windows, at most 512 - 2 (2^L - 1): 482 at the default L = 4,
    450 at L = 5): sizes with one block row past a window edge for either
    count, exact multiples, odd n; both generators and layouts (the
    interleaved batch is transposed first), m not a multiple of 32."""
    m = 5
    for kind, layout in ((0, I), (1, C)):
        _say("refined-window", n, m, kind, layout)
        dev = gpu.generate(kind, 900 + n, n, m, layout)
        ins = [_np(t).copy() for t in dev]
        x, sw = _refined(gpu, dev, n, m, layout)
        ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
        err = oracle_lib.normwise_relerr(x, ref, n, m, layout)
        assert err <= 1e-12, (n, kind, layout, err, sw)
        assert sw.min() >= 1 and sw.max() <= 3, sw


# ---------------- Heston ADI driver (hipGraph time steps) ----------------

def _heston_gpu(gpu, precision, use_graph, **kw):
    import torch
    p = gpu.heston_params(**kw)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        u = gpu.heston(p, precision=precision, use_graph=use_graph, stream=s)
    s.synchronize()
    return _np(u).reshape(p.nv, p.ns), p


def test_heston_matches_cpu_restatement(gpu, oracle_lib):
    """Small grid, FP64: the GPU Douglas steps equal a numpy restatement whose
    line solves use the CPU oracle (normwise relative difference <= 1e-12;
    the stencil sums round in a different order)."""
    import heston_ref as H
    kw = dict(ns=64, nv=32, steps=20)
    _say("heston-small")
    U, p = _heston_gpu(gpu, gpu.PREC_F64, False, **kw)
    d = dict(gpu.HESTON_GRID)
    d.update(kw)
    ref = H.adi(d, oracle_lib)
    err = np.abs(U - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


def test_heston_full_grid_f64_matches_cpu_restatement(gpu, oracle_lib):
    """The Heston grid itself (Ns = 1024, Nv = 512), FP64, first 4 Douglas
    steps: the GPU driver equals the numpy restatement with oracle line
    solves to 1e-12 normwise (the stencil sums round in another order)."""
    import heston_ref as H
    kw = dict(steps=4)
    _say("heston-full-grid-f64", kw)
    U, p = _heston_gpu(gpu, gpu.PREC_F64, False, **kw)
    d = dict(gpu.HESTON_GRID)
    d.update(kw)
    ref = H.adi(d, oracle_lib)
    err = np.abs(U - ref).max() / np.abs(ref).max()
    print("heston full grid FP64 vs restatement", err, flush=True)
    assert err <= 1e-12, err


def _heston_line_rhs(oracle_lib, d):
    """Line systems of the Heston grid and the right-hand sides of a Douglas
    step a few steps in (S-lines: R1; v-lines: X = Y1 + W, with Y1 from the
    FP64 oracle), built on the CPU."""
    import heston_ref as H
    S, V = H.line_systems(d)
    ns, nv = d["ns"], d["nv"]
    U = H.adi(dict(d, steps=3), oracle_lib)
    R1, W = H.explicit(d, U, 3)
    Y1 = oracle_lib.pcr_batch(*[t.ravel() for t in S], R1.ravel(), ns, nv,
                              oracle_lib.CONTIGUOUS).reshape(nv, ns)
    return S, V, R1, Y1 + W


@pytest.mark.parametrize("lines", ["S", "v"])
def test_heston_line_solves_refined_vs_dgbsv(gpu, oracle_lib, lines):
    """Every Heston line system (S-lines: 512 contiguous systems of n = 1024,
    v-lines: 1024 interleaved systems of n = 512) solved by FP32 + FP64
    refinement at the driver's settings is within 1e-12 of dgbsv, normwise,
    system by system."""
    import torch
    d = dict(gpu.HESTON_GRID)
    S, V, R1, X = _heston_line_rhs(oracle_lib, d)
    ns, nv = d["ns"], d["nv"]
    if lines == "S":
        diags, rhs, n, m, layout = S, R1, ns, nv, C
    else:
        diags, rhs, n, m, layout = V, X, nv, ns, I
    ins = [np.ascontiguousarray(t).ravel() for t in diags] + [np.ascontiguousarray(rhs).ravel()]
    _say("heston-lines-refined", lines, n, m, layout)
    dev = [torch.from_numpy(t.copy()).cuda() for t in ins]
    x, sw = _refined(gpu, dev, n, m, layout, max_sweeps=gpu.HESTON_MAX_SWEEPS,
                     tol=gpu.HESTON_REFINE_TOL)
    ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
    rs, ss = oracle_lib.strides(n, m, layout)
    errs = np.array([np.abs(x[j * ss + np.arange(n) * rs] - ref[j * ss + np.arange(n) * rs]).max()
                     / np.abs(ref[j * ss + np.arange(n) * rs]).max() for j in range(m)])
    print(f"heston {lines}-lines: max normwise error {errs.max():.3e} (median "
          f"{np.median(errs):.3e}), sweeps {np.bincount(sw).tolist()}", flush=True)
    assert errs.max() <= 1e-12, (lines, errs.max(), int(errs.argmax()))
    assert sw.min() >= 1 and sw.max() <= gpu.HESTON_MAX_SWEEPS, sw


def test_heston_graph_replay_equals_eager(gpu):
    _say("heston-graph")
    kw = dict(ns=128, nv=64, steps=30)
    for prec in (gpu.PREC_F64, gpu.PREC_F32_REFINE):
        for reuse in (0, 1):
            Ue, _ = _heston_gpu(gpu, prec, False, reuse_factor=reuse, **kw)
            Ug, _ = _heston_gpu(gpu, prec, True, reuse_factor=reuse, **kw)
            assert np.array_equal(Ue, Ug), (prec, reuse)


@pytest.mark.parametrize("prec", [0, 1])
def test_heston_folded_stage_equals_separate(gpu, prec):
    """The explicit stage computed by the S-line solve kernels while they load
    (two launches per step, the grid alternating between two buffers) gives
    the separate explicit-stage kernel's price grid bit for bit: both evaluate
    the same stencil expressions, so the solves see the same right-hand sides.
    Odd step counts (the last grid copied back), several replays of a captured
    chunk plus a shorter one, odd line lengths (rows loaded one at a time),
    eager steps, both reuse forms, and the full grid."""
    cases = (dict(ns=128, nv=64, steps=45), dict(ns=96, nv=40, steps=7),
             dict(ns=301, nv=37, steps=5), dict(steps=6))
    for kw in cases:
        for use_graph in (False, True):
            for reuse in (0, 1):
                _say("heston-folded", prec, kw, use_graph, reuse)
                U0, _ = _heston_gpu(gpu, prec, use_graph, reuse_factor=reuse,
                                    separate_explicit=1, **kw)
                U1, _ = _heston_gpu(gpu, prec, use_graph, reuse_factor=reuse,
                                    separate_explicit=0, **kw)
                assert np.array_equal(U0, U1), (kw, use_graph, reuse, np.abs(U0 - U1).max())


@pytest.mark.parametrize("prec", [0, 1])
def test_heston_factor_reuse_equals_full_solves(gpu, prec):
    """The Heston grid, all 200 steps: factoring the shared v-line matrix once
    gives the same price grid bit for bit as solving every line in full each
    step (the factored solves run the same operations)."""
    _say("heston-reuse", prec)
    U0, _ = _heston_gpu(gpu, prec, True, reuse_factor=0)
    U1, _ = _heston_gpu(gpu, prec, True, reuse_factor=1)
    assert np.array_equal(U0, U1), np.abs(U0 - U1).max()


def test_heston_price_and_refine(gpu):
    """Ns=1024, Nv=512, 200 steps: the FP64 price at S=100, v=0.04
    is within 2e-3 relative of the semi-analytic Heston price, and FP32+refine
    agrees with FP64 to 1e-10 normwise."""
    import heston_ref as H
    from scipy.interpolate import RectBivariateSpline
    _say("heston-1024x512")
    U64, p = _heston_gpu(gpu, gpu.PREC_F64, True)
    U32, _ = _heston_gpu(gpu, gpu.PREC_F32_REFINE, True)
    hS, hv = p.smax / (p.ns + 1), p.vmax / p.nv
    f = RectBivariateSpline(np.arange(p.nv) * hv, np.arange(1, p.ns + 1) * hS, U64)
    price = f(0.04, 100.0)[0, 0]
    ref = H.heston_call(100.0, 0.04, p.strike, p.steps * p.dt, p.r, p.kappa, p.theta_v, p.sigma,
                        p.rho)
    print("heston price", price, "semi-analytic", ref, flush=True)
    assert abs(price - ref) / ref < 2e-3
    assert np.abs(U32 - U64).max() / np.abs(U64).max() <= 1e-10


def test_split_autotune_keeps_results_exact(gpu, oracle_lib):
    import torch
    n, m = 5000, 4
    _say("split-autotune", n, m)
    saved = gpu.split_tune_get()
    try:
        dev = gpu.generate(0, 12, n, m, I)
        ins = [_np(t).copy() for t in dev]
        k, sh, ms = gpu.split_autotune(*dev, n, m, I, reps=1)
        assert 1 <= k <= 6 and ms > 0
        assert gpu.split_tune_get()[1:3] == (k, sh)
        assert gpu.split_tune_get()[0] in (256, 512, 1024)   # the LDS/HBM threshold is tuned too
        assert np.array_equal(_np(dev[5]), ins[5])
        x = _run_f64(gpu, dev, n, m, I)
        assert np.array_equal(x, oracle_lib.pcr_batch(*ins, n, m, I))
    finally:
        gpu.split_tune_set(*saved[:3])
        torch.cuda.synchronize()


# ---------------- FP32 refinement tunables ----------------

@pytest.mark.parametrize("levels", [2, 3, 4, 5, 6, 7])
def test_refined_window_levels_meet_1e12(gpu, oracle_lib, levels):
    """Every window level count (halo 2^L - 1 block rows per side) meets 1e-12
    against dgbsv: few levels make the weak generator's windows non-local, and
    those batches fall back to the split path."""
    saved = gpu.f32r_tune_get()
    try:
        gpu.f32r_tune_set(levels, saved[1])
        for kind, layout, n, m in ((0, I, 40000, 3), (1, C, 33001, 2), (1, I, 2000, 40)):
            _say("window-levels", levels, kind, layout, n, m)
            if n <= 2048:
                gpu.f32r_tune_set(levels, 0)   # the window path for this small batch too
            dev = gpu.generate(kind, 700 + levels, n, m, layout)
            ins = [_np(t).copy() for t in dev]
            x, sw = _refined(gpu, dev, n, m, layout)
            ref = oracle_lib.gbsv_batch(*ins, n, m, layout)
            err = oracle_lib.normwise_relerr(x, ref, n, m, layout)
            assert err <= 1e-12, (levels, kind, layout, n, err)
            assert sw.min() >= 1 and sw.max() <= 3, sw
            gpu.f32r_tune_set(levels, saved[1])
    finally:
        gpu.f32r_tune_set(*saved)


def test_f32r_autotune_keeps_results(gpu, oracle_lib, tmp_path):
    """The FP32 autotuner picks window levels and the window / one-workgroup
    choice on a real batch; the results stay within 1e-12, the tunables
    survive a save / load round trip."""
    saved_f = gpu.f32r_tune_get()
    saved_s = gpu.split_tune_get()
    try:
        for n, m in ((1800, 300), (6000, 6)):
            _say("f32r-autotune", n, m)
            dev = gpu.generate(0, 44 + n, n, m, I)
            ins = [_np(t).copy() for t in dev]
            wl, wp, ms = gpu.f32r_autotune(*dev, n, m, I, reps=1)
            assert 3 <= wl <= 6 and ms > 0 and wp in (0, 1)
            assert gpu.f32r_tune_get()[0] == wl
            if n <= 2048:
                assert (gpu.f32r_tune_get()[1] <= n * m) == bool(wp)
            assert np.array_equal(_np(dev[5]), ins[5])
            x, sw = _refined(gpu, dev, n, m, I)
            ref = oracle_lib.gbsv_batch(*ins, n, m, I)
            assert oracle_lib.normwise_relerr(x, ref, n, m, I) <= 1e-12
        path = str(tmp_path / "tune.txt")
        cur = gpu.f32r_tune_get()
        gpu.tune_save(path)
        gpu.f32r_tune_set(4, 123)
        gpu.tune_load(path)
        assert gpu.f32r_tune_get() == cur
    finally:
        gpu.f32r_tune_set(*saved_f)
        gpu.split_tune_set(*saved_s[:3])

