"""CPU-side checks of the C ABI: the library loads, exports every entry point
declared in include/pcr_penta.h, and rejects invalid requests with a status
code before touching the device (the reference's checkCuda exits instead,
performanalysis.h:12-20)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pcr_penta.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcrp_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import pcrpenta
    if not os.path.exists(pcrpenta.LIB_PATH):
        pcrpenta.build()
    return pcrpenta.load()


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert "pcrp_solve_f64" in names and len(names) >= 6
    for name in names:
        assert hasattr(lib, name), name


def test_abi_version_and_strerror(lib):
    import pcrpenta
    assert lib.pcrp_abi_version() == pcrpenta.ABI_VERSION == 2
    for st, name in pcrpenta.STATUS.items():
        assert pcrpenta.strerror(st)
    assert pcrpenta.strerror(999) == "unknown status"


def _fake(n):
    # never dereferenced: validation fails before any launch
    return [ctypes.c_void_p(0x1000 * (k + 1)) for k in range(n)]


@pytest.mark.parametrize("n,m", [(0, 4), (4, 0), (-1, 1), (1 << 16, 1 << 16)])
def test_invalid_sizes(lib, n, m):
    st = lib.pcrp_solve_f64(*_fake(6), n, m, 1, None, 0, None)
    assert st == 1
    st = lib.pcrp_solve_f32r_f64(*_fake(6), n, m, 1, 4, 1e-13, None, None, 0, None)
    assert st == 1


def test_null_pointer(lib):
    p = _fake(6)
    p[2] = None
    assert lib.pcrp_solve_f64(*p, 16, 4, 1, None, 0, None) == 2
    assert lib.pcrp_generate(0, 1, 16, 4, 1, *p, None) == 2


def test_bad_layout(lib):
    assert lib.pcrp_solve_f64(*_fake(6), 16, 4, 7, None, 0, None) == 3
    assert lib.pcrp_init_t2(16, 4, -1, *_fake(6), None) == 3


def test_bad_generator_kind(lib):
    assert lib.pcrp_generate(5, 1, 16, 4, 1, *_fake(6), None) == 6


def test_refine_argument_checks(lib):
    assert lib.pcrp_solve_f32r_f64(*_fake(6), 16, 4, 1, -1, 1e-13, None, None, 0, None) == 6
    assert lib.pcrp_solve_f32r_f64(*_fake(6), 16, 4, 1, 4, -1.0, None, None, 0, None) == 6


def test_workspace_is_reported(lib):
    for prec in (0, 1):
        for n in (1, 512, 1024, 2048, 4096, 1 << 20):
            assert lib.pcrp_workspace_size(n, 8, prec) >= 0


def test_workspace_too_small_is_rejected(lib):
    n, m = 1 << 14, 4
    need = lib.pcrp_workspace_size(n, m, 0)
    if need == 0:
        pytest.skip("no workspace needed at this size")
    assert lib.pcrp_solve_f64(*_fake(6), n, m, 1, _fake(1)[0], need - 1, None) == 4


def test_refine_split_workspace_is_required(lib):
    """n above the one-workgroup FP32 limit: the refinement needs a workspace."""
    n, m = 4096, 4
    need = lib.pcrp_workspace_size(n, m, 1)
    assert need > 0
    assert lib.pcrp_workspace_size(1024, m, 1) == 0
    assert lib.pcrp_solve_f32r_f64(*_fake(6), n, m, 1, 4, 1e-12, None, None, 0, None) == 4
    assert lib.pcrp_solve_f32r_f64(*_fake(6), n, m, 1, 4, 1e-12, None, _fake(1)[0], need - 1,
                                   None) == 4


def test_refine_window_route_workspace(lib):
    """n in (1024, 2048] may take the window path (from a tunable batch size
    on), so its workspace covers the transposed form (six system-contiguous
    FP64 arrays and two x buffers: 64 B per unknown) at every batch size; at
    n > 2048 the workspace covers the window path as well as the split path."""
    n = 1100
    m_small, m_big = (1 << 23) // n, (1 << 23) // n + 1
    assert lib.pcrp_workspace_size(n, m_small, 1) >= 64 * n * m_small
    assert lib.pcrp_workspace_size(n, m_big, 1) >= 64 * n * m_big
    assert lib.pcrp_workspace_size(1024, 1 << 14, 1) == 0      # n <= 1024: one workgroup
    assert lib.pcrp_workspace_size(2048, 4096, 1) >= 64 * 2048 * 4096
    assert lib.pcrp_workspace_size(3000, 7, 1) >= 64 * 3000 * 7


def test_uneven_argument_checks(lib):
    import numpy as np
    off = np.array([0, 10], dtype=np.int64)
    n = np.array([10, 0], dtype=np.int32)
    p = _fake(6)
    assert lib.pcrp_solve_uneven_f64(*p, off.ctypes.data, n.ctypes.data, 2, None, 0, None) == 1
    n[1] = 5
    assert lib.pcrp_solve_uneven_f64(*p, None, n.ctypes.data, 2, None, 0, None) == 2
    assert lib.pcrp_solve_uneven_f64(*p, off.ctypes.data, n.ctypes.data, 0, None, 0, None) == 1
    # small systems only: no pass arrays, but the plan tables still need workspace
    need = lib.pcrp_uneven_workspace_size(off.ctypes.data, n.ctypes.data, 2, 0)
    assert need > 0
    assert lib.pcrp_solve_uneven_f64(*p, off.ctypes.data, n.ctypes.data, 2, _fake(1)[0],
                                     need - 1, None) == 4
    big = np.array([5000], dtype=np.int32)
    assert lib.pcrp_uneven_workspace_size(off.ctypes.data, big.ctypes.data, 1, 0) >= 2 * 6 * 8 * 5000


def test_uneven_refined_argument_checks(lib):
    """The uneven FP32 + refinement entry point validates like the FP64 one,
    plus the refinement arguments; its workspace covers the block arrays."""
    import numpy as np
    off = np.array([0, 10], dtype=np.int64)
    n = np.array([10, 0], dtype=np.int32)
    p = _fake(6)
    f = lib.pcrp_solve_uneven_f32r_f64
    assert f(*p, off.ctypes.data, n.ctypes.data, 2, 4, 1e-12, None, None, 0, None) == 1
    n[1] = 5
    assert f(*p, None, n.ctypes.data, 2, 4, 1e-12, None, None, 0, None) == 2
    assert f(*p, off.ctypes.data, n.ctypes.data, 0, 4, 1e-12, None, None, 0, None) == 1
    assert f(*p, off.ctypes.data, n.ctypes.data, 2, -1, 1e-12, None, None, 0, None) == 6
    assert f(*p, off.ctypes.data, n.ctypes.data, 2, 4, -1.0, None, None, 0, None) == 6
    need = lib.pcrp_uneven_workspace_size(off.ctypes.data, n.ctypes.data, 2, 1)
    assert need > 0
    assert f(*p, off.ctypes.data, n.ctypes.data, 2, 4, 1e-12, None, _fake(1)[0], need - 1,
             None) == 4
    big = np.array([5000], dtype=np.int32)
    # y copy (8 B/row) + level-0 block rows and the one pass round's output set
    # (n = 5000: 2500 block rows, 2 HBM levels), 56 B per block row each
    assert lib.pcrp_uneven_workspace_size(off.ctypes.data, big.ctypes.data, 1, 1) >= \
        8 * 5000 + 2 * 56 * 2500


def test_factor_argument_checks(lib):
    """Factor-once entry points: sizes, null pointers, unsupported n, factor
    buffer size; all rejected before any launch."""
    p = _fake(5)
    F = _fake(1)[0]
    # (4 floor(log2 n) + 3) n doubles per system
    assert lib.pcrp_factor_size(512, 10, 0, 0) == 10 * (4 * 9 + 3) * 512 * 8
    assert lib.pcrp_factor_size(512, 10, 0, 1) == (4 * 9 + 3) * 512 * 8
    assert lib.pcrp_factor_size(1025, 10, 0, 0) == 0          # FP64 factor: n <= 1024
    assert lib.pcrp_factor_size(2048, 3, 1, 0) > 0 and lib.pcrp_factor_size(2049, 3, 1, 0) == 0
    assert lib.pcrp_factor_size(100, 4, 0, 2) == 0
    need = lib.pcrp_factor_size(100, 4, 0, 0)
    assert lib.pcrp_factor_f64(*p, 0, 4, 1, 0, F, need, None) == 1
    assert lib.pcrp_factor_f64(*p, 100, 4, 1, 0, None, need, None) == 2
    assert lib.pcrp_factor_f64(*p, 100, 4, 1, 0, F, need - 1, None) == 4
    assert lib.pcrp_factor_f64(*p, 100, 4, 7, 0, F, need, None) == 3
    assert lib.pcrp_factor_f64(*p, 100, 4, 1, 3, F, need, None) == 6
    assert lib.pcrp_factor_f64(*p, 1025, 4, 1, 0, F, 1 << 30, None) == 7
    assert lib.pcrp_solve_factored_f64(F, None, 100, 4, 1, 0, None) == 2
    assert lib.pcrp_solve_factored_f64(F, _fake(1)[0], 100, 0, 1, 0, None) == 1
    assert lib.pcrp_solve_factored_f64(F, _fake(1)[0], 100, 4, 9, 0, None) == 3
    need32 = lib.pcrp_factor_size(100, 4, 1, 0)
    assert lib.pcrp_factor_f32r(*p, 100, 4, 1, 0, F, need32 - 1, None) == 4
    assert lib.pcrp_factor_f32r(*p, 4000, 4, 1, 0, F, 1 << 30, None) == 7
    y = _fake(1)[0]
    f = lib.pcrp_solve_factored_f32r_f64
    assert f(F, *p, y, 100, 4, 1, 0, -1, 1e-12, None, None) == 6
    assert f(F, *p, None, 100, 4, 1, 0, 4, 1e-12, None, None) == 2
    assert f(F, *p, y, 100, 4, 1, 0, 4, float("nan"), None, None) == 6


def test_tuning_tables_are_consistent_under_concurrent_use(lib):
    """Writers alternate two complete settings of the split and FP32 tunables
    while readers run: every read sees one setting or the other, never a mix
    (the tables are read by every solve, possibly while another thread
    tunes). No GPU: the tunables are host state."""
    import ctypes
    import threading
    import pcrpenta
    split0, f32r0 = pcrpenta.split_tune_get()[:3], pcrpenta.f32r_tune_get()
    split_ok = {(1024, 4, 0), (512, 3, 2)}
    f32r_ok = {(4, 1 << 23), (6, 12345)}
    stop = threading.Event()
    bad = []

    def writer():
        k = 0
        while not stop.is_set():
            lm, kf, sh = sorted(split_ok)[k % 2]
            lib.pcrp_split_tune_set(0, lm, kf, sh)
            wl, wm = sorted(f32r_ok)[k % 2]
            lib.pcrp_f32r_tune_set(wl, wm)
            k += 1

    def reader():
        vals = [ctypes.c_int() for _ in range(4)]
        wl, wm = ctypes.c_int(), ctypes.c_longlong()
        for _ in range(20000):
            lib.pcrp_split_tune_get(0, *[ctypes.byref(v) for v in vals])
            if tuple(v.value for v in vals[:3]) not in split_ok:
                bad.append(tuple(v.value for v in vals[:3]))
            lib.pcrp_f32r_tune_get(ctypes.byref(wl), ctypes.byref(wm))
            if (wl.value, wm.value) not in f32r_ok:
                bad.append((wl.value, wm.value))

    try:
        pcrpenta.split_tune_set(1024, 4, 0)
        pcrpenta.f32r_tune_set(4, 1 << 23)
        ws = [threading.Thread(target=writer) for _ in range(2)]
        rs = [threading.Thread(target=reader) for _ in range(4)]
        for t in ws + rs:
            t.start()
        for t in rs:
            t.join()
        stop.set()
        for t in ws:
            t.join()
    finally:
        pcrpenta.split_tune_set(*split0)
        pcrpenta.f32r_tune_set(*f32r0)
    assert not bad, bad[:5]


def test_missing_library_fails_loudly(tmp_path):
    """Without the built library the binding raises a plain error that names
    the file (the product path never falls back to anything else)."""
    import pcrpenta
    saved_path, saved_lib = pcrpenta.LIB_PATH, pcrpenta._lib
    try:
        pcrpenta.LIB_PATH = str(tmp_path / "libpcrpenta.so")
        pcrpenta._lib = None
        with pytest.raises(FileNotFoundError, match="libpcrpenta.so missing"):
            pcrpenta.load()
        err = pcrpenta.PCRPError(5, "pcrp_solve_f64")   # no library: the status name only
        assert "PCRP_ERR_HIP" in str(err)
    finally:
        pcrpenta.LIB_PATH, pcrpenta._lib = saved_path, saved_lib
