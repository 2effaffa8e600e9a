"""ctypes binding of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, as the checker or as the timed CPU baseline. The shipped solver
never imports it.

Functions follow the reference as documented in pcr_oracle_body.inc:
``pcr`` restates pcr_penta (kernel.cu:32-154), ``gbsv`` is LAPACK dgbsv with
kl = ku = 2, ``gen_batch`` is the seeded generator and ``init_t2`` is the
reference's own test system (kernel.cu:5-26).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CONTIGUOUS = 0   # element i of system j at j*n + i   (kernel.cu:43)
INTERLEAVED = 1  # element i of system j at i*m + j   (Annexure A p.2)

_lib = None


def build(variant: str = "") -> str:
    """Compile liboracle.so (or a planted-mistake variant) with make."""
    target = "all" if not variant else os.path.join(HERE, f"liboracle_{variant}.so")
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)
    return LIB_PATH if not variant else target


_dp = ctypes.POINTER(ctypes.c_double)
_ldp = ctypes.POINTER(ctypes.c_longdouble)


def _declare(lib):
    lib.oracle_pcr_steps.argtypes = [ctypes.c_int]
    lib.oracle_pcr_steps.restype = ctypes.c_int
    lib.oracle_pcr_f64.argtypes = [_dp] * 7 + [ctypes.c_int, _dp]
    lib.oracle_pcr_f64.restype = None
    lib.oracle_pcr_f80.argtypes = [_ldp] * 7 + [ctypes.c_int, _ldp]
    lib.oracle_pcr_f80.restype = None
    lib.oracle_gbsv_f64.argtypes = [_dp] * 7 + [ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_gbsv_f64.restype = ctypes.c_int
    lib.oracle_gbsv_f80.argtypes = [_ldp] * 7 + [ctypes.c_int, _ldp, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_gbsv_f80.restype = ctypes.c_int
    lib.oracle_gen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
    lib.oracle_gen_uniform.restype = ctypes.c_double
    lib.oracle_gen_batch.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_long, ctypes.c_long] + [_dp] * 6
    lib.oracle_gen_batch.restype = None
    lib.oracle_init_t2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_long] + [_dp] * 6
    lib.oracle_init_t2.restype = None
    lib.oracle_pcr_batch_f64.argtypes = [_dp] * 7 + [ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                                      ctypes.c_long, ctypes.c_int]
    lib.oracle_pcr_batch_f64.restype = None
    lib.oracle_gbsv_batch_f64.argtypes = [_dp] * 7 + [ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                                       ctypes.c_long, ctypes.c_int]
    lib.oracle_gbsv_batch_f64.restype = ctypes.c_int
    lib.oracle_pcr_uneven_f64.argtypes = [_dp] * 7 + [ctypes.POINTER(ctypes.c_int64),
                                                       ctypes.POINTER(ctypes.c_int),
                                                       ctypes.c_int, ctypes.c_int]
    lib.oracle_pcr_uneven_f64.restype = None
    lib.oracle_gbsv_uneven_f64.argtypes = lib.oracle_pcr_uneven_f64.argtypes
    lib.oracle_gbsv_uneven_f64.restype = ctypes.c_int
    lib.oracle_max_threads.argtypes = []
    lib.oracle_max_threads.restype = ctypes.c_int
    return lib


def load(path: str | None = None):
    """Load liboracle.so (building it if absent). ``path`` selects a variant."""
    global _lib
    if path is not None:
        return _declare(ctypes.CDLL(path))
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def _p(x):
    return x.ctypes.data_as(_dp)


def _lp(x):
    return x.ctypes.data_as(_ldp)


def steps(n: int) -> int:
    return load().oracle_pcr_steps(int(n))


def pcr(a, b, c, d, e, y, lib=None):
    """PCR-Penta on one system in FP64 (bit-reproducible)."""
    lib = lib or load()
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d, e, y)]
    n = arrs[0].size
    x = np.zeros(n)
    work = np.empty(12 * max(n, 1))
    lib.oracle_pcr_f64(*[_p(v) for v in arrs], _p(x), n, _p(work))
    return x


def pcr80(a, b, c, d, e, y, lib=None):
    """The same PCR-Penta code in x87 80-bit long double; returns float64."""
    lib = lib or load()
    arrs = [np.ascontiguousarray(v, dtype=np.longdouble) for v in (a, b, c, d, e, y)]
    n = arrs[0].size
    x = np.zeros(n, dtype=np.longdouble)
    work = np.empty(12 * max(n, 1), dtype=np.longdouble)
    lib.oracle_pcr_f80(*[_lp(v) for v in arrs], _lp(x), n, _lp(work))
    return x.astype(np.float64)


def gbsv(a, b, c, d, e, y, lib=None, extended=False):
    """Banded LU with partial pivoting (dgbsv algorithm, kl = ku = 2)."""
    lib = lib or load()
    dt = np.longdouble if extended else np.float64
    arrs = [np.ascontiguousarray(v, dtype=dt) for v in (a, b, c, d, e, y)]
    n = arrs[0].size
    x = np.zeros(n, dtype=dt)
    ab = np.empty(7 * max(n, 1), dtype=dt)
    ipiv = np.empty(max(n, 1), dtype=np.int32)
    fn = lib.oracle_gbsv_f80 if extended else lib.oracle_gbsv_f64
    cv = _lp if extended else _p
    info = fn(*[cv(v) for v in arrs], cv(x), n, cv(ab), ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if info != 0:
        raise np.linalg.LinAlgError(f"dgbsv: zero pivot U({info},{info})")
    return x.astype(np.float64)


def strides(n: int, m: int, layout: int):
    """(row stride, system stride) of a batch layout."""
    return (1, n) if layout == CONTIGUOUS else (m, 1)


def gen_batch(kind: int, seed: int, n: int, m: int, layout: int = INTERLEAVED):
    """Six flat float64 arrays (a, b, c, d, e, y) of a seeded batch."""
    rs, ss = strides(n, m, layout)
    arrs = [np.empty(n * m) for _ in range(6)]
    load().oracle_gen_batch(kind, seed, n, m, rs, ss, *[_p(v) for v in arrs])
    return arrs


def init_t2(n: int, m: int, layout: int = INTERLEAVED):
    rs, ss = strides(n, m, layout)
    arrs = [np.empty(n * m) for _ in range(6)]
    load().oracle_init_t2(n, m, rs, ss, *[_p(v) for v in arrs])
    return arrs


def pcr_batch(a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED, nthreads: int = 0):
    rs, ss = strides(n, m, layout)
    x = np.zeros(n * m)
    load().oracle_pcr_batch_f64(*[_p(np.ascontiguousarray(v)) for v in (a, b, c, d, e, y)],
                                _p(x), n, m, rs, ss, nthreads)
    return x


def gbsv_batch(a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED, nthreads: int = 0):
    rs, ss = strides(n, m, layout)
    x = np.zeros(n * m)
    bad = load().oracle_gbsv_batch_f64(*[_p(np.ascontiguousarray(v)) for v in (a, b, c, d, e, y)],
                                       _p(x), n, m, rs, ss, nthreads)
    if bad:
        raise np.linalg.LinAlgError(f"dgbsv: {bad} singular systems")
    return x


def pcr_uneven(a, b, c, d, e, y, offsets, n_sys, nthreads: int = 0):
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    ns = np.ascontiguousarray(n_sys, dtype=np.int32)
    x = np.zeros(len(a))
    load().oracle_pcr_uneven_f64(*[_p(np.ascontiguousarray(v)) for v in (a, b, c, d, e, y)], _p(x),
                                 off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(ns), nthreads)
    return x


def gbsv_uneven(a, b, c, d, e, y, offsets, n_sys, nthreads: int = 0):
    """dgbsv (kl = ku = 2) of every system of an uneven batch."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    ns = np.ascontiguousarray(n_sys, dtype=np.int32)
    x = np.zeros(len(a))
    bad = load().oracle_gbsv_uneven_f64(
        *[_p(np.ascontiguousarray(v)) for v in (a, b, c, d, e, y)], _p(x),
        off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(ns), nthreads)
    if bad:
        raise np.linalg.LinAlgError(f"dgbsv: {bad} singular systems")
    return x


def uneven_relerr(x, ref, offsets, n_sys):
    """Per-system ||x - ref||_inf / ||ref||_inf of an uneven batch (array of m)."""
    off = np.asarray(offsets, dtype=np.int64)
    ns = np.asarray(n_sys, dtype=np.int64)
    end = int(off[-1] + ns[-1])
    num = np.maximum.reduceat(np.abs(x[:end] - ref[:end]), off)
    den = np.maximum.reduceat(np.abs(ref[:end]), off)
    # reduceat over a gap or an empty system would be wrong: systems are back to back
    assert np.array_equal(off[1:], off[:-1] + ns[:-1])
    return num / np.where(den == 0, 1.0, den)


def load_native(outdir: str):
    """Build this oracle for the host it runs on (gcc -O3 -march=native, used
    for the CPU baseline only: -march=native code cannot travel between
    machines) into outdir and make it the library every function here uses.
    The explicit fma() calls are still the only fused operations, so results
    keep the same bits."""
    global _lib
    path = os.path.join(outdir, "liboracle_native.so")
    subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={path}"], check=True)
    _lib = _declare(ctypes.CDLL(path))
    return path


def max_threads() -> int:
    return load().oracle_max_threads()


def system(arrs, j: int, n: int, m: int, layout: int = INTERLEAVED):
    """Extract system j of a strided batch as six contiguous arrays."""
    rs, ss = strides(n, m, layout)
    idx = j * ss + np.arange(n) * rs
    return [v[idx] for v in arrs]


def normwise_relerr(x, ref, n: int, m: int, layout: int = INTERLEAVED) -> float:
    """max over systems of ||x - ref||_inf / ||ref||_inf."""
    rs, ss = strides(n, m, layout)
    if layout == CONTIGUOUS:
        X, R = x.reshape(m, n), ref.reshape(m, n)
        num, den = np.abs(X - R).max(axis=1), np.abs(R).max(axis=1)
    else:
        X, R = x.reshape(n, m), ref.reshape(n, m)
        num, den = np.abs(X - R).max(axis=0), np.abs(R).max(axis=0)
    return float((num / np.where(den == 0, 1.0, den)).max())
