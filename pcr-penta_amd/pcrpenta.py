"""Python host binding of libpcrpenta.so (the gfx950 PCR-Penta solver).

Mirrors the reference's solve interface — six arrays (a, b, c, d, e, y) =
(sub-2, sub-1, main, super-1, super-2, rhs), y overwritten by the solution
(performanalysis.h:53, kernel.cu:32-39) — over torch device tensors. torch
only provides device memory and the stream; every solve runs the hand-written
HIP kernels through the C ABI of include/pcr_penta.h. There is no fallback:
if the library is missing or a call fails, a PCRPError is raised.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpcrpenta.so")

CONTIGUOUS = 0   # element i of system j at j*n + i  (kernel.cu:43)
INTERLEAVED = 1  # element i of system j at i*m + j  (Annexure A p.2)
PREC_F64 = 0
PREC_F32_REFINE = 1

STATUS = {
    0: "PCRP_OK", 1: "PCRP_ERR_INVALID_SIZE", 2: "PCRP_ERR_NULL", 3: "PCRP_ERR_LAYOUT",
    4: "PCRP_ERR_WORKSPACE", 5: "PCRP_ERR_HIP", 6: "PCRP_ERR_ARG", 7: "PCRP_ERR_UNSUPPORTED",
    8: "PCRP_ERR_SCRATCH",
}


class PCRPError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        # the library's own text when it is loaded (never load it from here)
        text = strerror(status) if _lib is not None else STATUS.get(status, str(status))
        super().__init__(f"{what}: {STATUS.get(status, status)} ({text})")


_lib = None
_vp = ctypes.c_void_p


def build() -> str:
    """Compile libpcrpenta.so for gfx950 with make (always recompiles stale objects)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)
    return LIB_PATH


def _declare(lib):
    lib.pcrp_abi_version.restype = ctypes.c_int
    lib.pcrp_strerror.argtypes = [ctypes.c_int]
    lib.pcrp_strerror.restype = ctypes.c_char_p
    lib.pcrp_workspace_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.pcrp_workspace_size.restype = ctypes.c_size_t
    lib.pcrp_solve_f64.argtypes = [_vp] * 6 + [ctypes.c_int] * 3 + [_vp, ctypes.c_size_t, _vp]
    lib.pcrp_solve_f64.restype = ctypes.c_int
    lib.pcrp_solve_f32r_f64.argtypes = ([_vp] * 6 + [ctypes.c_int] * 4 + [ctypes.c_double, _vp, _vp,
                                        ctypes.c_size_t, _vp])
    lib.pcrp_solve_f32r_f64.restype = ctypes.c_int
    lib.pcrp_generate.argtypes = ([ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int] + [_vp] * 6 + [_vp])
    lib.pcrp_generate.restype = ctypes.c_int
    lib.pcrp_init_t2.argtypes = [ctypes.c_int] * 3 + [_vp] * 6 + [_vp]
    lib.pcrp_init_t2.restype = ctypes.c_int
    lib.pcrp_lds_variant_count.argtypes = [ctypes.c_int]
    lib.pcrp_lds_variant_count.restype = ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int)
    lib.pcrp_lds_variant_info.argtypes = [ctypes.c_int, ctypes.c_int] + [ip] * 5
    lib.pcrp_lds_variant_info.restype = ctypes.c_int
    lib.pcrp_tune_set.argtypes = [ctypes.c_int] * 3
    lib.pcrp_tune_set.restype = ctypes.c_int
    lib.pcrp_tune_get.argtypes = [ctypes.c_int] * 2
    lib.pcrp_tune_get.restype = ctypes.c_int
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pcrp_autotune.argtypes = [ctypes.c_int] * 4 + [_vp] * 7 + [ctypes.c_int, _vp, ip, fp, fp]
    lib.pcrp_autotune.restype = ctypes.c_int
    lib.pcrp_split_tune_set.argtypes = [ctypes.c_int] * 4
    lib.pcrp_split_tune_set.restype = ctypes.c_int
    lib.pcrp_split_tune_get.argtypes = [ctypes.c_int] + [ip] * 4
    lib.pcrp_split_tune_get.restype = ctypes.c_int
    lib.pcrp_uneven_workspace_size.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int]
    lib.pcrp_uneven_workspace_size.restype = ctypes.c_size_t
    lib.pcrp_solve_uneven_f64.argtypes = [_vp] * 8 + [ctypes.c_int, _vp, ctypes.c_size_t, _vp]
    lib.pcrp_solve_uneven_f64.restype = ctypes.c_int
    lib.pcrp_solve_uneven_f32r_f64.argtypes = ([_vp] * 8 + [ctypes.c_int, ctypes.c_int,
                                                ctypes.c_double, _vp, _vp, ctypes.c_size_t, _vp])
    lib.pcrp_solve_uneven_f32r_f64.restype = ctypes.c_int
    lib.pcrp_generate_uneven.argtypes = [ctypes.c_int, ctypes.c_uint64] + [_vp] * 2 + \
        [ctypes.c_int] + [_vp] * 6 + [_vp]
    lib.pcrp_generate_uneven.restype = ctypes.c_int
    lib.pcrp_heston_workspace_size.argtypes = [_vp, ctypes.c_int]
    lib.pcrp_heston_workspace_size.restype = ctypes.c_size_t
    lib.pcrp_heston_solve.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_size_t, _vp]
    lib.pcrp_heston_solve.restype = ctypes.c_int
    lib.pcrp_split_autotune.argtypes = [ctypes.c_int] * 4 + [_vp] * 8 + \
        [ctypes.c_size_t, ctypes.c_int, _vp, ip, ip, fp]
    lib.pcrp_split_autotune.restype = ctypes.c_int
    lib.pcrp_factor_size.argtypes = [ctypes.c_int] * 4
    lib.pcrp_factor_size.restype = ctypes.c_size_t
    lib.pcrp_factor_f64.argtypes = [_vp] * 5 + [ctypes.c_int] * 4 + [_vp, ctypes.c_size_t, _vp]
    lib.pcrp_factor_f64.restype = ctypes.c_int
    lib.pcrp_solve_factored_f64.argtypes = [_vp, _vp] + [ctypes.c_int] * 4 + [_vp]
    lib.pcrp_solve_factored_f64.restype = ctypes.c_int
    lib.pcrp_factor_f32r.argtypes = lib.pcrp_factor_f64.argtypes
    lib.pcrp_factor_f32r.restype = ctypes.c_int
    lib.pcrp_solve_factored_f32r_f64.argtypes = ([_vp] * 7 + [ctypes.c_int] * 5 +
                                                 [ctypes.c_double, _vp, _vp])
    lib.pcrp_solve_factored_f32r_f64.restype = ctypes.c_int
    lib.pcrp_solve_f64_noalloc.argtypes = [_vp] * 6 + [ctypes.c_int] * 3 + [_vp]
    lib.pcrp_solve_f64_noalloc.restype = ctypes.c_int
    lib.pcrp_f32r_tune_set.argtypes = [ctypes.c_int, ctypes.c_longlong]
    lib.pcrp_f32r_tune_set.restype = ctypes.c_int
    lib.pcrp_f32r_tune_get.argtypes = [ip, ctypes.POINTER(ctypes.c_longlong)]
    lib.pcrp_f32r_tune_get.restype = ctypes.c_int
    lib.pcrp_f32r_autotune.argtypes = [ctypes.c_int] * 3 + [_vp] * 8 + \
        [ctypes.c_size_t, ctypes.c_int, _vp, ip, ip, fp]
    lib.pcrp_f32r_autotune.restype = ctypes.c_int
    lib.pcrp_plan_workspace_size.argtypes = [ctypes.c_int] * 3 + [_vp, _vp]
    lib.pcrp_plan_workspace_size.restype = ctypes.c_size_t
    lib.pcrp_plan_create.argtypes = ([ctypes.POINTER(_vp), ctypes.c_int] + [_vp] * 5 +
                                     [ctypes.c_int] * 3 + [_vp, _vp, _vp, ctypes.c_size_t, _vp])
    lib.pcrp_plan_create.restype = ctypes.c_int
    lib.pcrp_plan_execute.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp]
    lib.pcrp_plan_execute.restype = ctypes.c_int
    lib.pcrp_plan_route.argtypes = [_vp]
    lib.pcrp_plan_route.restype = ctypes.c_int
    lib.pcrp_plan_destroy.argtypes = [_vp]
    lib.pcrp_plan_destroy.restype = ctypes.c_int
    lib.pcrp_tune_save.argtypes = [ctypes.c_char_p]
    lib.pcrp_tune_save.restype = ctypes.c_int
    lib.pcrp_tune_load.argtypes = [ctypes.c_char_p]
    lib.pcrp_tune_load.restype = ctypes.c_int
    return lib


# PCRP_ABI_VERSION of include/pcr_penta.h this binding mirrors
ABI_VERSION = 2


def load():
    """Load the in-tree libpcrpenta.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing (run __graft_entry__.build())")
        lib = _declare(ctypes.CDLL(LIB_PATH))
        if lib.pcrp_abi_version() != ABI_VERSION:   # the structures mirrored below
            raise RuntimeError(f"{LIB_PATH}: ABI version {lib.pcrp_abi_version()}, "
                               f"this binding expects {ABI_VERSION} (rebuild the library)")
        _lib = lib
    return _lib


def strerror(status: int) -> str:
    return load().pcrp_strerror(int(status)).decode()


def _check(st: int, what: str):
    if st != 0:
        raise PCRPError(st, what)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _validate(arrs, n, m, dtype_name="float64"):
    import torch
    for t in arrs:
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError("pcrpenta: arrays must be CUDA (HIP) tensors")
        if str(t.dtype) != f"torch.{dtype_name}":
            raise TypeError(f"pcrpenta: arrays must be {dtype_name}")
        if not t.is_contiguous() or t.numel() != n * m:
            raise ValueError(f"pcrpenta: arrays must be contiguous with n*m = {n * m} elements")


def workspace_size(n: int, m: int, precision: int = PREC_F64) -> int:
    return int(load().pcrp_workspace_size(n, m, precision))


def _workspace(n, m, precision, device):
    import torch
    nb = workspace_size(n, m, precision)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=device)
    return ws, nb


def _keep_until_done(ws, stream):
    """A workspace allocated here on torch's current stream but used on another
    `stream` stays reserved by the caching allocator until that stream's work
    has completed (the solve is asynchronous; ws is dropped on return)."""
    if stream is not None:
        ws.record_stream(stream)


def solve(a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED, stream=None, ws=None):
    """Plain FP64 PCR-Penta: y <- A^{-1} y for m systems of size n (in place)."""
    _validate([a, b, c, d, e, y], n, m)
    own = ws is None
    if own:
        ws, nb = _workspace(n, m, PREC_F64, y.device)
    else:
        nb = ws.numel()
    st = load().pcrp_solve_f64(*[_ptr(t) for t in (a, b, c, d, e, y)], n, m, layout,
                               _ptr(ws), nb, _stream(stream))
    if own:
        _keep_until_done(ws, stream)
    _check(st, "pcrp_solve_f64")
    return y


def solve_refined(a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED, max_sweeps: int = 8,
                  tol: float = 1e-13, sweeps=None, stream=None, ws=None):
    """FP32 PCR-Penta + FP64 iterative refinement (in place on y).
    ``sweeps``: optional int32 device tensor of m entries for per-system sweep counts."""
    _validate([a, b, c, d, e, y], n, m)
    own = ws is None
    if own:
        ws, nb = _workspace(n, m, PREC_F32_REFINE, y.device)
    else:
        nb = ws.numel()
    sp = _ptr(sweeps) if sweeps is not None else None
    st = load().pcrp_solve_f32r_f64(*[_ptr(t) for t in (a, b, c, d, e, y)], n, m, layout,
                                    max_sweeps, tol, sp, _ptr(ws), nb, _stream(stream))
    if own:
        _keep_until_done(ws, stream)
    _check(st, "pcrp_solve_f32r_f64")
    return y


# ---- plan / execute ------------------------------------------------------------

ROUTE_LDS, ROUTE_SPLIT, ROUTE_WINDOW, ROUTE_UNEVEN_F64, ROUTE_UNEVEN_STAGE = range(5)


class Plan:
    """A solve plan (include/pcr_penta.h): matrix-dependent setup once, then
    executes that only launch kernels (capturable in a CUDA/HIP graph)."""

    def __init__(self, a, b, c, d, e, n: int = 0, m: int = 0, layout: int = INTERLEAVED,
                 precision: int = PREC_F64, offsets=None, sizes=None, stream=None):
        import numpy as np
        import torch
        lib = load()
        self._lib = lib
        self.precision = precision
        if offsets is not None:
            self._off = np.ascontiguousarray(offsets, dtype=np.int64)
            self._n = np.ascontiguousarray(sizes, dtype=np.int32)
            m = len(self._n)
            po, pn = self._off.ctypes.data, self._n.ctypes.data
        else:
            _validate([a, b, c, d, e], n, m)
            po = pn = None
        nb = int(lib.pcrp_plan_workspace_size(precision, n, m, po, pn))
        self.ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=a.device)
        self._keep = (a, b, c, d, e)   # the plan binds these arrays
        h = _vp()
        st = lib.pcrp_plan_create(ctypes.byref(h), precision, *[_ptr(t) for t in (a, b, c, d, e)],
                                  n, m, layout, po, pn, _ptr(self.ws), self.ws.numel(),
                                  _stream(stream))
        _check(st, "pcrp_plan_create")
        self._h = h
        self.n, self.m = n, m

    @property
    def route(self) -> int:
        return self._lib.pcrp_plan_route(self._h)

    def execute(self, y, max_sweeps: int = 8, tol: float = 1e-12, sweeps=None, stream=None):
        sp = _ptr(sweeps) if sweeps is not None else None
        st = self._lib.pcrp_plan_execute(self._h, _ptr(y), max_sweeps, tol, sp, _stream(stream))
        _check(st, "pcrp_plan_execute")
        return y

    def destroy(self):
        if getattr(self, "_h", None):
            self._lib.pcrp_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- factor once, solve many -------------------------------------------------

def factor_size(n: int, m: int, precision: int = PREC_F64, shared: bool = False) -> int:
    return int(load().pcrp_factor_size(n, m, precision, int(bool(shared))))


def factor(a, b, c, d, e, n: int, m: int, layout: int = INTERLEAVED, precision: int = PREC_F64,
           shared: bool = False, stream=None, out=None):
    """Factor m systems (or, with shared=True, the one system of n rows held by
    a..e) once; returns the factor as a uint8 device tensor for solve_factored."""
    import torch
    msys = 1 if shared else m
    _validate([a, b, c, d, e], n, msys)
    nb = factor_size(n, m, precision, shared)
    if nb == 0:
        raise PCRPError(7, f"factor: n = {n} not supported for precision {precision}")
    F = out if out is not None else torch.empty(nb, dtype=torch.uint8, device=a.device)
    fn = load().pcrp_factor_f64 if precision == PREC_F64 else load().pcrp_factor_f32r
    st = fn(*[_ptr(t) for t in (a, b, c, d, e)], n, m, layout, int(bool(shared)), _ptr(F),
            F.numel(), _stream(stream))
    _check(st, "pcrp_factor")
    return F


def solve_factored(F, y, n: int, m: int, layout: int = INTERLEAVED, shared: bool = False,
                   stream=None):
    """FP64: y <- A^{-1} y from a factor of factor(..., precision=PREC_F64)."""
    _validate([y], n, m)
    st = load().pcrp_solve_factored_f64(_ptr(F), _ptr(y), n, m, layout, int(bool(shared)),
                                        _stream(stream))
    _check(st, "pcrp_solve_factored_f64")
    return y


def solve_factored_refined(F, a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED,
                           shared: bool = False, max_sweeps: int = 8, tol: float = 1e-12,
                           sweeps=None, stream=None):
    """FP32 + FP64 refinement from a factor of factor(..., precision=PREC_F32_REFINE);
    a..e as given to factor() (the FP64 residuals need the matrix)."""
    _validate([y], n, m)
    _validate([a, b, c, d, e], n, 1 if shared else m)
    sp = _ptr(sweeps) if sweeps is not None else None
    st = load().pcrp_solve_factored_f32r_f64(_ptr(F), *[_ptr(t) for t in (a, b, c, d, e, y)], n, m,
                                             layout, int(bool(shared)), max_sweeps, tol, sp,
                                             _stream(stream))
    _check(st, "pcrp_solve_factored_f32r_f64")
    return y


def generate(kind: int, seed: int, n: int, m: int, layout: int = INTERLEAVED, device="cuda",
             stream=None):
    """Seeded batch on the device (same values as the oracle's generator)."""
    import torch
    arrs = [torch.empty(n * m, dtype=torch.float64, device=device) for _ in range(6)]
    st = load().pcrp_generate(kind, seed, n, m, layout, *[_ptr(t) for t in arrs], _stream(stream))
    _check(st, "pcrp_generate")
    return arrs


def init_t2(n: int, m: int, layout: int = INTERLEAVED, device="cuda", stream=None):
    """The reference's T^2 test batch (kernel.cu:5-26) on the device."""
    import torch
    arrs = [torch.empty(n * m, dtype=torch.float64, device=device) for _ in range(6)]
    st = load().pcrp_init_t2(n, m, layout, *[_ptr(t) for t in arrs], _stream(stream))
    _check(st, "pcrp_init_t2")
    return arrs


# ---- autotuner -------------------------------------------------------------

def variants(precision: int = PREC_F64):
    """List of dicts describing the compiled LDS-kernel variants."""
    lib = load()
    out = []
    for vid in range(lib.pcrp_lds_variant_count(precision)):
        vals = [ctypes.c_int() for _ in range(5)]
        _check(lib.pcrp_lds_variant_info(precision, vid, *[ctypes.byref(v) for v in vals]),
               "pcrp_lds_variant_info")
        kind, nt, rpt, max_n, lds = (v.value for v in vals)
        prefix = "paired-" if kind >= 8 else ("persistent-" if kind >= 4 else "")
        out.append({"id": vid, "kind": prefix + ("shared", "redundant", "inplace", "spaced")[kind & 3], "nt": nt, "rpt": rpt,
                    "max_n": max_n, "lds_bytes": lds})
    return out


def tune_set(n: int, variant_id: int, precision: int = PREC_F64):
    _check(load().pcrp_tune_set(precision, n, variant_id), "pcrp_tune_set")


def tune_get(n: int, precision: int = PREC_F64) -> int:
    return load().pcrp_tune_get(precision, n)


def autotune(a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED, reps: int = 3,
             precision: int = PREC_F64, stream=None):
    """Time every fitting variant on this batch; returns (best_id, best_ms, times_ms)."""
    import torch
    _validate([a, b, c, d, e, y], n, m)
    lib = load()
    work = torch.empty_like(y)
    nv = lib.pcrp_lds_variant_count(precision)
    times = (ctypes.c_float * max(nv, 1))()
    best = ctypes.c_int(-1)
    best_ms = ctypes.c_float(0)
    st = lib.pcrp_autotune(precision, n, m, layout, *[_ptr(t) for t in (a, b, c, d, e, y)],
                           _ptr(work), reps, _stream(stream), ctypes.byref(best),
                           ctypes.byref(best_ms), times)
    _check(st, "pcrp_autotune")
    return best.value, best_ms.value, [times[i] for i in range(nv)]


def split_autotune(a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED, reps: int = 2,
                   stream=None, ws=None):
    """Time every (steps per pass, tile shape) of the split path; keeps the fastest.
    Returns (k_fuse, shape, ms)."""
    import torch
    _validate([a, b, c, d, e, y], n, m)
    if ws is None:
        ws, nb = _workspace(n, m, PREC_F64, y.device)
    else:
        nb = ws.numel()
    work = torch.empty_like(y)
    k, sh, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
    st = load().pcrp_split_autotune(PREC_F64, n, m, layout,
                                    *[_ptr(t) for t in (a, b, c, d, e, y)], _ptr(work),
                                    _ptr(ws), nb, reps, _stream(stream), ctypes.byref(k),
                                    ctypes.byref(sh), ctypes.byref(ms))
    _check(st, "pcrp_split_autotune")
    return k.value, sh.value, ms.value


def f32r_tune_get():
    """(win_levels, window_min_unknowns) of the FP32 refinement paths."""
    wl, wm = ctypes.c_int(), ctypes.c_longlong()
    _check(load().pcrp_f32r_tune_get(ctypes.byref(wl), ctypes.byref(wm)), "pcrp_f32r_tune_get")
    return wl.value, wm.value


def f32r_tune_set(win_levels: int, window_min_unknowns: int):
    _check(load().pcrp_f32r_tune_set(win_levels, window_min_unknowns), "pcrp_f32r_tune_set")


def f32r_autotune(a, b, c, d, e, y, n: int, m: int, layout: int = INTERLEAVED, reps: int = 2,
                  stream=None, ws=None):
    """Time the window levels (and, for n <= 2048, the one-workgroup kernel) on this
    batch; keeps the fastest. Returns (win_levels, window_path, ms)."""
    import torch
    _validate([a, b, c, d, e, y], n, m)
    if ws is None:
        ws, nb = _workspace(n, m, PREC_F32_REFINE, y.device)
    else:
        nb = ws.numel()
    work = torch.empty_like(y)
    wl, wp, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
    st = load().pcrp_f32r_autotune(n, m, layout, *[_ptr(t) for t in (a, b, c, d, e, y)],
                                   _ptr(work), _ptr(ws), nb, reps, _stream(stream),
                                   ctypes.byref(wl), ctypes.byref(wp), ctypes.byref(ms))
    _check(st, "pcrp_f32r_autotune")
    return wl.value, wp.value, ms.value


def tune_save(path: str):
    _check(load().pcrp_tune_save(path.encode()), "pcrp_tune_save")


def tune_load(path: str):
    _check(load().pcrp_tune_load(path.encode()), "pcrp_tune_load")


def split_tune_get(precision: int = PREC_F64):
    """(lds_max, k_fuse, shape, n_shapes) of the split path."""
    vals = [ctypes.c_int() for _ in range(4)]
    _check(load().pcrp_split_tune_get(precision, *[ctypes.byref(v) for v in vals]),
           "pcrp_split_tune_get")
    return tuple(v.value for v in vals)


def split_tune_set(lds_max: int, k_fuse: int, shape: int, precision: int = PREC_F64):
    _check(load().pcrp_split_tune_set(precision, lds_max, k_fuse, shape), "pcrp_split_tune_set")


# ---- uneven batches ----------------------------------------------------------

def uneven_layout(sizes):
    """(offsets int64, n int32, total rows) for systems stored back to back."""
    import numpy as np
    n = np.ascontiguousarray(sizes, dtype=np.int32)
    off = np.zeros(len(n), dtype=np.int64)
    if len(n) > 1:
        off[1:] = np.cumsum(n[:-1], dtype=np.int64)
    return off, n, int(n.astype(np.int64).sum())


def uneven_workspace_size(offsets, sizes, precision: int = PREC_F64) -> int:
    import numpy as np
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n = np.ascontiguousarray(sizes, dtype=np.int32)
    return int(load().pcrp_uneven_workspace_size(off.ctypes.data, n.ctypes.data, len(n),
                                                 precision))


def solve_uneven(a, b, c, d, e, y, offsets, sizes, stream=None, ws=None):
    """Plain FP64 solve of an uneven batch (system j at rows offsets[j] ..
    offsets[j]+sizes[j]-1 of flat device arrays), y overwritten in place."""
    import numpy as np
    import torch
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n = np.ascontiguousarray(sizes, dtype=np.int32)
    for t in (a, b, c, d, e, y):
        if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
            raise TypeError("pcrpenta: arrays must be contiguous float64 device tensors")
    own = ws is None
    if own:
        nb = uneven_workspace_size(off, n)
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=y.device)
    st = load().pcrp_solve_uneven_f64(*[_ptr(t) for t in (a, b, c, d, e, y)], off.ctypes.data,
                                      n.ctypes.data, len(n), _ptr(ws), ws.numel(), _stream(stream))
    if own:
        _keep_until_done(ws, stream)
    _check(st, "pcrp_solve_uneven_f64")
    return y


def solve_uneven_refined(a, b, c, d, e, y, offsets, sizes, max_sweeps: int = 8,
                         tol: float = 1e-12, sweeps=None, stream=None, ws=None):
    """FP32 inner solve + FP64 refinement of an uneven batch, y overwritten in
    place. ``sweeps``: optional int32 device tensor of len(sizes) entries."""
    import numpy as np
    import torch
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n = np.ascontiguousarray(sizes, dtype=np.int32)
    for t in (a, b, c, d, e, y):
        if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
            raise TypeError("pcrpenta: arrays must be contiguous float64 device tensors")
    own = ws is None
    if own:
        nb = uneven_workspace_size(off, n, PREC_F32_REFINE)
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=y.device)
    sp = _ptr(sweeps) if sweeps is not None else None
    st = load().pcrp_solve_uneven_f32r_f64(*[_ptr(t) for t in (a, b, c, d, e, y)],
                                           off.ctypes.data, n.ctypes.data, len(n), max_sweeps,
                                           tol, sp, _ptr(ws), ws.numel(), _stream(stream))
    if own:
        _keep_until_done(ws, stream)
    _check(st, "pcrp_solve_uneven_f32r_f64")
    return y


def generate_uneven(kind: int, seed: int, offsets, sizes, device="cuda", stream=None):
    """Seeded uneven batch on the device (values as pcrp_generate for each system)."""
    import numpy as np
    import torch
    off = torch.from_numpy(np.ascontiguousarray(offsets, dtype=np.int64)).to(device)
    n = torch.from_numpy(np.ascontiguousarray(sizes, dtype=np.int32)).to(device)
    total = int(np.asarray(offsets, dtype=np.int64)[-1] + np.asarray(sizes, dtype=np.int64)[-1])
    arrs = [torch.empty(total, dtype=torch.float64, device=device) for _ in range(6)]
    st = load().pcrp_generate_uneven(kind, seed, _ptr(off), _ptr(n), len(n),
                                     *[_ptr(t) for t in arrs], _stream(stream))
    _check(st, "pcrp_generate_uneven")
    torch.cuda.current_stream().synchronize()
    return arrs


def uneven_log_uniform_sizes(m: int = 10000, seed: int = 1, lo: int = 64, hi: int = 131072):
    """Sizes of the uneven benchmark batch: n_j = floor(exp(U[ln lo, ln hi))), U drawn by the counter
    generator keyed (seed, 7, j, 0)."""
    import numpy as np
    M64 = (1 << 64) - 1

    def mix(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(M64)
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(M64)
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(M64)
        return z ^ (z >> np.uint64(31))
    with np.errstate(over="ignore"):
        key = mix(np.array([seed], dtype=np.uint64))[0]
        ctr = (np.uint64(7) << np.uint64(56)) | (np.arange(m, dtype=np.uint64) << np.uint64(24))
        u = mix((key + ctr) & np.uint64(M64))
    f = (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    ln_lo, ln_hi = np.log(lo), np.log(hi)
    return np.floor(np.exp(ln_lo + (ln_hi - ln_lo) * f)).astype(np.int32)


# ---- Heston ADI driver ---------------------------------------------------------

class HestonParams(ctypes.Structure):
    """Mirror of pcrp_heston_params (include/pcr_penta.h)."""
    _fields_ = [("ns", ctypes.c_int), ("nv", ctypes.c_int),
                ("smax", ctypes.c_double), ("vmax", ctypes.c_double),
                ("kappa", ctypes.c_double), ("theta_v", ctypes.c_double),
                ("sigma", ctypes.c_double), ("rho", ctypes.c_double),
                ("r", ctypes.c_double), ("strike", ctypes.c_double),
                ("dt", ctypes.c_double), ("steps", ctypes.c_int),
                ("theta_scheme", ctypes.c_double), ("reuse_factor", ctypes.c_int),
                ("separate_explicit", ctypes.c_int)]


# refinement settings of the FP32 line solves in the Heston driver
# (PCRP_HESTON_REFINE_TOL, PCRP_HESTON_MAX_SWEEPS in include/pcr_penta.h)
HESTON_REFINE_TOL = 1e-12
HESTON_MAX_SWEEPS = 8

# Heston ADI problem: Ns = 1024, Nv = 512 unknowns, 200 steps of 1/200
HESTON_GRID = dict(ns=1024, nv=512, smax=800.0, vmax=5.0, kappa=3.0, theta_v=0.04, sigma=0.15,
               rho=-0.7, r=0.03, strike=100.0, dt=1.0 / 200, steps=200, theta_scheme=0.5,
               reuse_factor=0, separate_explicit=0)


def heston_params(**kw) -> HestonParams:
    d = dict(HESTON_GRID)
    d.update(kw)
    return HestonParams(**d)


def heston(params: HestonParams, precision: int = PREC_F64, use_graph: bool = True, stream=None,
           ws=None, u=None):
    """Price grid u[j*Ns + i-1] at tau = steps*dt (device tensor)."""
    import torch
    lib = load()
    nb = int(lib.pcrp_heston_workspace_size(ctypes.byref(params), precision))
    if nb == 0:
        raise PCRPError(1, "pcrp_heston_workspace_size")
    own = ws is None
    if own:
        ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    if u is None:
        u = torch.empty(params.ns * params.nv, dtype=torch.float64, device="cuda")
    if use_graph and stream is None:
        raise ValueError("graph capture needs a non-default stream")
    st = lib.pcrp_heston_solve(ctypes.byref(params), precision, int(use_graph), _ptr(u), _ptr(ws),
                               ws.numel(), _stream(stream))
    if own:
        _keep_until_done(ws, stream)
    _check(st, "pcrp_heston_solve")
    return u
