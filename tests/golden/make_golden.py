"""Writes tests/golden/golden.npz: seeded pentadiagonal systems and their
LAPACK dgbsv (kl = ku = 2) solutions from scipy.linalg.lapack.dgbsv
(scipy 1.15.3, bundled OpenBLAS), plus tridiagonal cases solved by
scipy.linalg.solve_banded((1, 1)).

The PCR-Penta reference ships no test vectors, golden files or known-answer
tests (its driver never reads results back), so these fixtures are the
external anchor for the oracle. Inputs come from the oracle's counter-based
generator (oracle/pcr_oracle.c: oracle_gen_batch), re-derived here in numpy
so the fixture does not depend on the oracle build.

Run from the repo root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np
from scipy.linalg import solve_banded
from scipy.linalg.lapack import dgbsv

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def splitmix64(z: np.ndarray) -> np.ndarray:
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(M64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(M64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(M64)
    return z ^ (z >> np.uint64(31))


def gen_uniform(seed: int, diag: int, sys: np.ndarray, row: np.ndarray, lo: float, hi: float):
    """numpy restatement of oracle_gen_uniform (layout-free counter draw)."""
    with np.errstate(over="ignore"):
        key = splitmix64(np.array([seed], dtype=np.uint64))[0]
        ctr = (np.uint64(diag) << np.uint64(56)) | (sys.astype(np.uint64) << np.uint64(24)) \
            | row.astype(np.uint64)
        u = splitmix64((key + ctr) & np.uint64(M64))
    f = (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return lo + (hi - lo) * f


def gen_system(kind: int, seed: int, sys: int, n: int):
    off = 0.5 if kind == 1 else 1.0
    clo, chi = (2.2, 2.6) if kind == 1 else (4.5, 5.5)
    i = np.arange(n)
    s = np.full(n, sys)
    a = np.where(i >= 2, gen_uniform(seed, 0, s, i, -off, off), 0.0)
    b = np.where(i >= 1, gen_uniform(seed, 1, s, i, -off, off), 0.0)
    c = gen_uniform(seed, 2, s, i, clo, chi)
    d = np.where(i <= n - 2, gen_uniform(seed, 3, s, i, -off, off), 0.0)
    e = np.where(i <= n - 3, gen_uniform(seed, 4, s, i, -off, off), 0.0)
    y = gen_uniform(seed, 5, s, i, -1.0, 1.0)
    return [a, b, c, d, e, y]


def lapack_dgbsv(a, b, c, d, e, y):
    n = len(c)
    ab = np.zeros((7, n))
    i = np.arange(n)
    # ab[kl+ku+i-j, j] = A[i, j], kl = ku = 2
    ab[6, :n - 2] = a[2:]
    ab[5, :n - 1] = b[1:]
    ab[4, :] = c
    ab[3, 1:] = d[:n - 1]
    ab[2, 2:] = e[:n - 2]
    _, _, x, info = dgbsv(2, 2, ab, y.reshape(-1, 1).copy())
    assert info == 0, info
    del i
    return x[:, 0]


def main():
    cases = []
    # the m=256, n=512 dominant batch (seed 0), first 4 systems
    for j in range(4):
        cases.append(("c0", 0, 0, j, 512))
    # a spread of sizes, both generators, incl. non-powers of two (2x2 final branch)
    for n in (1, 2, 3, 4, 5, 6, 7, 8, 9, 15, 16, 17, 31, 33, 64, 100, 255, 257, 511, 1000, 1023,
              1024, 1025):
        cases.append(("dom", 0, 7, n, n))
        cases.append(("weak", 1, 11, n, n))
    out = {}
    names = []
    for idx, (tag, kind, seed, sysid, n) in enumerate(cases):
        arrs = gen_system(kind, seed, sysid, n)
        x = lapack_dgbsv(*arrs)
        key = f"{idx:03d}_{tag}_k{kind}_s{seed}_j{sysid}_n{n}"
        names.append(key)
        out[key + "_in"] = np.stack(arrs)
        out[key + "_x"] = x
        out[key + "_meta"] = np.array([kind, seed, sysid, n], dtype=np.int64)
    # tridiagonal special case (a = e = 0): textbook PCR, checked with solve_banded((1,1))
    for n in (2, 5, 16, 100, 513):
        arrs = gen_system(0, 5, n, n)
        arrs[0][:] = 0.0
        arrs[4][:] = 0.0
        a, b, c, d, e, y = arrs
        band = np.zeros((3, n))
        band[0, 1:] = d[:-1]
        band[1] = c
        band[2, :-1] = b[1:]
        x = solve_banded((1, 1), band, y)
        key = f"tri_n{n}"
        out[key + "_in"] = np.stack(arrs)
        out[key + "_x"] = x
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print(f"wrote {len(names)} dgbsv cases + 5 tridiagonal cases")


if __name__ == "__main__":
    main()
