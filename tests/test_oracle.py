"""Pins the CPU oracle (oracle/pcr_oracle_body.inc) before anything is checked
against it. The reference ships no tests or golden vectors, so the anchors are
LAPACK dgbsv / solve_banded fixtures made with scipy (tests/golden), closed
forms, special cases that reduce to textbook routines, and brute force.
Each check is chosen so that a dropped term, a wrong sign or a wrong index in
the oracle fails it; test_planted_mistakes_are_caught shows that directly."""
import numpy as np
import pytest


def _cases(golden):
    for key in golden["names"]:
        key = str(key)
        yield key, golden[key + "_in"], golden[key + "_x"]


def _relerr(x, ref):
    den = np.abs(ref).max()
    return float(np.abs(x - ref).max() / (den if den else 1.0))


def test_step_count_rule(oracle_lib):
    """main.cu:39 k2 = ceil((float)log(n)/log(2)); the loop at kernel.cu:63 runs
    k2-1 times. For every n <= 1024 that equals floor(log2 n)."""
    import math
    for n in range(2, 1025):
        k2 = math.ceil(float(np.float32(math.log(n))) / math.log(2))
        assert k2 - 1 == oracle_lib.steps(n), n
    # split path: s = log2(n/1024) stage-1 steps + k1-1 = 10 stage-2 steps (main.cu:19-38)
    for s in range(0, 8):
        assert oracle_lib.steps(1024 << s) == s + 10
    assert oracle_lib.steps(1) == 0


def test_generator_matches_numpy_restatement(oracle_lib):
    import make_golden as G
    for kind, seed, n, m in ((0, 0, 512, 3), (1, 11, 37, 5)):
        for layout in (oracle_lib.CONTIGUOUS, oracle_lib.INTERLEAVED):
            arrs = oracle_lib.gen_batch(kind, seed, n, m, layout)
            for j in range(m):
                ref = G.gen_system(kind, seed, j, n)
                got = oracle_lib.system(arrs, j, n, m, layout)
                for u, v in zip(got, ref):
                    assert np.array_equal(u, v)


@pytest.mark.parametrize("kind,off,clo,chi", [(0, 1.0, 4.5, 5.5), (1, 0.5, 2.2, 2.6)])
def test_generator_follows_config_distributions(oracle_lib, kind, off, clo, chi):
    """The seeded inputs against the configs' stated distributions, not
    against a copy of the generator's formula: off-diagonals U[-off, off]
    with out-of-range entries exactly zero, main diagonal U[clo, chi], RHS
    U[-1, 1] (Kolmogorov-Smirnov against each uniform, mean and variance),
    every row strictly diagonally dominant, systems and diagonals drawn
    independently, the same seed reproducing the same batch."""
    from scipy import stats
    n, m = 4096, 8
    arrs = oracle_lib.gen_batch(kind, 7, n, m, oracle_lib.INTERLEAVED)
    sys_ = [oracle_lib.system(arrs, j, n, m, oracle_lib.INTERLEAVED) for j in range(m)]
    for a, b, c, d, e, y in sys_:
        assert a[0] == 0.0 and a[1] == 0.0 and b[0] == 0.0
        assert d[n - 1] == 0.0 and e[n - 2] == 0.0 and e[n - 1] == 0.0
        assert np.all(np.abs(b) + np.abs(d) + np.abs(a) + np.abs(e) < c)
    pools = {
        "a": (np.concatenate([s[0][2:] for s in sys_]), -off, off),
        "b": (np.concatenate([s[1][1:] for s in sys_]), -off, off),
        "c": (np.concatenate([s[2] for s in sys_]), clo, chi),
        "d": (np.concatenate([s[3][:-1] for s in sys_]), -off, off),
        "e": (np.concatenate([s[4][:-2] for s in sys_]), -off, off),
        "y": (np.concatenate([s[5] for s in sys_]), -1.0, 1.0),
    }
    for name, (v, lo, hi) in pools.items():
        assert v.min() >= lo and v.max() <= hi, name
        assert stats.kstest(v, "uniform", args=(lo, hi - lo)).pvalue > 1e-4, name
        N = len(v)
        assert abs(v.mean() - (lo + hi) / 2) < 5 * (hi - lo) / np.sqrt(12 * N), name
        assert abs(v.var() / ((hi - lo) ** 2 / 12) - 1) < 0.05, name
    # independent systems and diagonals: correlations near zero
    assert abs(np.corrcoef(sys_[0][5], sys_[1][5])[0, 1]) < 0.08
    assert abs(np.corrcoef(sys_[0][1][1:], sys_[0][3][1:])[0, 1]) < 0.08
    again = oracle_lib.gen_batch(kind, 7, n, m, oracle_lib.INTERLEAVED)
    assert all(np.array_equal(u, v) for u, v in zip(arrs, again))
    other = oracle_lib.gen_batch(kind, 8, n, m, oracle_lib.INTERLEAVED)
    assert not np.array_equal(arrs[5], other[5])


def test_pcr80_matches_dgbsv(oracle_lib, golden):
    """The same oracle source in 80-bit arithmetic agrees with LAPACK dgbsv to
    ~1e-16: the recurrences (terms, signs, indices, guards, step count, final
    2x2/1x1 branches) are exact, and what FP64 adds is rounding only."""
    worst = 0.0
    for key, arrs, x in _cases(golden):
        got = oracle_lib.pcr80(*arrs)
        err = _relerr(got, x)
        worst = max(worst, err)
        assert err < 1e-14, (key, err)
    assert worst < 1e-14


def test_pcr64_within_rounding_of_dgbsv(oracle_lib, golden):
    for key, arrs, x in _cases(golden):
        err = _relerr(oracle_lib.pcr(*arrs), x)
        assert err < 1e-10, (key, err)


def test_gbsv_oracle_matches_scipy(oracle_lib, golden):
    """The C banded LU that the GPU tests use as their FP64 reference solution."""
    for key, arrs, x in _cases(golden):
        assert _relerr(oracle_lib.gbsv(*arrs), x) < 1e-14, key
        assert _relerr(oracle_lib.gbsv(*arrs, extended=True), x) < 1e-14, key


def test_tridiagonal_reduces_to_textbook_pcr(oracle_lib, golden):
    """a = e = 0: PCR-Penta is Hockney's PCR on a tridiagonal system."""
    for n in (2, 5, 16, 100, 513):
        arrs, x = golden[f"tri_n{n}_in"], golden[f"tri_n{n}_x"]
        assert _relerr(oracle_lib.pcr80(*arrs), x) < 1e-14, n
        assert _relerr(oracle_lib.pcr(*arrs), x) < 1e-12, n


def test_diagonal_system_is_exact(oracle_lib):
    """All off-diagonals 0: each step maps (c, y) -> (-c, -y) exactly, so rows
    that end in the 1x1 branch give x = y/c bit for bit; rows in the 2x2 Cramer
    branch (non-power-of-two n) give c'y/(c c') within 2 ulp."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 7, 64, 1000, 1024):
        c = rng.uniform(1, 2, n)
        y = rng.uniform(-1, 1, n)
        z = np.zeros(n)
        x = oracle_lib.pcr(z, z, c, z, z, y)
        delta = 1 << oracle_lib.steps(n)
        one = np.arange(n)
        one = (one >= n - delta) & (one < delta)
        assert np.array_equal(x[one], (y / c)[one]), n
        np.testing.assert_allclose(x, y / c, rtol=4.5e-16, atol=0)


def test_t2_closed_form(oracle_lib):
    """kernel.cu:19-24's T^2 x = 1 has x_k = k(N-k)(N^2+Nk-k^2+1)/24, k=i+1, N=n+1."""
    for n in (4, 8, 16, 33, 64):
        arrs = oracle_lib.init_t2(n, 1, oracle_lib.CONTIGUOUS)
        N = n + 1
        k = np.arange(1, n + 1, dtype=np.float64)
        xc = k * (N - k) * (N * N + N * k - k * k + 1) / 24
        assert _relerr(oracle_lib.pcr80(*arrs), xc) < 1e-12, n
        assert _relerr(oracle_lib.pcr(*arrs), xc) < 1e-9, n


def test_brute_force_dense_tiny(oracle_lib):
    """n <= 12, random (not dominant) systems: the 80-bit oracle equals a dense solve."""
    rng = np.random.default_rng(7)
    for n in range(1, 13):
        for _ in range(5):
            a, b, d, e = (rng.uniform(-1, 1, n) for _ in range(4))
            c = rng.uniform(2, 3, n) * rng.choice([-1, 1], n)
            y = rng.uniform(-1, 1, n)
            a[:2] = 0; b[:1] = 0; d[n - 1:] = 0; e[max(n - 2, 0):] = 0
            A = np.diag(c)
            for i in range(n):
                if i >= 2: A[i, i - 2] = a[i]
                if i >= 1: A[i, i - 1] = b[i]
                if i + 1 < n: A[i, i + 1] = d[i]
                if i + 2 < n: A[i, i + 2] = e[i]
            x = np.linalg.solve(A.astype(np.longdouble).astype(np.float64), y)
            cond = np.linalg.cond(A)
            assert _relerr(oracle_lib.pcr80(a, b, c, d, e, y), x) < 1e-13 * cond, (n, cond)


def test_batch_driver_matches_single(oracle_lib):
    n, m = 100, 7
    for layout in (oracle_lib.CONTIGUOUS, oracle_lib.INTERLEAVED):
        arrs = oracle_lib.gen_batch(0, 4, n, m, layout)
        x = oracle_lib.pcr_batch(*arrs, n, m, layout)
        for j in range(m):
            xs = oracle_lib.pcr(*oracle_lib.system(arrs, j, n, m, layout))
            assert np.array_equal(oracle_lib.system([x], j, n, m, layout)[0], xs)


def test_uneven_driver_matches_single(oracle_lib):
    sizes = [1, 5, 64, 1000, 3]
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sum(sizes))
    arrs = oracle_lib.gen_batch(0, 9, total, 1, oracle_lib.CONTIGUOUS)
    # zero the out-of-range entries of each system so every system is well formed
    for o, n in zip(off, sizes):
        arrs[0][o:o + min(2, n)] = 0; arrs[1][o] = 0
        arrs[3][o + n - 1] = 0; arrs[4][o + max(n - 2, 0):o + n] = 0
    x = oracle_lib.pcr_uneven(*arrs, off, sizes)
    for o, n in zip(off, sizes):
        xs = oracle_lib.pcr(*[v[o:o + n] for v in arrs])
        assert np.array_equal(x[o:o + n], xs)


@pytest.mark.parametrize("variant", ["plant1", "plant2", "plant3"])
def test_planted_mistakes_are_caught(oracle_lib, golden, variant):
    """Wrong index (1), dropped term (2), wrong sign (3) planted in the oracle
    source: the dgbsv golden check must fail for each."""
    path = oracle_lib.build(variant)
    lib = oracle_lib.load(path)
    worst = 0.0
    for key, arrs, x in _cases(golden):
        got = oracle_lib.pcr80(*arrs, lib=lib)
        with np.errstate(all="ignore"):
            err = _relerr(got, x)
        worst = max(worst, err if np.isfinite(err) else np.inf)
    assert worst > 1e-8, f"{variant} not detected (worst err {worst})"
