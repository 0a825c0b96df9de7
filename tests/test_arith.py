"""CPU checks of the arithmetic identities the HIP kernels rely on to stay
bit-exact with the reference's float32 operations."""
import numpy as np


def test_division_by_double_reciprocal_matches_fp32_division():
    """VIEWSHED computes (float)(n) / (float)a as (float)((double)n * (1.0/a))
    for |n| < 2^21, 1 <= a < 2^14: the two must agree bit for bit.  Checked
    on every a with a dense set of numerators, plus random pairs."""
    a = np.arange(1, 16384, dtype=np.int64)
    rd = 1.0 / a.astype(np.float64)
    rng = np.random.default_rng(1)
    nums = np.concatenate([np.arange(-2048, 2049), rng.integers(-(2 ** 21) + 1, 2 ** 21, 4000),
                           np.array([2 ** 21 - 1, -(2 ** 21) + 1, 65535, -65535, 1065535])])
    for n in nums:
        ref = np.float32(n) / a.astype(np.float32)
        got = (np.float64(n) * rd).astype(np.float32)
        assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), int(n)
    n = rng.integers(-(2 ** 21) + 1, 2 ** 21, 3_000_000)
    aa = rng.integers(1, 16384, n.size)
    ref = n.astype(np.float32) / aa.astype(np.float32)
    got = (n.astype(np.float64) * (1.0 / aa.astype(np.float64))).astype(np.float32)
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))


def test_round_half_away_by_trunc():
    """round_half_away(x) = trunc(x) + (frac >= .5) - (frac <= -.5) equals
    libm round() on float32 values, exact halves included."""
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-20000, 20000, 200000).astype(np.float32),
                        (np.arange(-4000, 4001) / 2).astype(np.float32),
                        np.array([np.nextafter(np.float32(0.5), np.float32(0)),
                                  np.nextafter(np.float32(-0.5), np.float32(0))])])
    x = x.astype(np.float32)
    t = np.trunc(x).astype(np.float32)
    f = (x - t).astype(np.float32)
    mine = t.astype(np.int64) + (f >= 0.5) - (f <= -0.5)
    ref = np.where(x >= 0, np.floor(x.astype(np.float64) + 0.5),
                   -np.floor(-x.astype(np.float64) + 0.5)).astype(np.int64)
    assert np.array_equal(mine, ref)


def test_float_index_split_exact():
    """SITE splits item indices with q = (int)((it + 0.5) * fl(1/nk)); exact
    for it < 2^21 and nk <= 64 (the kernel falls back to division beyond)."""
    for nk in range(1, 65):
        it = np.arange(0, 2 ** 21, dtype=np.int64)
        inv = np.float32(1.0) / np.float32(nk)
        q = ((it.astype(np.float32) + np.float32(0.5)) * inv).astype(np.int64)
        assert np.array_equal(q, it // nk), nk
