"""CPU checks that pin the oracle (oracle/, the C restatement of the
sequential reference) before it is trusted as the GPU path's checker.

Each function is pinned by something other than its own formula: the
reference's build where it exists (golden fixtures and, in a container with
the reference sources, its stage functions directly), closed-form known
answers, invariants, and brute force on tiny inputs.
"""
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import make_dem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- generator
def test_lcg_known_answers(orc):
    # glibc TYPE_0 step (siting.cpp:29-33)
    assert orc.lcg(0) == 12345
    assert orc.lcg(12345) == 1406932606
    assert orc.lcg(1) == 1103527590
    # the input is reduced mod 2^31 first
    assert orc.lcg(-1) == orc.lcg(0x7fffffff)


def test_lcg_jump_ahead_composition(orc):
    """r_m = (A_m r_0 + C_m) mod 2^31 with A, C iterated as the device's
    jump table does; the 2-step pair is a=1117952617, c=1406932606."""
    a, c = 1103515245, 12345
    A, C = a, c
    tab = []
    for m in range(1, 201):
        tab.append((A, C))
        C = (a * C + c) % 2 ** 32
        A = (a * A) % 2 ** 32
    assert tab[1][0] % 2 ** 31 == 1117952617 and tab[1][1] % 2 ** 31 == 1406932606
    rng = np.random.default_rng(0)
    for r0 in list(rng.integers(0, 2 ** 31, 20)) + [0, 1, 2 ** 31 - 1]:
        r = int(r0)
        for m in range(1, 201):
            r = orc.lcg(r)
            Am, Cm = tab[m - 1]
            assert r == ((Am * int(r0) + Cm) % 2 ** 32) & 0x7fffffff


# ---------------------------------------------------------------- derivation
@pytest.mark.parametrize("n,bw,nw0,bs,nb,nwpb,nwanted", [
    (1000, 10, 10000, 10.0, 100, 1, 10000),
    (4096, 16, 65536, 16.0, 256, 1, 65536),
    (2048, 8, 32768, 8.0, 256, 1, 65536),       # nwanted0 below nblocks^2
    (8192, 16, 131072, 16.0, 512, 1, 262144),
    (1000, 16, 3969, np.float32(1000 / 63), 63, 1, 3969),
    (998, 16, 5000, np.float32(998 / 62), 62, 2, 7688),
])
def test_derive_closed_form(orc, sa, n, bw, nw0, bs, nb, nwpb, nwanted):
    rc, d = orc.derive(n, bw, nw0)
    assert rc == 0
    assert np.float32(d.blocksize) == np.float32(bs)
    assert (d.nblocks, d.nwpb, d.nwanted) == (nb, nwpb, nwanted)
    # the library's own derivation (host code) agrees
    dd = sa.derive(sa.make_params(n, 10, 0, 1, bw, nw0, 0))
    assert (np.float32(dd.blocksize), dd.nblocks, dd.nwpb, dd.nwanted, dd.lastsize) == \
        (np.float32(d.blocksize), d.nblocks, d.nwpb, d.nwanted, d.lastsize)


def test_derive_small_last_block(orc, sa):
    # nwpb larger than the last block holds (siting.cpp:694-695)
    rc, d = orc.derive(25, 10, 20000)
    assert rc == 1
    with pytest.raises(sa.SitingError) as e:
        sa.derive(sa.make_params(25, 10, 0, 1, 10, 20000, 0))
    assert e.value.status == sa.ESMALLBLOCK


# ---------------------------------------------------------------- VIX
def _targets_in_terrain(n, roi, ntests, p):
    """Count of targets of point p that land inside the terrain, from the
    float32 target formula (siting.cpp:155-173) evaluated with numpy."""
    ox, oy = divmod(p, n)
    r = p
    span = np.float32(2 * roi) + np.float32(0.99999)
    cnt = 0
    for _ in range(ntests):
        r = (1103515245 * (r & 0x7fffffff) + 12345) & 0x7fffffff
        tx = int(np.float32(span * np.float32(r)) / np.float32(2147483648.0)) + ox - roi
        r = (1103515245 * (r & 0x7fffffff) + 12345) & 0x7fffffff
        ty = int(np.float32(span * np.float32(r)) / np.float32(2147483648.0)) + oy - roi
        if (tx - ox) ** 2 + (ty - oy) ** 2 > roi * roi:
            tx = ox + int((tx - ox) / 3)   # C truncation
            ty = oy + int((ty - oy) / 3)
        cnt += 0 <= tx < n and 0 <= ty < n
    return cnt


def test_vix_flat_terrain(orc):
    """On flat terrain every in-terrain target is visible, so vix is the
    closed-form byte of the in-terrain count: pins the generator, the target
    draw and the byte formula independently of the line-of-sight floats."""
    n, roi, nt = 40, 12, 23
    dem = np.full(n * n, 777, np.uint16)
    vix = orc.calc_vix(n, dem, roi, 10, 10, nt)
    for p in range(0, n * n, 7):
        k = _targets_in_terrain(n, roi, nt, p)
        v = np.float32(k) / np.float32(nt)
        assert vix[p] == min(255, int(np.float32(v * np.float32(255.999)))), p


def test_vix_byte_table():
    # (float)nvis/64 * 255.999f truncated: 0,3,7,11,...,247,251,255
    b = [min(255, int(np.float32(np.float32(k) / np.float32(64)) *
                      np.float32(255.999))) for k in range(65)]
    assert b[:4] == [0, 3, 7, 11] and b[-3:] == [247, 251, 255]


def test_vix_blocked_ridge(orc):
    """A wall between observer and target hides it; without the wall it is
    seen (hand-checked geometry)."""
    n = 32
    dem = np.zeros(n * n, np.uint16)
    t = orc.lib().so_test_one_target
    e = dem.ctypes.data
    assert t(n, e, 10, 10, 5, 10, 20, 5) == 1
    dem.reshape(n, n)[10, 12] = 100     # sample at offset 2 on the way
    assert t(n, e, 10, 10, 5, 10, 20, 5) == 0
    assert t(n, e, 10, 10, 5, 10, 11, 5) == 1   # adjacent: always visible


# ---------------------------------------------------------------- FINDMAX
def _hash(x, y):
    v = (x * (x + y) * 0o10101010101) & 0xffffffff
    return v - 2 ** 32 if v >= 2 ** 31 else v


def test_hash_known_answers():
    assert _hash(1, 0) == 1090785345
    assert _hash(2, 3) == -1977048438


def test_findmax_brute_force(orc):
    rng = np.random.default_rng(5)
    for n, bw, nw0 in [(23, 5, 60), (30, 7, 200), (17, 4, 50)]:
        rc, d = orc.derive(n, bw, nw0)
        assert rc == 0
        vix = rng.integers(0, 4, n * n).astype(np.uint8)   # many ties
        work = vix.copy()
        obs = orc.find_max(n, work, d.blocksize, d.nblocks, d.nwpb)
        bs = np.float32(d.blocksize)
        v2 = vix.copy()
        for b in range(d.nblocks ** 2):
            bx, by = divmod(b, d.nblocks)
            x0 = int(np.float32(bs * np.float32(bx)))
            x1 = min(int(np.float32(bs * np.float32(bx + 1))), n)
            y0 = int(np.float32(bs * np.float32(by)))
            y1 = min(int(np.float32(bs * np.float32(by + 1))), n)
            for k in range(d.nwpb):
                cells = [(x, y) for x in range(x0, x1) for y in range(y0, y1)]
                best = max(cells, key=lambda c: (v2[c[0] * n + c[1]], _hash(*c),
                                                 -(c[0] * n + c[1])))
                assert obs[b * d.nwpb + k] == best[0] * n + best[1]
                v2[best[0] * n + best[1]] = 0
        np.testing.assert_array_equal(work, v2)


# ---------------------------------------------------------------- VIEWSHED
def _shed_bits(shed, roi):
    nr = 2 * roi + 1
    nwpr = (nr + 63) // 64
    w = shed.reshape(nr, nwpr)
    bits = np.zeros((nr, nr), bool)
    for c in range(nr):
        bits[:, c] = (w[:, c // 64] >> np.uint64(63 - c % 64)) & np.uint64(1)
    pad = np.zeros((nr, nwpr * 64 - nr), bool)
    for c in range(nr, nwpr * 64):
        pad[:, c - nr] = (w[:, c // 64] >> np.uint64(63 - c % 64)) & np.uint64(1)
    return bits, pad


def test_viewshed_flat_is_disk(orc):
    """Flat terrain: every cell within roi and inside the terrain is seen,
    nothing else; pins the ray enumeration, rounding, clipping and the
    MSB-first layout."""
    n, roi = 50, 13
    dem = np.full(n * n, 500, np.uint16)
    obs = np.array([0, 5 * n + 7, 25 * n + 25, 49 * n + 30, 44 * n + 49], np.int32)
    sheds = orc.calc_sheds(n, dem, roi, 5, 5, obs)
    nr = 2 * roi + 1
    S = nr * ((nr + 63) // 64)
    for k, o in enumerate(obs):
        ox, oy = divmod(int(o), n)
        bits, pad = _shed_bits(sheds[k * S:(k + 1) * S], roi)
        assert not pad.any()
        for r in range(nr):
            for c in range(nr):
                x, y = ox - roi + r, oy - roi + c
                inside = 0 <= x < n and 0 <= y < n
                disk = (r - roi) ** 2 + (c - roi) ** 2 <= roi * roi
                assert bits[r, c] == (inside and disk), (k, r, c)


# ---------------------------------------------------------------- SITE
def _grid(shed, roi, n, o):
    ox, oy = divmod(int(o), n)
    bits, _ = _shed_bits(shed, roi)
    g = np.zeros((n, n), bool)
    nr = 2 * roi + 1
    for r in range(nr):
        for c in range(nr):
            if bits[r, c]:
                g[ox - roi + r, oy - roi + c] = True
    return g


def test_site_brute_force_greedy(orc, sa):
    """intervis=0: the oracle equals naive exact greedy max coverage over
    explicit boolean grids (max gain, lowest index, stop at 0 gain or >95%)."""
    for seed, n, roi, m in [(1, 40, 5, 60), (2, 36, 7, 45), (3, 30, 4, 80)]:
        dem = make_dem(sa, n, 0.6, 30000, seed, nhills=2, sigma=6.0)
        rng = np.random.default_rng(seed)
        obs = rng.choice(n * n, m, replace=False).astype(np.int32)
        sheds = orc.calc_sheds(n, dem, roi, 20, 20, obs)
        nr = 2 * roi + 1
        S = nr * ((nr + 63) // 64)
        grids = [_grid(sheds[k * S:(k + 1) * S], roi, n, obs[k]) for k in range(m)]
        cov = np.zeros((n, n), bool)
        used = np.zeros(m, bool)
        order = []
        while True:
            gains = [(-1 if used[k] else int((grids[k] & ~cov).sum())) for k in range(m)]
            best = max(range(m), key=lambda k: (gains[k], -k))
            if gains[best] <= 0:
                break
            cov |= grids[best]
            used[best] = True
            order.append(best)
            if 100.0 * cov.sum() / (n * n) > 95:
                break
        sel, ord_o, st = orc.site(n, roi, 0, obs, sheds)
        assert list(ord_o) == order
        np.testing.assert_array_equal(sel.astype(bool), used)
        assert st.cumarea == cov.sum()


def test_site_intervis_rule(orc, sa):
    """intervis=1: every observer after the first is covered by the union of
    those selected before it, or kept its unmasked first-iteration area
    (the stale rule); coverage still equals the union's popcount."""
    n, roi = 60, 6
    dem = make_dem(sa, n, 0.6, 30000, 9, nhills=2, sigma=6.0)
    rng = np.random.default_rng(9)
    obs = rng.choice(n * n, 150, replace=False).astype(np.int32)
    sheds = orc.calc_sheds(n, dem, roi, 20, 20, obs)
    nr = 2 * roi + 1
    S = nr * ((nr + 63) // 64)
    sel, order, st = orc.site(n, roi, 1, obs, sheds)
    cov = np.zeros((n, n), bool)
    for k in order:
        cov |= _grid(sheds[k * S:(k + 1) * S], roi, n, obs[k])
    assert st.cumarea == cov.sum()


# ---------------------------------------------------------------- goldens
def _load_readme_dem():
    with gzip.open(os.path.join(GOLDEN, "dem1000.u16.gz"), "rb") as f:
        raw = f.read()
    meta = json.load(open(os.path.join(GOLDEN, "readme_example.json")))
    assert hashlib.sha256(raw).hexdigest() == meta["dem_sha256"]
    return np.frombuffer(raw, np.uint16).copy(), meta


def test_golden_readme_example(orc):
    """The README example run of the reference's seqsite on its sample DEM."""
    dem, meta = _load_readme_dem()
    n, roi, ht, nt, bw, nw0, iv = meta["args"]
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv, nthreads=min(8, os.cpu_count()))
    text = "".join(f"{o // n},{o % n}\n" for o, s in zip(ref["obs"], ref["selected"]) if s)
    assert text == meta["observers"]
    assert ref["stats"].nused == meta["nusedsheds"]
    assert f"{ref['stats'].coverage:g}" == meta["coverage_printed"]


def test_golden_small_stages(orc, sa):
    z = np.load(os.path.join(GOLDEN, "stages_small.npz"))
    i = 0
    while f"c{i}_case" in z:
        n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = [int(v) if float(v).is_integer()
                                                      else float(v) for v in z[f"c{i}_case"]]
        dem = z[f"c{i}_dem"]
        np.testing.assert_array_equal(make_dem(sa, n, H, hmax, seed), dem)
        ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv)
        for key in ("vix", "vix_after", "obs", "sheds", "selected"):
            np.testing.assert_array_equal(ref[key], z[f"c{i}_{key}"], err_msg=key)
        i += 1
    assert i == 3


def test_golden_medium_digests(orc, sa):
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    for row in json.load(open(os.path.join(GOLDEN, "stages_digest.json"))):
        n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = row["case"]
        dem = make_dem(sa, n, H, hmax, seed)
        assert sha(dem) == row["dem"]
        ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv, nthreads=min(8, os.cpu_count()))
        for key in ("vix", "vix_after", "obs", "sheds", "selected"):
            assert sha(ref[key]) == row[key], (row["case"], key)


# ---------------------------------------------------------------- reference build
@pytest.mark.skipif(not __import__("oracle").ref_available(),
                    reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("case", [
    (90, 11, 100, 30, 10, 200, 0, 0.8, 20000, 301),
    (90, 11, 100, 30, 10, 200, 1, 0.8, 20000, 301),
    (128, 17, 50, 64, 8, 400, 1, 0.6, 30000, 302),
    (111, 64, 0, 5, 12, 100, 0, 0.5, 20000, 303),
    (140, 20, 100, 100, 10, 600, 0, 0.9, 20000, 304),
])
def test_oracle_matches_reference_build(orc, sa, case):
    n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = case
    dem = make_dem(sa, n, H, hmax, seed)
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv)
    R = orc.ref_lib()
    rc, d = orc.derive(n, bw, nw0)
    vix = np.empty(n * n, np.uint8)
    R.ref_calc_vix(n, dem.ctypes.data, roi, ht, ht, nt, vix.ctypes.data)
    np.testing.assert_array_equal(vix, ref["vix"])
    obs = np.empty(d.nwanted, np.int32)
    R.ref_find_max(n, vix.ctypes.data, d.blocksize, d.nwanted, d.nblocks, d.nwpb,
                   obs.ctypes.data)
    np.testing.assert_array_equal(obs, ref["obs"])
    # the -fwrapv build agrees: the signed-overflow hash is not exploited
    vw = ref["vix"].copy()
    ow = np.empty(d.nwanted, np.int32)
    orc.ref_lib(wrapv=True).ref_find_max(n, vw.ctypes.data, d.blocksize, d.nwanted,
                                         d.nblocks, d.nwpb, ow.ctypes.data)
    np.testing.assert_array_equal(ow, obs)
    sheds = np.empty_like(ref["sheds"])
    R.ref_calc_sheds(n, dem.ctypes.data, roi, ht, ht, d.nwanted, obs.ctypes.data,
                     sheds.ctypes.data)
    np.testing.assert_array_equal(sheds, ref["sheds"])
    sel = np.empty(d.nwanted, np.uint8)
    R.ref_site_it(n, roi, iv, d.nwanted, obs.ctypes.data, sheds.ctypes.data,
                  sel.ctypes.data)
    np.testing.assert_array_equal(sel, ref["selected"])


def test_openmp_variant_equals_sequential(orc, sa):
    n = 160
    dem = make_dem(sa, n, 0.7, 20000, 77)
    a = orc.run(n, dem, 16, 100, 40, 8, 400, 1, nthreads=1)
    b = orc.run(n, dem, 16, 100, 40, 8, 400, 1, nthreads=4)
    for key in ("vix", "obs", "sheds", "selected", "order"):
        np.testing.assert_array_equal(a[key], b[key])
