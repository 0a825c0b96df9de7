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
    """The tests model the reference's round() (siting.cpp:95,330) as
    trunc(x) + (frac >= .5) - (frac <= -.5); pinned here against libm
    round() on float32 values, exact halves included."""
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


def _wave_rays(roi):
    """The rays of VIEWSHED's wavefronts, as (g, lane, dx, dy, valid): each
    side of the box is dealt in whole wavefronts of 64 lanes, the rows
    dy = -roi and dy = roi without corners (y-major), then the columns
    dx = -roi and dx = roi with them (x-major).  A lane past the end of its
    side (valid False) repeats the ray 64 places back when the side has at
    least 64 rays, else ray j0 mod m of its own wavefront -- the kernel's
    rule (viewshed.hip, the lane-to-ray map of the walk and of the table
    builder)."""
    my, mx = 2 * roi - 1, 2 * roi + 1
    wy, wx = -(-my // 64), -(-mx // 64)
    for g in range(2 * (wy + wx)):
        along_y = g < 2 * wy
        h = g if along_y else g - 2 * wy
        wside = wy if along_y else wx
        lo = h < wside
        m = my if along_y else mx
        for lane in range(64):
            j0 = (h if lo else h - wside) * 64 + lane
            j = j0 if j0 < m else (j0 - 64 if m >= 64 else j0 % m)
            if along_y:
                dx, dy = -roi + 1 + j, (-roi if lo else roi)
            else:
                dx, dy = (-roi if lo else roi), -roi + j
            yield g, lane, dx, dy, j0 < m


def _ray(roi, dx, dy):
    """A ray's (along_y, len, dir, float32 slope) (siting.cpp:292-323)."""
    along_y = abs(dx) < abs(dy)
    major, minor = (dy, dx) if along_y else (dx, dy)
    slope = np.float32(minor) / np.float32(major)
    return along_y, abs(major), (1 if major > 0 else -1), slope


def _ray_cells(roi):
    """The valid rays of the kernel's wavefronts, in order."""
    return [_ray(roi, dx, dy) for _, _, dx, dy, ok in _wave_rays(roi) if ok]


def _offsets(dirn, slope, a):
    x = (np.float32(dirn) * a.astype(np.float32)) * slope
    x = x.astype(np.float32)
    t = np.trunc(x).astype(np.float32)
    f = (x - t).astype(np.float32)
    return t.astype(np.int64) + (f >= 0.5) - (f <= -0.5)


def test_viewshed_ray_prefix_by_bisection():
    """VIEWSHED finds the last step of a ray by bisection instead of testing
    every step (siting.cpp:333-334 break at the first cell outside the
    terrain or the roi disk).  That needs the valid steps to be a prefix:
    checked against the step-by-step walk for every ray of several roi, with
    observers at corners, edges and inside small terrains."""
    for roi in (1, 2, 3, 5, 7, 13, 31, 64, 100, 150):
        rays = _ray_cells(roi)
        for n in (1, 2, 7, 40, 301):
            obs = {(0, 0), (0, n - 1), (n - 1, 0), (n - 1, n - 1), (n // 2, n // 2),
                   (0, n // 2), (n // 3, n - 1), (min(n - 1, roi), min(n - 1, 2))}
            for ox, oy in obs:
                for ray, (along_y, length, dirn, slope) in enumerate(rays):
                    a = np.arange(1, length + 1, dtype=np.int64)
                    off = _offsets(dirn, slope, a)
                    i = dirn * a
                    qx, qy = (off, i) if along_y else (i, off)
                    px, py = ox + qx, oy + qy
                    ok = (px >= 0) & (px < n) & (py >= 0) & (py < n) & (qx * qx + qy * qy <= roi * roi)
                    walk = int(np.argmin(ok)) if not ok.all() else length
                    # bisection as in the kernel: cell(last) inside, cell(hi) not
                    last, hi = 0, length + 1
                    while hi - last > 1:
                        mid = (last + hi) // 2
                        if ok[mid - 1]:
                            last = mid
                        else:
                            hi = mid
                    assert last == walk, (roi, n, ox, oy, ray)
                    assert ok[:last].all()


def test_ray_set_is_the_box_perimeter():
    """The kernel's valid lanes cover every perimeter cell of the (2roi+1)^2
    box exactly once, as the reference's four sides do (siting.cpp:292-304);
    every wavefront walks one major axis in one direction; a padding lane
    repeats a valid ray of its own side -- one that a different wavefront
    walks when the side has at least 64 rays (so a duplicate never ORs the
    same bit as its original in one instruction), else one of its own."""
    for roi in (1, 2, 3, 25, 31, 32, 33, 50, 100, 150, 333):
        waves = {}
        side_rays = {}
        for g, lane, dx, dy, ok in _wave_rays(roi):
            along_y, length, dirn, _ = _ray(roi, dx, dy)
            waves.setdefault(g, []).append(((along_y, dirn), (dx, dy), ok))
            if ok:
                side_rays.setdefault((along_y, dirn), set()).add((dx, dy))
        for lanes in waves.values():
            assert len(lanes) == 64 and len({c for c, _, _ in lanes}) == 1
            assert lanes[0][2]
            cls = lanes[0][0]
            own = {r for _, r, ok in lanes if ok}
            pad = {r for _, r, ok in lanes if not ok}
            assert pad <= side_rays[cls]
            if len(side_rays[cls]) >= 64:
                assert not (pad & own)
            else:
                assert pad <= own
        cells = []
        for along_y, length, dirn, slope in _ray_cells(roi):
            end = _offsets(dirn, slope, np.array([length]))[0]
            cells.append((int(end), dirn * length) if along_y else (dirn * length, int(end)))
        want = {(x, y) for x in range(-roi, roi + 1) for y in range(-roi, roi + 1)
                if max(abs(x), abs(y)) == roi}
        assert len(cells) == len(set(cells)) == 8 * roi
        assert set(cells) == want


def test_half_away_rounding_by_biased_truncation():
    """VIEWSHED and VIX round a ray's minor offset fl(i * slope) half away
    from zero (siting.cpp:95,330) as sign * trunc(fl(|fl(i * slope)| +
    (0.5 - 2^-25))), with |fl(i * slope)| = fl(|i| * |slope|).
    The magnitude identity floor(x + 0.5) == trunc(fl(x + 0.49999997f)) is
    checked here for every float32 x in [0, 2^15) (offsets never exceed
    roi < 2^14), and the whole offset against the reference formula for
    every step of every ray of several roi."""
    h = np.float32(0.49999997)
    assert h.view(np.uint32) == 0x3EFFFFFF
    top = int(np.float32(32768.0).view(np.uint32))
    chunk = 1 << 25
    for b0 in range(0, top, chunk):
        x = np.arange(b0, min(top, b0 + chunk), dtype=np.uint32).view(np.float32)
        got = (x + h).astype(np.int32)
        want = np.floor(x.astype(np.float64) + 0.5).astype(np.int32)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, float(x[bad[0]])
    for roi in (1, 2, 7, 25, 100, 150, 333):
        for ray, (along_y, length, dirn, slope) in enumerate(_ray_cells(roi)):
            a = np.arange(1, length + 1, dtype=np.int64)
            ref = _offsets(dirn, slope, a)
            sgn = 1 if (dirn > 0) == (slope >= 0) else -1
            mag = (a.astype(np.float32) * np.float32(abs(slope))).astype(np.float32)
            mine = sgn * (mag + h).astype(np.float32).astype(np.int64)
            assert np.array_equal(mine, ref), (roi, ray)


def test_div3_by_float_reciprocal():
    """VIX pulls a far target a third of the way in with C's truncating
    x / 3 (siting.cpp:164-165), computed as trunc(fl(x * fl(1/3))) with
    fl(1/3) = 0x3eaaaaab (just above 1/3).  Checked for every |x| <= 2^16;
    the kernel's x is a target offset, |x| <= 2 roi + 1 < 2^15."""
    third = np.float32(0.33333334)
    assert third.view(np.uint32) == 0x3EAAAAAB
    assert np.float64(third) > 1.0 / 3.0
    x = np.arange(-(1 << 16), (1 << 16) + 1, dtype=np.int64)
    got = np.trunc(x.astype(np.float32) * third).astype(np.int64)
    want = np.sign(x) * (np.abs(x) // 3)
    assert np.array_equal(got, want)


def test_target_draw_with_folded_power_of_two():
    """VIX draws a target offset as (int)(fl(fl(span * (float)r) / 2^31))
    (siting.cpp:157,160), computed as (int)fl(span31 * (float)r) with
    span31 = span * 2^-31: scaling by a power of two is exact and commutes
    with rounding for normal floats.  Checked for r at the ends of its range,
    around powers of two and at random, for small, common and maximal roi."""
    rng = np.random.default_rng(3)
    two31 = np.float32(2.0 ** -31)
    edge = np.array([0, 1, 2, 3, (1 << 31) - 1, (1 << 31) - 2] +
                    [(1 << k) + d for k in range(1, 31) for d in (-1, 0, 1)], np.int64)
    for roi in (1, 2, 25, 50, 100, 150, 333, 4096, 16383):
        span = np.float32(np.float32(2 * roi) + np.float32(0.99999))
        r = np.concatenate([edge, rng.integers(0, 1 << 31, 400_000)]).astype(np.float32)
        ref = (np.float32(span * r) / np.float32(2147483648.0)).astype(np.int64)
        mine = (np.float32(span * two31) * r).astype(np.int64)
        assert np.array_equal(ref, mine), roi
