"""The tuning cache: its file format, the job classes it is keyed by, and how
the library finds it (SITING_TUNE_CACHE, "off", the default beside the
library).  Lookups need no GPU."""
import pytest


def _write(tmp_path, body):
    p = tmp_path / "tuning.txt"
    p.write_text(body)
    return str(p)


CACHE = """# comment line
gfx950:256 100 64 4096 32 512 2 3 4  # trailing comment
gfx950:256 25 32 1024 8 128 3 2 6 9
gfx950:256 200 64 2048 16 256 1 2 5
gfx942:304 100 64 4096 16 256 1 2 5
bad line with words
"""


def test_lookup_classes(sa, tmp_path, monkeypatch):
    monkeypatch.setenv("SITING_TUNE_CACHE", _write(tmp_path, CACHE))
    assert sa.tuning_cache_path().endswith("tuning.txt")
    # roi 110 -> class 100, 64 targets, edge 4000 -> class 4096
    t = sa.tuning_lookup("gfx950:256", sa.make_params(4000, 110, 100, 64, 16, 1000, 0))
    assert t.as_dict() == {"vix_halo": 32, "viewshed_threads": 512, "viewshed_mode": 2,
                           "site_q": 3, "site_tau_shift": 4, "site_walk_evals": 0}
    # roi 30 -> 25, 40 targets -> 32, edge 1500 -> 1024
    t = sa.tuning_lookup("gfx950:256", sa.make_params(1500, 30, 100, 40, 16, 1000, 0))
    assert (t.vix_halo, t.viewshed_threads, t.viewshed_mode, t.site_q, t.site_tau_shift,
            t.site_walk_evals) == (8, 128, 3, 2, 6, 9)
    # roi 250 -> 200, edge 3000 -> 2048
    t = sa.tuning_lookup("gfx950:256", sa.make_params(3000, 250, 100, 64, 16, 1000, 1))
    assert t.viewshed_mode == 1
    # another device's key; a class with no entry; roi beyond every class
    assert sa.tuning_lookup("gfx942:304", sa.make_params(4000, 100, 100, 64, 16, 1000, 0))
    assert sa.tuning_lookup("gfx950:128", sa.make_params(4000, 100, 100, 64, 16, 1000, 0)) \
        is None
    assert sa.tuning_lookup("gfx950:256", sa.make_params(4000, 50, 100, 64, 16, 1000, 0)) \
        is None
    assert sa.tuning_lookup("gfx950:256", sa.make_params(4000, 400, 100, 64, 16, 1000, 0)) \
        is None


def test_cache_off_and_missing(sa, tmp_path, monkeypatch):
    monkeypatch.setenv("SITING_TUNE_CACHE", "off")
    assert sa.tuning_cache_path() == ""
    p = sa.make_params(4000, 110, 100, 64, 16, 1000, 0)
    assert sa.tuning_lookup("gfx950:256", p) is None
    monkeypatch.setenv("SITING_TUNE_CACHE", str(tmp_path / "absent.txt"))
    assert sa.tuning_lookup("gfx950:256", p) is None
    # switching back to a real file reloads it
    monkeypatch.setenv("SITING_TUNE_CACHE", _write(tmp_path, CACHE))
    assert sa.tuning_lookup("gfx950:256", p) is not None


def test_default_cache_is_beside_the_library(sa, monkeypatch):
    monkeypatch.delenv("SITING_TUNE_CACHE", raising=False)
    import os
    path = sa.tuning_cache_path()
    assert os.path.normpath(path) == os.path.normpath(
        os.path.join(os.path.dirname(sa.LIB_PATH), "..", "tune", "tuning.txt"))


def test_committed_cache_parses(sa, monkeypatch):
    """Every line of the committed cache (when present) is well formed."""
    import os
    monkeypatch.delenv("SITING_TUNE_CACHE", raising=False)
    path = sa.tuning_cache_path()
    if not os.path.exists(path):
        pytest.skip("no tuning cache committed")
    rows = [ln.split("#")[0].split() for ln in open(path)]
    rows = [r for r in rows if r]
    assert rows
    for r in rows:
        assert len(r) in (9, 10), r
        roi, nt, edge, halo, th, mode, q, s = map(int, r[1:9])
        ev = int(r[9]) if len(r) == 10 else 0
        assert roi in (25, 50, 100, 150, 200) and nt in (32, 64) and edge in (1024, 2048, 4096)
        assert halo in (8, 16, 32) and th % 64 == 0 and mode in (1, 2, 3)
        assert 1 <= q <= 8 and 1 <= s <= 20 and 0 <= ev <= 33
        key = r[0]
        got = sa.tuning_lookup(key, sa.make_params(edge, roi, 100, nt, 16, 1000, 0))
        assert got.as_dict() == {"vix_halo": halo, "viewshed_threads": th,
                                 "viewshed_mode": mode, "site_q": q, "site_tau_shift": s,
                                 "site_walk_evals": ev}
