"""Failure handling of the SITE loop, forced from the host.

The loop's grid barrier gives up after a bounded number of polls, so a barrier
that cannot complete ends the loop with niter = -1 instead of hanging the
device, and the API returns an error.  Here the limit is set to 0 for one
call: every workgroup arrives, gives the loop up at once and exits normally.
The next call, with the default limit, must be correct again.
"""
import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu


def test_site_loop_gives_up_when_a_barrier_cannot_complete(gpu, orc, monkeypatch):
    n, roi = 180, 15  # a loop of 34 workgroups
    dem = make_dem(gpu, n, 0.7, 20000, 11)
    rng = np.random.default_rng(5)
    obs = rng.choice(n * n, size=700, replace=False).astype(np.int32)
    sheds = orc.calc_sheds(n, dem, roi, 100, 100, obs)
    monkeypatch.setenv("SITING_SITE_SPIN_LIMIT", "0")
    with pytest.raises(gpu.SitingError) as err:
        gpu.site_it(n, roi, 0, obs.size, obs, sheds)
    assert err.value.status == gpu.EDEVICE
    assert "grid barrier" in err.value.message
    monkeypatch.delenv("SITING_SITE_SPIN_LIMIT")
    sel, st = gpu.site_it(n, roi, 0, obs.size, obs, sheds)
    sel_r, _, st_r = orc.site(n, roi, 0, obs, sheds)
    np.testing.assert_array_equal(sel, sel_r)
    assert (st.nused, st.niter, st.cumarea) == (st_r.nused, st_r.niter, st_r.cumarea)


def test_site_loop_detects_diverging_picks(gpu, orc, monkeypatch):
    """Every workgroup of the loop derives a round's picks from the same
    published keys and publishes a hash of them; a mismatch ends the loop in
    all workgroups at the same point with an error.  The knob alters the
    last workgroup's hash of round 1 only (its picks stay right), so every
    workgroup still exits normally."""
    n, roi = 180, 15
    dem = make_dem(gpu, n, 0.7, 20000, 11)
    rng = np.random.default_rng(5)
    obs = rng.choice(n * n, size=700, replace=False).astype(np.int32)
    sheds = orc.calc_sheds(n, dem, roi, 100, 100, obs)
    sel_r, _, st_r = orc.site(n, roi, 0, obs, sheds)
    assert st_r.nused > 40  # many rounds
    monkeypatch.setenv("SITING_SITE_HASH_FAULT", "1")
    with pytest.raises(gpu.SitingError) as err:
        gpu.site_it(n, roi, 0, obs.size, obs, sheds)
    assert err.value.status == gpu.EDEVICE
    assert "disagreed" in err.value.message
    monkeypatch.delenv("SITING_SITE_HASH_FAULT")
    sel, st = gpu.site_it(n, roi, 0, obs.size, obs, sheds)
    np.testing.assert_array_equal(sel, sel_r)
    assert (st.nused, st.niter, st.cumarea) == (st_r.nused, st_r.niter, st_r.cumarea)
