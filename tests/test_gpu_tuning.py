"""GPU: kernel shape choices change speed only.  Every tunable choice (VIX
halo, VIEWSHED threads and bitset mode, SITE column classes and stale
shift) gives the built-in rules' bits in every stage; and a job created
under a tuning cache takes the cache's entry for its class and still gives
the rules' bits."""
import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu

# (n, roi, ntests, bw, nwanted0, intervis, rough_h, seed)
CASES = [(300, 40, 64, 10, 900, 0, 0.8, 91),
         (260, 100, 32, 16, 264, 1, 0.6, 92),
         (240, 150, 32, 16, 225, 0, 0.7, 93),
         (400, 25, 64, 8, 2500, 1, 0.6, 94)]

CHOICES = [dict(vix_halo=8), dict(vix_halo=16), dict(vix_halo=32),
           dict(viewshed_threads=128), dict(viewshed_threads=512),
           dict(viewshed_mode=1), dict(viewshed_mode=2), dict(viewshed_mode=3),
           dict(site_q=2), dict(site_q=3), dict(site_q=4), dict(site_tau_shift=1),
           dict(site_tau_shift=9), dict(site_walk_evals=1), dict(site_walk_evals=3),
           dict(site_walk_evals=33),
           dict(vix_halo=32, viewshed_threads=128, viewshed_mode=2, site_q=3,
                site_tau_shift=4, site_walk_evals=9)]


def _outputs(job):
    st = job.stats()
    return (job.download_vix(), job.download_obs(), job.download_sheds(),
            job.download_selected(), job.download_order(st.nused),
            (st.nused, st.niter, st.cumarea, st.sum_u))


def _equal(a, b):
    for x, y in zip(a[:5], b[:5]):
        np.testing.assert_array_equal(x, y)
    assert a[5] == b[5]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}_roi{c[1]}_iv{c[5]}")
def test_every_choice_gives_the_rules_bits(gpu, case, monkeypatch):
    monkeypatch.setenv("SITING_TUNE_CACHE", "off")
    n, roi, nt, bw, nw0, iv, H, seed = case
    print(f"[case] tuning choices {case}", flush=True)
    dem = make_dem(gpu, n, H, 20000, seed)
    p = gpu.make_params(n, roi, 100, nt, bw, nw0, iv)
    with gpu.Job(p) as job:
        job.upload_dem(dem)
        assert job.get_tuning().as_dict() == gpu.Tuning().as_dict()
        job.run()
        job.sync()
        ref = _outputs(job)
        for c in CHOICES:
            job.set_tuning(gpu.Tuning(**c))
            job.run()
            job.sync()
            _equal(_outputs(job), ref)


def test_job_takes_the_cache_entry(gpu, tmp_path, monkeypatch):
    n, roi, nt, bw, nw0, iv, H, seed = CASES[0]
    key = gpu.device_key(0)
    cache = tmp_path / "tuning.txt"
    cache.write_text(f"{key} 50 64 1024 8 128 3 3 6\n")
    monkeypatch.setenv("SITING_TUNE_CACHE", str(cache))
    dem = make_dem(gpu, n, H, 20000, seed)
    p = gpu.make_params(n, roi, 100, nt, bw, nw0, iv)
    with gpu.Job(p) as job:
        assert job.get_tuning().as_dict() == {"vix_halo": 8, "viewshed_threads": 128,
                                              "viewshed_mode": 3, "site_q": 3,
                                              "site_tau_shift": 6, "site_walk_evals": 0}
        job.upload_dem(dem)
        job.run()
        job.sync()
        tuned = _outputs(job)
    monkeypatch.setenv("SITING_TUNE_CACHE", "off")
    with gpu.Job(p) as job:
        assert job.get_tuning().as_dict() == gpu.Tuning().as_dict()
        job.upload_dem(dem)
        job.run()
        job.sync()
        _equal(tuned, _outputs(job))
