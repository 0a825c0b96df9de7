"""GPU parity on medium terrains: seeded draws from the uneven batch's and the
served stream's parameter ranges, against the CPU oracle, bit for bit.

The small cases elsewhere keep SITE on one or two workgroups.  Here the
terrains are 256..1100 cells wide, so SITE runs on many workgroups with
multi-pick rounds and stale bounds, VIX takes both LDS windows, and
VIEWSHED's per-roi tables cover every roi class of the batch
(25, 50, 100, 150) and of the stream (200).  About a fifth of the draws set
intervis.  The oracle runs on 8 threads; every thread count gives the
sequential reference's bits.
"""
import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu


def _draw(k):
    """Case k: edge, roi, targets, bw, intervis and roughness from one seed."""
    r = np.random.default_rng(7000 + k)
    n = int(r.integers(400, 1101))
    roi = int(r.choice([25, 50, 100, 150]))
    nt = int(r.choice([32, 64]))
    bw = int(r.choice([8, 16]))
    iv = int(r.random() < 0.25)
    h = float(np.round(r.uniform(0.5, 0.9), 3))
    return (n, roi, 100, nt, bw, n * n // 256, iv, h, 20000, 7000 + k)


def _draw_served(k):
    """Served-stream case k: roi from {25, 50, 100, 200}, 64 targets, bw 16,
    nwanted n^2/512, intervis with probability 0.2."""
    r = np.random.default_rng(8000 + k)
    n = int(r.integers(256, 1101))
    roi = int(r.choice([25, 50, 100, 200]))
    iv = int(r.random() < 0.2)
    h = float(np.round(r.uniform(0.5, 0.9), 3))
    return (n, roi, 100, 64, 16, n * n // 512, iv, h, 20000, 8000 + k)


# (n, roi, height, ntests, bw, nwanted0, intervis, rough_h, hmax, seed)
CASES = [_draw(k) for k in range(8)] + [_draw_served(k) for k in range(8)]


def _case_id(c):
    return "n{}_roi{}_t{}_bw{}_iv{}".format(c[0], c[1], c[3], c[4], c[6])


@pytest.mark.parametrize("case", CASES, ids=[_case_id(c) for c in CASES])
def test_medium_job_matches_oracle(gpu, orc, case):
    n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = case
    print(f"[case] sweep {_case_id(case)}", flush=True)
    dem = make_dem(gpu, n, H, hmax, seed)
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv, nthreads=8)
    p = gpu.make_params(n, roi, ht, nt, bw, nw0, iv)
    with gpu.Job(p) as job:
        job.upload_dem(dem)
        job.run()
        job.sync()
        st = job.stats()
        np.testing.assert_array_equal(job.download_vix(), ref["vix_after"])
        np.testing.assert_array_equal(job.download_obs(), ref["obs"])
        np.testing.assert_array_equal(job.download_sheds(), ref["sheds"])
        np.testing.assert_array_equal(job.download_selected(), ref["selected"])
        np.testing.assert_array_equal(job.download_order(st.nused), ref["order"])
        rs = ref["stats"]
        assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
            (rs.nused, rs.niter, rs.cumarea, rs.sum_u)
        assert st.coverage == rs.coverage
