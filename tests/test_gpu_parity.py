"""GPU parity: every stage of the HIP path against the CPU oracle, bit for bit.

The oracle (oracle/, the C restatement of the sequential reference) runs in
the same process on the same seeded DEM.  Cases cover the reference's CLI
example shape, bw below the CLI's minimum (library calls), intervis 0 and 1,
several tentative observers per block, more targets than a wavefront, an roi
wider than the terrain, tiny terrains and ragged block edges.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu

# (n, roi, height, ntests, bw, nwanted0, intervis, rough_h, hmax, seed)
CASES = [
    (200, 20, 100, 32, 10, 400, 0, 0.8, 20000, 1),
    (200, 20, 100, 32, 10, 400, 1, 0.8, 20000, 1),
    (257, 13, 50, 50, 8, 1000, 1, 0.6, 30000, 3),
    (300, 37, 100, 64, 16, 300, 0, 0.8, 20000, 4),
    (130, 9, 10, 100, 10, 800, 0, 0.5, 20000, 5),     # nwpb = 5, ntests > 64
    (64, 80, 100, 16, 10, 100, 1, 0.9, 20000, 6),     # roi wider than terrain
    (97, 5, 0, 1, 7, 200, 0, 0.7, 20000, 7),          # one target, ragged blocks
    (33, 3, 100, 8, 3, 1000, 1, 0.6, 30000, 8),       # tiny blocks, nwpb 9
    (150, 64, 100, 20, 25, 60, 0, 0.8, 20000, 9),     # nwpr exactly 3 words
    (120, 260, 100, 8, 40, 9, 0, 0.7, 20000, 10),     # bitset in one LDS array
    (130, 400, 50, 8, 43, 9, 1, 0.7, 20000, 11),      # bitset in HBM
    (240, 70, 100, 64, 12, 400, 0, 0.8, 20000, 12),   # VIX window of 32 cells
]


def _case_id(c):
    return "n{}_roi{}_h{}_t{}_bw{}_nw{}_iv{}".format(*c[:7])


@pytest.mark.parametrize("case", CASES, ids=[_case_id(c) for c in CASES])
def test_stages_host_api(gpu, orc, case):
    n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = case
    print(f"[case] host api {_case_id(case)}", flush=True)
    dem = make_dem(gpu, n, H, hmax, seed)
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv)
    d = ref["derived"]

    vix = gpu.calc_vix(n, dem, roi, ht, ht, nt)
    np.testing.assert_array_equal(vix, ref["vix"])

    obs = gpu.find_max(n, vix, d.blocksize, d.nwanted, d.nblocks, d.nwpb)
    np.testing.assert_array_equal(obs, ref["obs"])
    np.testing.assert_array_equal(vix, ref["vix_after"])  # mutated as cpp:231

    sheds = gpu.calc_sheds(n, dem, roi, ht, ht, d.nwanted, obs)
    np.testing.assert_array_equal(sheds, ref["sheds"])

    sel, st = gpu.site_it(n, roi, iv, d.nwanted, obs, sheds)
    rs = ref["stats"]
    np.testing.assert_array_equal(sel, ref["selected"])
    assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
        (rs.nused, rs.niter, rs.cumarea, rs.sum_u)
    assert st.coverage == rs.coverage


@pytest.mark.parametrize("case", CASES, ids=[_case_id(c) for c in CASES])
def test_job_resident(gpu, orc, case):
    n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = case
    print(f"[case] job {_case_id(case)}", flush=True)
    dem = make_dem(gpu, n, H, hmax, seed)
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv)
    p = gpu.make_params(n, roi, ht, nt, bw, nw0, iv)
    with gpu.Job(p) as job:
        job.upload_dem(dem)
        job.vix()
        job.sync()
        np.testing.assert_array_equal(job.download_vix(), ref["vix"])
        job.findmax()
        job.viewshed()
        job.site()
        job.sync()
        st = job.stats()
        np.testing.assert_array_equal(job.download_obs(), ref["obs"])
        np.testing.assert_array_equal(job.download_sheds(), ref["sheds"])
        np.testing.assert_array_equal(job.download_selected(), ref["selected"])
        np.testing.assert_array_equal(job.download_order(st.nused), ref["order"])
        rs = ref["stats"]
        assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
            (rs.nused, rs.niter, rs.cumarea, rs.sum_u)


def test_site_arbitrary_observer_order(gpu, orc):
    """site_it takes any observer list (the reference's signature), not only
    FINDMAX's block order: shuffle obs and sheds together."""
    n, roi = 180, 15
    dem = make_dem(gpu, n, 0.7, 20000, 11)
    rng = np.random.default_rng(11)
    obs = rng.choice(n * n, size=700, replace=False).astype(np.int32)
    obs = np.concatenate([obs, obs[:40]])  # duplicate observers too
    sheds = orc.calc_sheds(n, dem, roi, 100, 100, obs)
    for iv in (0, 1):
        sel_r, order_r, st_r = orc.site(n, roi, iv, obs, sheds)
        sel, st = gpu.site_it(n, roi, iv, obs.size, obs, sheds)
        np.testing.assert_array_equal(sel, sel_r)
        assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
            (st_r.nused, st_r.niter, st_r.cumarea, st_r.sum_u)


def test_repeat_deterministic(gpu):
    """Same job twice on the device: identical bits (no race in SITE)."""
    n = 256
    dem = make_dem(gpu, n, 0.6, 30000, 21)
    p = gpu.make_params(n, 30, 50, 40, 8, 1000, 1)
    outs = []
    for _ in range(2):
        with gpu.Job(p) as job:
            job.upload_dem(dem)
            job.run()
            job.sync()
            st = job.stats()
            outs.append((job.download_selected(), job.download_order(st.nused),
                         st.cumarea))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


def test_cli_drop_in(gpu, orc, tmp_path):
    """hipsite: same observers file and SITE summary as the sequential path."""
    n, roi, ht, nt, bw, nw0, iv = 240, 24, 100, 40, 10, 576, 1
    dem = make_dem(gpu, n, 0.8, 20000, 31)
    dem_path = tmp_path / "dem.bin"
    dem.tofile(dem_path)
    out_path = tmp_path / "observers.txt"
    cmd = [gpu.CLI_PATH, str(n), str(roi), str(ht), str(nt), str(bw), str(nw0),
           str(iv), str(dem_path), str(out_path)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv)
    expect = "".join(f"{o // n},{o % n}\n"
                     for o, s in zip(ref["obs"], ref["selected"]) if s)
    assert out_path.read_text() == expect
    st = ref["stats"]
    line = r.stdout.strip()
    assert line.startswith("input:")
    keys = [tok.split(":")[0] for tok in line.split()]
    assert keys == ["input", "vix", "findmax", "viewshed", "nusedsheds", "coverage",
                    "site", "output", "total"]
    assert f" nusedsheds:{st.nused} coverage:{st.coverage:g} " in line


@pytest.mark.parametrize("iv", [0, 1])
def test_site_global_candidate_mode(gpu, orc, monkeypatch, iv):
    """SITE's global-memory candidate state: a loop workgroup keeps its
    candidates' ids, positions, areas and flags in LDS when they fit (about
    9600 with the loop's 132 KiB), else in a global slice of the same layout.
    One workgroup (the SITING_SITE_GROUPS tuning cap) over 19044 candidates
    (1100^2, bw 8) takes the global slice."""
    n, roi, ht, nt, bw, nw0 = 1100, 30, 100, 32, 8, 19044
    print(f"[case] SITE global candidate mode iv={iv}", flush=True)
    dem = make_dem(gpu, n, 0.7, 20000, 71 + iv)
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv, nthreads=8)
    assert ref["derived"].nwanted == 19044
    monkeypatch.setenv("SITING_SITE_GROUPS", "1")
    with gpu.Job(gpu.make_params(n, roi, ht, nt, bw, nw0, iv)) as job:
        job.upload_dem(dem)
        job.run()
        job.sync()
        st = job.stats()
        np.testing.assert_array_equal(job.download_selected(), ref["selected"])
        np.testing.assert_array_equal(job.download_order(st.nused), ref["order"])
        rs = ref["stats"]
        assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
            (rs.nused, rs.niter, rs.cumarea, rs.sum_u)
