"""GPU parity at the five benchmark workloads, at their full sizes.

C1  the reference's own example: its sample terrain dem.bin with the README
    parameters (1000^2, roi 50, height 100, 50 targets, bw 10, nwanted
    10000), through the drop-in CLI and through the stage functions on host
    arrays, against the sequential reference build's observers file and
    summary (tests/golden/readme_example.json, written by seqsite); and the
    generated 1000^2 terrain of the same parameters.
C2  4096^2, roi 100, 64 targets, bw 16, nwanted 65536.
C3  2048^2, H 0.6, heights 0..30000, roi 80, height 50, 100 targets, bw 8
    (below the CLI's minimum, so through the library), nwanted 32768 ->
    65536, intervis 1.
C4  the 64-job uneven batch, through one batch object (stream pool, HIP
    graph, work-queue SITE with re-planned slots), every job checked.
C5  the first 32 requests of the served stream through the request loop,
    arriving at 8 per second, every request checked.

Expected outputs are the oracle's (oracle/, the C restatement of the
sequential reference), computed once per workload by
tests/golden/make_config_digests.py and committed as sha256 digests of every
stage output and SITE's exact counters; C1, C2 and C3 there were also checked
word for word against the reference's own stage functions.  The DEM digest
is checked first, so a generator difference is not mistaken for a kernel
one.  Bit-exact: no tolerance.
"""
import gzip
import hashlib
import json
import os
import subprocess
import threading
import time

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def sha(a: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(memoryview(np.ascontiguousarray(a)).cast("B"))
    return h.hexdigest()


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLDEN, "configs_digest.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def readme():
    with open(os.path.join(GOLDEN, "readme_example.json")) as f:
        g = json.load(f)
    with gzip.open(os.path.join(GOLDEN, "dem1000.u16.gz"), "rb") as f:
        raw = f.read()
    assert hashlib.sha256(raw).hexdigest() == g["dem_sha256"]
    return g, raw


def _params(gpu, w):
    return gpu.make_params(w["nrows"], w["roi"], w["height"], w["ntests"], w["bw"],
                           w["nwanted0"], w["intervis"])


def _check_site(st, e, tag):
    assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
        (e["nused"], e["niter"], e["cumarea"], e["sum_u"]), tag


def _check_job(j, e, tag, sheds=True):
    st = j.stats()
    _check_site(st, e, tag)
    assert sha(j.download_vix()) == e["vix_after"], tag
    assert sha(j.download_obs()) == e["obs"], tag
    assert sha(j.download_selected()) == e["selected"], tag
    assert sha(j.download_order(st.nused)) == e["order"], tag
    if sheds:
        assert sha(j.download_sheds()) == e["sheds"], tag


def test_c1_readme_example_cli(gpu, readme, tmp_path):
    g, raw = readme
    print("[case] C1 README example, hipsite", flush=True)
    dem_path = tmp_path / "dem.bin"
    dem_path.write_bytes(raw)
    out_path = tmp_path / "observers.txt"
    cmd = [gpu.CLI_PATH] + [str(v) for v in g["args"]] + [str(dem_path), str(out_path)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert out_path.read_text() == g["observers"]
    assert f" nusedsheds:{g['nusedsheds']} coverage:{g['coverage_printed']} " in r.stdout


def test_c1_readme_example_stage_functions(gpu, orc, readme):
    g, raw = readme
    print("[case] C1 README example, stage functions", flush=True)
    n, roi, ht, nt, bw, nw0, iv = g["args"]
    dem = np.frombuffer(raw, dtype=np.uint16)[: n * n].copy()
    d = gpu.derive(gpu.make_params(n, roi, ht, nt, bw, nw0, iv))
    vix = gpu.calc_vix(n, dem, roi, ht, ht, nt)
    ref = orc.run(n, dem, roi, ht, nt, bw, nw0, iv, nthreads=8)
    np.testing.assert_array_equal(vix, ref["vix"])
    obs = gpu.find_max(n, vix, d.blocksize, d.nwanted, d.nblocks, d.nwpb)
    np.testing.assert_array_equal(obs, ref["obs"])
    np.testing.assert_array_equal(vix, ref["vix_after"])
    sheds = gpu.calc_sheds(n, dem, roi, ht, ht, d.nwanted, obs)
    np.testing.assert_array_equal(sheds, ref["sheds"])
    sel, st = gpu.site_it(n, roi, iv, d.nwanted, obs, sheds)
    text = "".join(f"{o // n},{o % n}\n" for o, s in zip(obs, sel) if s)
    assert text == g["observers"]
    assert st.nused == g["nusedsheds"]
    assert f"{st.coverage:g}" == g["coverage_printed"]


@pytest.mark.parametrize("name", ["C1", "C2", "C3"])
def test_single_job_workload(gpu, digests, name):
    """Every stage of one job, through the device-resident job API."""
    w = {"C1": bench.C1, "C2": bench.SINGLE, "C3": bench.C3}[name]
    e = digests[name]
    assert e["work"] == {k: w[k] for k in e["work"]}
    print(f"[case] {name} {w}", flush=True)
    dem = bench.make_dem(gpu, w)
    assert sha(dem) == e["dem"], "terrain generator differs"
    with gpu.Job(_params(gpu, w)) as job:
        job.upload_dem(dem)
        job.vix()
        job.sync()
        assert sha(job.download_vix()) == e["vix"]
        job.findmax()
        job.viewshed()
        job.site()
        job.sync()
        _check_job(job, e, name)


def test_c4_uneven_batch(gpu, digests):
    """The 64-job batch: two runs (uniform slots, then re-planned ones from the
    first run's loop times, through the HIP graph), every job's outputs."""
    work = bench.uneven_batch()
    es = digests["C4"]
    assert len(es) == len(work)
    dems = []
    for i, w in enumerate(work):
        assert es[i]["work"] == {k: w[k] for k in es[i]["work"]}
        dems.append(bench.make_dem(gpu, w))
        assert sha(dems[-1]) == es[i]["dem"], f"terrain generator differs, job {i}"
    print("[case] C4 uneven batch of 64 jobs", flush=True)
    with gpu.Batch([_params(gpu, w) for w in work], use_graph=True) as b:
        for i, dem in enumerate(dems):
            b.upload_dem(i, dem)
        for rep in range(2):
            b.run()
            b.sync()
            print(f"[case] C4 run {rep} done", flush=True)
            for i in range(len(work)):
                _check_job(b.job(i), es[i], f"C4 job {i} run {rep}", sheds=rep == 1)


def test_c5_served_stream(gpu, digests):
    """The served stream's first 32 requests, submitted at their Poisson
    arrival times (8 per second), so that requests join jobs already on the
    device; every result equals the job's alone (the oracle's)."""
    reqs = bench.serve_stream(32)
    es = digests["C5"]
    dems = []
    for i, w in enumerate(reqs):
        assert es[i]["work"] == {k: w[k] for k in es[i]["work"]}
        dems.append(bench.make_dem(gpu, w))
        assert sha(dems[-1]) == es[i]["dem"], f"terrain generator differs, request {i}"
    print("[case] C5 served stream, 32 requests", flush=True)
    ids = [None] * len(reqs)
    with gpu.Server(max_active=8, site_groups=32) as srv:
        def submitter():
            t0 = time.perf_counter()
            for i, w in enumerate(reqs):
                delay = w["at"] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[i] = srv.submit(_params(gpu, w), dems[i])

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        overlapped = 0
        infos = []
        for i, jid in enumerate(ids):
            info = srv.wait(jid, 110.0)
            assert info.state == 2, (i, info.state, info.status)  # DONE
            infos.append(info)
            sel, obs = srv.result(jid)
            e = es[i]
            tag = f"C5 request {i}"
            assert sha(obs) == e["obs"], tag
            assert sha(sel) == e["selected"], tag
            _check_site(info.stats, e, tag)
            srv.release(jid)
        # continuous batching: some request was admitted while another ran
        for i, x in enumerate(infos):
            if any(y.t_admit < x.t_admit < y.t_done for j, y in enumerate(infos) if j != i):
                overlapped += 1
        assert overlapped > 0
