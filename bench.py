"""Siting throughput benchmark on one MI355X.

Workloads, all on synthetic seeded terrain (project generator: diamond-square
roughness H, linear rescale to 0..hmax, 8 Gaussian hills of peak 5000 and
sigma 80 cells):

  uneven batch  64 independent jobs: edge log-uniform in [512, 4096], H in
                [0.5, 0.9], heights 0..20000, roi in {25,50,100,150},
                targets in {32,64}, bw in {8,16}, nwanted = n^2/256,
                height 100, intervis 0.  A step is the whole batch through
                VIX -> FINDMAX -> VIEWSHED -> SITE.  Timed batched (one
                batch object: stream pool, one work-queue SITE launch, HIP
                graph replay), with the DEMs in HBM (`value`) and from host
                DEMs to host `selected` masks (PCIe included); and plain (the
                jobs one by one, each stage synced).
  single jobs   1000^2 (roi 50, 50 targets, bw 10, nwanted 10000), 4096^2
                (roi 100, 64 targets, bw 16, nwanted 65536) and the
                inter-visible 2048^2 (H 0.6, heights 0..30000, roi 80,
                height 50, 100 targets, bw 8, nwanted 32768, intervis 1):
                points per second on the plain path, per-stage HIP-event
                times of the 4096^2 job and its outputs' digests against the
                oracle's (tests/golden/configs_digest.json).
  served stream Poisson arrivals of jobs with edge log-uniform in
                [256, 8192], roi in {25,50,100,200}, 64 targets, bw 16,
                nwanted n^2/512, intervis 1 with p = 0.2, through the request
                loop (admission by HBM need, jobs sharing the device): 256
                requests at 8 per second (throughput, p50/p99 latency from
                submit to results), and a sweep of the offered rate to
                saturation.
  cpu baseline  the reference's own OpenMP code (siting.cpp built with its
                makefile's flags, oracle/_ref) on this host's cores: its
                stage functions on a sample of the uneven batch, and its
                ompsite program on the 1000^2 workload.

Prints one JSON line; `value` is the batched throughput (terrain points per
second over the whole batch, DEMs resident in HBM).
"""
import argparse
import concurrent.futures as cf
import hashlib
import json
import math
import os
import subprocess
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("terrain pts sited/sec batched vs plain GPU (bit-exact); SITE HBM GB/s; "
          "p99 job lat")

# The single-job workloads: the CPU-sized one (1000^2), the large plain one
# (4096^2, timed here stage by stage) and the inter-visible one (2048^2,
# bw 8, which the reference's CLI rejects, so it runs through the library).
C1 = dict(nrows=1000, roi=50, height=100, ntests=50, bw=10, nwanted0=10000,
          intervis=0, rough_h=0.8, hmax=20000, seed=1)
SINGLE = dict(nrows=4096, roi=100, height=100, ntests=64, bw=16, nwanted0=65536,
              intervis=0, rough_h=0.8, hmax=20000, seed=2)
C3 = dict(nrows=2048, roi=80, height=50, ntests=100, bw=8, nwanted0=32768,
          intervis=1, rough_h=0.6, hmax=30000, seed=3)

PMC_SUMMARY = os.path.join(ROOT, "profiles", "site_loop_queue_uneven_batch64_pmc.md")
DIGESTS = os.path.join(ROOT, "tests", "golden", "configs_digest.json")


def uneven_batch(njobs=64, seed=2024):
    """The uneven batch: one seeded draw per job."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        n = int(round(math.exp(rng.uniform(math.log(512), math.log(4096)))))
        jobs.append(dict(nrows=n, rough_h=float(rng.uniform(0.5, 0.9)), hmax=20000,
                         roi=int(rng.choice([25, 50, 100, 150])),
                         ntests=int(rng.choice([32, 64])), bw=int(rng.choice([8, 16])),
                         nwanted0=n * n // 256, height=100, intervis=0, seed=5000 + j))
    return jobs


def serve_stream(nreq=32, qps=8.0, seed=77):
    """The served stream: Poisson arrivals at a fixed rate; one seeded draw per
    request (edge log-uniform in [256, 8192], roi in {25,50,100,200}, 64
    targets, bw 16, nwanted n^2/512, intervis 1 with p = 0.2).  The requests'
    parameters do not depend on the rate, only their arrival times do."""
    import numpy as np
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for j in range(nreq):
        t += float(rng.exponential(1.0 / qps))
        n = int(round(math.exp(rng.uniform(math.log(256), math.log(8192)))))
        reqs.append(dict(at=t, nrows=n, rough_h=0.8, hmax=20000,
                         roi=int(rng.choice([25, 50, 100, 200])), ntests=64, bw=16,
                         nwanted0=n * n // 512, height=100,
                         intervis=int(rng.random() < 0.2), seed=9000 + j))
    return reqs


def make_dem(sa, w):
    return sa.dem_generate(sa.TerrainSpec(nrows=w["nrows"], rough_h=w["rough_h"], hmin=0,
                                          hmax=w["hmax"], nhills=8, hill_peak=5000.0,
                                          hill_sigma=80.0, seed=w["seed"]))


def make_dems(sa, works):
    """The DEMs of many workloads, generated on host threads (the generator
    is a C call that releases the interpreter lock)."""
    threads = min(16, len(os.sched_getaffinity(0)))
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda w: make_dem(sa, w), works))


def params_of(sa, w):
    return sa.make_params(w["nrows"], w["roi"], w["height"], w["ntests"], w["bw"],
                          w["nwanted0"], w["intervis"])


def algorithmic_bytes(n, d, st):
    """Algorithmic HBM bytes per stage for one job: the bytes the reference's
    sequential algorithm moves.  VIX reads the DEM and writes vix; FINDMAX
    scans vix once per observer wanted per block (nwpb passes) and writes
    obs; VIEWSHED writes every shed and reads the DEM; SITE reads one shed
    per union and per update-list re-evaluation (sum_u + nused of them) and
    scans area (4 B) and used (1 B) of every candidate in each iteration
    (siting.cpp:540-562, the loop of site_it)."""
    S = 8 * d.shed_words
    nsheds = d.nwanted
    return {
        "vix": 2 * n * n + n * n,                       # DEM in, vix out
        "findmax": d.nwpb * n * n + 4 * nsheds,         # vix scans, obs out
        "viewshed": nsheds * S + 2 * n * n,             # sheds out, DEM in
        "site": S * (st.sum_u + st.nused) + 5 * nsheds * st.niter,
    }


def pmc_traffic():
    """HBM bytes per launch of the batch's SITE work-queue kernel, from the two
    rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE, scaled by a plain copy)
    recorded in the profile summary; None when it is absent."""
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith("| traffic per launch |"):
                return int(float(line.split("|")[2].split()[0]) * 1e9)
    except OSError:
        pass
    return None


def sha(a):
    h = hashlib.sha256()
    h.update(memoryview(a).cast("B"))
    return h.hexdigest()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# uneven batch
# ---------------------------------------------------------------------------

def run_batch(sa, np, args):
    batch = uneven_batch(args.jobs)
    dems = make_dems(sa, batch)
    pts = sum(w["nrows"] ** 2 for w in batch)
    log(f"uneven batch: {len(batch)} jobs, {pts} points, DEMs generated")
    params = [params_of(sa, w) for w in batch]
    b = sa.Batch(params, use_graph=not args.no_graph)
    for i, dem in enumerate(dems):
        b.upload_dem(i, dem)
    for i in range(args.warmup):
        b.run()
        b.sync()
        log(f"batch warmup {i} done")
    # each step is synced so that the HIP events around the SITE work-queue
    # launch of every step can be read
    site_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
        b.sync()
        site_ms.append(b.site_ms())
    t_batch = (time.perf_counter() - t0) / args.steps
    dev_batch_ms = b.elapsed_ms()
    site_ms_avg = sum(site_ms) / len(site_ms)
    log(f"batched: {1000 * t_batch:.1f} ms/step (device {dev_batch_ms:.1f} ms last step, "
        f"SITE work queue {site_ms_avg:.1f} ms)")
    # the same, from DEMs in host memory to selected masks in host memory
    host_steps = max(1, min(args.steps, 5))
    sel_host = [np.empty(sa.derive(p).nwanted, np.uint8) for p in params]
    t0 = time.perf_counter()
    for _ in range(host_steps):
        for i, dem in enumerate(dems):
            b.upload_dem(i, dem)
        b.run()
        b.sync()
        for i in range(len(batch)):
            b.job(i).download_into(sa.BUF_SELECTED, sel_host[i])
    t_host = (time.perf_counter() - t0) / host_steps
    log(f"batched, host to host: {1000 * t_host:.1f} ms/step")
    out = []
    for i in range(len(batch)):
        j = b.job(i)
        st = j.stats()
        out.append((j.download_selected(), j.download_order(st.nused), j.download_obs(),
                    st))
    site_bytes = sum(algorithmic_bytes(w["nrows"], sa.derive(params[i]), out[i][3])["site"]
                     for i, w in enumerate(batch))
    host_same = all(np.array_equal(sel_host[i], out[i][0]) for i in range(len(batch)))
    b.close()

    # plain: the same jobs one at a time, each stage synced as the CLI does
    jobs = []
    for i, w in enumerate(batch):
        j = sa.Job(params[i])
        j.upload_dem(dems[i])
        jobs.append(j)
    for j in jobs[:4]:  # warm the code paths
        j.run()
        j.sync()
    t0 = time.perf_counter()
    for j in jobs:
        j.vix()
        j.sync()
        j.findmax()
        j.sync()
        j.viewshed()
        j.sync()
        j.site()
        j.sync()
    t_plain = time.perf_counter() - t0
    same = host_same
    for i, j in enumerate(jobs):
        st = j.stats()
        same = same and (np.array_equal(j.download_selected(), out[i][0]) and
                         np.array_equal(j.download_order(st.nused), out[i][1]) and
                         np.array_equal(j.download_obs(), out[i][2]) and
                         (st.nused, st.niter, st.cumarea, st.sum_u) ==
                         (out[i][3].nused, out[i][3].niter, out[i][3].cumarea,
                          out[i][3].sum_u))
        j.close()
    log(f"plain one-by-one: {1000 * t_plain:.1f} ms; batch==plain: {same}")
    return {"points": pts, "t_batch": t_batch, "t_host": t_host, "t_plain": t_plain,
            "dev_ms": dev_batch_ms, "site_ms": site_ms_avg, "site_bytes": site_bytes,
            "same": same, "njobs": len(batch), "batch": batch}


# ---------------------------------------------------------------------------
# single jobs
# ---------------------------------------------------------------------------

def tuning_report(sa, params, used):
    """The kernel shape choices a job ran with, and whether they came from
    the tuning cache (zeros: the built-in rules)."""
    cached = sa.tuning_lookup(sa.device_key(0), params)
    return {"from_cache": cached is not None and cached.as_dict() == used.as_dict(),
            **used.as_dict()}


def run_single(sa, np, w, steps, digest=None, stages=False):
    n = w["nrows"]
    dem = make_dem(sa, w)
    job = sa.Job(params_of(sa, w))
    job.upload_dem(dem)
    d = job.derived
    job.run()
    job.sync()
    stage_ms = {"vix": 0.0, "findmax": 0.0, "viewshed": 0.0, "site": 0.0}
    t0 = time.perf_counter()
    for _ in range(steps):
        job.run()
        job.sync()
        st = job.stats()
        stage_ms["vix"] += st.ms_vix
        stage_ms["findmax"] += st.ms_findmax
        stage_ms["viewshed"] += st.ms_viewshed
        stage_ms["site"] += st.ms_site
    t = (time.perf_counter() - t0) / steps
    res = {"points_per_s": round(n * n / t, 1), "ms_per_step": round(1000 * t, 3),
           "site": {"nused": st.nused, "niter": st.niter, "sum_u": st.sum_u,
                    "rounds": st.rounds, "coverage": st.coverage},
           "tuning": tuning_report(sa, job.params, job.get_tuning())}
    if stages:
        abytes = algorithmic_bytes(n, d, st)
        res["stages"] = {}
        for k in stage_ms:
            ms = stage_ms[k] / steps
            res["stages"][k] = {"ms": round(ms, 4), "alg_bytes": abytes[k],
                                "alg_gb_s": round(abytes[k] / (ms * 1e-3) / 1e9, 2)
                                if ms > 0 else None}
        # VIEWSHED writes every shed word once and reads the DEM: its
        # algorithmic bytes are its HBM traffic
        vs = res["stages"]["viewshed"]
        vs["hbm_frac"] = round(vs["alg_gb_s"] / HBM_PEAK_GBS, 4) if vs["alg_gb_s"] else None
        # SITE: the reference algorithm's bytes per second of the stage (the
        # GPU loop reads a fraction of them; see the roofline entry)
        res["stages"]["site"]["ref_work_gb_s"] = res["stages"]["site"].pop("alg_gb_s")
    if digest is not None:
        res["parity_vs_oracle"] = bool(
            sha(job.download_vix()) == digest["vix_after"] and
            sha(job.download_obs()) == digest["obs"] and
            sha(job.download_sheds()) == digest["sheds"] and
            sha(job.download_selected()) == digest["selected"] and
            sha(job.download_order(st.nused)) == digest["order"] and
            (st.nused, st.niter, st.cumarea, st.sum_u) ==
            (digest["nused"], digest["niter"], digest["cumarea"], digest["sum_u"]))
    job.close()
    return res


# ---------------------------------------------------------------------------
# served stream
# ---------------------------------------------------------------------------

def serve_once(sa, np, reqs, dems, qps):
    """Submits reqs at their arrival times for rate qps (Poisson, fixed seed)
    and waits for all; returns throughput and latency figures."""
    sched = serve_stream(len(reqs), qps)
    srv = sa.Server(max_active=8, site_groups=32)
    ids = [None] * len(reqs)

    def submitter():
        t0 = time.perf_counter()
        for i, w in enumerate(reqs):
            delay = sched[i]["at"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids[i] = srv.submit(params_of(sa, w), dems[i])

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    infos = [srv.wait(i, 600.0) for i in ids]
    srv.close()
    done = [x for x in infos if x.state == 2]
    lat = np.array([x.t_done - x.t_submit for x in done])
    qwait = np.array([x.t_admit - x.t_submit for x in done])
    span = max(x.t_done for x in done) - min(x.t_submit for x in done)
    pts = sum(w["nrows"] ** 2 for w, x in zip(reqs, infos) if x.state == 2)
    return {"requests": len(reqs), "done": len(done), "qps_offered": qps,
            "points_per_s": round(pts / span, 1), "jobs_per_s": round(len(done) / span, 3),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "p50_queue_s": round(float(np.percentile(qwait, 50)), 4),
            "max_latency_s": round(float(lat.max()), 4), "span_s": round(span, 3)}


def run_serve(sa, np, nreq, qps, sweep, sweep_req):
    reqs = [{k: v for k, v in w.items() if k != "at"}
            for w in serve_stream(max(nreq, sweep_req))]
    dems = make_dems(sa, reqs)
    log(f"serve: {len(reqs)} request DEMs generated")
    main = serve_once(sa, np, reqs[:nreq], dems[:nreq], qps)
    log(f"serve at {qps}/s: {main}")
    rows = []
    for q in sweep:
        r = serve_once(sa, np, reqs[:sweep_req], dems[:sweep_req], q)
        log(f"serve sweep {q}/s: {r}")
        rows.append(r)
    main["sweep"] = rows
    # the knee: the highest offered rate the loop still keeps up with
    # (achieved jobs/s at least 90 % of offered); saturation: the best rate
    kept = [r for r in rows if r["jobs_per_s"] >= 0.9 * r["qps_offered"]]
    knee = max(kept, key=lambda r: r["qps_offered"]) if kept else None
    best = max(rows, key=lambda r: r["points_per_s"]) if rows else None
    main["knee"] = {"qps_offered": knee["qps_offered"], "jobs_per_s": knee["jobs_per_s"],
                    "points_per_s": knee["points_per_s"],
                    "p99_latency_s": knee["p99_latency_s"]} if knee else None
    main["saturation_points_per_s"] = best["points_per_s"] if best else None
    main["saturation_jobs_per_s"] = best["jobs_per_s"] if best else None
    return main


# ---------------------------------------------------------------------------
# CPU baseline: the reference's OpenMP code
# ---------------------------------------------------------------------------

def cpu_sample(batch, points=24_000_000):
    """The first jobs of the uneven batch (its draw order is random) up to
    about `points` terrain points."""
    out, tot = [], 0
    for w in batch:
        if tot >= points:
            break
        out.append(w)
        tot += w["nrows"] ** 2
    return out


def run_cpu_baseline(batch, threads):
    sample = cpu_sample(batch)
    req = {"threads": threads, "jobs": sample, "cli": C1}
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", json.dumps(req)],
                       cwd=ROOT, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    line = [x for x in r.stdout.splitlines() if x.startswith("__CPU_BASELINE__ ")][-1]
    res = json.loads(line.split(" ", 1)[1])
    secs = sum(sum(j["stages_s"]) for j in res["jobs"])
    pts = sum(j["nrows"] ** 2 for j in res["jobs"])
    cli = res["cli"]
    cli_secs = cli["vix"] + cli["findmax"] + cli["viewshed"] + cli["site"]
    return {
        "value": round(pts / secs, 1), "unit": "points/s", "cores": threads,
        "kind": "reference",
        "sample": f"the first {len(sample)} jobs of the uneven batch ({pts} points), "
                  "through the reference's OpenMP stage functions (siting.cpp built "
                  f"g++ -O2 -fopenmp), {threads} threads; stage wall times summed",
        "seconds": round(secs, 3),
        "host_nproc": res["nproc"], "host_affinity_cpus": res["affinity"],
        "ompsite_1000": {
            "workload": "1000^2, roi 50, height 100, 50 targets, bw 10, nwanted 10000",
            "threads": 50, "stage_s": {k: cli[k] for k in ("vix", "findmax", "viewshed",
                                                           "site")},
            "points_per_s": round(1000 * 1000 / cli_secs, 1),
            "nusedsheds": int(cli["nusedsheds"])},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-singles", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-qps", type=float, default=8.0)
    ap.add_argument("--serve-sweep", type=str, default="16,32,64,128")
    ap.add_argument("--serve-sweep-requests", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU design is job-level data parallelism, not built here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import siting_amd as sa

    bt = run_batch(sa, np, args)
    value = bt["points"] / bt["t_batch"]

    digests = None
    if os.path.exists(DIGESTS):
        with open(DIGESTS) as f:
            digests = json.load(f)
    singles = {}
    if not args.no_singles:
        singles["4096"] = run_single(sa, np, SINGLE, args.steps,
                                     digests["C2"] if digests else None, stages=True)
        singles["4096"]["workload"] = ("4096^2 fractal+hills DEM, roi 100, height 100, "
                                       "targets 64, bw 16, nwanted 65536, intervis 0")
        log(f"single 4096^2: {singles['4096']}")
        singles["2048_intervis"] = run_single(sa, np, C3, args.steps,
                                              digests["C3"] if digests else None)
        singles["2048_intervis"]["workload"] = (
            "2048^2 H 0.6 heights 0..30000, roi 80, height 50, targets 100, bw 8, "
            "nwanted 32768, intervis 1")
        log(f"single 2048^2 intervis: {singles['2048_intervis']}")
        singles["1000"] = run_single(sa, np, C1, args.steps,
                                     digests["C1"] if digests else None)
        singles["1000"]["workload"] = ("1000^2, roi 50, height 100, targets 50, bw 10, "
                                       "nwanted 10000, intervis 0")
        log(f"single 1000^2: {singles['1000']}")

    serve = None
    if args.serve_requests > 0:
        sweep = [float(x) for x in args.serve_sweep.split(",") if x]
        serve = run_serve(sa, np, args.serve_requests, args.serve_qps, sweep,
                          args.serve_sweep_requests)

    # the dominant kernel: the batch's SITE work-queue launch.  achieved = its
    # measured HBM traffic per launch (counter runs, profiles/) over the
    # launch time measured here with HIP events
    traffic = pmc_traffic()
    ms = bt["site_ms"]
    ach = traffic / (ms * 1e-3) / 1e9 if traffic else None
    roofline = {"kernel": "site_loop_queue (uneven batch, one launch per step)",
                "bound": "hbm", "achieved": round(ach, 2) if ach else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4) if ach else None,
                "traffic": traffic, "ms_per_launch": round(ms, 3),
                "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT),
                "ref_work_bytes_per_launch": bt["site_bytes"],
                "ref_work_gb_s": round(bt["site_bytes"] / (ms * 1e-3) / 1e9, 2)}

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * bt["t_batch"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u64",
        "data": "synthetic",
        "config": {"workload": f"uneven batch of {bt['njobs']} siting jobs (edge 512..4096 "
                               "log-uniform, roi 25..150, targets 32/64, bw 8/16, "
                               "nwanted n^2/256, intervis 0)",
                   "points": bt["points"],
                   "batched_points_per_s": round(value, 1),
                   "batched_host_to_host_points_per_s": round(bt["points"] / bt["t_host"], 1),
                   "batched_host_to_host_ms_per_step": round(1000 * bt["t_host"], 3),
                   "plain_one_by_one_points_per_s": round(bt["points"] / bt["t_plain"], 1),
                   "batched_over_plain": round(bt["t_plain"] / bt["t_batch"], 3),
                   "batch_equals_plain": bt["same"],
                   "batch_device_ms": round(bt["dev_ms"], 3),
                   "site_queue_ms": round(bt["site_ms"], 3),
                   "graph": not args.no_graph,
                   "tuning_cache": os.path.relpath(sa.tuning_cache_path(), ROOT)
                   if sa.tuning_cache_path() else None,
                   "device_key": sa.device_key(0),
                   "batch_jobs_tuned_from_cache": sum(
                       sa.tuning_lookup(sa.device_key(0), params_of(sa, w)) is not None
                       for w in bt["batch"])},
        "single_jobs": singles,
        "serve": serve,
        "p99_job_latency_s": serve["p99_latency_s"] if serve else None,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline (reference OpenMP build) with {threads} threads")
        result["cpu_baseline"] = run_cpu_baseline(bt["batch"], threads)
        log(f"cpu baseline: {result['cpu_baseline']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
