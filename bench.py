"""Siting throughput benchmark on one MI355X.

Two workloads, both on synthetic seeded terrain (project generator:
diamond-square roughness H, linear rescale to 0..hmax, 8 Gaussian hills of
peak 5000 and sigma 80 cells), DEMs in HBM before any timed region:

  uneven batch  64 independent jobs: edge log-uniform in [512, 4096], H in
                [0.5, 0.9], heights 0..20000, roi in {25,50,100,150},
                targets in {32,64}, bw in {8,16}, nwanted = n^2/256,
                height 100, intervis 0.  A step is the whole batch through
                VIX -> FINDMAX -> VIEWSHED -> SITE.  Timed batched (one
                batch object: stream pool, one work-queue SITE launch, HIP
                graph replay) and plain (the jobs one by one, each synced).
  single job    4096^2, H 0.8, roi 100, height 100, 64 targets, bw 16,
                nwanted 65536, intervis 0: per-stage HIP-event times,
                algorithmic bytes and the roofline entry of the dominant
                stage; bit-exactness against the CPU oracle; the oracle's
                OpenMP variant timed on this host's cores (cpu_baseline).

  served stream  Poisson arrivals at a fixed rate of jobs with edge
                log-uniform in [256, 8192], roi in {25,50,100,200}, 64 targets,
                bw 16, nwanted n^2/512, intervis 1 with p = 0.2, through the
                request loop (admission by HBM need, jobs sharing the device):
                throughput and p50/p99 latency (submit to results).

Prints one JSON line; `value` is the batched throughput (terrain points per
second over the whole batch).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("terrain pts sited/sec batched vs plain GPU (bit-exact); SITE HBM GB/s; "
          "p99 job lat")

SINGLE = dict(nrows=4096, roi=100, height=100, ntests=64, bw=16, nwanted0=65536,
              intervis=0, rough_h=0.8, hmax=20000, seed=2)


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
    targets, bw 16, nwanted n^2/512, intervis 1 with p = 0.2)."""
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


def run_serve(sa, np, nreq, qps):
    import threading
    reqs = serve_stream(nreq, qps)
    dems = [make_dem(sa, w) for w in reqs]
    log(f"serve: {len(reqs)} requests at {qps}/s, DEMs generated")
    srv = sa.Server(max_active=8, site_groups=32)
    ids = [None] * len(reqs)

    def submitter():
        t0 = time.perf_counter()
        for i, w in enumerate(reqs):
            delay = w["at"] - (time.perf_counter() - t0)
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
            "max_latency_s": round(float(lat.max()), 4)}


def make_dem(sa, w):
    return sa.dem_generate(sa.TerrainSpec(nrows=w["nrows"], rough_h=w["rough_h"], hmin=0,
                                          hmax=w["hmax"], nhills=8, hill_peak=5000.0,
                                          hill_sigma=80.0, seed=w["seed"]))


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


PMC_SUMMARY = os.path.join(ROOT, "profiles", "site_loop_queue_uneven_batch64_pmc.md")


def pmc_traffic():
    """Bytes per launch of the batch's SITE work-queue kernel, as measured by
    the two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE, scaled by a plain
    copy) recorded in the profile summary; None when it is absent."""
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith("| traffic per launch |"):
                return int(float(line.split("|")[2].split()[0]) * 1e9)
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-qps", type=float, default=8.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU design is job-level data parallelism, not built here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import siting_amd as sa

    # ---------------- uneven batch: batched vs plain ----------------------
    batch = uneven_batch(args.jobs)
    dems = [make_dem(sa, w) for w in batch]
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
    batch_sel = [b.job(i).download_selected() for i in range(len(batch))]
    batch_stats = [b.job(i).stats() for i in range(len(batch))]
    batch_cum = [st.cumarea for st in batch_stats]
    site_bytes = sum(algorithmic_bytes(w["nrows"], sa.derive(params[i]), batch_stats[i])["site"]
                     for i, w in enumerate(batch))
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
    same = all(np.array_equal(jobs[i].download_selected(), batch_sel[i]) and
               jobs[i].stats().cumarea == batch_cum[i] for i in range(len(jobs)))
    for j in jobs:
        j.close()
    log(f"plain one-by-one: {1000 * t_plain:.1f} ms; batch==plain: {same}")

    # ---------------- single job: stages, roofline, oracle ------------------
    w = SINGLE
    n = w["nrows"]
    dem = sa.dem_generate(sa.TerrainSpec(nrows=n, rough_h=w["rough_h"], hmin=0,
                                         hmax=w["hmax"], nhills=8, hill_peak=5000.0,
                                         hill_sigma=80.0, seed=w["seed"]))
    p = sa.make_params(n, w["roi"], w["height"], w["ntests"], w["bw"], w["nwanted0"],
                       w["intervis"])
    job = sa.Job(p)
    job.upload_dem(dem)
    d = job.derived
    job.run()
    job.sync()
    stage_ms = {"vix": 0.0, "findmax": 0.0, "viewshed": 0.0, "site": 0.0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.run()
        job.sync()
        st = job.stats()
        stage_ms["vix"] += st.ms_vix
        stage_ms["findmax"] += st.ms_findmax
        stage_ms["viewshed"] += st.ms_viewshed
        stage_ms["site"] += st.ms_site
    t_single = (time.perf_counter() - t0) / args.steps
    for k in stage_ms:
        stage_ms[k] /= args.steps
    sel = job.download_selected()
    obs = job.download_obs()
    abytes = algorithmic_bytes(n, d, st)
    stages = {}
    for k, ms in stage_ms.items():
        gbs = abytes[k] / (ms * 1e-3) / 1e9 if ms > 0 else None
        stages[k] = {"ms": round(ms, 4), "bytes": abytes[k],
                     "gb_s": round(gbs, 2) if gbs else None}
    site_gbs = abytes["site"] / (stage_ms["site"] * 1e-3) / 1e9
    # the batch's SITE work-queue launch: algorithmic SITE bytes of all jobs
    # over the launch's HIP-event time
    ach = site_bytes / (site_ms_avg * 1e-3) / 1e9
    roofline = {"kernel": "site_loop_queue (uneven batch, one launch per step)",
                "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                "bytes_per_launch": site_bytes, "ms_per_launch": round(site_ms_avg, 3),
                "traffic": pmc_traffic()}
    log(f"single job: {1000 * t_single:.2f} ms/step, stages {stage_ms}")

    serve = None
    if args.serve_requests > 0:
        serve = run_serve(sa, np, args.serve_requests, args.serve_qps)
        log(f"serve: {serve}")

    value = pts / t_batch
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * t_batch, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u64",
        "data": "synthetic",
        "config": {"workload": f"uneven batch of {len(batch)} siting jobs (edge 512..4096 "
                               "log-uniform, roi 25..150, targets 32/64, bw 8/16, "
                               "nwanted n^2/256, intervis 0)",
                   "points": pts,
                   "batched_points_per_s": round(value, 1),
                   "plain_one_by_one_points_per_s": round(pts / t_plain, 1),
                   "batched_over_plain": round(t_plain / t_batch, 3),
                   "batch_equals_plain": same,
                   "batch_device_ms": round(dev_batch_ms, 3),
                   "site_queue_ms": round(site_ms_avg, 3),
                   "graph": not args.no_graph},
        "single_job": {"workload": "4096^2 fractal+hills DEM, roi 100, height 100, "
                                   "targets 64, bw 16, nwanted 65536, intervis 0",
                       "points_per_s": round(n * n / t_single, 1),
                       "ms_per_step": round(1000 * t_single, 3),
                       "stages": stages,
                       "site_hbm_gb_s": round(site_gbs, 2),
                       "site": {"nused": st.nused, "niter": st.niter, "sum_u": st.sum_u,
                                "coverage": st.coverage}},
        "serve": serve,
        "p99_job_latency_s": serve["p99_latency_s"] if serve else None,
        "roofline": roofline,
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline with {threads} threads")
        import oracle
        t0 = time.perf_counter()
        ref = oracle.run(n, dem, w["roi"], w["height"], w["ntests"], w["bw"],
                         w["nwanted0"], w["intervis"], nthreads=threads)
        dt = time.perf_counter() - t0
        parity = bool(np.array_equal(sel, ref["selected"]) and
                      np.array_equal(obs, ref["obs"]) and
                      st.cumarea == ref["stats"].cumarea)
        result["single_job"]["parity_vs_oracle"] = parity
        result["cpu_baseline"] = {
            "value": round(n * n / dt, 1), "unit": "points/s", "cores": threads,
            "kind": "port",
            "sample": "the single-job workload once (4096^2, all four stages), OpenMP "
                      "variant of the C restatement of the sequential reference",
            "seconds": round(dt, 3)}
    job.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
