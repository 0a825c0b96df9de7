"""Siting throughput benchmark on one MI355X.

A step is one pass of the four-stage pipeline (VIX -> FINDMAX -> VIEWSHED ->
SITE) over the resident DEM of the plain single-job workload: a 4096x4096
uint16 fractal+hills terrain (H 0.8, heights 0..20000, 8 Gaussian hills of
peak 5000 and sigma 80 cells, seed 2), roi 100, height 100, 64 targets,
bw 16, nwanted 65536, intervis 0.  The DEM is in HBM before the timed region
starts; every stage output stays in HBM.

Prints one JSON line.  Besides throughput it reports, per stage, the HIP
event time measured on the job's stream and the stage's algorithmic bytes,
a roofline entry for the dominant stage, a bit-exactness check of the GPU
result against the CPU oracle, and the oracle's OpenMP variant timed on this
host's cores on the same workload (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(nrows=4096, roi=100, height=100, ntests=64, bw=16, nwanted0=65536,
                intervis=0, rough_h=0.8, hmax=20000, seed=2)


def algorithmic_bytes(n, d, st):
    """Algorithmic HBM bytes per stage for one job (formulas in DESIGN.md)."""
    S = 8 * d.shed_words
    nsheds = d.nwanted
    return {
        "vix": 2 * n * n + n * n,                       # DEM in, vix out
        "findmax": d.nwpb * n * n + 4 * nsheds,         # vix scans, obs out
        "viewshed": nsheds * S + 2 * n * n,             # sheds out, DEM in
        "site": S * (st.sum_u + st.nused) + 5 * nsheds * st.niter,
    }


def cpu_baseline(dem, w, threads):
    import oracle
    t0 = time.perf_counter()
    ref = oracle.run(w["nrows"], dem, w["roi"], w["height"], w["ntests"], w["bw"],
                     w["nwanted0"], w["intervis"], nthreads=threads)
    dt = time.perf_counter() - t0
    return ref, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU design is job-level data parallelism, not built here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import siting_amd as sa

    w = WORKLOAD
    n = w["nrows"]
    dem = sa.dem_generate(sa.TerrainSpec(nrows=n, rough_h=w["rough_h"], hmin=0,
                                         hmax=w["hmax"], nhills=8, hill_peak=5000.0,
                                         hill_sigma=80.0, seed=w["seed"]))
    p = sa.make_params(n, w["roi"], w["height"], w["ntests"], w["bw"], w["nwanted0"],
                       w["intervis"])
    job = sa.Job(p)
    job.upload_dem(dem)
    d = job.derived
    print(f"[bench] workload n={n} nwanted={d.nwanted} job bytes={job.bytes_needed(p)}",
          file=sys.stderr, flush=True)

    for i in range(args.warmup):
        job.run()
        job.sync()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)

    stage_ms = {"vix": 0.0, "findmax": 0.0, "viewshed": 0.0, "site": 0.0}
    job.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.run()
        job.sync()  # per-step sync so every step's stage events are read
        st = job.stats()
        stage_ms["vix"] += st.ms_vix
        stage_ms["findmax"] += st.ms_findmax
        stage_ms["viewshed"] += st.ms_viewshed
        stage_ms["site"] += st.ms_site
    job.sync()
    elapsed = time.perf_counter() - t0
    for k in stage_ms:
        stage_ms[k] /= args.steps
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n * n * args.steps / elapsed
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.2f} ms/step", file=sys.stderr,
          flush=True)

    sel = job.download_selected()
    obs = job.download_obs()
    abytes = algorithmic_bytes(n, d, st)
    stages = {}
    for k, ms in stage_ms.items():
        gbs = abytes[k] / (ms * 1e-3) / 1e9 if ms > 0 else None
        stages[k] = {"ms": round(ms, 4), "bytes": abytes[k],
                     "gb_s": round(gbs, 2) if gbs else None}
    dom = max(stage_ms, key=stage_ms.get)
    ach = abytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    roofline = {"stage": dom, "bound": "hbm", "achieved": round(ach, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                "traffic": None}

    result = {
        "metric": "terrain points sited per second (plain single-job GPU path, "
                  "bit-exact vs sequential reference)",
        "value": round(value, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u64",
        "data": "synthetic",
        "config": {"workload": "plain single-job: 4096^2 uint16 fractal+hills DEM, "
                               "roi 100, height 100, targets 64, bw 16, nwanted 65536, "
                               "intervis 0",
                   "nwanted": d.nwanted, "shed_bytes": 8 * d.shed_words},
        "site": {"nused": st.nused, "niter": st.niter, "sum_u": st.sum_u,
                 "coverage": st.coverage},
        "stages": stages,
        "roofline": roofline,
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"[bench] cpu baseline with {threads} threads", file=sys.stderr, flush=True)
        ref, dt = cpu_baseline(dem, w, threads)
        parity = bool(np.array_equal(sel, ref["selected"]) and
                      np.array_equal(obs, ref["obs"]) and
                      st.cumarea == ref["stats"].cumarea)
        result["parity_vs_oracle"] = parity
        result["cpu_baseline"] = {
            "value": round(n * n / dt, 1), "unit": "points/s", "cores": threads,
            "kind": "port",
            "sample": "the full workload once (all four stages), OpenMP variant of "
                      "the C restatement of the sequential reference",
            "seconds": round(dt, 3)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
