"""CPU baseline: the reference's own OpenMP path, timed -- TEST INFRASTRUCTURE.

Run as a child process by bench.py's cpu_baseline leg (the reference's
stage functions print to stdout, which the bench keeps for its JSON line):

    python -m oracle.cpu_baseline '<json request>'

Request fields:
  threads    OpenMP threads for the stage functions
  jobs       workloads (bench.py dicts: nrows, roi, height, ntests, bw,
             nwanted0, intervis, rough_h, hmax, seed); their DEMs are
             generated here with the project's seeded generator
  cli        optional workload run through the reference's ompsite program
             itself (siting.cpp's main, 50 threads fixed at siting.cpp:635),
             whose stdout stage timers are parsed

What runs is oracle/_ref/libsiting_ref_omp.so and oracle/_ref/ompsite: the
untouched siting.cpp built with the reference makefile's OpenMP flags
(makefile:8, `g++ -O2 -fopenmp`).  Its SITE argmax reduction is
nondeterministic in tie order (siting.cpp:565-587), so these runs are timed,
never used for parity.  Prints one JSON line: per job the four stage times
(wall clock around each stage call, as siting.cpp:735-759 times them) and
nusedsheds, and the CLI's parsed line.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_OMP_LIB = os.path.join(HERE, "_ref", "libsiting_ref_omp.so")
REF_OMPSITE = os.path.join(HERE, "_ref", "ompsite")


def _dem(sa, w):
    return sa.dem_generate(sa.TerrainSpec(nrows=w["nrows"], rough_h=w["rough_h"], hmin=0,
                                          hmax=w["hmax"], nhills=8, hill_peak=5000.0,
                                          hill_sigma=80.0, seed=w["seed"]))


def _derive(w):
    """siting.cpp:689-692 (the oracle's restatement)."""
    import oracle
    rc, d = oracle.derive(w["nrows"], w["bw"], w["nwanted0"])
    if rc:
        raise ValueError(f"derive failed for {w}")
    return d.blocksize, d.nblocks, d.nwpb, d.nwanted


def run_jobs(jobs, threads):
    import siting_amd as sa
    L = ctypes.CDLL(REF_OMP_LIB)
    VP, I = ctypes.c_void_p, ctypes.c_int
    L.ref_set_threads.argtypes = [I]
    L.ref_calc_vix.argtypes = [I, VP, I, I, I, I, VP]
    L.ref_find_max.argtypes = [I, VP, ctypes.c_float, I, I, I, VP]
    L.ref_calc_sheds.argtypes = [I, VP, I, I, I, I, VP, VP]
    L.ref_site_it.argtypes = [I, I, I, I, VP, VP, VP]
    L.ref_set_threads(threads)
    out = []
    for w in jobs:
        n, roi, ht = w["nrows"], w["roi"], w["height"]
        dem = _dem(sa, w)
        bs, nb, nwpb, nwanted = _derive(w)
        nr = 2 * roi + 1
        nwpr = (nr + 63) // 64
        vix = np.empty(n * n, np.uint8)
        obs = np.empty(nwanted, np.int32)
        sheds = np.empty(nwanted * nr * nwpr, np.uint64)
        sel = np.empty(nwanted, np.uint8)
        t = [time.perf_counter()]
        L.ref_calc_vix(n, dem.ctypes.data, roi, ht, ht, w["ntests"], vix.ctypes.data)
        t.append(time.perf_counter())
        L.ref_find_max(n, vix.ctypes.data, bs, nwanted, nb, nwpb, obs.ctypes.data)
        t.append(time.perf_counter())
        L.ref_calc_sheds(n, dem.ctypes.data, roi, ht, ht, nwanted, obs.ctypes.data,
                         sheds.ctypes.data)
        t.append(time.perf_counter())
        L.ref_site_it(n, roi, w["intervis"], nwanted, obs.ctypes.data, sheds.ctypes.data,
                      sel.ctypes.data)
        t.append(time.perf_counter())
        out.append({"nrows": n, "roi": roi, "bw": w["bw"], "ntests": w["ntests"],
                    "stages_s": [round(t[k + 1] - t[k], 4) for k in range(4)],
                    "nused": int(sel.sum())})
    return out


def run_cli(w):
    import siting_amd as sa
    dem = _dem(sa, w)
    with tempfile.TemporaryDirectory() as td:
        dp, op = os.path.join(td, "dem.bin"), os.path.join(td, "observers.txt")
        dem.tofile(dp)
        args = [str(w[k]) for k in ("nrows", "roi", "height", "ntests", "bw", "nwanted0",
                                    "intervis")]
        r = subprocess.run([REF_OMPSITE] + args + [dp, op], capture_output=True, text=True,
                           check=True, timeout=600)
    toks = dict(t.split(":") for t in r.stdout.split() if ":" in t)
    return {k: float(v) for k, v in toks.items()}


def main():
    req = json.loads(sys.argv[1])
    res = {"threads": req["threads"], "nproc": os.cpu_count(),
           "affinity": len(os.sched_getaffinity(0))}
    if req.get("jobs"):
        res["jobs"] = run_jobs(req["jobs"], req["threads"])
    if req.get("cli"):
        res["cli"] = run_cli(req["cli"])
    sys.stdout.flush()
    print("\n__CPU_BASELINE__ " + json.dumps(res), flush=True)


if __name__ == "__main__":
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    main()
