"""Writes tests/golden/configs_digest.json: the oracle's outputs on the five
benchmark workloads, as sha256 digests and counters.

    python tests/golden/make_config_digests.py [--threads 8] [--ref]

Workloads (definitions in bench.py, all on the project's seeded generator):
  C1  1000^2, H 0.8, roi 50, height 100, 50 targets, bw 10, nwanted 10000
  C2  4096^2, H 0.8, roi 100, height 100, 64 targets, bw 16, nwanted 65536
  C3  2048^2, H 0.6, heights 0..30000, roi 80, height 50, 100 targets, bw 8,
      nwanted 32768, intervis 1
  C4  the 64-job uneven batch (bench.uneven_batch)
  C5  the first 32 requests of the served stream (bench.serve_stream)

Each entry holds the digest of the generated DEM (so that a generator
difference is reported as such), of vix before and after FINDMAX, obs,
every shed word, the selected mask and the selection order, and SITE's
counters (nused, niter, cumarea, sum_u).  The oracle is oracle/liboracle.so
(the C restatement of the sequential reference; its OpenMP variant returns
the sequential bits).  With --ref, C1, C2 and C3 are also run through the
reference's own stage functions (oracle/_ref/libsiting_ref.so, built from
the untouched siting.cpp) and must agree with the oracle word for word; the
file records which workloads were so checked.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import oracle  # noqa: E402
import siting_amd as sa  # noqa: E402

OUT = os.path.join(HERE, "configs_digest.json")


def sha(a: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(memoryview(np.ascontiguousarray(a)).cast("B"))
    return h.hexdigest()


def entry(w, nthreads):
    n = w["nrows"]
    dem = bench.make_dem(sa, w)
    ref = oracle.run(n, dem, w["roi"], w["height"], w["ntests"], w["bw"], w["nwanted0"],
                     w["intervis"], nthreads=nthreads)
    st = ref["stats"]
    e = {"work": {k: w[k] for k in ("nrows", "roi", "height", "ntests", "bw", "nwanted0",
                                    "intervis", "rough_h", "hmax", "seed")},
         "dem": sha(dem), "vix": sha(ref["vix"]), "vix_after": sha(ref["vix_after"]),
         "obs": sha(ref["obs"]), "sheds": sha(ref["sheds"]),
         "selected": sha(ref["selected"]), "order": sha(ref["order"]),
         "nused": int(st.nused), "niter": int(st.niter), "cumarea": int(st.cumarea),
         "sum_u": int(st.sum_u), "coverage": float(st.coverage)}
    return e, dem, ref


def ref_check(w, dem, ref):
    """The reference's own stage functions on the same DEM agree with the
    oracle on every output."""
    from make_golden import ref_stages
    vix0, vix1, obs, sheds, sel = ref_stages(w["nrows"], dem, w["roi"], w["height"],
                                             w["ntests"], w["bw"], w["nwanted0"],
                                             w["intervis"])
    for name, a, b in (("vix", vix0, ref["vix"]), ("vix_after", vix1, ref["vix_after"]),
                       ("obs", obs, ref["obs"]), ("sheds", sheds, ref["sheds"]),
                       ("selected", sel, ref["selected"])):
        if not np.array_equal(a, b):
            raise SystemExit(f"reference and oracle differ on {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--ref", action="store_true")
    args = ap.parse_args()
    if args.ref:
        oracle.build(ref=True)
    # entries are saved as they are made, so an interrupted run resumes
    part = OUT + ".partial"
    out = {"threads": args.threads, "ref_checked": [], "C4": [], "C5": []}
    if os.path.exists(part):
        with open(part) as f:
            out = json.load(f)

    def save():
        with open(part, "w") as f:
            json.dump(out, f)

    t0 = time.time()
    for name, w in (("C1", bench.C1), ("C2", bench.SINGLE), ("C3", bench.C3)):
        if name in out:
            continue
        e, dem, ref = entry(w, args.threads)
        if args.ref:
            ref_check(w, dem, ref)
            out["ref_checked"].append(name)
        out[name] = e
        save()
        print(f"{name} done ({time.time() - t0:.0f} s)", flush=True)
    for i, w in enumerate(bench.uneven_batch()):
        if i < len(out["C4"]):
            continue
        out["C4"].append(entry(w, args.threads)[0])
        save()
        print(f"C4 job {i} done ({time.time() - t0:.0f} s)", flush=True)
    for i, w in enumerate(bench.serve_stream(32)):
        if i < len(out["C5"]):
            continue
        w = {k: v for k, v in w.items() if k != "at"}
        out["C5"].append(entry(w, args.threads)[0])
        save()
        print(f"C5 request {i} done ({time.time() - t0:.0f} s)", flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    os.remove(part)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
