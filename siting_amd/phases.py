"""Phase profile of the SITE greedy loop: where a round's time goes.

    python -m siting_amd.phases [--jobs single,iv,b42,b49,...] [--groups G,...]

Runs jobs through the profiling build of the library (libsiting_amd_prof.so,
the same kernels with device-clock phase counters compiled in) and prints,
per job, the loop's time per round and phase, averaged over the loop's
workgroups (thread 0 of each, 100 MHz device clock):

  gather    every workgroup reads the published unit keys
  sort      the walk's 256 entries merged into one wavefront and sorted
  walk      the pick walk down the sorted entries (one wavefront; its
            exact area computations are shown separately)
  book      counters, pick hashes and, when the walk read cum, the walk
            barrier
  union     the picks' sheds ORed into cum
  update    update lists of the picks' neighbourhoods (decrements)
  reeval    stale candidates re-counted against cum
  publish   dirty units re-rank and publish their top keys
  endbar    end-of-round grid barrier (waiting for the slowest workgroup)

and why each round's walk ended.  Selections are checked against the
oracle's digests when those are given (bench.py's single-job
workloads and the batch's jobs, tests/golden/configs_digest.json), so a
profile is of the same work as the product's run.

Job names: "single" (4096^2, roi 100, 64 targets, bw 16, nwanted 65536),
"iv" (2048^2 inter-visible, roi 80, bw 8), "c1" (1000^2), and "bN" for job N
of the uneven 64-job batch (bench.py's seeded draw).  --groups runs each job
at those loop workgroup caps (SITING_SITE_GROUPS), --near-evals with those
budgets of the walk's exact computations per round (SITING_SITE_NEAR_EVALS),
instead of the defaults.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys

import numpy as np

import siting_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PHASES = ("walk", "book", "union", "update", "reeval", "publish", "endbar", "gather")
SORT = 20  # the counter of the walk's sort
TRACE_SHAPE = (8, 256, 128)  # marks x rounds x workgroups of the round trace
ENDS = ("budget", "rekeyed_full", "bound_entry", "stale_key", "no_area", "pick_cap",
        "no_pick")
NWORDS = 32


def _workloads(names):
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    batch = None
    out = []
    for nm in names:
        if nm == "single":
            out.append((nm, bench.SINGLE))
        elif nm == "iv":
            out.append((nm, bench.C3))
        elif nm == "c1":
            out.append((nm, bench.C1))
        elif nm.startswith("b") and nm[1:].isdigit():
            batch = batch or bench.uneven_batch()
            out.append((nm, batch[int(nm[1:])]))
        else:
            raise SystemExit(f"unknown job {nm!r}")
    return bench, out


def profile(job: sa.Job, lib) -> dict:
    """Runs the job twice (the first warms up) and returns the second run's
    phase profile."""
    buf = np.zeros(NWORDS, np.uint64)
    trace = np.zeros(TRACE_SHAPE, np.uint64)
    job.run()
    job.sync()
    sa._check(lib.siting_phase_profile_read(buf.ctypes.data, NWORDS, 1))
    sa._check(lib.siting_phase_trace_read(trace.ctypes.data, trace.size))
    job.run()
    job.sync()
    sa._check(lib.siting_phase_profile_read(buf.ctypes.data, NWORDS, 1))
    sa._check(lib.siting_phase_trace_read(trace.ctypes.data, trace.size))
    st = job.stats()
    groups = max(1, int(buf[8]))
    rounds = max(1, st.rounds)
    per = groups * rounds
    us = {nm: round(float(buf[q]) * 0.01 / per, 2) for q, nm in enumerate(PHASES)}
    us["sort"] = round(float(buf[SORT]) * 0.01 / per, 2)
    evals = float(buf[9]) / per
    steps = float(buf[11]) / per
    walk_eval_us = float(buf[10]) * 0.01 / per
    return {
        "site_ms": round(st.ms_site, 3),
        "rounds": st.rounds,
        "groups": groups,
        "picks_per_round": round(float(buf[12]) / per, 2),
        "us_per_round": us,
        "round_us": round(sum(us.values()), 1),
        "walk_computations_per_round": round(evals, 2),
        "walk_computation_us": round(walk_eval_us / evals, 2) if evals else 0.0,
        "walk_steps_per_round": round(steps, 2),
        "walk_step_us": round((us["walk"] - walk_eval_us) / steps, 3) if steps else 0.0,
        "walk_ends": {nm: round(float(buf[13 + q]) / per, 3) for q, nm in enumerate(ENDS)},
        "drift_us": drift(trace, min(groups, TRACE_SHAPE[2]), st.rounds),
        "selected_sha256": _sha(job.download_selected()),
    }


def drift(trace: np.ndarray, groups: int, rounds: int) -> dict:
    """How far the workgroups drift apart within a round, from the round
    trace (device clock at each round's start, key gather end, walk end,
    publish end, and the ends of the bookkeeping, union, update lists and
    re-evaluation, per workgroup), averaged over the traced rounds, in us:
    the spread (latest minus earliest workgroup) at the first four marks,
    the walk's longest duration over its mean, the same for the work after
    the walk (book to publish) and for each phase of it."""
    n = min(rounds, TRACE_SHAPE[1])
    t = trace[:, :n, :groups].astype(np.float64) * 0.01
    ok = (trace[:, :n, :groups] != 0).all(axis=(0, 2))
    if not ok.any():
        return {}
    t = t[:, ok, :]
    walk = t[2] - t[1]
    after = t[3] - t[2]
    spread = t.max(axis=2) - t.min(axis=2)
    # the phases after the walk: (start mark, end mark)
    post = {"book": (2, 4), "union": (4, 5), "update": (5, 6), "reeval": (6, 7),
            "publish": (7, 3)}
    over = {nm: round(float(((t[b] - t[a]).max(axis=1) - (t[b] - t[a]).mean(axis=1)).mean()), 2)
            for nm, (a, b) in post.items()}
    return {
        "rounds_traced": int(ok.sum()),
        "start_spread": round(float(spread[0].mean()), 2),
        "gather_end_spread": round(float(spread[1].mean()), 2),
        "walk_end_spread": round(float(spread[2].mean()), 2),
        "publish_end_spread": round(float(spread[3].mean()), 2),
        "walk_max_over_mean": round(float((walk.max(axis=1) - walk.mean(axis=1)).mean()), 2),
        "after_walk_max_over_mean": round(float((after.max(axis=1) - after.mean(axis=1)).mean()), 2),
        "phase_max_over_mean": over,
        # the workgroups most often slowest in the update lists and the
        # re-evaluation, with the share of rounds (a share near 1/G: random)
        "slowest_update_groups": _most_often((t[6] - t[5]).argmax(axis=1)),
        "slowest_reeval_groups": _most_often((t[7] - t[6]).argmax(axis=1)),
        "last_at_walk_end_is_last_at_publish": round(float(
            (t[2].argmax(axis=1) == t[3].argmax(axis=1)).mean()), 2),
    }


def _most_often(idx: np.ndarray, k: int = 3) -> list:
    vals, counts = np.unique(idx, return_counts=True)
    order = np.argsort(-counts)[:k]
    return [[int(vals[o]), round(float(counts[o]) / len(idx), 2)] for o in order]


def _sha(a):
    h = hashlib.sha256()
    h.update(memoryview(np.ascontiguousarray(a)).cast("B"))
    return h.hexdigest()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--jobs", default="single,iv,b42,b49,b46,b1,b39")
    ap.add_argument("--groups", default="", help="comma-separated loop workgroup caps")
    ap.add_argument("--near-evals", default="",
                    help="comma-separated budgets of the walk's exact computations per "
                         "round (SITING_SITE_NEAR_EVALS)")
    ap.add_argument("--json", action="store_true", help="one JSON object per line")
    args = ap.parse_args(argv)
    lib = sa.use_profiling_build()
    bench, works = _workloads(args.jobs.split(","))
    digests = {}
    try:
        with open(bench.DIGESTS) as f:
            digests = json.load(f)
    except OSError:
        pass
    caps = [int(x) for x in args.groups.split(",") if x] or [0]
    budgets = [int(x) for x in args.near_evals.split(",") if x] or [-1]
    runs = [(c, e) for c in caps for e in budgets]
    for name, w in works:
        dem = bench.make_dem(sa, w)
        with sa.Job(bench.params_of(sa, w)) as job:
            job.upload_dem(dem)
            for cap, ev in runs:
                if cap:
                    os.environ["SITING_SITE_GROUPS"] = str(cap)
                if ev >= 0:
                    os.environ["SITING_SITE_NEAR_EVALS"] = str(ev)
                r = profile(job, lib)
                os.environ.pop("SITING_SITE_GROUPS", None)
                os.environ.pop("SITING_SITE_NEAR_EVALS", None)
                want = None
                ref = {"single": "C2", "iv": "C3", "c1": "C1"}.get(name)
                if ref and ref in digests:
                    want = digests[ref].get("selected")
                elif name.startswith("b") and int(name[1:]) < len(digests.get("C4", [])):
                    want = digests["C4"][int(name[1:])].get("selected")
                if want:
                    r["selected_matches_oracle"] = r["selected_sha256"] == want
                r.update(job=name, nrows=w["nrows"], roi=w["roi"], intervis=w["intervis"],
                         group_cap=cap or None, near_evals=ev if ev >= 0 else None)
                if args.json:
                    print(json.dumps(r), flush=True)
                else:
                    u = r["us_per_round"]
                    print(f"{name:>6} n={w['nrows']} roi={w['roi']} G={r['groups']} "
                          f"SITE {r['site_ms']} ms, {r['rounds']} rounds of "
                          f"{r['round_us']} us ({r['picks_per_round']} picks): "
                          + " ".join(f"{k} {u[k]}" for k in PHASES + ("sort",))
                          + f" | walk: {r['walk_computations_per_round']} computations x "
                          f"{r['walk_computation_us']} us, {r['walk_steps_per_round']} steps x "
                          f"{r['walk_step_us']} us | ends {r['walk_ends']} | drift {r['drift_us']}",
                          flush=True)


if __name__ == "__main__":
    main()
