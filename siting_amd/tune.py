"""Tuner: measures kernel shape choices per class of job and writes the
tuning cache the library reads (include/siting.h, "tuning cache").

    python -m siting_amd.tune [--out PATH] [--reps 3] [--device 0]

For each class -- roi 25/50/100/150/200, targets 32/64, edge 1024/2048/4096
-- a seeded synthetic terrain of the class's edge (roughness 0.7, heights
0..20000, bw 16, nwanted n^2/256) runs once with the built-in rules, then
each stage is timed (HIP events, best of --reps) under every choice of its
own knobs, the others held at the rules:

  VIX       LDS halo 8, 16, 32
  VIEWSHED  threads per observer 128, 256, 512, times bitset mode (two LDS
            bitsets, one LDS bitset, HBM; the modes whose bitsets fit)
  SITE      column classes 2, 3, 4, times stale threshold shift 2 to 6;
            and, on the same terrain with intervis 1, column classes 2 to 4,
            times the pick walk's exact area computations per round: none,
            2, 8 or 16.  The column count kept is the one whose best
            intervis-0 and intervis-1 times sum least (both kinds of job
            share it), with its own best shift and computations

A choice replaces the rule's only when it is at least 3 % faster.  The job
is then run again with the chosen values, and its vix, observers, viewshed
bitsets and selection must equal the rules' bit for bit, or the tuner stops.
Each line of the cache: device key, roi class, targets class, edge class,
halo, threads, mode, column classes, shift, 1 + walk computations, then the
measured stage times as a comment.
"""
from __future__ import annotations

import argparse
import hashlib
import os
import sys
import time

import numpy as np

import siting_amd as sa

ROIS = (25, 50, 100, 150, 200)
NTESTS = (32, 64)
EDGES = (1024, 2048, 4096)
QS = (2, 3, 4)  # SITE column classes tried
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune", "tuning.txt")


def _sha(a: np.ndarray) -> str:
    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()


def _outputs(job: sa.Job):
    st = job.stats()
    return (_sha(job.download_vix()), _sha(job.download_obs()), _sha(job.download_sheds()),
            _sha(job.download_selected()), (st.nused, st.niter, st.cumarea, st.sum_u))


def _time_stage(job: sa.Job, stage: str, reps: int) -> float:
    best = float("inf")
    for _ in range(reps):
        getattr(job, stage)()
        job.sync()
        st = job.stats()
        best = min(best, {"vix": st.ms_vix, "viewshed": st.ms_viewshed,
                          "site": st.ms_site}[stage])
    return best


def _pick(times: dict, rule_key) -> tuple:
    """The fastest choice, unless the rule's is within 3 %."""
    best = min(times, key=times.get)
    if rule_key in times and times[best] > 0.97 * times[rule_key]:
        return rule_key
    return best


def viewshed_modes(roi: int) -> list[int]:
    """Bitset modes whose bitsets fit (1 split pair, 2 one LDS bitset, 3 HBM),
    as the library decides (common.hpp vs_split; one bitset up to 8192
    words)."""
    nr = 2 * roi + 1
    nwpr = (nr + 63) // 64
    modes = []
    if 2 * 4 * nr * (2 * nwpr + 1) <= 48 * 1024:
        modes.append(1)
    if nr * nwpr <= 8192:
        modes.append(2)
    modes.append(3)
    return modes


def rule_choice(roi: int, ntests: int, n: int, nsheds: int) -> dict:
    """The built-in rules' values for a job (vix.hip, viewshed.hip, site.hip)."""
    modes = viewshed_modes(roi)
    q_list = 3.14159265358979 * 4.0 * roi * roi * nsheds / (n * n)
    return {"vix_halo": 32 if roi >= 64 and ntests >= 64 else 16,
            "viewshed_threads": 256, "viewshed_mode": modes[0],
            "site_q": 3 if q_list >= 3000.0 else 2, "site_tau_shift": 3,
            "site_walk_evals": 17 if roi <= 40 else 9 if roi > 75 else 1}


def tune_class(key: str, roi: int, ntests: int, edge: int, reps: int, log):
    n = edge
    w = dict(nrows=n, roi=roi, height=100, ntests=ntests, bw=16, nwanted0=n * n // 256)
    p = sa.make_params(n, roi, 100, ntests, 16, n * n // 256, 0)
    d = sa.derive(p)
    dem = sa.dem_generate(sa.TerrainSpec(nrows=n, rough_h=0.7, hmin=0, hmax=20000,
                                         seed=31000 + roi * 10 + ntests + edge))
    rule = rule_choice(roi, ntests, n, d.nwanted)
    with sa.Job(p) as job:
        job.upload_dem(dem)
        base = sa.Tuning(**rule)
        job.set_tuning(base)
        job.run()
        job.sync()
        ref = _outputs(job)
        choice = dict(rule)
        times = {}
        # VIX
        tv = {}
        for h in (8, 16, 32):
            job.set_tuning(sa.Tuning(**dict(rule, vix_halo=h)))
            tv[h] = _time_stage(job, "vix", reps)
        choice["vix_halo"] = _pick(tv, rule["vix_halo"])
        times["vix"] = tv[choice["vix_halo"]]
        job.set_tuning(base)
        job.vix()
        job.findmax()
        job.sync()
        # VIEWSHED
        tw = {}
        for th in (128, 256, 512):
            for m in viewshed_modes(roi):
                job.set_tuning(sa.Tuning(**dict(rule, viewshed_threads=th, viewshed_mode=m)))
                tw[(th, m)] = _time_stage(job, "viewshed", reps)
        th, m = _pick(tw, (rule["viewshed_threads"], rule["viewshed_mode"]))
        choice["viewshed_threads"], choice["viewshed_mode"] = th, m
        times["viewshed"] = tw[(th, m)]
        # SITE: each column count's best stale threshold shift
        ts = {}
        for q in QS:
            for s in (2, 3, 4, 5, 6):
                job.set_tuning(sa.Tuning(**dict(rule, site_q=q, site_tau_shift=s)))
                ts[(q, s)] = _time_stage(job, "site", reps)
        best_s = {q: _pick({s: t for (qq, s), t in ts.items() if qq == q},
                           rule["site_tau_shift"]) for q in QS}
        rule_ms = {"vix": tv[rule["vix_halo"]],
                   "viewshed": tw[(rule["viewshed_threads"], rule["viewshed_mode"])],
                   "site": ts[(rule["site_q"], rule["site_tau_shift"])]}
    # intervis, on the same terrain: each column count's best number of the
    # pick walk's exact computations
    pi = sa.make_params(n, roi, 100, ntests, 16, n * n // 256, 1)
    with sa.Job(pi) as job:
        job.upload_dem(dem)
        job.set_tuning(sa.Tuning(**choice))
        job.run()
        job.sync()
        ref_iv = _outputs(job)
        te = {}
        for q in QS:
            for ev in (1, 3, 9, 17):
                job.set_tuning(sa.Tuning(**dict(choice, site_q=q, site_walk_evals=ev)))
                te[(q, ev)] = _time_stage(job, "site", max(1, reps - 1))
        best_ev = {q: _pick({ev: t for (qq, ev), t in te.items() if qq == q},
                            rule["site_walk_evals"]) for q in QS}
        # one column count serves both kinds of job: the one whose intervis-0
        # and intervis-1 loops take least time together (measured on the
        # 2048^2 roi-80 intervis job: 88.9 ms with 3, 100.8 with 2, where the
        # intervis-0 loop alone had chosen 2)
        both = {q: ts[(q, best_s[q])] + te[(q, best_ev[q])] for q in QS}
        q = _pick(both, rule["site_q"])
        choice["site_q"] = q
        choice["site_tau_shift"] = best_s[q]
        choice["site_walk_evals"] = best_ev[q]
        times["site"] = ts[(q, best_s[q])]
        times["site_iv"] = te[(q, best_ev[q])]
        rule_ms["site_iv"] = te[(rule["site_q"], rule["site_walk_evals"])]
        job.set_tuning(sa.Tuning(**choice))
        job.site()
        job.sync()
        if _outputs(job) != ref_iv:
            raise SystemExit(f"tuned intervis run differs from the rules' for {w}")
    # the chosen values must give the rules' bits
    with sa.Job(p) as job:
        job.upload_dem(dem)
        job.set_tuning(sa.Tuning(**choice))
        job.run()
        job.sync()
        if _outputs(job) != ref:
            raise SystemExit(f"tuned run differs from the rules' for {w}")
    log(f"{key} roi {roi} targets {ntests} edge {edge}: {choice} "
        f"ms {times} (rules {rule_ms})")
    return choice, times, rule_ms


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--rois", default=",".join(map(str, ROIS)))
    ap.add_argument("--edges", default=",".join(map(str, EDGES)))
    args = ap.parse_args(argv)
    key = sa.device_key(args.device)

    def log(msg):
        print(f"[tune] {msg}", file=sys.stderr, flush=True)

    lines = ["# siting tuning cache, written by python -m siting_amd.tune",
             "# key roi targets edge vix_halo viewshed_threads viewshed_mode site_q "
             "site_tau_shift site_walk_evals  # tuned ms (rules' ms): vix viewshed site "
             "site_intervis"]
    t0 = time.time()
    for roi in [int(x) for x in args.rois.split(",")]:
        for ntests in NTESTS:
            for edge in [int(x) for x in args.edges.split(",")]:
                c, t, r = tune_class(key, roi, ntests, edge, args.reps, log)
                lines.append(
                    f"{key} {roi} {ntests} {edge} {c['vix_halo']} {c['viewshed_threads']} "
                    f"{c['viewshed_mode']} {c['site_q']} {c['site_tau_shift']} "
                    f"{c['site_walk_evals']}  # "
                    f"{t['vix']:.3f} {t['viewshed']:.3f} {t['site']:.3f} {t['site_iv']:.3f} "
                    f"({r['vix']:.3f} {r['viewshed']:.3f} {r['site']:.3f} "
                    f"{r['site_iv']:.3f})")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    log(f"wrote {args.out} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
