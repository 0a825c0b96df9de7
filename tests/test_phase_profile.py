"""The profiling build of the library (the greedy loop's device-clock phase
counters, siting_amd.phases).

CPU: it exports everything the product library exports, plus the counter
read.  GPU: a job run through it (in its own process, since one process
loads one build) selects exactly what the oracle's digest says, and its
counters are consistent: the phases and the walk ran, and no round's walk
ended twice.
"""
import ctypes
import json
import os
import subprocess
import sys

import pytest

from test_cli_abi import declared_functions

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_profiling_build_exports(sa):
    prof = ctypes.CDLL(sa.PROF_LIB_PATH)
    missing = [fn for fn in declared_functions() if not hasattr(prof, fn)]
    assert not missing
    assert hasattr(prof, "siting_phase_profile_read")
    assert not hasattr(sa.lib(), "siting_phase_profile_read")


@pytest.mark.gpu
def test_phase_profile_of_the_1000_job(gpu):
    out = subprocess.run([sys.executable, "-m", "siting_amd.phases", "--jobs", "c1,iv",
                          "--json"], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    rows = [json.loads(line) for line in out.stdout.splitlines() if line.startswith("{")]
    assert [r["job"] for r in rows] == ["c1", "iv"]
    for r in rows:
        assert r["selected_matches_oracle"], r["job"]
        u = r["us_per_round"]
        assert all(v >= 0 for v in u.values())
        assert u["walk"] > 0 and u["endbar"] > 0 and u["gather"] > 0
        assert r["walk_steps_per_round"] >= r["picks_per_round"] > 0
        # each round's walk ends once: by a counted reason, or (intervis)
        # by a masked candidate's bound, which is not counted
        ends = dict(r["walk_ends"])
        assert ends.pop("no_pick") <= 1.0
        assert sum(ends.values()) <= 1.0 + 1e-6
