"""The SITE loop's measurement knobs give the oracle's selection.

The loop keeps, behind environment knobs, the variants its defaults were
measured against: decrements instead of exact counts in the update lists
(SITING_SITE_FULL_AT), the walk's exact computations per round
(SITING_SITE_NEAR_EVALS, up to 32, where a round can step past the 64
sorted entries the walk holds), the stale-bound threshold
(SITING_SITE_TAU_SHIFT, 0 turning stale bounds off), the column classes per
cell row (SITING_SITE_Q) and the loop's workgroups (SITING_SITE_GROUPS, 2
giving each workgroup update lists of several windows).  Whatever they
choose, every selection, count and area must equal the oracle's.
"""
import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu

KNOBS = [
    {},
    {"SITING_SITE_FULL_AT": "2"},
    {"SITING_SITE_FULL_AT": "1000"},
    {"SITING_SITE_FULL_AT": "1000", "SITING_SITE_GROUPS": "2"},
    {"SITING_SITE_GROUPS": "2"},
    {"SITING_SITE_NEAR_EVALS": "32"},
    {"SITING_SITE_NEAR_EVALS": "0"},
    {"SITING_SITE_TAU_SHIFT": "0"},
    {"SITING_SITE_TAU_SHIFT": "2"},
    {"SITING_SITE_Q": "1"},
    {"SITING_SITE_Q": "4"},
]


@pytest.fixture(scope="module")
def case(sa, orc):
    n, roi = 300, 20
    dem = make_dem(sa, n, 0.7, 20000, 17)
    rng = np.random.default_rng(23)
    obs = rng.choice(n * n, size=4000, replace=False).astype(np.int32)
    sheds = orc.calc_sheds(n, dem, roi, 100, 100, obs)
    refs = {iv: orc.site(n, roi, iv, obs, sheds) for iv in (0, 1)}
    return n, roi, obs, sheds, refs


@pytest.mark.parametrize("intervis", [0, 1])
@pytest.mark.parametrize("knobs", KNOBS, ids=lambda k: ",".join(f"{a[12:]}={b}" for a, b in k.items()) or "default")
def test_site_knobs_match_oracle(gpu, case, knobs, intervis, monkeypatch):
    n, roi, obs, sheds, refs = case
    for name, value in knobs.items():
        monkeypatch.setenv(name, value)
    print(f"[case] n {n} roi {roi} intervis {intervis} knobs {knobs}", flush=True)
    sel, st = gpu.site_it(n, roi, intervis, obs.size, obs, sheds)
    sel_r, order_r, st_r = refs[intervis]
    assert st_r.nused > 50  # many rounds
    np.testing.assert_array_equal(sel, sel_r)
    assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
        (st_r.nused, st_r.niter, st_r.cumarea, st_r.sum_u)


# roi 130: a candidate's shed box has 261 rows of 5-6 terrain words, more
# (row, word) items than the fifteen helper wavefronts' 960 lanes, so each of
# the walk's exact computations takes them in two passes
@pytest.fixture(scope="module")
def case_wide(sa, orc):
    n, roi = 700, 130
    dem = make_dem(sa, n, 0.7, 20000, 29)
    rng = np.random.default_rng(31)
    obs = rng.choice(n * n, size=3000, replace=False).astype(np.int32)
    sheds = orc.calc_sheds(n, dem, roi, 100, 100, obs)
    refs = {iv: orc.site(n, roi, iv, obs, sheds) for iv in (0, 1)}
    return n, roi, obs, sheds, refs


@pytest.mark.parametrize("intervis", [0, 1])
@pytest.mark.parametrize("knobs", [{}, {"SITING_SITE_NEAR_EVALS": "32"}],
                         ids=lambda k: ",".join(f"{a[12:]}={b}" for a, b in k.items()) or "default")
def test_site_wide_boxes_match_oracle(gpu, case_wide, knobs, intervis, monkeypatch):
    n, roi, obs, sheds, refs = case_wide
    for name, value in knobs.items():
        monkeypatch.setenv(name, value)
    print(f"[case] n {n} roi {roi} intervis {intervis} knobs {knobs}", flush=True)
    sel, st = gpu.site_it(n, roi, intervis, obs.size, obs, sheds)
    sel_r, order_r, st_r = refs[intervis]
    assert st_r.nused > 500
    np.testing.assert_array_equal(sel, sel_r)
    assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
        (st_r.nused, st_r.niter, st_r.cumarea, st_r.sum_u)
