"""GPU: batched jobs of uneven size and parameters give, for every job, the
same bits as the CPU oracle (and so as running the job alone), with and
without HIP-graph replay, and run after run."""
import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu

# (n, roi, height, ntests, bw, nwanted0, intervis, rough_h, hmax, seed)
JOBS = [
    (300, 25, 100, 32, 8, 300 * 300 // 256, 0, 0.5, 20000, 41),
    (180, 50, 100, 64, 16, 180 * 180 // 256, 0, 0.9, 20000, 42),
    (256, 30, 50, 48, 10, 400, 1, 0.6, 30000, 43),
    (97, 12, 100, 32, 8, 200, 0, 0.7, 20000, 44),
    (400, 40, 100, 64, 16, 400 * 400 // 256, 0, 0.8, 20000, 45),
    (64, 100, 100, 16, 10, 100, 1, 0.7, 20000, 46),
    (222, 18, 0, 1, 7, 700, 0, 0.8, 20000, 47),
]


@pytest.fixture(scope="module")
def refs(sa, orc):
    out = []
    for (n, roi, ht, nt, bw, nw0, iv, H, hmax, seed) in JOBS:
        dem = make_dem(sa, n, H, hmax, seed)
        out.append((dem, orc.run(n, dem, roi, ht, nt, bw, nw0, iv)))
    return out


@pytest.mark.parametrize("use_graph,slot,fixed", [(False, 16, True), (True, 16, False),
                                                 (True, 4, False), (False, 8, False)])
def test_batch_matches_oracle(gpu, refs, use_graph, slot, fixed):
    print(f"[case] batch graph={use_graph} slot={slot} fixed={fixed}", flush=True)
    params = [gpu.make_params(*c[:7]) for c in JOBS]
    with gpu.Batch(params, use_graph=use_graph, slot=slot, fixed_slots=fixed) as b:
        for i, (dem, _) in enumerate(refs):
            b.upload_dem(i, dem)
        # replays: state is rebuilt each run, and after the first the slots
        # may be re-planned from measured times
        for rep in range(3):
            b.run()
            b.sync()
            for i, (dem, ref) in enumerate(refs):
                j = b.job(i)
                st = j.stats()
                rs = ref["stats"]
                np.testing.assert_array_equal(j.download_obs(), ref["obs"], err_msg=str(i))
                np.testing.assert_array_equal(j.download_selected(), ref["selected"],
                                              err_msg=str(i))
                assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
                    (rs.nused, rs.niter, rs.cumarea, rs.sum_u), (i, rep)
        assert b.elapsed_ms() > 0
