"""GPU: siting_site_it_noalloc, SITE on device arrays with its working memory
in the library's fixed device array, gives the same selection, order and
counters as siting_site_it (the host-array entry point) -- on two calls in a
row, the second on a smaller input, so the shared array carries the first
call's contents into the second.  A call beyond the array's size is
rejected with SITING_ENOMEM and a message naming the size.

The inputs come from a job's device buffers (its obs and sheds after VIX,
FINDMAX and VIEWSHED), the outputs go to a second job's selected and order
buffers: the test itself allocates no device memory outside the library.
"""
import ctypes

import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu

# (n, roi, height, ntests, bw, nwanted0, intervis, rough_h, seed): the second
# case is smaller in every dimension
CASES = [(420, 30, 100, 32, 10, 1764, 0, 0.8, 81),
         (160, 12, 100, 32, 8, 400, 1, 0.6, 82)]


def _dev(gpu, job, which):
    p = ctypes.c_void_p()
    nb = ctypes.c_size_t()
    rc = gpu.lib().siting_job_device_ptr(job._h, which, ctypes.byref(p), ctypes.byref(nb))
    assert rc == gpu.OK
    return p.value, nb.value


def _run_case(gpu, case):
    n, roi, ht, nt, bw, nw0, iv, H, seed = case
    print(f"[case] noalloc n{n} roi{roi} iv{iv}", flush=True)
    dem = make_dem(gpu, n, H, 20000, seed)
    p = gpu.make_params(n, roi, ht, nt, bw, nw0, iv)
    d = gpu.derive(p)
    with gpu.Job(p) as src, gpu.Job(p) as out:
        src.upload_dem(dem)
        src.vix()
        src.findmax()
        src.viewshed()
        src.sync()
        obs = src.download_obs()
        sheds = src.download_sheds()
        # the existing entry point, on host copies of the same inputs
        sel_ref, st_ref = gpu.site_it(n, roi, iv, d.nwanted, obs, sheds)
        obs_d, _ = _dev(gpu, src, gpu.BUF_OBS)
        sheds_d, _ = _dev(gpu, src, gpu.BUF_SHEDS)
        sel_d, _ = _dev(gpu, out, gpu.BUF_SELECTED)
        order_d, _ = _dev(gpu, out, gpu.BUF_ORDER)
        st = gpu.Stats()
        rc = gpu.lib().siting_site_it_noalloc(n, roi, iv, d.nwanted, obs_d, sheds_d, sel_d,
                                              order_d, ctypes.byref(st), None)
        assert rc == gpu.OK, gpu.lib().siting_status_str(rc)
        sel = out.download_selected()
        order = out.download_order(st.nused)
    np.testing.assert_array_equal(sel, sel_ref)
    assert (st.nused, st.niter, st.cumarea, st.sum_u) == \
        (st_ref.nused, st_ref.niter, st_ref.cumarea, st_ref.sum_u)
    assert st.coverage == st_ref.coverage
    # the selection order: every selected candidate once
    assert sorted(order.tolist()) == np.flatnonzero(sel_ref).tolist()
    return order


def test_noalloc_matches_site_it_twice(gpu, orc):
    first = _run_case(gpu, CASES[0])
    second = _run_case(gpu, CASES[1])
    # the order is the sequential reference's (the oracle's)
    for case, order in ((CASES[0], first), (CASES[1], second)):
        n, roi, ht, nt, bw, nw0, iv, H, seed = case
        ref = orc.run(n, make_dem(gpu, n, H, 20000, seed), roi, ht, nt, bw, nw0, iv)
        np.testing.assert_array_equal(order, ref["order"])


def test_noalloc_rejects_beyond_its_array(gpu):
    # roi 10000 needs about 1.6 GB of per-round pick bits
    p = ctypes.c_void_p(1)  # never dereferenced: the size check comes first
    st = gpu.Stats()
    rc = gpu.lib().siting_site_it_noalloc(16384, 10000, 0, 65536, p, p, p, None,
                                          ctypes.byref(st), None)
    assert rc == gpu.ENOMEM
    msg = gpu.lib().siting_status_str(rc).decode()
    assert "siting_site_it_noalloc" in msg and "bytes" in msg
