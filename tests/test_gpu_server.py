"""GPU: the request loop.  A mix of jobs submitted at once (so they share
the device) gives, for every job, the same bits as the CPU oracle, i.e. as
the job run alone; bad requests are rejected without stopping the server;
a tight memory budget forces queueing and backfill, with the same results."""
import numpy as np
import pytest

from conftest import make_dem

pytestmark = pytest.mark.gpu

# (n, roi, height, ntests, bw, nwanted0, intervis, rough_h, hmax, seed)
MIX = [
    (256, 25, 100, 64, 16, 256 * 256 // 512, 0, 0.8, 20000, 61),
    (300, 50, 100, 64, 16, 300 * 300 // 512, 1, 0.7, 20000, 62),
    (180, 20, 100, 64, 16, 180 * 180 // 512, 0, 0.6, 20000, 63),
    (400, 30, 100, 64, 16, 400 * 400 // 512, 1, 0.9, 20000, 64),
    (128, 60, 100, 64, 16, 128 * 128 // 512, 0, 0.5, 20000, 65),
    (350, 25, 100, 64, 16, 350 * 350 // 512, 0, 0.8, 20000, 66),
]


@pytest.fixture(scope="module")
def mix(sa, orc):
    out = []
    for (n, roi, ht, nt, bw, nw0, iv, H, hmax, seed) in MIX:
        dem = make_dem(sa, n, H, hmax, seed)
        out.append((sa.make_params(n, roi, ht, nt, bw, nw0, iv), dem,
                    orc.run(n, dem, roi, ht, nt, bw, nw0, iv)))
    return out


def _check_all(srv, mix, ids):
    for jid, (p, dem, ref) in zip(ids, mix):
        info = srv.wait(jid, 120.0)
        assert info.state == 2, (jid, info.state, info.status)  # DONE
        sel, obs = srv.result(jid)
        np.testing.assert_array_equal(obs, ref["obs"])
        np.testing.assert_array_equal(sel, ref["selected"])
        rs = ref["stats"]
        assert (info.stats.nused, info.stats.cumarea, info.stats.sum_u) == \
            (rs.nused, rs.cumarea, rs.sum_u)
        assert info.t_submit <= info.t_admit <= info.t_done
        srv.release(jid)


def test_server_mix_equals_alone(gpu, mix):
    print("[case] server mix", flush=True)
    with gpu.Server(max_active=6, site_groups=16) as srv:
        ids = [srv.submit(p, dem) for (p, dem, _) in mix]
        _check_all(srv, mix, ids)


def test_server_tight_memory_backfill(gpu, mix):
    print("[case] server tight budget", flush=True)
    need = max(gpu.Job.bytes_needed(p) for (p, _, _) in mix)
    with gpu.Server(hbm_bytes=int(need * 1.6), max_active=4, site_groups=8,
                    backfill_ms=50) as srv:
        ids = [srv.submit(p, dem) for (p, dem, _) in mix]
        _check_all(srv, mix, ids)


def test_server_rejects_bad_requests(gpu, mix):
    print("[case] server rejects", flush=True)
    p, dem, ref = mix[0]
    with gpu.Server(hbm_bytes=64 << 20, max_active=2) as srv:
        bad = gpu.make_params(p.nrows, 0, 100, 64, 16, 100, 0)  # roi 0
        with pytest.raises(gpu.SitingError) as e:
            srv.submit(bad, dem)
        assert e.value.status == gpu.EINVAL
        big = gpu.make_params(4096, 200, 100, 64, 16, 65536, 0)
        with pytest.raises(gpu.SitingError) as e:
            srv.submit(big, np.zeros(4096 * 4096, np.uint16))
        assert e.value.status == gpu.ENOMEM
        jid = srv.submit(p, dem)  # still serving
        info = srv.wait(jid, 120.0)
        assert info.state == 2
        sel, _ = srv.result(jid)
        np.testing.assert_array_equal(sel, ref["selected"])


def test_server_destroy_wakes_blocked_waiters(gpu, mix):
    """Destroying the server while other threads are blocked waiting on its
    requests: each waiter returns with its request DONE or FAILED (dropped
    from the queue), and destroy returns only after they have left."""
    import threading
    print("[case] server destroy with blocked waiters", flush=True)
    import time
    srv = gpu.Server(max_active=1, site_groups=8)
    ids = [srv.submit(p, dem) for _ in range(4) for (p, dem, _) in mix]
    states, errors = {}, []

    def waiter(jid):
        try:
            states[jid] = srv.wait(jid, 60.0).state
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    ths = [threading.Thread(target=waiter, args=(j,)) for j in ids]
    for t in ths:
        t.start()
    time.sleep(0.05)  # every waiter is blocked in wait; most jobs still queued
    srv.close()
    for t in ths:
        t.join(30.0)
        assert not t.is_alive()
    assert not errors, errors
    assert sorted(states) == sorted(ids)
    assert all(s in (2, 3) for s in states.values())  # DONE or FAILED
