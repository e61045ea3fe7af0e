"""The serving loop (fv_server_*): runs that arrive while others are on the
GPU are admitted as device memory frees up, share the batch rounds with the
runs already there, and each result is bit-identical to the same run made
alone (fv_run to its stop time) -- a stricter bar than any tolerance."""
import numpy as np
import pytest

import fosite_amd
from fosite_amd import setups as S


def alone(cfg):
    init = S.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    sim.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    sim.first_step()
    while sim.run(256) == 256:
        pass
    cv, _ = sim.state()
    sc = sim.scalars()
    sim.close()
    return cv, sc


def requests(n, seed):
    out = []
    for t, cfg in S.serving_requests(n=n, seed=seed, nmin=24, nmax=72, orbits=0.5):
        out.append(cfg)
    return out


@pytest.mark.gpu
def test_gpu_server_results_equal_runs_alone():
    cfgs = requests(10, 7)
    # budget for the two largest runs at once: the first six do not fit
    # together, so some queue and are admitted as others finish
    sizes = sorted(fosite_amd.device_bytes(c) for c in cfgs)
    budget = sizes[-1] + sizes[-2]
    assert sum(fosite_amd.device_bytes(c) for c in cfgs[:6]) > budget
    print("server test", [(c["inum"], c["dtype"]) for c in cfgs], "budget %.1f MB" % (budget / 1e6), flush=True)
    srv = fosite_amd.Server(budget, round_steps=4, autotune=False)
    ids = [srv.submit(c, S.initial(c)) for c in cfgs[:6]]
    rounds, saw_queue = 0, False
    while True:
        running, queued = srv.step()
        saw_queue |= queued > 0
        rounds += 1
        if rounds == 3:  # arrivals while others are on the device
            ids += [srv.submit(c, S.initial(c)) for c in cfgs[6:]]
        st = srv.stats()
        if st["completed"] + st["failed"] == len(cfgs):
            break
        assert rounds < 5000
    st = srv.stats()
    print("server stats", st, flush=True)
    assert st["failed"] == 0 and saw_queue and st["peak_running"] >= 2
    assert st["peak_device_bytes"] <= budget
    for rid, cfg in zip(ids, cfgs):
        info = srv.poll(rid)
        cv, sc = srv.result(rid)
        cv0, sc0 = alone(cfg)
        assert info["t_done"] >= info["t_admit"] >= info["t_submit"]
        assert sc == sc0, (rid, sc, sc0)
        I = (slice(None), slice(2, -2), slice(2, -2))
        assert np.array_equal(cv[I], cv0[I]), rid
    srv.close()


@pytest.mark.gpu
def test_gpu_server_backfill_admits_small_runs_past_a_blocked_head():
    """With backfilling, a queued run that fits the budget is admitted while
    an earlier one that does not fit waits (for at most `backfill` rounds of
    being passed over); every result stays bit-identical to the run alone."""
    small = [dict(c, stoptime=c["stoptime"] * 0.2) for c in requests(4, 9)]
    big = dict(S.kepler_disk_config(ni=128, nj=256), stoptime=0.05)
    sizes = [fosite_amd.device_bytes(c) for c in small]
    budget = fosite_amd.device_bytes(big) + 0.5 * min(sizes)  # big fits only alone
    assert fosite_amd.device_bytes(big) >= 2 * max(sizes)  # two small runs fit together
    srv = fosite_amd.Server(budget, round_steps=4, autotune=False, backfill=1000)
    first = srv.submit(small[0], S.initial(small[0]))
    srv.step()  # small[0] admitted and running
    ids = [srv.submit(big, S.initial(big))] + [srv.submit(c, S.initial(c)) for c in small[1:]]
    rounds = 0
    while srv.stats()["completed"] + srv.stats()["failed"] < 1 + len(ids):
        srv.step()
        rounds += 1
        assert rounds < 5000
    assert srv.stats()["failed"] == 0
    t_big = srv.poll(ids[0])["t_admit"]
    passed = [srv.poll(rid)["t_admit"] < t_big for rid in ids[1:]]
    print("backfill: admitted before the big run", passed, flush=True)
    assert any(passed)
    for rid, cfg in zip([first] + ids, [small[0], big] + small[1:]):
        cv, sc = srv.result(rid)
        cv0, sc0 = alone(cfg)
        assert sc == sc0, (rid, sc, sc0)
        I = (slice(None), slice(2, -2), slice(2, -2))
        assert np.array_equal(cv[I], cv0[I]), rid
    srv.close()
