"""Batches (fv_batch_*): contexts advanced together, every kernel of a step one
launch over all members, must give exactly the bits of each context run alone
with fv_run, under every tiling, across halts (rejected steps, stop times),
membership changes and both dtypes.  Runs alone are themselves checked
against the oracle in test_gpu_parity.py; one member is also checked against
the oracle directly here."""
import numpy as np
import pytest

import fosite_amd
import oracle
from fosite_amd import setups as S
from tests.parity import rel_l1

I = (slice(None), slice(2, -2), slice(2, -2))


def make(cfg):
    init = S.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    sim.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    return sim


def alone(cfg, chunks):
    sim = make(cfg)
    sim.first_step()
    taken = sum(sim.run(n) for n in chunks)
    cv, _ = sim.state()
    sc = sim.scalars()
    bf = sim.boundary_flux()
    sim.close()
    return taken, cv, sc, bf


def members(dtype="fp64"):
    D = S.kepler_disk_config
    out = [
        dict(D(ni=40, nj=64), seed=11),
        dict(D(ni=72, nj=144, cs_ratio=0.03), seed=12),
        dict(D(ni=56, nj=80), fargo=0, seed=13),
        dict(D(ni=33, nj=70, cs_ratio=0.08), limiter=S.MINMOD, seed=14),
        dict(D(ni=100, nj=90), bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC], seed=15),
        dict(D(ni=48, nj=64), cfl=0.9, seed=16),  # rejected steps: halts redone on the host path
    ]
    return [dict(c, dtype=dtype) for c in out]


def run_batch(cfgs, chunks, **opts):
    sims = [make(c) for c in cfgs]
    for s in sims:
        s.first_step()
    b = fosite_amd.Batch(sims, **opts)
    taken = [0] * len(sims)
    for n in chunks:
        taken = [t + k for t, k in zip(taken, b.run(n))]
    res = [(taken[k], sims[k].state()[0], sims[k].scalars(), sims[k].boundary_flux()) for k in range(len(sims))]
    info = b.info()
    b.close()
    for s in sims:
        s.close()
    return res, info


def check_same(cfgs, res, chunks):
    for k, cfg in enumerate(cfgs):
        t0, cv0, sc0, bf0 = alone(cfg, chunks)
        t1, cv1, sc1, bf1 = res[k]
        assert t1 == t0, (k, t1, t0)
        assert sc1 == sc0, (k, sc1, sc0)
        assert np.array_equal(cv1[I], cv0[I]), (k, np.abs(cv1[I] - cv0[I]).max())
        # boundary fluxes (Fluxes%bxflux/byflux) accumulated by the batch
        assert np.array_equal(bf1[0], bf0[0]) and np.array_equal(bf1[1], bf0[1]), k


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
@pytest.mark.parametrize("tiling", [0, 1, 2, 3])
def test_gpu_batch_equals_members_alone(dtype, tiling):
    cfgs = members(dtype)
    chunks = [25, 31]
    print("batch", dtype, "tiling", tiling, [(c["inum"], c["jnum"]) for c in cfgs], flush=True)
    res, info = run_batch(cfgs, chunks, autotune=False, tiling=tiling)
    print("batch info", info, flush=True)
    assert info["tiling"] == tiling and info["device_steps"] > 0 and info["halts"] > 0
    check_same(cfgs, res, chunks)


@pytest.mark.gpu
def test_gpu_batch_autotune_and_short_graphs():
    """The autotuner times all four tilings on real steps (no work is lost);
    graphs of at most 7 steps give the same bits."""
    cfgs = members("fp32")[:4]
    chunks = [40, 13]
    res, info = run_batch(cfgs, chunks, autotune=True, max_graph_steps=7)
    print("autotune info", info, flush=True)
    assert info["tuned"] == 1 and all(t > 0 for t in info["tune_ms"])
    assert info["tiling"] == int(np.argmin(info["tune_ms"]))
    check_same(cfgs, res, chunks)


@pytest.mark.gpu
def test_gpu_batch_reuses_the_tuning_of_its_class():
    """With reuse_tuning, the first batch of a class (element type, fields,
    stages, total cells and member count in powers of two) tunes by trials
    and a later batch of the same class takes its tiling without trials; the
    bits stay those of the runs alone."""
    cfgs = [dict(c, cfl=0.4) for c in members("fp64")[1:4]]
    chunks = [12, 9]
    res1, info1 = run_batch(cfgs, chunks, autotune=True, reuse_tuning=True)
    res2, info2 = run_batch([dict(c, seed=c["seed"] + 100) for c in cfgs], chunks, autotune=True,
                            reuse_tuning=True)
    print("reuse tuning", info1, info2, flush=True)
    assert info1["tuned"] in (1, 2)  # an earlier batch of the class may have tuned already
    assert info2["tuned"] == 2 and info2["tiling"] == info1["tiling"]
    assert all(t == 0 for t in info2["tune_ms"])
    check_same(cfgs, res1, chunks)


@pytest.mark.gpu
def test_gpu_batch_member_reaching_its_stop_time():
    cfgs = members("fp64")[:3]
    cfgs[1] = dict(cfgs[1], stoptime=0.3)  # ends inside the run
    chunks = [60]
    res, _ = run_batch(cfgs, chunks, autotune=False, tiling=1)
    assert res[1][0] < 60
    check_same(cfgs, res, chunks)


@pytest.mark.gpu
def test_gpu_batch_euler_dp5_members():
    """Euler members with the embedded-error controller and absorbing
    boundaries (device log/exp): same bits as alone."""
    cfgs = [S.euler_disk_config(ni=40, nj=72), dict(S.euler_disk_config(ni=64, nj=48), seed=5),
            dict(S.euler_disk_config(ni=24, nj=96), cfl=0.9, seed=6)]
    chunks = [20, 20]
    res, info = run_batch(cfgs, chunks, autotune=False, tiling=3)
    print("euler batch info", info, flush=True)
    check_same(cfgs, res, chunks)


@pytest.mark.gpu
def test_gpu_batch_membership_changes():
    cfgs = members("fp64")[:4]
    sims = [make(c) for c in cfgs]
    for s in sims:
        s.first_step()
    b = fosite_amd.Batch(sims[:3], autotune=False, tiling=0)
    b.run(20)                 # members 0, 1, 2
    b.remove(sims[1])
    sims[1].run(7)            # alone, outside the batch
    b.add(sims[3])
    b.run(15)                 # members 0, 2, 3
    b.add(sims[1])
    b.run(10)                 # members 0, 2, 3, 1
    b.close()
    plan = {0: [20, 15, 10], 1: [20, 7, 10], 2: [20, 15, 10], 3: [15, 10]}
    for k, cfg in enumerate(cfgs):
        t0, cv0, sc0, bf0 = alone(cfg, plan[k])
        cv1, _ = sims[k].state()
        bf1 = sims[k].boundary_flux()
        assert sims[k].scalars() == sc0, k
        assert np.array_equal(cv1[I], cv0[I]), k
        assert np.array_equal(bf1[0], bf0[0]) and np.array_equal(bf1[1], bf0[1]), k
        sims[k].close()


@pytest.mark.gpu
def test_gpu_batch_member_matches_oracle():
    cfgs = members("fp64")[:3]
    res, _ = run_batch(cfgs, [30], autotune=False, tiling=1)
    cfg = cfgs[0]
    init = S.initial(cfg)
    o = oracle.Oracle(cfg)
    o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    o.first_step()
    for _ in range(30):
        o.step()
    rc, _, rs = o.state()
    _, cv, sc, bf = res[0]
    assert sc["iter"] == rs["iter"] and sc["n_adj"] == rs["n_adj"]
    assert (rel_l1(cv, rc) <= 1e-12).all(), rel_l1(cv, rc)
    rbx, rby = o.bflux()
    for got, ref in ((bf[0], rbx), (bf[1], rby)):
        assert np.abs(got - ref).sum() <= 1e-12 * np.abs(ref).sum()


@pytest.mark.gpu
@pytest.mark.parametrize("rs", [1, 2, 3, 4])
def test_gpu_short_rounds_across_the_output_time(rs):
    """Rounds of rs device steps that end just before an output-clipped step:
    the next round must start with the clipped step (as the host path does at
    the start of every Step).  Batched members, and a single context run in
    chunks of rs steps, both against the oracle run to the same stop time."""
    cfgs = [dict(S.kepler_disk_config(ni=24, nj=48), stoptime=1.7, seed=21),
            dict(S.kepler_disk_config(ni=40, nj=64, cs_ratio=0.04), stoptime=2.3, seed=22)]
    sims = [make(c) for c in cfgs]
    for s in sims:
        s.first_step()
    b = fosite_amd.Batch(sims, autotune=False, tiling=1, max_graph_steps=rs)
    for _ in range(400):
        if sum(b.run(rs)) == 0:
            break
    b.close()
    for cfg, sim in zip(cfgs, sims):
        init = S.initial(cfg)
        o = oracle.Oracle(cfg)
        o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
        o.first_step()
        while not o.step():
            pass
        rc, _, rs_ = o.state()
        cv, _ = sim.state()
        sc = sim.scalars()
        assert sc["iter"] == rs_["iter"] and sc["n_adj"] == rs_["n_adj"] and sc["brk"] == rs_["brk"], (sc, rs_)
        assert abs(sc["time"] - rs_["time"]) <= 1e-12 * rs_["time"]
        assert (rel_l1(cv, rc) <= 1e-12).all(), rel_l1(cv, rc)
        # one context alone, device batches of rs steps
        one = make(cfg)
        one.first_step()
        while one.run(rs) == rs:
            pass
        cv1, _ = one.state()
        assert one.scalars() == sc and np.array_equal(cv1[I], cv[I])
        one.close()
        sim.close()
