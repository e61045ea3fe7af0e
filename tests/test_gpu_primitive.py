"""PRIMITIVE reconstruction (fluxes_base.f90:378-381: the limited slopes act
on pvar, the face states are converted to conservative ones) on the HIP path
against the oracle, which matches the compiled reference bit for bit
(tests/test_oracle_golden.py, the euler_prim_* / disk_prim* fixtures).

Bit for bit without absorbing boundaries; with them (LOG/EXP on the path)
relative L1 <= 1e-12 per conserved field, as test_gpu_parity.py.  Host steps,
device-controlled batches and batches that mix PRIMITIVE and CONSERVATIVE
members (which then share the run-time-limiter kernels) are covered; the
FP32 form with PRIMITIVE reconstruction is a case of test_gpu_precision.py."""
import numpy as np
import pytest

import fosite_amd
import oracle
from fosite_amd import setups as S
from tests import golden_cases as G
from tests.parity import rel_l1

I = (slice(None), slice(2, -2), slice(2, -2))
E = S.riemann2d_config(n=64)
D = S.kepler_disk_config(ni=40, nj=64)
NG = [S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]

CASES = {
    "euler_vanleer": (dict(E, recvars=S.PRIMITIVE), 25, 0.0),
    "euler_minmod_dp5": (dict(E, recvars=S.PRIMITIVE, limiter=S.MINMOD, method=S.DORMAND_PRINCE, order=5,
                              tol_rel=1e-4), 15, 0.0),
    "euler_mc_ragged_50x36": (dict(E, recvars=S.PRIMITIVE, limiter=S.MONOCENT, theta=1.3, inum=50, jnum=36,
                                   ymax=36.0 / 50.0), 20, 0.0),
    "disk_nograd_fargo2": (dict(D, recvars=S.PRIMITIVE, bc=NG), 25, 0.0),
    "disk_nograd_fargo1_rejects": (dict(D, recvars=S.PRIMITIVE, bc=NG, fargo=1, cfl=0.9), 15, 0.0),
    "disk_absorbing": (dict(D, recvars=S.PRIMITIVE), 25, 1e-12),
    "euler_disk_absorbing": (dict(S.euler_disk_config(ni=32, nj=64), recvars=S.PRIMITIVE), 12, 1e-12),
}


def pair(cfg, dtype=None):
    init = S.initial(cfg)
    sim = fosite_amd.Fosite(dict(cfg, dtype=dtype) if dtype else cfg)
    ref = oracle.Oracle(cfg)
    w = init.get("w") if cfg.get("fargo", 0) == 2 else None
    sim.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), w)
    ref.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), w)
    sim.first_step()
    ref.first_step()
    return sim, ref


def check(sim, ref, tol, what):
    cv, _ = sim.state()
    rc, _, rs = ref.state()
    sc = sim.scalars()
    assert sc["n_adj"] == rs["n_adj"] and sc["iter"] == rs["iter"], (what, sc, rs)
    if tol == 0.0:
        assert sc["time"] == rs["time"] and sc["dt"] == rs["dt"], (what, sc, rs)
        assert np.array_equal(cv[I], rc[I]), (what, np.abs(cv[I] - rc[I]).max())
    else:
        assert abs(sc["dt"] - rs["dt"]) <= tol * rs["dt"], (what, sc, rs)
        assert (rel_l1(cv, rc) <= tol).all(), (what, rel_l1(cv, rc))
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_primitive_steps_match_oracle(name):
    cfg, nsteps, tol = CASES[name]
    print("primitive step()", name, cfg["inum"], cfg["jnum"], nsteps, flush=True)
    sim, ref = pair(cfg)
    check(sim, ref, tol, (name, 0))
    for n in range(1, nsteps + 1):
        assert sim.step() == ref.step()
        sc = check(sim, ref, tol, (name, n))
    if "rejects" in name:
        assert sc["n_adj"] > 0
    sim.close()


@pytest.mark.gpu
def test_gpu_primitive_differs_from_conservative():
    """The PRIMITIVE path is really taken: the same run with CONSERVATIVE
    reconstruction gives different bits."""
    cfg = CASES["euler_vanleer"][0]
    a, _ = pair(cfg)
    b, _ = pair(dict(cfg, recvars=S.CONSERVATIVE))
    a.run(10)
    b.run(10)
    assert np.abs(a.state()[0] - b.state()[0]).max() > 1e-6
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["euler_vanleer", "disk_nograd_fargo2", "disk_absorbing"])
def test_gpu_primitive_device_batches_match_oracle(name):
    cfg, _, tol = CASES[name]
    print("primitive run()", name, flush=True)
    sim, ref = pair(cfg)
    done = 0
    for chunk in (1, 6, 13, 10):
        done += sim.run(chunk)
        for _ in range(chunk):
            ref.step()
        check(sim, ref, tol, (name, done))
    assert done == 30
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in G.cases() if G.case(n)[0].get("recvars", 2) == S.PRIMITIVE])
def test_gpu_primitive_matches_reference_fixture(name):
    cfg, steps, _, states = G.case(name)
    print("primitive fixture", name, flush=True)
    sim, _ = pair(cfg)
    tol = 1e-12 if S.ABSORBING in cfg["bc"] else 0.0
    for n in range(1, steps + 1):
        sim.step()
        if n in states:
            cv, _ = sim.state()
            if tol == 0.0:
                assert np.array_equal(cv[I], states[n][I]), (name, n)
            else:
                assert (rel_l1(cv, states[n]) <= tol).all(), (name, n, rel_l1(cv, states[n]))
    sim.close()


@pytest.mark.gpu
def test_gpu_batch_mixing_primitive_and_conservative_is_bit_identical():
    """A batch whose members differ in the reconstruction variables runs the
    run-time-limiter kernels for all of them; each member keeps the bits of
    its run alone."""
    cfgs = [dict(D, recvars=S.PRIMITIVE, bc=NG, seed=21), dict(D, seed=22),
            dict(S.kepler_disk_config(ni=56, nj=80), recvars=S.PRIMITIVE, limiter=S.MINMOD, seed=23)]
    alone = []
    for c in cfgs:
        s, _ = pair(c)
        s.run(12)
        alone.append(s.state()[0])
        s.close()
    sims = [pair(c)[0] for c in cfgs]
    b = fosite_amd.Batch(sims)
    assert b.run(12) == [12, 12, 12]
    for s, a in zip(sims, alone):
        assert np.array_equal(s.state()[0], a)
    b.close()
    for s in sims:
        s.close()
