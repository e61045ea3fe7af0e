"""The rotating frame (sources/sources_rotframe.f90:73-247: centrifugal and
Coriolis accelerations of a frame turning with Mesh%omega about the origin,
after point-mass gravity in the source list) on the HIP path against the
oracle, which matches the compiled reference bit for bit
(tests/test_oracle_golden.py, the *rotframe* fixtures), and against those
fixtures directly.

Bit for bit without absorbing boundaries; with them (LOG/EXP on the path)
relative L1 <= 1e-12 per conserved field.  Covered: logcylindrical disks with
gravity, FARGO types 1 and 2 and none, Coriolis only (disable_centaccel),
an Euler Cartesian mesh (both centrifugal components non-zero), an Euler disk
with PRIMITIVE reconstruction, host steps, device rounds and a batch whose
members differ in the frame."""
import numpy as np
import pytest

import fosite_amd
from fosite_amd import setups as S
from tests import golden_cases as G
from tests.parity import rel_l1
from tests.test_gpu_primitive import I, check, pair

E = S.riemann2d_config(n=64)
D = S.kepler_disk_config(ni=40, nj=64)
NG = [S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]

CASES = {
    "disk_nograd_fargo2": (dict(D, bc=NG, omega=0.3, rotframe=1), 25, 0.0),
    "disk_nograd_fargo1_nocent": (dict(D, bc=NG, fargo=1, omega=0.5, rotframe=2), 20, 0.0),
    "disk_nograd_no_fargo": (dict(D, bc=NG, fargo=0, omega=0.3, rotframe=1), 20, 0.0),
    "disk_absorbing_fargo2": (dict(D, omega=0.3, rotframe=1), 20, 1e-12),
    "euler_cartesian": (dict(E, omega=2.0, rotframe=1), 25, 0.0),
    "euler_cartesian_nocent_dp5": (dict(E, omega=2.0, rotframe=2, method=S.DORMAND_PRINCE, order=5,
                                        tol_rel=1e-4), 15, 0.0),
    "euler_disk_prim_absorbing": (dict(S.euler_disk_config(ni=32, nj=64), omega=0.2, rotframe=1,
                                       recvars=S.PRIMITIVE), 12, 1e-12),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_rotframe_steps_match_oracle(name):
    cfg, nsteps, tol = CASES[name]
    print("rotframe step()", name, cfg["inum"], cfg["jnum"], nsteps, flush=True)
    sim, ref = pair(cfg)
    check(sim, ref, tol, (name, 0))
    for n in range(1, nsteps + 1):
        assert sim.step() == ref.step()
        check(sim, ref, tol, (name, n))
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["disk_nograd_fargo2", "disk_nograd_fargo1_nocent", "euler_cartesian",
                                  "disk_absorbing_fargo2"])
def test_gpu_rotframe_device_rounds_match_oracle(name):
    cfg, _, tol = CASES[name]
    print("rotframe run()", name, flush=True)
    sim, ref = pair(cfg)
    done = 0
    for chunk in (1, 6, 13):
        done += sim.run(chunk)
        for _ in range(chunk):
            ref.step()
        check(sim, ref, tol, (name, done))
    assert done == 20
    sim.close()


@pytest.mark.gpu
def test_gpu_rotframe_source_is_applied():
    """The frame's terms are really evaluated: the same run in the inertial
    frame, and with Coriolis terms only, gives different bits."""
    cfg = CASES["euler_cartesian"][0]
    runs = []
    for c in (cfg, dict(cfg, rotframe=2), dict(cfg, rotframe=0)):
        s, _ = pair(c)
        s.run(8)
        runs.append(s.state()[0])
        s.close()
    assert np.abs(runs[0] - runs[1]).max() > 1e-6
    assert np.abs(runs[0] - runs[2]).max() > 1e-6
    assert np.abs(runs[1] - runs[2]).max() > 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in G.cases() if G.case(n)[0].get("rotframe", 0)])
def test_gpu_rotframe_matches_reference_fixture(name):
    cfg, steps, _, states = G.case(name)
    print("rotframe fixture", name, flush=True)
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
def test_gpu_batch_with_rotating_and_inertial_members_is_bit_identical():
    """Members of one batch with different frames (omega, disable_centaccel,
    none) each keep the bits of their run alone."""
    cfgs = [dict(D, bc=NG, omega=0.3, rotframe=1, seed=31), dict(D, bc=NG, seed=32),
            dict(S.kepler_disk_config(ni=56, nj=80), bc=NG, omega=0.7, rotframe=2, seed=33)]
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
