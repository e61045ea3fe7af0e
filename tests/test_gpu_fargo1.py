"""FARGO type 1 (the background velocity w recomputed at every step from the
mean azimuthal velocity of each ring, timedisc_base.f90:2243-2326, called
from fosite.f90:305-311 and timedisc_base.f90:664-665) on the HIP path
against the oracle, which matches the compiled reference bit for bit
(tests/test_oracle_golden.py).

Bit for bit without absorbing boundaries; with them (LOG/EXP on the path)
relative L1 <= 1e-12 per conserved field and for w, as test_gpu_parity.py."""
import numpy as np
import pytest

import fosite_amd
import oracle
from fosite_amd import setups as S
from tests.parity import rel_l1

D = S.kepler_disk_config(ni=40, nj=64)
I = (slice(None), slice(2, -2), slice(2, -2))

CASES = {
    "nograd_radial": (dict(D, fargo=1, bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 25, 0.0),
    "nograd_cfl_0.9_rejects": (dict(D, fargo=1, cfl=0.9, bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC,
                                                             S.PERIODIC]), 20, 0.0),
    "absorbing": (dict(D, fargo=1), 25, 1e-12),
    "absorbing_ragged_70x90": (dict(S.kepler_disk_config(ni=70, nj=90), fargo=1), 15, 1e-12),
    # more than 256 rows and not a multiple of 256: the ring sums stage three
    # chunks (double-buffered), the last one partial
    "nograd_24x600": (dict(S.kepler_disk_config(ni=24, nj=600), fargo=1,
                           bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 12, 0.0),
    "absorbing_24x600": (dict(S.kepler_disk_config(ni=24, nj=600), fargo=1), 12, 1e-12),
}


def pair(cfg):
    init = S.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    ref = oracle.Oracle(cfg)
    sim.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), None)
    ref.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), None)
    sim.first_step()
    ref.first_step()
    return sim, ref


def check(sim, ref, tol, what):
    cv, _ = sim.state()
    rc, _, rs = ref.state()
    sc = sim.scalars()
    w, rw = sim.background_velocity(), ref.background_velocity()
    assert sc["n_adj"] == rs["n_adj"] and sc["iter"] == rs["iter"], (what, sc, rs)
    if tol == 0.0:
        assert sc["time"] == rs["time"] and sc["dt"] == rs["dt"], (what, sc, rs)
        assert np.array_equal(w, rw), (what, np.abs(w - rw).max())
        assert np.array_equal(cv[I], rc[I]), (what, np.abs(cv[I] - rc[I]).max())
    else:
        assert abs(sc["dt"] - rs["dt"]) <= tol * rs["dt"], (what, sc, rs)
        assert np.abs(w - rw).sum() <= tol * np.abs(rw).sum(), what
        assert (rel_l1(cv, rc) <= tol).all(), (what, rel_l1(cv, rc))
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_fargo1_matches_oracle(name):
    cfg, nsteps, tol = CASES[name]
    print("fargo1", name, cfg["inum"], cfg["jnum"], nsteps, flush=True)
    sim, ref = pair(cfg)
    check(sim, ref, tol, (name, 0))
    for n in range(1, nsteps + 1):
        assert sim.step() == ref.step()
        sc = check(sim, ref, tol, (name, n))
    if "rejects" in name:
        assert sc["n_adj"] > 0
    # w is not the Keplerian start value: the type-1 path really recomputes it
    wk = S.initial(dict(cfg, fargo=2))["w"]
    assert np.abs(sim.background_velocity() - wk).max() > 1e-6
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["nograd_radial", "nograd_cfl_0.9_rejects", "absorbing", "nograd_24x600",
                                  "absorbing_24x600"])
def test_gpu_fargo1_device_rounds_match_oracle(name):
    """Fosite.run() on a type-1 context: device-controlled steps with
    CalcBackgroundVelocity in the step graph (frame change, ordered ring sums
    into the other ring arrays, frame change back).  Rounds of several
    lengths; a rejected step halts the round and the host redoes it from the
    start-of-step w and velocities."""
    cfg, _, tol = CASES[name]
    print("fargo1 run()", name, flush=True)
    sim, ref = pair(cfg)
    done = 0
    for chunk in (1, 2, 7, 12, 8):
        done += sim.run(chunk)
        for _ in range(chunk):
            ref.step()
        sc = check(sim, ref, tol, (name, done))
    assert done == 30
    if "rejects" in name:
        assert sc["n_adj"] > 0
    sim.close()


@pytest.mark.gpu
def test_gpu_fargo1_batch_member_matches_run_alone():
    """In a batch a type-1 member takes host-controlled steps (the batch's
    shared launches have no CalcBackgroundVelocity); its bits equal those of
    the same context run alone in device rounds."""
    cfg = dict(CASES["nograd_radial"][0], seed=31)
    other = dict(S.kepler_disk_config(ni=48, nj=64), seed=32)
    alone, _ = pair(cfg)
    alone.run(15)
    ref_cv = alone.state()[0]
    alone.close()
    a, _ = pair(cfg)
    init = S.initial(other)  # a FARGO type-2 member beside it
    b = fosite_amd.Fosite(other)
    b.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"])
    b.first_step()
    batch = fosite_amd.Batch([a, b])
    assert batch.run(15) == [15, 15]
    assert np.array_equal(a.state()[0], ref_cv)
    batch.close()
    a.close()
    b.close()
