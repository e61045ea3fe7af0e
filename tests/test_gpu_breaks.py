"""Branches of the step that end or redirect a run, on the HIP path against
the oracle (bit for bit: Euler on a Cartesian box, no LOG/EXP on the path):

* CheckData breaks (timedisc_base.f90:1666-1742): NaN and negative density in
  the initial data (FirstStep), and a density falling below rhomin during the
  run -- host-controlled steps and device rounds, where the end-of-step
  flags halt the round (H_FLAGS) and the host redoes the step;
* the pressure floor pmin of ComputeRHS (timedisc_base.f90:1150-1164) when it
  clips;
* the dtlimit break after SolveODE (timedisc_base.f90:686-713), and a dtlimit
  close enough to the steps that device rounds halt on their conservative
  test (H_DTLIMIT) and the host redoes those steps without a break;
* the tableaux no other GPU test runs: SSPRK order 5 (the reference's
  default order, timedisc_ssprk.f90:77), RK-Fehlberg order 3 and Cash-Karp
  order 5 (timedisc_cashkarp.f90:101-113).

The rarefaction used below (two halves of a box moving apart) empties the
centre of the box step by step, so its minimum density and pressure fall
through the chosen limits at known steps of the oracle's run."""
import numpy as np
import pytest

import fosite_amd
import oracle
from fosite_amd import setups as S
from tests.parity import rel_l1

I = (slice(None), slice(2, -2), slice(2, -2))
E = S.riemann2d_config(n=32)
R = S.rarefaction_config(n=32)


def same(sim, ref, what, state=True):
    cv, _ = sim.state()
    rc, _, rs = ref.state()
    sc = sim.scalars()
    for k in ("iter", "n_adj", "brk", "time", "dt", "dtold", "maxerrold"):
        assert sc[k] == rs[k], (what, k, sc, rs)
    if state:
        assert np.array_equal(cv[I], rc[I], equal_nan=True), (what, rel_l1(cv, rc))
    return sc


def pair(cfg, pv):
    sim, ref = fosite_amd.Fosite(cfg), oracle.Oracle(cfg)
    sim.set_initial(pv)
    ref.set_initial(pv)
    return sim, ref


def run_both(cfg, pv, nsteps, chunk):
    """chunk 0: one Step() at a time; otherwise Run(chunk) (device rounds)
    against the oracle stepping alone.  Returns (steps taken, ended)."""
    sim, ref = pair(cfg, pv)
    assert sim.first_step() == bool(ref.first_step())
    same(sim, ref, "first step")
    n, ended = 0, False
    while n < nsteps and not ended:
        if chunk == 0:
            ended = sim.step()
            assert ended == bool(ref.step()), n
            n += 0 if ended else 1
        else:
            k = min(chunk, nsteps - n)
            taken = sim.run(k)
            for _ in range(taken):
                assert not ref.step()
            n += taken
            if taken < k:
                assert ref.step(), "the oracle did not end where the HIP run ended"
                ended = True
        same(sim, ref, "step %d" % n)
    sim.close()
    return n, ended


@pytest.mark.gpu
@pytest.mark.parametrize("what", ["nan", "negative density"])
def test_first_step_breaks_on_invalid_initial_data(what):
    pv = S.initial(R)["pvar"]
    pv[0, 10, 12] = np.nan if what == "nan" else -0.5
    print("case", what, flush=True)
    sim, ref = pair(R, pv)
    assert sim.first_step() and ref.first_step()
    assert sim.scalars()["brk"] == ref.state()[2]["brk"] == 1
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [0, 10])
def test_density_below_rhomin_breaks_the_run(chunk):
    cfg = dict(R, rhomin=0.05)  # the oracle breaks at step 17 (no rejections after step 9)
    print("case rhomin chunk", chunk, flush=True)
    n, ended = run_both(cfg, S.initial(cfg)["pvar"], 40, chunk)
    assert ended and n == 16


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [0, 7])
def test_pressure_floor_clips(chunk):
    cfg = dict(R, pmin=0.02)  # the minimum pressure reaches the floor after ~24 steps
    print("case pmin chunk", chunk, flush=True)
    n, ended = run_both(cfg, S.initial(cfg)["pvar"], 40, chunk)
    assert n == 40 and not ended


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [0, 3])
def test_dtlimit_break(chunk):
    cfg = dict(R, dtlimit=0.0019)  # the adjusted step falls under it at step 4
    print("case dtlimit break chunk", chunk, flush=True)
    n, ended = run_both(cfg, S.initial(cfg)["pvar"], 40, chunk)
    assert ended and n == 3


@pytest.mark.gpu
def test_dtlimit_halts_device_rounds_without_a_break():
    # 0.003 is below every adjusted step of the four-quadrant problem (the
    # smallest is 0.00323) but above 0.9*dt, the conservative bound a device
    # round tests: rounds halt and the host takes those steps exactly
    cfg = dict(E, dtlimit=0.003)
    n, ended = run_both(cfg, S.initial(cfg)["pvar"], 60, 20)
    assert n == 60 and not ended


TABLEAUX = {
    "ssprk5": dict(method=S.SSPRK, order=5),
    "rkf3": dict(method=S.RK_FEHLBERG, order=3),
    "cashkarp5": dict(method=S.CASH_KARP, order=5),
}


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [0, 12])
@pytest.mark.parametrize("name", list(TABLEAUX))
def test_tableaux_bitexact(name, chunk):
    print("case", name, "chunk", chunk, flush=True)
    cfg = dict(E, **TABLEAUX[name])
    n, _ = run_both(cfg, S.initial(cfg)["pvar"], 30, chunk)
    assert n == 30
    # a disk with rejections (cfl 0.9) and periodic radial boundaries: FARGO
    # frame changes and the reject path with every tableau
    D = dict(S.kepler_disk_config(ni=40, nj=64), cfl=0.9,
             bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC], **TABLEAUX[name])
    init = S.initial(D)
    sim, ref = fosite_amd.Fosite(D), oracle.Oracle(D)
    for o in (sim, ref):
        o.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"])
        o.first_step()
    if chunk:
        taken = sim.run(20)
        for _ in range(taken):
            ref.step()
    else:
        for _ in range(20):
            sim.step()
            ref.step()
    same(sim, ref, name + " disk")
    sim.close()
