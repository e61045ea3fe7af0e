"""Point-mass accretion (sources/gravity_pointmass.f90:383-428: at every
ComputeRHS the mass that crossed the outflow boundary since the last
evaluation, GetBoundaryFlux's density sum, is added to the point mass and the
acceleration is rescaled by new/old mass) on the HIP path against the oracle,
which matches the compiled reference bit for bit (tests/test_oracle_golden.py).

Bit for bit with no-gradients radial boundaries; with absorbing ones (LOG/EXP
on the path) relative L1 <= 1e-12 per conserved field and relative 1e-12 for
the point mass, as test_gpu_parity.py.  Host-controlled steps (step()) and
device-controlled batches (run(), where the step's boundary fluxes are
accumulated before the end-of-step evaluation) are both covered."""
import numpy as np
import pytest

import fosite_amd
import oracle
from fosite_amd import setups as S
from tests import golden_cases as G
from tests.parity import rel_l1

D = S.kepler_disk_config(ni=40, nj=64)
NG = [S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]
I = (slice(None), slice(2, -2), slice(2, -2))

CASES = {
    "nograd_west": (dict(D, outbound=1, bc=NG), 25, 0.0),
    "nograd_west_no_fargo_acclimit": (dict(D, outbound=1, bc=NG, fargo=0, acclimit=0.001), 25, 0.0),
    "nograd_east_cfl_0.9_rejects": (dict(D, outbound=2, bc=NG, cfl=0.9), 20, 0.0),
    "nograd_fargo1": (dict(D, outbound=1, bc=NG, fargo=1), 15, 0.0),
    "absorbing_west": (dict(D, outbound=1), 25, 1e-12),
    "euler_absorbing_west": (dict(S.euler_disk_config(ni=32, nj=64), outbound=1), 12, 1e-12),
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
    pm, rpm = sim.point_mass(), ref.pointmass()
    assert sc["n_adj"] == rs["n_adj"] and sc["iter"] == rs["iter"], (what, sc, rs)
    if tol == 0.0:
        assert sc["time"] == rs["time"] and sc["dt"] == rs["dt"], (what, sc, rs)
        assert np.array_equal(pm, rpm), (what, pm, rpm)
        assert np.array_equal(cv[I], rc[I]), (what, np.abs(cv[I] - rc[I]).max())
    else:
        assert abs(sc["dt"] - rs["dt"]) <= tol * rs["dt"], (what, sc, rs)
        assert abs(pm[0] - rpm[0]) <= tol * abs(rpm[0]), (what, pm, rpm)
        assert abs(pm[3] - rpm[3]) <= tol * abs(rpm[0]), (what, pm, rpm)
        assert (rel_l1(cv, rc) <= tol).all(), (what, rel_l1(cv, rc))
    return sc, pm


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_accretion_steps_match_oracle(name):
    cfg, nsteps, tol = CASES[name]
    print("accretion step()", name, cfg["inum"], cfg["jnum"], nsteps, flush=True)
    sim, ref = pair(cfg)
    check(sim, ref, tol, (name, 0))
    for n in range(1, nsteps + 1):
        assert sim.step() == ref.step()
        sc, pm = check(sim, ref, tol, (name, n))
    assert abs(pm[3]) > 1e-6, pm  # mass really crossed the boundary
    if "rejects" in name:
        assert sc["n_adj"] > 0
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["nograd_west", "nograd_west_no_fargo_acclimit", "nograd_east_cfl_0.9_rejects",
                                  "absorbing_west"])
def test_gpu_accretion_device_batches_match_oracle(name):
    """run(): device-controlled steps; a rejected step halts the batch and is
    redone by the host with the point mass of the step's start."""
    cfg, nsteps, tol = CASES[name]
    print("accretion run()", name, flush=True)
    sim, ref = pair(cfg)
    done = 0
    for chunk in (1, 7, 12, 10):
        done += sim.run(chunk)
        for _ in range(chunk):
            ref.step()
        check(sim, ref, tol, (name, done))
    assert done == 30
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in G.cases() if G.case(n)[0].get("outbound", 0)])
def test_gpu_accretion_matches_reference_fixture(name):
    """The reference's own point mass after every step of the accretion
    fixtures (absorbing boundaries: relative 1e-12)."""
    cfg, steps, _, states = G.case(name)
    pm_ref, _ = G.pointmass(name)
    print("accretion fixture", name, flush=True)
    sim, _ = pair(cfg)
    for n in range(1, steps + 1):
        sim.step()
        pm = sim.point_mass()
        assert abs(pm[0] - pm_ref[n][0]) <= 1e-12 * abs(pm_ref[n][0]), (name, n, pm, pm_ref[n])
        assert abs(pm[3] - pm_ref[n][3]) <= 1e-12 * abs(pm_ref[n][0]), (name, n, pm, pm_ref[n])
        if n in states:
            cv, _ = sim.state()
            assert (rel_l1(cv, states[n]) <= 1e-12).all(), (name, n, rel_l1(cv, states[n]))
    sim.close()


@pytest.mark.gpu
def test_gpu_accretion_fp32_form_tracks_fp64():
    """Precision form: the accreted mass (the telescoped boundary-flux sum)
    against the FP64 oracle, to the bar the FP32 boundary fluxes meet
    (tests/test_gpu_precision.py): relative 1e-4 of the mass change."""
    cfg, _, _ = CASES["nograd_west"]
    sim, ref = pair(cfg, dtype="fp32")
    sim.run(30)
    for _ in range(30):
        ref.step()
    pm, rpm = sim.point_mass(), ref.pointmass()
    print("fp32 accretion", pm, rpm, flush=True)
    assert abs((pm[0] - 1.0) - (rpm[0] - 1.0)) <= 1e-4 * abs(rpm[0] - 1.0), (pm, rpm)
    sim.close()


@pytest.mark.gpu
def test_gpu_point_mass_without_accretion():
    cfg = dict(D, bc=NG)
    sim, _ = pair(cfg)
    sim.run(5)
    assert np.array_equal(sim.point_mass(), [1.0, 0.0, 0.0, 0.0])
    sim.close()
    nograv = fosite_amd.Fosite(S.riemann2d_config(n=16))
    with pytest.raises(RuntimeError, match="point mass"):
        nograv.point_mass()
    nograv.close()


BATCH = [dict(D, outbound=1, bc=NG, seed=41), dict(D, bc=NG, seed=42),
         dict(S.kepler_disk_config(ni=56, nj=80), outbound=2, bc=NG, fargo=0, acclimit=0.001, seed=43),
         dict(S.kepler_disk_config(ni=48, nj=96), outbound=1, seed=44)]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
@pytest.mark.parametrize("tiling", [1, 2])
def test_gpu_batch_with_accreting_members_is_bit_identical(dtype, tiling):
    """Members with point-mass accretion (west and east boundaries, an
    accretion limit, absorbing boundaries) share the batch rounds with one
    without: every member keeps the state, boundary fluxes and point mass of
    its run alone, step for step over rounds of several sizes."""
    cfgs = [dict(c, dtype=dtype) for c in BATCH]
    print("accretion batch", dtype, "tiling", tiling, flush=True)
    alone = []
    for c in cfgs:
        s, _ = pair(c)
        s.run(14)
        alone.append((s.state()[0], s.point_mass(), s.boundary_flux(), s.scalars()))
        s.close()
    sims = [pair(c)[0] for c in cfgs]
    b = fosite_amd.Batch(sims, autotune=False, tiling=tiling)
    for k in (1, 5, 8):
        assert b.run(k) == [k] * len(sims)
    for s, (cv, pm, bf, sc) in zip(sims, alone):
        assert s.scalars() == sc
        assert np.array_equal(s.state()[0], cv)
        assert np.array_equal(s.point_mass(), pm)
        for x, y in zip(s.boundary_flux(), bf):
            assert np.array_equal(x, y)
    assert b.info()["device_steps"] > 0
    b.close()
    for s in sims:
        s.close()


@pytest.mark.gpu
def test_gpu_served_accretion_runs_equal_runs_alone():
    """Disks set up with the reference's default accretion through the
    western boundary (gravity_pointmass.f90:291-299) in the serving loop,
    beside runs without: each result bit-identical to the run alone."""
    reqs = [cfg for _, cfg in S.serving_requests(n=4, seed=11, nmin=24, nmax=64, orbits=0.3)]
    cfgs = [dict(c, outbound=1 if k % 2 == 0 else 0, bc=NG) for k, c in enumerate(reqs)]
    srv = fosite_amd.Server(1e10, round_steps=4, autotune=False)
    ids = [srv.submit(c, S.initial(c)) for c in cfgs]
    rounds = 0
    while True:
        srv.step()
        st = srv.stats()
        if st["completed"] + st["failed"] == len(cfgs):
            break
        rounds += 1
        assert rounds < 5000
    assert srv.stats()["failed"] == 0
    for rid, c in zip(ids, cfgs):
        cv, sc = srv.result(rid)
        init = S.initial(c)
        sim = fosite_amd.Fosite(c)
        sim.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
        sim.first_step()
        while sim.run(256) == 256:
            pass
        assert sc == sim.scalars(), rid
        assert np.array_equal(cv[I], sim.state()[0][I]), rid
        sim.close()
    srv.close()
