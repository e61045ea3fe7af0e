"""The viscosity source (sources/sources_viscosity.f90: constant MOLECULAR and
Shakura-Sunyaev ALPHA viscosity, stresses and their tensor divergence, the
Euler energy term, the update-only-when-the-time-changes rule and the viscous
time-step limit) on the HIP path against the oracle, which matches the
compiled reference bit for bit (tests/test_oracle_golden.py, the *viscous* and
*alpha* fixtures), and against those fixtures directly.

Bit for bit without absorbing boundaries; with them (LOG/EXP on the path)
relative L1 <= 1e-12 per conserved field.  Covered: periodic shear boxes
(Euler over two tiles in x, Euler with FARGO type 2 on a ragged mesh with a
bulk viscosity, isothermal with FARGO type 1 limited by the viscous time
step), Riemann problems (one rejecting steps until it breaks, periodic with
Dormand-Prince whose last two stages share a time, one limited by the
viscous step: dtcause 3), an Euler disk with FARGO type 1 and PRIMITIVE
reconstruction, and isothermal ALPHA disks with the rotating frame and FARGO
type 1, without FARGO and limited by the viscous step, with SSPRK(5) on two
tiles, and with absorbing boundaries; ALPHA on Euler disks (Dormand-Prince;
FARGO type 1 in a rotating frame), whose stages take the sound speed of the
step's initial state."""
import ctypes

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
    "shear_euler_molecular_2x_tiles": (S.shear_box_config(n=100, jnum=48, fargo=0, vismodel=1, dynconst=1e-3),
                                       20, 0.0),
    "shear_euler_molecular_fargo2_ragged": (S.shear_box_config(n=50, jnum=36, fargo=2, vismodel=1, dynconst=2e-3,
                                                               bulkconst=1e-3), 20, 0.0),
    "shear_iso_molecular_fargo1_limited": (S.shear_box_config(n=72, jnum=40, fargo=1, physics=S.EULER_ISOTHERM,
                                                              vismodel=1, dynconst=5e-3), 20, 0.0),
    # a Riemann problem whose error control rejects steps over and over
    # until the step falls below dtlimit and the run breaks
    "euler_molecular_rejections_break": (dict(E, inum=50, jnum=36, ymax=36.0 / 50.0, vismodel=1, dynconst=2e-3,
                                              limiter=S.MONOCENT, theta=1.3), 8, 0.0),
    "euler_molecular_bulk_periodic_dp5": (dict(E, vismodel=1, dynconst=2e-3, bulkconst=5e-4, bc=[S.PERIODIC] * 4,
                                               method=S.DORMAND_PRINCE, order=5, tol_rel=1e-4), 12, 0.0),
    "euler_molecular_viscous_cfl": (dict(E, vismodel=1, dynconst=2e-2, cvis=0.1), 15, 0.0),
    "euler_disk_molecular_fargo1_prim": (dict(S.euler_disk_config(ni=32, nj=64), fargo=1, recvars=S.PRIMITIVE,
                                              vismodel=1, dynconst=1e-4, bc=NG), 12, 0.0),
    "disk_alpha_fargo1_rotframe": (dict(D, fargo=1, vismodel=2, dynconst=0.01, omega=0.3, rotframe=1, bc=NG),
                                   20, 0.0),
    "disk_alpha_no_fargo_limited": (dict(D, fargo=0, vismodel=2, dynconst=0.3, cvis=0.2, bc=NG), 20, 0.0),
    "disk_alpha_ssprk5_wide": (dict(S.kepler_disk_config(ni=72, nj=64), fargo=0, vismodel=2, dynconst=0.02,
                                    method=S.SSPRK, order=5, bc=NG), 15, 0.0),
    "disk_alpha_dp5_absorbing": (dict(D, fargo=0, vismodel=2, dynconst=0.05, method=S.DORMAND_PRINCE, order=5,
                                      tol_rel=1e-4), 12, 1e-12),
    # ALPHA with Euler physics: the sound speed of Timedisc%pvar in the stages
    "euler_disk_alpha_dp5": (dict(S.euler_disk_config(ni=32, nj=64), vismodel=2, dynconst=0.02, bc=NG), 12, 0.0),
    "euler_disk_alpha_fargo1_rotframe": (dict(S.euler_disk_config(ni=32, nj=64), vismodel=2, dynconst=0.01,
                                              fargo=1, omega=0.3, rotframe=1, method=S.SSPRK, order=3, bc=NG),
                                         12, 0.0),
    # BETA: beta*rho*r^2*|Omega| (an isothermal disk; an Euler disk with FARGO type 1)
    "disk_beta_ssprk3": (dict(D, fargo=0, vismodel=3, dynconst=1e-3, bc=NG), 15, 0.0),
    "euler_disk_beta_fargo1": (dict(S.euler_disk_config(ni=32, nj=64), vismodel=3, dynconst=1e-3, fargo=1,
                                    method=S.SSPRK, order=3, bc=NG), 12, 0.0),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_viscosity_steps_match_oracle(name):
    cfg, nsteps, tol = CASES[name]
    print("viscosity step()", name, cfg["inum"], cfg["jnum"], nsteps, flush=True)
    sim, ref = pair(cfg)
    check(sim, ref, tol, (name, 0))
    causes = set()
    for n in range(1, nsteps + 1):
        assert sim.step() == ref.step()
        sc = check(sim, ref, tol, (name, n))
        rs = ref.state()[2]
        assert sc["dtcause"] == rs["dtcause"], (name, n, sc, rs)
        causes.add(sc["dtcause"])
    if name.endswith("limited") or name.endswith("viscous_cfl"):
        assert 3 in causes, causes  # the viscous limit set the step
    if name.endswith("break"):
        assert sim.scalars()["brk"] == 1
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["shear_euler_molecular_2x_tiles", "disk_alpha_fargo1_rotframe",
                                  "euler_molecular_bulk_periodic_dp5", "disk_alpha_no_fargo_limited",
                                  "shear_iso_molecular_fargo1_limited", "euler_disk_molecular_fargo1_prim",
                                  "disk_alpha_dp5_absorbing", "euler_disk_alpha_dp5",
                                  "euler_disk_alpha_fargo1_rotframe", "disk_beta_ssprk3", "euler_disk_beta_fargo1"])
def test_gpu_viscosity_device_rounds_match_oracle(name):
    """fv_run with viscosity: device-controlled steps (k_visc before every
    k_rhs, UpdateViscosity's time test and the viscous limit of the next step
    decided on the device, the host replaying the scalars), identical to the
    oracle's steps; the profiled chunk shows that device rounds ran."""
    cfg, _, tol = CASES[name]
    print("viscosity run()", name, flush=True)
    sim, ref = pair(cfg)
    done = 0
    for k, chunk in enumerate((1, 6, 7, 5)):
        if k == 3:
            sim.profile(True)
        done += sim.run(chunk)
        for _ in range(chunk):
            ref.step()
        sc = check(sim, ref, tol, (name, done))
        assert sc["dtcause"] == ref.state()[2]["dtcause"], (name, done)
    assert done == 19
    assert sim.kernel_stats("control")[0] > 0 and sim.kernel_stats("visc")[0] > 0
    sim.profile(False)
    # CalcTimestep right after a device round: the step the round's control
    # formed (viscous limit included), as the oracle's
    dt, cause = ctypes.c_double(), ctypes.c_int()
    assert sim.L.fv_calc_timestep(sim._h, ctypes.byref(dt), ctypes.byref(cause)) == 0
    rs = ref.state()[2]
    if tol == 0.0:
        assert dt.value == rs["dt"] and cause.value == rs["dtcause"], (name, dt.value, cause.value, rs)
    sim.close()


@pytest.mark.gpu
def test_gpu_viscosity_alpha_fargo2_disk_breaks_like_the_oracle():
    """At time 0 CalcTimestep's UpdateViscosity sees the comoving velocities
    of Timedisc%pvar: with FARGO type 2 on a Keplerian disk |e_phi/r . v| is
    near zero there and the viscous step collapses; the device run stops at
    the same step, with the same time step, as the oracle's."""
    cfg = dict(D, fargo=2, vismodel=2, dynconst=0.01, bc=NG)
    print("viscosity alpha fargo2", flush=True)
    sim, ref = pair(cfg)
    sc, rs = sim.scalars(), ref.state()[2]
    assert sc["dt"] == rs["dt"] and sc["dtcause"] == rs["dtcause"] == 3, (sc, rs)
    for n in range(1, 6):
        d1, d2 = sim.step(), ref.step()
        assert d1 == d2, n
        sc, rs = sim.scalars(), ref.state()[2]
        assert sc["brk"] == rs["brk"] and sc["iter"] == rs["iter"] and sc["dt"] == rs["dt"], (n, sc, rs)
        if d1:
            break
    sim.close()


@pytest.mark.gpu
def test_gpu_viscosity_source_is_applied():
    """The viscous terms are really evaluated: the same run without
    viscosity, and with another bulk viscosity, gives different bits."""
    cfg = CASES["shear_euler_molecular_2x_tiles"][0]
    runs = []
    for c in (cfg, dict(cfg, bulkconst=0.0), dict(cfg, vismodel=0)):
        s, _ = pair(c)
        for _ in range(5):
            s.step()
        runs.append(s.state()[0])
        s.close()
    assert np.abs(runs[0] - runs[1]).max() > 1e-9
    assert np.abs(runs[0] - runs[2]).max() > 1e-9


@pytest.mark.gpu
def test_gpu_viscosity_is_refused_in_a_batch():
    sim, _ = pair(CASES["disk_alpha_fargo1_rotframe"][0])
    L = fosite_amd.load_library()
    b = ctypes.c_void_p()
    assert L.fv_batch_create(ctypes.byref(b), ctypes.byref(fosite_amd.FvBatchOpts(0, 0, 0, 0))) == 0
    assert L.fv_batch_add(b, sim._h) == -1
    assert b"viscosity" in L.fv_batch_last_error(b)
    assert L.fv_batch_destroy(b) == 0
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in G.cases() if G.case(n)[0].get("vismodel", 0)])
def test_gpu_viscosity_matches_reference_fixture(name):
    cfg, steps, scalars, states = G.case(name)
    print("viscosity fixture", name, flush=True)
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
