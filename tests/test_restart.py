"""Restarts from a dataset (fosite_amd.datafile.load_restart,
examples/restart.f90:89-170): the configuration, start time and initial data
read back from a dataset this writer produced at an output time, and the
restarted run of the oracle against the restarted run of the compiled
reference (oracle/_ref/fv_driver with the same configuration, start time and
initial data -- restart.f90's steps: the file's /config with
/timedisc/starttime = the file's time, pvar on the interior cells,
Convert2Conservative, FirstStep, Step()) bit for bit."""
import os
import tempfile

import numpy as np
import pytest

import oracle
from fosite_amd import datafile as DF
from fosite_amd import setups as S
from oracle import refrun

NG = [S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]
CASES = {
    # an isothermal disk with FARGO 2 and a point mass, 4 outputs; restarted from the 2nd
    "disk_fargo2": (dict(S.kepler_disk_config(ni=24, nj=40), bc=NG, stoptime=0.8, count=4), 2),
    # the Euler four-quadrant problem, 3 outputs; restarted from the 1st
    "riemann": (dict(S.riemann2d_config(n=24), stoptime=0.06, count=3), 1),
    # FARGO type 1 with Dormand-Prince and absorbing boundaries; restarted from the 1st of 2
    "disk_fargo1_dp5": (dict(S.kepler_disk_config(ni=20, nj=32), fargo=1, method=S.DORMAND_PRINCE, order=5,
                             stoptime=0.5, count=2), 1),
    # an Euler Riemann problem in a rotating frame (Coriolis terms only); restarted from the 2nd of 3
    "riemann_rotframe": (dict(S.riemann2d_config(n=24), stoptime=0.06, count=3, omega=2.0, rotframe=2), 2),
    # an ALPHA-viscous disk in a rotating frame; restarted from the 2nd of 4
    "disk_alpha_rotframe": (dict(S.kepler_disk_config(ni=24, nj=40), bc=NG, stoptime=0.8, count=4, omega=0.3,
                                 rotframe=1, vismodel=2, dynconst=0.01), 2),
}


def run_to_output(cfg, k):
    """The oracle run until its k-th output (time = starttime + k*span/count,
    state moved to the real frame by the output's frame change); returns
    (oracle, time, pvar)."""
    init = S.initial(cfg)
    o = oracle.Oracle(cfg)
    o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"),
                  init.get("w") if cfg.get("fargo", 0) == 2 else None)
    o.first_step()
    t_out = k * cfg["stoptime"] / cfg["count"]
    for _ in range(10000):
        _, _, sc = o.state()
        if abs(sc["time"] - t_out) <= 1e-5 * t_out:
            break
        assert not o.step()
    _, pv, sc = o.state()
    return o, sc["time"], pv


def write_restart_file(cfg, k, d):
    o, time, pv = run_to_output(cfg, k)
    path = os.path.join(d, "run_%04d.bin" % k)
    pm = o.pointmass() if cfg.get("gravity", 0) else None
    DF.write(path, DF.dataset_entries(cfg, time, pv, pm, DF.filename_of(path)))
    return path, time, pv


@pytest.mark.parametrize("name", list(CASES))
def test_load_restart_reads_back_configuration_time_and_data(name):
    cfg, k = CASES[name]
    with tempfile.TemporaryDirectory() as d:
        path, time, pv = write_restart_file(cfg, k, d)
        r = DF.load_restart(path)
    full = S.full(cfg)
    assert r["time"] == time and r["cfg"]["starttime"] == time and r["step"] == k
    for key in ("physics", "geometry", "inum", "jnum", "xmin", "xmax", "ymin", "ymax", "gparam", "limiter",
                "recvars", "method", "order", "cfl", "stoptime", "tol_rel", "beta", "count", "fargo", "omega",
                "rotframe", "vismodel"):
        assert r["cfg"][key] == full[key], key
    assert list(r["cfg"]["bc"]) == list(full["bc"])
    if full["vismodel"]:
        assert all(r["cfg"][key] == full[key] for key in ("dynconst", "bulkconst", "cvis"))
    V = pv.shape[0]
    assert np.array_equal(np.asarray(r["cfg"]["tol_abs"][:V]), np.asarray(full["tol_abs"][:V]))
    I = (slice(None), slice(2, -2), slice(2, -2))
    assert np.array_equal(r["pvar"][I], pv[I])
    assert not r["pvar"][:, :2].any() and not r["pvar"][:, :, :2].any()  # ghosts zero, as LoadData leaves them


@pytest.mark.skipif(not refrun.available(), reason="reference build absent")
@pytest.mark.parametrize("name", list(CASES))
def test_oracle_restart_matches_reference_restart(name):
    cfg, k = CASES[name]
    with tempfile.TemporaryDirectory() as d:
        path, _, _ = write_restart_file(cfg, k, d)
        r = DF.load_restart(path)
    rcfg = r["cfg"]
    init = S.initial(cfg)  # the sound speeds and w the restart program sets itself
    init = dict(init, pvar=r["pvar"])
    w = init.get("w") if rcfg.get("fargo", 0) == 2 else None
    o = oracle.Oracle(rcfg)
    o.set_initial(r["pvar"], init.get("bccsound"), init.get("fcsound"), w)
    o.first_step()
    n = 12
    _, recs = refrun.run(rcfg, init, n)
    I = (slice(None), slice(2, -2), slice(2, -2))
    for rec in recs:
        if rec["step"] > 0 and o.step():
            break
        cv, _, sc = o.state()
        for key in ("iter", "n_adj", "brk", "time", "dt"):
            assert sc[key] == rec[key], (name, rec["step"], key, sc[key], rec[key])
        assert np.array_equal(cv[I], rec["cvar"][I]), (name, rec["step"])
    assert recs[0]["time"] == r["time"] and len(recs) > 3
