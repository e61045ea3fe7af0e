"""Output times and a start time on the HIP path against the oracle (which
matches the compiled reference on the golden cases euler_count5 and
disk_starttime_count3, tests/test_oracle_golden.py).

With a datafile count > 1 (io/fileio_base.f90:420-470), fosite%Step clips
the time step before each output time (AdjustTimestep, fileio_base.f90:
604-620), writes the dataset (FARGO: the state moves to the real frame,
fileio_base.f90:523-533) and advances the output clock by
|stoptime - starttime| / count (IncTime, 591-600).  A run that starts at
t0 > 0 (timedisc_base.f90:295; a restart sets it, examples/restart.f90:
103-106) starts the clock at t0.  Device rounds halt at every output time
(H_OUTPUT) and the host takes that step.  Bit for bit without absorbing
boundaries; with them relative L1 <= 1e-12 (LOG/EXP, see test_gpu_parity)."""
import numpy as np
import pytest

import fosite_amd
import oracle
from fosite_amd import setups as S
from tests.parity import rel_l1

I = (slice(None), slice(2, -2), slice(2, -2))
E = S.riemann2d_config(n=40)
D = S.kepler_disk_config(ni=40, nj=64)
ND = dict(D, bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC])

CASES = {
    "euler_count7": (dict(E, stoptime=0.08, count=7), 0.0),
    "disk_count4_fargo2": (dict(ND, stoptime=3.0, count=4), 0.0),
    "disk_starttime_count3_fargo1": (dict(ND, starttime=2.0, stoptime=5.0, count=3, fargo=1), 0.0),
    "disk_starttime_absorbing": (dict(D, starttime=1.5, stoptime=4.0, count=5), 1e-12),
}


def check(sim, ref, tol, what):
    cv, pv = sim.state()
    rc, rp, rs = ref.state()
    sc = sim.scalars()
    for k in ("iter", "n_adj", "brk"):
        assert sc[k] == rs[k], (what, k, sc, rs)
    if tol == 0.0:
        for k in ("time", "dt", "dtold", "maxerrold"):
            assert sc[k] == rs[k], (what, k, sc, rs)
        assert np.array_equal(cv[I], rc[I]), (what, np.abs(cv[I] - rc[I]).max())
    else:
        for k in ("time", "dt"):
            assert abs(sc[k] - rs[k]) <= tol * abs(rs[k]), (what, k, sc, rs)
        assert (rel_l1(cv, rc) <= tol).all(), (what, rel_l1(cv, rc))
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [0, 9])
@pytest.mark.parametrize("name", list(CASES))
def test_outputs_and_start_time_match_oracle(name, chunk):
    cfg, tol = CASES[name]
    print("case", name, "chunk", chunk, flush=True)
    init = S.initial(cfg)
    sim, ref = fosite_amd.Fosite(cfg), oracle.Oracle(cfg)
    for o in (sim, ref):
        o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
        o.first_step()
    check(sim, ref, tol, "first step")
    outputs, n, done = 0, 0, False
    t_out = cfg["starttime"] + abs(cfg["stoptime"] - cfg["starttime"]) / cfg["count"]
    while not done and n < 400:
        if chunk:
            taken = sim.run(chunk)
            for _ in range(taken):
                assert not ref.step()
            n += taken
            done = taken < chunk
            if done:
                assert ref.step()
        else:
            done = sim.step()
            assert done == bool(ref.step())
            n += 0 if done else 1
        sc = check(sim, ref, tol, "step %d" % n)
        while sc["time"] >= t_out * (1 - 1e-12) and outputs < cfg["count"]:
            outputs += 1
            t_out += abs(cfg["stoptime"] - cfg["starttime"]) / cfg["count"]
    assert done and abs(sim.scalars()["time"] - cfg["stoptime"]) <= 1e-5 * cfg["stoptime"]
    assert outputs == cfg["count"]
    sim.close()
