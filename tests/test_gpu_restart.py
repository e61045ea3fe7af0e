"""A run restarted from a dataset on the HIP path (fosite_amd.restart, the
steps of examples/restart.f90) against the oracle restarted from the same
file, which matches the compiled reference's restart (tests/test_restart.py):
bit for bit without absorbing boundaries, relative L1 <= 1e-12 with them, on
host steps and device rounds, through the outputs that follow the restart."""
import os
import tempfile

import numpy as np
import pytest

import fosite_amd
import oracle
from fosite_amd import datafile as DF
from fosite_amd import setups as S
from tests.parity import rel_l1
from tests.test_restart import CASES, write_restart_file

I = (slice(None), slice(2, -2), slice(2, -2))


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [0, 5])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_restart_matches_oracle_restart(name, chunk):
    cfg, k = CASES[name]
    print("restart", name, "chunk", chunk, flush=True)
    with tempfile.TemporaryDirectory() as d:
        path, _, _ = write_restart_file(cfg, k, d)
        init = S.initial(cfg)
        w = init.get("w") if cfg.get("fargo", 0) == 2 else None
        sim = fosite_amd.restart(path, init.get("bccsound"), init.get("fcsound"), w)
        r = DF.load_restart(path)
    ref = oracle.Oracle(r["cfg"])
    ref.set_initial(r["pvar"], init.get("bccsound"), init.get("fcsound"), w)
    assert sim.first_step() == bool(ref.first_step())
    tol = 1e-12 if S.ABSORBING in r["cfg"]["bc"] else 0.0
    n, ended = 0, False
    while n < 24 and not ended:
        if chunk:
            taken = sim.run(chunk)
            for _ in range(taken):
                assert not ref.step()
            ended = taken < chunk
            if ended:
                assert ref.step()
            n += taken
        else:
            ended = sim.step()
            assert ended == bool(ref.step())
            n += 1
        cv, _ = sim.state()
        rc, _, rs = ref.state()
        sc = sim.scalars()
        assert sc["iter"] == rs["iter"] and sc["n_adj"] == rs["n_adj"], (n, sc, rs)
        if tol == 0.0:
            assert sc["time"] == rs["time"] and sc["dt"] == rs["dt"], (n, sc, rs)
            assert np.array_equal(cv[I], rc[I]), (n, rel_l1(cv, rc))
        else:
            assert abs(sc["time"] - rs["time"]) <= 1e-10 * rs["time"], (n, sc, rs)
            assert (rel_l1(cv, rc) <= tol).all(), (n, rel_l1(cv, rc))
    assert sim.scalars()["time"] > r["time"]
    sim.close()
