"""Datasets written from a GPU run (Fosite.write_dataset) against the ones
the reference wrote for the same runs (tests/golden/ref_dataset_*.bin).
Every key of the reference's file is written, in its order; mesh, time,
point mass and the configuration tree are bit-exact, and so are
the primitive variables -- Timedisc%pvar as the last ComputeRHS left it,
after Euler's PMIN clamp (timedisc_base.f90:1150-1164) -- except where the
disk's absorbing boundaries (LOG/EXP on the path) bound them to relative L1
1e-12."""
import numpy as np
import pytest

import fosite_amd
from fosite_amd import datafile as DF
from fosite_amd import setups as S
from tests.dataset_cases import CASES
from tests.test_datafile import compare, ref_filename, reference


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_dataset_equals_the_reference_dataset(name, tmp_path):
    cfg, nsteps = CASES[name]
    print("dataset", name, flush=True)
    init = S.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    sim.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"),
                    init.get("w") if cfg.get("fargo", 0) == 2 else None)
    sim.first_step()
    for _ in range(nsteps):
        if sim.step():
            break
    path = str(tmp_path / "gpu.bin")
    ref = reference(name)
    sim.write_dataset(path, ref_filename(ref))
    sim.close()
    tol = 1e-12 if S.ABSORBING in cfg["bc"] else 0.0
    compare(DF.read(path), ref, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [S.riemann2d_config(n=48),
                                 dict(S.kepler_disk_config(ni=32, nj=64),
                                      bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC])],
                         ids=["euler", "disk_fargo"])
def test_gpu_state_pvar_is_the_reference_pvar(cfg):
    """fv_get_state's pvar after host-controlled steps: Timedisc%pvar of the
    last ComputeRHS on every cell, ghost cells included, bit for bit with the
    oracle (Euler: after the PMIN clamp; FARGO: in the comoving frame)."""
    import oracle
    init = S.initial(cfg)
    w = init.get("w") if cfg.get("fargo", 0) == 2 else None
    sim = fosite_amd.Fosite(cfg)
    ref = oracle.Oracle(cfg)
    for o in (sim, ref):
        o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), w)
        o.first_step()
    for n in range(8):
        sim.step()
        ref.step()
        _, pv = sim.state()
        _, rpv, _ = ref.state()
        assert np.array_equal(pv, rpv), (n, np.abs(pv - rpv).max())
    sim.close()
