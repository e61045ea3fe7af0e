"""Datasets written from a GPU run (Fosite.write_dataset) against the ones
the reference wrote for the same runs (tests/golden/ref_dataset_*.bin).
Mesh, time, point mass and configuration entries are bit-exact.  The
primitive variables: the disk's absorbing boundaries (LOG/EXP on the path)
bound them to relative L1 1e-12.  For Euler physics they are
Convert2Primitive of the output state, while the reference's Timedisc%pvar
is the clamped pvar before the PMIN round trip Prim2Cons
(timedisc_base.f90:1150-1164) -- the two differ by an ulp here and there:
relative L1 1e-15."""
import numpy as np
import pytest

import fosite_amd
from fosite_amd import datafile as DF
from fosite_amd import setups as S
from tests.dataset_cases import CASES
from tests.test_datafile import compare, reference


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
    sim.write_dataset(path)
    sim.close()
    tol = 1e-12 if S.ABSORBING in cfg["bc"] else (1e-15 if cfg["physics"] == S.EULER else 0.0)
    compare(DF.read(path), reference(name), tol)
