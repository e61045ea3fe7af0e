"""The Fortran 2008 host (fosite_gpu_mod over the ISO_C_BINDING layer):
on the CPU, that it is built, links the C ABI library and stops with the
library's message on a configuration the library rejects."""
import os

import pytest

from fosite_amd import setups as S
from tests import fortran_host as F


def test_fortran_host_is_built():
    assert os.access(F.FV_RUN, os.X_OK), "run __graft_entry__.build() first"


@pytest.mark.parametrize("bad, message", [
    (dict(physics=99), "physics must be"),
    (dict(bc=[S.PERIODIC, S.NO_GRADIENTS, S.NO_GRADIENTS, S.NO_GRADIENTS]), "opposite boundary"),
])
def test_fortran_host_reports_rejected_configuration(tmp_path, bad, message):
    cfg = dict(S.riemann2d_config(n=8), **bad)
    if cfg["physics"] == 99:  # initial data layout of the Euler case
        path = F.write_inputs(dict(cfg, physics=S.EULER), 1, str(tmp_path))
        txt = open(path).read().replace("k_physics = 17", "k_physics = 99")
        open(path, "w").write(txt)
        import subprocess
        r = subprocess.run([F.FV_RUN, path], cwd=str(tmp_path), capture_output=True, text=True)
    else:
        r, _ = F.run(cfg, 1, str(tmp_path))
    assert r.returncode != 0
    assert message in (r.stdout + r.stderr)
    assert "fosite_gpu::Setup" in r.stdout
