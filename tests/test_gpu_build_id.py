"""The library the GPU tests load was built from the sources that travel with
it: its build id (SHA-256 of the sources and compile flags) equals the one
the tree's sources give, so the GPU results belong to this code."""
import pytest

import fosite_amd
from fosite_amd import build


@pytest.mark.gpu
def test_gpu_library_matches_its_sources():
    print("build id", fosite_amd.build_id(), flush=True)
    assert fosite_amd.build_id() == build.source_id()
