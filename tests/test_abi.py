"""The C ABI library on the CPU: it loads, exports every entry point that
include/fosite_fv.h declares, computes the host-side mesh metric exactly like
the reference, and rejects unsupported configurations before touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import fosite_amd
from fosite_amd import setups as S
from tests import golden_cases as G
from tests.test_oracle_golden import MESH_ARRAYS

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "fosite_fv.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fv_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_python_export_list():
    assert declared() == sorted(fosite_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = fosite_amd.load_library()
    for name in declared():
        assert hasattr(L, name), name


def test_library_was_built_from_this_tree():
    """The library in the tree carries the SHA-256 of the sources and flags
    it was compiled from; it must be this tree's (a stale or foreign prebuilt
    library fails here)."""
    from fosite_amd import build
    assert fosite_amd.build_id() == build.source_id()


@pytest.mark.parametrize("name", G.meshes())
def test_host_mesh_matches_reference_fixture(name):
    cfg, arrays = G.mesh(name)
    for ours, (key, idx) in MESH_ARRAYS.items():
        ref = arrays[key] if idx is None else arrays[key][idx]
        assert np.array_equal(fosite_amd.mesh_array(cfg, ours), ref), (name, ours)


BAD = {
    "physics": dict(physics=99),
    "geometry": dict(geometry=3),
    "reconstruction variables": dict(recvars=3),
    "outflow boundary": dict(outbound=5),
    "accretion without a point mass": dict(outbound=1),
    "limiter": dict(limiter=6),
    "unpaired periodic": dict(bc=[S.PERIODIC, S.NO_GRADIENTS, S.NO_GRADIENTS, S.NO_GRADIENTS]),
    "absorbing in y": dict(bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.ABSORBING, S.NO_GRADIENTS]),
    "euler fargo in fp32": dict(fargo=2, dtype=1),
    "tableau": dict(method=S.SSPRK, order=4),
    "tiny mesh": dict(inum=3),
    "dtype": dict(dtype=1),
    "vismodel": dict(vismodel=4),
    "viscosity in fp32": dict(vismodel=1, dtype="fp32"),
}


@pytest.mark.parametrize("what", list(BAD))
def test_create_rejects_unsupported_configuration(what):
    L = fosite_amd.load_library()
    cfg = dict(S.riemann2d_config(n=16), **BAD[what])
    c = fosite_amd.FvConfig.from_dict(cfg)
    if what == "dtype":
        c.dtype = 2
    h = ctypes.c_void_p()
    assert L.fv_create(ctypes.byref(h), ctypes.byref(c)) == -1
    assert not h.value


def test_batch_create_rejects_bad_options():
    """Batch options are checked before any device resource is created."""
    L = fosite_amd.load_library()
    h = ctypes.c_void_p()
    for o in ((0, 4, 0), (0, -1, 0), (1, 0, 513), (0, 0, -2)):
        assert L.fv_batch_create(ctypes.byref(h), ctypes.byref(fosite_amd.FvBatchOpts(*o))) == -1, o
        assert not h.value
    assert L.fv_batch_run(None, 1, None) == -1
    assert L.fv_batch_add(None, None) == -1
    assert L.fv_batch_remove(None, None) == -1
    assert L.fv_batch_destroy(None) == 0


def test_server_create_rejects_bad_options():
    L = fosite_amd.load_library()
    h = ctypes.c_void_p()
    for o in ((0.0, 16, 0, 0), (1e9, 0, 0, 0), (1e9, 513, 1, 0), (1e9, 16, 0, -1)):
        assert L.fv_server_create(ctypes.byref(h), ctypes.byref(fosite_amd.FvServerOpts(*o))) == -1, o
        assert not h.value
    assert L.fv_server_step(None, None, None) == -1
    assert L.fv_server_destroy(None) == 0


def test_device_bytes_accounts_for_states_and_dtype():
    """Footprint: 4 state buffers + (stages-1) stage rhs + 2 coeff(1) buffers
    + the pvar readout buffer, ghost-inclusive and row-padded, plus small
    per-run arrays."""
    cfg = S.kepler_disk_config(ni=512, nj=1024)
    b64 = fosite_amd.device_bytes(cfg)
    b32 = fosite_amd.device_bytes(dict(cfg, dtype="fp32"))
    state64 = 8 * 3 * (1024 + 4) * 528  # NJ x padded NI
    assert 9 * state64 <= b64 <= 9 * state64 * 1.01   # SSPRK3: 4 + 2 + 2 + 1 buffers
    assert 0.49 < b32 / b64 < 0.52
    dp5 = fosite_amd.device_bytes(S.euler_disk_config(ni=512, nj=1024))
    assert dp5 > 13 * 8 * 4 * 1028 * 528              # DP5: 4 + 6 + 2 + 1 buffers of 4 fields


def test_read_back_entry_points_reject_a_null_context():
    """The boundary-flux and background-velocity read-backs fail cleanly
    (no device call) on a null context or a bad direction."""
    import ctypes
    L = fosite_amd.load_library()
    buf = (ctypes.c_double * 16)()
    assert L.fv_get_boundary_flux(None, buf, buf) == -1
    assert L.fv_get_boundary_flux_total(None, 1, buf) == -1
    assert L.fv_get_background_velocity(None, buf) == -1
