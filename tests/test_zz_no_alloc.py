"""fv_step_noalloc, the allocation-free variant of fv_step, through the C
ABI: its reduction scratch is the library's static device array instead of
the context's own buffers.  Two calls in a row, on a larger and then on a
smaller mesh, each compared bit for bit with fv_step on an identical
context (Euler on a Cartesian box: no LOG/EXP on the path), and a mesh whose
RHS launch has more workgroups than the array holds, which the entry point
refuses before any work."""
import ctypes

import numpy as np
import pytest

import fosite_amd
from fosite_amd import setups as S

LARGE = S.kepler_disk_config(ni=96, nj=160)
SMALL = S.riemann2d_config(n=40)


def make(cfg):
    init = S.initial(cfg)
    o = fosite_amd.Fosite(cfg)
    o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    return o


def noalloc_step(o):
    L = o.L
    L.fv_step_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    d = ctypes.c_int()
    rc = L.fv_step_noalloc(o._h, ctypes.byref(d))
    if rc:
        raise RuntimeError("fv_step_noalloc failed (%d): %s" % (rc, L.fv_last_error(o._h).decode()))
    return bool(d.value)


def same(a, b, what):
    ca, pa = a.state()
    cb, pb = b.state()
    assert a.scalars() == b.scalars(), (what, a.scalars(), b.scalars())
    assert np.array_equal(ca, cb), (what, np.abs(ca - cb).max())
    assert np.array_equal(pa, pb), what
    assert np.array_equal(a.boundary_flux()[0], b.boundary_flux()[0]), what


@pytest.mark.gpu
def test_noalloc_step_equals_step_on_two_meshes_in_a_row():
    print("fv_step_noalloc: disk 96x160 (absorbing, FARGO), then Riemann 40x40", flush=True)
    big, big_ref = make(LARGE), make(LARGE)
    small, small_ref = make(SMALL), make(SMALL)
    for n in range(6):
        # the larger mesh first, then the smaller one: the second call reads
        # only the partials it wrote itself
        assert noalloc_step(big) == big_ref.step()
        assert noalloc_step(small) == small_ref.step()
        same(big, big_ref, ("disk", n))
        same(small, small_ref, ("riemann", n))
    for o in (big, big_ref, small, small_ref):
        o.close()


@pytest.mark.gpu
def test_noalloc_step_refuses_meshes_beyond_the_scratch():
    cfg = S.riemann2d_config(n=16)
    cfg.update(inum=2048, jnum=4200)  # 32 x 525 = 16800 workgroups of 64 x 8 cells
    o = fosite_amd.Fosite(cfg)
    o.L.fv_step_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    d = ctypes.c_int()
    rc = o.L.fv_step_noalloc(o._h, ctypes.byref(d))
    assert rc != 0
    assert b"static scratch" in o.L.fv_last_error(o._h)
    o.close()


@pytest.mark.gpu
def test_noalloc_step_with_the_viscosity_source():
    """The viscous time-step limit reduces through the same static scratch:
    an ALPHA-viscous disk, then a smaller MOLECULAR-viscous shear box, each
    equal to fv_step bit for bit."""
    print("fv_step_noalloc: viscous disk 64x96, then viscous shear box 40x40", flush=True)
    disk = dict(S.kepler_disk_config(ni=64, nj=96), vismodel=2, dynconst=0.01, omega=0.3, rotframe=1)
    box = S.shear_box_config(n=40, vismodel=1, dynconst=1e-3)
    big, big_ref = make(disk), make(disk)
    small, small_ref = make(box), make(box)
    for n in range(5):
        assert noalloc_step(big) == big_ref.step()
        assert noalloc_step(small) == small_ref.step()
        same(big, big_ref, ("viscous disk", n))
        same(small, small_ref, ("viscous box", n))
    for o in (big, big_ref, small, small_ref):
        o.close()
