"""The Timedisc-level binding compiles and links against fosite's own
classes: fosite_amd/fortran/timedisc_gpu.f90 (timedisc_gpu EXTENDS
timedisc_base, new_timedisc_device, fosite_device EXTENDS fosite) with the
module files of the reference build, and the reference's setup program
(oracle/drivers/fv_driver.f90 with FV_DEVICE_TIMEDISC) linked with the
reference library and libfosite_fv.so.  CPU only: nothing is run."""
import os
import subprocess
import tempfile

import pytest

from oracle import build_ref

REF = build_ref.OUT
MOD = os.path.join(REF, "mod")
LIB = os.path.join(REF, "libfosite_ref.a")
HAVE = os.path.isdir(MOD) and os.path.exists(LIB) and os.path.exists(build_ref.LIBFV)


@pytest.mark.skipif(not HAVE, reason="reference build (module files, library) or libfosite_fv.so absent")
def test_device_timedisc_compiles_and_links_against_fosite():
    with tempfile.TemporaryDirectory() as d:
        objs = [build_ref.compile_one(os.path.join(build_ref.FORTRAN, f), d, MOD)
                for f in ("fosite_fv_c.f90", "timedisc_gpu.f90")]
        drv = os.path.join(build_ref.HERE, "drivers", "fv_driver.f90")
        obj = os.path.join(d, "drv.o")
        subprocess.check_call([build_ref.FC] + build_ref.FFLAGS + ["-DFV_DEVICE_TIMEDISC", "-J", d, "-I", MOD,
                                                                   "-c", drv, "-o", obj])
        exe = os.path.join(d, "fv_driver_device")
        subprocess.check_call([build_ref.FC, obj] + objs + [LIB, "-L" + os.path.dirname(build_ref.LIBFV),
                                                            "-lfosite_fv", "-o", exe])
        syms = subprocess.run(["nm", exe], capture_output=True, text=True).stdout
        # the overriding procedures and the factory are linked in, beside the
        # reference's own timedisc_base procedures
        for name in ("integrationstep_gpu", "computerhs_gpu", "calctimestep_gpu", "new_timedisc_device",
                     "setup_device"):
            assert name in syms.lower(), name
        assert "timedisc_base_mod" in syms.lower()
        # the C ABI entry points it binds are undefined in it, resolved by libfosite_fv.so
        for name in ("fv_first_rhs", "fv_integration_step", "fv_calc_timestep", "fv_add_background_velocity"):
            assert (" U " + name) in syms, name
