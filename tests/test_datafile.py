"""fosite_amd.datafile against datasets the reference itself wrote
(io/fileio_binary.f90; tests/golden/ref_dataset_*.bin from
tests/golden/make_datasets.py): every key the writer emits is in the
reference's file, in the same order, with the same type and bits -- mesh
coordinates, time, primitive variables, point mass and the configuration
values -- when the state is the oracle's (bit-exact with the reference)."""
import os

import numpy as np
import pytest

import oracle
from fosite_amd import datafile as DF
from fosite_amd import setups as S
from tests.dataset_cases import CASES

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference(name):
    return DF.read(os.path.join(GOLDEN, "ref_dataset_%s.bin" % name))


def oracle_output(cfg, nsteps):
    init = S.initial(cfg)
    o = oracle.Oracle(cfg)
    w = init.get("w") if cfg.get("fargo", 0) == 2 else None
    o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), w)
    o.first_step()
    for _ in range(nsteps):
        if o.step():
            break
    _, pv, sc = o.state()
    return sc["time"], pv, (o.pointmass() if cfg.get("gravity", 0) else None)


def ref_filename(ref):
    """/config/datafile/filename of a reference dataset (the path prefix the
    reference run was given), as a string."""
    return ref["/config/datafile/filename"][1].decode().rstrip(" ")


def compare(ours, ref, tol=0.0):
    assert list(ours) == list(ref)  # every key, in the reference's order
    for k, (t, v) in ours.items():
        rt, rv = ref[k]
        assert t == rt, (k, t, rt)
        if isinstance(v, np.ndarray):
            assert v.shape == rv.shape, (k, v.shape, rv.shape)
            if tol == 0.0 or not k.startswith("/timedisc/"):
                assert np.array_equal(v, rv), (k, np.abs(v - rv).max())
            else:
                assert np.abs(v - rv).sum() <= tol * np.abs(rv).sum(), k
        else:
            assert v == rv, (k, v, rv)


@pytest.mark.parametrize("name", list(CASES))
def test_dataset_of_the_oracle_state_equals_the_reference_dataset(name, tmp_path):
    cfg, nsteps = CASES[name]
    time, pv, pm = oracle_output(cfg, nsteps)
    path = str(tmp_path / "out.bin")
    ref = reference(name)
    DF.write(path, DF.dataset_entries(cfg, time, pv, pm, ref_filename(ref)))
    ours = DF.read(path)
    assert ours["/timedisc/time"][1] == cfg["stoptime"]
    compare(ours, ref)


@pytest.mark.parametrize("name", list(CASES))
def test_configuration_tree_depends_on_the_configuration(name):
    """Values that follow the configuration (not fixed defaults) move with
    it: dtlimit is EPSILON*stoptime unless given (timedisc_base.f90:308),
    the stop times, tol_abs has VNUM entries, the geometry's and FARGO's
    keys."""
    cfg, _ = CASES[name]
    base = dict(DF.config_entries(S.full(cfg)))
    alt = dict(DF.config_entries(S.full(dict(cfg, stoptime=2.5, dtlimit=-1.0))))
    assert alt["/config/timedisc/dtlimit"] == np.finfo(np.float64).eps * 2.5
    assert alt["/config/timedisc/stoptime"] == alt["/config/datafile/stoptime"] == 2.5
    given = DF.config_entries(S.full(dict(cfg, dtlimit=1e-9)))
    assert dict(given)["/config/timedisc/dtlimit"] == 1e-9
    keys = [k for k, _ in given]
    assert keys.index("/config/timedisc/dtlimit") < keys.index("/config/timedisc/starttime")
    assert len(base["/config/timedisc/tol_abs"]) == (3 if cfg["physics"] == S.EULER_ISOTHERM else 4)
    assert ("/config/mesh/dz" in base) == (cfg["geometry"] == S.CARTESIAN)
    keys = [k for k, _ in DF.config_entries(S.full(dict(cfg, fargo=1)))]
    assert keys.index("/config/mesh/fargo/method") < keys.index("/config/mesh/shearingbox")
    keys = [k for k, _ in DF.config_entries(S.full(dict(cfg, fargo=0)))]
    assert keys.index("/config/mesh/fargo/method") > keys.index("/config/mesh/rotation_center")
    assert ("/config/sources/grav/pmass/mass" in base) == bool(cfg.get("gravity", 0))


def test_reader_round_trip_and_header(tmp_path):
    a3 = np.arange(24.0).reshape(1, 4, 6)
    a4 = -np.arange(48.0).reshape(2, 1, 4, 6)
    path = str(tmp_path / "x.bin")
    DF.write(path, [("/n", None), ("/n/i", 7), ("/n/r", 0.25), ("/n/v", np.array([1.0, 2.0])),
                    ("/n/a", a3), ("/n/b", a4)])
    raw = open(path, "rb").read()
    assert raw[:13] == b"FOSITEII\x00 8 4"
    d = DF.read(path)
    assert d["/n"] == (DF.NODE, None) and d["/n/i"] == (DF.INT, 7) and d["/n/r"] == (DF.REAL, 0.25)
    assert np.array_equal(d["/n/v"][1], [1.0, 2.0])
    assert d["/n/a"][0] == DF.REAL_3D and np.array_equal(d["/n/a"][1], a3)
    assert d["/n/b"][0] == DF.REAL_4D and np.array_equal(d["/n/b"][1], a4)
    assert DF.filename_of("/tmp/x/run_data_0001.bin") == "/tmp/x/run_data"
    assert DF.filename_of("out.bin") == "out" and DF.filename_of("out") == "out"
    DF.write(path, [("/s", "run_data")])
    assert DF.read(path)["/s"] == (DF.CHAR, b"run_data".ljust(DF.MAX_CHAR_LEN))
    with pytest.raises(ValueError):
        DF.write(path, [("/s", "x" * (DF.MAX_CHAR_LEN + 1))])
