"""Host-side pieces: configuration marshalling and the seeded initial data."""
import numpy as np
import pytest

import fosite_amd
from fosite_amd import setups as S


def test_splitmix64_known_answer():
    # SplitMix64 from state 0: first output 0xE220A8397B1DCDAF (published test vector)
    u = S.splitmix64_uniform(0, 2)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert 0.0 <= u.min() and u.max() < 1.0


def test_config_struct_carries_every_field():
    cfg = S.kepler_disk_config(ni=40, nj=64)
    c = fosite_amd.FvConfig.from_dict(cfg)
    assert (c.physics, c.geometry, c.inum, c.jnum) == (S.EULER_ISOTHERM, S.LOGCYLINDRICAL, 40, 64)
    assert list(c.bc) == cfg["bc"] and c.fargo == 2 and c.gravity == 1 and c.limiter == S.VANLEER
    assert c.cfl == 0.3 and c.stoptime == 1e3 and list(c.tol_abs) == [0.001] * 4


def test_disk_initial_data_matches_its_definition():
    cfg = S.kepler_disk_config(ni=32, nj=64)
    init = S.initial(cfg)
    xc, _, dx, _ = S.coords(cfg)
    r = np.exp(xc)
    rho = init["pvar"][0]
    # Sigma ~ 1/r with 1% noise, Keplerian azimuthal velocity, no radial velocity
    assert np.all(np.abs(rho * r[None, :] - 1.0) <= 0.01 + 1e-15)
    np.testing.assert_allclose(init["pvar"][2], np.broadcast_to(1 / np.sqrt(r), rho.shape), rtol=1e-15)
    assert not init["pvar"][1].any()
    np.testing.assert_allclose(init["bccsound"][0], 0.05 / np.sqrt(r), rtol=1e-15)
    assert np.array_equal(init["w"], init["pvar"][2][0])


def test_riemann_quadrants():
    init = S.initial(S.riemann2d_config(n=8))
    pv = init["pvar"]
    assert tuple(pv[:, 10, 10]) == (1.5, 0.0, 0.0, 1.5)      # NE
    assert tuple(pv[:, 10, 1]) == (0.5323, 1.206, 0.0, 0.3)  # NW
    assert tuple(pv[:, 1, 1]) == (0.138, 1.206, 1.206, 0.029)  # SW
    assert tuple(pv[:, 1, 10]) == (0.5323, 0.0, 1.206, 0.3)  # SE


def test_initial_arrays_are_checked_against_the_abi_sizes():
    V, NJ, NI = 3, 12, 10
    pv = np.zeros((V, NJ, NI))
    b, f, w, p = fosite_amd.init_arrays(V, NJ, NI, np.ones((NJ, NI)), np.ones((4, NJ, NI)), np.ones(NI), pv)
    assert (b.size, f.size, w.size, p.size) == (NJ * NI, 4 * NJ * NI, NI, V * NJ * NI)
    bad = [dict(bccsound=np.ones(NJ * NI - 1), fcsound=np.ones((4, NJ, NI))),
           dict(bccsound=np.ones((NJ, NI)), fcsound=np.ones((3, NJ, NI))),
           dict(bccsound=np.ones((NJ, NI))),
           dict(w=np.ones(NI - 1)),
           dict(pvar=np.zeros((V, NJ, NI - 1)))]
    for kw in bad:
        with pytest.raises(ValueError):
            fosite_amd.init_arrays(V, NJ, NI, **kw)
