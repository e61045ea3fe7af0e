"""The x <-> y transposition invariant of the reference's own tests
(tests/KHI2D.f90:90-150 without FARGO): a run on a Cartesian mesh and the run
on the transposed mesh from the transposed initial data (u <-> v) give
transposed results, bit for bit.  The unsplit scheme adds the x and y flux
differences, sources and CFL terms with commutative FP additions, so the
reference holds it exactly; on the HIP path the x and y directions take
different code (64-wide tile rows against 8-row tile columns, different tile
edges on the two meshes), so the invariant checks that both directions
compute the same arithmetic.  Host steps and device rounds, periodic and
no-gradient boundaries, van Leer and minmod, SSPRK3 and DP5."""
import numpy as np
import pytest

import fosite_amd
from fosite_amd import setups as S
from fosite_amd.setups import splitmix64_uniform

I = (slice(None), slice(2, -2), slice(2, -2))


def config(ni, nj, **kw):
    # cells of 2^-6 in both directions, so dx = dy exactly
    c = dict(S.riemann2d_config(n=16), inum=ni, jnum=nj, xmin=0.0, xmax=ni / 64.0, ymin=0.0, ymax=nj / 64.0,
             bc=[S.PERIODIC] * 4, stoptime=10.0)
    c.update(kw)
    return c


def shear_layer(ni, nj, seed):
    """a Kelvin-Helmholtz-like shear layer along x with seeded noise, primitive
    variables on all cells (ghosts are set by the boundaries)"""
    pv = np.zeros((4, nj + 4, ni + 4))
    y = (np.arange(nj + 4) - 1.5) / nj
    u = splitmix64_uniform(seed, 3 * (nj + 4) * (ni + 4)).reshape(3, nj + 4, ni + 4)
    band = (np.abs(y - 0.5) < 0.25)[:, None]
    pv[0] = np.where(band, 2.0, 1.0) + 0.02 * (u[0] - 0.5)
    pv[1] = np.where(band, 0.5, -0.5) + 0.02 * (u[1] - 0.5)
    pv[2] = 0.05 * (u[2] - 0.5)
    pv[3] = 2.5
    return pv


def transposed(pv):
    """x <-> y: fields transposed, x and y velocity (or momentum) swapped"""
    out = np.empty((pv.shape[0],) + pv.shape[:0:-1])
    out[0] = pv[0].T
    out[1] = pv[2].T
    out[2] = pv[1].T
    out[3] = pv[3].T
    return out


CASES = {
    "periodic_vanleer_ssprk3": (dict(), 40),
    "nograd_minmod": (dict(bc=[S.NO_GRADIENTS] * 4, limiter=S.MINMOD), 30),
    "periodic_dp5": (dict(method=S.DORMAND_PRINCE, order=5, tol_rel=1e-4), 20),
}


@pytest.mark.gpu
@pytest.mark.parametrize("device_rounds", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_transposed_run_is_the_transpose(name, device_rounds):
    extra, nsteps = CASES[name]
    ni, nj = 136, 72  # tile edges fall differently on the two meshes
    pv = shear_layer(ni, nj, 20261205)
    print("transpose", name, "device rounds" if device_rounds else "host steps", flush=True)
    runs = []
    for cfg, p in ((config(ni, nj, **extra), pv), (config(nj, ni, **extra), transposed(pv))):
        sim = fosite_amd.Fosite(cfg)
        sim.set_initial(p)
        sim.first_step()
        if device_rounds:
            for k in (1, nsteps // 2, nsteps - 1 - nsteps // 2):
                assert sim.run(k) == k
        else:
            for _ in range(nsteps):
                assert not sim.step()
        runs.append((sim.state()[0], sim.scalars()))
        sim.close()
    (a, sa), (b, sb) = runs
    assert sa["time"] == sb["time"] and sa["dt"] == sb["dt"] and sa["n_adj"] == sb["n_adj"], (sa, sb)
    assert np.array_equal(transposed(a)[I], b[I]), np.abs(transposed(a)[I] - b[I]).max()
    assert np.abs(a[1][I[1:]]).max() > 0.1  # the shear flow is still there


def test_oracle_transposed_run_is_the_transpose():
    """The same invariant on the oracle (the reference's algorithm holds it)."""
    import oracle
    ni, nj = 40, 24
    pv = shear_layer(ni, nj, 7)
    res = []
    for cfg, p in ((config(ni, nj), pv), (config(nj, ni), transposed(pv))):
        o = oracle.Oracle(cfg)
        o.set_initial(p)
        o.first_step()
        for _ in range(15):
            o.step()
        cv, _, sc = o.state()
        res.append((cv, sc))
    (a, sa), (b, sb) = res
    assert sa["time"] == sb["time"]
    assert np.array_equal(transposed(a)[I], b[I])
