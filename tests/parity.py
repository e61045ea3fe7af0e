"""Shared helpers for parity tests: run the HIP path and the oracle side by
side on the same initial data and compare step by step."""
import numpy as np

from fosite_amd import setups


def rel_l1(a, b):
    """relative L1 per conserved field over interior cells"""
    out = []
    for l in range(a.shape[0]):
        x, y = a[l, 2:-2, 2:-2], b[l, 2:-2, 2:-2]
        den = np.abs(y).sum()
        num = np.abs(x - y).sum()
        out.append(num / den if den > 0 else num)
    return np.array(out)


def field_errors(a, b):
    """Relative L1 over interior cells of density, momentum (the vector:
    sum of |dm_x| + |dm_y| over sum of |m_x| + |m_y|) and energy (when
    present), plus each momentum component on its own."""
    I = (slice(2, -2), slice(2, -2))
    d = np.abs(a - b)

    def r(num, den):
        return float(num / den) if den > 0 else float(num)
    out = {"density": r(d[0][I].sum(), np.abs(b[0][I]).sum()),
           "momentum": r(d[1][I].sum() + d[2][I].sum(), np.abs(b[1][I]).sum() + np.abs(b[2][I]).sum()),
           "x_momentum": r(d[1][I].sum(), np.abs(b[1][I]).sum()),
           "y_momentum": r(d[2][I].sum(), np.abs(b[2][I]).sum())}
    if a.shape[0] == 4:
        out["energy"] = r(d[3][I].sum(), np.abs(b[3][I]).sum())
    return out


def run_pair(cfg, nsteps, every=1):
    """Yields (step, gpu_cvar, gpu_scalars, oracle_cvar, oracle_scalars)."""
    import fosite_amd
    import oracle
    init = setups.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    ref = oracle.Oracle(cfg)
    args = (init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    sim.set_initial(*args)
    ref.set_initial(*args)
    sim.first_step()
    ref.first_step()
    cv, _ = sim.state()
    rc, _, rs = ref.state()
    yield 0, cv, sim.scalars(), rc, rs
    for n in range(1, nsteps + 1):
        d1 = sim.step()
        d2 = ref.step()
        if n % every == 0 or n == nsteps or d1 or d2:
            cv, _ = sim.state()
            rc, _, rs = ref.state()
            yield n, cv, sim.scalars(), rc, rs
        if d1 or d2:
            assert d1 == d2
            break
    sim.close()


def run_batched(cfg, nsteps, chunk):
    """Like run_pair, but the HIP side advances with Fosite.run(chunk) (the
    device-controlled batches) while the oracle steps one at a time.
    Yields (steps so far, gpu_cvar, gpu_scalars, oracle_cvar, oracle_scalars,
    run ended)."""
    import fosite_amd
    import oracle
    init = setups.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    ref = oracle.Oracle(cfg)
    args = (init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    sim.set_initial(*args)
    ref.set_initial(*args)
    sim.first_step()
    ref.first_step()
    n = 0
    while n < nsteps:
        k = min(chunk, nsteps - n)
        taken = sim.run(k)
        for _ in range(taken):
            assert not ref.step()
        n += taken
        ended = taken < k
        if ended:
            assert ref.step(), "oracle did not end where the HIP run ended"
        cv, _ = sim.state()
        rc, _, rs = ref.state()
        yield n, cv, sim.scalars(), rc, rs, ended
        if ended:
            break
    sim.close()
