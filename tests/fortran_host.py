"""Drive the Fortran 2008 host program (fosite_amd/fortran/build/fv_run)
with a configuration and initial data, and read back its state records."""
import os
import subprocess

import numpy as np

from fosite_amd import setups
from oracle import refrun

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FV_RUN = os.path.join(ROOT, "fosite_amd", "fortran", "build", "fv_run")


def write_inputs(cfg, nsteps, tmp, batch=0, dump_every=1):
    init = setups.initial(cfg)
    infile = os.path.join(tmp, "init.bin")
    with open(infile, "wb") as f:
        f.write(np.ascontiguousarray(init["pvar"], dtype="<f8").tobytes())
        if cfg["physics"] == setups.EULER_ISOTHERM:
            f.write(np.ascontiguousarray(init["bccsound"], dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(init["fcsound"], dtype="<f8").tobytes())
        if cfg.get("fargo", 0) == 2:
            f.write(np.ascontiguousarray(init["w"], dtype="<f8").tobytes())
    nml = refrun._nml(cfg, nsteps, dump_every, infile, os.path.join(tmp, "run"))
    # the reference driver's dataset switch is not a key of this program
    nml = "\n".join(l for l in nml.split("\n") if not l.strip().startswith("k_binary"))
    nml = nml.replace("\n/\n", "\n k_batch = %d\n/\n" % batch)
    path = os.path.join(tmp, "cfg.nml")
    with open(path, "w") as f:
        f.write(nml)
    return path


def run(cfg, nsteps, tmp, batch=0, dump_every=1, timeout=300):
    """Returns (completed process, records)."""
    path = write_inputs(cfg, nsteps, tmp, batch, dump_every)
    r = subprocess.run([FV_RUN, path], cwd=tmp, capture_output=True, text=True, timeout=timeout)
    recs = []
    out = os.path.join(tmp, "run_steps.bin")
    bout = os.path.join(tmp, "run_bflux.bin")
    if r.returncode == 0 and os.path.exists(out):
        NI, NJ = cfg["inum"] + 4, cfg["jnum"] + 4
        V = 4 if cfg["physics"] == setups.EULER else 3
        N = NI * NJ
        raw = open(out, "rb").read()
        rec, off = 48 + 2 * V * N * 8, 0
        while off + rec <= len(raw):
            ib = np.frombuffer(raw, "<i4", 4, off)
            rb = np.frombuffer(raw, "<f8", 4, off + 16)
            cv = np.frombuffer(raw, "<f8", V * N, off + 48).reshape(V, NJ, NI).copy()
            recs.append(dict(step=int(ib[0]), iter=int(ib[1]), n_adj=int(ib[2]), brk=int(ib[3]),
                             time=rb[0], dt=rb[1], dtold=rb[2], maxerrold=rb[3], cvar=cv))
            off += rec
        # boundary fluxes of every record (GetBoundaryFluxes, GetBoundaryFlux x4)
        if os.path.exists(bout):
            braw = np.fromfile(bout, dtype="<f8")
            grav = 4 if cfg.get("gravity", 0) else 0  # GetPointMass after the totals
            brec = 2 * V * NJ + 2 * V * NI + 4 * 4 + grav
            for n, d in enumerate(recs):
                b = braw[n * brec:(n + 1) * brec]
                d["bx"] = b[:2 * V * NJ].reshape(V, 2, NJ)
                d["by"] = b[2 * V * NJ:2 * V * (NJ + NI)].reshape(V, 2, NI)
                d["bsum"] = b[2 * V * (NJ + NI):2 * V * (NJ + NI) + 16].reshape(4, 4)[:, :V]
                if grav:
                    d["pointmass"] = b[-4:].copy()
    return r, recs
