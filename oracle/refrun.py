"""Run the compiled reference (oracle/_ref/fv_driver) -- TEST INFRASTRUCTURE.

Writes a namelist and the initial-data file, runs the reference's
FirstStep + nsteps x Step (fosite.f90:258-464) through our setup program
oracle/drivers/fv_driver.f90, and parses the dumped records.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_ref", "fv_driver")
# the same program with fosite's Timedisc replaced by timedisc_gpu (the HIP
# library through fosite_amd/fortran/timedisc_gpu.f90); needs a GPU
DEVICE_DRIVER = os.path.join(HERE, "_ref", "fv_driver_device")


def available(device=False):
    return os.path.exists(DEVICE_DRIVER if device else DRIVER)


def _nml(cfg, nsteps, dump_every, infile, prefix):
    ta = list(cfg.get("tol_abs", [0.001] * 4)) + [0.001] * 4
    vals = dict(k_physics=cfg["physics"], k_geometry=cfg["geometry"], k_inum=cfg["inum"],
                k_jnum=cfg["jnum"], k_gamma=cfg.get("gamma", 1.4), k_xmin=cfg["xmin"],
                k_xmax=cfg["xmax"], k_ymin=cfg["ymin"], k_ymax=cfg["ymax"],
                k_gparam=cfg.get("gparam", 1.0), k_bcw=cfg["bc"][0], k_bce=cfg["bc"][1],
                k_bcs=cfg["bc"][2], k_bcn=cfg["bc"][3], k_limiter=cfg.get("limiter", 1),
                k_recvars=cfg.get("recvars", 2), k_theta=cfg.get("theta", 1.0),
                k_method=cfg.get("method", 5), k_order=cfg.get("order", 5),
                k_cfl=cfg.get("cfl", 0.4), k_tolrel=cfg.get("tol_rel", 0.01),
                k_stoptime=cfg["stoptime"], k_dtlimit=cfg.get("dtlimit", -1.0),
                k_beta=cfg.get("beta", 0.08), k_fargo=cfg.get("fargo", 0),
                k_gravity=cfg.get("gravity", 0), k_units=3 if cfg.get("gravity", 0) else 0,
                k_mass=cfg.get("mass", 1.0), k_outbound=cfg.get("outbound", 0),
                k_acclimit=float(cfg.get("acclimit", -1.0)), k_binary=int(cfg.get("binary_output", 0)),
                k_nsteps=nsteps, k_dumpevery=dump_every, k_count=int(cfg.get("count", 1)),
                k_pmin=float(cfg.get("pmin", -1.0)), k_rhomin=float(cfg.get("rhomin", -1.0)),
                k_starttime=float(cfg.get("starttime", 0.0)), k_omega=float(cfg.get("omega", 0.0)),
                k_rotframe=int(cfg.get("rotframe", 0)), k_vismodel=int(cfg.get("vismodel", 0)),
                k_dynconst=float(cfg.get("dynconst", 0.1)), k_cvis=float(cfg.get("cvis", 0.5)),
                k_setbulk=int(cfg.get("bulkconst") is not None),
                k_bulkconst=float(cfg["bulkconst"]) if cfg.get("bulkconst") is not None else 0.0)
    lines = ["&fvcfg"]
    for k, v in vals.items():
        lines.append(" %s = %s" % (k, repr(float(v)) if isinstance(v, float) else v))
    lines.append(" k_tolabs = " + ", ".join(repr(float(x)) for x in ta[:4]))
    lines.append(' k_infile = "%s"' % infile)
    lines.append(' k_prefix = "%s"' % prefix)
    lines.append("/")
    return "\n".join(lines) + "\n"


def run(cfg, init, nsteps, dump_every=1, workdir=None, device=False):
    """Returns (mesh dict, list of step records {step, iter, n_adj, brk, time,
    dt, dtold, maxerrold, cvar, pvar}).  With cfg["binary_output"] = 1 the
    reference also writes its binary datasets (run_data_NNNN.bin, at time 0
    and at the stop time) into workdir.  device=True runs fv_driver_device."""
    NI, NJ = cfg["inum"] + 4, cfg["jnum"] + 4
    N = NI * NJ
    V = 4 if cfg["physics"] == 17 else 3
    tmp = workdir or tempfile.mkdtemp(prefix="fvref_")
    infile = os.path.join(tmp, "init.bin")
    prefix = os.path.join(tmp, "run")
    with open(infile, "wb") as f:
        f.write(np.ascontiguousarray(init["pvar"], dtype="<f8").tobytes())
        if V == 3:
            f.write(np.ascontiguousarray(init["bccsound"], dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(init["fcsound"], dtype="<f8").tobytes())
        if cfg.get("fargo", 0) == 2:
            f.write(np.ascontiguousarray(init["w"], dtype="<f8").tobytes())
    nml = os.path.join(tmp, "cfg.nml")
    with open(nml, "w") as f:
        f.write(_nml(cfg, nsteps, dump_every, infile, prefix))
    r = subprocess.run([DEVICE_DRIVER if device else DRIVER, nml], cwd=tmp, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference driver failed:\n" + r.stdout[-3000:] + r.stderr[-3000:])
    step_seconds = None
    for line in r.stdout.splitlines():
        if line.startswith("FV_DRIVER_STEP_SECONDS"):
            _, ns, sec = line.split()
            step_seconds = (int(ns), float(sec))
    m = np.fromfile(prefix + "_mesh.bin", dtype="<f8")
    o = 4
    mesh = dict(dx=m[0], dy=m[1], invdx=m[2], invdy=m[3])
    def take(name, k=1):
        nonlocal o
        a = m[o:o + k * N].reshape((k, NJ, NI) if k > 1 else (NJ, NI))
        o += k * N
        mesh[name] = a
    take("dAx", 2); take("dAy", 2)
    for n in ("dydzdV", "dzdxdV", "dxdydV", "dlx", "dly", "cxyx", "cyxy", "czxz", "czyz"):
        take(n)
    for n in ("rdx_w", "rdx_e", "rdx_s", "rdx_n", "radius", "radius_w", "radius_e",
              "radius_s", "radius_n", "posvec_x", "posvec_y"):
        take(n)
    recs = []
    raw = open(prefix + "_steps.bin", "rb").read()
    off = 0
    rec = 48 + 2 * V * N * 8
    while off + rec <= len(raw):
        ib = np.frombuffer(raw, "<i4", 4, off)
        rb = np.frombuffer(raw, "<f8", 4, off + 16)
        cv = np.frombuffer(raw, "<f8", V * N, off + 48).reshape(V, NJ, NI).copy()
        pv = np.frombuffer(raw, "<f8", V * N, off + 48 + V * N * 8).reshape(V, NJ, NI).copy()
        recs.append(dict(step=int(ib[0]), iter=int(ib[1]), n_adj=int(ib[2]), brk=int(ib[3]),
                         time=rb[0], dt=rb[1], dtold=rb[2], maxerrold=rb[3], cvar=cv, pvar=pv))
        off += rec
    # boundary fluxes of every record: bxflux(JG,1,2,V), byflux(1,IG,2,V),
    # then GetBoundaryFlux of W, E, S, N (V values each)
    braw = np.fromfile(prefix + "_bflux.bin", dtype="<f8")
    brec = 2 * V * NJ + 2 * V * NI + 4 * V
    for n, r in enumerate(recs):
        b = braw[n * brec:(n + 1) * brec]
        r["bx"] = b[:2 * V * NJ].reshape(V, 2, NJ).copy()
        r["by"] = b[2 * V * NJ:2 * V * (NJ + NI)].reshape(V, 2, NI).copy()
        r["bsum"] = b[2 * V * (NJ + NI):].reshape(4, V).copy()
    if cfg.get("fargo", 0):
        wraw = np.fromfile(prefix + "_w.bin", dtype="<f8")
        for n, r in enumerate(recs):
            r["w"] = wraw[n * NI:(n + 1) * NI].copy()
    if cfg.get("gravity", 0):
        # point mass of every record: mass, accrate, massloss, mdot, then
        # accel(:,:,:,1:2) (gravity_pointmass.f90:79-85)
        graw = np.fromfile(prefix + "_grav.bin", dtype="<f8")
        grec = 4 + 2 * N
        for n, r in enumerate(recs):
            g = graw[n * grec:(n + 1) * grec]
            r["pointmass"] = g[:4].copy()
            r["accel"] = g[4:].reshape(2, NJ, NI).copy()
    # scalars of every step (written also between state dumps)
    spath = prefix + "_scal.bin"
    mesh["scalars"] = []
    if os.path.exists(spath):
        sraw = open(spath, "rb").read()
        for off in range(0, len(sraw) - 47, 48):
            ib = np.frombuffer(sraw, "<i4", 4, off)
            rb = np.frombuffer(sraw, "<f8", 4, off + 16)
            mesh["scalars"].append(dict(step=int(ib[0]), iter=int(ib[1]), n_adj=int(ib[2]), brk=int(ib[3]),
                                        time=rb[0], dt=rb[1], dtold=rb[2], maxerrold=rb[3]))
    mesh["step_seconds"] = step_seconds
    if workdir is None:
        shutil.rmtree(tmp, ignore_errors=True)
    return mesh, recs
