"""Write the golden fixtures of the FV step from the compiled reference.

Runs oracle/_ref/fv_driver -- the reference's own Fortran sources compiled
by oracle/build_ref.py, driven through FirstStep + Step() (fosite.f90:258-464)
-- on small seeded configurations and stores, per case, the configuration,
the step scalars after every step, the conservative state and the
accumulated boundary fluxes (with their GetBoundaryFlux sums) after chosen
steps (and, with accretion, the point mass per step and its acceleration
after the chosen steps), plus the reference's mesh arrays for one Cartesian and one
logcylindrical mesh.  Output: tests/golden/fv_golden.npz (data only).

Usage (needs the reference checkout and its build):
    python oracle/build_ref.py && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fosite_amd import setups as S  # noqa: E402
from oracle import refrun  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fv_golden.npz")

E = S.riemann2d_config(n=16)
D = S.kepler_disk_config(ni=16, nj=32)

# name: (config, steps, steps whose state is stored)
CASES = {
    "euler_vanleer_ssprk3_24": (S.riemann2d_config(n=24), 12, (1, 6, 12)),
    "euler_minmod_ssprk5": (dict(E, limiter=S.MINMOD, order=5), 6, (1, 6)),
    "euler_mc_1.2": (dict(E, limiter=S.MONOCENT, theta=1.2), 6, (6,)),
    "euler_sweby_1.5": (dict(E, limiter=S.SWEBY, theta=1.5), 6, (6,)),
    "euler_superbee": (dict(E, limiter=S.SUPERBEE), 6, (6,)),
    "euler_ospre": (dict(E, limiter=S.OSPRE), 6, (6,)),
    "euler_nolimit": (dict(E, limiter=S.NOLIMIT), 3, (3,)),
    "euler_periodic_rkf3": (dict(E, bc=[S.PERIODIC] * 4, method=S.RK_FEHLBERG, order=3), 8, (8,)),
    "euler_dp5": (dict(E, method=S.DORMAND_PRINCE, order=5, tol_rel=1e-4), 6, (6,)),
    "euler_rkf5": (dict(E, method=S.RK_FEHLBERG, order=5), 6, (6,)),
    "euler_ragged_20x12": (dict(E, inum=20, jnum=12), 8, (8,)),
    "disk_fargo_absorbing_16x32": (D, 10, (1, 10)),
    "disk_no_fargo": (dict(D, fargo=0), 8, (8,)),
    "disk_nograd_radial": (dict(D, bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (8,)),
    "disk_cfl_0.9": (dict(D, cfl=0.9), 12, (12,)),
    "euler_disk_absorbing_dp5_24x48": (S.euler_disk_config(ni=24, nj=48), 12, (1, 12)),
    "euler_disk_absorbing_ssprk3": (dict(S.euler_disk_config(ni=16, nj=32), method=S.SSPRK, order=3,
                                         tol_rel=0.01), 10, (10,)),
    # FARGO type 1: background velocity recomputed every step
    "disk_fargo1_absorbing": (dict(D, fargo=1), 10, (1, 10)),
    "disk_fargo1_cfl_0.9_rejects": (dict(D, fargo=1, cfl=0.9), 12, (12,)),
    # point-mass accretion through the western boundary (outbound = WEST)
    "disk_accretion_west": (dict(D, outbound=1), 10, (1, 10)),
    "disk_accretion_no_fargo_acclimit": (dict(D, outbound=1, fargo=0, acclimit=0.001), 8, (8,)),
    "disk_accretion_fargo1": (dict(D, outbound=1, fargo=1), 8, (8,)),
    "disk_accretion_cfl_0.9_rejects": (dict(D, outbound=1, cfl=0.9), 12, (12,)),
    "disk_accretion_east_no_fargo": (dict(D, outbound=2, fargo=0), 8, (8,)),
    "euler_disk_accretion": (dict(S.euler_disk_config(ni=16, nj=32), outbound=1), 8, (8,)),
    # PRIMITIVE reconstruction (fluxes_base.f90:378-381)
    "euler_prim_vanleer": (dict(E, recvars=S.PRIMITIVE), 10, (1, 10)),
    "euler_prim_minmod_dp5": (dict(E, recvars=S.PRIMITIVE, limiter=S.MINMOD, method=S.DORMAND_PRINCE, order=5,
                                   tol_rel=1e-4), 8, (8,)),
    "disk_prim": (dict(D, recvars=S.PRIMITIVE), 10, (10,)),
    "disk_prim_nograd_fargo1": (dict(D, recvars=S.PRIMITIVE, fargo=1,
                                     bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (8,)),
    "euler_disk_prim": (dict(S.euler_disk_config(ni=16, nj=32), recvars=S.PRIMITIVE), 8, (8,)),
    # Cash-Karp order 5 (timedisc_cashkarp.f90:101-113)
    "euler_cashkarp5": (dict(E, method=S.CASH_KARP, order=5), 6, (6,)),
    # CheckData limits and the pressure floor (timedisc_base.f90:353-359,
    # 1150-1164, 1666-1742): a density break at step 12, a floor that clips
    "rarefaction_rhomin_break": (S.rarefaction_config(n=16, rhomin=0.1), 12, (12,)),
    "rarefaction_pmin_floor": (S.rarefaction_config(n=16, pmin=0.03), 24, (24,)),
    # dtlimit break after SolveODE (timedisc_base.f90:686-713)
    "rarefaction_dtlimit_break": (S.rarefaction_config(n=16, dtlimit=0.0045), 3, (3,)),
    # output times: count = 5 before the stop time (fileio_base.f90:591-620)
    "euler_count5": (dict(E, stoptime=0.06, count=5), 10, (10,)),
    # a start time, as a restart sets it (timedisc_base.f90:295,
    # fileio_base.f90:474-487), with three outputs left
    "disk_starttime_count3": (dict(D, starttime=2.0, stoptime=8.0, count=3, fargo=1), 12, (12,)),
    # FARGO with Euler physics (energy terms, physics_euler.f90:2003-2047,
    # 2159-2203, 2368-2400) and on Cartesian meshes
    "shear_euler_fargo2": (S.shear_box_config(n=16), 10, (1, 10)),
    "shear_euler_fargo1": (S.shear_box_config(n=16, fargo=1), 10, (10,)),
    "shear_iso_fargo2": (S.shear_box_config(n=16, physics=S.EULER_ISOTHERM), 10, (10,)),
    "euler_disk_fargo2": (dict(S.euler_disk_config(ni=16, nj=32), fargo=2), 8, (8,)),
    "euler_disk_fargo1_prim": (dict(S.euler_disk_config(ni=16, nj=32), fargo=1, recvars=S.PRIMITIVE), 8, (8,)),
    # the rotating frame (sources_rotframe.f90:73-247), after gravity in the
    # source list; centrifugal and Coriolis terms, or Coriolis only
    # (disable_centaccel)
    "disk_rotframe": (dict(D, omega=0.3, rotframe=1), 8, (1, 8)),
    "disk_rotframe_nocent_fargo1": (dict(D, omega=0.5, rotframe=2, fargo=1), 8, (8,)),
    "disk_rotframe_nograd_no_fargo": (dict(D, omega=0.3, rotframe=1, fargo=0,
                                           bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (8,)),
    "euler_rotframe": (dict(E, omega=2.0, rotframe=1), 8, (1, 8)),
    "euler_disk_rotframe_prim": (dict(S.euler_disk_config(ni=16, nj=32), omega=0.2, rotframe=1,
                                      recvars=S.PRIMITIVE), 8, (8,)),
    # the viscosity source (sources_viscosity.f90): constant (MOLECULAR) and
    # Shakura-Sunyaev (ALPHA) viscosity, its time-step limit and the
    # update-only-when-the-time-changes rule (Dormand-Prince's stages 6, 7)
    "euler_viscous_molecular": (dict(E, vismodel=1, dynconst=1e-3), 8, (1, 8)),
    "euler_viscous_molecular_bulk_periodic": (dict(E, vismodel=1, dynconst=2e-3, bulkconst=5e-4,
                                                   bc=[S.PERIODIC] * 4), 8, (8,)),
    "euler_viscous_molecular_dp5": (dict(E, vismodel=1, dynconst=5e-3, cvis=0.2, method=S.DORMAND_PRINCE,
                                         order=5, tol_rel=1e-4), 8, (8,)),
    "disk_alpha_fargo1_rotframe": (dict(D, fargo=1, vismodel=2, dynconst=0.01, omega=0.3, rotframe=1,
                                        bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (1, 8)),
    "disk_alpha_no_fargo_limited": (dict(D, fargo=0, vismodel=2, dynconst=0.3, cvis=0.2,
                                         bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (8,)),
    "disk_alpha_dp5_absorbing": (dict(D, fargo=0, vismodel=2, dynconst=0.05, method=S.DORMAND_PRINCE, order=5,
                                      tol_rel=1e-4), 8, (8,)),
    # ALPHA with Euler physics: the sound speed of Timedisc%pvar, updated by
    # every ComputeRHS of the main state (timedisc_base.f90:1166-1174), also
    # in the stage evaluations
    "euler_disk_alpha": (dict(S.euler_disk_config(ni=16, nj=32), vismodel=2, dynconst=0.02,
                              bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (1, 8)),
    "euler_disk_alpha_fargo1_rotframe": (dict(S.euler_disk_config(ni=16, nj=32), vismodel=2, dynconst=0.01,
                                              fargo=1, omega=0.3, rotframe=1, method=S.SSPRK, order=3,
                                              bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]),
                                         8, (8,)),
    # BETA (Duschl): beta*rho*r^2*|Omega| with the bary-centre radius
    "disk_beta": (dict(D, fargo=0, vismodel=3, dynconst=1e-3,
                       bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (1, 8)),
    "euler_disk_beta_fargo1": (dict(S.euler_disk_config(ni=16, nj=32), vismodel=3, dynconst=1e-3, fargo=1,
                                    method=S.SSPRK, order=3,
                                    bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 8, (8,)),
}
MESHES = {"cartesian_16": E, "logcyl_16x32": D}
SCALARS = ("iter", "n_adj", "brk", "time", "dt", "dtold", "maxerrold")


def main():
    if not refrun.available():
        sys.exit("reference build missing: run oracle/build_ref.py first")
    # --only NAME ...: recompute those cases and keep the others from the
    # existing file
    only = sys.argv[sys.argv.index("--only") + 1:] if "--only" in sys.argv else None
    out = {}
    if only:
        with np.load(OUT, allow_pickle=False) as old:
            out = {k: old[k] for k in old.files if k.split("/")[0] not in only}
    for name, (cfg, steps, keep) in CASES.items():
        if only and name not in only:
            continue
        init = S.initial(cfg)
        _, recs = refrun.run(cfg, init, steps, dump_every=1)
        assert [r["step"] for r in recs] == list(range(steps + 1)), name
        out[name + "/config"] = np.frombuffer(json.dumps(
            {k: v for k, v in cfg.items()}).encode(), dtype=np.uint8)
        out[name + "/steps"] = np.array([steps])
        for k in SCALARS:
            out[name + "/" + k] = np.array([r[k] for r in recs])
        if cfg.get("outbound", 0):  # point mass: mass, accrate, massloss, mdot per step
            out[name + "/pointmass"] = np.array([r["pointmass"] for r in recs])
        for s in keep:
            out["%s/cvar_%d" % (name, s)] = recs[s]["cvar"]
            # boundary fluxes bxflux(JG,1,2,V) / byflux(1,IG,2,V) as [l][side][j|i]
            out["%s/bx_%d" % (name, s)] = recs[s]["bx"]
            out["%s/by_%d" % (name, s)] = recs[s]["by"]
            out["%s/bsum_%d" % (name, s)] = recs[s]["bsum"]
            if cfg.get("fargo", 0):  # Timedisc%w(IGMIN:IGMAX)
                out["%s/w_%d" % (name, s)] = recs[s]["w"]
            if cfg.get("outbound", 0):  # accel(:,:,:,1:2) of the point mass
                out["%s/accel_%d" % (name, s)] = recs[s]["accel"]
        print(name, "n_adj", recs[-1]["n_adj"], "time %.6e" % recs[-1]["time"])
    for name, cfg in (MESHES.items() if not only else ()):
        mesh, _ = refrun.run(cfg, S.initial(cfg), 0, dump_every=1)
        for k, v in mesh.items():
            if isinstance(v, np.ndarray):
                out["mesh_%s/%s" % (name, k)] = v
        out["mesh_%s/config" % name] = np.frombuffer(json.dumps(cfg).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
