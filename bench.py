#!/usr/bin/env python3
"""Benchmark: fosite's unsplit FV step on one MI355X.

Workload (N = 1): the plain FP64 drop-in at production size -- 2-D isothermal
Keplerian disk, logcylindrical 512 (r in [1,10]) x 1024 (phi), SSPRK3,
CFL 0.3, van Leer, KT, point-mass gravity, FARGO, absorbing/periodic
boundaries.  A "step" is one fosite Step() (one accepted time step of the
whole grid).  Inputs are resident in HBM before the timed region.

Prints one JSON line:
  value      cell-updates/s (interior cells x accepted steps / device time)
  roofline   the dominant kernel (largest share of the profiled device time):
             algorithmic bytes per launch / mean launch time against 8 TB/s,
             with its measured HBM traffic and its FP64 VALU-issue floor
  cpu_baseline  the compiled reference (oracle/_ref/fv_driver), or the C
             oracle if the reference build is absent, on a bounded sample
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch on this workload: rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE
# x1.0 (factors from a 1 GiB copy kernel), KiB per dispatch
# (profiles/k_rhs_disk512x1024_hbm_traffic.md)
TRAFFIC_BYTES_PER_LAUNCH = {"rhs": (2.0 * 13548 + 12420) * 1024, "rhs_combine": (2.0 * 21428 + 12523) * 1024,
                            "remap": (2.0 * 6633 + 12384) * 1024, "rhs_end": (2.0 * 7733 + 28220) * 1024}
# VALU-issue floor per launch in us: SQ_INSTS_VALU x 4 cycles + SQ_INSTS_VALU_TRANS_F64
# x 12 more per wave, x waves per SIMD, at 2.4 GHz
# (profiles/k_rhs_disk512x1024_sq_counters.md)
VALU_FLOOR_US = {"rhs": 9.55, "rhs_combine": 12.87, "remap": 2.69, "rhs_end": 10.89}
KERNEL_NAMES = {
    "rhs": "fv::k_rhs<double,3,VANLEER,64,3,M_STAGE>: stage 2 (stage sum over k_1, ghost cells, RHS)",
    "rhs_combine": "fv::k_rhs<double,3,VANLEER,64,3,M_LAST>: stage 3 with the final combination and "
                   "embedded error",
    "rhs_end": "fv::k_rhs<double,3,VANLEER,64,3,M_END>: end-of-step RHS (CFL, data checks, new state)",
    "remap": "fv::k_remap<double,3>: FARGO azimuthal remap",
}
METRIC = "cell-updates/s & HBM GB/s vs plain FP64 at stated err; p50/p99 run latency"


def b_alg(V=3, w=8, s=3, fargo=True):
    """Algorithmic HBM bytes per cell-update of one s-stage step as this
    implementation moves them: stage m (2..s-1) reads the state and m-1 stage
    slopes and writes one slope (m+1 words); the last stage also forms the
    final combination, reading the state and s-1 slopes and writing the new
    state (s+1); the FARGO remap reads and writes the state (2); the
    end-of-step RHS reads it and writes the first slope of the next step and
    the stored state (3).  w bytes per word, V conserved fields."""
    return w * V * (sum(m + 1 for m in range(2, s)) + (s + 1) + (2 if fargo else 0) + 3)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    """CPUs this process may run on (the scheduler's affinity mask), capped
    at 16, the host share of one GPU on the measuring machines."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_throughput(cfg, steps, procs):
    """Throughput mode of the serial reference: `procs` independent runs of
    the same workload on different seeds at once, one process each; the
    aggregate is the sum of their cell-updates over the wall clock of the
    slowest Step() loop."""
    from concurrent.futures import ThreadPoolExecutor
    from fosite_amd import setups
    from oracle import refrun
    ni, nj = cfg["inum"], cfg["jnum"]

    def one(k):
        c = dict(cfg, seed=cfg.get("seed", 20261202) + 1000 + k)
        mesh, _ = refrun.run(c, setups.initial(c), steps, dump_every=10 ** 9)
        return mesh["step_seconds"]

    with ThreadPoolExecutor(procs) as ex:
        res = list(ex.map(one, range(procs)))
    total = sum(ni * nj * n for n, _ in res)
    slowest = max(sec for _, sec in res)
    return dict(value=total / slowest, unit="cell-updates/s", processes=procs,
                per_process=[ni * nj * n / sec for n, sec in res],
                sample="%d concurrent serial reference runs of %d Step() calls of the %dx%d disk, "
                       "seeds differing; cell-updates of all / Step-loop wall clock of the slowest"
                       % (procs, steps, ni, nj))


def cpu_baseline(cfg, init, steps, throughput=True):
    from oracle import refrun
    ni, nj = cfg["inum"], cfg["jnum"]
    if refrun.available():
        mesh, _ = refrun.run(cfg, init, steps, dump_every=10 ** 9)
        n, sec = mesh["step_seconds"]
        out = dict(value=ni * nj * n / sec, unit="cell-updates/s", cores=1, kind="reference",
                   sample="%d Step() calls of the %dx%d disk (serial reference build, "
                          "amdflang -O3, wall clock of the Step loop)" % (n, ni, nj),
                   cpu_model=cpu_model(), host_cpus=os.cpu_count())
        if throughput:
            # the reference has no threading and no MPI library is present:
            # its multi-core figure is independent runs side by side
            out["throughput_mode"] = cpu_throughput(cfg, steps, host_cores())
        return out
    import oracle
    o = oracle.Oracle(cfg)
    o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    o.first_step()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
    sec = time.perf_counter() - t0
    return dict(value=ni * nj * steps / sec, unit="cell-updates/s", cores=1, kind="port",
                sample="%d steps of the %dx%d disk (C restatement, gcc -O2)" % (steps, ni, nj))


def run_until(sim, nsteps=None):
    """Advance in device-controlled batches until nsteps steps were taken or
    the run ended; returns (steps, device ms)."""
    taken, ms = 0, 0.0
    while nsteps is None or taken < nsteps:
        want = 512 if nsteps is None else min(512, nsteps - taken)
        k = sim.run(want)
        taken += k
        ms += sim.last_run_ms()
        if k < want:
            break
    return taken, ms


def field_errors(a, b):
    """Relative L1 over interior cells: density, momentum (vector), energy and
    the two momentum components."""
    import numpy as np
    I = (slice(2, -2), slice(2, -2))
    d, r = np.abs(a - b), np.abs(b)
    out = {"density": float(d[0][I].sum() / r[0][I].sum()),
           "momentum": float((d[1][I].sum() + d[2][I].sum()) / (r[1][I].sum() + r[2][I].sum())),
           "x_momentum": float(d[1][I].sum() / r[1][I].sum()),
           "y_momentum": float(d[2][I].sum() / r[2][I].sum())}
    if a.shape[0] == 4:
        out["energy"] = float(d[3][I].sum() / r[3][I].sum())
    return out


def precision_form(ni, nj, steps):
    """The precision form on the Euler-disk workload: an FP64 run of `steps`
    steps, the FP32 form run to the same simulation time, and an FP64 run with
    a 100x tighter tolerance as the accuracy reference.  Reports both
    throughputs, the relative L1 error of the FP32 run against the FP64 run
    and of both against the reference, and the relative mass drift of each run
    (FP64 sums of density x cell volume over the interior)."""
    import fosite_amd
    from fosite_amd import setups
    cfg64 = setups.euler_disk_config(ni=ni, nj=nj, steps=steps)
    init = setups.initial(cfg64)
    vol = 1.0 / fosite_amd.mesh_array(cfg64, "dxdydV")[2:-2, 2:-2]

    def mass(cv):
        return float((cv[0, 2:-2, 2:-2] * vol).sum())

    def run(cfg, nsteps=None):
        sim = fosite_amd.Fosite(cfg)
        sim.set_initial(init["pvar"])
        sim.first_step()
        cv0, _ = sim.state()
        n, ms = run_until(sim, nsteps)
        cv, _ = sim.state()
        sc = sim.scalars()
        sim.close()
        return cv0, cv, sc, n, ms

    cv0, cv64, sc64, n64, ms64 = run(cfg64, steps)
    t_end = sc64["time"]
    _, cv32, sc32, n32, ms32 = run(dict(cfg64, dtype="fp32", stoptime=t_end))
    _, cvref, scref, nref, _ = run(dict(cfg64, stoptime=t_end, tol_rel=cfg64["tol_rel"] * 1e-2))
    m0, m64, m32 = mass(cv0), mass(cv64), mass(cv32)
    cells = ni * nj
    return {"workload": "Euler disk, logcylindrical %dx%d, Dormand-Prince rtol 1e-4, absorbing r / "
                        "periodic phi, point-mass gravity, %d FP64 steps" % (ni, nj, steps),
            "fp64_cell_updates_per_s": cells * n64 / (ms64 * 1e-3),
            "fp32_cell_updates_per_s": cells * n32 / (ms32 * 1e-3),
            "speedup_fp32_over_fp64": (n32 / ms32) / (n64 / ms64),
            "steps_fp64": n64, "steps_fp32": n32, "steps_reference": nref,
            "t_end": t_end, "t_end_fp32": sc32["time"], "t_end_reference": scref["time"],
            "rel_l1_fp32_vs_fp64": field_errors(cv32, cv64),
            "rel_l1_fp64_vs_reference": field_errors(cv64, cvref),
            "rel_l1_fp32_vs_reference": field_errors(cv32, cvref),
            "reference": "FP64, rtol 1e-6",
            "mass_drift_fp64": (m64 - m0) / m0, "mass_drift_fp32": (m32 - m0) / m0}


def batched_form(n=64, steps=500, alone_steps=50, plain_fp64=None, err_members=4, err_steps=20):
    """The batched form: n independent isothermal disks of uneven sizes in
    one batch (every kernel of a step one launch over all disks, each round
    one HIP graph launch, tiling chosen by the autotuner), `steps` SSPRK3
    steps with FARGO.  Beside it: the same disks as an FP64 batch, each disk
    alone on the plain FP64 path (fv_run, `alone_steps` steps each), the
    plain FP64 rate of the 512x1024 disk (plain_fp64, cell-updates/s), and
    the FP32 members' error: the first err_members disks after err_steps
    steps in FP32 and in FP64, relative L1 per field (momentum as a vector)."""
    import fosite_amd
    from fosite_amd import setups

    def make(cfgs):
        sims = []
        for c in cfgs:
            init = setups.initial(c)
            sim = fosite_amd.Fosite(c)
            sim.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"])
            sim.first_step()
            sims.append(sim)
        return sims

    out = {}
    for dtype in ("fp32", "fp64"):
        cfgs = setups.batch_disks(n=n, steps=steps, dtype=dtype)
        sims = make(cfgs)
        b = fosite_amd.Batch(sims, autotune=True)
        taken = b.run(steps)
        info = b.info()
        upd = sum(c["inum"] * c["jnum"] * k for c, k in zip(cfgs, taken))
        out[dtype] = {"cell_updates_per_s": upd / (info["last_run_ms"] * 1e-3), "ms": info["last_run_ms"],
                      "member_steps": sum(taken), "tiling": ["64 row-major", "64 XCD-chunked", "32 row-major",
                                                             "32 XCD-chunked"][info["tiling"]],
                      "tune_ms_per_step": info["tune_ms"], "device_steps": info["device_steps"],
                      "host_steps": info["host_steps"], "halts": info["halts"]}
        b.close()
        for sim in sims:
            sim.close()
    cfgs = setups.batch_disks(n=n, steps=steps, dtype="fp64")
    upd, ms = 0, 0.0
    for c in cfgs:
        (sim,) = make([c])
        k = sim.run(alone_steps)
        ms += sim.last_run_ms()
        upd += c["inum"] * c["jnum"] * k
        sim.close()
    out["fp64_alone"] = {"cell_updates_per_s": upd / (ms * 1e-3), "ms": ms, "steps_per_disk": alone_steps}
    cells = sum(c["inum"] * c["jnum"] for c in cfgs)
    # FP32 members against the same members in FP64 (inertial frame, interior cells)
    import numpy as np
    errs = []
    for c32, c64 in zip(setups.batch_disks(n=n, steps=steps, dtype="fp32")[:err_members], cfgs[:err_members]):
        st = []
        for c in (c32, c64):
            (sim,) = make([c])
            sim.run(err_steps)
            cv, _ = sim.state()
            if c.get("fargo", 0):  # the stored azimuthal momentum is relative to rho*w
                cv[2] = cv[2] + cv[0] * sim.background_velocity()[None, :]
            st.append((cv[:, 2:-2, 2:-2], sim.scalars()["time"]))
            sim.close()
        (a, ta), (b, tb) = st
        mom = np.sqrt(b[1] ** 2 + b[2] ** 2)
        errs.append({"inum": c64["inum"], "jnum": c64["jnum"], "time_fp32": ta, "time_fp64": tb,
                     "density": float(np.abs(a[0] - b[0]).sum() / np.abs(b[0]).sum()),
                     "momentum": float(np.sqrt((a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2).sum() / mom.sum())})
    fp32 = out["fp32"]["cell_updates_per_s"]
    res = {"workload": "%d isothermal Keplerian disks, radial cells from {128,256,384,512}, azimuthal = 2x, "
                       "cs/vK U(0.03,0.08), FARGO, SSPRK3, %d steps (%d cells in all)" % (n, steps, cells),
           "fp32_batch": out["fp32"], "fp64_batch": out["fp64"], "fp64_alone": out["fp64_alone"],
           "speedup_fp32_batch_over_fp64_alone": fp32 / out["fp64_alone"]["cell_updates_per_s"],
           "speedup_fp32_batch_over_fp64_batch": fp32 / out["fp64"]["cell_updates_per_s"],
           "fp32_error_vs_fp64": {"members": err_members, "steps": err_steps, "per_member": errs,
                                  "max_density": max(e["density"] for e in errs),
                                  "max_momentum": max(e["momentum"] for e in errs)}}
    if plain_fp64:
        res["plain_fp64_512x1024"] = plain_fp64
        res["speedup_fp32_batch_over_plain_fp64_512x1024"] = fp32 / plain_fp64
    return res


def serving_form(n=300, rate=20.0, nmin=64, nmax=256, orbits=10.0, budget_gb=64.0, round_steps=32):
    """The serving loop: n disk runs arriving as a Poisson stream (mean `rate`
    per second), radial cells uniform in [nmin, nmax] (azimuthal 2x), FP32:FP64
    3:1, surface-density slope U(-1.5,-0.5), each integrated for `orbits` inner
    orbits; admitted by device-memory budget, continuous batching.  Latency of
    a run = completion - arrival (server clock); throughput = all cell updates
    / (last completion - first arrival)."""
    import time
    import numpy as np
    import fosite_amd
    from fosite_amd import setups
    reqs = setups.serving_requests(n=n, rate=rate, nmin=nmin, nmax=nmax, orbits=orbits)
    srv = fosite_amd.Server(budget_gb * 1e9, round_steps=round_steps, autotune=True)
    c0 = srv.stats()["clock"]
    t0 = time.perf_counter()
    ids, k, done = [], 0, 0
    while done < n:
        now = time.perf_counter() - t0
        while k < n and reqs[k][0] <= now:
            cfg = reqs[k][1]
            ids.append(srv.submit(cfg, setups.initial(cfg)))
            k += 1
        st = srv.stats()
        if st["running"] == 0 and st["queued"] == 0 and k < n:
            time.sleep(max(0.0, min(0.01, reqs[k][0] - (time.perf_counter() - t0))))
            continue
        srv.step()
        done = srv.stats()["completed"] + srv.stats()["failed"]
    st = srv.stats()
    lat, upd, t_end = [], 0.0, 0.0
    for rid, (arr, cfg) in zip(ids, reqs):
        info = srv.poll(rid)
        lat.append(info["t_done"] - (c0 + arr))
        upd += cfg["inum"] * cfg["jnum"] * (info["steps"] + 1)
        t_end = max(t_end, info["t_done"])
    lat = np.array(lat)
    wall = t_end - (c0 + reqs[0][0])
    srv.close()
    return {"workload": "%d disk runs, Poisson arrivals %.0f/s, radial cells U[%d,%d] (azimuthal 2x), FP32:FP64 "
                        "3:1, slope U(-1.5,-0.5), %g inner orbits, FARGO, SSPRK3" % (n, rate, nmin, nmax, orbits),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "cell_updates_per_s": upd / wall, "wall_s": wall,
            "completed": st["completed"], "failed": st["failed"], "peak_running": st["peak_running"],
            "peak_device_gb": st["peak_device_bytes"] / 1e9, "budget_gb": budget_gb, "rounds": st["rounds"],
            "gpu_busy_share": st["busy_ms"] * 1e-3 / wall, "cell_updates": upd,
            "arrival_span_s": reqs[-1][0] - reqs[0][0]}


def accretion_variant(steps=300):
    """The headline disk with point-mass accretion through the western
    boundary (outbound = 1): device-controlled steps with the accretion update
    (boundary-flux accumulation, the ordered density sum, rescaled gravity)
    before every end-of-step evaluation.  ms per step with and without it, and
    the update's launch time from the library's HIP-event profiling."""
    import fosite_amd
    from fosite_amd import setups
    out = {"workload": "isothermal Keplerian disk, logcylindrical 512x1024, FP64, SSPRK3, FARGO; "
                       "accretion through the western boundary"}
    for name, extra in (("no_accretion", {}), ("accretion", {"outbound": 1})):
        cfg = dict(setups.kepler_disk_config(ni=512, nj=1024), **extra)
        init = setups.initial(cfg)
        sim = fosite_amd.Fosite(cfg)
        sim.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"])
        sim.first_step()
        sim.run(20)
        taken = sim.run(steps)
        out[name + "_ms_per_step"] = sim.last_run_ms() / max(taken, 1)
        if extra:
            sim.profile(True)
            sim.run(50)
            n, ms, _ = sim.kernel_stats("accrete")
            sim.profile(False)
            out["accretion_update_us"] = 1e3 * ms / max(n, 1)
            out["point_mass"] = [float(x) for x in sim.point_mass()]
        sim.close()
    return out


def fargo1_variant(steps=300):
    """The headline disk with FARGO type 1 (the background velocity
    recomputed from the mean azimuthal velocity of every ring at each step)
    instead of the user-supplied one: ms per step of both, device-controlled."""
    import fosite_amd
    from fosite_amd import setups
    out = {"workload": "isothermal Keplerian disk, logcylindrical 512x1024, FP64, SSPRK3; FARGO type 1 vs type 2"}
    for name, fargo in (("type2", 2), ("type1", 1)):
        cfg = dict(setups.kepler_disk_config(ni=512, nj=1024), fargo=fargo)
        init = setups.initial(cfg)
        sim = fosite_amd.Fosite(cfg)
        sim.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"] if fargo == 2 else None)
        sim.first_step()
        sim.run(20)
        taken = sim.run(steps)
        out[name + "_ms_per_step"] = sim.last_run_ms() / max(taken, 1)
        sim.close()
    return out


def viscous_variant(steps=60, alpha=0.01, omega=0.3):
    """The headline disk in a rotating frame with ALPHA viscosity:
    device-controlled steps (the viscous source before every RHS evaluation,
    its time-step limit reduced in the step control), ms per step, and the
    viscous-source kernel's mean launch time and HBM rate (its algorithmic
    bytes: the stage inputs read and the source written) from the library's
    HIP-event profiling."""
    import fosite_amd
    from fosite_amd import setups
    cfg = dict(setups.kepler_disk_config(ni=512, nj=1024), vismodel=2, dynconst=alpha, omega=omega, rotframe=1)
    out = {"workload": "isothermal Keplerian disk, logcylindrical 512x1024, FP64, SSPRK3, FARGO; rotating frame "
                       "(omega %g), ALPHA viscosity (alpha %g), device-controlled steps" % (omega, alpha)}
    init = setups.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    sim.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"])
    sim.first_step()
    sim.run(5)
    taken = sim.run(steps)
    out["ms_per_step"] = sim.last_run_ms() / max(taken, 1)
    out["dtcause"] = sim.scalars()["dtcause"]
    sim.profile(True)
    sim.run(10)
    for k in ("visc", "rhs", "rhs_combine", "rhs_end"):
        n, ms, alg = sim.kernel_stats(k)
        out[k + "_us"] = 1e3 * ms / max(n, 1)
        if k == "visc" and ms > 0:
            out["visc_GBps"] = alg / (ms * 1e-3) / 1e9
            out["visc_hbm_frac"] = out["visc_GBps"] / 8000.0
    sim.profile(False)
    sim.close()
    return out


def serving_form_full(n=40):
    """The serving stream at the stated grid sizes, radial cells U[64,1024]
    (azimuthal 2x), 20 requests/s: the offered load (all cell updates of the
    n runs / their arrival span) is several times what one GPU serves, so the
    queue grows and the latencies measure the backlog.  A bounded prefix of n
    requests; the object adds the offered load and the share of it served."""
    out = serving_form(n=n, nmin=64, nmax=1024)
    span = out["arrival_span_s"]
    out["offered_cell_updates_per_s"] = out["cell_updates"] / span if span > 0 else None
    out["served_share"] = (out["cell_updates_per_s"] / out["offered_cell_updates_per_s"]
                           if span > 0 else None)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-throughput", action="store_true",
                    help="time the reference on one core only (no concurrent runs)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the precision, batched and serving forms")
    ap.add_argument("--form-steps", type=int, default=500)
    ap.add_argument("--only-form", choices=["precision", "batched", "serving", "serving_full", "accretion",
                                             "fargo1", "viscous"],
                    help="run one form alone and print its object (for profiling)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; the FV step would shard by radial slabs "
              "with ghost-ring exchange, but no multi-GPU run is implemented", file=sys.stderr)
        return 2
    if a.only_form:
        f = {"precision": lambda: precision_form(1024, 2048, a.form_steps), "batched": batched_form,
             "serving": serving_form, "serving_full": serving_form_full,
             "accretion": accretion_variant, "fargo1": fargo1_variant,
             "viscous": viscous_variant}[a.only_form]
        print(json.dumps(f()))
        return 0

    import fosite_amd
    from fosite_amd import setups
    cfg = setups.kepler_disk_config(ni=512, nj=1024)
    init = setups.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    t0 = time.perf_counter()
    sim.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"])
    upload_ms = (time.perf_counter() - t0) * 1e3  # host arrays -> HBM (marray layout)
    sim.first_step()
    sim.run(a.warmup)
    # timed region: device-controlled batches (HIP graphs), no per-kernel events
    taken = sim.run(a.steps)
    ms = sim.last_run_ms()
    sc = sim.scalars()
    cells = cfg["inum"] * cfg["jnum"]
    value = cells * taken / (ms * 1e-3)
    t0 = time.perf_counter()
    sim.state()
    download_ms = (time.perf_counter() - t0) * 1e3  # HBM -> host arrays (cvar and pvar)
    # per-kernel durations: the same device-controlled steps continued, each
    # launch bracketed by HIP events on the simulation's stream
    sim.profile(True)
    sim.run(a.profile_steps)
    stats = {k: sim.kernel_stats(k) for k in ("rhs", "rhs_combine", "rhs_end", "remap", "control")}
    sim.profile(False)
    kernels = {k: {"launches": n, "mean_launch_us": 1e3 * t / max(n, 1),
                   "alg_bytes_per_launch": b / max(n, 1),
                   "alg_GBps": b / (t * 1e-3) / 1e9 if t > 0 else None}
               for k, (n, t, b) in stats.items()}
    # the dominant kernel: the largest share of the profiled device time
    dom = max(TRAFFIC_BYTES_PER_LAUNCH, key=lambda k: stats[k][1])
    n_dom, ms_dom, bytes_dom = stats[dom]
    mean_us = 1e3 * ms_dom / max(n_dom, 1)
    achieved = bytes_dom / (ms_dom * 1e-3) / 1e9
    floor_us = VALU_FLOOR_US[dom]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": taken,
        "warmup": a.warmup,
        "ms_per_step": ms / max(taken, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 surface-density noise)",
        "config": {"workload": "isothermal Keplerian disk, logcylindrical 512x1024, FP64, SSPRK3, "
                               "CFL 0.3, van Leer + KT, point-mass gravity, FARGO, absorbing r / "
                               "periodic phi", "inum": 512, "jnum": 1024, "fields": 3,
                   "b_alg_bytes_per_cell_update": b_alg(),
                   "alg_GBps_whole_step": value * b_alg() / 1e9,
                   "sim_time_reached": sc["time"], "rejected_steps": sc["n_adj"],
                   "host_io_ms": {"upload": upload_ms, "download": download_ms},
                   "value_with_host_io": cells * taken / ((ms + upload_ms + download_ms) * 1e-3)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES_PER_LAUNCH[dom],
                     "kernel": KERNEL_NAMES[dom],
                     "launches": n_dom, "mean_launch_us": mean_us,
                     "alg_bytes_per_launch": bytes_dom / max(n_dom, 1),
                     "device_time_share": ms_dom / sum(t for _, t, _ in stats.values()),
                     # second ceiling: FP64 VALU issue (IEEE divisions keep the
                     # reciprocals on the transcendental unit)
                     "valu_issue_floor_us": floor_us, "valu_issue_frac": floor_us / mean_us,
                     "valu_ceiling_GBps": bytes_dom / max(n_dom, 1) / (floor_us * 1e-6) / 1e9,
                     "timing": "HIP events around each launch on the simulation stream, "
                               "%d device-controlled steps after the timed region" % a.profile_steps},
        "kernels": kernels,
    }
    sim.close()
    if not a.no_forms:
        out["precision_form"] = precision_form(1024, 2048, a.form_steps)
        out["batched_form"] = batched_form(plain_fp64=value)
        out["serving_form"] = serving_form()
        out["serving_form_full_sizes"] = serving_form_full()
        out["accretion_variant"] = accretion_variant()
        out["fargo1_variant"] = fargo1_variant()
        out["viscous_variant"] = viscous_variant()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, init, a.cpu_steps, throughput=not a.no_cpu_throughput)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
