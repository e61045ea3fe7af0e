#!/usr/bin/env python3
"""Benchmark: fosite's unsplit FV step on one MI355X.

Workload (N = 1): the plain FP64 drop-in at production size -- 2-D isothermal
Keplerian disk, logcylindrical 512 (r in [1,10]) x 1024 (phi), SSPRK3,
CFL 0.3, van Leer, KT, point-mass gravity, FARGO, absorbing/periodic
boundaries.  A "step" is one fosite Step() (one accepted time step of the
whole grid).  Inputs are resident in HBM before the timed region.

Prints one JSON line (see DESIGN.md for the roofline accounting):
  value      cell-updates/s (interior cells x accepted steps / device time)
  roofline   fused RHS kernel: algorithmic bytes per launch / mean launch time
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
# HBM bytes per launch of the stage-2 RHS kernel on this workload: rocprofv3
# FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (factors from a 1 GiB copy kernel),
# 15702 KiB and 13056 KiB per dispatch (profiles/k_rhs_disk512x1024_hbm_traffic.md)
TRAFFIC_BYTES_PER_LAUNCH = 45.5e6
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


def cpu_baseline(cfg, init, steps):
    from oracle import refrun
    ni, nj = cfg["inum"], cfg["jnum"]
    if refrun.available():
        mesh, _ = refrun.run(cfg, init, steps, dump_every=10 ** 9)
        n, sec = mesh["step_seconds"]
        return dict(value=ni * nj * n / sec, unit="cell-updates/s", cores=1, kind="reference",
                    sample="%d Step() calls of the %dx%d disk (serial reference build, "
                           "amdflang -O3, wall clock of the Step loop)" % (n, ni, nj))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the FV step shards by radial slabs "
              "(see DESIGN.md) but no multi-GPU run is part of it", file=sys.stderr)
        return 2

    import fosite_amd
    from fosite_amd import setups
    cfg = setups.kepler_disk_config(ni=512, nj=1024)
    init = setups.initial(cfg)
    sim = fosite_amd.Fosite(cfg)
    sim.set_initial(init["pvar"], init["bccsound"], init["fcsound"], init["w"])
    sim.first_step()
    sim.run(a.warmup)
    # timed region: device-controlled batches (HIP graphs), no per-kernel events
    taken = sim.run(a.steps)
    ms = sim.last_run_ms()
    sc = sim.scalars()
    cells = cfg["inum"] * cfg["jnum"]
    value = cells * taken / (ms * 1e-3)
    # per-kernel durations: the same device-controlled steps continued, each
    # launch bracketed by HIP events on the simulation's stream
    sim.profile(True)
    sim.run(a.profile_steps)
    stats = {k: sim.kernel_stats(k) for k in ("rhs", "rhs_combine", "rhs_end", "remap", "control")}
    sim.profile(False)
    n_rhs, ms_rhs, bytes_rhs = stats["rhs"]
    achieved = bytes_rhs / (ms_rhs * 1e-3) / 1e9
    kernels = {k: {"launches": n, "mean_launch_us": 1e3 * t / max(n, 1),
                   "alg_bytes_per_launch": b / max(n, 1),
                   "alg_GBps": b / (t * 1e-3) / 1e9 if t > 0 else None}
               for k, (n, t, b) in stats.items()}
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
                   "sim_time_reached": sc["time"], "rejected_steps": sc["n_adj"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES_PER_LAUNCH,
                     "kernel": "fv::k_rhs<3,BY,VANLEER> on a stage state (fused stage sum, ghost cells, RHS)",
                     "launches": n_rhs, "mean_launch_us": 1e3 * ms_rhs / max(n_rhs, 1),
                     "alg_bytes_per_launch": bytes_rhs / max(n_rhs, 1),
                     "timing": "HIP events around each launch on the simulation stream, "
                               "%d device-controlled steps after the timed region" % a.profile_steps},
        "kernels": kernels,
    }
    sim.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, init, a.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
