"""Fingerprints of the compiled reference at the full workload sizes.

Runs oracle/_ref/fv_driver (the reference's own Fortran sources, compiled by
oracle/build_ref.py; FirstStep + Step(), fosite.f90:258-464) on the
workloads at their stated sizes and stores, per workload, the scalars of
every step (iter, n_adj, break, time, dt, dtold, maxerrold) and the
fingerprint (tests/fingerprint.py: per-field L1 norms, seeded weighted
projections, SHA-256 of the interior state) after chosen steps.  The states
themselves (12.7 MB and more each) are not stored.  Output:
tests/golden/fingerprints.npz (data only).

Usage (needs the reference checkout and its build; several minutes):
    python oracle/build_ref.py && python tests/golden/make_fingerprints.py
"""
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fosite_amd import setups as S  # noqa: E402
from oracle import refrun  # noqa: E402
from tests import fingerprint as F  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fingerprints.npz")

# name: (config, steps, dump interval)
CASES = {
    "riemann256_200": (S.riemann2d_config(n=256), 200, 50),
    "disk512x1024_1000": (S.kepler_disk_config(ni=512, nj=1024), 1000, 250),
    "euler_disk1024x2048_20": (S.euler_disk_config(ni=1024, nj=2048), 20, 10),
}
SCALARS = ("iter", "n_adj", "brk", "time", "dt", "dtold", "maxerrold")


def main():
    if not refrun.available():
        sys.exit("reference build missing: run oracle/build_ref.py first")
    out = {}
    names = sys.argv[1:] or list(CASES)
    if os.path.exists(OUT):
        with np.load(OUT) as z:
            out.update({k: z[k] for k in z.files})
    for name in names:
        cfg, steps, every = CASES[name]
        init = S.initial(cfg)
        with tempfile.TemporaryDirectory(prefix="fvfp_") as tmp:
            mesh, recs = refrun.run(cfg, init, steps, dump_every=every, workdir=tmp)
        sc = mesh["scalars"]
        assert [r["step"] for r in sc] == list(range(steps + 1)), name
        for k in list(out):
            if k.startswith(name + "/"):
                del out[k]
        out[name + "/config"] = np.frombuffer(json.dumps(cfg).encode(), dtype=np.uint8)
        out[name + "/steps"] = np.array([steps])
        out[name + "/step_seconds"] = np.array([mesh["step_seconds"][1]])
        for k in SCALARS:
            out[name + "/" + k] = np.array([r[k] for r in sc])
        w = None
        dumped = []
        for r in recs:
            fp = F.fingerprint(r["cvar"], w)
            n = r["step"]
            dumped.append(n)
            out["%s/l1_%d" % (name, n)] = fp["l1"]
            out["%s/proj_%d" % (name, n)] = fp["proj"]
            out["%s/sha256_%d" % (name, n)] = np.frombuffer(fp["sha256"].encode(), dtype=np.uint8)
        out[name + "/dumped"] = np.array(dumped)
        print(name, "steps", steps, "n_adj", sc[-1]["n_adj"], "time %.6e" % sc[-1]["time"],
              "reference step seconds %.1f" % mesh["step_seconds"][1], flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
