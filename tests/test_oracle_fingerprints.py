"""The oracle against the compiled reference at the full workload sizes,
through the fingerprints the reference's own runs left
(tests/golden/fingerprints.npz, written by tests/golden/make_fingerprints.py):
every step's scalars and the SHA-256 digest of the interior state.  The
oracle and the reference build call the same C library's LOG/EXP on this
machine, so the digests must be equal, not just close.

Sizes: the four-quadrant problem at 256^2 for all 200 steps, the 512x1024
disk for its first 250 steps, the 1024x2048 Euler disk for 10 steps (the
GPU tests compare the remaining steps through the same fingerprints)."""
import json
import os

import numpy as np
import pytest

import oracle
from fosite_amd import setups as S
from tests import fingerprint as F

FP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fingerprints.npz")
SCALARS = ("iter", "n_adj", "brk", "time", "dt", "dtold", "maxerrold")


def load(name):
    z = np.load(FP)
    out = {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}
    out["config"] = json.loads(bytes(out["config"]).decode())
    return out


@pytest.mark.parametrize("name, steps", [("riemann256_200", 200), ("disk512x1024_1000", 250),
                                         ("euler_disk1024x2048_20", 10)])
def test_oracle_matches_reference_fingerprints(name, steps):
    ref = load(name)
    cfg = ref["config"]
    init = S.initial(cfg)
    o = oracle.Oracle(cfg)
    o.set_initial(init["pvar"], init.get("bccsound"), init.get("fcsound"), init.get("w"))
    o.first_step()
    for n in range(1, steps + 1):
        o.step()
        cv, _, sc = o.state() if n in list(ref["dumped"]) or n == steps else (None, None, None)
        if sc is None:
            continue
        for k in SCALARS:
            assert sc[k] == ref[k][n], (name, n, k, sc[k], ref[k][n])
        if n in list(ref["dumped"]):
            fp = F.fingerprint(cv)
            assert fp["sha256"] == bytes(ref["sha256_%d" % n]).decode(), (name, n)
            assert np.array_equal(fp["l1"], ref["l1_%d" % n]), (name, n)
