"""Access to the golden fixtures written by tests/golden/make_golden.py from
the compiled reference (data only: configurations, per-step scalars, states,
mesh arrays)."""
import json
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fv_golden.npz")
SCALARS = ("iter", "n_adj", "brk", "time", "dt", "dtold", "maxerrold")

_data = None


def data():
    global _data
    if _data is None:
        _data = dict(np.load(PATH, allow_pickle=False))
    return _data


def cases():
    return sorted({k.split("/")[0] for k in data() if not k.startswith("mesh_")})


def case(name):
    d = data()
    cfg = json.loads(bytes(d[name + "/config"]).decode())
    steps = int(d[name + "/steps"][0])
    scalars = {k: d[name + "/" + k] for k in SCALARS}
    states = {int(k.split("_")[-1]): d[k] for k in d if k.startswith(name + "/cvar_")}
    return cfg, steps, scalars, states


def background_velocity(name):
    """{step: Timedisc%w} of the stored steps of a FARGO case"""
    d = data()
    return {int(k.split("_")[-1]): d[k] for k in d if k.startswith(name + "/w_")}


def bflux(name):
    """{step: (bx [V,2,NJ], by [V,2,NI], sums [4,V] of W,E,S,N)} of the
    steps whose state is stored."""
    d = data()
    out = {}
    for k in d:
        if k.startswith(name + "/bx_"):
            s = int(k.split("_")[-1])
            out[s] = (d[k], d["%s/by_%d" % (name, s)], d["%s/bsum_%d" % (name, s)])
    return out


def pointmass(name):
    """(per-step [mass, accrate, massloss, mdot] array, {step: accel [2,NJ,NI]})
    of an accretion case"""
    d = data()
    acc = {int(k.split("_")[-1]): d[k] for k in d if k.startswith(name + "/accel_")}
    return d[name + "/pointmass"], acc


def meshes():
    return sorted({k.split("/")[0][5:] for k in data() if k.startswith("mesh_")})


def mesh(name):
    d = data()
    cfg = json.loads(bytes(d["mesh_%s/config" % name]).decode())
    arrays = {k.split("/")[1]: d[k] for k in d if k.startswith("mesh_%s/" % name) and not k.endswith("/config")}
    return cfg, arrays
