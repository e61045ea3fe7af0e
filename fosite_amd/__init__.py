"""fosite_amd -- fosite's unsplit finite-volume step on AMD MI355X (gfx950).

Python host for ``libfosite_fv.so`` (C ABI declared in include/fosite_fv.h).
The object API follows fosite's driver (fosite.f90): a ``Fosite`` is built
from a configuration dict (see ``fosite_amd.setups``), receives its initial
data (InitData), then runs FirstStep / Step / Run.  All array work happens on
the GPU; this module only marshals configuration and host arrays.

The library is built in-tree by ``fosite_amd.build.build()``; importing this
package never compiles anything and raises if the library is missing.
"""
import ctypes
import os

import numpy as np

from . import setups

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("FOSITE_FV_LIB") or os.path.join(HERE, "libfosite_fv.so")

__all__ = ["Fosite", "Batch", "Server", "device_bytes", "FvConfig", "mesh_array", "setups", "load_library", "LIBPATH",
           "restart"]


class FvConfig(ctypes.Structure):
    """Mirror of ``fv_config`` (include/fosite_fv.h)."""
    _fields_ = [("physics", ctypes.c_int), ("gamma", ctypes.c_double),
                ("geometry", ctypes.c_int), ("inum", ctypes.c_int), ("jnum", ctypes.c_int),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
                ("gparam", ctypes.c_double), ("bc", ctypes.c_int * 4),
                ("limiter", ctypes.c_int), ("theta", ctypes.c_double),
                ("recvars", ctypes.c_int), ("method", ctypes.c_int), ("order", ctypes.c_int),
                ("cfl", ctypes.c_double), ("tol_rel", ctypes.c_double),
                ("tol_abs", ctypes.c_double * 4), ("stoptime", ctypes.c_double),
                ("dtlimit", ctypes.c_double), ("beta", ctypes.c_double),
                ("fargo", ctypes.c_int), ("gravity", ctypes.c_int),
                ("mass", ctypes.c_double), ("GN", ctypes.c_double), ("dtype", ctypes.c_int),
                ("outbound", ctypes.c_int), ("acclimit", ctypes.c_double),
                ("pmin", ctypes.c_double), ("rhomin", ctypes.c_double),
                ("count", ctypes.c_int), ("starttime", ctypes.c_double),
                ("omega", ctypes.c_double), ("rotframe", ctypes.c_int),
                ("vismodel", ctypes.c_int), ("dynconst", ctypes.c_double), ("bulkconst", ctypes.c_double),
                ("cvis", ctypes.c_double)]

    @classmethod
    def from_dict(cls, cfg):
        cfg = setups.full(cfg)
        c = cls()
        for name, _ in cls._fields_:
            if name in ("bc", "tol_abs"):
                continue
            if name == "dtype":
                c.dtype = {"fp64": 0, 0: 0, "fp32": 1, 1: 1}.get(cfg.get("dtype", 0), -1)
                continue
            setattr(c, name, cfg[name])
        for d in range(4):
            c.bc[d] = cfg["bc"][d]
        ta = list(cfg["tol_abs"]) + [0.001] * 4
        for l in range(4):
            c.tol_abs[l] = ta[l]
        return c


class FvScalars(ctypes.Structure):
    _fields_ = [("iter", ctypes.c_int), ("n_adj", ctypes.c_int), ("brk", ctypes.c_int),
                ("dtcause", ctypes.c_int), ("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("dtold", ctypes.c_double), ("maxerrold", ctypes.c_double)]


# every entry point of include/fosite_fv.h
EXPORTS = ["fv_mesh_array", "fv_create", "fv_destroy", "fv_dims", "fv_set_sound_speeds",
           "fv_set_background_velocity", "fv_set_pvar", "fv_set_cvar", "fv_first_step", "fv_step",
           "fv_run", "fv_get_state", "fv_get_scalars", "fv_get_boundary_flux", "fv_get_boundary_flux_total",
           "fv_get_background_velocity", "fv_get_point_mass",
           "fv_last_run_ms", "fv_last_error", "fv_profile",
           "fv_kernel_stats", "fv_batch_create", "fv_batch_add", "fv_batch_remove", "fv_batch_run",
           "fv_batch_get_info", "fv_batch_destroy", "fv_batch_last_error", "fv_device_bytes", "fv_build_id",
           "fv_server_create", "fv_server_submit", "fv_server_step", "fv_server_poll", "fv_server_result",
           "fv_server_get_stats", "fv_server_destroy", "fv_server_last_error",
           "fv_first_rhs", "fv_integration_step", "fv_calc_timestep", "fv_add_background_velocity",
           "fv_step_noalloc"]


class FvServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_double), ("round_steps", ctypes.c_int), ("autotune", ctypes.c_int),
                ("backfill", ctypes.c_int)]


class FvRequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("steps", ctypes.c_long), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double), ("device_bytes", ctypes.c_double)]


class FvServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int), ("completed", ctypes.c_int), ("failed", ctypes.c_int),
                ("running", ctypes.c_int), ("queued", ctypes.c_int), ("peak_running", ctypes.c_int),
                ("device_bytes", ctypes.c_double), ("peak_device_bytes", ctypes.c_double),
                ("rounds", ctypes.c_long), ("busy_ms", ctypes.c_double), ("clock", ctypes.c_double)]


class FvBatchOpts(ctypes.Structure):
    """Mirror of ``fv_batch_opts``."""
    _fields_ = [("autotune", ctypes.c_int), ("tiling", ctypes.c_int), ("max_graph_steps", ctypes.c_int),
                ("reuse_tuning", ctypes.c_int)]


class FvBatchInfo(ctypes.Structure):
    """Mirror of ``fv_batch_info``."""
    _fields_ = [("members", ctypes.c_int), ("tiling", ctypes.c_int), ("tuned", ctypes.c_int),
                ("tune_ms", ctypes.c_double * 4), ("device_steps", ctypes.c_long),
                ("host_steps", ctypes.c_long), ("halts", ctypes.c_long), ("rounds", ctypes.c_long),
                ("last_run_ms", ctypes.c_double)]

_lib = None


def load_library():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError("libfosite_fv.so not built: run fosite_amd.build.build() "
                           "(or __graft_entry__.build()) first")
    L = ctypes.CDLL(LIBPATH)
    P = ctypes.POINTER(ctypes.c_double)
    V = ctypes.c_void_p
    L.fv_mesh_array.argtypes = [ctypes.POINTER(FvConfig), ctypes.c_int, P]
    L.fv_create.argtypes = [ctypes.POINTER(V), ctypes.POINTER(FvConfig)]
    L.fv_destroy.argtypes = [V]
    L.fv_dims.argtypes = [V, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_int)]
    L.fv_set_sound_speeds.argtypes = [V, P, P]
    L.fv_set_background_velocity.argtypes = [V, P]
    L.fv_set_pvar.argtypes = [V, P]
    L.fv_set_cvar.argtypes = [V, P]
    L.fv_first_step.argtypes = [V, ctypes.POINTER(ctypes.c_int)]
    L.fv_step.argtypes = [V, ctypes.POINTER(ctypes.c_int)]
    L.fv_step_noalloc.argtypes = [V, ctypes.POINTER(ctypes.c_int)]
    L.fv_run.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.fv_first_rhs.argtypes = [V, ctypes.POINTER(ctypes.c_int)]
    L.fv_integration_step.argtypes = [V, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.fv_calc_timestep.argtypes = [V, P, ctypes.POINTER(ctypes.c_int)]
    L.fv_add_background_velocity.argtypes = [V]
    L.fv_get_state.argtypes = [V, P, P]
    L.fv_get_scalars.argtypes = [V, ctypes.POINTER(FvScalars)]
    L.fv_get_boundary_flux.argtypes = [V, P, P]
    L.fv_get_background_velocity.argtypes = [V, P]
    L.fv_get_boundary_flux_total.argtypes = [V, ctypes.c_int, P]
    L.fv_get_point_mass.argtypes = [V, P]
    L.fv_last_run_ms.argtypes = [V]
    L.fv_last_run_ms.restype = ctypes.c_double
    L.fv_last_error.argtypes = [V]
    L.fv_last_error.restype = ctypes.c_char_p
    L.fv_profile.argtypes = [V, ctypes.c_int]
    L.fv_kernel_stats.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.fv_batch_create.argtypes = [ctypes.POINTER(V), ctypes.POINTER(FvBatchOpts)]
    L.fv_batch_add.argtypes = [V, V]
    L.fv_batch_remove.argtypes = [V, V]
    L.fv_batch_run.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.fv_batch_get_info.argtypes = [V, ctypes.POINTER(FvBatchInfo)]
    L.fv_batch_destroy.argtypes = [V]
    L.fv_batch_last_error.argtypes = [V]
    L.fv_batch_last_error.restype = ctypes.c_char_p
    L.fv_device_bytes.argtypes = [ctypes.POINTER(FvConfig), P]
    L.fv_build_id.argtypes = []
    L.fv_build_id.restype = ctypes.c_char_p
    L.fv_server_create.argtypes = [ctypes.POINTER(V), ctypes.POINTER(FvServerOpts)]
    L.fv_server_submit.argtypes = [V, ctypes.POINTER(FvConfig), P, P, P, P, ctypes.POINTER(ctypes.c_int)]
    L.fv_server_step.argtypes = [V, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.fv_server_poll.argtypes = [V, ctypes.c_int, ctypes.POINTER(FvRequestInfo)]
    L.fv_server_result.argtypes = [V, ctypes.c_int, P, ctypes.POINTER(FvScalars)]
    L.fv_server_get_stats.argtypes = [V, ctypes.POINTER(FvServerStats)]
    L.fv_server_destroy.argtypes = [V]
    L.fv_server_last_error.argtypes = [V]
    L.fv_server_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _sized(name, a, shape):
    """a as a contiguous FP64 array of exactly shape.size values (any shape
    with that many values is the same memory layout), else ValueError."""
    x = np.ascontiguousarray(a, dtype=np.float64)
    n = int(np.prod(shape))
    if x.size != n:
        raise ValueError("%s: %d values given, %d expected (shape %s)" % (name, x.size, n, shape))
    return x


def init_arrays(V, NJ, NI, bccsound=None, fcsound=None, w=None, pvar=None):
    """The initial-data arrays of fv_set_sound_speeds, fv_set_background_velocity
    and fv_set_pvar, checked against the sizes the C ABI reads: bccsound NJ*NI,
    fcsound 4*NJ*NI (both or neither), w NI, pvar V*NJ*NI.  Returns
    (bccsound, fcsound, w, pvar), None where not given."""
    if (bccsound is None) != (fcsound is None):
        raise ValueError("bccsound and fcsound go together")
    b = None if bccsound is None else _sized("bccsound", bccsound, (NJ, NI))
    f = None if fcsound is None else _sized("fcsound", fcsound, (4, NJ, NI))
    ww = None if w is None else _sized("w", w, (NI,))
    p = None if pvar is None else _sized("pvar", pvar, (V, NJ, NI))
    return b, f, ww, p


def mesh_array(cfg, name):
    """Host-side mesh metric array (no GPU needed), shape (NJ, NI)."""
    ids = ["dAx_w", "dAx_e", "dAy_s", "dAy_n", "dydzdV", "dzdxdV", "dxdydV", "dlx", "dly",
           "cxyx", "cyxy", "czxz", "czyz", "rdx_w", "rdx_e", "rdx_s", "rdx_n", "radius",
           "radius_w", "radius_e", "radius_s", "radius_n", "posvec_x", "posvec_y",
           "bcenter_x", "bcenter_y", "accel_x", "accel_y"]
    L = load_library()
    c = FvConfig.from_dict(cfg)
    out = np.empty((cfg["jnum"] + 4, cfg["inum"] + 4))
    rc = L.fv_mesh_array(ctypes.byref(c), ids.index(name), _ptr(out))
    if rc:
        raise ValueError("fv_mesh_array failed (%d)" % rc)
    return out


def restart(path, bccsound=None, fcsound=None, w=None, **changes):
    """A simulation restarted from a dataset in fosite's binary format, as
    examples/restart.f90:89-170 restarts one: the file's /config tree with the
    start time set to the file's time, its primitive variables as the initial
    data (ghost cells from the boundaries at the first step), configuration
    keys in ``changes`` overriding the file's.  The isothermal sound speeds
    and the FARGO type-2 background velocity are not in the file: the caller
    gives them, as restart.f90 computes its own.  Returns a Fosite ready for
    first_step()."""
    from . import datafile
    r = datafile.load_restart(path)
    cfg = dict(r["cfg"], **changes)
    sim = Fosite(cfg)
    sim.set_initial(r["pvar"], bccsound, fcsound, w)
    return sim


class Fosite:
    """One simulation on the current HIP device (fosite.f90 ``fosite`` type)."""

    def __init__(self, cfg):
        self.L = load_library()
        self.cfg = setups.full(cfg)
        c = FvConfig.from_dict(self.cfg)
        h = ctypes.c_void_p()
        rc = self.L.fv_create(ctypes.byref(h), ctypes.byref(c))
        if rc:
            raise RuntimeError("fv_create failed (%d); see stderr" % rc)
        self._h = h
        ni, nj, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.L.fv_dims(h, ctypes.byref(ni), ctypes.byref(nj), ctypes.byref(v))
        self.NI, self.NJ, self.V = ni.value, nj.value, v.value

    def _chk(self, rc, what):
        if rc:
            raise RuntimeError("%s failed (%d): %s" % (what, rc,
                               self.L.fv_last_error(self._h).decode()))

    def close(self):
        if getattr(self, "_h", None):
            self.L.fv_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_initial(self, pvar, bccsound=None, fcsound=None, w=None):
        """InitData: primitive state (V, NJ, NI), isothermal sound speeds
        at the bary-centres (NJ, NI) and the faces (4, NJ, NI), FARGO
        background velocity (NI,).  Arrays of another size raise ValueError
        before anything reaches the library."""
        b, f, ww, p = init_arrays(self.V, self.NJ, self.NI, bccsound, fcsound,
                                  w if self.cfg.get("fargo", 0) == 2 else None, pvar)
        if b is not None:
            self._chk(self.L.fv_set_sound_speeds(self._h, _ptr(b), _ptr(f)), "fv_set_sound_speeds")
        if ww is not None:
            self._chk(self.L.fv_set_background_velocity(self._h, _ptr(ww)),
                      "fv_set_background_velocity")
        self._chk(self.L.fv_set_pvar(self._h, _ptr(p)), "fv_set_pvar")

    def set_conservative(self, cvar):
        """Timedisc%cvar as given, all cells (V, NJ, NI), no conversion
        (fv_set_cvar); sound speeds and w as for set_initial, beforehand."""
        c = np.ascontiguousarray(cvar, dtype=np.float64)
        if c.shape != (self.V, self.NJ, self.NI):
            raise ValueError("cvar must have shape %s, not %s" % ((self.V, self.NJ, self.NI), c.shape))
        self._chk(self.L.fv_set_cvar(self._h, _ptr(c)), "fv_set_cvar")

    def first_step(self):
        b = ctypes.c_int()
        self._chk(self.L.fv_first_step(self._h, ctypes.byref(b)), "fv_first_step")
        return bool(b.value)

    def step(self):
        d = ctypes.c_int()
        self._chk(self.L.fv_step(self._h, ctypes.byref(d)), "fv_step")
        return bool(d.value)

    def run(self, nsteps):
        t = ctypes.c_int()
        self._chk(self.L.fv_run(self._h, int(nsteps), ctypes.byref(t)), "fv_run")
        return t.value

    def last_run_ms(self):
        return self.L.fv_last_run_ms(self._h)

    def profile(self, enable=True):
        self._chk(self.L.fv_profile(self._h, int(enable)), "fv_profile")

    KERNELS = {"rhs": 0, "rhs_end": 1, "rhs_combine": 2, "frame": 3, "control": 4, "remap": 5, "accrete": 6, "visc": 7}

    def kernel_stats(self, kernel):
        n, ms, b = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.fv_kernel_stats(self._h, self.KERNELS[kernel], ctypes.byref(n),
                                         ctypes.byref(ms), ctypes.byref(b)), "fv_kernel_stats")
        return n.value, ms.value, b.value

    def state(self):
        cv = np.empty((self.V, self.NJ, self.NI))
        pv = np.empty((self.V, self.NJ, self.NI))
        self._chk(self.L.fv_get_state(self._h, _ptr(cv), _ptr(pv)), "fv_get_state")
        return cv, pv

    def background_velocity(self):
        """Timedisc%w: FARGO background velocity per x-ring (NI values)."""
        w = np.empty(self.NI)
        self._chk(self.L.fv_get_background_velocity(self._h, _ptr(w)), "fv_get_background_velocity")
        return w

    def boundary_flux(self):
        """Fluxes%bxflux / byflux (fluxes_base.f90:142-143): bx[l, side, j]
        (V x 2 x NJ) and by[l, side, i] (V x 2 x NI), side 0 west/south,
        1 east/north."""
        bx = np.empty((self.V, 2, self.NJ))
        by = np.empty((self.V, 2, self.NI))
        self._chk(self.L.fv_get_boundary_flux(self._h, _ptr(bx), _ptr(by)), "fv_get_boundary_flux")
        return bx, by

    WEST, EAST, SOUTH, NORTH = 1, 2, 3, 4

    def boundary_flux_total(self, direction):
        """Fluxes%GetBoundaryFlux (fluxes_base.f90:220-300): VNUM sums over
        one boundary; direction WEST=1, EAST=2, SOUTH=3, NORTH=4."""
        out = np.empty(self.V)
        self._chk(self.L.fv_get_boundary_flux_total(self._h, int(direction), _ptr(out)),
                  "fv_get_boundary_flux_total")
        return out

    def write_dataset(self, path, filename=None):
        """The current state as a dataset in fosite's binary format
        (io/fileio_binary.f90; see fosite_amd.datafile): mesh coordinates,
        time, primitive variables, point mass, configuration tree.  filename
        is stored as /config/datafile/filename; by default it is path without
        its extension and step number, as fosite names its files
        (<filename>_<step 4 digits>.bin)."""
        from . import datafile
        if filename is None:
            filename = datafile.filename_of(path)
        _, pv = self.state()
        pm = self.point_mass() if self.cfg.get("gravity", 0) else None
        datafile.write(path, datafile.dataset_entries(self.cfg, self.scalars()["time"], pv, pm, filename))

    def point_mass(self):
        """The point mass (gravity_pointmass.f90:79-85): array [mass, accrate,
        massloss, mdot]; with accretion (config key ``outbound``) updated by
        every ComputeRHS as UpdateGravity_single does (383-428)."""
        out = np.empty(4)
        self._chk(self.L.fv_get_point_mass(self._h, _ptr(out)), "fv_get_point_mass")
        return out

    def scalars(self):
        s = FvScalars()
        self._chk(self.L.fv_get_scalars(self._h, ctypes.byref(s)), "fv_get_scalars")
        return {k: getattr(s, k) for k, _ in FvScalars._fields_}


class Batch:
    """Many ``Fosite`` simulations advanced together (``fv_batch_*``): every
    kernel of a step is one launch over all members, the steps of a round one
    HIP graph.  Results are bit-identical to running each member alone.

    tiling: 0/1 = 64-wide RHS tiles in row-major / XCD-chunked order, 2/3 =
    the same with 32-wide tiles; with ``autotune`` the batch times all four
    on its first device steps and keeps the fastest."""

    def __init__(self, sims=(), autotune=True, tiling=0, max_graph_steps=0, reuse_tuning=False):
        self.L = load_library()
        o = FvBatchOpts(int(bool(autotune)), int(tiling), int(max_graph_steps), int(bool(reuse_tuning)))
        h = ctypes.c_void_p()
        rc = self.L.fv_batch_create(ctypes.byref(h), ctypes.byref(o))
        if rc:
            raise RuntimeError("fv_batch_create failed (%d); see stderr" % rc)
        self._h = h
        self.members = []
        for s in sims:
            self.add(s)

    def _chk(self, rc, what):
        if rc:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, self.L.fv_batch_last_error(self._h).decode()))

    def add(self, sim):
        self._chk(self.L.fv_batch_add(self._h, sim._h), "fv_batch_add")
        self.members.append(sim)

    def remove(self, sim):
        self._chk(self.L.fv_batch_remove(self._h, sim._h), "fv_batch_remove")
        self.members.remove(sim)

    def run(self, nsteps):
        """nsteps Steps of every member; returns the steps each member took."""
        t = (ctypes.c_int * max(1, len(self.members)))()
        self._chk(self.L.fv_batch_run(self._h, int(nsteps), t), "fv_batch_run")
        return [t[k] for k in range(len(self.members))]

    def info(self):
        i = FvBatchInfo()
        self._chk(self.L.fv_batch_get_info(self._h, ctypes.byref(i)), "fv_batch_get_info")
        d = {k: getattr(i, k) for k, _ in FvBatchInfo._fields_}
        d["tune_ms"] = list(i.tune_ms)
        return d

    def last_run_ms(self):
        return self.info()["last_run_ms"]

    def close(self):
        if getattr(self, "_h", None):
            self.L.fv_batch_destroy(self._h)
            self._h = None
            self.members = []

    def __del__(self):
        self.close()


def build_id():
    """SHA-256 of the sources and flags the loaded library was built from
    (fv_build_id); fosite_amd.build.source_id() computes it for the tree."""
    return load_library().fv_build_id().decode()


def device_bytes(cfg):
    """Device memory a run with this configuration holds (fv_device_bytes)."""
    L = load_library()
    b = ctypes.c_double()
    if L.fv_device_bytes(ctypes.byref(FvConfig.from_dict(cfg)), ctypes.byref(b)):
        raise ValueError("unsupported configuration")
    return b.value


class Server:
    """Continuous batching of runs to their stop times (fv_server_*): runs are
    admitted in arrival order while they fit the device memory budget, stay
    resident, and advance in the FP32 and FP64 batches, round by round.  With
    backfill = k > 0, queued runs that fit may pass a head that does not,
    until that head has been passed over in k rounds."""

    STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "taken"}

    def __init__(self, hbm_budget_bytes, round_steps=64, autotune=True, backfill=0):
        self.L = load_library()
        o = FvServerOpts(float(hbm_budget_bytes), int(round_steps), int(bool(autotune)), int(backfill))
        h = ctypes.c_void_p()
        rc = self.L.fv_server_create(ctypes.byref(h), ctypes.byref(o))
        if rc:
            raise RuntimeError("fv_server_create failed (%d); see stderr" % rc)
        self._h = h
        self.shapes = {}

    def _chk(self, rc, what):
        if rc:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, self.L.fv_server_last_error(self._h).decode()))

    def submit(self, cfg, init):
        """Queue a run of cfg (to cfg["stoptime"]) from initial data init (as
        returned by setups.initial); returns the request id."""
        cfg = setups.full(cfg)
        c = FvConfig.from_dict(cfg)
        V, NJ, NI = 4 if cfg["physics"] == setups.EULER else 3, cfg["jnum"] + 4, cfg["inum"] + 4
        if init.get("pvar") is None:
            raise ValueError("pvar is required")
        b, f, ww, p = init_arrays(V, NJ, NI, init.get("bccsound"), init.get("fcsound"),
                                  init.get("w") if cfg.get("fargo", 0) == 2 else None, init["pvar"])
        ptrs = [None if a is None else _ptr(a) for a in (p, b, f, ww)]
        i = ctypes.c_int()
        self._chk(self.L.fv_server_submit(self._h, ctypes.byref(c), *ptrs, ctypes.byref(i)), "fv_server_submit")
        self.shapes[i.value] = (V, NJ, NI)
        return i.value

    def step(self):
        """One scheduling round; returns (running, queued)."""
        r, q = ctypes.c_int(), ctypes.c_int()
        self._chk(self.L.fv_server_step(self._h, ctypes.byref(r), ctypes.byref(q)), "fv_server_step")
        return r.value, q.value

    def poll(self, rid):
        o = FvRequestInfo()
        self._chk(self.L.fv_server_poll(self._h, int(rid), ctypes.byref(o)), "fv_server_poll")
        d = {k: getattr(o, k) for k, _ in FvRequestInfo._fields_}
        d["state"] = self.STATES[d["state"]]
        return d

    def result(self, rid):
        cv = np.empty(self.shapes[rid])
        s = FvScalars()
        self._chk(self.L.fv_server_result(self._h, int(rid), _ptr(cv), ctypes.byref(s)), "fv_server_result")
        return cv, {k: getattr(s, k) for k, _ in FvScalars._fields_}

    def stats(self):
        o = FvServerStats()
        self._chk(self.L.fv_server_get_stats(self._h, ctypes.byref(o)), "fv_server_get_stats")
        return {k: getattr(o, k) for k, _ in FvServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self.L.fv_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
