"""Datasets in fosite's self-describing binary format (io/fileio_binary.f90).

Layout of a file (fileio_binary.f90:36-49, WriteHeader 215-247, WriteKey
298-362, WriteNode 398-575):

    header: "FOSITE" + endianness ("II" little) + version byte 0 + "%2d%2d" %
            (real size, integer size)                         13 bytes
    record: key length (int32), key (ASCII), type (int32), data length
            (int32, including the dimensions), [dimensions (int32 x 3, 4 or
            5)], data

Types are common_dict.f90:93-105 (0 a node of the key tree, 1 int, 2 real,
3 character, 5 1-D real, 7 3-D real, 8 4-D real).  Arrays are stored as
Fortran stores them: first index fastest.  The writer emits, in the
reference's order and under its keys, the mesh coordinates (mesh_base.f90:835-862), the time and
the primitive variables of the Timedisc (timedisc_base.f90:566-580), the
point mass (gravity_pointmass.f90:152-195) and the configuration tree: the
values a setup passes and every default that fosite's Init routines fill in.
"""
import math
import re
import struct

import numpy as np

from . import setups

NODE, INT, REAL, CHAR, REAL_1D, REAL_3D, REAL_4D = 0, 1, 2, 3, 5, 7, 8
MAX_CHAR_LEN = 128  # common_dict.f90:92; strings are blank-padded to it
_NDIMS = {6: 3, 7: 3, 8: 4, 12: 5}


def _record(key, typ, payload, dims=None):
    kb = key.encode()
    dd = b"" if dims is None else struct.pack("<%di" % len(dims), *dims)
    return struct.pack("<i", len(kb)) + kb + struct.pack("<ii", typ, len(dd) + len(payload)) + dd + payload


def write(path, entries):
    """entries: (key, value) pairs in file order; value None is a node, int an
    integer, float a real, str a blank-padded character value of MAX_CHAR_LEN
    bytes, a 1-D array a REAL_ONED, an array with .ndim 3 or 4
    (Fortran index order, i.e. numpy shape reversed) a 3-D / 4-D real array."""
    out = [b"FOSITE" + b"II" + bytes([0]) + b"%2d%2d" % (8, 4)]
    for key, v in entries:
        if v is None:
            out.append(_record(key, NODE, b""))
        elif isinstance(v, (bool, np.bool_)):
            raise TypeError("logical values are not written")
        elif isinstance(v, (int, np.integer)):
            out.append(_record(key, INT, struct.pack("<i", int(v))))
        elif isinstance(v, (float, np.floating)):
            out.append(_record(key, REAL, struct.pack("<d", float(v))))
        elif isinstance(v, str):
            b = v.encode()
            if len(b) > MAX_CHAR_LEN:
                raise ValueError("string value of %s longer than %d bytes" % (key, MAX_CHAR_LEN))
            out.append(_record(key, CHAR, b.ljust(MAX_CHAR_LEN, b" ")))
        else:
            a = np.ascontiguousarray(v, dtype="<f8")
            if a.ndim == 1:
                out.append(_record(key, REAL_1D, a.tobytes()))
            elif a.ndim in (3, 4):
                out.append(_record(key, REAL_3D if a.ndim == 3 else REAL_4D, a.tobytes(), a.shape[::-1]))
            else:
                raise ValueError("unsupported rank for " + key)
    with open(path, "wb") as f:
        f.write(b"".join(out))


def read(path):
    """{key: (type, value)}: value None for nodes, int, float, bytes for other
    scalar types, numpy arrays (shape = Fortran dims reversed) for arrays."""
    b = open(path, "rb").read()
    if b[:6] != b"FOSITE" or b[6:8] != b"II" or int(b[9:11]) != 8 or int(b[11:13]) != 4:
        raise ValueError("not a little-endian double-precision fosite dataset")
    o, out = 13, {}
    while o < len(b):
        (kl,) = struct.unpack_from("<i", b, o)
        key = b[o + 4:o + 4 + kl].decode()
        o += 4 + kl
        typ, nb = struct.unpack_from("<ii", b, o)
        o += 8
        if typ in _NDIMS and nb > 0:
            nd = _NDIMS[typ]
            dims = struct.unpack_from("<%di" % nd, b, o)
            n = (nb - 4 * nd) // 8
            val = np.frombuffer(b, "<f8", n, o + 4 * nd).reshape(dims[::-1]).copy()
        elif typ == NODE:
            val = None
        elif typ == INT:
            (val,) = struct.unpack_from("<i", b, o)
        elif typ == REAL:
            (val,) = struct.unpack_from("<d", b, o)
        elif typ == REAL_1D:
            val = np.frombuffer(b, "<f8", nb // 8, o).copy()
        else:
            val = b[o:o + nb]
        out[key] = (typ, val)
        o += nb
    return out


def filename_of(path):
    """The /config/datafile/filename value for an output file at path: path
    without its extension and step number, fosite's files being named
    <filename>_<step, 4 digits>.bin (fileio_base.f90:803-812)."""
    return re.sub(r"(_\d{4})?\.bin$", "", path)


def _cartesian(cfg, x, y, z):
    """Convert2Cartesian_coords (geometry_logcylindrical.f90:344-356; the
    cartesian geometry is the identity)"""
    if cfg["geometry"] == setups.CARTESIAN:
        return x, y, z
    gp = cfg.get("gparam", 1.0)
    r = np.array([gp * math.exp(float(v)) for v in x.ravel()]).reshape(x.shape)
    cx = r * np.array([math.cos(float(v)) for v in y.ravel()]).reshape(y.shape)
    cy = r * np.array([math.sin(float(v)) for v in y.ravel()]).reshape(y.shape)
    return cx, cy, z


def mesh_entries(cfg):
    """/mesh: cell corners (grid_x/y/z, corner 1 of cells IMIN..IMAX+1,
    JMIN..JMAX+1) and bary centres, curvilinear and cartesian
    (mesh_base.f90:646-671, 835-862)."""
    ni, nj = cfg["inum"], cfg["jnum"]
    dx = (cfg["xmax"] - cfg["xmin"]) / ni
    dy = (cfg["ymax"] - cfg["ymin"]) / nj
    dz = 0.0  # knum = 1, zmin = zmax = 0
    i = np.arange(1, ni + 2)[None, :]
    j = np.arange(1, nj + 2)[:, None]
    xc = cfg["xmin"] + ((2 * i - 1) * 0.5) * dx
    yc = cfg["ymin"] + ((2 * j - 1) * 0.5) * dy
    zc = 0.0 + ((2 * 1 - 1) * 0.5) * dz
    gx = np.broadcast_to(xc + (-0.5 * dx), (nj + 1, ni + 1))
    gy = np.broadcast_to(yc + (-0.5 * dy), (nj + 1, ni + 1))
    gz = np.full((nj + 1, ni + 1), zc + (-0.5 * dz))
    cx, cy, cz = _cartesian(cfg, np.array(gx), np.array(gy), gz)
    bx = np.broadcast_to(xc[:, :ni], (nj, ni))
    by = np.broadcast_to(yc[:nj, :], (nj, ni))
    bz = np.full((nj, ni), zc)
    bcx, bcy, bcz = _cartesian(cfg, np.array(bx), np.array(by), bz)
    return [("/mesh", None),
            ("/mesh/grid_x", cx[None]), ("/mesh/grid_y", cy[None]), ("/mesh/grid_z", cz[None]),
            ("/mesh/bary_curv", np.stack([bx, by, bz])[:, None]),
            ("/mesh/bary_centers", np.stack([bcx, bcy, bcz])[:, None])]


PVAR_NAMES = {setups.EULER_ISOTHERM: ("density", "xvelocity", "yvelocity"),
              setups.EULER: ("density", "xvelocity", "yvelocity", "pressure")}


def dataset_entries(cfg, time, pvar, point_mass=None, filename=""):
    """The entries of one output: pvar with ghost cells ([V][NJ][NI], the
    reference's marray layout), the time, the point mass ([mass, accrate,
    massloss, mdot]: mass and accrate, and massloss when an accretion limit is
    set) when the configuration has one, and the configuration
    tree; filename is the value of /config/datafile/filename (the output
    file's name without its step number and extension)."""
    cfg = setups.full(cfg)
    e = mesh_entries(cfg)
    e += [("/timedisc", None), ("/timedisc/time", float(time))]
    for l, name in enumerate(PVAR_NAMES[cfg["physics"]]):
        e.append(("/timedisc/" + name, np.asarray(pvar)[l, 2:-2, 2:-2][None]))
    if cfg.get("gravity", 0) and point_mass is not None:
        e += [("/sources", None), ("/sources/grav", None), ("/sources/grav/pmass", None),
              ("/sources/grav/pmass/mass", float(point_mass[0])),
              ("/sources/grav/pmass/accrate", float(point_mass[1]))]
        if cfg.get("acclimit", -1.0) > 0.0:  # gravity_pointmass.f90:193-194
            e.append(("/sources/grav/pmass/massloss", float(point_mass[2])))
    e += config_entries(cfg, filename)
    return e


CVAR_NAMES = {setups.EULER_ISOTHERM: ("density", "xmomentum", "ymomentum"),
              setups.EULER: ("density", "xmomentum", "ymomentum", "energy")}
VERSION = "0.8.3"  # CMakeLists.txt:18
BINARY, LINEAR, KT, GRAVITY, POINTMASS, NEWTON = 1, 2, 1, 1, 1, 1
ROTATING_FRAME = 20  # sources_generic.f90:77
VISCOSITY = 3  # sources_generic.f90:74
SI, GEOMETRICAL = 1, 3  # physics units (physics_base.f90:554)
CHECK_ALL = -1


def _tree(prefix, pairs):
    """(key, value) pairs under prefix; a key "a/b" is preceded by its node
    "a" the first time "a" occurs, as SetAttr creates it."""
    out, seen = [], set()
    for k, v in pairs:
        parts = k.split("/")
        for d in range(1, len(parts)):
            node = "/".join(parts[:d])
            if node not in seen:
                seen.add(node)
                out.append((prefix + node, None))
        seen.add(k)
        out.append((prefix + k, v))
    return out


def config_entries(cfg, filename=""):
    """/config: the configuration Dict of a run, in the order its keys enter
    it -- first the values a setup program passes (as tests/golden's driver
    builds them, like tests/riemann2d.f90:586-670), then each default that an
    Init routine fills in with GetAttr, in call order:
      mesh       geometry_cartesian.f90:91 (dz), mesh_base.f90:324-378
                 (shearingbox, omega, Qshear, rotation_center, fargo/method),
                 mesh_base.f90:836-920 (output/*)
      physics    physics_base.f90:554-590 (units, mu, advanced_wave_speeds,
                 softening, output/*), physics_eulerisotherm.f90:176 (cs)
      fluxes     fluxes_base.f90 output/*
      timedisc   timedisc_base.f90:295-359 (starttime .. rhomin),
                 SetOutput 534-646 (output/*), timedisc_rkfehlberg.f90:121
      datafile   fileio_base.f90:420-470
      sources    gravity_pointmass.f90, sources_gravity.f90, sources_rotframe.f90,
                 sources_viscosity.f90
      version    fosite.f90:250
    The values are the ones fosite 0.8.3 writes (checked key by key against
    datasets the reference wrote)."""
    cfg = setups.full(cfg)
    isot = cfg["physics"] == setups.EULER_ISOTHERM
    V = 3 if isot else 4
    fargo = int(cfg.get("fargo", 0))
    stop = float(cfg["stoptime"])
    mesh = [("meshtype", 1), ("geometry", int(cfg["geometry"])), ("inum", int(cfg["inum"])),
            ("jnum", int(cfg["jnum"])), ("knum", 1),
            ("xmin", float(cfg["xmin"])), ("xmax", float(cfg["xmax"])),
            ("ymin", float(cfg["ymin"])), ("ymax", float(cfg["ymax"])),
            ("zmin", 0.0), ("zmax", 0.0), ("gparam", float(cfg["gparam"]))]
    if fargo:
        mesh.append(("fargo/method", fargo))
        if cfg["geometry"] == setups.CARTESIAN:  # along y (mesh_base.f90:390-404)
            mesh.append(("fargo/direction", 2))
    omega = float(cfg.get("omega", 0.0))
    if omega != 0.0:  # given by the setup
        mesh.append(("omega", omega))
    if cfg["geometry"] == setups.CARTESIAN:
        mesh.append(("dz", 1.0))
    mesh.append(("shearingbox", 0))
    if omega == 0.0:
        mesh.append(("omega", 0.0))
    mesh += [("Qshear", 1.5), ("rotation_center", np.zeros(3))]
    if not fargo:
        mesh.append(("fargo/method", 0))
    mesh += [("output/" + k, v) for k, v in (
        ("corners", 0), ("grid", 1), ("bary_curv", 1), ("bary", 1), ("rotation", 0), ("dl", 0),
        ("bh", 0), ("commutator", 0), ("volume", 0), ("dA", 0), ("radius", 0),
        ("position_vector", 0))]
    phys = [("problem", int(cfg["physics"]))]
    if not isot:
        phys.append(("gamma", float(cfg["gamma"])))
    phys += [("units", GEOMETRICAL if cfg.get("gravity", 0) else SI), ("mu", 0.029),
             ("advanced_wave_speeds", 0), ("softening", 1.0)]
    if isot:
        phys.append(("cs", 0.0))
    phys += [("output/bccsound", 0), ("output/fcsound", 0)]
    bnd = [(name, int(cfg["bc"][d])) for d, name in
           enumerate(("western", "eastern", "southern", "northern"))]
    bnd += [("bottomer", setups.NO_GRADIENTS), ("topper", setups.NO_GRADIENTS)]
    flx = [("order", LINEAR), ("fluxtype", KT), ("variables", int(cfg["recvars"])),
           ("limiter", int(cfg["limiter"])), ("theta", float(cfg["theta"]))]
    flx += [("output/" + k, 0) for k in ("pstates", "cstates", "pfluxes", "wave_speeds", "slopes")]
    td = [("method", int(cfg["method"])), ("order", int(cfg["order"])), ("cfl", float(cfg["cfl"])),
          ("stoptime", stop), ("tol_rel", float(cfg["tol_rel"])), ("beta", float(cfg["beta"])),
          ("maxiter", 100000000), ("tol_abs", np.array([float(x) for x in cfg["tol_abs"]][:V]))]
    # a given dtlimit is the setup's value; otherwise the default follows starttime
    # keys the setup gives come first (in the order a setup program sets
    # them), the defaults InitTimedisc fills in follow in its call order
    start = float(cfg.get("starttime", 0.0))
    tiny = float(np.finfo(np.float64).tiny)
    given = dict(dtlimit=float(cfg.get("dtlimit", -1.0)), pmin=float(cfg.get("pmin", -1.0)),
                 rhomin=float(cfg.get("rhomin", -1.0)), starttime=start)
    setup = {k: v for k, v in given.items() if (v != 0.0 if k == "starttime" else v > 0)}
    td += list(setup.items())
    defaults = [("starttime", 0.0), ("dtlimit", float(np.finfo(np.float64).eps * stop)), ("dtmax", 5.0),
                ("rhstype", 0), ("always_update_bccsound", 1), ("checkdata", CHECK_ALL), ("pmin", tiny),
                ("tmin", tiny), ("rhomin", tiny)]
    td += [(k, v) for k, v in defaults if k not in setup]
    pn, cn = PVAR_NAMES[cfg["physics"]], CVAR_NAMES[cfg["physics"]]
    td += [("output/" + pn[0], 1), ("output/error", 0), ("output/solution", 0), ("output/time", 1)]
    for l in range(V):
        if l:
            td.append(("output/" + pn[l], 1))
        td.append(("output/" + (cn[l] + "_cvar" if cn[l] == pn[l] else cn[l]), 0))
        if not l:
            td += [("output/" + k, 0) for k in ("geometrical_sources", "external_sources", "rhs", "fluxes")]
    td += [("output/bflux", 0), ("ShowButcherTableau", 0)]
    count = int(cfg.get("count", 1))
    dfile = [("fileformat", BINARY), ("filename", filename), ("count", count if count > 0 else 1), ("filepath", ""),
             ("filecycles", 2), ("unit", 11), ("dtwall", 3600), ("stoptime", stop),
             ("walltime", float(np.finfo(np.float64).max)), ("step", 0), ("cartcoords", 1)]
    e = [("/config", None)]
    e += _tree("/config/", [("mesh/" + k, v) for k, v in mesh] + [("physics/" + k, v) for k, v in phys]
               + [("boundary/" + k, v) for k, v in bnd] + [("fluxes/" + k, v) for k, v in flx]
               + [("timedisc/" + k, v) for k, v in td] + [("datafile/" + k, v) for k, v in dfile])
    src = []
    if cfg.get("gravity", 0):
        pm = [("gtype", POINTMASS), ("potential", NEWTON), ("mass", float(cfg["mass"])),
              ("outbound", int(cfg.get("outbound", 0))), ("acclimit", float(cfg.get("acclimit", -1.0))),
              ("output/accel", 0), ("output/position", 0), ("output/potential", 0),
              ("x", 0.0), ("y", 0.0), ("z", 0.0), ("softening", 0.0), ("switchon", -1.0)]
        grav = [("stype", GRAVITY)] + [("pmass/" + k, v) for k, v in pm]
        grav += [("update_disk_height", 0), ("energy", 1), ("output/height", 0), ("output/accel", 0)]
        src += [("grav/" + k, v) for k, v in grav]
    if cfg.get("vismodel", 0):  # sources_viscosity.f90:147-165, 271-289
        bulk = cfg.get("bulkconst")
        bulk = -(2.0 / 3.0) * float(cfg.get("dynconst", 0.1)) if bulk is None else bulk
        src += [("vis/stype", VISCOSITY), ("vis/vismodel", int(cfg["vismodel"])),
                ("vis/dynconst", float(cfg.get("dynconst", 0.1))), ("vis/bulkconst", float(bulk)),
                ("vis/cvis", float(cfg.get("cvis", 0.5))), ("vis/exponent", 0.0)]
        src += [("vis/output/" + k, 0) for k in ("stress", "dynvis", "kinvis", "bulkvis")]
    if cfg.get("rotframe", 0):  # sources_rotframe.f90:86-89
        src += [("rotframe/stype", ROTATING_FRAME),
                ("rotframe/disable_centaccel", 1 if cfg["rotframe"] == 2 else 0)]
    e += _tree("/config/", [("sources/" + k, v) for k, v in src])
    e.append(("/config/version", VERSION))
    return e


def _val(d, key, default=None):
    v = d.get(key)
    if v is None:
        return default
    typ, x = v
    if typ == CHAR:
        return x.decode().rstrip()
    return x


def load_restart(path):
    """A restart from a dataset, as examples/restart.f90:89-170 makes one:
    the configuration is the file's /config tree with /timedisc/starttime set
    to the file's /timedisc/time (restart.f90:100-106), the initial data its
    primitive variables (interior cells; the ghost cells follow from the
    boundaries at the first step) and the datafile step the number in the
    file name (restart.f90:115-127).  Returns {"cfg", "pvar" ([V][NJ][NI],
    ghosts zero), "time", "step"}.  What the file does not hold -- the
    isothermal sound speeds, the FARGO type-2 background velocity -- the
    caller supplies, as restart.f90 sets its sound speeds itself."""
    d = read(path)
    g = lambda k, dflt=None: _val(d, "/config/" + k, dflt)  # noqa: E731
    physics = int(g("physics/problem"))
    cfg = dict(physics=physics, geometry=int(g("mesh/geometry")), inum=int(g("mesh/inum")),
               jnum=int(g("mesh/jnum")), xmin=float(g("mesh/xmin")), xmax=float(g("mesh/xmax")),
               ymin=float(g("mesh/ymin")), ymax=float(g("mesh/ymax")), gparam=float(g("mesh/gparam")),
               bc=[int(g("boundary/" + k)) for k in ("western", "eastern", "southern", "northern")],
               limiter=int(g("fluxes/limiter")), theta=float(g("fluxes/theta")),
               recvars=int(g("fluxes/variables")), method=int(g("timedisc/method")),
               order=int(g("timedisc/order")), cfl=float(g("timedisc/cfl")),
               stoptime=float(g("timedisc/stoptime")), tol_rel=float(g("timedisc/tol_rel")),
               beta=float(g("timedisc/beta")), dtlimit=float(g("timedisc/dtlimit")),
               pmin=float(g("timedisc/pmin")), rhomin=float(g("timedisc/rhomin")),
               count=int(g("datafile/count", 1)), fargo=int(g("mesh/fargo/method", 0)))
    tol = g("timedisc/tol_abs")
    cfg["tol_abs"] = [float(x) for x in np.atleast_1d(tol)] + [0.001] * (4 - np.atleast_1d(tol).size)
    if physics == setups.EULER:
        cfg["gamma"] = float(g("physics/gamma"))
    if _val(d, "/config/sources/grav/pmass/mass") is not None:
        cfg.update(gravity=1, mass=float(g("sources/grav/pmass/mass")),
                   outbound=int(g("sources/grav/pmass/outbound", 0)),
                   acclimit=float(g("sources/grav/pmass/acclimit", -1.0)))
        # the restart reads pvar only (restart.f90:373-720 LoadData): the
        # point mass starts again from the configured mass
    cfg["omega"] = float(g("mesh/omega", 0.0))
    if _val(d, "/config/sources/rotframe/stype") is not None:
        cfg["rotframe"] = 2 if int(g("sources/rotframe/disable_centaccel", 0)) else 1
    if _val(d, "/config/sources/vis/stype") is not None:
        cfg.update(vismodel=int(g("sources/vis/vismodel")), dynconst=float(g("sources/vis/dynconst", 0.1)),
                   cvis=float(g("sources/vis/cvis", 0.5)))
        if _val(d, "/config/sources/vis/bulkconst") is not None:
            cfg["bulkconst"] = float(g("sources/vis/bulkconst"))
    time = float(_val(d, "/timedisc/time"))
    cfg["starttime"] = time
    V = 4 if physics == setups.EULER else 3
    NI, NJ = cfg["inum"] + 4, cfg["jnum"] + 4
    pvar = np.zeros((V, NJ, NI))
    for l, name in enumerate(PVAR_NAMES[physics]):
        pvar[l, 2:-2, 2:-2] = np.asarray(_val(d, "/timedisc/" + name)).reshape(cfg["jnum"], cfg["inum"])
    m = re.search(r"_(\d{4})\.bin$", path)
    step = int(m.group(1)) if m else 0
    return {"cfg": setups.full(cfg), "pvar": pvar, "time": time, "step": step}
