"""Datasets in fosite's self-describing binary format (io/fileio_binary.f90).

Layout of a file (fileio_binary.f90:36-49, WriteHeader 215-247, WriteKey
298-362, WriteNode 398-575):

    header: "FOSITE" + endianness ("II" little) + version byte 0 + "%2d%2d" %
            (real size, integer size)                         13 bytes
    record: key length (int32), key (ASCII), type (int32), data length
            (int32, including the dimensions), [dimensions (int32 x 3, 4 or
            5)], data

Types are common_dict.f90:93-105 (0 a node of the key tree, 1 int, 2 real,
5 1-D real, 7 3-D real, 8 4-D real).  Arrays are stored as Fortran stores
them: first index fastest.  The writer emits, in the reference's order and
under its keys, the mesh coordinates (mesh_base.f90:835-862), the time and
the primitive variables of the Timedisc (timedisc_base.f90:566-580), the
point mass (gravity_pointmass.f90:152-195) and the configuration values that
a context holds (the reference also stores every default it filled in;
those are not repeated here).
"""
import math
import struct

import numpy as np

from . import setups

NODE, INT, REAL, REAL_1D, REAL_3D, REAL_4D = 0, 1, 2, 5, 7, 8
_NDIMS = {6: 3, 7: 3, 8: 4, 12: 5}


def _record(key, typ, payload, dims=None):
    kb = key.encode()
    dd = b"" if dims is None else struct.pack("<%di" % len(dims), *dims)
    return struct.pack("<i", len(kb)) + kb + struct.pack("<ii", typ, len(dd) + len(payload)) + dd + payload


def write(path, entries):
    """entries: (key, value) pairs in file order; value None is a node, int an
    integer, float a real, a 1-D array a REAL_ONED, an array with .ndim 3 or 4
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


def dataset_entries(cfg, time, pvar, point_mass=None):
    """The entries of one output: pvar with ghost cells ([V][NJ][NI], the
    reference's marray layout), the time, and the point mass ([mass,
    accrate, massloss, mdot]) when the configuration has one."""
    cfg = setups.full(cfg)
    e = mesh_entries(cfg)
    e += [("/timedisc", None), ("/timedisc/time", float(time))]
    for l, name in enumerate(PVAR_NAMES[cfg["physics"]]):
        e.append(("/timedisc/" + name, np.asarray(pvar)[l, 2:-2, 2:-2][None]))
    if cfg.get("gravity", 0) and point_mass is not None:
        e += [("/sources", None), ("/sources/grav", None), ("/sources/grav/pmass", None),
              ("/sources/grav/pmass/mass", float(point_mass[0])),
              ("/sources/grav/pmass/accrate", float(point_mass[1]))]
    e += config_entries(cfg)
    return e


def config_entries(cfg):
    """/config: the Dict values a context is built from, under the reference's
    keys (e.g. tests/riemann2d.f90:586-670)."""
    e = [("/config", None), ("/config/mesh", None), ("/config/mesh/meshtype", 1),
         ("/config/mesh/geometry", int(cfg["geometry"])), ("/config/mesh/inum", int(cfg["inum"])),
         ("/config/mesh/jnum", int(cfg["jnum"])), ("/config/mesh/knum", 1),
         ("/config/mesh/xmin", float(cfg["xmin"])), ("/config/mesh/xmax", float(cfg["xmax"])),
         ("/config/mesh/ymin", float(cfg["ymin"])), ("/config/mesh/ymax", float(cfg["ymax"])),
         ("/config/mesh/zmin", 0.0), ("/config/mesh/zmax", 0.0), ("/config/mesh/gparam", float(cfg["gparam"]))]
    if cfg.get("fargo", 0):
        e += [("/config/mesh/fargo", None), ("/config/mesh/fargo/method", int(cfg["fargo"]))]
    e += [("/config/physics", None), ("/config/physics/problem", int(cfg["physics"]))]
    if cfg["physics"] == setups.EULER:
        e.append(("/config/physics/gamma", float(cfg["gamma"])))
    e += [("/config/boundary", None)]
    for d, name in enumerate(("western", "eastern", "southern", "northern")):
        e.append(("/config/boundary/" + name, int(cfg["bc"][d])))
    e += [("/config/fluxes", None), ("/config/fluxes/variables", int(cfg["recvars"])),
          ("/config/fluxes/limiter", int(cfg["limiter"])), ("/config/fluxes/theta", float(cfg["theta"])),
          ("/config/timedisc", None), ("/config/timedisc/method", int(cfg["method"])),
          ("/config/timedisc/order", int(cfg["order"])), ("/config/timedisc/cfl", float(cfg["cfl"])),
          ("/config/timedisc/stoptime", float(cfg["stoptime"])), ("/config/timedisc/tol_rel", float(cfg["tol_rel"])),
          ("/config/timedisc/beta", float(cfg["beta"]))]
    return e
