"""Configurations and initial data for the finite-volume workloads.

A configuration is a plain dict whose keys follow the reference's Dict keys
(mesh/physics/fluxes/timedisc/boundary/sources sub-dictionaries flattened):

  physics   16 EULER_ISOTHERM | 17 EULER        (physics_base.f90:496-498)
  geometry  1 CARTESIAN | 31 LOGCYLINDRICAL     (geometry_base.f90:345-357)
  inum, jnum, xmin, xmax, ymin, ymax, gparam    (mesh_base.f90:295-372)
  bc        [western, eastern, southern, northern] boundary ids
            1 NO_GRADIENTS, 2 PERIODIC, 13 ABSORBING (boundary_base.f90:110-130)
  limiter, theta, recvars                       (reconstruction_linear.f90:71-78,
                                                 reconstruction_base.f90:117)
  method, order, cfl, tol_rel, tol_abs, stoptime, dtlimit, beta
                                                (timedisc_base.f90:187-191, 315-352)
  fargo     0 off | 1 per-ring mean velocity, every step | 2 user-supplied
            background velocity (mesh_base.f90:372-423)
  gravity, mass, GN                             (gravity_pointmass.f90:132-301)
  pmin, rhomin, starttime                       (timedisc_base.f90:295, 353-359)
  omega     angular speed of the rotating frame  (mesh_base.f90:338)
  rotframe  rotating-frame source: 0 off, 1 on, 2 with disable_centaccel
            (sources_rotframe.f90:86-93, 212-247)
  count     datafile outputs between starttime and stoptime (fileio_base.f90:420-470)
  vismodel  viscosity source: 0 off, 1 MOLECULAR, 2 ALPHA, 3 BETA (sources_viscosity.f90:73-577)
  dynconst, bulkconst, cvis
            viscosity constants and viscous Courant number; bulkconst defaults
            to -2/3 dynconst (sources_viscosity.f90:152-165)

Initial data is returned as primitive variables on all cells (ghosts
included), the isothermal sound speeds at cell centres and faces, and the
FARGO background velocity per x-ring, i.e. exactly what a fosite setup
program assigns in its InitData (e.g. examples/mmsn.f90:258-269).
"""
import math

import numpy as np

EULER_ISOTHERM, EULER = 16, 17
CARTESIAN, LOGCYLINDRICAL = 1, 31
NO_GRADIENTS, PERIODIC, ABSORBING = 1, 2, 13
MINMOD, MONOCENT, SWEBY, SUPERBEE, OSPRE, VANLEER, NOLIMIT = 1, 2, 3, 4, 5, 7, 8
PRIMITIVE, CONSERVATIVE = 1, 2
RK_FEHLBERG, CASH_KARP, DORMAND_PRINCE, SSPRK = 2, 3, 4, 5
GNUM = 2

DEFAULTS = dict(gamma=1.4, gparam=1.0, limiter=MINMOD, theta=1.0, recvars=CONSERVATIVE,
                method=SSPRK, order=5, cfl=0.4, tol_rel=0.01, tol_abs=[0.001] * 4,
                dtlimit=-1.0, beta=0.08, fargo=0, gravity=0, mass=1.0, GN=1.0,
                # point-mass accretion off unless named: the reference's default
                # on cylindrical meshes is WEST (gravity_pointmass.f90:285-299)
                outbound=0, acclimit=-1.0,
                # CheckData limits (<= 0: TINY, timedisc_base.f90:353-359),
                # datafile count, start time (timedisc_base.f90:295)
                pmin=-1.0, rhomin=-1.0, count=1, starttime=0.0,
                # rotating frame (mesh_base.f90:338, sources_rotframe.f90)
                omega=0.0, rotframe=0,
                # viscosity (sources_viscosity.f90:147-165)
                vismodel=0, dynconst=0.1, cvis=0.5)


def full(cfg):
    """Configuration with the reference's defaults filled in."""
    out = dict(DEFAULTS)
    out.update(cfg)
    if out.get("bulkconst") is None:  # sources_viscosity.f90:158
        out["bulkconst"] = -(2.0 / 3.0) * out["dynconst"]
    return out


def coords(cfg):
    """Cell-centre curvilinear coordinates per ring/column (ghosts included),
    the reference's formula xmin + (2i-1)/2 * dx (mesh_base.f90:650-653)."""
    dx = (cfg["xmax"] - cfg["xmin"]) / cfg["inum"]
    dy = (cfg["ymax"] - cfg["ymin"]) / cfg["jnum"]
    fi = np.arange(cfg["inum"] + 2 * GNUM) - 1
    fj = np.arange(cfg["jnum"] + 2 * GNUM) - 1
    xc = cfg["xmin"] + ((2 * fi - 1) * 0.5) * dx
    yc = cfg["ymin"] + ((2 * fj - 1) * 0.5) * dy
    return xc, yc, dx, dy


def _exp(a):
    # libm exp per value: the reference's scale factors use EXP from libm
    return np.array([math.exp(float(v)) for v in a])


def splitmix64_uniform(seed, n):
    """U(0,1) samples u[k] for k = 0..n-1 from SplitMix64 keyed by (seed, k)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64)
             * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# ----------------------------------------------------------------- config 1

def riemann2d_config(n=256, steps=200):
    """2-D Euler four-quadrant Riemann problem on [0,1]^2 (Lax-Liu config 3)."""
    return full(dict(physics=EULER, gamma=1.4, geometry=CARTESIAN, inum=n, jnum=n,
                     xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0,
                     bc=[NO_GRADIENTS] * 4, limiter=VANLEER, recvars=CONSERVATIVE,
                     method=SSPRK, order=3, cfl=0.4, stoptime=1.0, steps=steps))


def riemann2d_initial(cfg):
    """Quadrant states (rho,u,v,p) split at (0.5,0.5) by cell centre."""
    xc, yc, _, _ = coords(cfg)
    X, Y = np.meshgrid(xc, yc)
    east, north = X >= 0.5, Y >= 0.5
    states = {  # (east, north): (rho, u, v, p)
        (True, True): (1.5, 0.0, 0.0, 1.5),
        (False, True): (0.5323, 1.206, 0.0, 0.3),
        (False, False): (0.138, 1.206, 1.206, 0.029),
        (True, False): (0.5323, 0.0, 1.206, 0.3),
    }
    pv = np.zeros((4,) + X.shape)
    for (e, nn), st in states.items():
        m = (east == e) & (north == nn)
        for l in range(4):
            pv[l][m] = st[l]
    return dict(pvar=pv)


# ----------------------------------------------------------------- config 2

def kepler_disk_config(ni=512, nj=1024, steps=1000, cs_ratio=0.05):
    """Isothermal Keplerian disk on a logcylindrical grid, r in [1,10]."""
    return full(dict(physics=EULER_ISOTHERM, geometry=LOGCYLINDRICAL, inum=ni, jnum=nj,
                     xmin=math.log(1.0), xmax=math.log(10.0), ymin=0.0, ymax=2 * math.pi,
                     gparam=1.0, bc=[ABSORBING, ABSORBING, PERIODIC, PERIODIC],
                     limiter=VANLEER, recvars=CONSERVATIVE, method=SSPRK, order=3,
                     cfl=0.3, stoptime=1.0e3, fargo=2, gravity=1, mass=1.0, GN=1.0,
                     steps=steps, cs_ratio=cs_ratio, seed=20261202))


def kepler_disk_initial(cfg):
    """Sigma ~ r^slope (1 + 0.01 U(-1,1)) (slope: cfg "sigma_slope", default
    -1), v_r = 0, v_phi = sqrt(GM/r),
    cs = cs_ratio * sqrt(GM/r) at centres and faces, w(i) = sqrt(GM/r_i)."""
    xc, yc, dx, _ = coords(cfg)
    gp, GM = cfg["gparam"], cfg["GN"] * cfg["mass"]
    NI, NJ = len(xc), len(yc)
    r = gp * _exp(xc)
    rw = gp * _exp(xc + (-(0.5 * dx)))
    re = gp * _exp(xc + 0.5 * dx)
    u = splitmix64_uniform(cfg.get("seed", 20261202), NI * NJ).reshape(NJ, NI)
    slope = cfg.get("sigma_slope", -1.0)
    prof = 1.0 / r if slope == -1.0 else r ** slope
    sigma = prof[None, :] * (1.0 + 0.01 * (2.0 * u - 1.0))
    vk = np.sqrt(GM / r)
    pv = np.zeros((3, NJ, NI))
    pv[0] = sigma
    pv[2] = vk[None, :]
    q = cfg.get("cs_ratio", 0.05)
    cs = np.broadcast_to((q * np.sqrt(GM / r))[None, :], (NJ, NI)).copy()
    fcs = np.empty((4, NJ, NI))
    fcs[0] = (q * np.sqrt(GM / rw))[None, :]
    fcs[1] = (q * np.sqrt(GM / re))[None, :]
    fcs[2] = cs
    fcs[3] = cs
    return dict(pvar=pv, bccsound=cs, fcsound=fcs, w=vk.copy())


# ----------------------------------------------------------------- config 3

def euler_disk_config(ni=1024, nj=2048, steps=500, h=0.05, dtype="fp64"):
    """Euler (gamma 1.4) Keplerian disk on a logcylindrical grid, r in [1,10],
    Dormand-Prince with rtol 1e-4, point-mass gravity, absorbing radial and
    periodic azimuthal boundaries, no FARGO (the precision-form workload)."""
    return full(dict(physics=EULER, gamma=1.4, geometry=LOGCYLINDRICAL, inum=ni, jnum=nj,
                     xmin=math.log(1.0), xmax=math.log(10.0), ymin=0.0, ymax=2 * math.pi,
                     gparam=1.0, bc=[ABSORBING, ABSORBING, PERIODIC, PERIODIC],
                     limiter=VANLEER, recvars=CONSERVATIVE, method=DORMAND_PRINCE, order=5,
                     tol_rel=1e-4, cfl=0.4, stoptime=1.0e3, fargo=0, gravity=1, mass=1.0, GN=1.0,
                     steps=steps, h=h, seed=20261203, dtype=dtype, problem="euler_disk"))


def standard_normal(seed, n):
    """N(0,1) samples by the Box-Muller transform of SplitMix64 uniforms
    (pair k uses uniforms 2k, 2k+1)."""
    u = splitmix64_uniform(seed, 2 * n)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def euler_disk_initial(cfg):
    """Sigma = r^-1.5 exp(0.01 N(0,1)), v_r = 0, v_phi = sqrt(GM/r),
    p = Sigma (h v_phi)^2 at the cell centres; FARGO background velocity
    w(i) = sqrt(GM/r_i)."""
    xc, _, _, _ = coords(cfg)
    GM = cfg["GN"] * cfg["mass"]
    NI, NJ = len(xc), cfg["jnum"] + 2 * GNUM
    r = cfg["gparam"] * _exp(xc)
    z = standard_normal(cfg.get("seed", 20261203), NI * NJ).reshape(NJ, NI)
    sigma = (r ** -1.5)[None, :] * np.exp(0.01 * z)
    vk = np.sqrt(GM / r)
    pv = np.zeros((4, NJ, NI))
    pv[0] = sigma
    pv[2] = vk[None, :]
    pv[3] = sigma * ((cfg.get("h", 0.05) * vk) ** 2)[None, :]
    return dict(pvar=pv, w=vk.copy())


# ----------------------------------------------------------------- config 4

def batch_disks(n=64, seed=20261204, sizes=(128, 256, 384, 512), dtype="fp32", steps=500):
    """n independent isothermal Keplerian disks of uneven sizes (the batched
    workload): radial cells drawn from sizes, azimuthal cells = 2x radial,
    sound-speed ratio U(0.03, 0.08), surface-density noise seeded per disk,
    FARGO on, SSPRK3."""
    u = splitmix64_uniform(seed, 2 * n)
    out = []
    for k in range(n):
        ni = sizes[min(int(u[2 * k] * len(sizes)), len(sizes) - 1)]
        cfg = kepler_disk_config(ni=ni, nj=2 * ni, steps=steps, cs_ratio=0.03 + 0.05 * float(u[2 * k + 1]))
        cfg.update(seed=seed + 1 + k, dtype=dtype)
        out.append(cfg)
    return out


# ----------------------------------------------------------------- config 5

def serving_requests(n=2000, rate=20.0, seed=20261205, nmin=64, nmax=1024, orbits=10.0, fp32_share=0.75,
                     cs_ratio=0.05):
    """Disk runs for the serving loop: Poisson arrivals (mean rate per second),
    radial cells uniform in [nmin, nmax] (azimuthal = 2x), FP32 with
    probability fp32_share else FP64, surface-density slope U(-1.5, -0.5),
    1% seeded noise, FARGO, SSPRK3, integrated for `orbits` orbits of the
    inner edge (r = 1).  Returns [(arrival_s, cfg), ...] in arrival order."""
    u = splitmix64_uniform(seed, 4 * n)
    t, out = 0.0, []
    for k in range(n):
        t += -math.log1p(-float(u[4 * k])) / rate
        ni = nmin + min(int(u[4 * k + 1] * (nmax - nmin + 1)), nmax - nmin)
        cfg = kepler_disk_config(ni=ni, nj=2 * ni, cs_ratio=cs_ratio)
        cfg.update(stoptime=orbits * 2.0 * math.pi, seed=seed + 1 + k,
                   dtype="fp32" if u[4 * k + 2] < fp32_share else "fp64",
                   sigma_slope=-1.5 + float(u[4 * k + 3]))
        out.append((t, cfg))
    return out


def rarefaction_config(n=32, **kw):
    """Two halves of the unit box moving apart (Euler): the centre empties
    step by step, so the minimum density and pressure fall through any
    chosen limit (the CheckData and pressure-floor workload)."""
    return full(dict(riemann2d_config(n=n), problem="rarefaction", **kw))


def rarefaction_initial(cfg, u0=2.0):
    """rho = 1, p = 0.4, u = -u0 for x < 0.5 and +u0 beyond, v = 0.1 sin(2 pi y)."""
    xc, yc, _, _ = coords(cfg)
    X, Y = np.meshgrid(xc, yc)
    pv = np.zeros((4,) + X.shape)
    pv[0] = 1.0
    pv[1] = np.where(X < 0.5, -u0, u0)
    pv[2] = 0.1 * np.sin(2 * np.pi * Y)
    pv[3] = 0.4
    return dict(pvar=pv)


def shear_box_config(n=32, physics=EULER, fargo=2, **kw):
    """A periodic Cartesian box with a shear layer: a dense band |x| < 0.25
    moving in +y through a lighter medium moving in -y, with a small
    transverse perturbation (the Kelvin-Helmholtz setup of tests/KHI2D.f90,
    here with FARGO along y and a piecewise background velocity)."""
    cfg = dict(physics=physics, gamma=1.4, geometry=CARTESIAN, inum=n, jnum=n,
               xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, bc=[PERIODIC] * 4,
               limiter=MONOCENT, theta=1.2, recvars=CONSERVATIVE, method=SSPRK, order=3,
               cfl=0.4, stoptime=1.0, fargo=fargo, problem="shear_box")
    cfg.update(kw)
    return full(cfg)


def shear_box_initial(cfg):
    """rho = 2 in the band, 1 outside; v_y = +0.5 / -0.5; v_x = 0.01 sin(2 pi y)
    sin(pi x); p = 2.5 (Euler) or cs = 1 (isothermal); w(i) = v_y of column i."""
    xc, yc, _, _ = coords(cfg)
    X, Y = np.meshgrid(xc, yc)
    band = np.abs(X) < 0.25
    V = 4 if cfg["physics"] == EULER else 3
    pv = np.zeros((V,) + X.shape)
    pv[0] = np.where(band, 2.0, 1.0)
    pv[1] = 0.01 * np.sin(2 * math.pi * Y) * np.sin(math.pi * X)
    pv[2] = np.where(band, 0.5, -0.5)
    out = dict(pvar=pv, w=np.where(np.abs(xc) < 0.25, 0.5, -0.5))
    if V == 4:
        pv[3] = 2.5
    else:
        out["bccsound"] = np.ones(X.shape)
        out["fcsound"] = np.ones((4,) + X.shape)
    return out


def initial(cfg):
    if cfg.get("problem") == "shear_box":
        return shear_box_initial(cfg)
    if cfg.get("problem") == "euler_disk":
        return euler_disk_initial(cfg)
    if cfg.get("problem") == "rarefaction":
        return rarefaction_initial(cfg)
    if cfg["physics"] == EULER:
        return riemann2d_initial(cfg)
    return kepler_disk_initial(cfg)
