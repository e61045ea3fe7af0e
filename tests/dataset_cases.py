"""Runs whose final output the reference wrote as a binary dataset
(tests/golden/ref_dataset_<name>.bin, by tests/golden/make_datasets.py):
configuration and the Step() calls that reach its stop time."""
from fosite_amd import setups as S

CASES = {
    # isothermal disk: FARGO type 2, point mass, absorbing radial boundaries
    "disk16x32": (dict(S.kepler_disk_config(ni=16, nj=32), stoptime=0.05), 3),
    # Euler, four-quadrant Riemann problem, no-gradient boundaries
    "riemann16": (dict(S.riemann2d_config(n=16), stoptime=0.02), 8),
    # Euler disk: Dormand-Prince, primitive reconstruction, accretion through
    # the western boundary with an accretion limit, no-gradient radial
    # boundaries, no FARGO
    "euler_disk16x32": (dict(S.euler_disk_config(ni=16, nj=32), stoptime=0.02,
                             recvars=S.PRIMITIVE, outbound=1, acclimit=0.001,
                             bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 40),
    # isothermal disk: FARGO type 1, a given dtlimit, accretion through the
    # eastern boundary
    "disk_fargo1_16x32": (dict(S.kepler_disk_config(ni=16, nj=32), stoptime=0.05, fargo=1,
                               dtlimit=1e-10, outbound=2,
                               bc=[S.NO_GRADIENTS, S.NO_GRADIENTS, S.PERIODIC, S.PERIODIC]), 20),
    # isothermal disk in a rotating frame (centrifugal and Coriolis terms)
    # with the point mass and FARGO type 2
    "disk_rotframe16x32": (dict(S.kepler_disk_config(ni=16, nj=32), stoptime=0.05, omega=0.3, rotframe=1), 3),
    # Euler Riemann problem in a rotating frame, Coriolis terms only
    # (disable_centaccel), no gravity
    "riemann_rotframe16": (dict(S.riemann2d_config(n=16), stoptime=0.02, omega=2.0, rotframe=2), 8),
    # isothermal disk with ALPHA viscosity in a rotating frame, the point
    # mass and FARGO type 2
    "disk_alpha16x32": (dict(S.kepler_disk_config(ni=16, nj=32), stoptime=0.05, omega=0.3, rotframe=1, vismodel=2,
                             dynconst=0.01), 3),
    # Euler Riemann problem with MOLECULAR viscosity and a given bulk viscosity
    "riemann_viscous16": (dict(S.riemann2d_config(n=16), stoptime=0.02, vismodel=1, dynconst=1e-3, bulkconst=5e-4),
                          8),
}
