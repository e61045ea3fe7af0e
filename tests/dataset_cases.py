"""Runs whose final output the reference wrote as a binary dataset
(tests/golden/ref_dataset_<name>.bin, by tests/golden/make_datasets.py):
configuration and the Step() calls that reach its stop time."""
from fosite_amd import setups as S

CASES = {
    # isothermal disk: FARGO type 2, point mass, absorbing radial boundaries
    "disk16x32": (dict(S.kepler_disk_config(ni=16, nj=32), stoptime=0.05), 3),
    # Euler, four-quadrant Riemann problem, no-gradient boundaries
    "riemann16": (dict(S.riemann2d_config(n=16), stoptime=0.02), 8),
}
