"""Write the reference's own binary datasets (io/fileio_binary.f90) of two
small runs, as fixtures for fosite_amd.datafile: the compiled reference
(oracle/_ref/fv_driver, built by oracle/build_ref.py) runs to the stop time
and writes its output there.  Output: tests/golden/ref_dataset_<name>.bin
(data only).

Usage (needs the reference checkout and its build):
    python oracle/build_ref.py && python tests/golden/make_datasets.py [NAME ...]
(all cases when no NAME is given)
"""
import os
import shutil
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refrun  # noqa: E402
from tests.dataset_cases import CASES  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    if not refrun.available():
        sys.exit("reference build missing: run oracle/build_ref.py first")
    from fosite_amd import setups as S
    for name in sys.argv[1:] or list(CASES):
        cfg, nsteps = CASES[name]
        tmp = tempfile.mkdtemp(prefix="fvds_")
        refrun.run(dict(cfg, binary_output=1), S.initial(cfg), nsteps, workdir=tmp)
        out = os.path.join(HERE, "ref_dataset_%s.bin" % name)
        shutil.copyfile(os.path.join(tmp, "run_data_0001.bin"), out)
        shutil.rmtree(tmp, ignore_errors=True)
        print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
