import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def h5_values(path, dset):
    """Values of one HDF5 dataset, in file order, read with h5dump."""
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", dset, str(path)], capture_output=True, text=True, check=True).stdout
    data = re.sub(r"\(\d+(,\d+)*\):", " ", out[out.index("DATA {") + 6:])
    return np.array([float(v) for v in data.replace("}", " ").replace(",", " ").split()])


def demean(a):
    a = np.asarray(a, dtype=np.float64)
    return a - a.mean()


# Small grids used throughout: (nxm, nym, nzm, istr3, str3).  Lateral sizes
# cover radix 2, 3, 4, 5 and an odd y length.
SMALL_CASES = [
    dict(nxm=24, nym=16, nzm=20, istr3=4, str3=3.0, ylen=2.0, zlen=1.5, ray=1e5, pra=0.7),
    dict(nxm=17, nym=30, nzm=12, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e6, pra=1.0),
    dict(nxm=31, nym=15, nzm=18, istr3=4, str3=2.5, ylen=1.0, zlen=2.0, ray=2e6, pra=4.4),
]


def case_id(c):
    return f"{c['nxm']}x{c['nym']}x{c['nzm']}-s{c['istr3']}"
