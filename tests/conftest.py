import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def fit17():
    import numpy as np

    return np.load(os.path.join(GOLDEN, "fit17_bso4.npz"))


@pytest.fixture(scope="session")
def fit19():
    import numpy as np

    return np.load(os.path.join(GOLDEN, "fit19_bso4_3ch.npz"))


def structures(z):
    off = z["atom_off"]
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        yield str(z["names"][s]), z["pos"][a:b], z["cell"][s], z["itype"][a:b], z["B"][a:b]


def fit19_weights():
    """weighted_auto weights for Ta Ti V W (read_ml_file.F90:1642-1678, build_subdata.F90:2966-2987):
    channel 2 = Z / sqrt(mean Z^2), channel 3 = (mass^2 + rcov^2) / mean, with the
    masses taken as the single-precision literals the reference stores."""
    import numpy as np
    from mld_amd import io

    els = ["Ta", "Ti", "V", "W"]
    Z = np.array([io.element_z(e) for e in els])
    m = np.array([io.element_mass(e) for e in els])
    rc = np.array([io.element_rcov(e) for e in els])
    w1 = Z / np.sqrt((Z ** 2).mean())
    w3 = m ** 2 + rc ** 2
    return w1, w3 / w3.mean()
