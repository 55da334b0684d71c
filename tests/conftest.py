"""Shared fixtures: generated coupling directories and the oracle/library handles."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


# small configurations that the oracle finishes in seconds
SMALL = {
    # a few splines, 2 kappa shells, 2 multipoles: every code path, tiny sizes
    "tiny": dict(nsplines=10, nkappa=2, nmu=0, nl=2, rbox=10.0, Z=1, intensity=10, omega=50.0, cycles=1),
    # the dipole production-test size (40 blocks x 110 radial)
    "dipole": dict(nsplines=50, nkappa=4, nmu=0, nl=1, rbox=30.0, Z=1, intensity=10, omega=50.0, cycles=1),
    # non-dipole with a mu cut: empty blocks are removed and the Hall perm is non-trivial
    "nondipole_small": dict(nsplines=20, nkappa=4, nmu=2, nl=4, rbox=20.0, Z=1, intensity=10, omega=50.0, cycles=2),
}


@pytest.fixture(scope="session")
def datadir(tmp_path_factory):
    from dirac_hydrogen_code_amd.couplings import generate
    cache = {}

    def get(name):
        if name not in cache:
            d = str(tmp_path_factory.mktemp(name))
            generate(d, **SMALL[name])
            cache[name] = d
        return cache[name]

    return get


def dt_of(cfg, ntime=200):
    c = SMALL[cfg]
    return (c["cycles"] * 2 * np.pi / c["omega"]) / ntime
