import importlib.util
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950) device")


def load_package():
    """Import the product package from disco-eb_amd/ under the module name ``discoeb_amd``."""
    if "discoeb_amd" in sys.modules:
        return sys.modules["discoeb_amd"]
    pkg_dir = os.path.join(ROOT, "disco-eb_amd")
    spec = importlib.util.spec_from_file_location("discoeb_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["discoeb_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def oracle():
    lib = os.path.join(ROOT, "oracle", "_build", "libdeb_oracle.so")
    if not os.path.exists(lib):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")], stdout=subprocess.DEVNULL)
    from oracle import oracle as O
    return O


@pytest.fixture(scope="session")
def golden():
    def _load(name):
        return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    return _load


def disco_notebook():
    """Cosmology of the reference's tests/test_cmb.py:72-90."""
    return dict(Omegam=0.3099, Omegab=0.0488911, H0=67.742, n_s=0.96822, A_s=2.1064e-09, mnu=0.06, Tcmb=2.7255, YHe=0.248,
                Neff=2.046, Nmnu=1.0, w_DE_0=-0.99, w_DE_a=0.0, cs2_DE=1.0, Omegak=0.0, k_p=0.05)
