"""C ABI checks that need no GPU: the library loads, exports what include/discoeb_amd.h
declares, and its host-side layout helpers agree with the Python mirror."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, load_package


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def test_library_exports_header_symbols(pkg):
    L = pkg._lib.lib()
    header = open(os.path.join(ROOT, "include", "discoeb_amd.h")).read()
    declared = set(re.findall(r"\b(deb_[a-z_0-9]+)\s*\(", header))
    assert declared == set(pkg._lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_layout_helpers(pkg):
    L = pkg._lib.lib()
    assert L.deb_version().decode().startswith("discoeb_amd")
    for N in (256, 1024):
        size = L.deb_tables_size(N)
        assert size == 64 + 7 * 512 + (len(pkg._lib.ARRAY_SLOTS) - 7) * max(N, 512)
        assert L.deb_table_offset(pkg._lib.ARRAY_SLOTS.index("aexp"), N) == 64 + 7 * 512
    assert L.deb_nvar(12, 12, 12, 17, 8) == 192 == pkg.nvar(12, 12, 12, 17, 8)
    assert L.deb_nvar(12, 12, 12, 8, 3) == 75
    assert L.deb_nvar(24, 24, 12, 8, 3) == 99


def test_argument_errors_without_gpu(pkg):
    """Invalid descriptors are rejected on the host before any launch."""
    L = pkg._lib.lib()
    prob = pkg._lib.DebProblem(2, 11, 11, 8, 3, 0)
    cfg = pkg._lib.DebSolverCfg(1e-4, 1e-4, 0.25, 0.8, 0.0, 20.0, 0.3, 2048)
    rc = L.deb_evolve(None, 256, None, None, 4, None, 1, ctypes.byref(prob), ctypes.byref(cfg), None, None, None, None, None, None, None, None)
    assert rc < 0 and b">= 3" in L.deb_last_error()
    prob = pkg._lib.DebProblem(11, 11, 11, 8, 17, 0)
    rc = L.deb_evolve(None, 256, None, None, 4, None, 1, ctypes.byref(prob), ctypes.byref(cfg), None, None, None, None, None, None, None, None)
    assert rc < 0 and b"nqmax" in L.deb_last_error()
    assert L.deb_background(None, 1, 4, None, None) < 0
    assert L.deb_tau_out(None, 256, 1, None, 3, None, None) < 0
    # the adjoint: at most 64 output times, 256 threads per unit, a mesh stride that holds the mesh
    prob = pkg._lib.DebProblem(12, 12, 12, 17, 8, 0)
    args = lambda nout, stride, maxlen: (None, 256, 1, None, None, None, None, 4, None, None, nout, ctypes.byref(prob), None, stride,
                                         maxlen, None, None, None, None, None, None, None, None, None)
    assert L.deb_evolve_adjoint(*args(65, 10, 10)) < 0 and b"64" in L.deb_last_error()
    assert L.deb_evolve_adjoint(*args(1, 5, 10)) < 0 and b"mesh_stride" in L.deb_last_error()
    assert L.deb_evolve_adjoint(*args(1, 10, 10)) < 0 and b"NULL" in L.deb_last_error()
    # per unit: 11 checkpointed states and 10 step records (13 n + nch + the reduced dense part, rounded to 4)
    rec = 13 * 192 + 11 + (19 * 19 + 4 * 24 + 17 * 24 + 9 * 8 + 2 * 19)
    rec = (rec + 3) // 4 * 4
    assert L.deb_adjoint_workspace_doubles(4, 1, 256, 10, ctypes.byref(prob)) == 32 + 6 * L.deb_tables_size(256) + 4 * (11 * 192 + 10 * rec)
    # large-scale structure: rejected before any launch
    assert L.deb_savgol_coefficients(3, 3, None, None) < 0 and b"exceed" in L.deb_last_error()
    assert L.deb_savgol_coefficients(11, 9, None, None) < 0
    assert L.deb_power_smoothed(None, None, 100, 20, 4, 2e-9, 0.96, 0.05, 4, None, None, None) < 0 and b"odd" in L.deb_last_error()
    assert L.deb_xi_from_power(None, None, 511, 0, None, None, None, None) < 0 and b"even" in L.deb_last_error()
    assert L.deb_power_kaiser(None, None, 10, 20, None, None, 1.0, 2e-9, 0.96, 0.05, 0, 1, None, None, None) < 0


def test_k_grid_matches_reference_rules(pkg, oracle):
    k = pkg.k_grid({}, 1e-4, 1.0, 64, "log")
    assert abs(k[0] - 1e-4) / 1e-4 < 1e-14 and abs(k[-1] - 1.0) < 1e-14
    with pytest.raises(ValueError):
        pkg.k_grid({}, 1e-4, 1.0, 64, "camb")
    with pytest.raises(ValueError):
        pkg.k_grid({}, 1e-4, 1.0, 64, "cubic")
    kc = pkg.k_grid({"tau_maxvis": 280.0}, 1e-4, 1.0, 100, "camb")
    assert kc.size == 100 and (abs(kc[20] - 8.0 / 280.0) < 1e-15) and abs(kc[60] - 30.0 / 280.0) < 1e-15
    import numpy as np
    np.testing.assert_allclose(k, oracle.geomspace(1e-4, 1.0, 64), rtol=0, atol=0)
