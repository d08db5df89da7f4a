"""CPU-side checks of the drop-in boundary: the C ABI library loads and
exports every entry point include/mpm_abi.h declares, and the Python host
mirrors the reference's constants and argument checks."""
import ctypes
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "mpm_abi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mpm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    syms = _declared_symbols()
    for name in ("mpm_create", "mpm_destroy", "mpm_add_particles", "mpm_substeps",
                 "mpm_particle_info", "mpm_set_gravity", "mpm_add_surface_collider"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    from taichi_elements_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmpm_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in _declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(_declared_symbols()) == bound


def test_config_struct_layout_matches_header():
    from taichi_elements_amd import _lib
    # 8 int32 fields then 8 doubles
    assert ctypes.sizeof(_lib.MpmConfig) == 8 * 4 + 8 * 8
    assert _lib.MpmConfig.dx.offset == 32


def test_unsupported_options_fail_before_touching_the_gpu():
    from taichi_elements_amd import MPMSolver
    with pytest.raises(ValueError):  # its dt would be 0 and step() would never end
        MPMSolver(res=(16, 16, 16), use_adaptive_dt=True, g2p2g_allowed_cfl=0.0)
    with pytest.raises(AssertionError):
        MPMSolver(res=(16, ))


def test_reference_constants(oracle):
    """Constants of MPMSolver.__init__ (mpm_solver.py:82-87, 200-210)."""
    k = oracle.reference_constants((128, 128, 128))
    assert k["dx"] == 1 / 128 and k["default_dt"] == 2e-2 / 128
    assert math.isclose(k["mu_0"], 416666.6666666667)
    assert math.isclose(k["lambda_0"], 277777.7777777778)
    assert math.isclose(k["alpha"], 0.5036, rel_tol=1e-4)
    assert k["p_mass"] == (1 / 128) ** 3 * 1000


def test_step_substep_count_rule():
    """step() runs int(frame_dt/default_dt)+1 substeps of dt = frame_dt/substeps,
    but its float64 while-loop can run one more (mpm_solver.py:752-771)."""
    def loop_count(frame_dt, default_dt):
        substeps = int(frame_dt / default_dt) + 1
        dt = frame_dt / substeps
        left, count = frame_dt, 0
        while left > 0:
            count += 1
            left -= dt
        return substeps, count
    extra = 0
    for frame_dt in (4e-3, 8e-3, 1 / 60, 1 / 24, 3e-3, 1e-2):
        for res in (24, 64, 128, 256):
            substeps, count = loop_count(frame_dt, 2e-2 / res)
            assert count in (substeps, substeps + 1)
            extra += count - substeps
    assert extra > 0  # the quirk is real for some of these frame times


@pytest.mark.parametrize("dim", [2, 3])
def test_engine_svd_on_host_against_lapack(dim):
    """The kernels' SVD (replacement of ti.svd) evaluated on the host through
    mpm_host_svd: reconstruction, orthogonality, det = +1, singular values
    against numpy.linalg.svd and U V^T against scipy.linalg.polar."""
    import numpy as np
    import scipy.linalg
    from taichi_elements_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmpm_hip.so not built")
    L = _lib.lib()
    rng = np.random.default_rng(20 + dim)
    for t in range(1500):
        if t % 3 == 0:
            A = np.eye(dim) + 0.05 * rng.normal(size=(dim, dim))
        elif t % 3 == 1:
            A = rng.normal(size=(dim, dim))
        else:
            A = np.diag(rng.uniform(0.5, 1.5, dim))
        A = A.astype(np.float32)
        U = np.zeros((dim, dim), np.float32)
        V = np.zeros((dim, dim), np.float32)
        s = np.zeros(dim, np.float32)
        assert L.mpm_host_svd(dim, _lib.fptr(A), _lib.fptr(U), _lib.fptr(s), _lib.fptr(V)) == 0
        A64, U, V, s = A.astype(float), U.astype(float), V.astype(float), s.astype(float)
        scale = max(1.0, np.abs(A64).max())
        assert np.abs((U * s) @ V.T - A64).max() <= 2e-5 * scale
        assert np.abs(U.T @ U - np.eye(dim)).max() <= 2e-6
        assert np.abs(V.T @ V - np.eye(dim)).max() <= 2e-6
        assert abs(np.linalg.det(U) - 1) <= 1e-5 and abs(np.linalg.det(V) - 1) <= 1e-5
        ref = np.linalg.svd(A64, compute_uv=False)
        assert np.abs(np.abs(s) - ref).max() <= 2e-5 * ref[0]
        if np.linalg.det(A64) > 1e-2:
            R, _ = scipy.linalg.polar(A64)
            assert np.abs(U @ V.T - R).max() <= 1e-4
