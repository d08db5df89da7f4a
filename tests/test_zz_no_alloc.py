"""The allocation-free differentiable substep, mpm_diff_substep_noalloc,
through the C ABI against mpm_diff_substep on the same inputs.

The two run the same kernels in the same order, so their output states are
compared bitwise.  Calls follow each other on the one static workspace: a
3D scene, then a smaller 2D scene (fewer particles, a smaller grid; the
workspace still holds the first call's contents), then the 3D scene again.
The device's free memory is the same before and after the calls, and a scene
whose workspace exceeds the static array is refused with MPM_EINVAL.
"""
import ctypes

import numpy as np
import pytest

from tests.helpers import cube_particles, deformed_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _solver(dim, res, n, seed):
    """A differentiable handle with n pre-deformed particles of the four
    materials; its state tensor and per-particle Young's moduli."""
    from taichi_elements_amd.diff import DiffMPMSolver, n_fields
    rng = np.random.default_rng(seed)
    x = cube_particles(rng, n, [0.3] * dim, [0.35] * dim, dim)
    mat = rng.integers(0, 4, n).astype(np.int32)
    st = deformed_state(rng, x, mat, dim)
    ds = DiffMPMSolver((res, ) * dim)
    ds.setup(mat)
    rows = [st["x"].T, st["v"].T, st["C"].reshape(n, dim * dim).T,
            st["F"].reshape(n, dim * dim).T, st["Jp"][None]]
    # (NF, n) rows, C order: the C ABI reads dense rows of n floats
    s0 = torch.tensor(np.ascontiguousarray(np.concatenate(rows, 0), np.float32),
                      device=ds.device)
    assert s0.is_contiguous()
    assert s0.shape == (n_fields(dim), n)
    E = torch.tensor(rng.uniform(5e5, 2e6, n).astype(np.float32), device=ds.device)
    return ds, s0, E


def _step_both(ds, s0, E, nu=0.2):
    """(mpm_diff_substep's state, mpm_diff_substep_noalloc's state)."""
    from taichi_elements_amd import _lib
    from taichi_elements_amd._lib import check
    from taichi_elements_amd.diff import _ptr, _stream
    L = _lib.lib()
    ref = torch.empty_like(s0)
    got = torch.full_like(s0, float("nan"))
    dt = ctypes.c_float(ds.default_dt)
    check(L.mpm_diff_substep_noalloc(ds._handle, dt, _ptr(s0), _ptr(got), _ptr(E), nu,
                                     _stream()))
    check(L.mpm_diff_sync(ds._handle, _stream()))
    check(L.mpm_diff_substep(ds._handle, dt, _ptr(s0), _ptr(ref), _ptr(E), nu, None, _stream()))
    check(L.mpm_diff_sync(ds._handle, _stream()))
    return ref.cpu().numpy(), got.cpu().numpy()


def _free_bytes():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info(0)[0]  # hipMemGetInfo of the process's runtime


def test_noalloc_substep_matches_the_allocating_one():
    print("noalloc: 3D 32^3 12000 particles, then 2D 64^2 3000, then 3D again", flush=True)
    big = _solver(3, 32, 12000, 11)
    small = _solver(2, 64, 3000, 12)
    for k, (ds, s0, E) in enumerate((big, small, big)):
        ref, got = _step_both(ds, s0, E)
        assert np.isfinite(ref).all()
        moved = np.abs(ref - s0.cpu().numpy()).max()
        print(f"call {k}: {s0.shape[1]} particles, max state change {moved:.3e}", flush=True)
        assert moved > 0
        np.testing.assert_array_equal(got, ref)


def test_noalloc_substep_allocates_nothing():
    from taichi_elements_amd import _lib
    from taichi_elements_amd._lib import check
    from taichi_elements_amd.diff import _ptr, _stream
    ds, s0, E = _solver(3, 32, 8000, 13)
    out = torch.empty_like(s0)
    L = _lib.lib()
    dt = ctypes.c_float(ds.default_dt)
    check(L.mpm_diff_substep_noalloc(ds._handle, dt, _ptr(s0), _ptr(out), _ptr(E), 0.2,
                                     _stream()))
    check(L.mpm_diff_sync(ds._handle, _stream()))
    before = _free_bytes()
    for _ in range(3):
        check(L.mpm_diff_substep_noalloc(ds._handle, dt, _ptr(s0), _ptr(out), _ptr(E), 0.2,
                                         _stream()))
    check(L.mpm_diff_sync(ds._handle, _stream()))
    after = _free_bytes()
    print(f"free device memory before {before} after {after}", flush=True)
    assert after == before


def test_noalloc_refuses_a_scene_beyond_the_workspace():
    from taichi_elements_amd import _lib
    from taichi_elements_amd._lib import MPM_EINVAL
    from taichi_elements_amd.diff import _ptr, _stream
    ds, s0, E = _solver(3, 256, 1000, 14)  # a 69^3-block window: 337 MB of grid alone
    out = torch.empty_like(s0)
    L = _lib.lib()
    rc = L.mpm_diff_substep_noalloc(ds._handle, ctypes.c_float(ds.default_dt), _ptr(s0),
                                    _ptr(out), _ptr(E), 0.2, _stream())
    msg = L.mpm_last_error().decode()
    print("refused:", rc, msg, flush=True)
    assert rc == MPM_EINVAL
    assert "static workspace" in msg
