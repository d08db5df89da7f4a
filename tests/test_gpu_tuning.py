"""Launch tuning and the autotuner on the GPU: autotune() leaves the scene
exactly as it found it (with use_g2p2g also the grid the next substep
reads), and a tuned solver matches an untuned one up to float32 summation
order: norm-wise relative 1e-5 after one substep, and the parity tests'
drift bound 1e-4 after eleven (the scene has random per-particle
velocities, which amplify rounding differences)."""
import numpy as np
import pytest

from tests.helpers import cube_particles, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _scene(seed=11, use_g2p2g=False):
    from taichi_elements_amd import MPMSolver
    rng = np.random.default_rng(seed)
    s = MPMSolver(res=(48, 48, 48), use_g2p2g=use_g2p2g)
    n = 40000
    x = cube_particles(rng, n, [0.2] * 3, [0.5] * 3, 3)
    v = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    s.read_restart(n, x, v, rng.integers(0, 4, n).astype(np.int32), np.zeros(n, np.int32))
    return s


@pytest.mark.parametrize("use_g2p2g", [False, True])
def test_autotune_restores_the_scene_and_keeps_results(use_g2p2g):
    s = _scene(use_g2p2g=use_g2p2g)
    for _ in range(3):
        s.substep(s.default_dt)
    before = s.get_state()
    t0, k0 = s.t, s.total_substeps
    res = s.autotune(substeps=2, chunks=(64, 256), groups=(512, 2048))
    assert len(res) == 4 and all(v > 0 for v in res.values())
    assert s.tuning() == min(res, key=res.get)
    assert set(s.last_autotune_g2p_waves) == {0, 6}
    assert s.g2p_waves() == min(s.last_autotune_g2p_waves, key=s.last_autotune_g2p_waves.get)
    after = s.get_state()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    assert s.t == t0 and s.total_substeps == k0
    ref = _scene(use_g2p2g=use_g2p2g)
    for _ in range(3):
        ref.substep(ref.default_dt)
    # one substep: only the float32 summation order differs
    s.substep(s.default_dt)
    ref.substep(ref.default_dt)
    a, b = s.get_state(), ref.get_state()
    for k in ("x", "v", "F", "Jp"):
        e = rel_err(a[k], b[k])
        print("1 substep", k, f"{e:.2e}", flush=True)
        assert e <= 1e-5, k
    # ten more: the drift bound of the parity tests
    for _ in range(10):
        s.substep(s.default_dt)
        ref.substep(ref.default_dt)
    a, b = s.get_state(), ref.get_state()
    for k in ("x", "v", "F", "Jp"):
        e = rel_err(a[k], b[k])
        print("11 substeps", k, f"{e:.2e}", flush=True)
        assert e <= 1e-4, k


def test_set_tuning_checks_its_arguments():
    s = _scene()
    with pytest.raises(ValueError):
        s.set_tuning(chunk=8)
    with pytest.raises(ValueError):
        s.set_tuning(chunk=2048)
    with pytest.raises(ValueError):
        s.set_tuning(groups=10)
    s.set_tuning(chunk=128, groups=1024)
    assert s.tuning() == (128, 1024)
    for w in (6, 0):
        s.set_tuning(g2p_waves=w)
        assert s.g2p_waves() == w and s.tuning() == (128, 1024)
    with pytest.raises(ValueError):
        s.set_tuning(g2p_waves=5)
    s.substep(s.default_dt)
    assert s.last_counts()["valid"] == 40000


@pytest.mark.parametrize("use_g2p2g", [False, True])
def test_snapshot_restore_and_its_errors(use_g2p2g):
    """mpm_snapshot: substeps after a restore continue from the saved scene
    (within the summation-order bound of one substep: the restore re-bins
    from scratch); restoring after particles were added is refused."""
    from taichi_elements_amd import _lib
    from taichi_elements_amd.mpm_solver import (SNAPSHOT_RELEASE, SNAPSHOT_RESTORE,
                                                SNAPSHOT_SAVE)
    L = _lib.lib()
    s = _scene(use_g2p2g=use_g2p2g)
    s.substep(s.default_dt)
    s.substep(s.default_dt)
    assert L.mpm_snapshot(s._h(), SNAPSHOT_SAVE) == 0
    saved = s.get_state()
    s.substep(s.default_dt)
    first = s.get_state()
    for _ in range(5):
        s.substep(s.default_dt)
    assert L.mpm_snapshot(s._h(), SNAPSHOT_RESTORE) == 0
    for k, a in s.get_state().items():
        assert np.array_equal(a, saved[k]), k
    s.substep(s.default_dt)
    again = s.get_state()
    for k in ("x", "v", "F", "Jp"):
        e = rel_err(again[k], first[k])
        print("after restore", k, f"{e:.2e}", flush=True)
        assert e <= 1e-5, k
    s.add_particles(np.full((1, 3), 0.5, np.float32), 0)
    assert L.mpm_snapshot(s._h(), SNAPSHOT_RESTORE) == -1
    assert L.mpm_snapshot(s._h(), SNAPSHOT_RELEASE) == 0
    assert L.mpm_snapshot(s._h(), SNAPSHOT_RESTORE) == -1


def test_g2p_occupancy_changes_nothing_but_time():
    """g2p at six waves per SIMD (fewer registers) runs the same arithmetic:
    ten substeps bitwise equal to the default build's."""
    a, b = _scene(seed=12), _scene(seed=12)
    b.set_tuning(g2p_waves=6)
    for _ in range(10):
        a.substep(a.default_dt)
        b.substep(b.default_dt)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
