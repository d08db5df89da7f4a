"""The scenes of the reference's own test scripts (tests/test_2d.py,
tests/test_3d.py, tests/test_3d_surface_collider.py, tests/test_3d_collider.py),
run through the drop-in with the same calls, checked frame by frame against
the oracle.

The reference's scripts only run the scenes (they hold no assertions).  Here
every frame's step() is compared with the oracle run from the same state
(read from the engine at the frame's start, after the script's seeding call):
the substep count and dt of step()'s float64 rule (mpm_solver.py:752-771),
then x and v within 1e-4 (the bound of the other multi-substep tests).
Particles are seeded from numpy's PCG64, not ti.random, so the scenes match
the reference's in shape, not particle for particle.
"""
import math

import numpy as np
import pytest

from tests.helpers import oracle_cfg_for, rel_err

pytestmark = pytest.mark.gpu

DRIFT_TOL = 1e-4


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _frame(oracle, s, frame_dt, label):
    """s.step(frame_dt) against the oracle from the same state."""
    st = s.get_state()
    n_before = s.total_substeps
    s.step(frame_dt)
    substeps = int(frame_dt / s.default_dt) + 1  # mpm_solver.py:752-771
    dt = frame_dt / substeps
    left, count = frame_dt, 0
    while left > 0:
        left -= dt
        count += 1
    assert s.total_substeps - n_before == count
    if len(st["x"]) == 0:
        return
    run = oracle.Substepper(oracle_cfg_for(oracle, s), st, np.float32)
    for _ in range(count):
        run.substep(float(np.float32(dt)))
    got = s.get_state()
    for f in ("x", "v"):
        e = rel_err(got[f], run.s[f])
        assert e <= DRIFT_TOL, (label, f, e)


def test_2d_script(oracle):
    """tests/test_2d.py: res 24^2; per frame step(8e-3), then a thin water
    cube with velocity (sin(0.1 frame), 0), then particle_info()."""
    from taichi_elements_amd import MPMSolver
    mpm = MPMSolver(res=(24, 24))
    for frame in range(5):
        _frame(oracle, mpm, 8e-3, f"2d frame {frame}")
        mpm.add_cube(lower_corner=[0.3, 0.7], cube_size=[0.2, 0.01],
                     material=MPMSolver.material_water, velocity=[math.sin(frame * 0.1), 0])
        particles = mpm.particle_info()
        assert len(particles["position"]) == mpm.n_particles[None]


def test_3d_script(oracle):
    """tests/test_3d.py: res 24^3, size 10, a snow ellipsoid falling at 10
    onto an elastic box, gravity -50, five frames of step(4e-3)."""
    from taichi_elements_amd import MPMSolver
    mpm = MPMSolver(res=(24, 24, 24), size=10)
    mpm.add_ellipsoid(center=[2, 4, 3], radius=1, material=MPMSolver.material_snow,
                      velocity=[0, -10, 0])
    mpm.add_cube(lower_corner=[2, 6, 3], cube_size=[1, 1, 3], material=MPMSolver.material_elastic)
    mpm.set_gravity((0, -50, 0))
    for frame in range(5):
        _frame(oracle, mpm, 4e-3, f"3d frame {frame}")
        particles = mpm.particle_info()
        np_x = particles["position"] / 10.0
        assert np.isfinite(np_x).all()


def test_3d_surface_collider_script(oracle):
    """tests/test_3d_surface_collider.py: a tilted sticky plane, water cubes
    added every frame, gravity -20."""
    from taichi_elements_amd import MPMSolver
    mpm = MPMSolver(res=(24, 24, 24), size=1)
    mpm.set_gravity((0, -20, 0))
    mpm.add_surface_collider((0.5, 0.5, 0.5), (1.0, 1.0, 0.0))
    for frame in range(5):
        mpm.add_cube((0.2, 0.8, 0.45), (0.1, 0.03, 0.1), mpm.material_water, color=0x8888FF)
        _frame(oracle, mpm, 4e-3, f"surface frame {frame}")
        mpm.particle_info()


def test_3d_sphere_collider_script(oracle):
    """tests/test_3d_collider.py: slip, sticky and separate spheres under
    water cubes, gravity -20."""
    from taichi_elements_amd import MPMSolver
    mpm = MPMSolver(res=(24, 24, 24), size=1)
    mpm.set_gravity((0, -20, 0))
    mpm.add_sphere_collider(center=(0.25, 0.5, 0.5), radius=0.1, surface=mpm.surface_slip)
    mpm.add_sphere_collider(center=(0.5, 0.5, 0.5), radius=0.1, surface=mpm.surface_sticky)
    mpm.add_sphere_collider(center=(0.75, 0.5, 0.5), radius=0.1, surface=mpm.surface_separate)
    for frame in range(5):
        mpm.add_cube((0.2, 0.8, 0.45), (0.1, 0.03, 0.1), mpm.material_water, color=0x8888FF)
        mpm.add_cube((0.45, 0.8, 0.45), (0.1, 0.03, 0.1), mpm.material_water, color=0xFF8888)
        mpm.add_cube((0.7, 0.8, 0.45), (0.1, 0.03, 0.1), mpm.material_water, color=0xFFFFFF)
        _frame(oracle, mpm, 4e-3, f"sphere frame {frame}")
        mpm.particle_info()
