"""The 2D seeders add_ngon (mpm_solver.py:886-941) and add_texture_2d
(mpm_solver.py:943-957) through the engine.

Particle positions come from numpy's PCG64 here, not Taichi's ti.random, so
the checks are the properties the reference's seeders fix: the particle
count formula, every particle inside the polygon (half-plane test against
its vertices, independent of the rejection rule used to seed), uniform
density (the share in the inscribed circle and in each of the polygon's
sectors within 4 standard deviations of the area ratios), and for textures
one particle per texel above 0.1 at offset + (i, j) dx exactly.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


@pytest.mark.parametrize("sides,angle", [(3, 0.0), (5, 0.3), (6, -1.0), (12, 2.0)])
def test_add_ngon(sides, angle):
    from taichi_elements_amd import MPMSolver
    s = MPMSolver(res=(128, 128))
    center, radius = np.array([0.5, 0.45]), 0.3
    s.add_ngon(sides, center, radius, angle, s.material_elastic, color=0x336699,
               velocity=[1.5, -2.0])
    n = s.n_particles[None]
    want = int(math.ceil(0.5 * (radius * 128)**2 * math.sin(2 * math.pi / sides) * sides * 4))
    assert n == want
    info = s.particle_info()
    p = info["position"].astype(np.float64) - center
    # inside every edge of the polygon (vertices at angle + 2 pi k / sides)
    a = angle + 2 * np.pi * np.arange(sides + 1) / sides
    vx, vy = radius * np.cos(a), radius * np.sin(a)
    for k in range(sides):
        ex, ey = vx[k + 1] - vx[k], vy[k + 1] - vy[k]
        cross = ex * (p[:, 1] - vy[k]) - ey * (p[:, 0] - vx[k])
        assert (cross >= -1e-6).all(), k
    # uniform: inscribed circle share and sector shares
    area = 0.5 * sides * math.sin(2 * math.pi / sides)  # circumradius 1
    r_in = math.cos(math.pi / sides)

    def close(count, prob):
        assert abs(count - n * prob) <= 4 * math.sqrt(n * prob * (1 - prob)) + 1, (count, n * prob)

    close(int((np.hypot(p[:, 0], p[:, 1]) < radius * r_in).sum()), math.pi * r_in**2 / area)
    sector = np.floor(np.mod(np.arctan2(p[:, 1], p[:, 0]) - angle, 2 * np.pi) /
                      (2 * np.pi / sides)).astype(int)
    for k in range(sides):
        close(int((sector == k).sum()), 1 / sides)
    np.testing.assert_array_equal(info["velocity"], np.tile([1.5, -2.0], (n, 1)).astype(np.float32))
    assert (info["material"] == s.material_elastic).all() and (info["color"] == 0x336699).all()
    s.step(1e-3)  # the seeded particles run
    with pytest.raises(ValueError):
        MPMSolver(res=(16, 16, 16)).add_ngon(4, [0.5] * 3, 0.2, 0.0, 0)


def test_add_texture_2d():
    from taichi_elements_amd import MPMSolver
    s = MPMSolver(res=(64, 64))
    rng = np.random.default_rng(3)
    tex = rng.random((20, 30)).astype(np.float32)
    s.add_cube((0.1, 0.1), (0.05, 0.05), s.material_water, velocity=[0.5, 0.25])
    n0 = s.n_particles[None]
    s.add_texture_2d(0.3, 0.2, tex, s.material_snow, 0xABCDEF)
    info = s.particle_info()
    i, j = np.nonzero(tex > 0.1)
    assert s.n_particles[None] - n0 == len(i)
    dx = np.float32(1 / 64)
    want = np.stack([np.float32(0.3) + i.astype(np.float32) * dx,
                     np.float32(0.2) + j.astype(np.float32) * dx], 1)
    np.testing.assert_array_equal(info["position"][n0:], want)
    # the velocity of the last seeding call (the reference's source_velocity)
    np.testing.assert_array_equal(info["velocity"][n0:],
                                  np.tile([0.5, 0.25], (len(i), 1)).astype(np.float32))
    assert (info["material"][n0:] == s.material_snow).all()
    assert (info["color"][n0:] == 0xABCDEF).all()
    s.step(1e-3)


def test_particle_field_proxies():
    """mpm.x, mpm.v, mpm.F, ... read and write the particles as the
    reference's fields do (to_numpy / from_numpy / shape / [i])."""
    from taichi_elements_amd import MPMSolver
    s = MPMSolver(res=(32, 32, 32))
    s.add_cube((0.3, 0.3, 0.3), (0.1, 0.1, 0.1), s.material_snow, color=0x112233,
               velocity=[1.0, 0.0, -1.0])
    n = s.n_particles[None]
    st = s.get_state()
    assert s.x.shape == (n, 3) and s.F.shape == (n, 3, 3) and s.Jp.shape == (n, )
    np.testing.assert_array_equal(s.x.to_numpy(), st["x"])
    np.testing.assert_array_equal(s.v[5], [1.0, 0.0, -1.0])
    assert (s.material.to_numpy() == s.material_snow).all()
    assert (s.color.to_numpy() == 0x112233).all()
    F = np.tile(np.diag([1.01, 0.99, 1.0]).astype(np.float32), (n, 1, 1))
    s.F.from_numpy(F)
    np.testing.assert_array_equal(s.F.to_numpy(), F)
    s.step(1e-3)
