"""write_particles / write_particles_ply on the GPU against the oracle.

The quantisation (ranges, packed x_and_v, colour bytes) runs in HIP kernels;
the file must be bit-identical to the oracle's restatement of
engine/particle_io.py:12-76 applied to the same particles (particle_info()).
"""
import os

import numpy as np
import pytest

from oracle import particle_io as OP
from tests.helpers import cube_particles

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


@pytest.mark.parametrize("dim", [2, 3])
def test_write_particles_matches_oracle(tmp_path, dim):
    from taichi_elements_amd import MPMSolver
    from taichi_elements_amd.particle_io import ParticleIO
    rng = np.random.default_rng(40 + dim)
    s = MPMSolver(res=(48, ) * dim)
    n = 30000
    x = cube_particles(rng, n, [0.2] * dim, [0.5] * dim, dim)
    v = rng.uniform(-2, 2, (n, dim)).astype(np.float32)
    v[:, 0] = 0.0  # a constant component: widened range
    col = rng.integers(0, 1 << 24, n).astype(np.int32)
    s.read_restart(n, x, v, rng.integers(0, 4, n).astype(np.int32), col)
    for _ in range(5):
        s.substep(s.default_dt)  # physically reordered slots, insertion order restored
    info = s.particle_info()
    fn = os.path.join(tmp_path, "frame.npz")
    s.write_particles(fn)
    got = np.load(fn)
    ranges, xv, rgb = OP.quantize(info["position"], info["velocity"], info["color"])
    assert np.array_equal(got["ranges"], ranges)
    assert np.array_equal(got["x_and_v"], xv)
    assert np.array_equal(got["color"], rgb)
    assert got["x_and_v"].dtype == np.uint32 and got["color"].dtype == np.uint8
    xr, _, _ = ParticleIO.read_particles(fn, dim)
    assert np.abs(xr - info["position"]).max() < 1e-6


def test_write_particles_edge_cases(tmp_path):
    from taichi_elements_amd import MPMSolver
    s = MPMSolver(res=(32, 32, 32))
    with pytest.raises(ValueError):
        s.write_particles(os.path.join(tmp_path, "empty.npz"))
    s.add_particles(np.array([[0.5, 0.5, 0.5]], np.float32), s.material_elastic, color=0xABCDEF)
    fn = os.path.join(tmp_path, "one.npz")
    s.write_particles(fn)
    got = np.load(fn)
    info = s.particle_info()
    ranges, xv, rgb = OP.quantize(info["position"], info["velocity"], info["color"])
    assert np.array_equal(got["ranges"], ranges) and np.array_equal(got["x_and_v"], xv)
    assert got["color"].tolist() == [[0xAB, 0xCD, 0xEF]]


def test_write_particles_ply(tmp_path):
    from taichi_elements_amd import MPMSolver
    rng = np.random.default_rng(5)
    s = MPMSolver(res=(32, 32, 32))
    s.add_cube((0.3, 0.3, 0.3), (0.2, 0.2, 0.2), s.material_water, color=0x102030)
    s.substep(s.default_dt)
    fn = os.path.join(tmp_path, "p.ply")
    s.write_particles_ply(fn)
    info = s.particle_info()
    body = open(fn, "rb").read().split(b"end_header\n", 1)[1]
    rec = np.frombuffer(body, np.float32).reshape(-1, 4)
    assert np.array_equal(rec[:, :3], info["position"])
    assert (np.frombuffer(body, np.uint8).reshape(-1, 16)[:, 12:] ==
            np.array([0x30, 0x20, 0x10, 0], np.uint8)).all()
    del rng
