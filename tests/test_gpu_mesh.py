"""Mesh voxelization and seeding on the GPU against the oracle.

mpm_voxelize must reproduce oracle/voxelizer.py's column-fill counts
bit for bit (the kernel computes in float64 with no contraction, the
fills are integer atomics), and mpm_add_mesh must seed exactly the
particles of the oracle's seeding restatement for the same seed: same
count, same float32 positions, in voxel order."""
import ctypes

import numpy as np
import pytest

from oracle import voxelizer as V
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _gpu_voxelize(tri, res, dx, super_sample=2, padding=3):
    from taichi_elements_amd import _lib
    from taichi_elements_amd.mpm_solver import check
    L = _lib.lib()
    tri = np.ascontiguousarray(tri, np.float64).reshape(-1, 9)
    r = (ctypes.c_int32 * 3)(*res)
    vres = (ctypes.c_int32 * 3)()
    check(L.mpm_voxelize(0, r, float(dx), super_sample, padding, len(tri),
                         tri.ctypes.data_as(_lib._D3), vres, None))
    out = np.zeros(tuple(vres), np.int32)
    check(L.mpm_voxelize(0, r, float(dx), super_sample, padding, len(tri),
                         tri.ctypes.data_as(_lib._D3), vres, _lib.iptr(out)))
    return out


def _meshes():
    rng = np.random.default_rng(11)
    rot = H.random_rotations(rng, 2, 3)
    # open, overlapping, some nearly vertical: signed counts
    soup = (rng.uniform(0.15, 0.85, (300, 1)).repeat(9, 1)
            + rng.uniform(-0.12, 0.12, (300, 9)))
    return {
        "box": H.triangles_of(*H.box_mesh((0.21, 0.33, 0.4), (0.77, 0.52, 0.61))),
        "sphere": H.triangles_of(*H.icosphere(3), scale=0.3, center=(0.5, 0.47, 0.52),
                                 rot=rot[0]),
        "torus": H.triangles_of(*H.torus(0.25, 0.09, 32, 16), center=(0.49, 0.52, 0.5),
                                rot=rot[1]),
        "soup": soup,
        "empty": np.zeros((0, 9)),
    }


@pytest.mark.parametrize("name", ["box", "sphere", "torus", "soup", "empty"])
@pytest.mark.parametrize("res,ss", [((24, 24, 24), 2), ((20, 33, 17), 3)])
def test_voxelize_matches_oracle(name, res, ss):
    tri = _meshes()[name]
    dx = 1.0 / res[0]
    vres, vdx = V.voxel_grid(res, dx, ss)
    got = _gpu_voxelize(tri, res, dx, ss)
    assert got.shape == vres
    np.testing.assert_array_equal(got, V.voxelize(tri, vres, vdx))


@pytest.mark.parametrize("sample_density,ss", [(8, 2), (12, 2), (4, 2), (3, 1)])
def test_add_mesh_seeds_the_oracle_particles(sample_density, ss):
    from taichi_elements_amd import MPMSolver, _lib
    from taichi_elements_amd.mpm_solver import check
    res = (32, 32, 32)
    s = MPMSolver(res=res, voxelizer_super_sample=ss)
    tri = _meshes()["sphere"]
    before = np.asarray([[0.5, 0.9, 0.5]], np.float32)
    s.add_particles(before, MPMSolver.material_water)
    vel = np.asarray([0.5, -1.0, 0.25], np.float32)
    shift = np.asarray([0.01, 0.02, -0.015], np.float32)
    added = ctypes.c_int64()
    t = np.ascontiguousarray(tri, np.float64)
    check(_lib.lib().mpm_add_mesh(s._h(), len(t), t.ctypes.data_as(_lib._D3),
                                  MPMSolver.material_elastic, 0x00FF80, sample_density,
                                  _lib.fptr(vel), _lib.fptr(shift), ss, 987654321, added))
    vres, vdx = V.voxel_grid(res, s.dx, ss)
    want = V.seed_from_voxels(V.voxelize(tri, vres, vdx), sample_density, ss, 987654321, vdx,
                              shift)
    assert added.value == len(want) > 0
    info = s.particle_info()
    assert len(info["position"]) == 1 + len(want)
    np.testing.assert_array_equal(info["position"][1:], want)
    np.testing.assert_array_equal(info["position"][:1], before)
    np.testing.assert_array_equal(info["velocity"][1:], np.broadcast_to(vel, want.shape))
    assert np.all(info["material"][1:] == MPMSolver.material_elastic)
    assert np.all(info["color"][1:] == 0x00FF80)
    st = s.get_state()
    np.testing.assert_array_equal(st["F"][1:], np.broadcast_to(np.eye(3, dtype=np.float32),
                                                               (len(want), 3, 3)))


def test_add_mesh_api_then_simulate():
    from taichi_elements_amd import MPMSolver
    s = MPMSolver(res=(32, 32, 32), seed=4)
    tri = H.triangles_of(*H.icosphere(2), scale=0.2, center=(0.5, 0.5, 0.5)).astype(np.float32)
    n = s.add_mesh(triangles=tri, material=MPMSolver.material_elastic, color=0xFFFF00,
                   velocity=(0, -1, 0), translation=(0.0, 0.1, 0.0), emmiter_id=3)
    vol = H.mesh_volume(tri)
    # about sample_density / super_sample^3 = 1 particle per voxel of (dx/2)^3
    assert abs(n * (s.dx / 2)**3 / vol - 1) < 0.1
    x = s.particle_info()["position"]
    assert np.all(np.abs(np.linalg.norm(x - [0.5, 0.6, 0.5], axis=1)) < 0.2 + s.dx)
    s.set_gravity((0, -20, 0))
    for _ in range(3):
        s.step(4e-3)
    x2 = s.particle_info()["position"]
    assert np.all(np.isfinite(x2)) and x2[:, 1].mean() < x[:, 1].mean()
    with pytest.raises(Exception):
        MPMSolver(res=(16, 16)).add_mesh(triangles=tri, material=0)
