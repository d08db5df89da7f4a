"""Pins of the voxelizer and seeding restatement (oracle/voxelizer.py)
against what the geometry fixes: closed forms for boxes, the generalised
winding number of the mesh at every voxel centre (an independent
point-in-solid test), the divergence-theorem volume, orientation reversal
and degenerate triangles.  The reference's own mesh test (tests/test_3d_mesh.py)
downloads its mesh at run time, so these synthetic meshes stand in for it."""
import numpy as np
import pytest

from oracle import voxelizer as V
from tests import helpers as H


def test_voxel_grid_rounds_each_axis_to_a_power_of_two():
    assert V.voxel_grid((24, 24, 24), 1 / 24) == ((64, 64, 64), 1 / 48)
    assert V.voxel_grid((100, 50, 33), 0.01, super_sample=3) == ((384, 192, 192), 0.01 / 3)
    assert V.voxel_grid((1, 2, 64), 1.0, super_sample=1)[0] == (1, 2, 64)


@pytest.mark.parametrize("lo,hi", [((0.3, 0.3, 0.3), (0.6, 0.6, 0.6)),
                                   ((0.21, 0.33, 0.4), (0.77, 0.52, 0.61)),
                                   ((0.1011, 0.1, 0.099), (0.2, 0.9, 0.3))])
def test_box_matches_the_voxel_centres_inside(lo, hi):
    vres, vdx = V.voxel_grid((24, 24, 24), 1 / 24)
    vox = V.voxelize(H.triangles_of(*H.box_mesh(lo, hi)), vres, vdx)
    c = (np.arange(vres[0]) + 0.5) * vdx
    ins = [(c >= lo[a]) & (c <= hi[a]) for a in range(3)]
    expect = ins[0][:, None, None] & ins[1][None, :, None] & ins[2][None, None, :]
    assert vox.min() >= 0 and vox.max() <= 1
    np.testing.assert_array_equal(vox > 0, expect)


def _centres(vres, vdx):
    g = np.stack(np.meshgrid(*[(np.arange(n) + 0.5) * vdx for n in vres], indexing="ij"), -1)
    return g.reshape(-1, 3)


@pytest.mark.parametrize("shape", ["sphere", "torus"])
def test_inside_matches_the_winding_number(shape):
    rng = np.random.default_rng(3)
    rot = H.random_rotations(rng, 1, 3)[0]
    if shape == "sphere":
        tri = H.triangles_of(*H.icosphere(2), scale=0.27, center=(0.51, 0.48, 0.5), rot=rot)
    else:
        tri = H.triangles_of(*H.torus(0.25, 0.09, 20, 10), center=(0.49, 0.52, 0.5), rot=rot)
    vres, vdx = V.voxel_grid((16, 16, 16), 1 / 16)
    vox = V.voxelize(tri, vres, vdx)
    w = H.winding_number(tri, _centres(vres, vdx)).reshape(vres)
    assert np.all(np.abs(w - np.round(w)) < 1e-6)  # no centre on the surface
    np.testing.assert_array_equal(vox, np.round(w).astype(np.int32))


def test_volume_and_reversed_orientation():
    tri = H.triangles_of(*H.icosphere(3), scale=0.3, center=(0.5, 0.5, 0.5))
    vres, vdx = V.voxel_grid((24, 24, 24), 1 / 24)
    vox = V.voxelize(tri, vres, vdx)
    vol = (vox > 0).sum() * vdx**3
    assert abs(vol / H.mesh_volume(tri) - 1) < 0.02
    flipped = tri.reshape(-1, 3, 3)[:, [0, 2, 1]].reshape(-1, 9)
    np.testing.assert_array_equal(V.voxelize(flipped, vres, vdx), -vox)


def test_vertical_outside_and_empty_inputs_fill_nothing():
    vres, vdx = V.voxel_grid((8, 8, 8), 1 / 8)
    vertical = np.array([[0.3, 0.3, 0.2, 0.6, 0.6, 0.2, 0.3, 0.3, 0.7]])
    outside = H.triangles_of(*H.box_mesh((1.5, 1.5, 1.5), (1.9, 1.9, 1.9)))
    below_pad = H.triangles_of(*H.box_mesh((0.3, 0.3, 0.0), (0.6, 0.6, 0.03)))
    for t in (vertical, outside, below_pad, np.zeros((0, 9))):
        assert not V.voxelize(t, vres, vdx).any()


def test_uniform_hash_statistics():
    u = V.uniform01(77, np.arange(200000), 5, 2)
    assert u.dtype == np.float32 and u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 5 * (1 / 12 / len(u))**0.5
    assert abs(u.var() - 1 / 12) < 0.002
    hist = np.bincount((u * 16).astype(int), minlength=16)
    chi2 = ((hist - len(u) / 16)**2 / (len(u) / 16)).sum()
    assert chi2 < 50  # 15 degrees of freedom
    # distinct streams per component, l and seed
    for other in (V.uniform01(77, np.arange(200000), 5, 3), V.uniform01(77, np.arange(200000), 6, 2),
                  V.uniform01(78, np.arange(200000), 5, 2)):
        assert abs(np.corrcoef(u, other)[0, 1]) < 0.01


def test_seeding_density_and_cells():
    vres, vdx = V.voxel_grid((24, 24, 24), 1 / 24)
    vox = V.voxelize(H.triangles_of(*H.box_mesh((0.3, 0.3, 0.3), (0.6, 0.6, 0.6))), vres, vdx)
    inside = int((vox > 0).sum())
    # sample_density 8, super_sample 2: s = 1, one particle per voxel
    pos = V.seed_from_voxels(vox, 8, 2, 5, vdx, (0, 0, 0))
    assert len(pos) == inside
    cell = np.floor(pos / np.float32(vdx)).astype(int)
    assert np.all(vox[cell[:, 0], cell[:, 1], cell[:, 2]] > 0)
    # s = 1.5: one particle always, a second with probability 1/2
    pos = V.seed_from_voxels(vox, 12, 2, 5, vdx, (0.01, -0.02, 0.03))
    assert abs(len(pos) - 1.5 * inside) < 5 * (0.25 * inside)**0.5
    # s = 0.5
    pos = V.seed_from_voxels(vox, 4, 2, 9, vdx, (0, 0, 0))
    assert abs(len(pos) - 0.5 * inside) < 5 * (0.25 * inside)**0.5
    assert len(V.seed_from_voxels(vox, 0, 2, 9, vdx, (0, 0, 0))) == 0
