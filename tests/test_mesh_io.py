"""PLY mesh input and point-cloud output (engine/mesh_io.py), on the CPU."""
import numpy as np
import pytest

from taichi_elements_amd import mesh_io
from tests import helpers as H


def test_binary_round_trip_with_scale_and_offset(tmp_path):
    v, f = H.icosphere(1)
    fn = str(tmp_path / "sphere.ply")
    mesh_io.write_mesh_ply(fn, v, f)
    tri = mesh_io.load_mesh(fn, scale=(0.3, 0.2, 0.1), offset=(0.5, 0.4, 0.3))
    assert tri.dtype == np.float32 and tri.shape == (len(f), 9)
    v32 = v.astype(np.float32).astype(np.float64)
    want = (v32[f] * [0.3, 0.2, 0.1] + [0.5, 0.4, 0.3]).reshape(-1, 9).astype(np.float32)
    np.testing.assert_array_equal(tri, want)
    tri1 = mesh_io.load_mesh(fn, scale=2)
    np.testing.assert_array_equal(tri1, (v32[f] * 2).reshape(-1, 9).astype(np.float32))


def test_ascii_and_big_endian(tmp_path):
    ascii_ply = ("ply\nformat ascii 1.0\ncomment unit tetrahedron\nelement vertex 4\n"
                 "property float x\nproperty float y\nproperty float z\n"
                 "property uchar red\nelement face 4\nproperty list uchar int vertex_indices\n"
                 "end_header\n0 0 0 1\n1 0 0 2\n0 1 0 3\n0 0 1 4\n3 0 2 1\n3 0 1 3\n"
                 "3 0 3 2\n3 1 2 3\n")
    fn = tmp_path / "tet.ply"
    fn.write_text(ascii_ply)
    tri = mesh_io.load_mesh(str(fn))
    assert tri.shape == (4, 9)
    assert H.mesh_volume(tri) == pytest.approx(1 / 6)
    # the same tetrahedron, binary big-endian with double vertices
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], ">f8")
    header = ("ply\nformat binary_big_endian 1.0\nelement vertex 4\nproperty double x\n"
              "property double y\nproperty double z\nelement face 4\n"
              "property list uchar uint vertex_indices\nend_header\n").encode()
    faces = b"".join(bytes([3]) + np.asarray(t, ">u4").tobytes()
                     for t in [(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)])
    fb = tmp_path / "tet_be.ply"
    fb.write_bytes(header + v.tobytes() + faces)
    np.testing.assert_array_equal(mesh_io.load_mesh(str(fb)), tri)


def test_non_triangle_faces_are_rejected(tmp_path):
    quad = ("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
            "property float z\nelement face 1\nproperty list uchar int vertex_indices\n"
            "end_header\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n4 0 1 2 3\n")
    fn = tmp_path / "quad.ply"
    fn.write_text(quad)
    with pytest.raises(AssertionError):
        mesh_io.load_mesh(str(fn))
    (tmp_path / "bad.ply").write_bytes(b"not a ply\n")
    with pytest.raises(ValueError):
        mesh_io.load_mesh(str(tmp_path / "bad.ply"))


def test_write_point_cloud(tmp_path):
    pc = np.zeros((5, 4), np.float32)
    pc[:, :3] = np.arange(15).reshape(5, 3)
    fn = tmp_path / "pc.ply"
    mesh_io.write_point_cloud(str(fn), pc)
    data = fn.read_bytes()
    head, body = data.split(b"end_header\n")
    assert b"element vertex 5" in head and b"property uchar placeholder" in head
    assert body == pc.tobytes()
