"""Particle files: the oracle's quantisation pinned by hand-computed values
and invariants, and the host-side readers / writers of
taichi_elements_amd.particle_io (no GPU needed)."""
import os

import numpy as np
import pytest

from oracle import particle_io as OP


def test_quantize_known_answers():
    # x component 0 spans [0, 1]: 1 / (1 - 0) = 1, so q = f32(c * 16777215 + 0.499):
    #   0.0  -> 0.499          -> 0
    #   0.5  -> 8388607.5 rounds to 8388608 (ties to even), + 0.499 -> 8388608
    #   1.0  -> 16777215.499 rounds to 16777215
    # v component 0 spans [-2, 2]: inv = 0.25, q = ((v + 2) * 0.25) * 255 + 0.499:
    #   -2 -> 0,   0 -> 127.5 + 0.499 -> 127,   2 -> 255
    x = np.array([[0.0, 0.25], [0.5, 0.25], [1.0, 0.25]], np.float32)
    v = np.array([[-2.0, 0.0], [0.0, 0.0], [2.0, 0.0]], np.float32)
    color = np.array([0x123456, 0xFFFFFF, 0], np.int32)
    ranges, xv, rgb = OP.quantize(x, v, color)
    assert ranges[0, 0].tolist() == [0.0, 1.0] and ranges[1, 0].tolist() == [-2.0, 2.0]
    assert (xv[:, 0] >> 8).tolist() == [0, 8388608, 16777215]
    assert (xv[:, 0] & 255).tolist() == [0, 127, 255]
    # constant components: hi widened to lo + 1e-5 and every particle maps to 0
    assert ranges[0, 1, 1] == np.float32(float(np.float32(0.25)) + 1e-5)
    assert ranges[1, 1, 1] == np.float32(1e-5)
    assert (xv[:, 1] == 0).all()
    assert rgb.tolist() == [[0x12, 0x34, 0x56], [255, 255, 255], [0, 0, 0]]


def test_narrow_range_widening_is_numpy_version_independent():
    """hi = max(lo + 1e-5, hi): NumPy 1.x adds the float32 lo and the Python
    float in float64, NumPy 2 in float32.  The two sums round to the same
    float32 over whole binades (lo + f32(1e-5) has a fixed residue modulo the
    binade's spacing), checked exhaustively on [0.5, 1) and [2^-10, 2^-9)."""
    for first in (0x3F000000, 0x3A800000):
        lo = np.arange(first, first + (1 << 23), dtype=np.uint32).view(np.float32)
        f64 = (lo.astype(np.float64) + 1e-5).astype(np.float32)
        f32 = lo + np.float32(1e-5)
        assert np.array_equal(f64, f32)


def test_quantize_invariants_and_round_trip():
    rng = np.random.default_rng(1)
    n = 20000
    x = rng.uniform(-0.3, 1.7, (n, 3)).astype(np.float32)
    v = rng.normal(0, 3, (n, 3)).astype(np.float32)
    ranges, xv, _ = OP.quantize(x, v, np.zeros(n, np.int32))
    qx, qv = xv >> 8, xv & 255
    for a in range(3):
        assert qx[np.argmin(x[:, a]), a] == 0 and qv[np.argmin(v[:, a]), a] == 0
        assert abs(int(qx[np.argmax(x[:, a]), a]) - (2**24 - 1)) <= 1
        assert qv[np.argmax(v[:, a]), a] == 255
        # dequantisation error: half a level from the rounding, and for the
        # 24-bit x about a level per float32 operation (2^24 levels span the
        # float32 mantissa), so at most 3 levels; v has 8 bits and rounds to
        # within half a level
        step_x = (ranges[0, a, 1] - ranges[0, a, 0]) / (2**24 - 1)
        step_v = (ranges[1, a, 1] - ranges[1, a, 0]) / 255
        xr = qx[:, a] / (2**24 - 1) * (ranges[0, a, 1] - ranges[0, a, 0]) + ranges[0, a, 0]
        vr = qv[:, a] / 255 * (ranges[1, a, 1] - ranges[1, a, 0]) + ranges[1, a, 0]
        assert np.abs(xr - x[:, a]).max() <= 3 * step_x
        assert np.abs(vr - v[:, a]).max() <= 0.51 * step_v
    # monotone: sorting by the component sorts the codes
    o = np.argsort(x[:, 0], kind="stable")
    assert (np.diff(qx[o, 0].astype(np.int64)) >= 0).all()


def test_quantize_rejects_no_particles():
    with pytest.raises(ValueError):
        OP.quantize(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32),
                    np.zeros(0, np.int32))


def test_read_particles_inverts_the_file(tmp_path):
    from taichi_elements_amd.particle_io import ParticleIO
    rng = np.random.default_rng(2)
    for dim in (2, 3):
        n = 5000
        x = rng.uniform(0.1, 0.9, (n, dim)).astype(np.float32)
        v = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
        col = rng.integers(0, 1 << 24, n).astype(np.int32)
        ranges, xv, rgb = OP.quantize(x, v, col)
        fn = os.path.join(tmp_path, f"p{dim}.npz")
        np.savez(fn, ranges=ranges, x_and_v=xv, color=rgb)
        xr, vr, cr = (ParticleIO.read_particles_3d if dim == 3 else
                      ParticleIO.read_particles_2d)(fn)
        assert xr.dtype == np.float32 and xr.shape == (n, dim)
        for a in range(dim):
            step = (ranges[0, a, 1] - ranges[0, a, 0]) / (2**24 - 1)
            assert np.abs(xr[:, a] - x[:, a]).max() <= 3 * step
            stepv = (ranges[1, a, 1] - ranges[1, a, 0]) / 255
            assert np.abs(vr[:, a] - v[:, a]).max() <= 0.51 * stepv + 1e-7
        assert np.array_equal(cr, rgb)
        # float32 arithmetic in the reader: q / levels * width + lo per element
        q = (xv[:, 0] >> 8).astype(np.float32) / np.float32(2**24 - 1)
        assert np.array_equal(xr[:, 0], q * np.float32(ranges[0, 0, 1] - ranges[0, 0, 0])
                              + ranges[0, 0, 0])


def test_point_cloud_layout(tmp_path):
    from taichi_elements_amd.particle_io import ParticleIO, write_point_cloud
    pos = np.array([[1.0, 2.0, 3.0], [-1.0, 0.5, 0.25]], np.float32)
    col = np.array([0x0A0B0C, 0xFF0001], np.uint32)
    fn = os.path.join(tmp_path, "a.ply")
    write_point_cloud(fn, np.hstack([pos, col[:, None].view(np.float32)]))
    raw = open(fn, "rb").read()
    head, body = raw.split(b"end_header\n", 1)
    lines = head.decode().splitlines()
    assert lines[:4] == ["ply", "format binary_little_endian 1.0", "comment Created by taichi",
                         "element vertex 2"]
    assert lines[4:] == ["property float x", "property float y", "property float z",
                         "property uchar red", "property uchar green", "property uchar blue",
                         "property uchar placeholder"]
    assert len(body) == 2 * 16
    rec = np.frombuffer(body, np.float32).reshape(2, 4)
    assert np.array_equal(rec[:, :3], pos)
    assert np.frombuffer(body, np.uint8).reshape(2, 16)[0, 12:].tolist() == [0x0C, 0x0B, 0x0A, 0]
    # convert_particle_to_ply writes R, G, B, 0
    ranges, xv, rgb = OP.quantize(pos, np.zeros_like(pos), col.astype(np.int32))
    npz = os.path.join(tmp_path, "b.npz")
    np.savez(npz, ranges=ranges, x_and_v=xv, color=rgb)
    ParticleIO.convert_particle_to_ply([npz])
    body = open(npz + ".ply", "rb").read().split(b"end_header\n", 1)[1]
    assert np.frombuffer(body, np.uint8).reshape(2, 16)[0, 12:].tolist() == [0x0A, 0x0B, 0x0C, 0]


def test_blender_cache_round_trip(tmp_path):
    from taichi_elements_amd import particle_io as P
    rng = np.random.default_rng(3)
    n = 100
    data = dict(np_x=rng.random((n, 3), dtype=np.float32),
                np_v=rng.random((n, 3), dtype=np.float32),
                np_color=rng.integers(0, 1 << 24, n).astype(np.int32),
                np_material=rng.integers(0, 4, n).astype(np.int32),
                np_emitters=np.zeros(n, np.int32))
    folder = str(tmp_path)
    path = P.save_particles(folder, 7, **data)
    assert os.path.basename(path) == "particles_000007.bin"
    raw = open(path, "rb").read()
    assert raw[:8] == (1).to_bytes(4, "little") + n.to_bytes(4, "little")
    name0 = b"particles_000007_pos.bin"
    assert raw[8:12] == len(name0).to_bytes(4, "little") and raw[12:12 + len(name0)] == name0
    caches = {}
    P.read_pars(path, caches, folder)
    got = caches[folder]
    assert np.array_equal(got[P.POS], data["np_x"].ravel())
    assert np.array_equal(got[P.VEL], data["np_v"].ravel())
    assert np.array_equal(got[P.COL], data["np_color"])
    assert np.array_equal(got[P.MAT], data["np_material"])
    assert np.array_equal(got[P.EMT], data["np_emitters"])
