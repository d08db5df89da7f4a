"""numpy restatement of the mesh voxelizer (engine/voxelizer.py).

TEST INFRASTRUCTURE: this module is the oracle for the mesh seeding path.
Only ``tests/`` may import it, as the checker; the product package never does.

``voxel_grid`` follows Voxelizer.__init__ (engine/voxelizer.py:18-38: each
axis rounded up to a power of two, times the super-sampling factor; voxel
size dx / super_sample) and ``voxelize`` follows voxelize_triangles
(engine/voxelizer.py:45-121) in float64, the reference's default precision:
every triangle is jittered by a fixed vector (1e-8 scale), its xy bounding
box is clamped to [padding, res - padding), and every column (p, q) whose
centre lies inside the xy projection (either winding) receives +1 (normal
z > 0) or -1 below the triangle's height at the column centre,
``int(-n.(c - a) / n_z * inv_dx + 0.5)`` truncated toward zero and capped at
res[1] - padding, from z = padding up (engine/voxelizer.py:41-43).  A voxel
is inside a closed, outward-oriented mesh when its count is positive.

The normal's ``normalized()`` is Taichi's (third-party, taichi==1.1.0):
restated as n / sqrt(n . n); the heights only depend on it through the
ratio n.(c - a) / n_z, where the scale cancels up to rounding.
"""
import numpy as np

JITTER = np.array([-0.057616723909439505, -0.25608986292614977, 0.06716309129743714])


def voxel_grid(res, dx, super_sample=2):
    """(voxel resolution per axis, voxel size)."""
    out = []
    for r in res:
        p = 1
        while p < r:
            p *= 2
        out.append(p * super_sample)
    return tuple(out), dx / super_sample


def _cross2d(ax, ay, bx, by):
    return ax * by - ay * bx


def _inside_ccw(px, py, a, b, c):
    """engine/voxelizer.py:10-13 on arrays of points (px, py)."""
    return ((_cross2d(a[0] - px, a[1] - py, b[0] - px, b[1] - py) >= 0)
            & (_cross2d(b[0] - px, b[1] - py, c[0] - px, c[1] - py) >= 0)
            & (_cross2d(c[0] - px, c[1] - py, a[0] - px, a[1] - py) >= 0))


def voxelize(triangles, vres, vdx, padding=3):
    """Signed column-fill counts (vres[0], vres[1], vres[2]) int32 of the
    triangles ((n, 9) float64: a, b, c).  Each triangle's columns are
    evaluated as numpy arrays; every operation is the same float64 IEEE
    operation, in the same order, as the per-column loop of
    engine/voxelizer.py:94-109."""
    tri = np.asarray(triangles, np.float64).reshape(-1, 9)
    vox = np.zeros(vres, np.int32)
    inv_dx = 1.0 / vdx
    jitter = JITTER * 1e-8
    z_index = np.arange(vres[2])
    for t in tri:
        a = t[0:3] + jitter
        b = t[3:6] + jitter
        c = t[6:9] + jitter
        lo = np.minimum(np.minimum(a, b), c)
        hi = np.maximum(np.maximum(a, b), c)
        p_min = max(padding, int(np.floor(lo[0] * inv_dx)))
        p_max = min(vres[0] - padding, int(np.floor(hi[0] * inv_dx)) + 1)
        q_min = max(padding, int(np.floor(lo[1] * inv_dx)))
        q_max = min(vres[1] - padding, int(np.floor(hi[1] * inv_dx)) + 1)
        u, w = b - a, c - a
        n = np.array([u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                      u[0] * w[1] - u[1] * w[0]])
        n = n * (1.0 / np.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]))
        if abs(n[2]) < 1e-10 or p_min >= p_max or q_min >= q_max:
            continue
        inc = 1 if n[2] > 0 else -1
        P, Q = np.meshgrid(np.arange(p_min, p_max), np.arange(q_min, q_max), indexing="ij")
        P, Q = P.reshape(-1), Q.reshape(-1)
        px, py = (P + 0.5) * vdx, (Q + 0.5) * vdx
        m = _inside_ccw(px, py, a, b, c) | _inside_ccw(px, py, a, c, b)
        P, Q, px, py = P[m], Q[m], px[m], py[m]
        dot = n[0] * (px - a[0]) + n[1] * (py - a[1]) + n[2] * (0.0 - a[2])
        h = np.trunc(-dot / n[2] * inv_dx + 0.5)  # int(): toward zero
        h = np.minimum(h, vres[1] - padding).astype(np.int64)
        keep = h > padding
        P, Q, h = P[keep], Q[keep], h[keep]
        # fill(p, q, h, inc): voxels padding <= z < h (engine/voxelizer.py:41-44)
        fill = (z_index[None, :] >= padding) & (z_index[None, :] < h[:, None])
        np.add.at(vox, (P, Q), (inc * fill).astype(np.int32))
    return vox


# --------------------------------------------------------------------------
# particle seeding (MPMSolver.seed_from_voxels, mpm_solver.py:1017-1047)
# --------------------------------------------------------------------------
_M32 = np.uint64(0xFFFFFFFF)


def _mix32(x):
    """lowbias32 integer hash (C. Wellons, hash-prospector), on uint64
    arrays holding 32-bit values."""
    x = x & _M32
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & _M32
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & _M32
    x ^= x >> np.uint64(16)
    return x


def uniform01(seed, voxel, l, comp):
    """float32 in [0, 1) on a 2^-24 grid hashed from (seed, voxel index,
    l, component): the engine's stand-in for ti.random in the seeding."""
    voxel = np.asarray(voxel, np.uint64)
    h = _mix32(np.uint64(l * 4 + comp + 0x9E3779B9) + np.zeros_like(voxel))
    h = _mix32(h ^ (voxel & _M32))
    h = _mix32(h ^ (voxel >> np.uint64(32)))
    h = _mix32(h ^ np.uint64(seed))
    return (h >> np.uint64(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def seed_from_voxels(vox, sample_density, super_sample, seed, voxel_dx, translation,
                     p_limit=2045):
    """Positions (n, 3) float32 seeded in the voxels with a positive count,
    in voxel C order then l: particle l of voxel (i, j, k) exists when
    r + l < sample_density / super_sample^3 and sits at
    ((r0, r1, r2) + (i, j, k)) * voxel_dx + translation in float32
    (mpm_solver.py:1026-1038; the reference's inside test looks at i only,
    against the bounded grid's half extent minus the padding)."""
    vres = vox.shape
    idx = np.flatnonzero(vox.reshape(-1) > 0)
    ijk = np.stack(np.unravel_index(idx, vres), 1)
    keep = ijk[:, 0] < p_limit
    idx, ijk = idx[keep], ijk[:, :][keep]
    s = np.float32(sample_density) / np.float32(super_sample**3)
    pdx = np.float32(voxel_dx)
    shift = np.asarray(translation, np.float32)
    rows, order = [], []
    for l in range(sample_density + 1):
        r3 = uniform01(seed, idx, l, 3)
        m = (r3 + np.float32(l)) < s
        if not m.any():
            continue
        pos = np.empty((int(m.sum()), 3), np.float32)
        for a in range(3):
            r = uniform01(seed, idx[m], l, a)
            pos[:, a] = (r + ijk[m, a].astype(np.float32)) * pdx + shift[a]
        rows.append(pos)
        order.append(np.stack([idx[m], np.full(int(m.sum()), l)], 1))
    if not rows:
        return np.zeros((0, 3), np.float32)
    pos = np.concatenate(rows)
    key = np.concatenate(order)
    return pos[np.lexsort((key[:, 1], key[:, 0]))]
