"""Shared scene builders and error metrics for the parity tests."""
import numpy as np
import scipy.linalg

SAND = 3


def rel_err(a, b, floor=1e-6):
    """Norm-wise relative error ||a - b|| / max(||b||, floor)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), floor))


def random_rotations(rng, n, dim):
    Q = np.linalg.qr(rng.normal(size=(n, dim, dim)))[0]
    det = np.linalg.det(Q)
    Q[det < 0, :, 0] *= -1
    return Q


def deformed_state(rng, x, material, dim, stretch=0.05, c_scale=5.0):
    """Pre-deformed particle state so that one substep exercises every
    constitutive branch: F = R diag(1 +- stretch) Q^T, C ~ N(0, c_scale),
    snow Jp in [0.9, 1.1], sand Jp in [-0.05, 0], water Jp in [0.95, 1.05];
    elastic and stationary Jp in [0.9, 1.1] as well (a caller may set them:
    stationary particles are hardened by exp(10 (1 - Jp)), elastic ones are
    not, mpm_solver.py:515-521)."""
    n = len(x)
    material = np.asarray(material)
    R = random_rotations(rng, n, dim)
    Q = random_rotations(rng, n, dim)
    s = 1 + stretch * rng.uniform(-1, 1, size=(n, dim))
    F = np.einsum("nij,nj,nkj->nik", R, s, Q).astype(np.float32)
    C = (c_scale * rng.normal(size=(n, dim, dim))).astype(np.float32)
    Jp = np.ones(n, np.float32)
    Jp[material == 2] = rng.uniform(0.9, 1.1, (material == 2).sum())
    Jp[material == 3] = rng.uniform(-0.05, 0.0, (material == 3).sum())
    Jp[material == 0] = rng.uniform(0.95, 1.05, (material == 0).sum())
    v = rng.uniform(-1, 1, size=(n, dim)).astype(np.float32)
    for m in (1, 4):
        Jp[material == m] = rng.uniform(0.9, 1.1, (material == m).sum())
    return dict(x=np.asarray(x, np.float32), v=v, C=C, F=F, Jp=Jp,
                material=np.asarray(material, np.int32))


def cube_particles(rng, n, lo, size, dim):
    return (np.asarray(lo, np.float32)
            + rng.random((n, dim), dtype=np.float32) * np.asarray(size, np.float32))


def oracle_cfg_for(O, solver, post=None, gravity=None):
    """Oracle configuration matching a taichi_elements_amd MPMSolver."""
    dim = solver.dim
    kw = dict(size=solver.size, padding=solver.padding, water_density=solver.water_density)
    if gravity is None:
        gravity = solver.gravity[None]
    if post is None:
        post = []
        for kind, args in solver.grid_postprocess:
            if kind == "bounding_box":
                post.append({"kind": "bbox"})
            elif kind == "sphere":
                post.append({"kind": "sphere", **args})
            else:
                post.append({"kind": "plane", **args})
    lo = [-8] * dim
    n = [int(solver.res[a]) + 16 for a in range(dim)]
    cfg = O.make_cfg(tuple(solver.res), gravity=gravity, post=post, grid_lo=lo, grid_n=n, **kw)
    for name in ("dx", "inv_dx", "p_vol", "p_mass", "mu_0", "lambda_0", "alpha"):
        setattr(cfg, name, getattr(solver, name))
    cfg.quant = 1 if getattr(solver, "quant", False) else 0
    cfg.no_plasticity = 0 if getattr(solver, "support_plasticity", True) else 1
    if any(kind == "bounding_box" and args.get("unbounded") for kind, args in
           solver.grid_postprocess):
        cfg.unbounded_half = unbounded_half(solver.res)
    return cfg


def unbounded_half(res):
    """grid_size / 2 of MPMSolver(unbounded=True): 4096 doubled while it is
    at most twice the largest resolution (mpm_solver.py:74, 143-148)."""
    g = 4096
    while g <= 2 * max(res):
        g *= 2
    return g // 2


# ---------------------------------------------------------------------------
# closed triangle meshes (outward-oriented) for the voxelizer tests
# ---------------------------------------------------------------------------
def box_mesh(lo, hi):
    """(12, 9) triangles of the axis-aligned box [lo, hi]."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
         (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return v, np.array(f)


def icosphere(subdiv):
    """Unit icosphere vertices and outward faces after subdiv subdivisions."""
    t = (1 + 5**0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v), np.array(f)


def torus(R, r, nu=24, nv=12):
    """Torus about the z axis, outward faces."""
    u = np.arange(nu) * 2 * np.pi / nu
    w = np.arange(nv) * 2 * np.pi / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    v = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U),
                  r * np.sin(W)], -1).reshape(-1, 3)
    f = []
    for i in range(nu):
        for j in range(nv):
            a, b = i * nv + j, ((i + 1) % nu) * nv + j
            c, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
            f += [(a, b, c), (a, c, d)]
    return v, np.array(f)


def triangles_of(v, f, scale=1.0, center=(0, 0, 0), rot=None):
    v = np.asarray(v, np.float64)
    if rot is not None:
        v = v @ np.asarray(rot).T
    v = v * scale + np.asarray(center, np.float64)
    return v[np.asarray(f)].reshape(-1, 9)


def winding_number(tri, pts):
    """Generalised winding number of the closed triangle mesh at pts: the
    sum of the signed solid angles of the triangles (Van Oosterom and
    Strackee's formula) over 4 pi; 1 inside an outward mesh, 0 outside."""
    tri = np.asarray(tri, np.float64).reshape(-1, 3, 3)
    out = np.zeros(len(pts))
    for a, b, c in tri:
        A, B, C = a - pts, b - pts, c - pts
        la, lb, lc = (np.linalg.norm(X, axis=1) for X in (A, B, C))
        num = np.einsum("ij,ij->i", A, np.cross(B, C))
        den = (la * lb * lc + np.einsum("ij,ij->i", A, B) * lc
               + np.einsum("ij,ij->i", B, C) * la + np.einsum("ij,ij->i", C, A) * lb)
        out += 2 * np.arctan2(num, den)
    return out / (4 * np.pi)


def mesh_volume(tri):
    """Signed volume of a closed triangle mesh (divergence theorem)."""
    tri = np.asarray(tri, np.float64).reshape(-1, 3, 3)
    return float(np.einsum("ij,ij->i", tri[:, 0], np.cross(tri[:, 1], tri[:, 2])).sum() / 6)
