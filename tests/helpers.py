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
    snow Jp in [0.9, 1.1], sand Jp in [-0.05, 0], water Jp in [0.95, 1.05]."""
    n = len(x)
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
    return cfg
