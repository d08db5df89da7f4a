"""Golden fixtures for one MLS-MPM substep, from an independent float64
brute-force restatement of the reference (engine/mpm_solver.py).

This script is test infrastructure.  It shares no code with the oracle
(oracle/mpm_oracle_impl.inc): it is a plain per-particle Python loop over a
dictionary grid, with the singular value decomposition taken from LAPACK
(numpy.linalg.svd) and brought to the reference's convention (U and V
rotations, the sign of det F on the last singular value).  It restates:

  constants          mpm_solver.py:82-87, 200-210
  p2g                mpm_solver.py:487-584 (sand_projection 318-340)
  grid operator      mpm_solver.py:586-598 (no g2p2g clamp), bounding box
                     walls 600-616
  g2p                mpm_solver.py:694-724

Each fixture holds the inputs (x, v, C, F, Jp, material, dt, gravity, res)
and the outputs of one substep (x, v, C, F, Jp and the grid nodes the
particles touch: integer node index, mass, velocity after the grid
operator).  Run from the repository root:

    python tests/golden/make_golden.py

which rewrites tests/golden/substep_{2,3}d.npz.  The states are
pre-deformed (F = R diag(1 +- 0.05) Q^T, C ~ N(0, 5^2), snow Jp in
[0.9, 1.1], sand Jp in [-0.05, 0], water Jp in [0.95, 1.05]) so that every
constitutive branch does work, and the particles sit next to the low walls
so that the bounding box acts.
"""
import math
import os

import numpy as np

WATER, ELASTIC, SNOW, SAND, STATIONARY = 0, 1, 2, 3, 4


def constants(res):
    dim = len(res)
    dx = 1.0 / res[0]
    E, nu = 1e6, 0.2
    sin_phi = math.sin(math.radians(45))
    return dict(dx=dx, inv_dx=1.0 / dx, p_vol=dx**dim, p_mass=dx**dim * 1000,
                mu_0=E / (2 * (1 + nu)), lambda_0=E * nu / ((1 + nu) * (1 - 2 * nu)),
                alpha=math.sqrt(2 / 3) * 2 * sin_phi / (3 - sin_phi), padding=3)


def svd_rotations(F):
    """F = U diag(sig) V^T with det U = det V = +1 (the last singular value
    carries the sign of det F)."""
    U, s, Vt = np.linalg.svd(F)
    V = Vt.T.copy()
    if np.linalg.det(U) < 0:
        U[:, -1] *= -1
        s[-1] *= -1
    if np.linalg.det(V) < 0:
        V[:, -1] *= -1
        s[-1] *= -1
    return U, s, V


def sand_projection(sig, Jp, k, dim):
    eps = np.log(np.maximum(np.abs(sig), 1e-4))
    tr = eps.sum() + Jp
    eps_hat = eps - tr / dim
    norm = np.linalg.norm(eps_hat) + 1e-20
    if tr >= 0.0:
        return np.ones(dim), tr
    dgamma = norm + (dim * k["lambda_0"] + 2 * k["mu_0"]) / (2 * k["mu_0"]) * tr * k["alpha"]
    return np.exp(eps - max(0.0, dgamma) / norm * eps_hat), 0.0


def substep(st, res, dt, gravity):
    dim = len(res)
    k = constants(res)
    x, v, C, F, Jp, mat = (st[f].copy() for f in ("x", "v", "C", "F", "Jp", "material"))
    inv_dx, dx = k["inv_dx"], k["dx"]
    grid = {}
    offsets = [np.array(o) for o in np.ndindex(*(3, ) * dim)]
    for p in range(len(x)):
        base = np.floor(x[p] * inv_dx - 0.5).astype(int)
        fx = x[p] * inv_dx - base
        w = [0.5 * (1.5 - fx)**2, 0.75 - (fx - 1)**2, 0.5 * (fx - 0.5)**2]
        Fp = F[p].copy()
        if mat[p] == WATER:
            Fp = np.eye(dim)
            Fp[0, 0] = Jp[p]
        Fp = (np.eye(dim) + dt * C[p]) @ Fp
        h = 1.0
        if mat[p] != WATER:
            h = math.exp(10 * (1.0 - Jp[p]))
        if mat[p] == ELASTIC:
            h = 0.3
        mu, la = k["mu_0"] * h, k["lambda_0"] * h
        if mat[p] == WATER:
            mu = 0.0
        U, sig, V = svd_rotations(Fp)
        J = 1.0
        if mat[p] != SAND:
            for d in range(dim):
                ns = sig[d]
                if mat[p] == SNOW:
                    ns = min(max(sig[d], 1 - 2.5e-2), 1 + 4.5e-3)
                Jp[p] *= sig[d] / ns
                sig[d] = ns
                J *= ns
        if mat[p] == WATER:
            Fp = np.eye(dim)
            Fp[0, 0] = J
            Jp[p] = J
        elif mat[p] == SNOW:
            Fp = U @ np.diag(sig) @ V.T
        if mat[p] != SAND:
            stress = 2 * mu * (Fp - U @ V.T) @ Fp.T + np.eye(dim) * la * J * (J - 1)
        else:
            sig, Jp[p] = sand_projection(sig, Jp[p], k, dim)
            Fp = U @ np.diag(sig) @ V.T
            ls = np.log(sig)
            center = np.diag(2.0 * k["mu_0"] * ls / sig + k["lambda_0"] * ls.sum() / sig)
            stress = U @ center @ V.T @ Fp.T
        F[p] = Fp
        stress = (-dt * k["p_vol"] * 4 * inv_dx**2) * stress
        mass = k["p_mass"]  # water_density 1
        affine = stress + mass * C[p]
        for o in offsets:
            weight = np.prod([w[o[d]][d] for d in range(dim)])
            node = tuple(base + o)
            m0, p0 = grid.get(node, (0.0, np.zeros(dim)))
            grid[node] = (m0 + weight * mass,
                          p0 + weight * (mass * v[p] + affine @ ((o - fx) * dx)))
    gv = {}
    for node, (m, mom) in grid.items():
        vel = mom.copy()
        if m > 0:
            vel = mom / m + dt * np.asarray(gravity, np.float64)
        for d in range(dim):
            if node[d] < k["padding"] and vel[d] < 0:
                vel[d] = 0.0
            if node[d] >= res[d] - k["padding"] and vel[d] > 0:
                vel[d] = 0.0
        gv[node] = vel
    for p in range(len(x)):
        base = np.floor(x[p] * inv_dx - 0.5).astype(int)
        fx = x[p] * inv_dx - base
        w = [0.5 * (1.5 - fx)**2, 0.75 - (fx - 1.0)**2, 0.5 * (fx - 0.5)**2]
        nv, nC = np.zeros(dim), np.zeros((dim, dim))
        for o in offsets:
            weight = np.prod([w[o[d]][d] for d in range(dim)])
            g = gv[tuple(base + o)]
            nv += weight * g
            nC += 4 * inv_dx * weight * np.outer(g, o - fx)
        if mat[p] != STATIONARY:
            v[p], C[p] = nv, nC
            x[p] = x[p] + dt * v[p]
    nodes = np.array(sorted(grid), np.int64)
    gm = np.array([grid[tuple(n)][0] for n in nodes])
    gvel = np.array([gv[tuple(n)] for n in nodes])
    return dict(x=x, v=v, C=C, F=F, Jp=Jp), nodes, gm, gvel


def rotations(rng, n, dim):
    Q = np.linalg.qr(rng.normal(size=(n, dim, dim)))[0]
    Q[np.linalg.det(Q) < 0, :, 0] *= -1
    return Q


def make(dim, seed, n=64):
    rng = np.random.default_rng(seed)
    res = (32, ) * dim if dim == 3 else (64, 64)
    # float32-representable inputs, so a float32 engine reads the same state
    x = rng.uniform(0.05, 0.25, (n, dim)).astype(np.float32).astype(np.float64)
    mat = np.arange(n) % 5
    R, Q = rotations(rng, n, dim), rotations(rng, n, dim)
    s = 1 + 0.05 * rng.uniform(-1, 1, (n, dim))
    F = np.einsum("nij,nj,nkj->nik", R, s, Q)
    C = 5.0 * rng.normal(size=(n, dim, dim))
    Jp = np.ones(n)
    Jp[mat == SNOW] = rng.uniform(0.9, 1.1, (mat == SNOW).sum())
    Jp[mat == SAND] = rng.uniform(-0.05, 0.0, (mat == SAND).sum())
    Jp[mat == WATER] = rng.uniform(0.95, 1.05, (mat == WATER).sum())
    v = rng.uniform(-1, 1, (n, dim))
    st = {f: a.astype(np.float32).astype(np.float64) for f, a in
          dict(x=x, v=v, C=C, F=F, Jp=Jp).items()}
    st["material"] = mat.astype(np.int32)
    dt = 2e-2 / res[0]
    gravity = [0.0, -9.8, 0.0][:dim]
    out, nodes, gm, gvel = substep(st, res, dt, gravity)
    return dict(res=np.array(res), dt=np.float64(dt), gravity=np.array(gravity),
                **{f"in_{f}": a for f, a in st.items()},
                **{f"out_{f}": a for f, a in out.items()},
                grid_nodes=nodes, grid_m=gm, grid_v=gvel)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for dim, seed in ((2, 20), (3, 30)):
        np.savez(os.path.join(here, f"substep_{dim}d.npz"), **make(dim, seed))


if __name__ == "__main__":
    main()
