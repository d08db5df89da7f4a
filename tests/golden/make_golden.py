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
  g2p2g (use_g2p2g)  mpm_solver.py:363-485, clamp 589, 596-598, loop 773-787

Each fixture holds the inputs (x, v, C, F, Jp, material, dt, gravity, res)
and the outputs of one substep (x, v, C, F, Jp and the grid nodes the
particles touch: integer node index, mass, velocity after the grid
operator).  Run from the repository root:

    python tests/golden/make_golden.py

which rewrites tests/golden/substep_{2,3}d.npz and g2p2g_{2,3}d.npz (three
use_g2p2g substeps).  The states are
pre-deformed (F = R diag(1 +- 0.05) Q^T, C ~ N(0, 5^2), snow, elastic and
stationary Jp in [0.9, 1.1], sand Jp in [-0.05, 0], water Jp in
[0.95, 1.05]) so that every constitutive branch does work (the hardening of
stationary particles included), and the particles sit next to the low walls
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


def constitutive_g2p2g(Fp, Jp, mat, C, dt, k, dim):
    """The constitutive half of g2p2g (mpm_solver.py:411-473): F is not
    rebuilt for water, the hardening applies to every material but elastic,
    and Jp of water is untouched (its volume stays in F[0][0])."""
    Fp = (np.eye(dim) + dt * C) @ Fp
    h = math.exp(10 * (1.0 - Jp))
    if mat == ELASTIC:
        h = 0.3
    mu, la = k["mu_0"] * h, k["lambda_0"] * h
    if mat == WATER:
        mu = 0.0
    U, sig, V = svd_rotations(Fp)
    J = 1.0
    if mat != SAND:
        for d in range(dim):
            ns = sig[d]
            if mat == SNOW:
                ns = min(max(sig[d], 1 - 2.5e-2), 1 + 4.5e-3)
            Jp *= sig[d] / ns
            sig[d] = ns
            J *= ns
    if mat == WATER:
        Fp = np.eye(dim)
        Fp[0, 0] = J
    elif mat == SNOW:
        Fp = U @ np.diag(sig) @ V.T
    if mat != SAND:
        stress = 2 * mu * (Fp - U @ V.T) @ Fp.T + np.eye(dim) * la * J * (J - 1)
    else:
        sig, Jp = sand_projection(sig, Jp, k, dim)
        Fp = U @ np.diag(sig) @ V.T
        ls = np.log(sig)
        center = np.diag(2.0 * k["mu_0"] * ls / sig + k["lambda_0"] * ls.sum() / sig)
        stress = U @ center @ V.T @ Fp.T
    return Fp, Jp, stress


def g2p2g_substeps(st, res, dt, gravity, nsub, cfl=0.9):
    """use_g2p2g substeps (mpm_solver.py:363-485, 586-598, 773-787): g2p from
    the input grid (none before the first substep: every particle is new and
    keeps its v with C = 0), advection, constitutive update and p2g into the
    output grid; the grid operator clamps the velocity to +-cfl dx / dt.
    Returns the state and output grid after each substep."""
    dim = len(res)
    k = constants(res)
    x, v, F, Jp, mat = (st[f].copy() for f in ("x", "v", "F", "Jp", "material"))
    inv_dx, dx = k["inv_dx"], k["dx"]
    offsets = [np.array(o) for o in np.ndindex(*(3, ) * dim)]
    grid_in = {}
    n_final = 0
    out = []
    for _ in range(nsub):
        grid = {}
        for p in range(len(x)):
            base = np.floor(x[p] * inv_dx - 0.5).astype(int)
            fx = x[p] * inv_dx - base
            w = [0.5 * (1.5 - fx)**2, 0.75 - (fx - 1.0)**2, 0.5 * (fx - 0.5)**2]
            nv, C = np.zeros(dim), np.zeros((dim, dim))
            for o in offsets:
                weight = np.prod([w[o[d]][d] for d in range(dim)])
                g = grid_in.get(tuple(base + o), np.zeros(dim))
                nv += weight * g
                C += 4 * inv_dx * weight * np.outer(g, o - fx)
            if p >= n_final:
                nv, C = v[p].copy(), np.zeros((dim, dim))
            if mat[p] != STATIONARY:
                v[p] = nv
                x[p] = x[p] + dt * v[p]
            base = np.floor(x[p] * inv_dx - 0.5).astype(int)
            fx = x[p] * inv_dx - base
            w2 = [0.5 * (1.5 - fx)**2, 0.75 - (fx - 1)**2, 0.5 * (fx - 0.5)**2]
            F[p], Jp[p], stress = constitutive_g2p2g(F[p], Jp[p], mat[p], C, dt, k, dim)
            affine = (-dt * k["p_vol"] * 4 * inv_dx**2) * stress + k["p_mass"] * C
            for o in offsets:
                weight = np.prod([w2[o[d]][d] for d in range(dim)])
                node = tuple(base + o)
                m0, p0 = grid.get(node, (0.0, np.zeros(dim)))
                grid[node] = (m0 + weight * k["p_mass"],
                              p0 + weight * (k["p_mass"] * v[p] + affine @ ((o - fx) * dx)))
        v_allowed = dx * cfl / dt
        grid_in = {}
        for node, (m, mom) in grid.items():
            vel = mom.copy()
            if m > 0:
                vel = mom / m + dt * np.asarray(gravity, np.float64)
            vel = np.minimum(np.maximum(vel, -v_allowed), v_allowed)
            for d in range(dim):
                if node[d] < k["padding"] and vel[d] < 0:
                    vel[d] = 0.0
                if node[d] >= res[d] - k["padding"] and vel[d] > 0:
                    vel[d] = 0.0
            grid_in[node] = vel
        n_final = len(x)
        nodes = np.array(sorted(grid), np.int64)
        out.append((dict(x=x.copy(), v=v.copy(), F=F.copy(), Jp=Jp.copy()), nodes,
                    np.array([grid[tuple(q)][0] for q in nodes]),
                    np.array([grid_in[tuple(q)] for q in nodes])))
    return out


def rotations(rng, n, dim):
    Q = np.linalg.qr(rng.normal(size=(n, dim, dim)))[0]
    Q[np.linalg.det(Q) < 0, :, 0] *= -1
    return Q


def initial_state(dim, seed, n=64):
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
    # elastic and stationary particles with Jp != 1: stationary ones are
    # hardened by it, elastic ones are not (mpm_solver.py:515-521, 419-424)
    for m in (ELASTIC, STATIONARY):
        Jp[mat == m] = rng.uniform(0.9, 1.1, (mat == m).sum())
    st = {f: a.astype(np.float32).astype(np.float64) for f, a in
          dict(x=x, v=v, C=C, F=F, Jp=Jp).items()}
    st["material"] = mat.astype(np.int32)
    return st, res, 2e-2 / res[0], [0.0, -9.8, 0.0][:dim]


def make(dim, seed):
    st, res, dt, gravity = initial_state(dim, seed)
    out, nodes, gm, gvel = substep(st, res, dt, gravity)
    return dict(res=np.array(res), dt=np.float64(dt), gravity=np.array(gravity),
                **{f"in_{f}": a for f, a in st.items()},
                **{f"out_{f}": a for f, a in out.items()},
                grid_nodes=nodes, grid_m=gm, grid_v=gvel)


def make_g2p2g(dim, seed, nsub=3):
    """Three use_g2p2g substeps; the state after each, and the output grid of
    the last.  The input velocities are kept below 0.6 cells per substep (a
    g2p2g particle may move at most one cell)."""
    st, res, dt, gravity = initial_state(dim, seed)
    st["v"] = (st["v"] * 0.3).astype(np.float32).astype(np.float64)
    out = g2p2g_substeps(st, res, dt, gravity, nsub)
    d = dict(res=np.array(res), dt=np.float64(dt), gravity=np.array(gravity), nsub=np.int64(nsub),
             **{f"in_{f}": a for f, a in st.items()})
    for i, (state, nodes, gm, gvel) in enumerate(out):
        d.update({f"out{i + 1}_{f}": a for f, a in state.items()})
    d.update(grid_nodes=nodes, grid_m=gm, grid_v=gvel)
    return d


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for dim, seed in ((2, 20), (3, 30)):
        np.savez(os.path.join(here, f"substep_{dim}d.npz"), **make(dim, seed))
        np.savez(os.path.join(here, f"g2p2g_{dim}d.npz"), **make_g2p2g(dim, seed + 1))


if __name__ == "__main__":
    main()
