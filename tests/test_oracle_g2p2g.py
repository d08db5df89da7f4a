"""Pins for the oracle's use_g2p2g mode (mpm_solver.py:363-485, 586-598).

Fixed by the mathematics and by the p2g path already pinned in
test_oracle_pins.py: on a first substep (no G2P) the fused transfer is p2g
of the advected particles with C = 0; an affine input grid is reproduced
exactly by its G2P and the C it computes is the one the F update uses; new
particles keep their velocity; the grid velocity clamp per node; the water
state of g2p2g (volume in F[0][0], Jp kept, hardening by Jp) in closed form.
"""
import numpy as np
import pytest

F64 = np.float64


def _node_coords(cfg):
    dim = cfg.dim
    axes = [(np.arange(cfg.grid_n[a]) + cfg.grid_lo[a]) * cfg.dx for a in range(dim)]
    grids = np.meshgrid(*axes[::-1], indexing="ij")
    return np.stack(grids[::-1], axis=-1)


def _state(oracle, dim, n, rng, materials):
    x = rng.uniform(0.35, 0.65, (n, dim))
    mat = np.asarray([materials[i % len(materials)] for i in range(n)], np.int32)
    st = oracle.rest_state(x, rng.uniform(-1, 1, (n, dim)), mat)
    st = {k: (np.asarray(v, F64) if k != "material" else v) for k, v in st.items()}
    Q = np.linalg.qr(rng.normal(size=(n, dim, dim)))[0]
    Q[np.linalg.det(Q) < 0, :, 0] *= -1
    s = 1 + 0.04 * rng.uniform(-1, 1, (n, dim))
    F = np.einsum("nij,nj,nkj->nik", Q, s, Q)
    deformed = mat != 0
    st["F"][deformed] = F[deformed]
    st["Jp"][mat == 2] = rng.uniform(0.95, 1.05, (mat == 2).sum())
    st["Jp"][mat == 3] = rng.uniform(-0.05, 0.0, (mat == 3).sum())
    return st


@pytest.mark.parametrize("dim", [2, 3])
def test_first_substep_is_p2g_of_the_advected_particles(oracle, dim):
    """n_final = 0: no G2P, x += dt v (not for stationary particles), C = 0,
    then p2g -- the grid equals the p2g path's on that state (water at rest
    volume, where both paths agree)."""
    cfg = oracle.make_cfg((32, ) * dim)
    rng = np.random.default_rng(30 + dim)
    st = _state(oracle, dim, 300, rng, (0, 1, 2, 3, 4))
    dt = 3e-4
    garbage = rng.normal(size=oracle.grid_shape(cfg) + (dim, ))
    gm, gp, new = oracle.g2p2g(cfg, st, garbage, dt, n_final=0, dtype=F64)
    adv = {k: v.copy() for k, v in st.items()}
    moving = st["material"] != 4
    adv["x"][moving] += dt * st["v"][moving]
    adv["C"][:] = 0
    gm_ref, gp_ref, ref = oracle.p2g(cfg, adv, dt, F64)
    np.testing.assert_allclose(gm, gm_ref, rtol=0, atol=1e-15)
    np.testing.assert_allclose(gp, gp_ref, rtol=0, atol=1e-12 * np.abs(gp_ref).max())
    np.testing.assert_allclose(new["x"], adv["x"], rtol=0, atol=0)
    nonwater = st["material"] != 0
    np.testing.assert_allclose(new["F"][nonwater], ref["F"][nonwater], atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_affine_input_grid(oracle, dim):
    """G2P of v(x) = A x + b gives v = A x_p + b and C = A; the elastic F
    update then is (I + dt A) F, the grid momentum is sum m v', and new
    particles (index >= n_final) keep their velocity."""
    cfg = oracle.make_cfg((32, ) * dim)
    rng = np.random.default_rng(40 + dim)
    A = 0.5 * rng.normal(size=(dim, dim))
    b = rng.normal(size=dim)
    gv_in = _node_coords(cfg) @ A.T + b
    st = _state(oracle, dim, 200, rng, (1, ))
    n_final = 150
    dt = 2e-4
    gm, gp, new = oracle.g2p2g(cfg, st, gv_in, dt, n_final=n_final, dtype=F64)
    vp = st["x"] @ A.T + b
    np.testing.assert_allclose(new["v"][:n_final], vp[:n_final], atol=1e-11)
    np.testing.assert_array_equal(new["v"][n_final:], st["v"][n_final:])
    np.testing.assert_allclose(new["x"], st["x"] + dt * new["v"], atol=1e-14)
    K = np.eye(dim) + dt * A
    np.testing.assert_allclose(new["F"][:n_final], K @ st["F"][:n_final], atol=1e-12)
    np.testing.assert_allclose(new["F"][n_final:], st["F"][n_final:], atol=1e-14)  # C = 0
    np.testing.assert_allclose(gp.reshape(-1, dim).sum(0), cfg.p_mass * new["v"].sum(0),
                               rtol=1e-10)
    np.testing.assert_allclose(gm.sum(), cfg.p_mass * len(new["v"]), rtol=1e-12)
    np.testing.assert_array_equal(new["C"], st["C"])  # the C field is not written


@pytest.mark.parametrize("dim", [2, 3])
def test_water_state_and_stress(oracle, dim):
    """Water in g2p2g: F' = (I + dt C) F with the stored F, J = det F',
    F_out = diag(J, 1..), Jp unchanged, tau = lambda_0 exp(10 (1 - Jp))
    J (J - 1) I (mpm:416-448), read from the grid's first moment
    (-dt p_vol tau, as for p2g)."""
    cfg = oracle.make_cfg((32, ) * dim)
    rng = np.random.default_rng(50 + dim)
    st = _state(oracle, dim, 1, rng, (0, ))
    J0 = 0.96
    st["F"][0] = np.eye(dim)
    st["F"][0][0, 0] = J0
    st["Jp"][0] = 1.02
    st["v"][:] = 0
    dt = 1e-3
    gin = np.zeros(oracle.grid_shape(cfg) + (dim, ))
    gm, gp, new = oracle.g2p2g(cfg, st, gin, dt, n_final=1, dtype=F64)  # G2P of zero: C = 0
    assert new["Jp"][0] == 1.02
    expect_F = np.eye(dim)
    expect_F[0, 0] = J0
    np.testing.assert_allclose(new["F"][0], expect_F, atol=1e-14)
    tau = cfg.lambda_0 * np.exp(10 * (1 - 1.02)) * J0 * (J0 - 1) * np.eye(dim)
    X = _node_coords(cfg).reshape(-1, dim)
    M1 = gp.reshape(-1, dim).T @ (X - new["x"][0])
    np.testing.assert_allclose(-M1 / (dt * cfg.p_vol), tau, rtol=0, atol=1e-9 * np.abs(tau).max())


def test_grid_velocity_clamp(oracle):
    """grid_normalization_and_gravity with the g2p2g clamp: every component
    of v = p / m + dt g limited to +-dx * cfl / dt (mpm:589-598), on nodes
    without mass too."""
    res = (16, 16, 16)
    cfg = oracle.with_clamp(oracle.make_cfg(res, post=[]), 0.9)
    rng = np.random.default_rng(60)
    shape = oracle.grid_shape(cfg)
    gm = rng.uniform(0, 2, shape)
    gm[rng.random(shape) < 0.3] = 0
    gv = rng.normal(0, 300, shape + (3, )) * (gm > 0)[..., None]
    dt = 2e-4
    m2, v2 = oracle.grid_op(cfg, gm, gv, dt, dtype=F64)
    va = cfg.dx * 0.9 / dt
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.where(gm[..., None] > 0, gv / gm[..., None] + dt * np.array([0, -9.8, 0]), 0)
    np.testing.assert_allclose(v2, np.clip(v, -va, va), rtol=1e-13, atol=1e-12)
    assert (np.abs(v2) <= va * (1 + 1e-12)).all() and (np.abs(v) > va).any()
