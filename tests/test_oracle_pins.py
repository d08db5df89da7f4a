"""Pins for the CPU restatement (oracle) of the MLS-MPM substep.

The reference publishes no golden vectors (its tests/*.py hold no asserts),
so every check here is fixed by mathematics: LAPACK SVD / polar
decompositions (numpy, scipy), closed forms of the constitutive models on
diagonal deformation gradients, conservation laws of the APIC transfers,
exactness of G2P on affine velocity fields, and per-node closed forms of the
grid operator and colliders.  A dropped term, a transposed operand or a wrong
sign in the restatement fails at least one of them.
"""
import math

import numpy as np
import pytest
import scipy.linalg

F32, F64 = np.float32, np.float64


# ----------------------------------------------------------------------------
# SVD (replaces ti.svd, mpm_solver.py:525)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("dtype,tol", [(F32, 2e-5), (F64, 1e-12)])
@pytest.mark.parametrize("dim", [2, 3])
def test_svd_against_lapack(oracle, dtype, tol, dim):
    rng = np.random.default_rng(10 + dim)
    for t in range(400):
        if t % 3 == 0:
            A = np.eye(dim) + 0.05 * rng.normal(size=(dim, dim))
        elif t % 3 == 1:
            A = rng.normal(size=(dim, dim))
        else:
            A = np.diag(rng.uniform(0.5, 1.5, dim))
        A = A.astype(dtype).astype(np.float64)
        U, s, V = oracle.svd(A, dtype)
        U, s, V = U.astype(np.float64), s.astype(np.float64), V.astype(np.float64)
        scale = max(1.0, np.abs(A).max())
        assert np.abs((U * s) @ V.T - A).max() <= tol * 10 * scale
        assert np.abs(U.T @ U - np.eye(dim)).max() <= tol * 10
        assert np.abs(V.T @ V - np.eye(dim)).max() <= tol * 10
        assert abs(np.linalg.det(U) - 1) <= tol * 10
        assert abs(np.linalg.det(V) - 1) <= tol * 10
        ref = np.linalg.svd(A, compute_uv=False)
        assert np.abs(np.abs(s) - ref).max() <= tol * 10 * ref[0]
        assert np.all(np.diff(np.abs(s)) <= tol * 10 * ref[0])  # decreasing
        assert np.all(s[:-1] >= 0)
        if abs(np.linalg.det(A)) > 1e-3:
            assert np.sign(s[-1]) == np.sign(np.linalg.det(A))
            if np.linalg.det(A) > 0:
                R, _ = scipy.linalg.polar(A)
                assert np.abs(U @ V.T - R).max() <= tol * 100


# ----------------------------------------------------------------------------
# Constitutive models (mpm_solver.py:507-568, 321-342)
# ----------------------------------------------------------------------------
def _cfg(oracle, dim=3, **kw):
    res = (64,) * dim
    return oracle.make_cfg(res, **kw)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("material", [0, 1, 2, 3])
def test_rest_state_has_zero_stress(oracle, dim, material):
    cfg = _cfg(oracle, dim)
    Jp = 0.0 if material == oracle.SAND else 1.0
    F, Jp2, st = oracle.constitutive(cfg, material, 1e-4, np.eye(dim), np.zeros((dim, dim)),
                                     Jp, F64)
    assert np.abs(st).max() <= 1e-9
    assert np.abs(F - np.eye(dim)).max() <= 1e-12
    assert Jp2 == Jp


def _fixed_corotated(F, mu, la):
    R, _ = scipy.linalg.polar(F)
    J = np.linalg.det(F)
    return 2 * mu * (F - R) @ F.T + la * J * (J - 1) * np.eye(len(F))


@pytest.mark.parametrize("dim", [2, 3])
def test_elastic_fixed_corotated(oracle, dim):
    """Elastic: h = 0.3 (mpm:519-521), tau = 2 mu (F - R) F^T + la J (J-1) I,
    with R from scipy's polar decomposition."""
    cfg = _cfg(oracle, dim)
    mu, la = 0.3 * cfg.mu_0, 0.3 * cfg.lambda_0
    rng = np.random.default_rng(3)
    for t in range(50):
        Q1 = scipy.linalg.qr(rng.normal(size=(dim, dim)))[0]
        Q2 = scipy.linalg.qr(rng.normal(size=(dim, dim)))[0]
        if np.linalg.det(Q1) < 0:
            Q1[:, 0] *= -1
        if np.linalg.det(Q2) < 0:
            Q2[:, 0] *= -1
        F0 = Q1 @ np.diag(rng.uniform(0.8, 1.2, dim)) @ Q2.T
        C = 10 * rng.normal(size=(dim, dim))
        dt = 1e-3
        Fn, Jp, st = oracle.constitutive(cfg, oracle.ELASTIC, dt, F0, C, 1.0, F64)
        F1 = (np.eye(dim) + dt * C) @ F0  # mpm:513
        np.testing.assert_allclose(Fn, F1, rtol=0, atol=1e-13)
        ref = _fixed_corotated(F1, mu, la)
        np.testing.assert_allclose(st, ref, rtol=0, atol=1e-9 * np.abs(ref).max())
        assert Jp == 1.0
        # Kirchhoff stress of an isotropic model is symmetric
        np.testing.assert_allclose(st, st.T, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_dilation_closed_form(oracle, dim):
    """F = s I: tau = (2 mu (s - 1) s + la s^d (s^d - 1)) I."""
    cfg = _cfg(oracle, dim)
    for s in (0.9, 0.97, 1.05):
        F0 = s * np.eye(dim)
        _, _, st = oracle.constitutive(cfg, oracle.ELASTIC, 0.0, F0, np.zeros((dim, dim)),
                                       1.0, F64)
        mu, la = 0.3 * cfg.mu_0, 0.3 * cfg.lambda_0
        J = s ** dim
        ref = (2 * mu * (s - 1) * s + la * J * (J - 1)) * np.eye(dim)
        np.testing.assert_allclose(st, ref, rtol=1e-12, atol=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("g2p2g", [0, 1])
def test_hardening_of_stationary_and_elastic(oracle, dim, g2p2g):
    """h = exp(10 (1 - Jp)) for every material but water in p2g and for
    every material in g2p2g (mpm:515-518, 419-421), then h = 0.3 for elastic
    (mpm:519-521).  A stationary particle at F = s I with Jp = j carries
    tau = h (2 mu_0 (s - 1) s + lambda_0 s^d (s^d - 1)) I and keeps Jp; an
    elastic one carries the h = 0.3 stress whatever its Jp."""
    cfg = _cfg(oracle, dim)
    for j in (0.9, 0.97, 1.0, 1.08):
        for s in (0.95, 1.03):
            J = s ** dim
            base = (2 * cfg.mu_0 * (s - 1) * s + cfg.lambda_0 * J * (J - 1)) * np.eye(dim)
            for mat, h in ((oracle.STATIONARY, math.exp(10 * (1 - j))), (oracle.ELASTIC, 0.3)):
                Fn, Jp, st = oracle.constitutive(cfg, mat, 0.0, s * np.eye(dim),
                                                 np.zeros((dim, dim)), j, F64, g2p2g=g2p2g)
                np.testing.assert_allclose(st, h * base, rtol=1e-12, atol=1e-6)
                np.testing.assert_allclose(Fn, s * np.eye(dim), rtol=1e-13)
                assert Jp == pytest.approx(j, rel=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_water_volume_state(oracle, dim):
    """Water keeps its volume in F[0,0] and Jp (mpm:508-511, 537-542): with
    J = Jp det(I + dt C), tau = lambda_0 J (J - 1) I (mu = 0, h = 1)."""
    cfg = _cfg(oracle, dim)
    rng = np.random.default_rng(4)
    for t in range(20):
        Jp0 = rng.uniform(0.9, 1.1)
        C = 20 * rng.normal(size=(dim, dim))
        dt = 1e-3
        Fjunk = rng.normal(size=(dim, dim))  # ignored for water
        Fn, Jp, st = oracle.constitutive(cfg, oracle.WATER, dt, Fjunk, C, Jp0, F64)
        J = Jp0 * np.linalg.det(np.eye(dim) + dt * C)
        np.testing.assert_allclose(Jp, J, rtol=1e-12)
        ref_F = np.eye(dim)
        ref_F[0, 0] = J
        np.testing.assert_allclose(Fn, ref_F, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(st, cfg.lambda_0 * J * (J - 1) * np.eye(dim),
                                   rtol=1e-9, atol=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
def test_snow_plasticity_closed_form(oracle, dim):
    """Snow (mpm:515-545): h = exp(10 (1 - Jp)) from the incoming Jp, singular
    values clamped to [1 - 2.5e-2, 1 + 4.5e-3], Jp *= sigma / clamped,
    F = clamped; evaluated on diagonal (and rotated diagonal) F."""
    cfg = _cfg(oracle, dim)
    rng = np.random.default_rng(5)
    lo, hi = 1 - 2.5e-2, 1 + 4.5e-3
    for t in range(30):
        sig = rng.uniform(0.95, 1.02, dim)
        Jp0 = rng.uniform(0.95, 1.05)
        Q = scipy.linalg.qr(rng.normal(size=(dim, dim)))[0]
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        F0 = Q @ np.diag(sig) @ Q.T
        Fn, Jp, st = oracle.constitutive(cfg, oracle.SNOW, 0.0, F0, np.zeros((dim, dim)),
                                         Jp0, F64)
        h = math.exp(10 * (1 - Jp0))
        mu, la = cfg.mu_0 * h, cfg.lambda_0 * h
        cl = np.clip(sig, lo, hi)
        np.testing.assert_allclose(Jp, Jp0 * np.prod(sig / cl), rtol=1e-10)
        Fe = Q @ np.diag(cl) @ Q.T
        np.testing.assert_allclose(Fn, Fe, atol=1e-10)
        J = np.prod(cl)
        ref = 2 * mu * (Fe - np.eye(dim)) @ Fe.T + la * J * (J - 1) * np.eye(dim)
        np.testing.assert_allclose(st, ref, rtol=0, atol=1e-8 * max(1, np.abs(ref).max()))


def _hencky(sig, mu, la, dim):
    ls = np.log(sig)
    return 2 * mu * ls + la * ls.sum()


@pytest.mark.parametrize("dim", [2, 3])
def test_sand_drucker_prager(oracle, dim):
    """Sand (mpm:552-566 with sand_projection mpm:321-342; Klar et al. 2016):
    (i) expansion (tr >= 0) projects to the cone tip: sigma = 1, Jp = tr,
    stress 0; (ii) inside the cone sigma is unchanged and the stress is the
    Hencky stress 2 mu ln(s) + la sum ln(s); (iii) otherwise the projected
    log-strain keeps its trace and lies on the yield surface
    |dev eps| + (d la + 2 mu)/(2 mu) tr(eps) alpha = 0.  Hardening is
    ignored (mu_0, la_0 are used)."""
    cfg = _cfg(oracle, dim)
    mu, la, alpha = cfg.mu_0, cfg.lambda_0, cfg.alpha
    coef = (dim * la + 2 * mu) / (2 * mu)
    # (i) tip
    sig = np.array([1.02, 1.01, 1.015][:dim])
    Fn, Jp, st = oracle.constitutive(cfg, oracle.SAND, 0.0, np.diag(sig), np.zeros((dim, dim)),
                                     0.0, F64)
    np.testing.assert_allclose(Jp, np.log(sig).sum(), rtol=1e-12)
    np.testing.assert_allclose(Fn, np.eye(dim), atol=1e-12)
    assert np.abs(st).max() < 1e-6
    # (ii) inside the cone: compression with a small deviatoric part
    sig = np.array([0.97, 0.969, 0.971][:dim])
    eps = np.log(sig)
    dev = eps - eps.sum() / dim
    assert np.linalg.norm(dev) + coef * eps.sum() * alpha < 0
    Fn, Jp, st = oracle.constitutive(cfg, oracle.SAND, 0.0, np.diag(sig), np.zeros((dim, dim)),
                                     0.0, F64)
    assert Jp == 0.0
    np.testing.assert_allclose(Fn, np.diag(sig), atol=1e-12)
    np.testing.assert_allclose(np.diag(st), _hencky(sig, mu, la, dim), rtol=1e-9)
    # (iii) outside the cone: large shear under compression
    rng = np.random.default_rng(6)
    for t in range(20):
        dev0 = np.array([0.15, -0.15]) if dim == 2 else np.array([0.15, 0.0, -0.15])
        sig = np.sort(0.97 * np.exp(dev0) * rng.uniform(0.98, 1.0, dim))[::-1]
        eps = np.log(sig)
        assert eps.sum() < 0
        assert (np.linalg.norm(eps - eps.sum() / dim) + coef * eps.sum() * alpha) > 0
        Q = scipy.linalg.qr(rng.normal(size=(dim, dim)))[0]
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        F0 = Q @ np.diag(sig) @ Q.T
        Fn, Jp, st = oracle.constitutive(cfg, oracle.SAND, 0.0, F0, np.zeros((dim, dim)),
                                         0.0, F64)
        s_out = np.linalg.svd(Fn, compute_uv=False)
        e_out = np.log(s_out)
        np.testing.assert_allclose(e_out.sum(), eps.sum(), rtol=1e-9)
        dev_out = e_out - e_out.sum() / dim
        np.testing.assert_allclose(np.linalg.norm(dev_out) + coef * e_out.sum() * alpha, 0,
                                   atol=1e-9)
        assert Jp == 0.0
        # stress = Hencky stress of the projected strain, rotated back
        ref = Q @ np.diag(_hencky(np.sort(s_out)[::-1], mu, la, dim)) @ Q.T
        np.testing.assert_allclose(st, ref, atol=1e-6 * np.abs(ref).max())


def test_sand_projection_function(oracle):
    """Direct call of sand_projection: max(|s|, 1e-4) guard and tip case."""
    s, jp = oracle.sand_projection(np.array([1e-6, 1.0, 1.0]), 0.0, 1.0, 1.0, 0.5, F64)
    # eps_0 = log(1e-4) < 0 -> compression; output finite
    assert np.all(np.isfinite(s)) and jp == 0.0
    s, jp = oracle.sand_projection(np.array([1.1, 1.0, 1.0]), 0.0, 1.0, 1.0, 0.5, F64)
    np.testing.assert_allclose(s, 1.0)
    np.testing.assert_allclose(jp, math.log(1.1))


# ----------------------------------------------------------------------------
# Transfers: P2G and G2P invariants (mpm_solver.py:487-584, 694-724)
# ----------------------------------------------------------------------------
def _random_state(oracle, dim, n, rng, materials=(0, 1, 2, 3), lo=0.3, hi=0.7):
    x = rng.uniform(lo, hi, size=(n, dim))
    v = rng.uniform(-1, 1, size=(n, dim))
    mat = rng.choice(materials, size=n).astype(np.int32)
    st = oracle.rest_state(x, v, mat)
    st["C"] = 5 * rng.normal(size=(n, dim, dim))
    F = np.empty((n, dim, dim))
    for p in range(n):
        Q = scipy.linalg.qr(rng.normal(size=(dim, dim)))[0]
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        F[p] = Q @ np.diag(rng.uniform(0.95, 1.05, dim)) @ Q.T
    st["F"] = F
    st["Jp"] = np.where(mat == 3, rng.uniform(-0.05, 0.0, n), rng.uniform(0.95, 1.05, n))
    st["x"] = x
    st["v"] = v
    return st


def _node_coords(cfg):
    dim = cfg.dim
    axes = [np.arange(cfg.grid_lo[a], cfg.grid_lo[a] + cfg.grid_n[a]) * cfg.dx
            for a in range(dim)]
    grids = np.meshgrid(*axes[::-1], indexing="ij")  # z, y, x order
    return np.stack(grids[::-1], axis=-1)  # (..., dim) with component a = axis a


@pytest.mark.parametrize("dim", [2, 3])
def test_p2g_conserves_mass_momentum_angular_momentum(oracle, dim):
    cfg = oracle.make_cfg((32,) * dim, water_density=1.5)
    rng = np.random.default_rng(7)
    st = _random_state(oracle, dim, 200, rng)
    dt = 2e-4
    gm, gp, new = oracle.p2g(cfg, st, dt, F64)
    mass = np.where(st["material"] == 0, cfg.p_mass * 1.5, cfg.p_mass)  # mpm:571-573
    np.testing.assert_allclose(gm.sum(), mass.sum(), rtol=1e-12)
    np.testing.assert_allclose(gp.reshape(-1, dim).sum(0), (mass[:, None] * st["v"]).sum(0),
                               rtol=1e-10, atol=1e-12)
    # APIC angular momentum: sum_i x_i ^ p_i = sum_p x_p ^ m v_p + dx^2/4 m eps_abc C_cb
    X = _node_coords(cfg).reshape(-1, dim)
    P = gp.reshape(-1, dim)
    D = cfg.dx ** 2 / 4
    C = st["C"]
    if dim == 3:
        Lg = np.cross(X, P).sum(0)
        Lp = np.cross(st["x"], mass[:, None] * st["v"]).sum(0)
        eps_term = np.stack([C[:, 2, 1] - C[:, 1, 2], C[:, 0, 2] - C[:, 2, 0],
                             C[:, 1, 0] - C[:, 0, 1]], -1)
        Lp = Lp + D * (mass[:, None] * eps_term).sum(0)
    else:
        Lg = (X[:, 0] * P[:, 1] - X[:, 1] * P[:, 0]).sum()
        Lp = (st["x"][:, 0] * mass * st["v"][:, 1] - st["x"][:, 1] * mass * st["v"][:, 0]).sum()
        Lp = Lp + D * (mass * (C[:, 1, 0] - C[:, 0, 1])).sum()
    np.testing.assert_allclose(Lg, Lp, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("material", [0, 1, 2, 3])
def test_p2g_stress_first_moment(oracle, dim, material):
    """With v = 0 and C = 0, sum_i p_i (x_i - x_p)^T = -dt p_vol tau (the APIC
    inertia tensor is dx^2/4 I), so the grid reproduces the constitutive
    stress; pins the stress scale -dt p_vol 4 inv_dx^2 (mpm:569)."""
    cfg = oracle.make_cfg((32,) * dim)
    rng = np.random.default_rng(8 + material)
    st = _random_state(oracle, dim, 1, rng, materials=(material,))
    st["v"][:] = 0
    st["C"][:] = 0
    dt = 1e-3
    gm, gp, new = oracle.p2g(cfg, st, dt, F64)
    _, _, tau = oracle.constitutive(cfg, material, dt, st["F"][0], st["C"][0],
                                    float(st["Jp"][0]), F64)
    X = _node_coords(cfg).reshape(-1, dim)
    M1 = gp.reshape(-1, dim).T @ (X - st["x"][0])
    np.testing.assert_allclose(-M1 / (dt * cfg.p_vol), tau, rtol=0,
                               atol=1e-7 * max(1.0, np.abs(tau).max()))


@pytest.mark.parametrize("dim", [2, 3])
def test_g2p_affine_exactness(oracle, dim):
    """An affine grid velocity v(x) = A x + b is reproduced exactly by G2P:
    v_p = A x_p + b, C_p = A, x += dt v (mpm:714-724); stationary particles
    keep x, v, C (mpm:722)."""
    cfg = oracle.make_cfg((32,) * dim)
    rng = np.random.default_rng(9)
    A = rng.normal(size=(dim, dim))
    b = rng.normal(size=dim)
    X = _node_coords(cfg)
    gv = X @ A.T + b
    st = _random_state(oracle, dim, 100, rng, materials=(0, 1, 4))
    dt = 1e-3
    new, mv = oracle.g2p(cfg, st, gv, dt, F64)
    moving = st["material"] != 4
    vp = st["x"] @ A.T + b
    np.testing.assert_allclose(new["v"][moving], vp[moving], atol=1e-11)
    np.testing.assert_allclose(new["C"][moving], np.broadcast_to(A, (moving.sum(), dim, dim)),
                               atol=1e-9)
    np.testing.assert_allclose(new["x"][moving], st["x"][moving] + dt * vp[moving], atol=1e-12)
    np.testing.assert_array_equal(new["x"][~moving], st["x"][~moving])
    np.testing.assert_array_equal(new["v"][~moving], st["v"][~moving])
    np.testing.assert_array_equal(new["C"][~moving], st["C"][~moving])
    np.testing.assert_allclose(mv, np.abs(new["v"]).max(), rtol=0)


# ----------------------------------------------------------------------------
# Grid operator and colliders (mpm_solver.py:586-687)
# ----------------------------------------------------------------------------
def test_grid_normalisation_gravity_and_walls(oracle):
    res = (16, 16, 16)
    cfg = oracle.make_cfg(res, padding=3, gravity=(0.5, -9.8, 1.0))
    rng = np.random.default_rng(11)
    shape = oracle.grid_shape(cfg)
    gm = np.where(rng.uniform(size=shape) < 0.7, rng.uniform(0.1, 2, shape), 0.0)
    gp = rng.normal(size=shape + (3,)) * (gm[..., None] > 0)
    dt = 1e-2
    m2, v2 = oracle.grid_op(cfg, gm, gp, dt, 0.0, F64)
    X = np.rint(_node_coords(cfg) / cfg.dx).astype(int)
    g = np.array([0.5, -9.8, 1.0])
    vexp = np.where(gm[..., None] > 0, gp / np.where(gm > 0, gm, 1)[..., None] + dt * g, gp)
    for a in range(3):
        low = (X[..., a] < 3) & (vexp[..., a] < 0)
        high = (X[..., a] >= res[a] - 3) & (vexp[..., a] > 0)
        vexp[..., a] = np.where(low | high, 0.0, vexp[..., a])
    np.testing.assert_allclose(v2, vexp, rtol=1e-14, atol=1e-14)
    np.testing.assert_array_equal(m2, gm)


@pytest.mark.parametrize("surface,friction", [(0, 0.0), (1, 0.0), (2, 0.0), (1, 0.4), (2, 0.4)])
def test_surface_collider(oracle, surface, friction):
    res = (16, 16, 16)
    point, normal = (0.5, 0.4, 0.5), (1.0, 2.0, 0.0)
    cfg = oracle.make_cfg(res, gravity=(0, 0, 0), post=[
        {"kind": "plane", "point": point, "normal": normal, "surface": surface,
         "friction": friction}])
    shape = oracle.grid_shape(cfg)
    rng = np.random.default_rng(12)
    gm = np.ones(shape)
    gv = rng.normal(size=shape + (3,))
    _, v2 = oracle.grid_op(cfg, gm, gv, 0.0, 0.0, F64)
    n = np.array(normal) / np.linalg.norm(normal)
    X = _node_coords(cfg)
    inside = (X - np.array(point)) @ n < 0
    np.testing.assert_array_equal(v2[~inside], gv[~inside])
    vi, vo = gv[inside], v2[inside]
    vn = vi @ n
    if surface == 0:
        assert np.abs(vo).max() == 0
        return
    vt = vi - np.outer(vn, n) if surface == 1 else vi - np.outer(np.minimum(vn, 0), n)
    # friction shrinks the post-projection velocity by friction*|v_n| when v_n < 0
    scale = np.ones(len(vn))
    neg = vn < 0
    nrm = np.linalg.norm(vt, axis=1)
    scale[neg] = np.maximum(0, nrm[neg] + vn[neg] * friction) / nrm[neg]
    np.testing.assert_allclose(vo, vt * scale[:, None], atol=1e-12)
    if surface == 1:
        np.testing.assert_allclose(vo @ n, 0, atol=1e-12)
    else:
        assert np.all(vo @ n >= -1e-12)


@pytest.mark.parametrize("surface", [0, 1, 2])
def test_sphere_collider(oracle, surface):
    res = (16, 16, 16)
    c, r = (0.5, 0.5, 0.5), 0.3
    cfg = oracle.make_cfg(res, gravity=(0, 0, 0), post=[
        {"kind": "sphere", "center": c, "radius": r, "surface": surface}])
    shape = oracle.grid_shape(cfg)
    rng = np.random.default_rng(13)
    gv = rng.normal(size=shape + (3,))
    _, v2 = oracle.grid_op(cfg, np.ones(shape), gv, 0.0, 0.0, F64)
    X = _node_coords(cfg) - np.array(c)
    d = np.linalg.norm(X, axis=-1)
    inside = d ** 2 < r * r
    np.testing.assert_array_equal(v2[~inside], gv[~inside])
    if surface == 0:
        assert np.abs(v2[inside]).max() == 0
        return
    nrm = X[inside] / (d[inside] + 1e-5)[:, None]
    vn = (gv[inside] * nrm).sum(1)
    proj = vn if surface == 1 else np.minimum(vn, 0)
    np.testing.assert_allclose(v2[inside], gv[inside] - nrm * proj[:, None], atol=1e-12)


# ----------------------------------------------------------------------------
# Binning (build_pid, mpm_solver.py:344-361): keys and stable permutation
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_block_keys_and_stable_sort(oracle, dim):
    rng = np.random.default_rng(14)
    n = 5000
    inv_dx = np.float32(96.0)
    x = rng.uniform(-0.05, 1.05, size=(n, dim)).astype(np.float32)
    x[:50] = (np.floor(x[:50] * 96) + 0.5) / np.float32(96)  # values at cell faces
    blo = np.array([-1] * dim, np.int32)
    bn = np.array([26] * dim, np.int32)
    keys = oracle.block_keys(x, float(inv_dx), 4, blo, bn)
    base = np.floor(x * inv_dx - np.float32(0.5)).astype(np.int64)  # numpy f32 ops, no FMA
    b = np.floor_divide(base, 4) - blo
    ok = np.all((b >= 0) & (b < bn), axis=1)
    lin = np.zeros(n, np.int64)
    for a in reversed(range(dim)):
        lin = lin * bn[a] + b[:, a]
    ref = np.where(ok, lin, np.prod(bn))
    np.testing.assert_array_equal(keys, ref)
    perm = oracle.sort_by_key(keys, int(np.prod(bn)))
    np.testing.assert_array_equal(perm, np.argsort(keys, kind="stable"))


# ----------------------------------------------------------------------------
# Whole substep
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_substep_free_fall_is_exact(oracle, dim):
    """Uniform velocity, F = I, C = 0, no walls: grid velocity is v0 + dt g
    everywhere, so G2P returns v = v0 + dt g, C = 0 (partition of unity and
    zero first moment of the weights)."""
    g = (0.3, -9.8, 0.7)[:dim]
    cfg = oracle.make_cfg((32,) * dim, gravity=g, post=[])
    rng = np.random.default_rng(15)
    n = 300
    x = rng.uniform(0.3, 0.7, size=(n, dim))
    v0 = np.array([0.2, -0.1, 0.3][:dim])
    st = oracle.rest_state(x, np.tile(v0, (n, 1)), np.ones(n, np.int32))
    st["x"] = x
    dt = 1e-3
    new, gm, gv, mv = oracle.substep(cfg, st, dt, 0.0, F64)
    v1 = v0 + dt * np.array(g)
    np.testing.assert_allclose(new["v"], np.tile(v1, (n, 1)), atol=1e-12)
    np.testing.assert_allclose(new["C"], 0, atol=1e-9)
    np.testing.assert_allclose(new["x"], x + dt * v1, atol=1e-14)
    np.testing.assert_allclose(mv, np.abs(v1).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_substep_conserves_momentum_without_gravity(oracle, dim):
    cfg = oracle.make_cfg((32,) * dim, gravity=(0, 0, 0)[:dim], post=[])
    rng = np.random.default_rng(16)
    st = _random_state(oracle, dim, 300, rng, materials=(1, 2, 3))
    mass = cfg.p_mass
    p0 = mass * st["v"].sum(0)
    new, gm, gv, mv = oracle.substep(cfg, st, 1e-4, 0.0, F64)
    np.testing.assert_allclose(mass * new["v"].sum(0), p0, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_substep_f32_tracks_f64(oracle, dim):
    cfg = oracle.make_cfg((32,) * dim)
    rng = np.random.default_rng(17)
    st = _random_state(oracle, dim, 400, rng)
    st32 = {k: np.asarray(v, np.float32 if k != "material" else np.int32)
            for k, v in st.items()}
    st64 = {k: np.asarray(v, np.float64) if k != "material" else v for k, v in st32.items()}
    a, *_ = oracle.substep(cfg, st32, 1e-4, 0.0, F32)
    b, *_ = oracle.substep(cfg, st64, 1e-4, 0.0, F64)
    for k in ("x", "v", "C", "F", "Jp"):
        err = np.linalg.norm(a[k] - b[k]) / max(np.linalg.norm(b[k]), 1e-6)
        assert err < 1e-5, (k, err)


@pytest.mark.parametrize("side", ["low", "high"])
def test_unbounded_walls_are_the_bounded_walls_moved(oracle, side):
    """unbounded=True walls at -grid_size/2 + padding and grid_size/2 -
    padding (mpm_solver.py:604-611) are the bounded walls at padding and
    res - padding moved by -grid_size/2 and grid_size/2 - res cells: an
    elastic blob thrown against either wall moves the same (float64) once
    translated by that many cells.  A wall one cell off changes the motion
    by ~1e-2."""
    res, half = (32, 32), 2048
    shift = -half if side == "low" else half - res[0]  # cells
    rng = np.random.default_rng(18)
    n = 600
    lo = 4 if side == "low" else 23
    x = np.stack([rng.uniform(lo, lo + 5, n), rng.uniform(12.8, 19.2, n)], 1) / res[0]
    v = np.zeros((n, 2))
    v[:, 0] = -20.0 if side == "low" else 20.0
    st = oracle.rest_state(x, v, np.ones(n, np.int32))
    runs = []
    for h, lo_x in ((0, -8), (half, -8 + shift)):
        cfg = oracle.make_cfg(res, post=[{"kind": "bbox"}], grid_lo=[lo_x, -8], grid_n=[48, 48])
        cfg.unbounded_half = h
        s = {k: (np.asarray(a, np.float64) if k != "material" else a) for k, a in st.items()}
        s["x"] = s["x"] + np.array([shift / res[0] if h else 0.0, 0.0])
        run = oracle.Substepper(cfg, s, F64)
        for _ in range(80):
            run.substep(cfg.dx * 2e-2)
        runs.append(run.s)
    bounded, unbounded = runs
    moved = unbounded["x"] - np.array([shift / res[0], 0.0])
    assert np.abs(bounded["x"][:, 0] * res[0] - lo).max() > 0  # it moved
    assert np.abs(bounded["x"][:, 0] * res[0]).min() > 1 and (bounded["v"][:, 0] != v[:, 0]).all()
    np.testing.assert_allclose(moved, bounded["x"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(unbounded["v"], bounded["v"], rtol=0, atol=1e-6)


def test_subnormal_node_mass_is_an_empty_node(oracle):
    """mpm_solver.py:591-593 normalises nodes with m > 0 as (1 / m) * p.  A
    float32 mass below the smallest normal float overflows 1/m; the
    reference on GPUs (flush-to-zero) reads it as 0 and leaves the node
    alone, and so do the oracle and the engine: such a node keeps its
    (subnormal) momentum and gets no gravity, a node of the smallest normal
    mass is normalised.  float64 normalises both (they are normal there)."""
    cfg = oracle.make_cfg((16, 16), padding=0, gravity=(0.0, -9.8), post=[])
    shape = oracle.grid_shape(cfg)
    dt = 1e-3
    for dtype in (np.float32, np.float64):
        gm = np.zeros(shape, dtype)
        gp = np.zeros(shape + (2, ), dtype)
        gm[8, 8], gp[8, 8] = 1e-40, (2e-40, -1e-40)        # subnormal in float32
        gm[8, 9], gp[8, 9] = 1.1754944e-38, (2.3509887e-38, 0)  # smallest normal
        m2, v2 = oracle.grid_op(cfg, gm, gp, dt, 0.0, dtype)
        assert np.isfinite(v2).all()
        np.testing.assert_allclose(v2[8, 9], [2.0, -9.8e-3], rtol=1e-6)
        if dtype == np.float32:
            np.testing.assert_array_equal(v2[8, 8], gp[8, 8])
        else:
            np.testing.assert_allclose(v2[8, 8], [2.0, -1.0 - 9.8e-3], rtol=1e-12)
