"""Pins of the oracle's quantised storage (MPMSolver(quant=True)) and of
support_plasticity=False.

The formats (oracle_quant_x / _f / _v) are checked against an independent
exact-rational restatement in Python integers (fractions.Fraction: the
nearest representable value, ties to the even code), against their stated
error bounds, saturation, idempotence and power-of-two scaling.  Taichi's
own quantised types are third-party (taichi==1.1.0, absent), so the
rounding mode is the engine's documented choice; the bit budgets (21-bit x
with max 2, 7-bit shared exponent with 19-bit mantissas, 16-bit F with max
F_bound + 0.1) are the reference's (mpm_solver.py:106-114)."""
import math
from fractions import Fraction

import numpy as np
import pytest

from tests.helpers import cube_particles, deformed_state

F_STEP = np.float32(4.1 / 32768)


def _round_half_even(q):
    """Nearest integer to the Fraction q, ties to even."""
    f = math.floor(q)
    r = q - f
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and f % 2 == 1):
        f += 1
    return f


def test_x_grid_bounds_ties_and_saturation(oracle):
    rng = np.random.default_rng(1)
    x = rng.uniform(-1.99, 1.99, 100000).astype(np.float32)
    q = oracle.quant_x(x)
    codes = q.astype(np.float64) * 2**19
    assert np.all(codes == np.round(codes))
    for xi, qi in zip(x[:2000], q[:2000]):
        assert Fraction(float(qi)) == Fraction(_round_half_even(Fraction(float(xi)) * 2**19), 2**19)
    assert np.abs(q.astype(np.float64) - x).max() <= 2.0**-20
    # ties go to the even code
    ties = (np.arange(-20, 20) + 0.5) / 2**19
    qt = oracle.quant_x(ties) * 2**19
    assert np.all(qt % 2 == 0)
    # saturation at [-2, 2 - 2^-19]
    np.testing.assert_array_equal(oracle.quant_x([-7.0, 2.0, 9.0]), [-2.0, 2 - 2.0**-19, 2 - 2.0**-19])
    np.testing.assert_array_equal(oracle.quant_x(q), q)


def test_f_grid_bounds_and_saturation(oracle):
    rng = np.random.default_rng(2)
    f = rng.uniform(-4.09, 4.09, 100000).astype(np.float32)
    q = oracle.quant_f(f)
    # half a step, plus the float32 rounding of f * (32768 / 4.1) near ties
    assert np.abs(q.astype(np.float64) - f).max() <= 0.503 * 4.1 / 32768
    np.testing.assert_array_equal(oracle.quant_f(q), q)  # idempotent
    np.testing.assert_array_equal(oracle.quant_f(-q), -q)
    codes = np.rint(q.astype(np.float64) / (4.1 / 32768))
    assert np.abs(q - codes.astype(np.float32) * F_STEP).max() == 0
    sweep = oracle.quant_f(np.linspace(-4.2, 4.2, 400001, dtype=np.float32))
    assert len(np.unique(sweep)) == 65536  # every 16-bit code reached
    assert sweep.min() == np.float32(-32768) * F_STEP and sweep.max() == np.float32(32767) * F_STEP
    assert oracle.quant_f([0.0])[0] == 0.0
    # the identity is not representable: 1 -> 7992 steps
    assert oracle.quant_f([1.0])[0] == np.float32(7992) * F_STEP


def _exact_v(v):
    """Exact-rational shared-exponent rounding of one vector."""
    fr = [Fraction(float(a)) for a in v]
    m = max(abs(a) for a in fr)
    if m == 0:
        return [0.0] * len(v)
    e = math.floor(math.log2(m)) + 1  # m < 2^e
    while Fraction(2)**e <= m:
        e += 1
    while Fraction(2)**(e - 1) > m:
        e -= 1
    if e < -62:
        return [0.0] * len(v)
    mags = [_round_half_even(abs(a) * Fraction(2)**(18 - e)) for a in fr]
    if max(mags) >= 2**18:
        e += 1
        mags = [_round_half_even(abs(a) * Fraction(2)**(18 - e)) for a in fr]
    return [float(math.copysign(mg * Fraction(2)**(e - 18), a) if mg else 0.0) for mg, a in zip(mags, v)]


@pytest.mark.parametrize("dim", [2, 3])
def test_v_shared_exponent_against_exact_rounding(oracle, dim):
    rng = np.random.default_rng(3 + dim)
    v = (rng.normal(size=(3000, dim)) * 10.0**rng.uniform(-8, 6, (3000, 1))).astype(np.float32)
    v[::7, 0] = 0.0
    v[::11] *= np.float32(1e-3)
    q = oracle.quant_v(v)
    for a, b in zip(v, q):
        assert list(b) == [np.float32(t) for t in _exact_v(a)], (a, b)
    np.testing.assert_array_equal(oracle.quant_v(q), q)  # idempotent
    m = np.abs(v).max(1, keepdims=True).astype(np.float64)
    assert np.all(np.abs(q - v) <= m * 2.0**-18 + 1e-45)
    # power-of-two scaling commutes with the rounding (inside the range)
    np.testing.assert_array_equal(oracle.quant_v(v * np.float32(8)), q * np.float32(8))
    np.testing.assert_array_equal(oracle.quant_v(np.zeros((2, dim))), np.zeros((2, dim)))


def test_v_saturates_instead_of_overflowing(oracle):
    q = oracle.quant_v([[np.inf, 1.0, -np.inf], [1e30, -1e30, 0.0]])
    top = np.float32(262143 * 2.0**46)
    np.testing.assert_array_equal(q[0], [top, 0.0, -top])
    np.testing.assert_array_equal(q[1], [top, -top, 0.0])


def _one(oracle, cfg, mat, F, C, Jp, dt=1e-4, g2p2g=False):
    import ctypes
    d = cfg.dim
    F = np.array(F, np.float32).reshape(d, d).copy()
    C = np.array(C, np.float32).reshape(d, d)
    jp = np.array([Jp], np.float32)
    st = np.zeros((d, d), np.float32)
    P = ctypes.POINTER(ctypes.c_float)
    if g2p2g:
        pytest.skip("g2p2g constitutive is exercised through g2p2g()")
    oracle.lib().oracle_constitutive_f32(ctypes.byref(cfg), mat, ctypes.c_float(dt),
                                         F.ctypes.data_as(P), C.ctypes.data_as(P),
                                         jp.ctypes.data_as(P), st.ctypes.data_as(P))
    return F, st, float(jp[0])


def test_without_plasticity_closed_forms(oracle):
    cfg = oracle.with_storage(oracle.make_cfg((32, 32, 32)), plasticity=False)
    rng = np.random.default_rng(5)
    C = rng.normal(size=(3, 3)).astype(np.float32)
    dt = 1e-3
    G = np.eye(3, dtype=np.float32) + np.float32(dt) * C
    Fin = np.diag([1.02, 0.99, 1.01]).astype(np.float32)
    # sand: no stress, F = (I + dt C) F, Jp untouched
    F, st, jp = _one(oracle, cfg, 3, Fin, C, -0.3, dt)
    assert np.all(st == 0) and jp == np.float32(-0.3)
    np.testing.assert_allclose(F, G.astype(np.float64) @ Fin, rtol=1e-6, atol=1e-7)
    # water restarts from F = I whatever Jp holds: tau = la J (J - 1) I, J = det(I + dt C)
    F, st, jp = _one(oracle, cfg, 0, Fin, C, 0.5, dt)
    J = np.linalg.det(G.astype(np.float64))
    assert jp == np.float32(0.5)
    np.testing.assert_allclose(np.diag(st), cfg.lambda_0 * J * (J - 1), rtol=2e-4)
    np.testing.assert_allclose(F, np.diag([J, 1, 1]), rtol=1e-6, atol=1e-7)
    # snow: h = 1 whatever Jp (with plasticity Jp = 0.9 would give h = e)
    zero = np.zeros((3, 3), np.float32)
    s = np.float32(1.003)
    _, st_a, jp_a = _one(oracle, cfg, 2, np.eye(3) * s, zero, 0.9, dt)
    _, st_b, _ = _one(oracle, cfg, 2, np.eye(3) * s, zero, 1.0, dt)
    np.testing.assert_array_equal(st_a, st_b)
    assert jp_a == np.float32(0.9)
    expect = (2 * cfg.mu_0 * (s - 1) * s + cfg.lambda_0 * s**3 * (s**3 - 1))
    np.testing.assert_allclose(np.diag(st_a), expect, rtol=1e-4)
    plastic = oracle.make_cfg((32, 32, 32))
    _, st_c, _ = _one(oracle, plastic, 2, np.eye(3) * s, zero, 0.9, dt)
    np.testing.assert_allclose(st_c, st_a * np.exp(10 * (1 - 0.9)), rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("g2p2g", [False, True])
def test_quantised_substep_stores_representable_values(oracle, g2p2g):
    dim = 3
    rng = np.random.default_rng(7)
    n = 2000
    x = cube_particles(rng, n, [0.3] * dim, [0.4] * dim, dim)
    mat = rng.integers(0, 5, n).astype(np.int32)
    st = deformed_state(rng, x, mat, dim, stretch=0.02, c_scale=1.0)
    st["x"] = oracle.quant_x(st["x"])
    st["v"] = oracle.quant_v(st["v"])
    st["F"] = oracle.quant_f(st["F"])
    cfg = oracle.with_storage(oracle.make_cfg((32, ) * dim), quant=True)
    dt = 2e-2 / 32
    if g2p2g:
        stepper = oracle.G2P2GStepper(oracle.with_clamp(cfg, 0.9), st)
        stepper.substep(dt)
        stepper.substep(dt)
        out = stepper.s
    else:
        out, *_ = oracle.substep(cfg, st, dt, 0.0, np.float32)
    np.testing.assert_array_equal(oracle.quant_x(out["x"]), out["x"])
    np.testing.assert_array_equal(oracle.quant_v(out["v"]), out["v"])
    np.testing.assert_array_equal(oracle.quant_f(out["F"]), out["F"])
    assert np.abs(out["F"]).max() <= oracle.quant_f([4.0])[0] or not g2p2g
    # and close to the float run: one rounding per store
    ref, *_ = oracle.substep(oracle.make_cfg((32, ) * dim), st, dt, 0.0, np.float32) \
        if not g2p2g else (None,)
    if ref is not None:
        assert np.abs(out["x"] - ref["x"]).max() <= 2.0**-20 + 1e-7
        assert np.abs(out["F"] - ref["F"]).max() <= 0.5 * 4.1 / 32768 + 1e-6


def test_g2p2g_quantised_f_is_clamped(oracle):
    """mpm:413-416: in g2p2g with quant the new F is clamped to +-F_bound
    before it is stored.  The input grid carries v(x) = A x, so the G2P
    returns C = A exactly (APIC) and F_00 = 3.9 grows by (1 + dt A_00) to
    4.29 > 4; without quant it is stored unclamped."""
    dim = 3
    cfg = oracle.make_cfg((32, ) * dim, gravity=(0, 0, 0))
    st = dict(x=np.array([[0.5, 0.5, 0.5]], np.float32), v=np.zeros((1, 3), np.float32),
              C=np.zeros((1, 3, 3), np.float32),
              F=np.array([np.diag([3.9, 1.0, 0.25])], np.float32),
              Jp=np.ones(1, np.float32), material=np.array([1], np.int32))
    A = np.diag([100.0, 0.0, 0.0])
    shape = oracle.grid_shape(cfg)  # (z, y, x)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1)[..., ::-1]
    pos = (idx + np.array([cfg.grid_lo[a] for a in range(dim)])) * cfg.dx
    grid = (pos @ A.T).astype(np.float32)
    dt = 1e-3
    _, _, plain = oracle.g2p2g(cfg, st, grid, dt, n_final=1)  # not a new particle
    assert plain["F"][0, 0, 0] > 4.25
    _, _, q = oracle.g2p2g(oracle.with_storage(cfg, quant=True), st, grid, dt, n_final=1)
    F = q["F"][0]
    four = oracle.quant_f([4.0])[0]  # 31969 steps, 4.000027
    assert np.abs(F).max() <= four and F[0, 0] == four
    np.testing.assert_array_equal(oracle.quant_f(F), F)
