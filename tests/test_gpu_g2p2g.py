"""use_g2p2g mode (fused g2p2g transfer, mpm_solver.py:363-485) on the GPU
against the oracle's G2P2GStepper (same seeded inputs, float32).

Tolerances: norm-wise relative 1e-5 on x, v, F, Jp and the grid after the
first substep (p2g of the advected particles) and 5e-5 after three (the
G2P's C feeds the F update, which amplifies grid rounding as in the p2g
path's C); the particles' C field is never written in this mode and must
stay bit-identical."""
import numpy as np
import pytest

from tests.helpers import cube_particles, deformed_state, oracle_cfg_for, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _setup(oracle, dim, cfl=0.9, vscale=1.0, n=6000, seed=0):
    from taichi_elements_amd import MPMSolver
    rng = np.random.default_rng(70 + dim + seed)
    s = MPMSolver(res=(48, ) * dim, use_g2p2g=True, g2p2g_allowed_cfl=cfl)
    x = cube_particles(rng, n, [0.3] * dim, [0.35] * dim, dim)
    mat = rng.integers(0, 5, n).astype(np.int32)
    st = deformed_state(rng, x, mat, dim)
    st["v"] = (vscale * st["v"]).astype(np.float32)
    # water keeps its volume in F[0][0] in this mode
    water = mat == 0
    st["F"][water] = np.eye(dim, dtype=np.float32)
    st["F"][water, 0, 0] = st["Jp"][water]
    st["Jp"][water] = 1.0
    s.read_restart(n, st["x"], st["v"], st["material"], np.zeros(n, np.int32))
    s.set_state(F=st["F"], C=st["C"], Jp=st["Jp"])
    cfg = oracle.with_clamp(oracle_cfg_for(oracle, s), cfl)
    ref = oracle.G2P2GStepper(cfg, st)
    return s, ref, st


def _compare(s, ref, st, tol, label):
    got = s.get_state()
    for k in ("x", "v", "F", "Jp"):
        e = rel_err(got[k], ref.s[k])
        print(label, k, f"{e:.2e}", flush=True)
        assert e <= tol, (label, k, e)
    assert np.array_equal(got["C"], st["C"]), "C must not be written"
    m, v = s.read_grid()
    lo = [-8] * s.dim
    sl = tuple(slice(8, 8 + s.res[a]) for a in reversed(range(s.dim)))
    em = rel_err(m, ref.gm[sl])
    ev = rel_err(v, ref.gv_in[sl])
    print(label, f"grid m {em:.2e} v {ev:.2e}", flush=True)
    assert em <= tol and ev <= tol
    del lo


@pytest.mark.parametrize("dim", [2, 3])
def test_g2p2g_substeps_against_oracle(oracle, dim):
    s, ref, st = _setup(oracle, dim)
    dt = s.default_dt
    for k in range(3):
        s.substep(dt)
        ref.substep(dt)
        _compare(s, ref, st, 1e-5 if k == 0 else 5e-5, f"{dim}D substep {k + 1}")


def test_g2p2g_clamp_and_new_particles(oracle):
    """A low allowed CFL makes the clamp bind; particles added between
    substeps skip the G2P once (they keep their velocity)."""
    s, ref, st = _setup(oracle, 3, cfl=0.2, vscale=20.0, n=3000, seed=5)
    dt = s.default_dt
    s.substep(dt)
    ref.substep(dt)
    vmax = s.get_state()["v"]
    va = s.dx * 0.2 / dt
    assert np.abs(ref.gv_in).max() <= va * (1 + 1e-6)
    rng = np.random.default_rng(9)
    extra = cube_particles(rng, 500, [0.45, 0.45, 0.45], [0.1, 0.1, 0.1], 3)
    vel = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    s.read_restart(500, extra, vel, np.full(500, 1, np.int32), np.zeros(500, np.int32))
    ref.s = {k: np.concatenate([ref.s[k], add]) for k, add in (
        ("x", extra), ("v", vel), ("C", np.zeros((500, 3, 3), np.float32)),
        ("F", np.tile(np.eye(3, dtype=np.float32), (500, 1, 1))),
        ("Jp", np.ones(500, np.float32)), ("material", np.full(500, 1, np.int32)))}
    ref.n = len(ref.s["Jp"])
    st = dict(st, C=np.concatenate([st["C"], np.zeros((500, 3, 3), np.float32)]))
    for k in range(2):
        s.substep(dt)
        ref.substep(dt)
        _compare(s, ref, st, 5e-5, f"added particles substep {k + 1}")
    del vmax


def test_g2p2g_too_fast_new_particles_are_reported():
    from taichi_elements_amd import MPMSolver
    from taichi_elements_amd._lib import DomainError
    s = MPMSolver(res=(32, 32, 32), use_g2p2g=True)
    rng = np.random.default_rng(3)
    x = cube_particles(rng, 100, [0.4] * 3, [0.1] * 3, 3)
    # 3 cells per substep on their first (G2P-free) substep
    v = np.full((100, 3), 3.0 / 32 / s.default_dt, np.float32)
    s.read_restart(100, x, v, np.full(100, 1, np.int32), np.zeros(100, np.int32))
    with pytest.raises(DomainError):
        s.substep(s.default_dt)
