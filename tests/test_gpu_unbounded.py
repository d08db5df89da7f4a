"""MPMSolver(unbounded=True) on the GPU against the oracle.

The reference (mpm_solver.py:74, 143-148, 604-611) widens its sparse grid so
that the bounding-box walls sit at -grid_size/2 + padding and
grid_size/2 - padding cells (grid_size: 4096, doubled while it is at most
twice the largest resolution).  Here the dense block window follows the
particles instead.  Checked:
  * a blob thrown out of the [0, 1)^d domain passes where the bounded walls
    would stop it and keeps flying for 240 substeps (36 blocks): against the
    float32 oracle on one grid window that spans the whole trajectory, drift
    <= 1e-4 on x and v (the bound of the other 100-substep tests), with and
    without HIP graph replay; the grid read back at the blob's final place
    holds all of its mass;
  * the same flight with use_g2p2g: the grid a g2p2g substep reads is
    carried into the moved window; and with quant=True (the window follows
    positions decoded from the packed words), x and v within the same bound
    (measured 2e-5 and 6e-5; F, whose 16-bit codes flip at rounding
    boundaries, is printed);
  * particles seeded beside the far wall at -grid_size/2 + padding are
    stopped by it as in the oracle (x within 1e-4; v, which the elastic
    bounce makes sensitive to rounding, at most 3x as far from the float64
    oracle as the float32 oracle is);
  * a scene spread over more blocks than the window can hold raises
    DomainError; batches and the differentiable mode refuse unbounded.
"""
import numpy as np
import pytest

from tests.helpers import cube_particles, oracle_cfg_for, rel_err, unbounded_half

pytestmark = pytest.mark.gpu

DRIFT_TOL = 1e-4


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _blob(dim, n, v, lo=0.35, size=0.2, seed=0, materials=(0, 1)):
    rng = np.random.default_rng(700 + 10 * dim + seed)
    x = cube_particles(rng, n, [lo] * dim, [size] * dim, dim)
    vel = (np.asarray(v, np.float32) + rng.uniform(-0.5, 0.5, (n, dim))).astype(np.float32)
    mat = rng.choice(np.asarray(materials, np.int32), n).astype(np.int32)
    return x, vel, mat


def _set_window(cfg, lo, n):
    for a in range(len(lo)):
        cfg.grid_lo[a] = lo[a]
        cfg.grid_n[a] = n[a]


def _flight(dim):
    """A mixed water / elastic blob of 32^d-resolution particles thrown at
    30 units/s along x (0.6 cells per substep at default_dt), upwards in y
    and, in 3D, through the z = 0 floor; the oracle's grid window spans the
    whole trajectory."""
    n = 3000 if dim == 2 else 6000
    v = [30.0, 5.0] + ([-10.0] if dim == 3 else [])
    lo = [-8, -8] + ([-56] if dim == 3 else [])
    cells = [190, 72] + ([88] if dim == 3 else [])
    return n, v, lo, cells


@pytest.mark.parametrize("dim,graphs", [(2, True), (3, True), (3, False)])
def test_blob_leaves_the_domain_like_the_oracle(oracle, dim, graphs):
    from taichi_elements_amd import MPMSolver
    n, v, lo, cells = _flight(dim)
    s = MPMSolver(res=(32, ) * dim, unbounded=True)
    s.set_graphs(graphs)
    x, vel, mat = _blob(dim, n, v)
    s.read_restart(n, x, vel, mat, np.zeros(n, np.int32))
    cfg = oracle_cfg_for(oracle, s)
    assert cfg.unbounded_half == unbounded_half((32, ) * dim) == 2048
    _set_window(cfg, lo, cells)
    ref = oracle.Substepper(cfg, oracle.rest_state(x, vel, mat), np.float32)
    dt = s.default_dt
    for k in range(4):  # calls of 60 substeps: the window moves inside and between calls
        s._run_substeps(60, dt)
        for _ in range(60):
            ref.substep(dt)
    got = s.get_state()
    errs = {f: rel_err(got[f], ref.s[f]) for f in ("x", "v", "C", "F", "Jp")}
    print(f"unbounded {dim}D graphs={graphs} vs f32 oracle",
          {f: f"{e:.2e}" for f, e in errs.items()}, flush=True)
    for f in ("x", "v"):
        assert errs[f] <= DRIFT_TOL, (f, errs[f])
    # far outside where the bounded walls (3 cells inside [0, 1)) would stand
    assert got["x"][:, 0].min() > 3.0
    if dim == 3:
        assert got["x"][:, 2].max() < 0.0
    # the grid after the last grid operator, read in global node coordinates
    base = np.floor(got["x"].min(0) * 32).astype(int) - 4
    span = (np.ceil(got["x"].max(0) * 32).astype(int) + 4 - base)
    m, _ = s.read_grid(base.tolist(), span.tolist())
    assert abs(m.sum() / (n * s.p_mass) - 1) < 1e-4


def test_blob_leaves_the_domain_with_g2p2g(oracle):
    from taichi_elements_amd import MPMSolver
    dim = 3
    n, v, lo, cells = _flight(dim)
    s = MPMSolver(res=(32, ) * dim, unbounded=True, use_g2p2g=True)
    x, vel, mat = _blob(dim, n, v, seed=1)
    s.read_restart(n, x, vel, mat, np.zeros(n, np.int32))
    cfg = oracle.with_clamp(oracle_cfg_for(oracle, s), s.g2p2g_allowed_cfl)
    _set_window(cfg, lo, cells)
    ref = oracle.G2P2GStepper(cfg, oracle.rest_state(x, vel, mat))
    dt = s.default_dt
    for _ in range(3):
        s._run_substeps(80, dt)
        for _ in range(80):
            ref.substep(dt)
    got = s.get_state()
    errs = {f: rel_err(got[f], ref.s[f]) for f in ("x", "v", "F", "Jp")}
    print("unbounded g2p2g vs f32 oracle", {f: f"{e:.2e}" for f, e in errs.items()}, flush=True)
    for f in ("x", "v"):
        assert errs[f] <= DRIFT_TOL, (f, errs[f])
    assert got["x"][:, 0].min() > 3.0


def test_packed_blob_leaves_the_domain(oracle):
    """quant=True: the window follows positions read from the packed x words
    (21-bit fixed point, |x| < 2).  A 3D blob thrown at 8 units/s leaves
    [0, 1) and ends near x = 1.6 after 240 substeps; against the oracle with
    the same storage flags (stored state as input) within the drift bound."""
    from taichi_elements_amd import MPMSolver
    dim = 3
    s = MPMSolver(res=(32, ) * dim, unbounded=True, quant=True)
    x, vel, mat = _blob(dim, 6000, [8.0, 2.0, -3.0], seed=2)
    s.read_restart(len(x), x, vel, mat, np.zeros(len(x), np.int32))
    st = s.get_state()  # the stored (rounded) state is the common input
    cfg = oracle_cfg_for(oracle, s)
    assert cfg.quant == 1
    _set_window(cfg, [-8, -8, -32], [72, 56, 64])
    ref = oracle.Substepper(cfg, st, np.float32)
    dt = s.default_dt
    for _ in range(4):
        s._run_substeps(60, dt)
        for _ in range(60):
            ref.substep(dt)
    got = s.get_state()
    errs = {f: rel_err(got[f], ref.s[f]) for f in ("x", "v", "F")}
    print("unbounded packed vs f32 oracle", {f: f"{e:.2e}" for f, e in errs.items()},
          "x range", got["x"].min(0), got["x"].max(0), flush=True)
    assert got["x"][:, 0].min() > 1.2
    for f in ("x", "v"):
        assert errs[f] <= DRIFT_TOL, (f, errs[f])


@pytest.mark.parametrize("dim", [2, 3])
def test_far_wall_stops_particles(oracle, dim):
    """Particles 5-10 cells inside the wall at -grid_size/2 + padding = -2045
    cells, moving towards it at 20 units/s (0.4 cells per substep)."""
    from taichi_elements_amd import MPMSolver
    s = MPMSolver(res=(32, ) * dim, unbounded=True)
    rng = np.random.default_rng(40 + dim)
    n = 2000
    x = cube_particles(rng, n, [-2040 / 32] + [0.4] * (dim - 1), [5 / 32] + [0.2] * (dim - 1), dim)
    vel = np.zeros((n, dim), np.float32)
    vel[:, 0] = -20.0
    mat = np.ones(n, np.int32)
    s.read_restart(n, x, vel, mat, np.zeros(n, np.int32))
    cfg = oracle_cfg_for(oracle, s)
    _set_window(cfg, [-2056] + [-8] * (dim - 1), [48] + [48] * (dim - 1))
    st = oracle.rest_state(x, vel, mat)
    ref = oracle.Substepper(cfg, st, np.float32)
    ref64 = oracle.Substepper(cfg, {k: (np.asarray(a, np.float64) if k != "material" else a)
                                    for k, a in st.items()}, np.float64)
    dt = s.default_dt
    s._run_substeps(80, dt)
    for _ in range(80):
        ref.substep(dt)
        ref64.substep(dt)
    got = s.get_state()
    errs = {f: rel_err(got[f], ref.s[f]) for f in ("x", "v")}
    e64 = {f: rel_err(got[f], ref64.s[f]) for f in ("x", "v")}
    e32_64 = rel_err(ref.s["v"], ref64.s["v"])
    print(f"far wall {dim}D vs f32 oracle", {f: f"{e:.2e}" for f, e in errs.items()},
          "vs f64", {f: f"{e:.2e}" for f, e in e64.items()}, f"f32 oracle v vs f64 {e32_64:.2e}",
          flush=True)
    assert errs["x"] <= DRIFT_TOL
    # the elastic bounce amplifies rounding: the float32 oracle's own v is
    # ~1e-2 away from the float64 one; the HIP path must be as close
    assert e64["v"] <= 3 * e32_64 + 1e-6
    # 32 cells of free flight would have taken them 22+ cells past the wall
    assert got["x"][:, 0].min() * 32 > -2048


def test_window_limit_and_refusals():
    from taichi_elements_amd import MPMSolver, MPMSolverBatch
    from taichi_elements_amd._lib import DomainError
    s = MPMSolver(res=(32, ) * 3, unbounded=True)
    # opposite corners 3000 cells apart: 750^3 blocks > 2^24
    x = np.array([[-45.0, -45.0, -45.0], [45.0, 45.0, 45.0]], np.float32)
    s.add_particles(x, s.material_elastic)
    with pytest.raises(DomainError):
        s.substep(s.default_dt)
    with pytest.raises(NotImplementedError):
        MPMSolverBatch([dict(res=(32, ) * 3, unbounded=True), dict(res=(32, ) * 3)])
