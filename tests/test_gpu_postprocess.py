"""Callables in grid_postprocess: the reference's extension point
(engine/mpm_solver.py:300, 794-795 calls every entry as p(t, dt, grid_v)
between grid_normalization_and_gravity and g2p).

  * a callable that changes nothing leaves every particle field bitwise as
    the fused substep does (2D/3D, plain and use_g2p2g; unbounded, packed
    storage), so splitting the substep around the callables changes nothing
    by itself;
  * a callable zeroing the velocity of the nodes below a plane gives bitwise
    the state of the built-in sticky surface collider on that plane (whose
    oracle parity is tested elsewhere): node coordinates and the write-back
    reach the right nodes;
  * the nodes handed over are those the grid keeps (mpm_read_grid of the
    same substep: masses and velocities bitwise), their masses add up to
    the particles' mass, and t runs as the reference's float64 frame loop;
  * a linear velocity field v_i = A x_i + b written by a callable reaches the
    particles as g2p's closed form says: quadratic B-spline weights
    reproduce linear functions and their second moment is dx^2 / 4, so
    v_p = A x_p + b and C_p = A (mpm_solver.py:694-724);
  * errors: a callable that leaves a wrong-shaped v raises ValueError and
    the solver runs on; other calls on the solver inside a callable raise
    ValueError;
  * the split-substep C entry points refuse calls out of order.
"""
import numpy as np
import pytest

from tests.helpers import cube_particles

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _scene(dim, seed=3, g2p2g=False, vy=0.0, **kw):
    from taichi_elements_amd import MPMSolver
    s = MPMSolver(res=(32, ) * dim, use_g2p2g=g2p2g, **kw)
    rng = np.random.default_rng(seed)
    n = 3000 if dim == 2 else 8000
    x = cube_particles(rng, n, [0.3] * dim, [0.3] * dim, dim)
    v = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
    v[:, 1] += vy
    mat = rng.integers(0, 4, n).astype(np.int32)
    s.read_restart(n, x, v, mat, np.zeros(n, np.int32))
    return s


def _same_state(a, b):
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


@pytest.mark.parametrize("g2p2g", [False, True])
@pytest.mark.parametrize("dim", [2, 3])
def test_noop_callable_is_the_fused_substep(dim, g2p2g):
    a = _scene(dim, g2p2g=g2p2g)
    b = _scene(dim, g2p2g=g2p2g)
    calls = []
    b.grid_postprocess.append(lambda t, dt, g: calls.append(len(g)))
    for _ in range(3):
        a.step(2e-3)
        b.step(2e-3)
    assert len(calls) == b.total_substeps and min(calls) > 0
    _same_state(a, b)
    assert a.all_time_max_velocity == b.all_time_max_velocity


@pytest.mark.parametrize("opts", [dict(unbounded=True), dict(quant=True),
                                  dict(quant=True, g2p2g=True)])
def test_noop_callable_with_storage_and_window_options(opts):
    a = _scene(3, **opts)
    b = _scene(3, **opts)
    calls = []
    b.grid_postprocess.append(lambda t, dt, g: calls.append(len(g)))
    for _ in range(3):
        a.step(2e-3)
        b.step(2e-3)
    assert len(calls) == b.total_substeps and min(calls) > 0
    _same_state(a, b)


@pytest.mark.parametrize("dim", [2, 3])
def test_sticky_floor_callable_equals_the_builtin_collider(dim):
    k = 11  # plane between node rows k and k + 1
    a = _scene(dim, vy=-3.0)
    b = _scene(dim, vy=-3.0)
    y0 = (k + 0.5) * a.dx
    point = [0.0] * dim
    point[1] = y0
    normal = [0.0] * dim
    normal[1] = 1.0
    a.add_surface_collider(point, normal, surface=a.surface_sticky)
    zeroed = []

    def floor(t, dt, g):
        below = g.I[:, 1] <= k
        zeroed.append(int(below.sum()))
        g.v[below] = 0.0

    b.grid_postprocess.append(floor)
    for _ in range(4):
        a.step(2e-3)
        b.step(2e-3)
    assert max(zeroed) > 0  # the particles reached the plane
    _same_state(a, b)


@pytest.mark.parametrize("g2p2g", [False, True])
@pytest.mark.parametrize("dim", [2, 3])
def test_nodes_are_the_grid_and_time_runs_as_the_reference(dim, g2p2g):
    s = _scene(dim, g2p2g=g2p2g)
    rec = []

    def probe(t, dt, g):
        rec.append((t, dt, g.I.cpu().numpy(), g.v.cpu().numpy(), g.m.cpu().numpy()))

    s.step(1e-3)
    s.grid_postprocess.append(probe)
    t0 = s.t
    s.step(2e-3)
    count = s.total_substeps
    # t of each call: the reference's float64 loop (self.t += dt after each
    # substep's post-processes)
    t = t0
    dt = rec[0][1]
    for r in rec:
        assert r[0] == t and r[1] == dt
        t += dt
    assert s.t == t and len(rec) > 1
    assert count > len(rec)
    I, v, m = rec[-1][2], rec[-1][3], rec[-1][4]
    n = s.n_particles[None]
    np.testing.assert_allclose(m.sum(dtype=np.float64), n * s.p_mass, rtol=1e-5)
    assert len(np.unique(I, axis=0)) == len(I)
    lo = I.min(0)
    span = I.max(0) - lo + 1
    gm, gv = s.read_grid(lo.tolist(), span.tolist())
    idx = tuple((I - lo)[:, a] for a in reversed(range(dim)))
    np.testing.assert_array_equal(gm[idx], m)
    np.testing.assert_array_equal(gv[idx], v)
    # nodes without mass carry no velocity
    assert np.all(v[m == 0] == 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_grid_field_reaches_the_particles(dim):
    s = _scene(dim, seed=9)
    rng = np.random.default_rng(5)
    A = rng.uniform(-2, 2, (dim, dim)).astype(np.float32)
    b = rng.uniform(-1, 1, dim).astype(np.float32)
    At = torch.tensor(A, device="cuda")
    bt = torch.tensor(b, device="cuda")

    def linear(t, dt, g):
        g.v.copy_(g.positions() @ At.T + bt)

    s.step(1e-3)
    x0 = s.get_state()["x"].astype(np.float64)
    s.grid_postprocess.append(linear)
    dt = s.default_dt
    s.substep(dt)
    st = s.get_state()
    v_expect = x0 @ A.T.astype(np.float64) + b
    scale = np.abs(v_expect).max()
    assert np.abs(st["v"] - v_expect).max() <= 1e-5 * scale
    C = st["C"].reshape(-1, dim, dim)
    assert np.abs(C - A).max() <= 1e-4 * np.abs(A).max()
    np.testing.assert_allclose(st["x"], x0 + dt * st["v"].astype(np.float64), rtol=0,
                               atol=1e-6)


def test_callable_errors():
    s = _scene(3)
    s.grid_postprocess.append(lambda t, dt, g: setattr(g, "v", g.v[:-1]))
    with pytest.raises(ValueError, match="grid_v.v"):
        s.step(1e-3)
    s.grid_postprocess.pop()

    def nested(t, dt, g):
        s.particle_info()

    s.grid_postprocess.append(nested)
    with pytest.raises(ValueError, match="mpm_substep_end"):
        s.step(1e-3)
    s.grid_postprocess.pop()
    s.step(1e-3)  # the solver runs on
    assert np.isfinite(s.get_state()["x"]).all()


def test_batch_refuses_callables():
    from taichi_elements_amd import MPMSolverBatch
    b = MPMSolverBatch([dict(res=(32, 32, 32)), dict(res=(32, 32, 32))])
    for sc in b:
        sc.add_cube((0.3, 0.3, 0.3), (0.2, 0.2, 0.2), sc.material_water)
    b[1].grid_postprocess.append(lambda t, dt, g: None)
    with pytest.raises(NotImplementedError):
        b.substeps(1)
    b[1].grid_postprocess.pop()
    b.substeps(1)
    b.close()


def test_split_substep_entry_points_check_their_order():
    """mpm_substep_begin / mpm_substep_end through the C ABI: a second begin,
    an end or a node query without a begin, a gather with the wrong node
    count and other calls on the handle in between are MPM_EINVAL; a
    begin / gather / scatter / end cycle then runs."""
    from taichi_elements_amd import _lib
    L = _lib.lib()
    s = _scene(3)
    h = s._h()
    n = _lib.ctypes.c_int64()
    assert L.mpm_substep_end(h) == _lib.MPM_EINVAL
    assert L.mpm_grid_node_count(h, _lib.ctypes.byref(n)) == _lib.MPM_EINVAL
    dt = s.default_dt
    assert L.mpm_substep_begin(h, dt) == _lib.MPM_OK
    assert L.mpm_substep_begin(h, dt) == _lib.MPM_EINVAL
    assert L.mpm_substeps(h, 1, dt, 0.0) == _lib.MPM_EINVAL
    assert L.mpm_grid_node_count(h, _lib.ctypes.byref(n)) == _lib.MPM_OK and n.value > 0
    I = torch.empty((n.value, 3), dtype=torch.int32, device="cuda")
    v = torch.empty((n.value, 3), dtype=torch.float32, device="cuda")
    m = torch.empty((n.value, ), dtype=torch.float32, device="cuda")
    ptr = lambda t: _lib.ctypes.c_void_p(t.data_ptr())
    torch.cuda.synchronize()
    assert L.mpm_grid_gather(h, ptr(I), ptr(v), ptr(m), n.value - 1) == _lib.MPM_EINVAL
    assert L.mpm_grid_gather(h, ptr(I), ptr(v), ptr(m), n.value) == _lib.MPM_OK
    assert L.mpm_grid_scatter(h, ptr(v), n.value) == _lib.MPM_OK
    assert L.mpm_substep_end(h) == _lib.MPM_OK
    assert L.mpm_substep_end(h) == _lib.MPM_EINVAL
    assert L.mpm_sync(h) == _lib.MPM_OK
    assert abs(float(m.sum()) - s.n_particles[None] * s.p_mass) <= 1e-5 * s.n_particles[None] * s.p_mass
