"""HIP graph replay of steady-state substep pairs: bitwise the same results
as kernel-by-kernel launches, across calls with host work and
synchronisations between the replays (plain, use_g2p2g and batched)."""
import numpy as np
import pytest

from tests.helpers import cube_particles, deformed_state

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _make(dim, graphs, **kw):
    from taichi_elements_amd import MPMSolver
    rng = np.random.default_rng(90 + dim)
    s = MPMSolver(res=(40, ) * dim, **kw)
    s.set_graphs(graphs)
    n = 20000
    x = cube_particles(rng, n, [0.25] * dim, [0.45] * dim, dim)
    mat = rng.integers(0, 5, n).astype(np.int32)
    # mild pre-deformation: the scene stays stable for the 74 substeps (a
    # stronger one blows up after ~50, in the oracle too)
    st = deformed_state(rng, x, mat, dim, stretch=0.02, c_scale=1.0)
    s.read_restart(n, st["x"], st["v"], st["material"], np.zeros(n, np.int32))
    s.set_state(F=st["F"], C=st["C"], Jp=st["Jp"])
    return s


def _run_calls(s, calls):
    from taichi_elements_amd import _lib
    out = []
    for k in calls:
        vmax = s._run_substeps(k, s.default_dt)
        out.append((s.get_state(), vmax.copy()))
    del _lib
    return out


@pytest.mark.parametrize("dim,mode", [(2, "plain"), (3, "plain"), (3, "g2p2g")])
def test_graph_replay_is_bitwise_eager(dim, mode):
    kw = dict(use_g2p2g=True) if mode == "g2p2g" else {}
    calls = (40, 7, 25, 2)  # crosses physical reorders, odd splits
    a = _run_calls(_make(dim, True, **kw), calls)
    b = _run_calls(_make(dim, False, **kw), calls)
    for i, ((sa, va), (sb, vb)) in enumerate(zip(a, b)):
        assert np.array_equal(va, vb), f"call {i}: max velocity"
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), f"call {i}: {k}"


def test_graph_replay_in_a_batch():
    from taichi_elements_amd import MPMSolverBatch
    rng = np.random.default_rng(7)
    states = []
    for k in range(3):
        n = 3000 + 2000 * k
        x = cube_particles(rng, n, [0.3] * 3, [0.3] * 3, 3)
        states.append((x, rng.uniform(-1, 1, (n, 3)).astype(np.float32),
                       rng.integers(0, 4, n).astype(np.int32)))
    results = []
    for graphs in (True, False):
        b = MPMSolverBatch([dict(res=(32 + 16 * k, ) * 3) for k in range(3)])
        b.set_tuning(0, 0)
        from taichi_elements_amd import _lib
        _lib.check(_lib.lib().mpm_set_graphs(b._handle, int(graphs)))
        for sc, (x, v, m) in zip(b, states):
            sc.read_restart(len(x), x, v, m, np.zeros(len(x), np.int32))
        got = []
        for n in (20, 13):
            vm = b.substeps(n)
            got.append((b._get_state_all(), vm.copy()))
        results.append(got)
    for (sa, va), (sb, vb) in zip(*results):
        assert np.array_equal(va, vb)
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), k
