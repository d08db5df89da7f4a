"""Packed particle storage (MPMSolver(quant=True)) and
support_plasticity=False on the GPU against the oracle.

* Storage: what the engine stores is bitwise the oracle's quantisation of
  the input (x on the 2^-19 grid, v with a shared exponent, F in 16-bit
  steps of 4.1/32768); after substeps every stored value is still
  representable (quantising it again changes nothing).
* One substep (p2g path, and the first use_g2p2g substep) from the stored
  state against the oracle with the same storage flags: norm-wise relative
  1e-5 on x, v, Jp (3e-5 on C, as the float parity), 2e-5 on F (a product
  landing within float rounding of a code boundary may round to the
  neighbouring 16-bit code, 1.25e-4 apart).  On later g2p2g substeps such
  one-code F differences feed the stress and dominate: at most 2% of the F
  entries may differ, by at most two codes, x stays within 1e-5, F within
  5e-5 and v within 5e-4.
* The physical reorder moves packed words only: results are bitwise the
  same with and without it; a batch of packed scenes is bitwise its own
  solvers.
"""
import numpy as np
import pytest

from tests.helpers import cube_particles, deformed_state, oracle_cfg_for, rel_err

pytestmark = pytest.mark.gpu

TOL = {"x": 1e-5, "v": 1e-5, "Jp": 1e-5, "C": 3e-5, "F": 2e-5}


@pytest.fixture(autouse=True)
def _quiet():
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False


def _make(dim, quant=True, plastic=True, g2p2g=False, n=6000, seed=0, res=32, **kw):
    from taichi_elements_amd import MPMSolver
    rng = np.random.default_rng(500 + 10 * dim + seed)
    s = MPMSolver(res=(res, ) * dim, quant=quant, support_plasticity=plastic, use_g2p2g=g2p2g,
                  **kw)
    x = cube_particles(rng, n, [0.2] * dim, [0.5] * dim, dim)
    mat = rng.integers(0, 5, n).astype(np.int32)
    st = deformed_state(rng, x, mat, dim, stretch=0.02, c_scale=1.0)
    if g2p2g:  # water keeps its volume in F[0][0] in this mode
        water = mat == 0
        st["F"][water] = np.eye(dim, dtype=np.float32)
        st["F"][water, 0, 0] = st["Jp"][water]
        st["Jp"][water] = 1.0
    s.read_restart(n, st["x"], st["v"], st["material"], np.zeros(n, np.int32))
    s.set_state(F=st["F"], C=st["C"], Jp=st["Jp"])
    return s, st


def _representable(oracle, got):
    np.testing.assert_array_equal(oracle.quant_x(got["x"]), got["x"])
    np.testing.assert_array_equal(oracle.quant_v(got["v"]), got["v"])
    np.testing.assert_array_equal(oracle.quant_f(got["F"]), got["F"])


@pytest.mark.parametrize("dim,g2p2g,plastic", [(2, False, True), (3, False, True),
                                               (3, True, False), (2, True, True)])
def test_storage_rounds_like_the_oracle(oracle, dim, g2p2g, plastic):
    s, st = _make(dim, g2p2g=g2p2g, plastic=plastic)
    got = s.get_state()
    np.testing.assert_array_equal(got["x"], oracle.quant_x(st["x"]))
    np.testing.assert_array_equal(got["v"], oracle.quant_v(st["v"]))
    np.testing.assert_array_equal(got["F"], oracle.quant_f(st["F"]))
    if g2p2g:  # no C field in this mode
        assert not got["C"].any()
    else:
        np.testing.assert_array_equal(got["C"], st["C"])
    np.testing.assert_array_equal(got["Jp"], st["Jp"] if plastic else np.ones_like(st["Jp"]))
    info = s.particle_info()
    np.testing.assert_array_equal(info["position"], got["x"])
    # packed words (x 2, v 2, F, C unless g2p2g, Jp if plastic) + material,
    # insertion index, scene + the host-side colour
    words = 4 + (dim * dim + 1) // 2 + (0 if g2p2g else dim * dim) + (1 if plastic else 0)
    assert s.particle_bytes() == 4 * (words + 3) + 4
    assert s.particle._cell_size_bytes == s.particle_bytes()  # as the demos print it
    # a fresh particle: F = I is stored as 7992 steps on the diagonal
    s2, _ = _make(dim, n=10, g2p2g=g2p2g, plastic=plastic)
    s2.add_particles(np.full((1, dim), 0.4, np.float32), 1)
    F = s2.get_state()["F"][-1]
    np.testing.assert_array_equal(F, oracle.quant_f(np.eye(dim)))


@pytest.mark.parametrize("dim,plastic", [(2, True), (3, True), (3, False)])
def test_one_packed_substep_against_oracle(oracle, dim, plastic):
    s, _ = _make(dim, plastic=plastic)
    st = s.get_state()  # the stored (rounded) state is the common input
    dt = s.default_dt
    s.substep(dt)
    got = s.get_state()
    _representable(oracle, got)
    ref, *_ = oracle.substep(oracle_cfg_for(oracle, s), st, dt, 0.0, np.float32)
    for k in ("x", "v", "C", "F", "Jp"):
        e = rel_err(got[k], ref[k])
        print(f"packed {dim}D plastic={plastic} {k}: {e:.2e}", flush=True)
        assert e <= TOL[k], k


@pytest.mark.parametrize("dim", [2, 3])
def test_float_storage_without_plasticity_against_oracle(oracle, dim):
    s, _ = _make(dim, quant=False, plastic=False)
    st = s.get_state()
    dt = s.default_dt
    for _ in range(2):
        s.substep(dt)
    ref = st
    cfg = oracle_cfg_for(oracle, s)
    for _ in range(2):
        ref, *_ = oracle.substep(cfg, ref, dt, 0.0, np.float32)
    got = s.get_state()
    for k in ("x", "v", "C", "F", "Jp"):
        e = rel_err(got[k], ref[k])
        print(f"float {dim}D no plasticity {k}: {e:.2e}", flush=True)
        assert e <= 2 * TOL[k], k
    np.testing.assert_array_equal(got["Jp"], st["Jp"])  # never written


def test_packed_g2p2g_without_plasticity_against_oracle(oracle):
    """The reference's large-scale configuration (quant, use_g2p2g, no
    plasticity: demo/demo_quantized_simulation_letters.py:88-98)."""
    s, _ = _make(3, g2p2g=True, plastic=False, res=48)
    st = s.get_state()
    ref = oracle.G2P2GStepper(oracle.with_clamp(oracle_cfg_for(oracle, s), 0.9), st)
    dt = s.default_dt
    step = 4.1 / 32768
    for k in range(3):
        s.substep(dt)
        ref.substep(dt)
        got = s.get_state()
        _representable(oracle, got)
        errs = {f: rel_err(got[f], ref.s[f]) for f in ("x", "v", "F")}
        codes = np.abs(got["F"].astype(np.float64) - ref.s["F"]) / step
        flips = float((codes > 0.5).mean())
        print(f"packed g2p2g substep {k + 1}: " + " ".join(f"{f} {e:.2e}" for f, e in errs.items())
              + f"; F entries one code apart {flips:.2%}, max {codes.max():.1f} codes", flush=True)
        if k == 0:
            for f, e in errs.items():
                assert e <= TOL[f], (k, f)
        else:
            # a product within float rounding of a code boundary rounds to the
            # neighbouring 16-bit code (1.25e-4 apart): such F flips, not the
            # arithmetic, dominate the later differences
            assert codes.max() <= 2.5 and flips <= 0.02, (k, flips)
            assert errs["x"] <= TOL["x"] and errs["F"] <= 5e-5 and errs["v"] <= 5e-4, (k, errs)


def test_packed_reorder_is_invisible():
    states = []
    for interval in (-1, 3):
        s, _ = _make(3, reorder_interval=interval, seed=4)
        for _ in range(10):
            s.substep(s.default_dt)
        states.append(s.get_state())
    for k in states[0]:
        assert np.array_equal(states[0][k], states[1][k]), k


def test_packed_batch_matches_its_solvers():
    from taichi_elements_amd import MPMSolver, MPMSolverBatch
    rng = np.random.default_rng(8)
    scenes = [dict(res=(32, ) * 3, quant=True), dict(res=(48, ) * 3, quant=True, E_scale=2)]
    parts = [(cube_particles(rng, 3000 + 1000 * k, [0.3] * 3, [0.3] * 3, 3),
              rng.integers(0, 4, 3000 + 1000 * k).astype(np.int32)) for k in range(2)]
    b = MPMSolverBatch(scenes)
    for sc, (x, m) in zip(b, parts):
        sc.add_particles(x, m)  # one material per call in the reference; here per particle
    b.substeps(12)
    for k, (kw, (x, m)) in enumerate(zip(scenes, parts)):
        kw = dict(kw)
        solo = MPMSolver(res=kw.pop("res"), **kw)
        solo.add_particles(x, m)
        for _ in range(12):
            solo.substep(solo.default_dt)
        a, c = b[k].get_state(), solo.get_state()
        for f in a:
            assert np.array_equal(a[f], c[f]), (k, f)
    with pytest.raises(ValueError):
        MPMSolverBatch([dict(res=(32, ) * 3, quant=True), dict(res=(32, ) * 3)])


def test_packed_particle_file_and_storage_switch(oracle, tmp_path):
    from oracle import particle_io as OP
    from taichi_elements_amd import _lib
    s, _ = _make(3, n=5000)
    for _ in range(3):
        s.substep(s.default_dt)
    fn = str(tmp_path / "q.npz")
    s.write_particles(fn)
    info = s.particle_info()
    got = np.load(fn)
    ranges, xv, rgb = OP.quantize(info["position"], info["velocity"], info["color"])
    assert np.array_equal(got["x_and_v"], xv) and np.array_equal(got["ranges"], ranges)
    with pytest.raises(ValueError):  # the layout cannot change under particles
        _lib.check(_lib.lib().mpm_set_storage(s._h(), 0, 1))
