"""Benchmark of the MLS-MPM substep on one MI355X.

Workload (the 3D four-material plain configuration): dense 128^3 grid,
dx = 1/128, 1,000,000 float32 particles in four seeded uniform-random cubes
of side 0.2 (water, fixed-corotated elastic, snow, Drucker-Prager sand),
initial velocities uniform in [-1, 1]^3, E = 1e6, nu = 0.2, gravity -9.8
along y, dt = default_dt = 1.5625e-4.  A "step" is one substep (binning,
p2g, grid operator, g2p) over all particles; inputs are resident in HBM
before the timed region starts.

Prints one JSON line:
  value      particle-substeps/s over the K timed substeps (wall clock,
             device synchronised on both sides)
  roofline   the dominant kernel (p2g): algorithmic bytes per launch
             (148 B per particle + 16 B per active grid node) divided by its
             mean duration from HIP events on the solver's stream, against
             MI355X's 8 TB/s; traffic = measured HBM bytes per launch from
             the committed rocprofv3 counter summary when present
  cpu_baseline  the oracle (C restatement, OpenMP) on the host cores over a
             bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_PER_CUBE = 250_000
RES = 128


def make_scene(seed=1):
    """Four cubes of side 0.2 with lower corners U[0.03, 0.77]^3, materials
    water / elastic / snow / sand, 250k particles each, v ~ U[-1, 1]^3."""
    rng = np.random.default_rng(seed)
    xs, vs, ms = [], [], []
    for mat in range(4):
        lo = rng.uniform(0.03, 0.77, 3).astype(np.float32)
        xs.append(lo + rng.random((N_PER_CUBE, 3), dtype=np.float32) * np.float32(0.2))
        vs.append(rng.uniform(-1, 1, (N_PER_CUBE, 3)).astype(np.float32))
        ms.append(np.full(N_PER_CUBE, mat, np.int32))
    return np.concatenate(xs), np.concatenate(vs), np.concatenate(ms)


def make_sparse_scene(seed=2, total=16_000_000, res=512, parts=False):
    """64 ellipsoids, centres U[0.1, 0.9]^3, semi-axes U[0.02, 0.08]^3, random
    material in {water, elastic, snow, sand}; per-ellipsoid counts from
    add_ellipsoid's rule at 8 particles per cell (mpm_solver.py:997-1005),
    scaled so that the total is exactly `total`; uniform interior sampling by
    rejection in the unit ball (mpm_solver.py:959-967); zero velocity.
    parts=True also returns the ellipsoids (centres, radii, materials,
    particle counts, in the order their particles appear)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.1, 0.9, (64, 3))
    radii = rng.uniform(0.02, 0.08, (64, 3))
    mats = rng.integers(0, 4, 64)
    counts = np.ceil(4 / 3 * np.pi * np.prod(radii * res, axis=1) * 8)
    counts = np.floor(counts * total / counts.sum()).astype(np.int64)
    counts[-1] += total - counts.sum()
    x = np.empty((total, 3), np.float32)
    m = np.empty(total, np.int32)
    at = 0
    for c, r, mat, k in zip(centres, radii, mats, counts):
        got = 0
        while got < k:
            cand = rng.random((int((k - got) * 2.0) + 64, 3), dtype=np.float32) * 2 - 1
            cand = cand[(cand * cand).sum(1) <= 1][:k - got]
            x[at + got:at + got + len(cand)] = c.astype(np.float32) + cand * r.astype(np.float32)
            got += len(cand)
        m[at:at + k] = mat
        at += k
    if parts:
        return x, np.zeros_like(x), m, dict(centres=centres, radii=radii, materials=mats,
                                            counts=counts)
    return x, np.zeros_like(x), m


def make_batched_scenes(seed=3, nscenes=64):
    """64 independent 3D scenes: N_k = round(exp(U[ln 1e4, ln 5e5])) particles,
    res_k uniform in {64, 96, 128}, 1-4 cubes of side U[0.1, 0.3] with lower
    corners U[0.05, 0.95 - side]^3 and one random material each (water /
    elastic / snow / sand), v ~ U[-1, 1]^3."""
    rng = np.random.default_rng(seed)
    scenes = []
    for _ in range(nscenes):
        n = int(round(np.exp(rng.uniform(np.log(1e4), np.log(5e5)))))
        res = int(rng.choice([64, 96, 128]))
        ncubes = int(rng.integers(1, 5))
        split = rng.multinomial(n, [1.0 / ncubes] * ncubes)
        xs, ms = [], []
        for c in range(ncubes):
            side = rng.uniform(0.1, 0.3)
            lo = rng.uniform(0.05, 0.95 - side, 3)
            xs.append((lo + rng.random((split[c], 3)) * side).astype(np.float32))
            ms.append(np.full(split[c], rng.integers(0, 4), np.int32))
        v = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        scenes.append(dict(res=res, x=np.concatenate(xs), v=v, m=np.concatenate(ms)))
    return scenes


def batched_bench(steps, warmup, autotune=True):
    """The 64-scene batch advanced by one engine (every stage one launch over
    all scenes) against a loop of plain solves of the same scenes, one
    MPMSolver per scene.  Both advance every scene by `steps` substeps of its
    own default_dt from the same state."""
    from taichi_elements_amd import MPMSolver, MPMSolverBatch
    from taichi_elements_amd._lib import check, fptr
    from taichi_elements_amd._lib import lib as _l
    L = _l()
    specs = make_batched_scenes()
    total = sum(len(sp["x"]) for sp in specs)
    b = MPMSolverBatch([dict(res=(sp["res"], ) * 3) for sp in specs])
    for sc, sp in zip(b, specs):
        sc.read_restart(len(sp["x"]), sp["x"], sp["v"], sp["m"], np.zeros(len(sp["x"]), np.int32))
    dts = np.ascontiguousarray(np.asarray(b.default_dts(), np.float32))
    tuning = None
    if autotune:
        # tune a copy, so that the measured batch and the separate solvers
        # start from the same particle order and stay comparable bit for bit
        probe = MPMSolverBatch([dict(res=(sp["res"], ) * 3) for sp in specs])
        for sc, sp in zip(probe, specs):
            sc.read_restart(len(sp["x"]), sp["x"], sp["v"], sp["m"],
                            np.zeros(len(sp["x"]), np.int32))
        probe.autotune(substeps=3)
        chunk, groups = probe.tuning()
        waves = probe.g2p_waves()
        probe.close()
        del probe
        b.set_tuning(chunk, groups, waves)
        tuning = dict(chunk=chunk, groups=groups, g2p_waves=waves)
    solos = []
    for sp in specs:
        s = MPMSolver(res=(sp["res"], ) * 3)
        s.read_restart(len(sp["x"]), sp["x"], sp["v"], sp["m"], np.zeros(len(sp["x"]), np.int32))
        if tuning:
            s.set_tuning(tuning["chunk"], tuning["groups"], tuning["g2p_waves"])
        solos.append(s)
    print(f"[bench] batched: {len(specs)} scenes, {total} particles", file=sys.stderr, flush=True)

    def run_batch(k):
        check(L.mpm_substeps_batch(b._handle, k, fptr(dts)))
        check(L.mpm_sync(b._handle))

    def run_loop(k):
        # one scene after the other, each solve finished before the next starts
        for s in solos:
            check(L.mpm_substeps(s._handle, k, s.default_dt, 0.0))
            check(L.mpm_sync(s._handle))

    def run_loop_async(k):
        # every solve enqueued on its own stream, one wait at the end
        for s in solos:
            check(L.mpm_substeps(s._handle, k, s.default_dt, 0.0))
        for s in solos:
            check(L.mpm_sync(s._handle))

    res = {}
    identical = None
    for name, fn in (("batch", run_batch), ("loop", run_loop), ("loop_async", run_loop_async)):
        fn(warmup)
        t0 = time.perf_counter()
        fn(steps)
        res[name] = time.perf_counter() - t0
        if name == "loop":
            # batch and loop have now run the same substeps from the same state
            got = b.particle_info()
            identical = all(np.array_equal(g["position"], s.particle_info()["position"]) and
                            np.array_equal(g["velocity"], s.particle_info()["velocity"])
                            for g, s in zip(got, solos))
    check(L.mpm_set_timing(b._handle, 1))
    run_batch(steps)
    names, calls, ms = b.kernel_times()
    check(L.mpm_set_timing(b._handle, 0))
    kern = {nm: round(1000 * t / max(c, 1), 2) for nm, c, t in zip(names, calls, ms)}
    counts = b.last_counts()
    out = dict(scenes=len(specs), particles=total, substeps=steps, tuning=tuning,
               kernels_us=kern, active_grid_blocks=counts["active_blocks"],
               chunks=counts["chunks"],
               value=total * steps / res["batch"],
               ms_per_substep=1000 * res["batch"] / steps,
               loop_value=total * steps / res["loop"],
               loop_ms_per_substep=1000 * res["loop"] / steps,
               loop_async_value=total * steps / res["loop_async"],
               speedup_vs_loop=res["loop"] / res["batch"],
               speedup_vs_loop_async=res["loop_async"] / res["batch"],
               identical_to_loop=identical,
               unit="particle-substeps/s",
               workload="64 independent 3D scenes, N log-uniform in [1e4, 5e5], res in "
                        "{64, 96, 128}, 1-4 random cubes of random material, per-scene "
                        "default_dt")
    b.close()
    del solos
    return out


def make_diff_scene(seed=4, n_el=100_000, n_wa=100_000):
    """The differentiable workload's particles: an elastic cube (lower corner
    (0.25, 0.2, 0.25), side 0.23) with E ~ U[5e5, 2e6] per particle and a
    water cube beside it (lower corner (0.48, 0.2, 0.25)), v ~ U[-0.5, 0.5]^3;
    the target of the loss is the elastic body's centre of mass moved by
    U[-0.05, 0.05]^3.  Returns x, v, material, E, target, n_el."""
    rng = np.random.default_rng(seed)
    x_el = (np.array([0.25, 0.2, 0.25], np.float32)
            + rng.random((n_el, 3), dtype=np.float32) * np.float32(0.23))
    x_wa = (np.array([0.48, 0.2, 0.25], np.float32)
            + rng.random((n_wa, 3), dtype=np.float32) * np.float32(0.23))
    x = np.concatenate([x_el, x_wa])
    mat = np.concatenate([np.full(n_el, 1, np.int32), np.zeros(n_wa, np.int32)])
    v = rng.uniform(-0.5, 0.5, x.shape).astype(np.float32)
    E = np.concatenate([rng.uniform(5e5, 2e6, n_el), np.full(n_wa, 1e6)]).astype(np.float32)
    target = x_el.mean(0) + rng.uniform(-0.05, 0.05, 3).astype(np.float32)
    return x, v, mat, E, target, n_el


def fwd_bwd_bench(nsub=64, checkpoint_every=None, repeats=3, seed=4):
    """Differentiable workload: 128^3 grid, an elastic cube of 100k particles
    with per-particle E ~ U[5e5, 2e6] and a water cube of 100k particles
    pressed against one of its faces (so that the elastic body's motion
    depends on its stiffness through the contact), v ~ U[-0.5, 0.5]^3,
    64 substeps at default_dt (every substep's input kept: DiffMPMSolver.run's
    default at this size; also timed with checkpoints every 8); loss = squared
    distance of the elastic body's centre of mass from a seeded target;
    gradients with respect to the initial velocities, E and nu.  Times the
    forward alone (no autograd) and forward + backward."""
    import torch  # before the engine library loads: one HIP runtime
    from taichi_elements_amd.diff import DiffMPMSolver
    x, v, mat, E, target, n_el = make_diff_scene(seed)
    target = torch.tensor(target, device="cuda")
    ds = DiffMPMSolver((128, 128, 128))
    ds.setup(mat, x)
    s0 = ds.state(x, v)
    E_t = torch.tensor(E, device="cuda")
    nu_t = torch.tensor(0.2, device="cuda")

    def forward_only():
        with torch.no_grad():
            s = ds.run(s0, E_t, nu_t, nsub, checkpoint_every=checkpoint_every)
            return ((ds.positions(s)[:n_el].mean(0) - target) ** 2).sum()

    def forward_backward():
        v_t = s0[3:6].clone().requires_grad_(True)
        e_t = E_t.clone().requires_grad_(True)
        n_t = nu_t.clone().requires_grad_(True)
        s = torch.cat([s0[:3], v_t, s0[6:]], 0)
        s = ds.run(s, e_t, n_t, nsub, checkpoint_every=checkpoint_every)
        loss = ((ds.positions(s)[:n_el].mean(0) - target) ** 2).sum()
        loss.backward()
        return loss, v_t.grad, e_t.grad, n_t.grad

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        return best, out

    t_f, loss = timed(forward_only)
    t_fb, (loss2, gv, ge, gn) = timed(forward_backward)
    chosen = ds.last_checkpoint_every  # 64 states of 200k particles: ~1.4 GB, all kept
    # the same with checkpoints every 8 substeps (segments recomputed)
    checkpoint_every = 8
    t_fb8, (_, gv8, _, _) = timed(forward_backward)
    ds.check()
    n = x.shape[0]
    return dict(particles=n, substeps=nsub, checkpoint_every=chosen,
                fwd_ms=1000 * t_f, fwd_bwd_ms=1000 * t_fb, ratio=t_fb / t_f,
                fwd_value=n * nsub / t_f, fwd_bwd_value=n * nsub / t_fb,
                checkpoints_every_8=dict(fwd_bwd_ms=1000 * t_fb8, ratio=t_fb8 / t_f,
                                         same_gradients=bool(torch.equal(gv8, gv))),
                unit="particle-substeps/s", loss=float(loss),
                grad_norms=dict(v0=float(gv.norm()), E=float(ge.norm()), nu=float(gn)),
                workload="128^3 grid, elastic cube 100k (E ~ U[5e5, 2e6]) + water cube 100k, "
                         "64 substeps, every substep's input kept (also: checkpoints every 8), "
                         "loss = |COM_elastic - target|^2, gradients w.r.t. v0, E, nu")


def variant_bench(x, v, m, steps, warmup, autotune, workload, **solver_kw):
    """The same 1M-particle workload with other solver options (use_g2p2g:
    one fused transfer per substep, grid velocity clamp at 0.9 cells; quant:
    packed particle storage; support_plasticity)."""
    from taichi_elements_amd import MPMSolver
    from taichi_elements_amd._lib import check
    from taichi_elements_amd._lib import lib as _l
    L = _l()
    s = MPMSolver(res=(RES, RES, RES), **solver_kw)
    s.read_restart(len(x), x, v, m, np.zeros(len(x), np.int32))
    if autotune:
        s.autotune(substeps=3)
    dt = s.default_dt

    def run(k):
        check(L.mpm_substeps(s._handle, k, dt, 0.0))
        check(L.mpm_sync(s._handle))

    run(warmup)
    t0 = time.perf_counter()
    run(steps)
    el = time.perf_counter() - t0
    check(L.mpm_set_timing(s._handle, 1))
    run(steps)
    names, calls, ms = s.kernel_times()
    check(L.mpm_set_timing(s._handle, 0))
    return dict(value=len(x) * steps / el, ms_per_substep=1000 * el / steps,
                unit="particle-substeps/s", tuning=dict(zip(("chunk", "groups"), s.tuning()), g2p_waves=s.g2p_waves()),
                bytes_per_particle=s.particle_bytes(),
                kernels_us={nm: round(1000 * t / max(c, 1), 2) for nm, c, t in zip(names, calls, ms)},
                workload=workload)


def g2p2g_bench(x, v, m, steps, warmup, autotune=True):
    return variant_bench(x, v, m, steps, warmup, autotune,
                         "3D four-material 1M particles, 128^3, use_g2p2g=True", use_g2p2g=True)


def quant_bench(x, v, m, steps, warmup, autotune=True):
    """Packed particle storage on the p2g / g2p path, and the reference's
    large-scale combination (quant, use_g2p2g, support_plasticity=False)."""
    out = variant_bench(x, v, m, steps, warmup, autotune,
                        "3D four-material 1M particles, 128^3, quant=True", quant=True)
    out["g2p2g_no_plasticity"] = variant_bench(
        x, v, m, steps, warmup, autotune,
        "3D four-material 1M particles, 128^3, quant=True, use_g2p2g=True, "
        "support_plasticity=False", quant=True, use_g2p2g=True, support_plasticity=False)
    return out


def small2d_bench(steps=100, repeats=3):
    """The reference's own small case: 2D, 128^2, 20,000 water particles in
    [0.2, 0.5] x [0.4, 0.8] (seed 0), 100 substeps at default_dt -- timed
    with HIP graph replay of substep pairs and kernel by kernel."""
    from taichi_elements_amd import MPMSolver
    from taichi_elements_amd._lib import check
    from taichi_elements_amd._lib import lib as _l
    L = _l()
    rng = np.random.default_rng(0)
    x = (np.array([0.2, 0.4], np.float32)
         + rng.random((20000, 2), dtype=np.float32) * np.array([0.3, 0.4], np.float32))
    out = {}
    for graphs in (True, False):
        s = MPMSolver(res=(128, 128))
        s.set_graphs(graphs)
        s.read_restart(len(x), x, np.zeros_like(x), np.zeros(len(x), np.int32),
                       np.zeros(len(x), np.int32))
        dt = s.default_dt
        best = None
        for _ in range(repeats + 1):
            t0 = time.perf_counter()
            check(L.mpm_substeps(s._handle, steps, dt, 0.0))
            check(L.mpm_sync(s._handle))
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        out["graphs" if graphs else "kernel_by_kernel"] = 1e6 * best / steps
        del s
    return dict(us_per_substep=out, speedup=out["kernel_by_kernel"] / out["graphs"],
                value=20000 / (out["graphs"] * 1e-6), unit="particle-substeps/s",
                workload="2D water, 128^2, 20k particles, default_dt, 100 substeps per call")


def algorithmic_bytes(n_particles, n_nodes):
    """Bytes a substep must move in the reference's field layout (float32):
    per particle 16 (binning) + 148 (p2g) + 80 (g2p), per active grid node
    (4 channels) 16 (p2g write) + 28 (grid op) + 12 (g2p read)."""
    return dict(substep=244 * n_particles + 56 * n_nodes,
                p2g=148 * n_particles + 16 * n_nodes)


def cpu_baseline(x, v, m, dt, seconds):
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cfg = O.make_cfg((RES, RES, RES), grid_lo=[-8] * 3, grid_n=[RES + 16] * 3)
    st = O.rest_state(x, v, m)
    run = O.Substepper(cfg, st, np.float32, nthreads=threads)
    run.substep(dt)  # first touch of the grid buffers
    t0 = time.perf_counter()
    k = 0
    while True:
        run.substep(dt)
        k += 1
        el = time.perf_counter() - t0
        if el >= seconds or k >= 100:
            break
    # the same restatement on one thread, over a few substeps
    one = O.Substepper(cfg, st, np.float32, nthreads=1)
    t1 = time.perf_counter()
    k1 = 0
    while True:
        one.substep(dt)
        k1 += 1
        el1 = time.perf_counter() - t1
        if el1 >= seconds / 4 or k1 >= 10:
            break
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=len(x) * k / el, unit="particle-substeps/s", cores=threads, kind="port",
                sample=f"{k} substeps of the full 1M-particle 128^3 workload "
                       f"(oracle C restatement, OpenMP atomics, float32), {el:.1f} s",
                single_thread=dict(value=len(x) * k1 / el1, substeps=k1, seconds=round(el1, 2)),
                host=dict(cpu_model=model, nproc=os.cpu_count()))


def copy_calibration(nbytes=256 << 20):
    """One device-to-device hipMemcpy of nbytes (reads nbytes, writes nbytes)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def measured_traffic(config):
    """HBM bytes per p2g launch of the workload `config` (dense1m or
    sparse16m) from the committed rocprofv3 counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "p2g_3d_1M_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        prof = json.load(f)
    if config == "sparse16m":
        prof = prof.get("sparse16m", {})
    return prof.get("p2g_hbm_bytes_per_launch")


def plain_leg(config, x, v, m, steps, warmup, autotune, reorder_interval=0, chunk=0):
    """The plain substep (binning, p2g, grid operator, g2p) of one workload:
    dense1m (128^3, 1M particles in four cubes) or sparse16m (512^3 window,
    16M particles in 64 ellipsoids, one slip plane at y = 0.05).  Wall-clock
    particle-substeps/s over `steps` substeps, then the same substeps with
    HIP events around each kernel on the solver's stream for the roofline."""
    from taichi_elements_amd import MPMSolver
    from taichi_elements_amd._lib import check
    from taichi_elements_amd._lib import lib as _l
    L = _l()
    res = 512 if config == "sparse16m" else RES
    n = len(x)
    s = MPMSolver(res=(res, res, res), reorder_interval=reorder_interval, chunk=chunk)
    if config == "sparse16m":
        s.add_surface_collider((0.0, 0.05, 0.0), (0.0, 1.0, 0.0), surface=s.surface_slip)
    s.read_restart(n, x, v, m, np.zeros(n, np.int32))
    dt = s.default_dt
    tuning = None
    if autotune and not chunk:
        cand = s.autotune(substeps=3)
        tuning = dict(chosen=dict(zip(("chunk", "groups"), s.tuning()), g2p_waves=s.g2p_waves()),
                      ms_per_substep={f"{c}/{g}": round(t, 4) for (c, g), t in cand.items()},
                      g2p_waves_ms={str(w): round(t, 4)
                                    for w, t in getattr(s, "last_autotune_g2p_waves", {}).items()})
        print(f"[bench] {config}: autotuned launch shape {tuning['chosen']}", file=sys.stderr,
              flush=True)

    def run(k):
        check(L.mpm_substeps(s._handle, k, dt, 0.0))
        check(L.mpm_sync(s._handle))

    print(f"[bench] {config}: {n} particles; warmup {warmup}", file=sys.stderr, flush=True)
    run(warmup)
    t0 = time.perf_counter()
    run(steps)
    t1 = time.perf_counter()
    ms_per_step = 1000 * (t1 - t0) / steps
    counts = s.last_counts()
    n_nodes = 64 * counts["active_blocks"]
    print(f"[bench] {config}: timed {steps} substeps: {ms_per_step:.4f} ms/substep",
          file=sys.stderr, flush=True)
    check(L.mpm_set_timing(s._handle, 1))
    t2 = time.perf_counter()
    run(steps)
    t3 = time.perf_counter()
    names, calls, ms = s.kernel_times()
    check(L.mpm_set_timing(s._handle, 0))
    kern = {nm: dict(calls=c, avg_us=1000 * t / max(c, 1)) for nm, c, t in zip(names, calls, ms)}
    ab = algorithmic_bytes(counts["valid"], n_nodes)
    achieved = ab["p2g"] / (kern["k_p2g"]["avg_us"] * 1e-6) / 1e9
    substep_gbs = ab["substep"] / (ms_per_step * 1e-3) / 1e9
    if config == "sparse16m":
        data = "synthetic (seeded numpy: 64 ellipsoids, 16M particles, at rest)"
        workload = ("3D sparse large domain, 512^3 window of 4^3 blocks, 16M particles in 64 "
                    "random ellipsoids (8 per cell), plane collider, dt=default_dt")
    else:
        data = "synthetic (seeded numpy: 4 cubes x 250k particles, v ~ U[-1,1]^3)"
        workload = ("3D four-material plain substep, 128^3 dense grid, dx=1/128, 1M particles "
                    "(water/elastic/snow/sand cubes of side 0.2), dt=default_dt")
    out = dict(
        value=n * steps / (t1 - t0), unit="particle-substeps/s", ms_per_substep=ms_per_step,
        steps=steps, dt=dt, data=data,
        config={"workload": workload, "res": res, "particles": n,
                "active_grid_blocks": counts["active_blocks"],
                "particle_blocks": counts["particle_blocks"], "chunks": counts["chunks"]},
        roofline={"bound": "hbm", "kernel": "k_p2g", "achieved": achieved,
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                  "traffic": measured_traffic(config),
                  "algorithmic_bytes_per_launch": ab["p2g"]},
        substep_hbm={"algorithmic_bytes": ab["substep"], "achieved_GBs": substep_gbs,
                     "frac": substep_gbs / HBM_PEAK_GBS},
        kernels_us={k: round(v_["avg_us"], 2) for k, v_ in kern.items()},
        timed_with_events_ms_per_step=1000 * (t3 - t2) / steps,
        tuning=tuning)
    del s
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reorder-interval", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--config", choices=["dense1m", "sparse16m", "fwdbwd200k", "g2p2g1m",
                                         "batched64", "quant1m"], default="dense1m",
                    help="dense1m: the 3D four-material 128^3 workload (default, the "
                         "headline line); sparse16m: 512^3 window, 16M particles in 64 "
                         "ellipsoids with a plane collider; fwdbwd200k / g2p2g1m / "
                         "batched64 / quant1m: only the differentiable, use_g2p2g, 64-scene "
                         "batched or packed-storage measurement")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of autotuning it on the scene")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 64-scene batched measurement")
    ap.add_argument("--no-small2d", action="store_true",
                    help="skip the 2D reference-sized scene (graphs vs kernel by kernel)")
    ap.add_argument("--no-quant", action="store_true",
                    help="skip the packed-storage (quant=True) legs")
    ap.add_argument("--no-g2p2g", action="store_true",
                    help="skip the use_g2p2g measurement")
    ap.add_argument("--no-fwd-bwd", action="store_true",
                    help="skip the differentiable (forward + backward) measurement")
    ap.add_argument("--no-sparse", action="store_true",
                    help="skip the sparse 512^3 / 16M-particle measurement")
    ap.add_argument("--sparse-steps", type=int, default=50,
                    help="timed substeps of the sparse leg (its configuration runs 50)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy (hipMemcpy) so that "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE can be scaled by a kernel of "
                         "known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the substep engine "
              "is single-GPU and its sharding is a design note, not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    if args.config in ("g2p2g1m", "batched64", "quant1m"):
        from taichi_elements_amd import mpm_solver
        mpm_solver.PRINT_PROGRESS = False
        if args.config in ("g2p2g1m", "quant1m"):
            x, v, m = make_scene()
            fn = g2p2g_bench if args.config == "g2p2g1m" else quant_bench
            leg = fn(x, v, m, args.steps, args.warmup, not args.no_autotune)
            data = "synthetic (seeded numpy: 4 cubes x 250k particles, v ~ U[-1,1]^3)"
        else:
            leg = batched_bench(args.steps, args.warmup, not args.no_autotune)
            data = "synthetic (seeded numpy: 64 scenes of 1-4 random cubes)"
        print(json.dumps({
            "metric": "particle-substeps/s (3D 1M, 4-material): plain/batched/fwd+bwd; % HBM peak",
            "value": leg["value"], "unit": "particle-substeps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": leg["ms_per_substep"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": data, "config": {"workload": leg["workload"]}, args.config: leg}), flush=True)
        return
    if args.config == "fwdbwd200k":
        import torch  # noqa: F401  (before the engine library loads)
        fb = fwd_bwd_bench()
        print(json.dumps({
            "metric": "particle-substeps/s (3D 1M, 4-material): plain/batched/fwd+bwd; % HBM peak",
            "value": fb["fwd_bwd_value"], "unit": "particle-substeps/s", "n_gpus": 1,
            "steps": fb["substeps"], "warmup": 1, "ms_per_step": fb["fwd_bwd_ms"] / fb["substeps"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded numpy: elastic + water cube, 100k particles each)",
            "config": {"workload": fb["workload"]}, "fwd_bwd": fb}), flush=True)
        return
    if not args.no_fwd_bwd and args.config == "dense1m":
        # torch (used by the differentiable leg) must be imported before the
        # engine library loads, so that the process runs one HIP runtime
        import torch  # noqa: F401
    from taichi_elements_amd import mpm_solver
    mpm_solver.PRINT_PROGRESS = False
    if args.config == "sparse16m":
        x, v, m = make_sparse_scene()
    else:
        x, v, m = make_scene()
    leg = plain_leg(args.config, x, v, m, args.steps, args.warmup, not args.no_autotune,
                    args.reorder_interval, args.chunk)
    dt = leg.pop("dt")
    out = {
        "metric": "particle-substeps/s (3D 1M, 4-material): plain/batched/fwd+bwd; % HBM peak",
        "value": leg["value"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": leg["ms_per_substep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": leg.pop("data"),
        "config": leg.pop("config"),
        "roofline": leg.pop("roofline"),
        "substep_hbm": leg.pop("substep_hbm"),
        "kernels_us": leg.pop("kernels_us"),
        "timed_with_events_ms_per_step": leg.pop("timed_with_events_ms_per_step"),
        "tuning": leg.pop("tuning"),
        "sparse16m": None,
        "g2p2g": None,
        "quant": None,
        "small2d": None,
        "batched": None,
        "fwd_bwd": None,
    }
    if args.copy_calibration:
        out["copy_calibration_bytes"] = copy_calibration()
    if args.config == "dense1m" and not args.no_small2d:
        out["small2d"] = small2d_bench()
    if not args.no_g2p2g and args.config == "dense1m":
        out["g2p2g"] = g2p2g_bench(x, v, m, args.steps, args.warmup, not args.no_autotune)
    if not args.no_quant and args.config == "dense1m":
        out["quant"] = quant_bench(x, v, m, args.steps, args.warmup, not args.no_autotune)
    if not args.no_batched and args.config == "dense1m":
        out["batched"] = batched_bench(args.steps, args.warmup, not args.no_autotune)
    if not args.no_fwd_bwd and args.config == "dense1m":
        out["fwd_bwd"] = fwd_bwd_bench()
    if not args.no_sparse and args.config == "dense1m":
        # the DRAM-honest workload: its 2.5 GB of particle and grid traffic per
        # substep does not fit in the 256 MB last-level cache as the 1M one does
        print("[bench] sparse16m", file=sys.stderr, flush=True)
        xs, vs, ms = make_sparse_scene()
        sp = plain_leg("sparse16m", xs, vs, ms, args.sparse_steps, args.warmup,
                       not args.no_autotune, args.reorder_interval, args.chunk)
        sp.pop("dt")
        sp.pop("data")
        out["sparse16m"] = sp
        del xs, vs, ms
    if not args.no_cpu_baseline and args.config == "dense1m":
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(x, v, m, dt, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
