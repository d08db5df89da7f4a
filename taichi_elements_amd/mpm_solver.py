"""Drop-in ``MPMSolver`` for taichi_elements (engine/mpm_solver.py) on MI355X.

The class keeps the reference's constructor, constants, attributes and
methods (``add_cube``, ``add_ellipsoid``, ``add_particles``,
``read_restart``, ``add_surface_collider``, ``add_sphere_collider``,
``set_gravity``, ``step``, ``particle_info`` ...).  Host-side bookkeeping
(constants, seeding, the frame loop of ``step``) is Python, as in the
reference; every substep runs in hand-written HIP kernels behind the C ABI of
``libmpm_hip.so`` (include/mpm_abi.h).  No Taichi is involved.

Differences from the reference, all explicit:
  * particles are seeded on the host with numpy's PCG64 (``seed`` argument)
    because Taichi's ``ti.random`` stream cannot be reproduced;
  * the grid is a dense window of 4^d blocks around the ``res`` domain (two
    blocks of margin below, >= 4 cells above the walls) instead of Taichi's
    4096^d sparse grid; particles that leave it raise ``DomainError``;
  * ``use_g2p2g=True`` runs the fused g2p2g transfer of the reference
    (mpm_solver.py:363-485) with its grid velocity clamp;
  * ``unbounded=True`` keeps the reference's walls at +-grid_size/2
    (mpm_solver.py:143-148, 604-611); the dense window then follows the
    particles, checked every 8 substeps;
  * ``quant=True`` stores particles in packed words of the same precision
    as the reference's quantised fields (mpm_solver.py:106-136);
  * ``use_adaptive_dt=True`` reads the grid's largest velocity back before
    every substep, as the reference does (one host synchronisation per
    substep);
  * ``grid_postprocess`` lists descriptions of the built-in post-processes
    (bounding box and colliders), which run inside the grid operator kernel,
    and any callables the caller appends.  A callable is called as the
    reference calls its entries, ``p(t, dt, grid_v)`` after the grid
    operator and before g2p (mpm_solver.py:794-795), but ``grid_v`` is a
    ``GridNodes`` of torch tensors on the device (node coordinates, writable
    velocities, masses of the active blocks' nodes) instead of a Taichi
    field; the built-ins run before every callable, wherever they sit in
    the list.
"""
import math
import numbers
import time

import numpy as np

from . import _lib
from ._lib import check

# Print one '.' per substep as the reference's step() does (mpm_solver.py:760).
PRINT_PROGRESS = True


class GridNodes:
    """The grid a callable of ``grid_postprocess`` receives as ``grid_v``.

    The reference hands its post-processes the sparse grid velocity field and
    they loop over its active nodes (``for I in ti.grouped(grid_v)``,
    mpm_solver.py:618-685).  Here the same nodes -- every node of the grid
    blocks the substep's particles touch, 4^dim nodes per block as the
    reference's leaf blocks -- are torch tensors on the solver's device:

      ``I``  (n, dim) int32 node indices (node position ``I * dx``),
      ``v``  (n, dim) float32 velocities after the grid operator; the
             callables change them in place (or assign a new (n, dim)
             float32 tensor to ``v``), and g2p reads what they leave,
      ``m``  (n,) float32 masses (read only: writes are not kept).

    Nodes without mass have v = 0, as in the reference.
    """

    def __init__(self, I, v, m, dx):
        self.I = I
        self.v = v
        self.m = m
        self.dx = dx

    def __len__(self):
        return int(self.m.shape[0])

    def positions(self):
        """Node positions ``I * dx`` (float32 (n, dim))."""
        return self.I.float() * self.dx


def _tptr(t):
    return _lib.ctypes.c_void_p(t.data_ptr())


class _Field0D:
    """Mimics a Taichi 0-D field: ``obj[None]`` reads the value."""

    def __init__(self, getter, setter=None):
        self._get = getter
        self._set = setter

    def __getitem__(self, key):
        if key is not None:
            raise IndexError("0-D field: index with None")
        return self._get()

    def __setitem__(self, key, value):
        if self._set is None:
            raise TypeError("read-only field")
        if key is not None:
            raise IndexError("0-D field: index with None")
        self._set(value)


class _ParticleField:
    """Stands in for one of the reference's per-particle fields (``mpm.x``,
    ``mpm.v``, ``mpm.C``, ``mpm.F``, ``mpm.Jp``, ``mpm.material``,
    ``mpm.color``; mpm_solver.py:100-141) for code that reads or writes them
    directly: ``to_numpy()`` / ``from_numpy()`` over the solver's particles
    in insertion order (one device transfer each), ``shape``, and
    ``field[i]`` (a copy of particle i's value)."""

    def __init__(self, solver, name):
        self._solver = solver
        self._name = name

    @property
    def shape(self):
        _, rank = self._solver._FIELDS[self._name]
        return self._solver._shape(rank, self._solver.n_particles[None])

    def to_numpy(self):
        return self._solver.get_state()[self._name]

    def from_numpy(self, array):
        self._solver.set_state(**{self._name: array})

    def __getitem__(self, i):
        return self.to_numpy()[i]


class _ParticleStorage:
    """Stands in for the reference's particle SNode, whose per-particle size
    the demos print (``mpm.particle._cell_size_bytes``,
    demo/demo_3d_bunnies.py:116)."""

    def __init__(self, solver):
        self._solver = solver

    @property
    def _cell_size_bytes(self):
        return self._solver.particle_bytes()


SNAPSHOT_SAVE, SNAPSHOT_RESTORE, SNAPSHOT_RELEASE = 0, 1, 2


def _autotune(obj, handle, dt, substeps, chunks, groups, verbose, dts):
    """Shared by MPMSolver and MPMSolverBatch: time each launch shape on
    the live particles.  Four untimed substeps after each change re-bin,
    reorder the slots and capture the shape's substep-pair graph; the
    shape's time is the best of three timed runs of ``substeps``.  The
    scene is put back from a device snapshot (mpm_snapshot) before every
    candidate and at the end: the particle slots and, with use_g2p2g, the
    grid the next substep reads and which particles are new."""
    L = _lib.lib()
    if obj.n_particles_total() == 0:
        return {}
    check(L.mpm_snapshot(handle, SNAPSHOT_SAVE))
    steps = np.ascontiguousarray(np.asarray(dts(), np.float32))
    batched = len(steps) > 1 or hasattr(obj, "scenes")

    def run(k):
        if batched:
            check(L.mpm_substeps_batch(handle, k, _lib.fptr(steps)))
        else:
            check(L.mpm_substeps(handle, k, float(dt), 0.0))
        check(L.mpm_sync(handle))

    if chunks is None:
        # the fused g2p2g transfer holds a chunk in one batch of 256
        chunks = (64, 128, 256) if getattr(obj, "use_g2p2g", False) else (128, 256, 512, 1024)
    results = {}
    for ch in chunks:
        for gr in groups:
            check(L.mpm_set_tuning(handle, int(ch), int(gr)))
            check(L.mpm_snapshot(handle, SNAPSHOT_RESTORE))
            run(4)
            best = None
            for _ in range(3):
                t0 = time.perf_counter()
                run(substeps)
                el = time.perf_counter() - t0
                best = el if best is None else min(best, el)
            results[(ch, gr)] = 1000 * best / substeps
            if verbose:
                print(f"chunk {ch:4d} groups {gr:5d}: {results[(ch, gr)]:.4f} ms/substep")
    best = min(results, key=results.get)
    check(L.mpm_set_tuning(handle, int(best[0]), int(best[1])))
    # g2p's occupancy at the chosen shape (results are bitwise the same)
    waves = {}
    for w in (0, 6):
        check(L.mpm_set_g2p_waves(handle, w))
        check(L.mpm_snapshot(handle, SNAPSHOT_RESTORE))
        run(4)
        el = None
        for _ in range(3):
            t0 = time.perf_counter()
            run(substeps)
            dt_w = time.perf_counter() - t0
            el = dt_w if el is None else min(el, dt_w)
        waves[w] = 1000 * el / substeps
        if verbose:
            print(f"chunk {best[0]:4d} groups {best[1]:5d} g2p waves {w}: {waves[w]:.4f} ms/substep")
    w_best = min(waves, key=waves.get)
    check(L.mpm_set_g2p_waves(handle, w_best))
    results[best] = min(results[best], waves[w_best])
    obj.last_autotune_g2p_waves = waves
    check(L.mpm_snapshot(handle, SNAPSHOT_RESTORE))
    check(L.mpm_snapshot(handle, SNAPSHOT_RELEASE))
    return results


class MPMSolver:
    material_water = 0
    material_elastic = 1
    material_snow = 2
    material_sand = 3
    material_stationary = 4
    materials = {
        'WATER': material_water,
        'ELASTIC': material_elastic,
        'SNOW': material_snow,
        'SAND': material_sand,
        'STATIONARY': material_stationary,
    }

    surface_sticky = 0
    surface_slip = 1
    surface_separate = 2
    surfaces = {
        'STICKY': surface_sticky,
        'SLIP': surface_slip,
        'SEPARATE': surface_separate
    }

    def __init__(self,
                 res,
                 quant=False,
                 use_voxelizer=True,
                 size=1,
                 max_num_particles=2**30,
                 padding=3,
                 unbounded=False,
                 dt_scale=1,
                 E_scale=1,
                 voxelizer_super_sample=2,
                 use_g2p2g=False,
                 v_clamp_g2p2g=True,
                 use_bls=True,
                 g2p2g_allowed_cfl=0.9,
                 water_density=1.0,
                 support_plasticity=True,
                 use_adaptive_dt=False,
                 use_ggui=False,
                 use_emitter_id=False,
                 *,
                 device=0,
                 seed=0,
                 reorder_interval=0,
                 chunk=0,
                 kernel_profiler=False):
        self.dim = len(res)
        assert self.dim in (2, 3), "MPM solver supports only 2D and 3D simulations."
        if use_adaptive_dt and not g2p2g_allowed_cfl > 0:
            # dt = min(dt, cfl dx / (v + 1e-6)) would be 0 and step() would
            # never finish (mpm_solver.py:762-771)
            raise ValueError("use_adaptive_dt needs g2p2g_allowed_cfl > 0")
        self.quant = quant
        self.use_g2p2g = use_g2p2g
        self.v_clamp_g2p2g = v_clamp_g2p2g
        self.use_bls = use_bls
        self.g2p2g_allowed_cfl = g2p2g_allowed_cfl
        self.water_density = water_density
        self.grid_size = 4096
        self.use_voxelizer = use_voxelizer
        self.voxelizer = None
        self.voxelizer_super_sample = voxelizer_super_sample
        self.use_ggui = use_ggui
        self.support_plasticity = support_plasticity
        self.use_adaptive_dt = use_adaptive_dt
        self.use_emitter_id = use_emitter_id

        # Host-side constants (mpm_solver.py:79-99, 198-210).
        self.t = 0.0
        self.res = res
        self.size = size
        self.dx = size / res[0]
        self.inv_dx = 1.0 / self.dx
        self.default_dt = 2e-2 * self.dx / size * dt_scale
        self.p_vol = self.dx**self.dim
        self.p_rho = 1000
        self.p_mass = self.p_vol * self.p_rho
        self.max_num_particles = max_num_particles
        self.all_time_max_velocity = 0
        self.F_bound = 4.0
        self.padding = padding
        self.E, self.nu = 1e6 * size * E_scale, 0.2
        self.mu_0, self.lambda_0 = self.E / (2 * (1 + self.nu)), self.E * self.nu / (
            (1 + self.nu) * (1 - 2 * self.nu))
        friction_angle = math.radians(45)
        sin_phi = math.sin(friction_angle)
        self.alpha = math.sqrt(2 / 3) * 2 * sin_phi / (3 - sin_phi)
        self.total_substeps = 0
        self.unbounded = unbounded
        self.offset = tuple(-self.grid_size // 2 for _ in range(self.dim))

        self._rng = np.random.default_rng(seed)
        self._device = device
        self._reorder_interval = reorder_interval
        self._chunk = chunk
        self._kernel_profiler = kernel_profiler
        self._emitter_ids = np.zeros(0, np.int32)
        self._source_velocity = [0.0] * self.dim
        self._gravity = [0.0] * self.dim
        self._handle = None
        self.n_particles = _Field0D(self._num_particles)
        self.particle = _ParticleStorage(self)
        for name in self._FIELDS:  # mpm.x, mpm.v, ... as in the reference
            setattr(self, name, _ParticleField(self, name))
        self.gravity = _Field0D(lambda: tuple(self._gravity), self.set_gravity)
        self.grid_postprocess = []
        self.writers = []
        if self._owns_handle:
            self._create_handle()
            self._init_scene()

    # A solver that shares the handle of a batch (taichi_elements_amd.batch)
    # sets this to False and is initialised by the batch.
    _owns_handle = True

    def _init_scene(self):
        """Gravity and bounding box of a new solver (mpm_solver.py:283-302)."""
        if self.dim == 2:
            self.set_gravity((0, -9.8))
        else:
            self.set_gravity((0, -9.8, 0))
        self.add_bounding_box(self.unbounded)

    # ------------------------------------------------------------------ #
    # handle management
    # ------------------------------------------------------------------ #
    def _h(self):
        return self._handle

    def _make_config(self):
        cfg = _lib.MpmConfig()
        cfg.dim = self.dim
        for a in range(3):
            cfg.res[a] = int(self.res[a]) if a < self.dim else 1
        cfg.padding = int(self.padding)
        cfg.device = int(self._device)
        cfg.reorder_interval = int(self._reorder_interval)
        cfg.chunk = int(self._chunk)
        cfg.dx, cfg.inv_dx, cfg.p_vol, cfg.p_mass = self.dx, self.inv_dx, self.p_vol, self.p_mass
        cfg.mu_0, cfg.lambda_0, cfg.alpha = self.mu_0, self.lambda_0, self.alpha
        cfg.water_density = self.water_density
        return cfg

    def _create_handle(self):
        L = _lib.lib()
        cfg = self._make_config()
        h = _lib._P()
        check(L.mpm_create(cfg, h))
        self._handle = h
        if self.unbounded:  # mpm_solver.py:143-148
            check(L.mpm_set_unbounded(h, 1))
        if self.quant or not self.support_plasticity:  # mpm_solver.py:106-136, 216-262
            check(L.mpm_set_storage(h, int(bool(self.quant)), int(bool(self.support_plasticity))))
        if self.use_g2p2g:
            check(L.mpm_set_g2p2g(h, 1, float(self.g2p2g_allowed_cfl), int(bool(self.v_clamp_g2p2g))))
        if self._kernel_profiler:
            check(L.mpm_set_timing(self._h(), 1))

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and self._owns_handle and _lib._lib is not None:
            _lib._lib.mpm_destroy(h)
            self._handle = None

    def n_particles_total(self):
        return self._num_particles()

    def _num_particles(self):
        n = _lib.ctypes.c_int64()
        check(_lib.lib().mpm_num_particles(self._h(), n))
        return int(n.value)

    # ------------------------------------------------------------------ #
    # configuration (mpm_solver.py:316-319, 618-692)
    # ------------------------------------------------------------------ #
    def set_gravity(self, g):
        assert isinstance(g, (tuple, list))
        assert len(g) == self.dim
        self._gravity = [float(v) for v in g]
        check(_lib.lib().mpm_set_gravity(self._h(), _lib.dvec(self._gravity)))

    def add_bounding_box(self, unbounded):
        check(_lib.lib().mpm_add_bounding_box(self._h(), int(bool(unbounded))))
        self.grid_postprocess.append(("bounding_box", {"unbounded": bool(unbounded)}))

    def add_sphere_collider(self, center, radius, surface=surface_sticky):
        center = list(center)
        check(_lib.lib().mpm_add_sphere_collider(self._h(), _lib.dvec(center), float(radius),
                                                 int(surface)))
        self.grid_postprocess.append(("sphere", dict(center=center, radius=radius,
                                                     surface=surface)))

    def add_surface_collider(self, point, normal, surface=surface_sticky, friction=0.0):
        point = list(point)
        normal_scale = 1.0 / math.sqrt(sum(x**2 for x in normal))
        normal = list(normal_scale * x for x in normal)
        if surface == self.surface_sticky and friction != 0:
            raise ValueError('friction must be 0 on sticky surfaces.')
        check(_lib.lib().mpm_add_surface_collider(self._h(), _lib.dvec(point),
                                                  _lib.dvec(normal), int(surface),
                                                  float(friction)))
        self.grid_postprocess.append(("surface", dict(point=point, normal=normal,
                                                      surface=surface, friction=friction)))

    def clear_grid_postprocess(self):
        check(_lib.lib().mpm_clear_grid_postprocess(self._h()))
        self.grid_postprocess.clear()

    # ------------------------------------------------------------------ #
    # seeding (mpm_solver.py:823-1015, 1081-1144)
    # ------------------------------------------------------------------ #
    def set_source_velocity(self, velocity):
        if velocity is not None:
            velocity = list(velocity)
            assert len(velocity) == self.dim
            self._source_velocity = [float(v) for v in velocity]
        else:
            self._source_velocity = [0.0] * self.dim

    def _append(self, x, v, material, color, emitter_id=0):
        n = len(x)
        if n == 0:
            return
        x = _lib.as_f32(x, (n, self.dim))
        v = _lib.as_f32(v, (n, self.dim)) if v is not None else None
        mat = _lib.as_i32(np.broadcast_to(np.asarray(material, np.int32), (n,)), (n,))
        col = _lib.as_i32(np.broadcast_to(np.asarray(color, np.int64).astype(np.int32), (n,)),
                          (n,))
        check(_lib.lib().mpm_add_particles(self._h(), n, _lib.fptr(x), _lib.fptr(v),
                                           _lib.iptr(mat), _lib.iptr(col)))
        self._emitter_ids = np.concatenate(
            [self._emitter_ids, np.full(n, emitter_id, np.int32)])

    def _uniform_velocity(self, n):
        return np.tile(np.asarray(self._source_velocity, np.float32), (n, 1))

    def add_cube(self,
                 lower_corner,
                 cube_size,
                 material,
                 color=0xFFFFFF,
                 sample_density=None,
                 velocity=None):
        if sample_density is None:
            sample_density = 2**self.dim
        vol = 1
        for i in range(self.dim):
            vol = vol * cube_size[i]
        num_new_particles = int(sample_density * vol / self.dx**self.dim + 1)
        assert self.n_particles[None] + num_new_particles <= self.max_num_particles
        self.set_source_velocity(velocity=velocity)
        lo = np.asarray(lower_corner[:self.dim], np.float32)
        sz = np.asarray(cube_size[:self.dim], np.float32)
        # x_k = lower_k + random() * size_k in f32 (mpm_solver.py:846-848)
        x = lo + self._rng.random((num_new_particles, self.dim), dtype=np.float32) * sz
        self._append(x, self._uniform_velocity(num_new_particles), material, color)

    def _random_point_in_unit_sphere(self, n):
        out = np.empty((n, self.dim), np.float32)
        got = 0
        while got < n:
            m = max(64, int((n - got) * 2.2))
            cand = self._rng.random((m, self.dim), dtype=np.float32) * 2 - 1
            cand = cand[(cand * cand).sum(1) <= 1]  # mpm_solver.py:962-966
            k = min(len(cand), n - got)
            out[got:got + k] = cand[:k]
            got += k
        return out

    def add_ellipsoid(self,
                      center,
                      radius,
                      material,
                      color=0xFFFFFF,
                      sample_density=None,
                      velocity=None):
        if sample_density is None:
            sample_density = 2**self.dim
        if isinstance(radius, numbers.Number):
            radius = [radius] * self.dim
        radius = list(radius)
        if self.dim == 2:
            num_particles = math.pi
        else:
            num_particles = 4 / 3 * math.pi
        for i in range(self.dim):
            num_particles *= radius[i] * self.inv_dx
        num_particles = int(math.ceil(num_particles * sample_density))
        self.set_source_velocity(velocity=velocity)
        assert self.n_particles[None] + num_particles <= self.max_num_particles
        c = np.asarray(center[:self.dim], np.float32)
        r = np.asarray(radius[:self.dim], np.float32)
        x = c + self._random_point_in_unit_sphere(num_particles) * r
        self._append(x, self._uniform_velocity(num_particles), material, color)

    def add_particles(self, particles, material, color=0xFFFFFF, velocity=None):
        self.set_source_velocity(velocity=velocity)
        particles = np.asarray(particles, np.float32).reshape(-1, self.dim)
        n = len(particles)
        assert self.n_particles[None] + n <= self.max_num_particles
        self._append(particles, self._uniform_velocity(n), material, color)

    def read_restart(self, num_particles, pos, vel, material, color):
        n = int(num_particles)
        assert self.n_particles[None] + n <= self.max_num_particles
        self._append(np.asarray(pos, np.float32)[:n], np.asarray(vel, np.float32)[:n],
                     np.asarray(material, np.int32)[:n], np.asarray(color)[:n])

    def clear_particles(self):
        """mpm_solver.py:1193-1196: drop every particle (keeps the settings)."""
        post = list(self.grid_postprocess)
        _lib.lib().mpm_destroy(self._h())
        self._handle = None
        self._create_handle()
        check(_lib.lib().mpm_set_gravity(self._h(), _lib.dvec(self._gravity)))
        self.grid_postprocess = []
        for entry in post:
            if callable(entry):
                self.grid_postprocess.append(entry)
                continue
            kind, kw = entry
            if kind == "bounding_box":
                self.add_bounding_box(kw.get("unbounded", False))
            elif kind == "sphere":
                self.add_sphere_collider(**kw)
            else:
                self.add_surface_collider(**kw)
        self._emitter_ids = np.zeros(0, np.int32)

    def add_mesh(self,
                 triangles,
                 material,
                 color=0xFFFFFF,
                 sample_density=None,
                 velocity=None,
                 translation=None,
                 emmiter_id=0):
        """mpm_solver.py:1049-1079: voxelize the closed triangle mesh
        ((n, 9): a, b, c per row; engine/voxelizer.py) and seed particles in
        the voxels inside it (seed_from_voxels, mpm_solver.py:1017-1047).
        Both run on the GPU; the random numbers come from a counter hash
        seeded by this solver's generator, in place of ti.random."""
        assert self.dim == 3
        if not self.use_voxelizer:
            raise ValueError("add_mesh needs use_voxelizer=True")
        if sample_density is None:
            sample_density = 2**self.dim
        self.set_source_velocity(velocity=velocity)
        assert isinstance(triangles, np.ndarray)
        tri = np.ascontiguousarray(triangles.astype(np.float64))  # engine/voxelizer.py:113
        assert tri.ndim == 2 and tri.shape[1] == 9
        shift = np.zeros(3, np.float32)
        if translation:
            shift[:] = [float(translation[i]) for i in range(3)]
        vel = np.asarray(self._source_velocity, np.float32)
        seed = int(self._rng.integers(0, 2**32))
        added = _lib.ctypes.c_int64()
        check(_lib.lib().mpm_add_mesh(
            self._h(), len(tri), tri.ctypes.data_as(_lib._D3), int(material),
            int(np.int64(color).astype(np.int32)), int(sample_density), _lib.fptr(vel),
            _lib.fptr(shift), int(self.voxelizer_super_sample), seed, added))
        self._emitter_ids = np.concatenate(
            [self._emitter_ids, np.full(int(added.value), emmiter_id, np.int32)])
        return int(added.value)

    def _random_point_in_unit_polygon(self, n, sides, angle):
        """Uniform points in the regular polygon of circumradius 1 with a
        vertex at ``angle`` from +x: rejection from [-1, 1]^2 by the
        reference's test (mpm_solver.py:918-931) -- the point's angle past
        the last vertex direction, theta = (atan2(y, x) - angle) mod (2 pi /
        sides), is inside when |p| < cos(phi) / cos(phi - theta), phi = pi /
        sides (the edge's distance along that direction)."""
        central = 2 * math.pi / sides
        phi = central / 2
        out = np.empty((n, 2), np.float32)
        got = 0
        while got < n:
            m = max(64, int((n - got) * 2.5))
            cand = self._rng.random((m, 2), dtype=np.float32) * 2 - 1
            c = cand.astype(np.float64)
            theta = np.mod(np.arctan2(c[:, 1], c[:, 0]) - angle, central)
            cand = cand[np.hypot(c[:, 0], c[:, 1]) < math.cos(phi) / np.cos(phi - theta)]
            k = min(len(cand), n - got)
            out[got:got + k] = cand[:k]
            got += k
        return out

    def add_ngon(self, sides, center, radius, angle, material, color=0xFFFFFF,
                 sample_density=None, velocity=None):
        """A regular polygon of particles (2D; mpm_solver.py:886-916):
        ceil(sample_density * sides/2 sin(2 pi / sides) (radius inv_dx)^2)
        particles, uniform in the polygon of circumradius ``radius`` around
        ``center`` with a vertex at ``angle`` radians from +x."""
        if self.dim != 2:
            raise ValueError("Add Ngon only works for 2D simulations")
        if sample_density is None:
            sample_density = 2**self.dim
        num_particles = 0.5 * (radius * self.inv_dx)**2 * math.sin(2 * math.pi / sides) * sides
        num_particles = int(math.ceil(num_particles * sample_density))
        self.set_source_velocity(velocity=velocity)
        assert self.n_particles[None] + num_particles <= self.max_num_particles
        pts = self._random_point_in_unit_polygon(num_particles, sides, angle)
        x = np.asarray(center[:2], np.float32) + pts * np.float32(radius)
        self._append(x, self._uniform_velocity(num_particles), material, color)

    def add_texture_2d(self, offset_x, offset_y, texture, new_material, color=0xFFFFFF):
        """One particle per texel above 0.1 (2D; mpm_solver.py:943-957): texel
        (i, j) at (offset_x + i dx, offset_y + j dx), with the velocity of
        the last seeding call.  Particles are stored in texel order (i major),
        where the reference's atomic counter leaves the order open."""
        if self.dim != 2:
            raise ValueError("add_texture_2d works for 2D simulations")
        tex = np.asarray(texture)
        if tex.ndim != 2:
            raise ValueError("texture must be a 2D array")
        i, j = np.nonzero(tex > 0.1)
        n = len(i)
        assert self.n_particles[None] + n <= self.max_num_particles
        x = np.stack([np.float32(offset_x) + i.astype(np.float32) * np.float32(self.dx),
                      np.float32(offset_y) + j.astype(np.float32) * np.float32(self.dx)], 1)
        self._append(x.astype(np.float32), self._uniform_velocity(n), new_material, color)

    def write_particles(self, fn, slice_size=1000000):
        """Quantised particle file (mpm_solver.py:1198-1200); the
        quantisation runs on the GPU (taichi_elements_amd.particle_io)."""
        from .particle_io import ParticleIO
        ParticleIO.write_particles(self, fn, slice_size)

    def write_particles_ply(self, fn):
        """PLY point cloud of positions and 0xRRGGBB colours as four
        little-endian bytes (mpm_solver.py:1202-1209)."""
        from .particle_io import write_point_cloud
        info = self.particle_info()
        write_point_cloud(fn, np.hstack([info['position'],
                                         info['color'].view(np.uint32)[:, None].view(np.float32)]))

    # ------------------------------------------------------------------ #
    # time stepping (mpm_solver.py:748-821)
    # ------------------------------------------------------------------ #
    def step(self, frame_dt, print_stat=False, smry_writer=None):
        begin_t = time.time()
        begin_substep = self.total_substeps

        substeps = int(frame_dt / self.default_dt) + 1
        dt = frame_dt / substeps
        frame_time_left = frame_dt
        if print_stat:
            print(f'needed substeps: {substeps}')
        if self.use_adaptive_dt:
            count, dt, maxv = self._adaptive_substeps(frame_dt, dt, smry_writer)
            if print_stat:
                self._print_step_stat(maxv, count, dt, begin_t, begin_substep)
            return
        count = 0
        while frame_time_left > 0:  # float64 loop; may run substeps + 1 times
            count += 1
            frame_time_left -= dt
        if PRINT_PROGRESS:
            print('.' * count, flush=True)
        maxv = self._run_substeps(count, dt)
        for k in range(count):
            self.total_substeps += 1
            self.t += dt
            cur_frame_velocity = float(maxv[k])
            if smry_writer is not None:
                smry_writer.add_scalar("substep_max_CFL", cur_frame_velocity * dt / self.dx,
                                       self.total_substeps)
            self.all_time_max_velocity = max(self.all_time_max_velocity, cur_frame_velocity)

        if print_stat:
            self._print_step_stat(maxv, count, dt, begin_t, begin_substep)

    def _print_step_stat(self, maxv, count, dt, begin_t, begin_substep):
        self.print_kernel_profiler_info()
        cur_frame_velocity = float(maxv[-1]) if count else 0.0
        print(f'CFL: {cur_frame_velocity * dt / self.dx}')
        print(f'num particles={self.n_particles[None]}')
        print(f'  frame time {time.time() - begin_t:.3f} s')
        print(f'  substep time {1000 * (time.time() - begin_t) / max(1, self.total_substeps - begin_substep):.3f} ms')

    def _adaptive_substeps(self, frame_dt, dt, smry_writer):
        """step()'s loop with use_adaptive_dt (mpm_solver.py:758-808): before
        each substep dt = min(dt, cfl dx / (max grid velocity + 1e-6), time
        left in the frame), in float64 as the reference's host loop; the
        substep itself runs with dt rounded to float32."""
        frame_time_left = frame_dt
        maxv = []
        while frame_time_left > 0:
            if PRINT_PROGRESS:
                print('.', end='', flush=True)
            self.total_substeps += 1
            max_grid_v = self.compute_max_grid_velocity()
            cfl_dt = self.g2p2g_allowed_cfl * self.dx / (max_grid_v + 1e-6)
            dt = min(dt, cfl_dt, frame_time_left)
            frame_time_left -= dt
            cur_frame_velocity = float(self._run_substeps(1, dt)[0])
            self.t += dt
            maxv.append(cur_frame_velocity)
            if smry_writer is not None:
                smry_writer.add_scalar("substep_max_CFL", cur_frame_velocity * dt / self.dx,
                                       self.total_substeps)
            self.all_time_max_velocity = max(self.all_time_max_velocity, cur_frame_velocity)
        if PRINT_PROGRESS:
            print(flush=True)
        return len(maxv), dt, maxv

    def _run_substeps(self, count, dt):
        hooks = [p for p in self.grid_postprocess if callable(p)]
        if hooks:
            return self._run_substeps_with_callables(count, dt, hooks)
        L = _lib.lib()
        check(L.mpm_substeps(self._h(), int(count), float(dt), float(self.t)))
        out = np.zeros(max(count, 1), np.float32)
        check(L.mpm_read_max_velocity(self._h(), int(count), _lib.fptr(out)))
        return out[:count]

    def _run_substeps_with_callables(self, count, dt, hooks):
        """count substeps, each split around the callables of
        grid_postprocess: the first half (binning, p2g, grid operator with
        the built-in post-processes) and the second (g2p) run in the HIP
        kernels, the callables on the active nodes in between, with the
        time t of the reference's step() loop (self.t before the substep,
        advanced by dt per substep in float64)."""
        import torch
        L = _lib.lib()
        h = self._h()
        dev = torch.device("cuda", int(self._device))
        out = np.zeros(max(count, 1), np.float32)
        t = self.t
        for k in range(count):
            check(L.mpm_substep_begin(h, float(dt)))
            failed = True
            try:
                n = _lib.ctypes.c_int64()
                check(L.mpm_grid_node_count(h, _lib.ctypes.byref(n)))
                n = n.value
                I = torch.empty((n, self.dim), dtype=torch.int32, device=dev)
                v = torch.empty((n, self.dim), dtype=torch.float32, device=dev)
                m = torch.empty((n,), dtype=torch.float32, device=dev)
                torch.cuda.synchronize(dev)  # the gather runs on the solver's own stream
                check(L.mpm_grid_gather(h, _tptr(I), _tptr(v), _tptr(m), n))
                nodes = GridNodes(I, v, m, self.dx)
                for fn in hooks:
                    fn(t, dt, nodes)
                nv = nodes.v
                if (not torch.is_tensor(nv) or nv.dtype != torch.float32
                        or tuple(nv.shape) != (n, self.dim) or nv.device != dev):
                    raise ValueError(f"grid_postprocess: grid_v.v must stay a float32 ({n}, "
                                     f"{self.dim}) tensor on {dev}")
                nv = nv.contiguous()
                # the callables' work on any stream, before the scatter reads v
                torch.cuda.synchronize(dev)
                check(L.mpm_grid_scatter(h, _tptr(nv), n))
                failed = False
            finally:
                rc = L.mpm_substep_end(h)
            if not failed:
                check(rc)
            mv = np.zeros(1, np.float32)
            check(L.mpm_read_max_velocity(h, 1, _lib.fptr(mv)))
            out[k] = mv[0]
            t += dt
        return out[:count]

    def substep(self, dt):
        """One substep with exactly ``dt`` (the body of step()'s loop).
        Returns that substep's max particle velocity component."""
        maxv = self._run_substeps(1, dt)
        self.total_substeps += 1
        self.t += dt
        self.all_time_max_velocity = max(self.all_time_max_velocity, float(maxv[0]))
        return float(maxv[0])

    def compute_max_grid_velocity(self, grid_v=None):
        """Largest velocity component magnitude on the grid the last substep
        left (mpm_solver.py:737-746); ``grid_v`` is accepted for the
        reference's signature and ignored (the engine holds one current
        grid)."""
        out = np.zeros(1, np.float32)
        check(_lib.lib().mpm_max_grid_velocity(self._h(), _lib.fptr(out)))
        return float(out[0])

    def compute_max_velocity(self):
        v = self.get_state()["v"]
        return float(np.abs(v).max()) if len(v) else 0.0

    def print_kernel_profiler_info(self):
        names, counts, ms = self.kernel_times()
        print(f"{'kernel':<16}{'calls':>8}{'total ms':>12}{'avg us':>12}")
        for nm, c, t in zip(names, counts, ms):
            print(f"{nm:<16}{c:>8}{t:>12.3f}{1000 * t / max(c, 1):>12.2f}")

    def kernel_times(self):
        L = _lib.lib()
        cap = 64
        names = _lib.ctypes.create_string_buffer(4096)
        counts = (_lib.ctypes.c_int32 * cap)()
        ms = (_lib.ctypes.c_double * cap)()
        nk = _lib.ctypes.c_int32()
        check(L.mpm_kernel_times(self._h(), names, 4096, counts, ms, cap, nk))
        nm = names.value.decode().split("\n")[:nk.value]
        return nm, list(counts[:nk.value]), list(ms[:nk.value])

    # ------------------------------------------------------------------ #
    # readout (mpm_solver.py:1172-1191) and state access
    # ------------------------------------------------------------------ #
    def particle_info(self):
        n = self.n_particles[None]
        np_x = np.empty((n, self.dim), dtype=np.float32)
        np_v = np.empty((n, self.dim), dtype=np.float32)
        np_material = np.empty((n, ), dtype=np.int32)
        np_color = np.empty((n, ), dtype=np.int32)
        if n:
            check(_lib.lib().mpm_particle_info(self._h(), _lib.fptr(np_x), _lib.fptr(np_v),
                                               _lib.iptr(np_material), _lib.iptr(np_color)))
        particles_data = {
            'position': np_x,
            'velocity': np_v,
            'material': np_material,
            'color': np_color
        }
        if self.use_emitter_id:
            particles_data['emitter_ids'] = self._emitter_ids.copy()
        return particles_data

    _FIELDS = {"x": (_lib.FIELD_X, 1), "v": (_lib.FIELD_V, 1), "C": (_lib.FIELD_C, 2),
               "F": (_lib.FIELD_F, 2), "Jp": (_lib.FIELD_JP, 0),
               "material": (_lib.FIELD_MATERIAL, 0), "color": (_lib.FIELD_COLOR, 0)}

    def _shape(self, rank, n):
        return (n, ) + (self.dim, ) * rank

    def get_state(self):
        """Full particle state in insertion order: x, v, C, F, Jp, material, color."""
        n = self.n_particles[None]
        out = {}
        for name, (fid, rank) in self._FIELDS.items():
            dt = np.int32 if name in ("material", "color") else np.float32
            a = np.empty(self._shape(rank, n), dt)
            if n:
                check(_lib.lib().mpm_get_state(self._h(), fid, a.ctypes.data))
            out[name] = a
        return out

    def set_state(self, **fields):
        """Overwrite fields (x, v, C, F, Jp, material, color) in insertion order."""
        n = self.n_particles[None]
        for name, value in fields.items():
            fid, rank = self._FIELDS[name]
            dt = np.int32 if name in ("material", "color") else np.float32
            a = np.ascontiguousarray(np.asarray(value, dt).reshape(self._shape(rank, n)))
            if n:
                check(_lib.lib().mpm_set_state(self._h(), fid, a.ctypes.data))

    def read_grid(self, lo=None, n=None):
        """Grid mass and velocity after the last grid operator over the node
        window [lo, lo + n) (default: the whole domain); arrays are indexed
        [z, y, x] (3D) or [y, x] (2D)."""
        if lo is None:
            lo = [0] * self.dim
        if n is None:
            n = [self.res[a] for a in range(self.dim)]
        lo3 = (_lib.ctypes.c_int32 * 3)(*(list(lo) + [0] * (3 - self.dim)))
        n3 = (_lib.ctypes.c_int32 * 3)(*(list(n) + [1] * (3 - self.dim)))
        shape = tuple(int(v) for v in reversed(list(n)))
        m = np.zeros(shape, np.float32)
        v = np.zeros(shape + (self.dim, ), np.float32)
        check(_lib.lib().mpm_read_grid(self._h(), lo3, n3, _lib.fptr(m), _lib.fptr(v)))
        return m, v

    def read_binning(self):
        """(keys, perm, slot_to_insertion) of the last substep's binning."""
        n = self.n_particles[None]
        keys = np.zeros(n, np.int32)
        perm = np.zeros(n, np.int32)
        orig = np.zeros(n, np.int32)
        check(_lib.lib().mpm_read_binning(self._h(), _lib.iptr(keys), _lib.iptr(perm),
                                          _lib.iptr(orig)))
        return keys, perm, orig

    def set_debug(self, flags):
        check(_lib.lib().mpm_set_debug(self._h(), int(flags)))

    # ------------------------------------------------------------------ #
    # launch tuning
    # ------------------------------------------------------------------ #
    def tuning(self):
        """(chunk, groups): particles per transfer work item and workgroups
        of the persistent transfer kernels."""
        ch, gr = _lib.ctypes.c_int32(), _lib.ctypes.c_int32()
        check(_lib.lib().mpm_get_tuning(self._h(), ch, gr))
        return int(ch.value), int(gr.value)

    def set_tuning(self, chunk=0, groups=0, g2p_waves=None):
        """chunk / groups as tuning() reports them (0 keeps a value);
        g2p_waves: 0 (the compiler's occupancy) or 6 waves per SIMD."""
        check(_lib.lib().mpm_set_tuning(self._h(), int(chunk), int(groups)))
        if g2p_waves is not None:
            check(_lib.lib().mpm_set_g2p_waves(self._h(), int(g2p_waves)))

    def g2p_waves(self):
        """g2p's occupancy target (0: the compiler's choice, or 6)."""
        w = _lib.ctypes.c_int32()
        check(_lib.lib().mpm_get_g2p_waves(self._h(), w))
        return int(w.value)

    def particle_bytes(self):
        """Bytes per particle in this storage mode (device state + colour)."""
        b = _lib.ctypes.c_int32()
        check(_lib.lib().mpm_particle_bytes(self._h(), b))
        return int(b.value)

    def set_graphs(self, enabled):
        """Replay steady-state substep pairs as HIP graphs (default on)."""
        check(_lib.lib().mpm_set_graphs(self._h(), int(bool(enabled))))

    def autotune(self, dt=None, substeps=4, chunks=None, groups=(1024, 2048, 4096),
                 verbose=False):
        """Times every (chunk, groups) candidate on the current particles
        and keeps the fastest (chunks: particles per transfer work item,
        default 128-1024, or 64-256 with use_g2p2g).  The particle state (with use_g2p2g also the
        grid the next substep reads), t and the substep count are restored
        afterwards; only float32 summation order depends on the choice.
        Returns {(chunk, groups): ms per substep}."""
        return _autotune(self, self._h(), dt if dt is not None else self.default_dt,
                         substeps, chunks, groups, verbose,
                         lambda: [dt if dt is not None else self.default_dt])

    def last_counts(self):
        out = np.zeros(4, np.int64)
        check(_lib.lib().mpm_last_counts(self._h(), out.ctypes.data_as(_lib._I64P)))
        return dict(valid=int(out[0]), chunks=int(out[1]), active_blocks=int(out[2]),
                    particle_blocks=int(out[3]))

    def block_window(self):
        """(lo, n) of the dense block window in blocks of 4 nodes per axis."""
        lo = [-2] * self.dim
        n = [(int(self.res[a]) + 3) // 4 + 5 for a in range(self.dim)]
        return lo, n
