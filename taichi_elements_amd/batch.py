"""Batched scenes: many independent ``MPMSolver`` scenes advanced together.

The reference simulates one scene per ``MPMSolver`` (engine/mpm_solver.py)
and a caller with many small scenes runs them one after another.  Here the
scenes share one engine handle: their block windows are stacked along the
last axis of one combined window and every stage of a substep (binning, p2g,
grid operator, g2p) is one launch over the work of all scenes, so small
scenes no longer leave the GPU idle between launches.  Each scene keeps its
own constants (``res``, ``size``, ``dt_scale``, ``E_scale``,
``water_density``, ``padding``), gravity, colliders and particles.

``MPMSolverBatch(scenes)`` takes one dict of ``MPMSolver`` constructor
arguments per scene; ``batch[k]`` is scene k with the reference's seeding,
collider, gravity and readout methods (``add_cube``, ``add_ellipsoid``,
``add_particles``, ``read_restart``, ``add_surface_collider``,
``add_sphere_collider``, ``set_gravity``, ``particle_info``, ``get_state``,
``set_state``, ``read_grid``).  Time stepping belongs to the batch:
``substep(dts)`` advances every scene by one substep of its own ``dt`` and
``step(frame_dt)`` applies the reference's substep-count rule
(mpm_solver.py:752-771) per scene, letting scenes that need fewer substeps
sit the remaining ones out.
"""
import numpy as np

from . import _lib
from ._lib import check
from .mpm_solver import MPMSolver, PRINT_PROGRESS


class BatchScene(MPMSolver):
    """One scene of an :class:`MPMSolverBatch` (the ``MPMSolver`` API on its
    share of the batch).  Created by the batch, not directly."""

    _owns_handle = False

    def __init__(self, batch, index, res, **kwargs):
        self._batch = batch
        self._index = index
        self._ids = np.zeros(0, np.int64)  # global insertion indices of this scene
        super().__init__(res, **kwargs)

    def _h(self):
        h = self._batch._handle
        if h is None:
            raise RuntimeError("the batch of this scene has been closed")
        check(_lib.lib().mpm_set_active_scene(h, self._index))
        return h

    def _num_particles(self):
        return len(self._ids)

    def _append(self, x, v, material, color, emitter_id=0):
        n0 = self._batch._total_particles()
        super()._append(x, v, material, color, emitter_id)
        n1 = self._batch._total_particles()
        self._ids = np.concatenate([self._ids, np.arange(n0, n1, dtype=np.int64)])

    def add_mesh(self, *args, **kwargs):
        n0 = self._batch._total_particles()
        added = super().add_mesh(*args, **kwargs)
        n1 = self._batch._total_particles()
        self._ids = np.concatenate([self._ids, np.arange(n0, n1, dtype=np.int64)])
        return added

    # readout: the library returns the whole batch in insertion order
    def particle_info(self):
        full = self._batch._particle_info_all()
        out = {k: v[self._ids] for k, v in full.items()}
        if self.use_emitter_id:
            out['emitter_ids'] = self._emitter_ids.copy()
        return out

    def get_state(self):
        return {k: v[self._ids] for k, v in self._batch._get_state_all().items()}

    def set_state(self, **fields):
        full = self._batch._get_state_all()
        n = len(self._ids)
        for name, value in fields.items():
            _, rank = self._FIELDS[name]
            a = full[name]
            a[self._ids] = np.asarray(value, a.dtype).reshape(self._shape(rank, n))
            self._batch._set_state_all(name, a)

    def read_binning(self):
        raise NotImplementedError("the binning of a batch is read with MPMSolverBatch.read_binning")

    def last_counts(self):
        raise NotImplementedError("bookkeeping of a batch is read with MPMSolverBatch.last_counts")

    # time stepping belongs to the batch
    def step(self, frame_dt, print_stat=False, smry_writer=None):
        raise NotImplementedError("scenes of a batch are advanced with MPMSolverBatch.step")

    def substep(self, dt):
        raise NotImplementedError("scenes of a batch are advanced with MPMSolverBatch.substep")

    def clear_particles(self):
        raise NotImplementedError("particles cannot be removed from one scene of a batch")


class MPMSolverBatch:
    """Independent MPM scenes advanced by one engine.

    ``scenes``: one dict per scene of keyword arguments of ``MPMSolver``
    (``res`` is required; all scenes must have the same dimension).  The
    keyword-only arguments apply to the whole batch; scene k seeds its
    particles from ``numpy.random.default_rng([seed, k])``.
    """

    def __init__(self, scenes, *, device=0, seed=0, reorder_interval=0, chunk=0,
                 kernel_profiler=False):
        scenes = [dict(s) for s in scenes]
        if not scenes:
            raise ValueError("a batch needs at least one scene")
        self._handle = None
        common = dict(device=device, reorder_interval=reorder_interval, chunk=chunk,
                      kernel_profiler=kernel_profiler)
        self.scenes = []
        for k, kw in enumerate(scenes):
            res = kw.pop("res")
            self.scenes.append(BatchScene(self, k, res, seed=[seed, k], **common, **kw))
        self.dim = self.scenes[0].dim
        if any(s.dim != self.dim for s in self.scenes):
            raise ValueError("all scenes of a batch must have the same dimension")
        if any(s.unbounded or s.use_adaptive_dt for s in self.scenes):
            raise NotImplementedError("unbounded and use_adaptive_dt scenes run in their own "
                                      "MPMSolver")
        cfgs = (_lib.MpmConfig * len(self.scenes))()
        for k, s in enumerate(self.scenes):
            cfgs[k] = s._make_config()
        L = _lib.lib()
        h = _lib._P()
        check(L.mpm_create_batch(cfgs, len(self.scenes), h))
        self._handle = h
        storage = {(bool(s.quant), bool(s.support_plasticity)) for s in self.scenes}
        if len(storage) > 1:
            raise ValueError("quant and support_plasticity must be the same for every scene "
                             "of a batch")
        quant, plastic = storage.pop()
        if quant or not plastic:
            check(L.mpm_set_storage(h, int(quant), int(plastic)))
        modes = {(bool(s.use_g2p2g), bool(s.v_clamp_g2p2g), float(s.g2p2g_allowed_cfl))
                 for s in self.scenes}
        if len(modes) > 1:
            raise ValueError("use_g2p2g, v_clamp_g2p2g and g2p2g_allowed_cfl must be the same "
                             "for every scene of a batch")
        g2p2g, clamp, cfl = modes.pop()
        if g2p2g:
            check(L.mpm_set_g2p2g(h, 1, cfl, int(clamp)))
        if kernel_profiler:
            check(L.mpm_set_timing(h, 1))
        for s in self.scenes:
            s._init_scene()
        self.total_substeps = 0

    def __len__(self):
        return len(self.scenes)

    def __getitem__(self, k):
        return self.scenes[k]

    def __iter__(self):
        return iter(self.scenes)

    def close(self):
        if self._handle is not None and _lib._lib is not None:
            _lib._lib.mpm_destroy(self._handle)
        self._handle = None

    def __del__(self):
        self.close()

    # ------------------------------------------------------------------ #
    def _total_particles(self):
        n = _lib.ctypes.c_int64()
        check(_lib.lib().mpm_num_particles(self._handle, n))
        return int(n.value)

    @property
    def n_particles(self):
        """Particles of every scene together."""
        return self._total_particles()

    def default_dts(self):
        return [s.default_dt for s in self.scenes]

    def substeps(self, n, dts=None):
        """``n`` substeps; scene k uses ``dts[k]`` (default: its default_dt;
        0 lets the scene sit the substeps out).  Returns the per-substep
        maximum particle velocity component over the scenes that moved."""
        n = int(n)
        if dts is None:
            dts = self.default_dts()
        dts = [float(dt) for dt in dts]
        if len(dts) != len(self.scenes):
            raise ValueError(f"expected {len(self.scenes)} time steps, got {len(dts)}")
        if any(callable(p) for sc in self.scenes for p in sc.grid_postprocess):
            raise NotImplementedError("callables in grid_postprocess run with a single "
                                      "MPMSolver; a batch runs the built-in post-processes only")
        dts32 = np.ascontiguousarray(np.asarray(dts, np.float32))
        L = _lib.lib()
        check(L.mpm_substeps_batch(self._handle, n, _lib.fptr(dts32)))
        out = np.zeros(max(n, 1), np.float32)
        check(L.mpm_read_max_velocity(self._handle, n, _lib.fptr(out)))
        vmax = float(out[:n].max(initial=0))
        for s, dt in zip(self.scenes, dts):
            if dt > 0:
                s.total_substeps += n
                for _ in range(n):
                    s.t += dt
                s.all_time_max_velocity = max(s.all_time_max_velocity, vmax)
        self.total_substeps += n
        return out[:n]

    def substep(self, dts=None):
        """One substep of every scene (see ``substeps``)."""
        return float(self.substeps(1, dts)[0])

    def step(self, frame_dt):
        """Advance every scene by ``frame_dt`` with the reference's rule:
        ``int(frame_dt / default_dt) + 1`` substeps of ``frame_dt / substeps``
        counted by the float64 loop of mpm_solver.py:752-771.  Scenes that need
        fewer substeps than others sit the remaining ones out."""
        plan = []
        for s in self.scenes:
            substeps = int(frame_dt / s.default_dt) + 1
            dt = frame_dt / substeps
            left, count = frame_dt, 0
            while left > 0:
                count += 1
                left -= dt
            plan.append((count, dt))
        if PRINT_PROGRESS:
            print('.' * max(c for c, _ in plan), flush=True)
        done = 0
        for count in sorted(set(c for c, _ in plan)):
            dts = [dt if c >= count else 0.0 for c, dt in plan]
            if count > done:
                self.substeps(count - done, dts)
            done = count

    # ------------------------------------------------------------------ #
    def _particle_info_all(self):
        n = self._total_particles()
        d = self.dim
        x = np.empty((n, d), np.float32)
        v = np.empty((n, d), np.float32)
        m = np.empty((n, ), np.int32)
        c = np.empty((n, ), np.int32)
        if n:
            check(_lib.lib().mpm_particle_info(self._handle, _lib.fptr(x), _lib.fptr(v),
                                               _lib.iptr(m), _lib.iptr(c)))
        return {'position': x, 'velocity': v, 'material': m, 'color': c}

    def particle_info(self):
        """``particle_info()`` of every scene from one readback."""
        full = self._particle_info_all()
        return [{k: v[s._ids] for k, v in full.items()} for s in self.scenes]

    def _get_state_all(self):
        n = self._total_particles()
        out = {}
        for name, (fid, rank) in MPMSolver._FIELDS.items():
            dt = np.int32 if name in ("material", "color") else np.float32
            a = np.empty((n, ) + (self.dim, ) * rank, dt)
            if n:
                check(_lib.lib().mpm_get_state(self._handle, fid, a.ctypes.data))
            out[name] = a
        return out

    def _set_state_all(self, name, a):
        fid, _ = MPMSolver._FIELDS[name]
        a = np.ascontiguousarray(a)
        if len(a):
            check(_lib.lib().mpm_set_state(self._handle, fid, a.ctypes.data))

    def read_binning(self):
        """(keys, perm, slot_to_insertion) of the last substep over the batch."""
        n = self._total_particles()
        keys = np.zeros(n, np.int32)
        perm = np.zeros(n, np.int32)
        orig = np.zeros(n, np.int32)
        check(_lib.lib().mpm_read_binning(self._handle, _lib.iptr(keys), _lib.iptr(perm),
                                          _lib.iptr(orig)))
        return keys, perm, orig

    def last_counts(self):
        out = np.zeros(4, np.int64)
        check(_lib.lib().mpm_last_counts(self._handle, out.ctypes.data_as(_lib._I64P)))
        return dict(valid=int(out[0]), chunks=int(out[1]), active_blocks=int(out[2]),
                    particle_blocks=int(out[3]))

    def sync(self):
        check(_lib.lib().mpm_sync(self._handle))

    def n_particles_total(self):
        return self._total_particles()

    def set_tuning(self, chunk=0, groups=0, g2p_waves=None):
        check(_lib.lib().mpm_set_tuning(self._handle, int(chunk), int(groups)))
        if g2p_waves is not None:
            check(_lib.lib().mpm_set_g2p_waves(self._handle, int(g2p_waves)))

    def tuning(self):
        return MPMSolver.tuning(self)

    def g2p_waves(self):
        return MPMSolver.g2p_waves(self)

    def autotune(self, substeps=4, chunks=None, groups=(1024, 2048, 4096),
                 verbose=False):
        """MPMSolver.autotune for the whole batch (each scene at its
        default_dt); the state of every scene is restored afterwards."""
        from .mpm_solver import _autotune

        return _autotune(self, self._handle, None, substeps, chunks, groups, verbose,
                         self.default_dts)

    def kernel_times(self):
        return MPMSolver.kernel_times(self)

    def _h(self):
        return self._handle
