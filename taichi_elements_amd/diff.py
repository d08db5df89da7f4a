"""Differentiable MLS-MPM substeps on the GPU through torch.autograd.

The reference's substep (engine/mpm_solver.py:487-735) is not
differentiable: its p2g overwrites F and Jp in place and its Young's modulus
is one global constant (mpm_solver.py:201).  Here one substep is a pure
function of a particle state held in a torch tensor, with a per-particle
Young's modulus E_p and a global Poisson ratio nu as differentiable inputs:

    state' = substep(state, E, nu)   (water, fixed-corotated elastic, snow, sand)

The forward runs the engine's HIP kernels (binning, p2g, grid operator,
g2p) on the caller's tensors; the backward runs hand-written adjoint kernels
(scatter of the g2p adjoint into LDS node tiles, the grid-operator adjoint
through the walls and the surface and sphere colliders, and a per-particle
kernel with the p2g/g2p gathers and the constitutive adjoint, including the
derivative of the polar rotation R = U V^T and, for snow and sand, of the
singular-value clamp / Drucker-Prager return map and the hardening), on the
forward's record of the substep (binning and raw grid sums).  Each
substep is one ``torch.autograd.Function``; ``run`` checkpoints the state
every ``checkpoint_every`` substeps (torch.utils.checkpoint) so only those
states are stored and each segment but the last is recomputed during the
backward pass.

State layout: float32 tensor of shape (NF, n), NF = 2d + 2d^2 + 1, rows
x[d], v[d], C[d*d] (row-major), F[d*d], Jp.

torch is imported before the engine library is loaded, so that the process
runs a single HIP runtime (torch's).
"""
import numpy as np
import torch
import torch.utils.checkpoint

from . import _lib
from ._lib import check
from .mpm_solver import MPMSolver


def n_fields(dim):
    return 2 * dim + 2 * dim * dim + 1


def _ptr(t):
    return _lib.ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return _lib.ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class _Substep(torch.autograd.Function):

    @staticmethod
    def forward(ctx, state, youngs, nu, solver, nu_value, dt):
        # the kernels read the state as dense (NF, n) rows and E as n floats
        if not (state.is_contiguous() and youngs.is_contiguous()):
            raise ValueError("state and youngs must be contiguous tensors")
        out = torch.empty_like(state)
        # the substep's record (binning and grid sums), kept for the backward
        # pass, which then runs only the adjoint kernels
        binning = torch.empty(solver._binning_bytes, dtype=torch.uint8, device=state.device)
        check(_lib.lib().mpm_diff_substep(solver._handle, float(dt), _ptr(state), _ptr(out),
                                          _ptr(youngs), float(nu_value), _ptr(binning),
                                          _stream()))
        ctx.mark_non_differentiable(binning)
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(state, youngs, binning)
        ctx.solver, ctx.nu_value, ctx.dt = solver, nu_value, dt
        ctx.nu_shape, ctx.nu_dtype = nu.shape, nu.dtype
        return out, binning

    @staticmethod
    def backward(ctx, grad_out, _grad_binning):
        state, youngs, binning = ctx.saved_tensors
        if grad_out is None:
            grad_out = torch.zeros_like(state)
        grad_out = grad_out.contiguous()
        grad_in = torch.empty_like(state)
        grad_e = torch.empty_like(youngs)
        grad_nu = torch.empty(1, dtype=torch.float32, device=state.device)
        check(_lib.lib().mpm_diff_substep_backward(
            ctx.solver._handle, float(ctx.dt), _ptr(state), _ptr(youngs), float(ctx.nu_value),
            _ptr(binning), _ptr(grad_out), _ptr(grad_in), _ptr(grad_e), _ptr(grad_nu),
            _stream()))
        return (grad_in, grad_e, grad_nu.reshape(ctx.nu_shape).to(ctx.nu_dtype), None, None,
                None)


class DiffMPMSolver:
    """A differentiable scene: the constants, gravity and walls of
    ``MPMSolver(res, size, padding, dt_scale, E_scale, water_density)``,
    particles in a torch state tensor.

    ``setup(material)`` fixes the particle count and materials (water /
    elastic); ``state(x, v)`` builds an initial state (F = I, C = 0,
    Jp = 1); ``substep`` / ``run`` advance it differentiably.
    """

    def __init__(self, res, size=1, padding=3, dt_scale=1, E_scale=1, water_density=1.0,
                 device=0, gravity=None):
        from . import mpm_solver
        mpm_solver.PRINT_PROGRESS = False
        self._solver = MPMSolver(res, size=size, padding=padding, dt_scale=dt_scale,
                                 E_scale=E_scale, water_density=water_density, device=device,
                                 reorder_interval=-1)
        if gravity is not None:
            self._solver.set_gravity(tuple(gravity))
        self.dim = self._solver.dim
        self.default_dt = self._solver.default_dt
        self.E = self._solver.E
        self.nu = self._solver.nu
        self.device = torch.device("cuda", device)
        self.n = 0
        self._order = self._inverse = None

    @property
    def _handle(self):
        return self._solver._handle

    @property
    def solver(self):
        """The underlying MPMSolver (constants, set_gravity)."""
        return self._solver

    def setup(self, material, positions=None):
        """Fix the particles' materials (water, elastic, snow or sand; not
        stationary).  With the initial
        ``positions`` given, the substeps work on the particles in a
        spatially sorted order (by 4^d grid block) so that the transfers
        read coherent memory; ``run`` maps to and from the caller's order
        with differentiable gathers."""
        material = np.ascontiguousarray(np.asarray(material, np.int32).reshape(-1))
        n = len(material)
        self._order = self._inverse = None
        caller_material = material
        if positions is not None:
            x = np.asarray(positions, np.float32).reshape(n, self.dim)
            base = np.floor(x * np.float32(self._solver.inv_dx) - np.float32(0.5)).astype(np.int64)
            blk = (base >> 2) + 1024
            key = np.zeros(n, np.int64)
            for a in reversed(range(self.dim)):
                key = key * 4096 + blk[:, a]
            order = np.argsort(key, kind="stable")
            self._order = torch.as_tensor(order, device=self.device)
            self._inverse = torch.empty_like(self._order)
            self._inverse[self._order] = torch.arange(n, device=self.device)
            material = np.ascontiguousarray(material[order])
        check(_lib.lib().mpm_diff_setup(self._handle, n, _lib.iptr(material)))
        nbytes = _lib.ctypes.c_int64()
        check(_lib.lib().mpm_diff_binning_bytes(self._handle, nbytes))
        self._binning_bytes = int(nbytes.value)
        self.n = n
        self.material = caller_material

    def state(self, x, v=None):
        """Initial state tensor from positions (and velocities): F = I,
        C = 0, Jp = 0 for sand and 1 otherwise (seed_particle,
        mpm_solver.py:823-838)."""
        d, n = self.dim, self.n
        x = torch.as_tensor(np.asarray(x, np.float32).reshape(n, d))
        s = torch.zeros(n_fields(d), n, dtype=torch.float32)
        s[0:d] = x.T
        if v is not None:
            s[d:2 * d] = torch.as_tensor(np.asarray(v, np.float32).reshape(n, d)).T
        f0 = 2 * d + d * d
        for a in range(d):
            s[f0 + a * d + a] = 1.0
        s[-1] = torch.as_tensor(np.where(self.material == 3, 0.0, 1.0).astype(np.float32))
        return s.to(self.device)

    def positions(self, state):
        return state[0:self.dim].T

    def substep(self, state, youngs, nu, dt=None, nu_value=None):
        """One differentiable substep.  ``youngs``: (n,) float32 tensor;
        ``nu``: 0-d tensor (its value is read on the host unless
        ``nu_value`` is given)."""
        return self.run(state, youngs, nu, 1, dt=dt, checkpoint_every=1, nu_value=nu_value)

    def _segment(self, state, youngs, nu, nu_value, dt, k):
        for _ in range(k):
            state, _ = _Substep.apply(state, youngs, nu, self, nu_value, dt)
        return state

    # Device memory that run() may fill with the inputs of every substep
    # before it falls back to checkpointing (a few per cent of an MI355X's
    # 288 GB of HBM).
    keep_bytes = 8 << 30

    def run(self, state, youngs, nu, nsub, dt=None, checkpoint_every=None, nu_value=None):
        """``nsub`` substeps; with gradients enabled and checkpoint_every > 1
        only every checkpoint_every-th state is kept and each segment but the
        last is recomputed in the backward pass (the last segment keeps its
        substeps' inputs: the backward pass would recompute them first, so
        the peak memory is the same).  checkpoint_every=None keeps every
        substep's input (no recomputation) when they fit in ``keep_bytes``,
        else checkpoints every 8 substeps.  The gradients are the same
        either way."""
        if dt is None:
            dt = self.default_dt
        if nu_value is None:
            nu_value = float(nu.detach()) if torch.is_tensor(nu) else float(nu)
        if self._order is not None:
            state = state[:, self._order]
            youngs = youngs[self._order]
        # a transposed or sliced tensor is copied into the (NF, n) row layout
        # the kernels read (the copy is part of the autograd graph)
        state = state.contiguous()
        youngs = youngs.contiguous()
        if checkpoint_every is None:
            per = state.numel() * state.element_size() + self._binning_bytes
            checkpoint_every = 1 if nsub * per <= self.keep_bytes else 8
        self.last_checkpoint_every = int(checkpoint_every)
        k = max(1, int(checkpoint_every))
        for t0 in range(0, nsub, k):
            m = min(k, nsub - t0)
            if k > 1 and torch.is_grad_enabled() and t0 + m < nsub:
                state = torch.utils.checkpoint.checkpoint(self._segment, state, youngs, nu,
                                                          nu_value, dt, m, use_reentrant=False)
            else:
                state = self._segment(state, youngs, nu, nu_value, dt, m)
        if self._inverse is not None:
            state = state[:, self._inverse]
        return state

    def check(self):
        """Waits for the work queued so far; raises DomainError if particles
        left the grid window or moved more than a block in one substep."""
        check(_lib.lib().mpm_diff_sync(self._handle, _stream()))
