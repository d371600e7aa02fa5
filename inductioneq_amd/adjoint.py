"""Gradients of a periodic Hall solve through the library's discrete adjoint.

``HallSolve.apply(b0, u, c_hall, solver)`` runs ``solver.nsteps`` SSPRK33 steps of the
fused periodic stage (6th/4th/2nd-order central SBP operators, HO dissipation, uniform
velocity ``u`` and Hall coefficient ``c_hall``) and returns B(T); its backward is the
hand-written discrete adjoint (transposed operators, linearised Hall term, reversed RK
stages) with optimal binomial checkpointing.  Gradients flow to ``b0`` (a ``(3, stride)``
device tensor in the library layout), ``u`` (3,) and ``c_hall`` () — the latter two as
FP64 tensors.

There is no counterpart in the reference; the forward is the time loop of
``matlab/+induction_eq/compute_numerical_solution.m:91-108`` with the SSPRK33 kernels of
``kernel/kernel_time_integrator.cl:224-303`` (periodic grid, uniform u and rho).
"""
from __future__ import annotations

import ctypes

from . import _capi
from .program import Program


class AdjointSolver:
    """Owns the library's adjoint object (registers and snapshot pool) for one program."""

    def __init__(self, program: Program, nsteps: int, snapshots: int = 15):
        self.program = program
        self.lib = program.lib
        self.nsteps = int(nsteps)
        self.snapshots = int(snapshots)
        program.use_current_stream()
        h = ctypes.c_void_p()
        _capi.check(self.lib.ieq_adjoint_create(ctypes.byref(h), program.handle, self.nsteps, self.snapshots),
                    program.handle)
        self.handle = h
        self.last_forward_seconds = 0.0
        self.last_backward_seconds = 0.0

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.ieq_adjoint_release(h)
            self.handle = None

    def _check(self, t, what):
        if (tuple(t.shape) != (3, self.program.stride) or t.dtype != self.program.torch_dtype or not t.is_cuda
                or not t.is_contiguous()):
            raise ValueError("%s must be a contiguous (3, %d) %s device tensor" % (what, self.program.stride,
                                                                                 self.program.torch_dtype))

    def forward(self, b0, u, c_hall: float):
        self._check(b0, "b0")
        bT = self.program.vector()
        uu = (ctypes.c_double * 3)(*[float(v) for v in u])
        sec = ctypes.c_double(0.0)
        rc = self.lib.ieq_adjoint_forward(self.handle, ctypes.c_void_p(b0.data_ptr()), ctypes.c_void_p(bT.data_ptr()),
                                          float(c_hall), uu, ctypes.byref(sec))
        _capi.check(rc, self.program.handle)
        self.last_forward_seconds = sec.value
        return bT

    def backward(self, lam_T):
        g = self.program.vector()
        grads = (ctypes.c_double * 4)()
        sec = ctypes.c_double(0.0)
        lam_T = lam_T.contiguous()
        self._check(lam_T, "lambda_T")
        rc = self.lib.ieq_adjoint_backward(self.handle, ctypes.c_void_p(lam_T.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                                           grads, ctypes.byref(sec))
        _capi.check(rc, self.program.handle)
        self.last_backward_seconds = sec.value
        return g, float(grads[0]), [float(grads[1]), float(grads[2]), float(grads[3])]

    def counts(self):
        f, a = ctypes.c_longlong(), ctypes.c_longlong()
        _capi.check(self.lib.ieq_adjoint_counts(self.handle, ctypes.byref(f), ctypes.byref(a)), self.program.handle)
        return int(f.value), int(a.value)


def _autograd_function():
    import torch

    class HallSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, b0, u, c_hall, solver):
            ctx.solver = solver
            ctx.u_dtype, ctx.c_dtype = u.dtype, c_hall.dtype
            return solver.forward(b0.detach(), u.detach().cpu().tolist(), float(c_hall.detach().cpu()))

        @staticmethod
        def backward(ctx, grad_bT):
            g, gc, gu = ctx.solver.backward(grad_bT)
            gu_t = torch.tensor(gu, dtype=ctx.u_dtype)
            gc_t = torch.tensor(gc, dtype=ctx.c_dtype)
            return g, gu_t, gc_t, None

    return HallSolve


_HALL_SOLVE = None


def hall_solve(b0, u, c_hall, solver: AdjointSolver):
    """Differentiable B(T) = solve(b0; u, c_hall) (see module docstring)."""
    global _HALL_SOLVE
    if _HALL_SOLVE is None:
        _HALL_SOLVE = _autograd_function()
    return _HALL_SOLVE.apply(b0, u, c_hall, solver)
