"""Ensembles of independent periodic solves advanced together.

Every member is a periodic box with its own grid, spacing, time step, uniform density
and uniform velocity; all members share one program's settings (REAL type, operator,
Hall term, dissipation, flux forms).  Each RK stage of the whole ensemble is one
persistent launch per fused kernel over a tile list that covers every member.  The tile
shape, z-chunk and tile order, and whether one step (two for the 2N schemes) is replayed
from a captured hipGraph or launched directly, are chosen by timing on the first step.

There is no counterpart in the reference, which builds one OpenCL program per solve
(``matlab/+induction_eq/compute_numerical_solution.m:23``); each member's arithmetic is
that of the fused single-solve path (``ieq_step``).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Sequence, Tuple

from . import _capi
from .program import Program

ENSEMBLE_INTEGRATORS = ("SSPRK33", "CarpenterKennedy2N54", "ToulorgeDesmet2N84F")


@dataclass
class MemberSpec:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dt: float
    rho: float = 1.0
    u: Tuple[float, float, float] = (0.0, 0.0, 0.0)


class Ensemble:
    """Device registers of every member plus the library's ensemble handle."""

    def __init__(self, program: Program, members: Sequence[MemberSpec], integrator: str = "SSPRK33"):
        import torch

        self.program = program
        self.lib = program.lib
        self.members: List[MemberSpec] = list(members)
        self.integrator = integrator
        dt = program.torch_dtype
        self.fields = []
        arr = (_capi.Member * len(self.members))()
        for i, m in enumerate(self.members):
            stride = int(self.lib.ieq_member_stride(m.nx, m.ny, m.nz))
            regs = [torch.zeros((3, stride), dtype=dt, device="cuda") for _ in range(4)]
            self.fields.append(regs)
            arr[i] = _capi.Member(m.nx, m.ny, m.nz, m.dx, m.dy, m.dz, m.dt, m.rho, m.u[0], m.u[1], m.u[2],
                                  regs[0].data_ptr(), regs[1].data_ptr(), regs[2].data_ptr(), regs[3].data_ptr())
        program.use_current_stream()
        h = ctypes.c_void_p()
        rc = self.lib.ieq_ensemble_create(ctypes.byref(h), program.handle, arr, len(self.members), integrator.encode())
        _capi.check(rc, program.handle)
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.ieq_ensemble_release(h)
            self.handle = None

    def set_state(self, i: int, b):
        """b: (3, nx, ny, nz) array indexed [c, x, y, z] (node index x + nx*(y + ny*z))."""
        import torch

        m = self.members[i]
        t = torch.as_tensor(b, dtype=self.program.torch_dtype)
        flat = t.permute(0, 3, 2, 1).reshape(3, -1)
        self.fields[i][0].zero_()
        self.fields[i][0][:, : m.nx * m.ny * m.nz].copy_(flat)

    def state(self, i: int):
        """Current b1 of member i as a (3, nx, ny, nz) CPU tensor."""
        m = self.members[i]
        f = self.fields[i][0][:, : m.nx * m.ny * m.nz].cpu()
        return f.reshape(3, m.nz, m.ny, m.nx).permute(0, 3, 2, 1).contiguous()

    def step(self, nsteps: int) -> float:
        """Advance every member by nsteps steps; returns the device time in seconds."""
        sec = ctypes.c_double(0.0)
        rc = self.lib.ieq_ensemble_step(self.handle, int(nsteps), ctypes.byref(sec))
        _capi.check(rc, self.program.handle)
        return sec.value

    def config(self) -> dict:
        ty, zc, order = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _capi.check(self.lib.ieq_ensemble_config(self.handle, ctypes.byref(ty), ctypes.byref(zc), ctypes.byref(order)),
                    self.program.handle)
        if ty.value in (1, 2):  # the one-kernel Hall stage (E formed in the stage kernel), two tiles
            f64 = self.program.info.real_bytes == 8
            tile = ((32, 8) if f64 else (64, 8)) if ty.value == 1 else ((16, 16) if f64 else (32, 8))
            name = "hall_%dx%d" % tile
        else:
            tile = {4: (64, 4), 8: (64, 8), 32: (32, 8)}.get(ty.value, (64, ty.value))  # tile shape code -> (x, y)
            name = "%dx%d" % tile
        return {"tile": name, "tile_rows": tile[1], "z_chunk": zc.value,
                "order": ("member", "largest_first")[order.value & 1], "graph": bool(order.value & 2)}
