"""Program handle: the C-ABI library driven with torch tensors as device memory.

``Program`` mirrors the reference's build-then-run contract
(``cl_run_kernel(device, kernel_path_list, settings)`` then
``cl_run_kernel(device, names, g_range, l_range, args..., 0)``,
``matlab/+induction_eq/compute_numerical_solution.m:23,39-41``): it is built from the
same settings string and module list, and runs kernel lists by reference name with the
reference argument order.  Device fields are torch tensors in this library's SoA
layout (``(3, stride)`` for vector fields, ``(stride,)`` for scalar fields).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Union

import numpy as np

from . import _capi
from .settings import ProgramConfig


class Program:
    def __init__(self, config: Union[ProgramConfig, str], modules: Optional[Sequence[str]] = None):
        self.lib = _capi.load()
        if isinstance(config, ProgramConfig):
            self.config = config
            settings, modules = config.settings_string(), config.modules()
        else:
            settings = config
            self.config = ProgramConfig.from_settings(settings, modules or [])
        self.settings, self.modules = settings, list(modules)
        h = ctypes.c_void_p()
        mods = (ctypes.c_char_p * len(self.modules))(*[m.encode() for m in self.modules])
        rc = self.lib.ieq_program_build(ctypes.byref(h), settings.encode(), mods, len(self.modules))
        _capi.check(rc, None)
        self.handle = h
        info = _capi.ProgramInfo()
        _capi.check(self.lib.ieq_program_get_info(self.handle, ctypes.byref(info)), self.handle)
        self.info = info
        self._stream = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.ieq_program_release(h)
            self.handle = None

    # --- properties ----------------------------------------------------------------
    @property
    def stride(self) -> int:
        return int(self.info.stride)

    @property
    def torch_dtype(self):
        import torch

        return torch.float64 if self.info.real_bytes == 8 else torch.float32

    @property
    def np_dtype(self):
        return np.float64 if self.info.real_bytes == 8 else np.float32

    # --- device fields (torch tensors) ------------------------------------------------
    def vector(self, device="cuda"):
        import torch

        return torch.zeros((3, self.stride), dtype=self.torch_dtype, device=device)

    def scalar(self, fill=0.0, device="cuda"):
        import torch

        return torch.full((self.stride,), fill, dtype=self.torch_dtype, device=device)

    def time_buffer(self, t0=0.0, device="cuda"):
        import torch

        return torch.tensor([t0, t0], dtype=self.torch_dtype, device=device)

    def reduce_output(self, device="cuda"):
        import torch

        return torch.zeros((int(self.info.num_groups),), dtype=self.torch_dtype, device=device)

    def use_current_stream(self):
        """Enqueue on torch's current stream (the library's work then orders with torch's)."""
        import torch

        s = torch.cuda.current_stream()
        self._stream = s
        _capi.check(self.lib.ieq_set_stream(self.handle, ctypes.c_void_p(s.cuda_stream)), self.handle)

    # --- host layout <-> device layout -------------------------------------------------
    def upload_real4(self, host_aos: np.ndarray, dst=None):
        """AoS REAL4 host array (n, 4) -> device vector field."""
        a = np.ascontiguousarray(host_aos, dtype=self.np_dtype)
        assert a.ndim == 2 and a.shape[1] == 4 and a.shape[0] <= self.stride
        dst = self.vector() if dst is None else dst
        _capi.check(self.lib.ieq_upload_real4(self.handle, ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(a.ctypes.data),
                                              a.shape[0]), self.handle)
        return dst

    def download_real4(self, src, n: Optional[int] = None) -> np.ndarray:
        n = int(self.info.num_nodes_pad) if n is None else n
        out = np.zeros((n, 4), self.np_dtype)
        _capi.check(self.lib.ieq_download_real4(self.handle, ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(src.data_ptr()),
                                                n), self.handle)
        return out

    def upload_real(self, host: np.ndarray, dst=None):
        a = np.ascontiguousarray(host, dtype=self.np_dtype).reshape(-1)
        dst = self.scalar() if dst is None else dst
        _capi.check(self.lib.ieq_upload_real(self.handle, ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(a.ctypes.data),
                                             a.size), self.handle)
        return dst

    def download_real(self, src, n: Optional[int] = None) -> np.ndarray:
        n = int(self.info.num_nodes_pad) if n is None else n
        out = np.zeros((n,), self.np_dtype)
        _capi.check(self.lib.ieq_download_real(self.handle, ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(src.data_ptr()),
                                               n), self.handle)
        return out

    # --- execution -----------------------------------------------------------------------
    @staticmethod
    def _ptrs(args):
        arr = (ctypes.c_void_p * len(args))()
        for i, a in enumerate(args):
            arr[i] = a.data_ptr() if hasattr(a, "data_ptr") else int(a)
        return arr

    def run(self, names, args, g=None, l=None, timed: bool = False):
        """Run one or a list of reference kernels on one argument list (MatCL style)."""
        if isinstance(names, str):
            names = [names]
        g = g or (self.info.nx, self.info.ny, self.info.nz)
        l = l or (self.info.w_size, 1, 1)
        G = (ctypes.c_uint32 * 3)(*[int(v) for v in g])
        L = (ctypes.c_uint32 * 3)(*[int(v) for v in l])
        cn = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        sec = ctypes.c_double(0.0)
        rc = self.lib.ieq_run(self.handle, cn, len(names), G, L, self._ptrs(args), len(args),
                              ctypes.byref(sec) if timed else None)
        _capi.check(rc, self.handle)
        return sec.value if timed else None

    def run_host(self, names, arrays, g=None, l=None) -> float:
        """MatCL's run call on host numpy arrays (ieq_run_host): (n, 4) arrays are REAL4 fields, 1-D arrays
        REAL buffers; every array is written back in place.  Returns the kernel time in seconds."""
        if isinstance(names, str):
            names = [names]
        g = g or (self.info.nx, self.info.ny, self.info.nz)
        l = l or (self.info.w_size, 1, 1)
        dt = np.float64 if self.info.real_bytes == 8 else np.float32
        hargs = (_capi.HostArg * len(arrays))()
        for i, a in enumerate(arrays):
            if a.dtype != dt or not a.flags.c_contiguous or a.ndim not in (1, 2) or (a.ndim == 2 and a.shape[1] != 4):
                raise ValueError("argument %d: need a C-contiguous %s array of shape (n, 4) or (n,)" % (i, dt.__name__))
            hargs[i] = _capi.HostArg(a.ctypes.data, a.shape[0], int(a.ndim == 2))
        G = (ctypes.c_uint32 * 3)(*[int(v) for v in g])
        L = (ctypes.c_uint32 * 3)(*[int(v) for v in l])
        cn = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        sec = ctypes.c_double(0.0)
        _capi.check(self.lib.ieq_run_host(self.handle, cn, len(names), G, L, hargs, len(arrays), ctypes.byref(sec)),
                    self.handle)
        return sec.value

    def step(self, integrator: str, nsteps: int, args, variable_u: bool = False, timed: bool = False, uniform=None):
        """Advance nsteps full time steps (fused kernels where the program qualifies).

        ``uniform = (u_x, u_y, u_z, rho)`` states that the u and rho fields hold these values at every
        node (ieq_step_uniform: no device check, no host synchronisation)."""
        sec = ctypes.c_double(0.0)
        if uniform is not None:
            u3 = (ctypes.c_double * 3)(*[float(v) for v in uniform[:3]])
            rc = self.lib.ieq_step_uniform(self.handle, integrator.encode(), int(nsteps), u3, float(uniform[3]),
                                           self._ptrs(args), len(args), ctypes.byref(sec) if timed else None)
        else:
            rc = self.lib.ieq_step(self.handle, integrator.encode(), int(nsteps), int(variable_u), self._ptrs(args),
                                   len(args), ctypes.byref(sec) if timed else None)
        _capi.check(rc, self.handle)
        return sec.value if timed else None

    def diagnostics(self, b, time: float = 0.0) -> dict:
        """FP64 energy, divergence norm and error norm of b (ieq_diagnostics), whatever REAL is."""
        out = (ctypes.c_double * 3)()
        _capi.check(self.lib.ieq_diagnostics(self.handle, ctypes.c_void_p(b.data_ptr()), float(time), out), self.handle)
        return {"energy": out[0], "divergence_norm": out[1], "error_norm": out[2]}

    @property
    def last_step_fused(self) -> bool:
        """True when the last run / step call ran the fused stage kernels (ieq_last_step_path)."""
        return self.lib.ieq_last_step_path(self.handle) == 1

    @property
    def last_step_path(self) -> str:
        """Path of the last run / step call (ieq_last_step_path): "fused", "tiled" (the reference kernel
        sequence with the z-marching stage kernel), "reference" (one work-item per node) or "none"."""
        return {1: "fused", 2: "tiled", 0: "reference"}.get(self.lib.ieq_last_step_path(self.handle), "none")

    def reduce(self, name: str, *fields) -> float:
        """Run a group-reduction kernel and return the host sum of its partials."""
        out = self.reduce_output()
        self.run(name, list(fields) + [out], g=(self.info.num_nodes_pad, 1, 1), l=(self.info.w_size, 1, 1))
        return float(out.double().sum().item())

    def clean_div_cg(self, b, phi=None, abs_threshold: float = 0.0, rel_tolerance: float = 0.0, max_iterations: int = 100):
        """Divergence cleaning of b in place (CG on -Lap phi = div B, then B += grad phi)."""
        phi = self.scalar() if phi is None else phi
        res = _capi.CgResult()
        rc = self.lib.ieq_clean_div_cg(self.handle, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(phi.data_ptr()),
                                       float(abs_threshold), float(rel_tolerance), int(max_iterations), ctypes.byref(res))
        _capi.check(rc, self.handle)
        return {"divergence_norm_initial": res.divergence_norm_initial, "divergence_norm_final": res.divergence_norm_final,
                "residual_norm": res.residual_norm, "iterations": res.iterations, "seconds": res.seconds, "phi": phi}

    def synchronize(self):
        _capi.check(self.lib.ieq_synchronize(self.handle), self.handle)
