"""ctypes front-end of the C oracle -- TEST INFRASTRUCTURE, NOT THE PRODUCT.

Only ``tests/``, ``__graft_entry__.smoke()`` and the benchmark's ``cpu_baseline`` leg
may import this module.  Fields are numpy arrays in the reference's host layout:
vector fields ``(num_nodes_pad, 4)`` AoS REAL4, scalar fields ``(num_nodes_pad,)``,
``time`` a REAL array of length >= 2.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class _Cfg(ctypes.Structure):
    _fields_ = [
        ("opset", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("xmin", ctypes.c_double), ("ymin", ctypes.c_double), ("zmin", ctypes.c_double),
        ("dt", ctypes.c_double), ("periodic", ctypes.c_int), ("hall", ctypes.c_int), ("dissipation", ctypes.c_int),
        ("ho_factor", ctypes.c_double), ("mp2min", ctypes.c_double), ("mp2max", ctypes.c_double),
        ("cmin", ctypes.c_double), ("cmax", ctypes.c_double),
        ("form_uibj", ctypes.c_int), ("form_source", ctypes.c_int), ("form_ujbi", ctypes.c_int),
        ("testcase", ctypes.c_int), ("laplace", ctypes.c_int), ("bnodes", ctypes.c_int),
        ("ux", ctypes.c_double), ("uy", ctypes.c_double), ("uz", ctypes.c_double), ("nthreads", ctypes.c_int),
    ]


_LIBS = {}


def build():
    """Compile the oracle libraries (gcc; seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(real: str):
    if real not in _LIBS:
        path = os.path.join(_HERE, "ieq_oracle_%s.so" % ("f64" if real == "double" else "f32"))
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.ieqo_run_kernel.restype = ctypes.c_int
        lib.ieqo_run_list.restype = ctypes.c_int
        lib.ieqo_rhs.restype = ctypes.c_int
        assert lib.ieqo_sizeof_cfg() == ctypes.sizeof(_Cfg), "oracle config layout mismatch"
        _LIBS[real] = lib
    return _LIBS[real]


class Oracle:
    """Runs reference kernels by name on host arrays (AoS REAL4 layout)."""

    def __init__(self, cfg, nthreads: int = 0):
        from inductioneq_amd.settings import (DISSIPATION_FORMS, SOURCE_FORMS, TESTCASES, UIBJ_FORMS, UJBI_FORMS)

        self.cfg = cfg
        self.dtype = np.float64 if cfg.real == "double" else np.float32
        self.lib = _lib(cfg.real)
        c = _Cfg()
        c.opset = cfg.opset_index
        c.nx, c.ny, c.nz = cfg.nx, cfg.ny, cfg.nz
        c.dx, c.dy, c.dz = cfg.dx, cfg.dy, cfg.dz
        c.xmin, c.ymin, c.zmin = cfg.xmin, cfg.ymin, cfg.zmin
        c.dt = cfg.dt
        c.periodic, c.hall = int(cfg.periodic), int(cfg.hall)
        c.dissipation = DISSIPATION_FORMS[cfg.dissipation]
        c.ho_factor, c.mp2min, c.mp2max, c.cmin, c.cmax = cfg.ho_factor, cfg.mp2min, cfg.mp2max, cfg.cmin, cfg.cmax
        c.form_uibj = UIBJ_FORMS.index(cfg.form_uibj)
        c.form_source = SOURCE_FORMS.index(cfg.form_source)
        c.form_ujbi = UJBI_FORMS.index(cfg.form_ujbi)
        c.testcase = TESTCASES.index(cfg.testcase)
        c.laplace = cfg.laplace_index
        c.bnodes = cfg.bnodes
        c.ux, c.uy, c.uz = cfg.ux, cfg.uy, cfg.uz
        c.nthreads = nthreads
        self._c = c

    # --- allocation helpers in the reference host layout ---------------------------
    def vec(self):
        return np.zeros((self.cfg.num_nodes_pad, 4), self.dtype)

    def scal(self, fill=0.0):
        return np.full((self.cfg.num_nodes_pad,), fill, self.dtype)

    def time(self, t0=0.0):
        return np.array([t0, t0, 0, 0], self.dtype)

    def _ptrs(self, args):
        arr = (ctypes.c_void_p * len(args))()
        for i, a in enumerate(args):
            assert isinstance(a, np.ndarray) and a.dtype == self.dtype and a.flags.c_contiguous, (i, a.dtype)
            arr[i] = a.ctypes.data
        return arr

    def run(self, names, args: Sequence[np.ndarray], g=None, l=None):
        """Run one kernel name or a list of names with one argument list (MatCL style)."""
        if isinstance(names, str):
            names = [names]
        g = g or (self.cfg.nx, self.cfg.ny, self.cfg.nz)
        l = l or (self.cfg.w_size, 1, 1)
        G = (ctypes.c_uint32 * 3)(*[int(v) for v in g])
        L = (ctypes.c_uint32 * 3)(*[int(v) for v in l])
        cn = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        rc = self.lib.ieqo_run_list(ctypes.byref(self._c), cn, len(names), G, L, self._ptrs(args), len(args))
        if rc != 0:
            raise RuntimeError("oracle kernel list %s failed: %d" % (names, rc))

    def rhs(self, b, J, u, t=0.0):
        out = self.vec()
        self.lib.ieqo_rhs(ctypes.byref(self._c), ctypes.c_void_p(b.ctypes.data), ctypes.c_void_p(J.ctypes.data),
                          ctypes.c_void_p(u.ctypes.data), ctypes.c_double(t), ctypes.c_void_p(out.ctypes.data))
        return out

    # --- executor protocol of inductioneq_amd.harness ---------------------------------
    def host_vec(self, f):
        return f[: self.cfg.num_nodes, :3].astype(np.float64)

    def dev_vec(self, arr3):
        out = self.vec()
        out[: arr3.shape[0], :3] = arr3
        return out

    def host_scal(self, f):
        return f[: self.cfg.num_nodes].astype(np.float64)

    def dev_scal(self, arr):
        out = self.scal()
        out[: len(arr)] = arr
        return out

    def coef_buffer(self):
        return np.zeros((1,), self.dtype)

    def set_coef(self, buf, v):
        buf[0] = v

    def reduce(self, name, *fields):
        """Run a reduction kernel and return the host-side sum of the group partials."""
        out = np.zeros((self.cfg.num_groups,), self.dtype)
        self.run(name, list(fields) + [out], g=(self.cfg.num_nodes_pad, 1, 1), l=(self.cfg.w_size, 1, 1))
        return float(np.sum(out.astype(np.float64)))
