"""ieq_diagnostics_noalloc (no allocation, results in device memory) against ieq_diagnostics.

Two calls in a row through the C ABI, the second on a smaller grid (another program) than the
first: each result triple equals the allocating entry point's bitwise (same kernels, same fixed
block count and summation order).  Run alone in a fresh process or after the other tests.
"""
import ctypes
import math

import numpy as np
import pytest

from inductioneq_amd.settings import ProgramConfig

pytestmark = pytest.mark.gpu


def _run(ops, n, tc, real, t):
    import torch

    from inductioneq_amd.program import Program

    lo, hi = (-1.0, 1.0)
    cfg = ProgramConfig(real=real, opset=ops, nx=n[0], ny=n[1], nz=n[2], xmin=lo, xmax=hi, ymin=lo, ymax=hi, zmin=lo,
                        zmax=hi, testcase=tc, w_size=64)
    prog = Program(cfg)
    b = prog.vector()
    prog.run("analytical_b", [b, prog.time_buffer(t)])
    rng = np.random.default_rng(n[0])
    b[:, : cfg.num_nodes] += torch.as_tensor(0.01 * rng.standard_normal((3, cfg.num_nodes)), dtype=b.dtype, device="cuda")
    out_dev = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    rc = prog.lib.ieq_diagnostics_noalloc(prog.handle, ctypes.c_void_p(b.data_ptr()), float(t),
                                          ctypes.c_void_p(out_dev.data_ptr()))
    assert rc == 0, prog.lib.ieq_last_error(prog.handle)
    prog.synchronize()
    got = out_dev.cpu().tolist()
    want = prog.diagnostics(b, t)
    return got, [want["energy"], want["divergence_norm"], want["error_norm"]]


def test_noalloc_diagnostics_equal_allocating_entry_point():
    # first call on a 64 x 60 x 62 grid (930 partial blocks of the 1024 the scratch holds), second on a smaller one
    for ops, n, tc, real, t in [("4thOrder", (64, 60, 62), "rotation_2D", "double", 0.3),
                                ("6thOrder", (20, 18, 22), "rotation_3D", "float", 0.1)]:
        got, want = _run(ops, n, tc, real, t)
        assert got == want, (n, got, want)
        assert all(math.isfinite(v) and v > 0 for v in got)


def test_noalloc_rejects_null_output():
    from inductioneq_amd.program import Program

    cfg = ProgramConfig(opset="2ndOrder", nx=8, ny=8, nz=8, testcase="rotation_2D", w_size=64)
    prog = Program(cfg)
    b = prog.vector()
    assert prog.lib.ieq_diagnostics_noalloc(prog.handle, ctypes.c_void_p(b.data_ptr()), 0.0, None) != 0
