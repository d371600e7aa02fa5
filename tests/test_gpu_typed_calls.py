"""Typed per-kernel entry points (include/inductioneq_kernels.h) on the GPU, one per buffer count.

ieq_<name>(prog, g, l, buffers..., t) must do exactly what ieq_run does for the one name: for
init_b (1 buffer), analytical_b and calc_div (2), norm2_diff (3) and the SSPRK33 stage list with
calc_curlB_rho_*_3args and calc_time_3_args (7; kernel_time_integrator.cl:168-303), the typed
calls and ieq_run on copies of the same inputs give bitwise identical outputs.
"""
import ctypes
import math

import pytest

from inductioneq_amd import _capi
from inductioneq_amd.settings import ProgramConfig

pytestmark = pytest.mark.gpu


def test_typed_calls_equal_ieq_run():
    import torch

    from inductioneq_amd.program import Program
    from inductioneq_amd.schedules import rk_schedule

    L = 4 * math.pi / 3
    cfg = ProgramConfig(opset="4thOrder", nx=20, ny=18, nz=16, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L,
                        periodic=True, hall=True, testcase="hall_periodic", w_size=64)
    cfg.dt = 1e-3
    prog = Program(cfg)
    lib = prog.lib
    G = (ctypes.c_uint32 * 3)(cfg.nx, cfg.ny, cfg.nz)
    Lr = (ctypes.c_uint32 * 3)(cfg.w_size, 1, 1)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def typed(name, *bufs, g=G):
        _capi.check(getattr(lib, "ieq_" + name)(prog.handle, g, Lr, *[P(b) for b in bufs], None), prog.handle)

    # 1 and 2 buffers
    b_t, b_r = prog.vector(), prog.vector()
    typed("init_b", b_t)
    prog.run("init_b", [b_r])
    tt, ana_t, ana_r = prog.time_buffer(0.3), prog.vector(), prog.vector()
    typed("analytical_b", ana_t, tt)
    prog.run("analytical_b", [ana_r, tt])
    div_t, div_r = prog.scalar(), prog.scalar()
    typed("calc_div", ana_t, div_t)
    prog.run("calc_div", [ana_r, div_r])
    # 3 buffers (a reduction over the padded node range, one partial per l[0] nodes: I_Tech('g_range'))
    gred = (cfg.num_nodes_pad, 1, 1)
    out_t, out_r = prog.reduce_output(), prog.reduce_output()
    typed("norm2_diff", b_t, ana_t, out_t, g=(ctypes.c_uint32 * 3)(*gred))
    prog.run("norm2_diff", [b_r, ana_r, out_r], g=gred)
    torch.cuda.synchronize()
    for x, y in ((b_t, b_r), (ana_t, ana_r), (div_t, div_r), (out_t, out_r)):
        assert torch.equal(x, y)
    assert out_t.abs().sum().item() > 0
    # 7 buffers: one SSPRK33 step (variable u) as typed calls vs one ieq_run on the list
    names, nf = rk_schedule("SSPRK33", True, True)
    assert nf == 3

    def regs(b0):
        u = prog.vector()
        prog.run("init_u", [u])
        return [b0.clone(), prog.vector(), prog.vector(), prog.vector(), u, prog.scalar(1.0), prog.time_buffer(0.0)]

    rt, rr = regs(b_t), regs(b_r)
    for name in names:
        typed(name, *rt)
    prog.run(names, rr)
    torch.cuda.synchronize()
    for x, y in zip(rt, rr):
        assert torch.equal(x, y)
    assert not torch.equal(rt[0], b_t)  # the step moved b1
