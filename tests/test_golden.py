"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces them.  GPU: the HIP library reproduces them through the
C ABI -- J and the right-hand side via the reference kernels, and one RK step -- to 1e-12
relative (FP64).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from golden_cases import CASES, case_config  # noqa: E402

from inductioneq_amd.schedules import rk_schedule  # noqa: E402


def _load(name):
    return dict(np.load(os.path.join(HERE, "golden", name + ".npz")))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(name):
    from pyoracle import Oracle

    g = _load(name)
    cfg = case_config(name)
    o = Oracle(cfg, nthreads=1)
    names, nf = rk_schedule(CASES[name]["integrator"], cfg.hall, False)
    J = o.vec()
    if cfg.hall:
        jargs = [g["B"].copy()] + [o.vec() for _ in range(nf - 1)] + [J, g["U"].copy(), g["rho"].copy(), o.time(0.0)]
        o.run("calc_curlB_rho_1_3args" if nf == 3 else "calc_curlB_rho_1_2args", jargs)
    assert _rel(o.rhs(g["B"], J, g["U"], 0.0), g["rhs"]) <= 1e-14
    args = [g["B"].copy()] + [o.vec() for _ in range(nf - 1)] + [o.vec(), g["U"].copy(), g["rho"].copy(), o.time(0.0)]
    o.run(names, args)
    assert _rel(args[0], g["step_b1"]) <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_hip_reproduces_golden(name):
    import torch

    from inductioneq_amd.program import Program

    g = _load(name)
    cfg = case_config(name)
    prog = Program(cfg)
    names, nf = rk_schedule(CASES[name]["integrator"], cfg.hall, False)
    n = cfg.num_nodes
    regs = [prog.upload_real4(g["B"])] + [prog.vector() for _ in range(nf - 1)]
    args = regs + [prog.vector(), prog.upload_real4(g["U"]), prog.upload_real(g["rho"]), prog.time_buffer(0.0)]
    prog.run(names, args)
    torch.cuda.synchronize()
    got = prog.download_real4(regs[0])[:n, :3]
    assert _rel(got, g["step_b1"][:n, :3]) <= 1e-12
    t = args[-1].double().cpu().numpy()
    assert abs(t[0] - g["time"][0]) <= 1e-15 and abs(t[1] - g["time"][1]) <= 1e-15
