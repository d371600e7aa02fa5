"""ieq_run_host: MatCL's run call on host arrays (copy in, run the kernel list, write every
argument back in place; compute_numerical_solution.m:39-41, initialize.m:88-93).

One SSPRK33 RK_Step kernel list (compute_numerical_solution_setup.m:142-174) of the periodic
Hall problem with HO dissipation on host REAL4 / REAL arrays: every argument written back
(b1 and the work registers b2, b3, curlB_rho, u, rho, time) equals the oracle's run of the
same list on the same arrays to 1e-12, and the call reports a positive kernel time.
"""
import math

import numpy as np
import pytest

from inductioneq_amd.settings import ProgramConfig

pytestmark = pytest.mark.gpu


def test_run_host_matches_oracle_and_writes_back():
    from pyoracle import Oracle

    from inductioneq_amd.problems import hall_dt, random_modes_field, to_aos4
    from inductioneq_amd.program import Program
    from inductioneq_amd.schedules import rk_schedule

    n, L = 16, 2 * math.pi
    cfg = ProgramConfig(opset="6thOrder", nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L, periodic=True,
                        hall=True, dissipation="HIGH_ORDER", testcase="uniform_flow", w_size=64)
    B = random_modes_field(n, seed=6)
    cfg.dt = hall_dt(B, (cfg.dx, cfg.dy, cfg.dz), 6)
    orc = Oracle(cfg)
    names, nf = rk_schedule("SSPRK33", True, False)
    want = [to_aos4(B, cfg.num_nodes_pad), orc.vec(), orc.vec(), orc.vec(), orc.vec(), orc.scal(1.0), orc.time(0.25)]
    got = [a.copy() for a in want]
    sec = Program(cfg).run_host(names, got)
    orc.run(names, want)
    assert sec > 0
    for k, (g, w) in enumerate(zip(got, want)):
        scale = max(np.abs(w).max(), 1e-300)
        assert np.abs(g - w).max() <= 1e-12 * scale, k
    assert not np.array_equal(got[0], to_aos4(B, cfg.num_nodes_pad))  # the step changed b1 in place
