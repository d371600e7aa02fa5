"""Write the golden fixtures: oracle outputs for small seeded inputs.

The reference ships no golden vectors (its only known answers are the analytic testcase
solutions and the coefficient tables), so these fixtures freeze the CPU oracle's outputs
on small grids after the oracle was pinned by the tests in tests/test_tables.py and
tests/test_oracle.py.  They let the GPU tests check the HIP library without rebuilding
the oracle and catch any later drift of the oracle itself.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from golden_cases import CASES, case_config, case_inputs  # noqa: E402
from inductioneq_amd.schedules import rk_schedule  # noqa: E402
from pyoracle import Oracle  # noqa: E402


def main():
    for name in CASES:
        cfg = case_config(name)
        o = Oracle(cfg, nthreads=1)
        B, U, rho = case_inputs(cfg, name)
        names, nf = rk_schedule(CASES[name]["integrator"], cfg.hall, False)
        regs = [B.copy()] + [o.vec() for _ in range(nf - 1)]
        args = regs + [o.vec(), U.copy(), rho.copy(), o.time(0.0)]
        # RHS of B: J = curl B / rho first (Hall), then DB (kernel_time_integrator.cl:9-31)
        J = o.vec()
        if cfg.hall:
            jargs = [B.copy()] + [o.vec() for _ in range(nf - 1)] + [J, U.copy(), rho.copy(), o.time(0.0)]
            o.run("calc_curlB_rho_1_3args" if nf == 3 else "calc_curlB_rho_1_2args", jargs)
        rhs = o.rhs(B, J, U, 0.0)
        o.run(names, args)  # one RK step of the reference schedule
        np.savez_compressed(os.path.join(HERE, name + ".npz"), B=B, U=U, rho=rho, rhs=rhs, step_b1=args[0],
                            time=args[-1])
        print("wrote", name, cfg.nx, cfg.ny, cfg.nz)


if __name__ == "__main__":
    main()
