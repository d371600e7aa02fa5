"""Parameter sweeps in the form of the reference's paper scripts.

``convergence_sweep`` restates the loop of paper/convergence_hall_periodic.m:99-152 (and its
siblings convergence_rotation_3D.m, convergence_confined_domain.m): for every grid size N,
SBP order and flux-form triple it sets CFL = cfl_factor / N, runs ``initialize`` and
``compute_numerical_solution``, and appends one line of the reference's table format
(``outputs.table_row``).  The executor factory defaults to the HIP library; any object with
the executor interface of ``harness`` works.
"""
from __future__ import annotations

import math
from dataclasses import replace
from typing import Callable, List, Optional, Sequence, Tuple

from .harness import compute_numerical_solution, initialize
from .outputs import table_row, write_convergence_table
from .settings import ProgramConfig

# the flux-form triples of convergence_hall_periodic.m:109-116 (USE_UIBJ_*, USE_SOURCE_*, USE_UJBI_*)
FORMS_PAPER: List[Tuple[str, str, str]] = [
    ("CENTRAL", "ZERO", "CENTRAL"),
    ("CENTRAL", "CENTRAL", "CENTRAL"),
    ("SPLIT", "CENTRAL", "SPLIT"),
    ("PRODUCT", "CENTRAL", "PRODUCT"),
    ("PRODUCT", "CENTRAL", "SPLIT"),
    ("PRODUCT", "CENTRAL", "CENTRAL"),
]

# convergence_hall_periodic.m:12-62: [0, 4 pi / 3]^3, periodic, Hall term, variable u, no dissipation
HALL_PERIODIC = ProgramConfig(opset="4thOrder", xmin=0.0, xmax=4 * math.pi / 3, ymin=0.0, ymax=4 * math.pi / 3,
                              zmin=0.0, zmax=4 * math.pi / 3, periodic=True, hall=True, dissipation=None,
                              testcase="hall_periodic")


def convergence_sweep(Ns: Sequence[int], orders: Sequence[int], forms: Sequence[Tuple[str, str, str]],
                      base: ProgramConfig = HALL_PERIODIC, final_time: float = 1.0, cfl_factor: float = 0.95,
                      integrator: str = "CarpenterKennedy2N54", variable_u: bool = True, extended: bool = False,
                      make_executor: Optional[Callable[[ProgramConfig], object]] = None, path: Optional[str] = None,
                      device: str = "gfx950") -> List[dict]:
    """Run the sweep; returns one dict per run (N, order, forms, result, row) and, with ``path``, writes the
    table there.  ``extended`` selects the *OrderExtended operators (I_RunOps('operator_form'))."""
    if make_executor is None:
        from .harness import HipExecutor

        make_executor = HipExecutor
    runs = []
    for n in Ns:
        for order in orders:
            for f in forms:
                opset = {2: "2ndOrder", 4: "4thOrder", 6: "6thOrder"}[order] + ("Extended" if extended else "")
                # dx = dy = dz = None: the spacing rule of initialize.m:12-21 runs again for this N
                cfg = replace(base, opset=opset, nx=n, ny=n, nz=n, dx=None, dy=None, dz=None, form_uibj=f[0],
                              form_source=f[1], form_ujbi=f[2])
                init = initialize(cfg, make_executor, cfl_factor / n, final_time)
                cfg.dt = init["dt"]
                res = compute_numerical_solution(cfg, make_executor(cfg), integrator, init["b"], init["u"], init["rho"],
                                                 init["num_steps"], final_time, variable_u=variable_u)
                names = ("USE_UIBJ_" + f[0], "USE_SOURCE_" + f[1], "USE_UJBI_" + f[2])
                runs.append({"N": n, "order": order, "forms": names, "result": res, "row": table_row(n, order, names, res)})
    if path:
        write_convergence_table(path, [r["row"] for r in runs], device)
    return runs


def observed_orders(runs: List[dict]) -> dict:
    """log2-style rates between consecutive N of each (order, forms): {(order, forms): [rate, ...]}."""
    by = {}
    for r in runs:
        by.setdefault((r["order"], r["forms"]), []).append((r["N"], r["result"]["abs_err"]))
    out = {}
    for k, v in by.items():
        v.sort()
        out[k] = [math.log(e0 / e1) / math.log(n1 / n0) for (n0, e0), (n1, e1) in zip(v, v[1:])]
    return out
