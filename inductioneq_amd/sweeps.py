"""Parameter sweeps in the form of the reference's paper scripts.

``convergence_sweep`` restates the loop of paper/convergence_hall_periodic.m:99-152 (and its
siblings convergence_rotation_3D.m, convergence_confined_domain.m:104-150 with ``base=CONFINED_DOMAIN``,
a fixed CFL and no variable velocity): for every grid size N, SBP order and flux-form triple it
sets the CFL, runs ``initialize`` and ``compute_numerical_solution``, and appends one line of the
reference's table format (``outputs.table_row``).

``divcleaning_sweep`` restates paper/divcleaning_confined_domain.m:130-212 (and
divcleaning_hall_outflow.m:137-200): for every order and form triple, a run with divergence
cleaning after every step for each -Laplacian, then one run without; each run saves its time
series (``outputs.write_time_series``, the reference's file names) and one table line.  The executor factory defaults to the HIP library; any object with
the executor interface of ``harness`` works.
"""
from __future__ import annotations

import math
from dataclasses import replace
from typing import Callable, List, Optional, Sequence, Tuple

from .harness import compute_numerical_solution, initialize
from .outputs import (DC_TABLE_HEADER, dc_series_name, dc_table_row, table_row, write_convergence_table,
                      write_time_series)
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


# convergence_confined_domain.m:16-94, divcleaning_confined_domain.m:24-94: [0, 1]^3, bounded, no Hall term,
# static velocity u = B0, no dissipation, CFL 0.95 (not scaled with N)
CONFINED_DOMAIN = ProgramConfig(opset="4thOrder", xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                                periodic=False, hall=False, dissipation=None, testcase="confined_domain")

# divcleaning_hall_outflow.m:24-112: hall_periodic's field and velocity on a bounded [0, 4 pi / 3]^3 grid (Hall
# outflow SATs), time-dependent velocity, CFL 0.95 / N
HALL_OUTFLOW = replace(HALL_PERIODIC, periodic=False)

# the -Laplacians of divcleaning_confined_domain.m:127-128 and their file labels
DC_FORMS: List[Tuple[str, str]] = [("WIDE_STENCIL_LNS", "WS_LN"), ("WIDE_STENCIL_DIRICHLET", "WS_D0"),
                                   ("NARROW_STENCIL_DIRICHLET", "NS_D0")]


def convergence_sweep(Ns: Sequence[int], orders: Sequence[int], forms: Sequence[Tuple[str, str, str]],
                      base: ProgramConfig = HALL_PERIODIC, final_time: float = 1.0, cfl_factor: float = 0.95,
                      integrator: str = "CarpenterKennedy2N54", variable_u: bool = True, extended: bool = False,
                      make_executor: Optional[Callable[[ProgramConfig], object]] = None, path: Optional[str] = None,
                      device: str = "gfx950", cfl_scaled: bool = True) -> List[dict]:
    """Run the sweep; returns one dict per run (N, order, forms, result, row) and, with ``path``, writes the
    table there.  ``extended`` selects the *OrderExtended operators (I_RunOps('operator_form')); ``cfl_scaled``
    sets CFL = cfl_factor / N (convergence_hall_periodic.m) rather than cfl_factor (convergence_confined_domain.m)."""
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
                init = initialize(cfg, make_executor, cfl_factor / n if cfl_scaled else cfl_factor, final_time)
                cfg.dt = init["dt"]
                res = compute_numerical_solution(cfg, make_executor(cfg), integrator, init["b"], init["u"], init["rho"],
                                                 init["num_steps"], final_time, variable_u=variable_u)
                names = ("USE_UIBJ_" + f[0], "USE_SOURCE_" + f[1], "USE_UJBI_" + f[2])
                runs.append({"N": n, "order": order, "forms": names, "result": res, "row": table_row(n, order, names, res)})
    if path:
        write_convergence_table(path, [r["row"] for r in runs], device)
    return runs


def divcleaning_sweep(n: int, orders: Sequence[int], forms: Sequence[Tuple[str, str, str]],
                      base: ProgramConfig = CONFINED_DOMAIN, final_time: float = 2.0, cfl: float = 0.95,
                      dc_forms: Sequence[Tuple[str, str]] = tuple(DC_FORMS), threshold: float = 1e-3,
                      max_iterations: int = 50, integrator: str = "CarpenterKennedy2N54", variable_u: bool = False,
                      make_executor: Optional[Callable[[ProgramConfig], object]] = None, path: Optional[str] = None,
                      series_dir: Optional[str] = None, prefix: str = "confined_domain",
                      device: str = "gfx950") -> List[dict]:
    """divcleaning_confined_domain.m:130-212: for every order and form triple, one run per -Laplacian with
    clean_div after every step (absolute threshold, max_iterations), then one run without cleaning; every run
    records the integrals over time.  Returns one dict per run (order, forms, dc_form, result, row); writes the
    table to ``path`` and the time series into ``series_dir``."""
    import os

    if make_executor is None:
        from .harness import HipExecutor

        make_executor = HipExecutor
    runs = []
    for order in orders:
        for f in forms:
            names = ("USE_UIBJ_" + f[0], "USE_SOURCE_" + f[1], "USE_UJBI_" + f[2])
            for clean in (True, False):
                for lap, label in (dc_forms if clean else dc_forms[:1]):
                    opset = {2: "2ndOrder", 4: "4thOrder", 6: "6thOrder"}[order]
                    cfg = replace(base, opset=opset, nx=n, ny=n, nz=n, dx=None, dy=None, dz=None, form_uibj=f[0],
                                  form_source=f[1], form_ujbi=f[2], laplace=lap if clean else None)
                    init = initialize(cfg, make_executor, cfl, final_time)
                    cfg.dt = init["dt"]
                    res = compute_numerical_solution(cfg, make_executor(cfg), integrator, init["b"], init["u"], init["rho"],
                                                     init["num_steps"], final_time, variable_u=variable_u,
                                                     save_integrals_over_time=True,
                                                     divergence_cleaning=(threshold, max_iterations) if clean else None)
                    dc_define = "USE_LAPLACE_" + lap if clean else ""
                    row = dc_table_row(n, order, names, res, dc_define)
                    run = {"N": n, "order": order, "forms": names, "dc_form": label if clean else "", "result": res,
                           "row": row}
                    if series_dir:
                        run["series"] = write_time_series(
                            os.path.join(series_dir, dc_series_name(prefix, n, order, names, label if clean else "")), res)
                    runs.append(run)
    if path:
        write_convergence_table(path, [r["row"] for r in runs], device, header=DC_TABLE_HEADER)
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
