"""Host harness: the reference's MATLAB driver flow over a kernel executor.

Restates, on top of any executor exposing the reference kernels by name,
* ``initialize``            -- matlab/+induction_eq/initialize.m:3-120 (grid spacing,
  ``init_b/u/rho``, ``dt = cfl*min(dx)/max|u|``, ``num_steps = ceil(T/dt)``, ``dt = T/num_steps``);
* ``compute_numerical_solution`` -- compute_numerical_solution.m:10-250 (the repeated
  ``RK_Step`` list, then ``analytical_b`` + ``norm2_diff`` for the error, ``calc_div`` +
  ``norm2_S`` for the divergence norm, ``norm2`` for the energy);
* ``clean_div_cg``          -- matlab/+divergence_cleaning/clean_div_CG.m:3-112 and
  +induction_eq/clean_div.m (CG on the -Laplacian with beta = rr/old, alpha = rr/(u.q),
  absolute stop rr <= thr^2, then the projection).

An executor provides ``vec() / scal(fill) / time(t0)`` allocations, ``run(names, args,
g, l)``, ``reduce(name, *fields) -> float`` (host sum of the group partials) and
``host_vec(field) -> (n, 3)`` / ``dev_vec(array)`` conversions.  :class:`HipExecutor`
wraps :class:`inductioneq_amd.program.Program`.
"""
from __future__ import annotations

import math
import time as _time
from typing import Callable, Dict

import numpy as np

from .schedules import rk_schedule
from .settings import ProgramConfig


class HipExecutor:
    """Executor over the HIP library (device fields are torch tensors)."""

    def __init__(self, cfg: ProgramConfig):
        from .program import Program

        self.cfg = cfg
        self.prog = Program(cfg)

    def vec(self):
        return self.prog.vector()

    def scal(self, fill=0.0):
        return self.prog.scalar(fill)

    def time(self, t0=0.0):
        return self.prog.time_buffer(t0)

    def run(self, names, args, g=None, l=None):
        return self.prog.run(names, args, g=g, l=l)

    def step(self, integrator, nsteps, args, variable_u=False):
        return self.prog.step(integrator, nsteps, args, variable_u=variable_u)

    def clean_div(self, b, threshold, max_iterations):
        """clean_div.m through the library's CG (ieq_clean_div_cg, absolute threshold like the reference)."""
        return self.prog.clean_div_cg(b, abs_threshold=threshold, max_iterations=max_iterations)

    def reduce(self, name, *fields):
        return self.prog.reduce(name, *fields)

    def host_vec(self, f):
        return f[:, : self.cfg.num_nodes].double().cpu().numpy().T.copy()

    def dev_vec(self, arr3):
        import torch

        t = self.prog.vector()
        t[:, : self.cfg.num_nodes] = torch.as_tensor(np.ascontiguousarray(arr3.T), dtype=t.dtype, device=t.device)
        return t

    def host_scal(self, f):
        return f[: self.cfg.num_nodes].double().cpu().numpy().copy()

    def dev_scal(self, arr):
        import torch

        t = self.prog.scalar()
        t[: self.cfg.num_nodes] = torch.as_tensor(np.asarray(arr), dtype=t.dtype, device=t.device)
        return t

    def set_time(self, tbuf, t0):
        tbuf.fill_(t0)

    def coef_buffer(self):
        return self.prog.scalar(0.0)[:1].clone()

    def set_coef(self, buf, v):
        buf.fill_(v)

    def sync(self):
        import torch

        torch.cuda.synchronize()


def initialize(cfg: ProgramConfig, make_executor: Callable[[ProgramConfig], object], cfl: float, final_time: float):
    """initialize.m:3-120: fields from the testcase, dt and num_steps from the CFL rule."""
    ex = make_executor(cfg)
    b, u, rho = ex.vec(), ex.vec(), ex.scal(1.0)
    ex.run("init_b", [b])
    ex.run("init_u", [u])
    ex.run("init_rho", [rho])
    uh = ex.host_vec(u)
    u_max = float(np.sqrt((uh**2).sum(axis=1)).max())
    dt = cfl * min(cfg.dx, cfg.dy, cfg.dz) / u_max
    num_steps = int(math.ceil(final_time / dt))
    dt = final_time / num_steps
    return {"b": ex.host_vec(b), "u": ex.host_vec(u), "rho": ex.host_scal(rho), "dt": dt, "num_steps": num_steps}


def compute_numerical_solution(cfg: ProgramConfig, ex, integrator: str, b0: np.ndarray, u0: np.ndarray, rho0: np.ndarray,
                               num_steps: int, final_time: float, variable_u: bool = False, fused: bool = False,
                               analytic: bool = True, save_integrals_over_time: bool = False,
                               divergence_cleaning=None) -> Dict[str, object]:
    """compute_numerical_solution.m: advance num_steps, then error / divergence / energy.

    save_integrals_over_time (I_RunOps('save_integrals_over_time'), compute_numerical_solution.m:15-19,
    69-89, 218-250): before every step record the energy (norm2 of b), the divergence norm
    (calc_div + norm2_S) and the error against analytical_b at the current time (written into the
    second field register, as the reference does, and compared with norm2_diff); the final values
    complete the series, returned as energy_over_time, L2error_B_over_time, L2error_divB_over_time
    (num_steps + 1 entries) and time = linspace(0, final_time, num_steps + 1).

    divergence_cleaning = (absolute_error_threshold, max_iterations) is I_DC('divergence_cleaning')
    (compute_numerical_solution.m:37-68,121-...): after every step the field is cleaned by clean_div
    (+induction_eq/clean_div.m:3-81: CG on the program's -Laplacian to the absolute threshold, then the
    projection), through the executor's own clean_div when it has one (the library's ieq_clean_div_cg),
    else through the reference kernel sequence of clean_div_cg below.  Requires a program with the
    divergence-cleaning module (cfg.laplace).  The number of CG iterations of every step is returned as
    cg_iterations."""
    names, nf = rk_schedule(integrator, cfg.hall, variable_u)
    fields = [ex.dev_vec(b0)] + [ex.vec() for _ in range(nf - 1)]
    J, u, rho, tb = ex.vec(), ex.dev_vec(u0), ex.dev_scal(rho0), ex.time(0.0)
    args = fields + [J, u, rho, tb]
    series = {"energy": [], "L2error_B": [], "L2error_divB": []}
    cg_iterations = []
    t0 = _time.perf_counter()
    if save_integrals_over_time or divergence_cleaning is not None:
        div_t = ex.scal()
        for _ in range(num_steps):
            if save_integrals_over_time:
                series["energy"].append(ex.reduce("norm2", fields[0]))
                ex.run("calc_div", [fields[0], div_t])
                series["L2error_divB"].append(math.sqrt(max(ex.reduce("norm2_S", div_t), 0.0)))
                ex.run("analytical_b", [fields[1], tb])
                series["L2error_B"].append(math.sqrt(max(ex.reduce("norm2_diff", fields[0], fields[1]), 0.0)))
            if fused:
                ex.step(integrator, 1, args, variable_u=variable_u)
            else:
                ex.run(names, args)
            if divergence_cleaning is not None:
                thr, max_it = divergence_cleaning
                if hasattr(ex, "clean_div"):
                    r = ex.clean_div(fields[0], thr, max_it)
                else:
                    r = clean_div_cg(cfg, ex, fields[0], thr, max_it)
                cg_iterations.append(r["iterations"])
    elif fused:
        ex.step(integrator, num_steps, args, variable_u=variable_u)
    else:
        block = 15000  # RK_block_size, compute_numerical_solution.m:14
        left = num_steps
        while left > 0:
            k = min(block, left)
            ex.run(names * k, args)
            left -= k
    if hasattr(ex, "sync"):
        ex.sync()
    runtime = _time.perf_counter() - t0
    res = {"runtime": runtime, "field_b": ex.host_vec(fields[0]), "time": ex.host_scal(tb)[:2] if hasattr(ex, "host_scal") else None}
    if divergence_cleaning is not None:
        res["cg_iterations"] = cg_iterations
    if analytic:
        ana, tt = ex.vec(), ex.time(final_time)
        ex.run("analytical_b", [ana, tt])
        abs_err = math.sqrt(max(ex.reduce("norm2_diff", fields[0], ana), 0.0))
        res["abs_err"] = abs_err
        res["rel_err"] = abs_err / math.sqrt(ex.reduce("norm2", ana))
    div = ex.scal()
    ex.run("calc_div", [fields[0], div])
    res["divergence_norm"] = math.sqrt(max(ex.reduce("norm2_S", div), 0.0))
    res["energy"] = ex.reduce("norm2", fields[0])
    if save_integrals_over_time:
        res["energy_over_time"] = np.array(series["energy"] + [res["energy"]])
        res["L2error_divB_over_time"] = np.array(series["L2error_divB"] + [res["divergence_norm"]])
        if analytic:
            res["L2error_B_over_time"] = np.array(series["L2error_B"] + [res["abs_err"]])
        res["time_series"] = np.linspace(0.0, final_time, num_steps + 1)
    return res


def clean_div_cg(cfg: ProgramConfig, ex, b, threshold: float, max_iterations: int) -> Dict[str, object]:
    """clean_div_CG.m:3-112 with the reference's kernels and host scalar loop."""
    bn = cfg.bnodes
    g_dc = (cfg.nx - 2 * bn, cfg.ny - 2 * bn, cfg.nz - 2 * bn)
    g_full = (cfg.nx, cfg.ny, cfg.nz)
    divB, phi, lap, u = ex.scal(), ex.scal(), ex.scal(), ex.scal()
    ex.run("calc_div_divcleaning", [b, divB], g=g_dc)
    r = divB  # field_r(:) = field_b_div(:); divB is not read again before the final calc_div
    rr = ex.reduce("dot_product", r, r)
    old = 1.0
    initial = math.sqrt(rr)
    tol = threshold * threshold
    it = 0
    coef = ex.coef_buffer()
    while rr > tol and it < max_iterations:
        ex.set_coef(coef, rr / old)
        ex.run("calc_axpy", [u, r, coef], g=g_dc)
        ex.run("calc_laplace_divcleaning", [u, lap], g=g_dc)
        alpha = rr / ex.reduce("dot_product", u, lap)
        ex.set_coef(coef, alpha)
        ex.run("calc_xpay", [phi, u, coef], g=g_dc)
        ex.set_coef(coef, -alpha)
        ex.run("calc_xpay", [r, lap, coef], g=g_dc)
        old = rr
        rr = ex.reduce("dot_product", r, r)
        it += 1
    ex.run("projector_divcleaning", [b, phi], g=g_full)
    ex.run("calc_div", [b, divB], g=g_full)
    final = math.sqrt(max(ex.reduce("norm2_S", divB), 0.0))
    return {"divergence_norm_initial": initial, "divergence_norm_final": final, "residual_norm": math.sqrt(rr),
            "iterations": it, "phi": phi, "b": b}
