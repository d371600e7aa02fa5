"""The z-marching stage kernel of the reference kernel sequence (programs without the Hall term) against
the CPU oracle, through the C ABI.

The tiled kernel replaces the one-work-item-per-node stage kernel whenever the NDRange covers the
grid: SBP boundary rows with the centre clamp, SAT inflow data, a velocity field, high-order or
first-order dissipation, bounded or periodic grids.  Each case runs whole RK steps (ieq_run with the
reference RK_Step list) and checks that the tiled path ran.  Bar: FP64 1e-12 relative (max norm per
register), FP32 2e-5 (1e-4 for work registers holding dt-scaled increments).  Grid sizes leave partial
32 x 8 tiles; forced z-chunks of 2 and 5 planes put chunk boundaries next to the boundary planes.
"""
import math
import os

import numpy as np
import pytest

from inductioneq_amd.schedules import INTEGRATORS, rk_schedule
from inductioneq_amd.settings import ProgramConfig

pytestmark = pytest.mark.gpu

TOL = {"double": 1e-12, "float": 2e-5}


def _cfg(ops, n, periodic, diss, forms, tc, real, lo=None, hi=None):
    if lo is None:
        lo, hi = (0.0, 2 * math.pi) if periodic else (-1.0, 1.0)
    cfg = ProgramConfig(real=real, opset=ops, nx=n[0], ny=n[1], nz=n[2], xmin=lo, xmax=hi, ymin=lo, ymax=hi, zmin=lo,
                        zmax=hi, periodic=periodic, hall=False, dissipation=diss, form_uibj=forms[0],
                        form_source=forms[1], form_ujbi=forms[2], testcase=tc, w_size=64, ho_factor=0.7)
    cfg.dt = 0.05 * min(cfg.dx, cfg.dy, cfg.dz)
    return cfg


def _fields(cfg, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(cfg.nx) * cfg.dx + cfg.xmin
    y = np.arange(cfg.ny) * cfg.dy + cfg.ymin
    z = np.arange(cfg.nz) * cfg.dz + cfg.zmin
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    B = np.stack([np.sin(X + 0.3) * np.cos(Y), np.cos(Z - 0.2) * np.sin(X), np.sin(Y + Z)], -1).reshape(-1, 3)
    B = B + 0.05 * rng.standard_normal(B.shape)
    U = np.stack([np.cos(Y) - 0.2, np.sin(Z) + 0.4, np.cos(X) * 0.3 - 0.1], -1).reshape(-1, 3)
    U = U + 0.05 * rng.standard_normal(U.shape)
    return B, U


def _aos(cfg, a3):
    out = np.zeros((cfg.num_nodes_pad, 4), cfg.dtype)
    out[: a3.shape[0], :3] = a3
    return out


def _run_both(cfg, integrator, steps=2, variable_u=False, seed=5, env=None):
    """Run `steps` RK steps on the oracle and through ieq_run; return (prog, device regs, host regs, path)."""
    import torch
    from pyoracle import Oracle

    from inductioneq_amd.program import Program

    prog, orc = Program(cfg), Oracle(cfg)
    names, nf = rk_schedule(integrator, False, variable_u)
    B, U = _fields(cfg, seed)
    hb = [_aos(cfg, B)] + [orc.vec() for _ in range(nf - 1)]
    hJ, hu, hr, ht = orc.vec(), _aos(cfg, U), np.ones(cfg.num_nodes_pad, cfg.dtype), orc.time(0.3)
    db = [prog.upload_real4(h) for h in hb]
    dJ, du, dr = prog.upload_real4(hJ), prog.upload_real4(hu), prog.upload_real(hr)
    dt_ = torch.tensor([0.3, 0.3], dtype=prog.torch_dtype, device="cuda")
    orc.run(names * steps, hb + [hJ, hu, hr, ht])
    old = {k: os.environ.get(k) for k in (env or {})}
    os.environ.update(env or {})
    try:
        prog.run(names * steps, db + [dJ, du, dr, dt_])
        torch.cuda.synchronize()
        path = prog.last_step_path
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    assert np.allclose(dt_.cpu().numpy()[:2], ht[:2], rtol=1e-12 if cfg.real == "double" else 1e-6, atol=0)
    return prog, db, hb, path


def _check(cfg, prog, db, hb, integrator, what):
    tol = TOL[cfg.real] * (20 if integrator == "SSPRK104" else 1)
    for i in range(len(db)):
        got = prog.download_real4(db[i])[: cfg.num_nodes, :3].astype(np.float64)
        want = hb[i][: cfg.num_nodes, :3].astype(np.float64)
        e = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
        reg_tol = tol if (i == 0 or cfg.real == "double") else 1e-4
        assert e <= reg_tol, "%s register %d: rel err %.3e > %.1e" % (what, i, e, reg_tol)


FORMS = [("PRODUCT", "CENTRAL", "CENTRAL"), ("SPLIT", "SPLIT", "SPLIT"), ("CENTRAL", "ZERO", "PRODUCT"),
         ("PRODUCT", "SPLIT", "CENTRAL"), ("CENTRAL", "CENTRAL", "SPLIT")]
OPSETS = ["2ndOrder", "4thOrder", "6thOrder", "2ndOrderExtended", "4thOrderExtended", "6thOrderExtended"]


@pytest.mark.parametrize("ops", OPSETS)
@pytest.mark.parametrize("diss", [None, "HIGH_ORDER", "FIRST_ORDER"])
def test_bounded_stage_all_opsets(ops, diss):
    """Bounded grid with SAT inflow data (rotation_2D), every operator set and dissipation, every flux form
    over the cases; 37 x 21 x 19 nodes (partial tiles in x and y, boundary rows on all faces)."""
    k = OPSETS.index(ops) + 3 * [None, "HIGH_ORDER", "FIRST_ORDER"].index(diss)
    cfg = _cfg(ops, (37, 21, 19), False, diss, FORMS[k % len(FORMS)], "rotation_2D", "double")
    prog, db, hb, path = _run_both(cfg, "SSPRK33")
    assert path == "tiled"
    _check(cfg, prog, db, hb, "SSPRK33", "%s %s" % (ops, diss))


@pytest.mark.parametrize("integrator", INTEGRATORS)
def test_bounded_stage_all_integrators(integrator):
    cfg = _cfg("4thOrder", (40, 40, 40), False, None, FORMS[0], "gaussian_rotation_3D", "double")
    prog, db, hb, path = _run_both(cfg, integrator)
    assert path == "tiled"
    _check(cfg, prog, db, hb, integrator, integrator)


@pytest.mark.parametrize("zc", ["1", "2", "5"])
@pytest.mark.parametrize("ops", ["4thOrder", "6thOrderExtended"])
def test_bounded_stage_forced_z_chunks(zc, ops):
    """Many z-chunks: chunk boundaries inside and next to the boundary planes of z."""
    cfg = _cfg(ops, (34, 17, 23), False, "HIGH_ORDER", FORMS[1], "rotation_3D", "double")
    prog, db, hb, path = _run_both(cfg, "CarpenterKennedy2N54", env={"IEQ_TILED_ZC": zc})
    assert path == "tiled"
    _check(cfg, prog, db, hb, "CarpenterKennedy2N54", "zc %s" % zc)


@pytest.mark.parametrize("ops", ["2ndOrderExtended", "6thOrder"])
def test_periodic_stage_variable_velocity(ops):
    """Periodic grid with a time-dependent velocity field (calc_u before every stage, alfven_periodic_2D)."""
    cfg = _cfg(ops, (33, 18, 20), True, "FIRST_ORDER", FORMS[4], "alfven_periodic_2D", "double")
    prog, db, hb, path = _run_both(cfg, "SSPRK33", variable_u=True)
    assert path == "tiled"
    _check(cfg, prog, db, hb, "SSPRK33", ops)


@pytest.mark.parametrize("ops", ["4thOrder", "6thOrderExtended"])
@pytest.mark.parametrize("diss", [None, "HIGH_ORDER", "FIRST_ORDER"])
def test_fp32_stage(ops, diss):
    cfg = _cfg(ops, (36, 20, 18), False, diss, FORMS[0], "rotation_2D", "float")
    prog, db, hb, path = _run_both(cfg, "SSPRK33")
    assert path == "tiled"
    _check(cfg, prog, db, hb, "SSPRK33", "fp32 %s %s" % (ops, diss))


def test_tiled_equals_per_node_kernel_at_160():
    """160^3 rotation workload (4th order, SATs, u = (-y, x, 0)): three SSPRK33 steps through the tiled
    kernel equal the same steps with one work-item per node (IEQ_NO_TILED=1) to 1e-13."""
    import torch

    from inductioneq_amd.program import Program

    cfg = _cfg("4thOrder", (160, 160, 160), False, None, FORMS[0], "gaussian_rotation_3D", "double")
    prog = Program(cfg)
    names, nf = rk_schedule("SSPRK33", False, False)
    B, _ = _fields(cfg, 9)
    x = np.arange(cfg.nx) * cfg.dx + cfg.xmin
    y = np.arange(cfg.ny) * cfg.dy + cfg.ymin
    Y, X = np.meshgrid(y, x, indexing="ij")
    U = np.zeros((cfg.num_nodes, 3))
    U[:, 0] = np.tile(-Y.reshape(-1), cfg.nz)
    U[:, 1] = np.tile(X.reshape(-1), cfg.nz)
    out = {}
    for mode in ("tiled", "reference"):
        os.environ["IEQ_NO_TILED"] = "1" if mode == "reference" else "0"
        try:
            regs = [prog.upload_real4(_aos(cfg, B))] + [prog.vector() for _ in range(nf - 1)]
            args = regs + [prog.vector(), prog.upload_real4(_aos(cfg, U)), prog.scalar(1.0), prog.time_buffer(0.1)]
            prog.run(names * 3, args)
            torch.cuda.synchronize()
            assert prog.last_step_path == mode
            out[mode] = regs[0][:, : cfg.num_nodes].double().cpu()
        finally:
            os.environ.pop("IEQ_NO_TILED", None)
    e = float((out["tiled"] - out["reference"]).abs().max() / out["reference"].abs().max())
    assert e <= 1e-13, e


def test_programs_outside_the_tiled_kernel_keep_the_per_node_kernel():
    """Adaptive dissipation and grids narrower than two boundary blocks stay on the one-work-item-per-node
    stage kernel (and still match the oracle)."""
    for ops, n, diss in [("4thOrder", (20, 18, 16), "ADAPTIVE"), ("6thOrder", (11, 30, 30), None)]:
        cfg = _cfg(ops, n, False, diss, FORMS[0], "rotation_2D", "double")
        prog, db, hb, path = _run_both(cfg, "SSPRK33")
        assert path == "reference", (ops, n, diss, path)
        _check(cfg, prog, db, hb, "SSPRK33", "%s %s" % (ops, diss))
