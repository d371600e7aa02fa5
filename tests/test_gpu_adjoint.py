"""Discrete adjoint of the periodic SSPRK33 Hall solve (ieq_adjoint_*).

Checks, all FP64 unless stated:
* the adjoint's forward equals ieq_step on the same settings (same fused kernels);
* gradients w.r.t. B0, u (3) and the Hall coefficient match central finite differences
  of the loss 0.5 h^3 |B(T) - B_target|^2 (eps = 1e-4 relative to the direction) to 1e-6;
* dot-product test <J v, w> = <v, J^T w> with J = dB(T)/dB0 (J v by central differences);
* the gradient does not depend on the number of snapshots (bitwise);
* recomputation counts follow the binomial schedule;
* the FP32 gradient agrees with the FP64 one to 1e-4 relative.
"""
import math

import numpy as np
import pytest

from inductioneq_amd.settings import ProgramConfig

pytestmark = pytest.mark.gpu

N = (16, 18, 14)


def _setup(real="double", source="CENTRAL", ops="6thOrder", nsteps=6, snapshots=2, hall=True, diss="HIGH_ORDER"):
    import torch

    from inductioneq_amd.adjoint import AdjointSolver
    from inductioneq_amd.problems import hall_dt
    from inductioneq_amd.program import Program

    L = 2 * math.pi
    cfg = ProgramConfig(real=real, opset=ops, nx=N[0], ny=N[1], nz=N[2], xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L,
                        periodic=True, hall=hall, dissipation=diss, ho_factor=0.5, testcase="uniform_flow",
                        form_source=source, w_size=64)
    rng = np.random.default_rng(3)
    x = np.arange(N[0]) * cfg.dx
    y = np.arange(N[1]) * cfg.dy
    z = np.arange(N[2]) * cfg.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    B = np.stack([np.sin(Y) + 0.3 * np.cos(Z), np.cos(X) * 0.7 + np.sin(Z), 1.0 + 0.5 * np.sin(X + Y)]).reshape(3, -1)
    B = B + 0.05 * rng.standard_normal(B.shape)
    cfg.dt = hall_dt(B.T, (cfg.dx, cfg.dy, cfg.dz), 6, c_hall=1.3, cfl=0.3)
    prog = Program(cfg)
    solver = AdjointSolver(prog, nsteps, snapshots)
    n = cfg.num_nodes
    b0 = prog.vector()
    b0[:, :n] = torch.as_tensor(B, dtype=prog.torch_dtype)
    tgt = prog.vector()
    tgt[:, :n] = torch.as_tensor(B + 0.1 * rng.standard_normal(B.shape), dtype=prog.torch_dtype)
    return cfg, prog, solver, b0, tgt


def _loss(cfg, bT, tgt):
    h3 = cfg.dx * cfg.dy * cfg.dz
    d = (bT - tgt)[:, : cfg.num_nodes].double()
    return 0.5 * h3 * float((d * d).sum())


def _grad(cfg, solver, b0, tgt, u, ch):
    bT = solver.forward(b0, u, ch)
    lam = (bT - tgt) * (cfg.dx * cfg.dy * cfg.dz)
    lam[:, cfg.num_nodes:] = 0
    return solver.backward(lam)


def test_adjoint_forward_equals_fused_step():
    import torch

    cfg, prog, solver, b0, tgt = _setup()
    u, ch = (0.3, -0.2, 0.1), 0.8
    bT = solver.forward(b0, u, ch)
    b1 = b0.clone()
    uf = prog.vector()
    uf[0, : cfg.num_nodes], uf[1, : cfg.num_nodes], uf[2, : cfg.num_nodes] = u
    args = [b1, prog.vector(), prog.vector(), prog.vector(), uf, prog.scalar(1.0 / ch), prog.time_buffer()]
    prog.step("SSPRK33", solver.nsteps, args)
    torch.cuda.synchronize()
    n = cfg.num_nodes
    err = float((bT[:, :n] - b1[:, :n]).abs().max() / b1[:, :n].abs().max())
    assert err <= 1e-14, err


@pytest.mark.parametrize("source", ["CENTRAL", "ZERO"])
@pytest.mark.parametrize("ops", ["6thOrder", "4thOrder"])
def test_adjoint_gradients_match_central_differences(source, ops):
    import torch

    cfg, prog, solver, b0, tgt = _setup(source=source, ops=ops)
    u, ch = [0.3, -0.2, 0.1], 0.8
    g, gc, gu = _grad(cfg, solver, b0, tgt, u, ch)
    n = cfg.num_nodes
    eps = 1e-4
    # direction in B0
    v = prog.vector()
    v[:, :n] = torch.as_tensor(np.random.default_rng(11).standard_normal((3, n)), dtype=torch.float64)
    lp = _loss(cfg, solver.forward(b0 + eps * v, u, ch), tgt)
    lm = _loss(cfg, solver.forward(b0 - eps * v, u, ch), tgt)
    fd = (lp - lm) / (2 * eps)
    ad = float((g[:, :n] * v[:, :n]).sum())
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
    # velocity components
    for k in range(3):
        up, um = list(u), list(u)
        up[k] += eps
        um[k] -= eps
        fd = (_loss(cfg, solver.forward(b0, up, ch), tgt) - _loss(cfg, solver.forward(b0, um, ch), tgt)) / (2 * eps)
        assert abs(fd - gu[k]) <= 1e-6 * max(abs(fd), 1e-3 * max(abs(v) for v in gu)), (k, fd, gu[k])
    # Hall coefficient
    fd = (_loss(cfg, solver.forward(b0, u, ch + eps), tgt) - _loss(cfg, solver.forward(b0, u, ch - eps), tgt)) / (2 * eps)
    assert abs(fd - gc) <= 1e-6 * abs(fd), (fd, gc)


def test_adjoint_gradient_at_zero_velocity_and_without_hall_term():
    """u = 0 (convection off in the forward) still has a velocity gradient; the linear problem too."""
    import torch

    for hall in (True, False):
        cfg, prog, solver, b0, tgt = _setup(hall=hall, nsteps=4, snapshots=1)
        u, ch = [0.0, 0.0, 0.0], 1.0
        g, gc, gu = _grad(cfg, solver, b0, tgt, u, ch)
        eps = 1e-4
        for k in range(3):
            up, um = list(u), list(u)
            up[k] += eps
            um[k] -= eps
            fd = (_loss(cfg, solver.forward(b0, up, ch), tgt) - _loss(cfg, solver.forward(b0, um, ch), tgt)) / (2 * eps)
            assert abs(fd - gu[k]) <= 1e-6 * max(abs(fd), 1e-3 * max(abs(v) for v in gu)), (hall, k, fd, gu[k])
        if not hall:
            assert gc == 0.0
        n = cfg.num_nodes
        v = prog.vector()
        v[:, :n] = torch.as_tensor(np.random.default_rng(5).standard_normal((3, n)), dtype=torch.float64)
        fd = (_loss(cfg, solver.forward(b0 + eps * v, u, ch), tgt) - _loss(cfg, solver.forward(b0 - eps * v, u, ch), tgt)) / (
            2 * eps)
        g, _, _ = _grad(cfg, solver, b0, tgt, u, ch)
        ad = float((g[:, :n] * v[:, :n]).sum())
        assert abs(fd - ad) <= 1e-6 * abs(fd), (hall, fd, ad)


def test_dot_product_test():
    import torch

    cfg, prog, solver, b0, _ = _setup(nsteps=5, snapshots=1)
    u, ch = [0.2, 0.1, -0.3], 1.3
    n = cfg.num_nodes
    rng = np.random.default_rng(21)
    v, w = prog.vector(), prog.vector()
    v[:, :n] = torch.as_tensor(rng.standard_normal((3, n)), dtype=torch.float64)
    w[:, :n] = torch.as_tensor(rng.standard_normal((3, n)), dtype=torch.float64)
    eps = 1e-5
    Jv = (solver.forward(b0 + eps * v, u, ch) - solver.forward(b0 - eps * v, u, ch)) / (2 * eps)
    solver.forward(b0, u, ch)
    JTw, _, _ = solver.backward(w)
    lhs = float((Jv[:, :n] * w[:, :n]).sum())
    rhs = float((v[:, :n] * JTw[:, :n]).sum())
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs), (lhs, rhs)


def test_gradient_independent_of_snapshot_count():
    import torch

    out = []
    counts = []
    for snaps in (0, 1, 3, 9):
        cfg, prog, solver, b0, tgt = _setup(nsteps=9, snapshots=snaps)
        g, gc, gu = _grad(cfg, solver, b0, tgt, [0.1, 0.2, 0.3], 0.9)
        torch.cuda.synchronize()
        out.append((g[:, : cfg.num_nodes].cpu(), gc, tuple(gu)))
        counts.append(solver.counts())
    for g, gc, gu in out[1:]:
        assert torch.equal(g, out[0][0]) and gc == out[0][1] and gu == out[0][2]
    # no snapshots: every step is recomputed from B0 (n(n-1)/2 extra forward steps); all stored: none
    assert counts[0] == (9 + 36, 9)
    assert counts[-1] == (9, 9)
    assert all(c[1] == 9 for c in counts)


def test_fp32_gradient_close_to_fp64():
    import torch

    res = {}
    for real in ("double", "float"):
        cfg, prog, solver, b0, tgt = _setup(real=real, nsteps=6, snapshots=2)
        g, gc, gu = _grad(cfg, solver, b0, tgt, [0.3, -0.2, 0.1], 0.8)
        res[real] = (g[:, : cfg.num_nodes].double().cpu(), gc, np.array(gu))
    g64, gc64, gu64 = res["double"]
    g32, gc32, gu32 = res["float"]
    assert float((g32 - g64).abs().max() / g64.abs().max()) <= 1e-4
    assert abs(gc32 - gc64) <= 1e-3 * abs(gc64)
    assert np.abs(gu32 - gu64).max() <= 1e-3 * np.abs(gu64).max()


def test_autograd_function_matches_solver():
    """hall_solve (torch.autograd.Function) returns the solver's gradients for b0, u and c_H."""
    import torch

    from inductioneq_amd.adjoint import hall_solve

    cfg, prog, solver, b0, tgt = _setup(nsteps=4, snapshots=1)
    u0, ch0 = [0.3, -0.2, 0.1], 0.8
    h3 = cfg.dx * cfg.dy * cfg.dz
    b = b0.clone().requires_grad_(True)
    u = torch.tensor(u0, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(ch0, dtype=torch.float64, requires_grad=True)
    bT = hall_solve(b, u, c, solver)
    loss = 0.5 * h3 * ((bT - tgt) ** 2).sum()
    loss.backward()
    g, gc, gu = _grad(cfg, solver, b0, tgt, u0, ch0)
    assert torch.equal(b.grad, g)
    assert float(c.grad) == gc and [float(v) for v in u.grad] == gu
    assert abs(float(loss) - _loss(cfg, solver.forward(b0, u0, ch0), tgt)) <= 1e-12 * float(loss)


def _cfg4(real, nsteps, snapshots):
    """The gradient workload at its full size: 96^3 periodic, 6th order + HO dissipation, SSPRK33, u = (0.2, -0.1,
    0.15), c_H = 1, B0 = (0,0,1) + random modes (seed 1), target from seed 2."""
    import torch

    from inductioneq_amd.adjoint import AdjointSolver
    from inductioneq_amd.problems import hall_dt, random_modes_field
    from inductioneq_amd.program import Program

    n, L = 96, 2 * math.pi
    B0, Bt = random_modes_field(n, seed=1), random_modes_field(n, seed=2)
    h = L / n
    cfg = ProgramConfig(real=real, opset="6thOrder", nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L,
                        periodic=True, hall=True, dissipation="HIGH_ORDER", testcase="uniform_flow", w_size=256,
                        dt=hall_dt(B0, (h, h, h), 6, c_hall=1.0, cfl=0.4))
    prog = Program(cfg)
    solver = AdjointSolver(prog, nsteps, snapshots)
    b0, tgt = prog.vector(), prog.vector()
    b0[:, : cfg.num_nodes] = torch.as_tensor(B0.T, dtype=prog.torch_dtype)
    tgt[:, : cfg.num_nodes] = torch.as_tensor(Bt.T, dtype=prog.torch_dtype)
    return cfg, prog, solver, b0, tgt


def test_dot_product_test_at_96():
    """<J v, w> = <v, J^T w> at the gradient workload's size (FP64, 5 steps, J v by central differences)."""
    import torch

    cfg, prog, solver, b0, _ = _cfg4("double", 5, 2)
    u, ch = [0.2, -0.1, 0.15], 1.0
    n = cfg.num_nodes
    rng = np.random.default_rng(31)
    v, w = prog.vector(), prog.vector()
    v[:, :n] = torch.as_tensor(rng.standard_normal((3, n)), dtype=torch.float64)
    w[:, :n] = torch.as_tensor(rng.standard_normal((3, n)), dtype=torch.float64)
    eps = 1e-5
    Jv = (solver.forward(b0 + eps * v, u, ch) - solver.forward(b0 - eps * v, u, ch)) / (2 * eps)
    solver.forward(b0, u, ch)
    JTw, _, _ = solver.backward(w)
    lhs = float((Jv[:, :n] * w[:, :n]).sum())
    rhs = float((v[:, :n] * JTw[:, :n]).sum())
    print("dot-product test 96^3: %.15e vs %.15e" % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs), (lhs, rhs)


def test_binomial_15_snapshots_equals_store_all_at_96():
    """200 steps at 96^3 (FP32): the gradient with 15 binomial snapshots is bitwise the one with every stage
    input stored, and the recomputation count is the binomial schedule's."""
    import torch

    u, ch = [0.2, -0.1, 0.15], 1.0
    out = []
    for snaps in (15, 200):
        cfg, prog, solver, b0, tgt = _cfg4("float", 200, snaps)
        g, gc, gu = _grad(cfg, solver, b0, tgt, u, ch)
        torch.cuda.synchronize()
        out.append((g.cpu(), gc, tuple(gu), solver.counts()))
        del solver, prog
    (g15, gc15, gu15, c15), (ga, gca, gua, ca) = out
    assert torch.equal(g15, ga) and gc15 == gca and gu15 == gua
    assert ca == (200, 200)
    assert c15[1] == 200 and 200 < c15[0] <= 3 * 200
    assert float(g15.abs().max()) > 0
