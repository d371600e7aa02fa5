"""Ensembles (several periodic solves in one persistent launch per stage) vs the oracle.

Each member is checked against the oracle running the reference stage schedule on
that member alone (its own grid, DT, uniform u and rho): FP64 to 1e-12 relative.
The FP32 ensemble is checked against the FP64 ensemble: 1e-5 relative after 10 steps.
"""
import math

import numpy as np
import pytest

from inductioneq_amd.schedules import rk_schedule
from inductioneq_amd.settings import ProgramConfig

pytestmark = pytest.mark.gpu

MEMBERS = [  # (nx, ny, nz, rho, u, dt factor)
    (24, 20, 18, 1.25, (0.0, 0.0, 0.0), 0.02),
    (16, 28, 12, 0.8, (0.3, -0.2, 0.1), 0.015),
    (70, 10, 9, 2.0, (0.0, 0.0, 0.0), 0.01),
]


def _cfg(ops, n, hall, diss, real="double"):
    nx, ny, nz, rho, u, f = n
    L = 2 * math.pi
    cfg = ProgramConfig(real=real, opset=ops, nx=nx, ny=ny, nz=nz, xmin=0, xmax=L, ymin=0, ymax=L * 0.75, zmin=0,
                        zmax=L * 1.25, periodic=True, hall=hall, dissipation=diss, testcase="uniform_flow", ux=u[0],
                        uy=u[1], uz=u[2], ho_factor=0.5, w_size=64)
    cfg.dt = f * min(cfg.dx, cfg.dy, cfg.dz) ** 2
    return cfg


def _field(cfg, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(cfg.nx) * cfg.dx
    y = np.arange(cfg.ny) * cfg.dy
    z = np.arange(cfg.nz) * cfg.dz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    ky, kz = 2 * math.pi / (cfg.ny * cfg.dy), 2 * math.pi / (cfg.nz * cfg.dz)
    B = np.stack([np.sin(X + 0.3) * np.cos(ky * Y), np.cos(kz * Z - 0.2) * np.sin(X), np.sin(ky * Y + kz * Z)])
    return B + 0.02 * rng.standard_normal(B.shape)  # (3, nx, ny, nz)


def _oracle_steps(cfg, integrator, B, rho, nsteps):
    from pyoracle import Oracle

    orc = Oracle(cfg)
    names, nf = rk_schedule(integrator, cfg.hall, False)
    h = np.zeros((cfg.num_nodes_pad, 4), cfg.dtype)
    h[: cfg.num_nodes, :3] = B.transpose(3, 2, 1, 0).reshape(-1, 3)
    U = np.zeros((cfg.num_nodes_pad, 4), cfg.dtype)
    U[: cfg.num_nodes, :3] = (cfg.ux, cfg.uy, cfg.uz)
    r = np.zeros((cfg.num_nodes_pad,), cfg.dtype)
    r[: cfg.num_nodes] = rho
    regs = [h] + [orc.vec() for _ in range(nf - 1)]
    orc.run(names * nsteps, regs + [orc.vec(), U, r, orc.time(0.0)])
    return h[: cfg.num_nodes, :3].reshape(cfg.nz, cfg.ny, cfg.nx, 3).transpose(3, 2, 1, 0)


def _ensemble(ops, hall, diss, integrator, real, seeds):
    from inductioneq_amd.ensemble import Ensemble, MemberSpec
    from inductioneq_amd.program import Program

    cfgs = [_cfg(ops, n, hall, diss, real) for n in MEMBERS]
    prog = Program(cfgs[0])
    specs = [MemberSpec(c.nx, c.ny, c.nz, c.dx, c.dy, c.dz, c.dt, n[3], n[4]) for c, n in zip(cfgs, MEMBERS)]
    ens = Ensemble(prog, specs, integrator)
    Bs = [_field(c, s) for c, s in zip(cfgs, seeds)]
    for i, B in enumerate(Bs):
        ens.set_state(i, B)
    return cfgs, ens, Bs


@pytest.mark.parametrize("shape", ["auto", "1", "2", "32"])
@pytest.mark.parametrize("integrator", ["SSPRK33", "CarpenterKennedy2N54", "ToulorgeDesmet2N84F"])
@pytest.mark.parametrize("ops,hall,diss", [("6thOrder", True, "HIGH_ORDER"), ("4thOrder", False, "HIGH_ORDER"),
                                           ("2ndOrder", True, None)])
def test_ensemble_members_match_oracle(integrator, ops, hall, diss, shape, monkeypatch):
    """Tile shape chosen by the tuner, or forced (IEQ_ENS_SHAPE): 1 and 2 the one-kernel Hall stage (E formed in
    the stage kernel; FP64 tiles 32 x 8 and 16 x 16), 32 the EMF kernel + stage kernel pair."""
    import torch

    if shape != "auto":
        monkeypatch.setenv("IEQ_ENS_SHAPE", shape)

    cfgs, ens, Bs = _ensemble(ops, hall, diss, integrator, "double", seeds=(1, 2, 3))
    nsteps = 2
    ens.step(nsteps)
    torch.cuda.synchronize()
    tile = ens.config()["tile"]
    assert tile in ("64x4", "64x8", "32x8", "hall_32x8", "hall_16x16")
    if shape in ("1", "2"):
        assert tile == (("hall_32x8" if shape == "1" else "hall_16x16") if hall else "64x8")
    elif shape == "32":
        assert tile == "32x8"
    for i, (cfg, B) in enumerate(zip(cfgs, Bs)):
        want = _oracle_steps(cfg, integrator, B, MEMBERS[i][3], nsteps)
        got = ens.state(i).numpy()
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= 1e-12, "member %d: rel err %.3e" % (i, err)


def test_tuner_overrides_pin_the_launch(monkeypatch):
    """IEQ_ENS_SHAPE / IEQ_ENS_ZC / IEQ_ENS_ORDER restrict the tuner to one launch form, z-chunk and tile order;
    the pinned launch reports them and its members still equal the oracle (FP64, 1e-12)."""
    monkeypatch.setenv("IEQ_ENS_SHAPE", "2")
    monkeypatch.setenv("IEQ_ENS_ZC", "8")
    monkeypatch.setenv("IEQ_ENS_ORDER", "1")
    cfgs, ens, Bs = _ensemble("6thOrder", True, "HIGH_ORDER", "SSPRK33", "double", seeds=(2, 3, 4))
    ens.step(2)
    c = ens.config()
    assert (c["tile"], c["z_chunk"], c["order"]) == ("hall_16x16", 8, "largest_first"), c
    for i, (cfg, B) in enumerate(zip(cfgs, Bs)):
        want = _oracle_steps(cfg, "SSPRK33", B, MEMBERS[i][3], 2)
        assert np.abs(ens.state(i).numpy() - want).max() / np.abs(want).max() <= 1e-12


def test_ensemble_repeated_steps_continue_the_solve():
    """Two calls of 2 steps equal one oracle run of 4 steps (the graph replays from b1)."""
    cfgs, ens, Bs = _ensemble("6thOrder", True, "HIGH_ORDER", "CarpenterKennedy2N54", "double", seeds=(4, 5, 6))
    ens.step(2)
    ens.step(2)
    for i, (cfg, B) in enumerate(zip(cfgs, Bs)):
        want = _oracle_steps(cfg, "CarpenterKennedy2N54", B, MEMBERS[i][3], 4)
        got = ens.state(i).numpy()
        assert np.abs(got - want).max() / np.abs(want).max() <= 1e-12


def test_ensemble_odd_step_count_rejected_for_2n():
    from inductioneq_amd._capi import IEQError

    _, ens, _ = _ensemble("4thOrder", False, None, "CarpenterKennedy2N54", "double", seeds=(1, 2, 3))
    with pytest.raises(IEQError):
        ens.step(3)


@pytest.mark.parametrize("shape", ["auto", "1", "2", "32"])
def test_ensemble_fp32_within_1e5_of_fp64(shape, monkeypatch):
    """Mixed-precision claim for batched runs: FP32 members stay within 1e-5 (relative, max norm) of FP64, for
    the tuner's choice and for each forced FP32 launch form (one-kernel Hall stage with 64 x 8 or 32 x 8 tiles,
    EMF + stage kernels with 32 x 8 tiles)."""
    _, e64, _ = _ensemble("6thOrder", True, "HIGH_ORDER", "SSPRK33", "double", seeds=(7, 8, 9))
    e64.step(10)
    if shape != "auto":
        monkeypatch.setenv("IEQ_ENS_SHAPE", shape)
    _, e32, _ = _ensemble("6thOrder", True, "HIGH_ORDER", "SSPRK33", "float", seeds=(7, 8, 9))
    e32.step(10)
    if shape != "auto":
        want = {"1": "hall_64x8", "2": "hall_32x8", "32": "32x8"}[shape]
        assert e32.config()["tile"] == want
    for i in range(len(MEMBERS)):
        a, b = e64.state(i).numpy(), e32.state(i).numpy().astype(np.float64)
        err = np.abs(a - b).max() / np.abs(a).max()
        assert err <= 1e-5, "member %d: %.3e" % (i, err)


@pytest.mark.parametrize("integrator,nrep", [("SSPRK33", 4), ("CarpenterKennedy2N54", 3)])
def test_graph_replays_each_match_oracle(integrator, nrep, monkeypatch):
    """The captured step forced into graph replay (IEQ_ENS_GRAPH=1): every replay, with eager device work and
    host read-backs between replays, equals the oracle run for that many steps (FP64, 1e-12)."""
    import torch

    monkeypatch.setenv("IEQ_ENS_GRAPH", "1")
    cfgs, ens, Bs = _ensemble("6thOrder", True, "HIGH_ORDER", integrator, "double", seeds=(7, 8, 9))
    per = 2 if integrator == "CarpenterKennedy2N54" else 1  # steps per captured graph
    for rep in range(1, nrep + 1):
        ens.step(per)
        assert ens.config()["graph"]
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64)  # eager work between replays
        assert torch.isfinite(junk.sum()).item()
        torch.cuda.synchronize()
        for i, (cfg, B) in enumerate(zip(cfgs, Bs)):
            want = _oracle_steps(cfg, integrator, B, MEMBERS[i][3], rep * per)
            got = ens.state(i).numpy()
            assert np.abs(got - want).max() / np.abs(want).max() <= 1e-12, (rep, i)
