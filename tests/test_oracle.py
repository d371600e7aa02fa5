"""The CPU oracle against what the mathematics fixes (no GPU).

* operators: calc_div / calc_curl / calc_grad equal Kronecker products of the 1-D table
  matrices (themselves pinned by test_tables.py), on periodic and bounded grids;
* right-hand side: uniform velocity reduces the flux-differencing form to
  -(u.grad)B (+ u div B for the ZERO source form); the Hall term equals
  -curl_D((curl_D B / rho) x B); the dissipation equals the assembled HOD matrices;
* time integration: one step on a single Fourier mode of uniform advection matches
  exp(lambda dt) to the method's order; time[0] advances by DT per step;
* analytic solutions (the reference's own testcases): relative L2 error converges at
  the expected rate (rotation_2D with SATs; hall_periodic with the Hall term);
* CG divergence cleaning: the converged potential equals a sparse direct solve of the
  same assembled -Laplacian, and the projection removes the divergence.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from inductioneq_amd.harness import clean_div_cg, compute_numerical_solution, initialize
from inductioneq_amd.schedules import INTEGRATORS, STAGES_PER_STEP, rk_schedule
from inductioneq_amd.settings import ProgramConfig
from pyoracle import Oracle
from sbp_matrices import ORDER, assemble, hod_matrix, load_tables, norm_matrix

T = load_tables()
OPS_NAME = {"2ndOrder": "2nd", "4thOrder": "4th", "6thOrder": "6th", "2ndOrderExtended": "2nd_ext",
            "4thOrderExtended": "4th_ext", "6thOrderExtended": "6th_ext"}


def kron3(Ax, Ay, Az):
    """Operators on x-fastest flattened fields: (Az x Ay x Ax)."""
    nx, ny, nz = Ax.shape[0], Ay.shape[0], Az.shape[0]
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    return (sp.kron(sp.kron(Iz, Iy), sp.csr_matrix(Ax)), sp.kron(sp.kron(Iz, sp.csr_matrix(Ay)), Ix),
            sp.kron(sp.kron(sp.csr_matrix(Az), Iy), Ix))


def _cfg(ops, n, periodic, **kw):
    lo, hi = (0.0, 2 * math.pi) if periodic else (-1.0, 1.0)
    return ProgramConfig(opset=ops, nx=n[0], ny=n[1], nz=n[2], xmin=lo, xmax=hi, ymin=lo, ymax=hi, zmin=lo, zmax=hi,
                         periodic=periodic, w_size=16, **kw)


def _rand(cfg, seed, c=3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((cfg.num_nodes, c))


def _aos(cfg, a3):
    out = np.zeros((cfg.num_nodes_pad, 4))
    out[: a3.shape[0], :3] = a3
    return out


@pytest.mark.parametrize("ops", list(OPS_NAME))
@pytest.mark.parametrize("periodic", [False, True])
def test_operators_equal_kronecker_matrices(ops, periodic):
    n = (18, 17, 19)
    cfg = _cfg(ops, n, periodic, testcase="rotation_2D")
    o = Oracle(cfg)
    t = T[OPS_NAME[ops]]
    Dx, Dy, Dz = kron3(assemble(t["diff"], n[0], periodic) / cfg.dx, assemble(t["diff"], n[1], periodic) / cfg.dy,
                       assemble(t["diff"], n[2], periodic) / cfg.dz)
    B = _rand(cfg, 1)
    b = _aos(cfg, B)
    div, curl = o.scal(), o.vec()
    o.run("calc_div", [b, div])
    o.run("calc_curl", [b, curl])
    want_div = Dx @ B[:, 0] + Dy @ B[:, 1] + Dz @ B[:, 2]
    want_curl = np.stack([Dy @ B[:, 2] - Dz @ B[:, 1], Dz @ B[:, 0] - Dx @ B[:, 2], Dx @ B[:, 1] - Dy @ B[:, 0]], 1)
    assert np.abs(div[: cfg.num_nodes] - want_div).max() < 1e-11 * np.abs(want_div).max()
    assert np.abs(curl[: cfg.num_nodes, :3] - want_curl).max() < 1e-11 * np.abs(want_curl).max()
    s = np.zeros(cfg.num_nodes_pad)
    s[: cfg.num_nodes] = B[:, 0]
    g = o.vec()
    o.run("calc_grad", [s, g])
    want = np.stack([Dx @ B[:, 0], Dy @ B[:, 0], Dz @ B[:, 0]], 1)
    assert np.abs(g[: cfg.num_nodes, :3] - want).max() < 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("ops", ["2ndOrder", "4thOrder", "6thOrder", "4thOrderExtended"])
@pytest.mark.parametrize("source", ["ZERO", "CENTRAL", "SPLIT"])
@pytest.mark.parametrize("periodic", [True, False])
def test_uniform_velocity_convection_reduces_to_advection(ops, source, periodic):
    n = (16, 14, 15)
    u = np.array([0.7, -0.4, 0.25])
    cfg = _cfg(ops, n, periodic, testcase="uniform_flow", ux=u[0], uy=u[1], uz=u[2], form_uibj="SPLIT",
               form_source=source, form_ujbi="PRODUCT")
    o = Oracle(cfg)
    t = T[OPS_NAME[ops]]
    D = kron3(assemble(t["diff"], n[0], periodic) / cfg.dx, assemble(t["diff"], n[1], periodic) / cfg.dy,
              assemble(t["diff"], n[2], periodic) / cfg.dz)
    B = _rand(cfg, 2)
    U = np.tile(u, (cfg.num_nodes, 1))
    rhs = o.rhs(_aos(cfg, B), o.vec(), _aos(cfg, U))[: cfg.num_nodes, :3]
    want = -sum(u[d] * np.stack([D[d] @ B[:, c] for c in range(3)], 1) for d in range(3))
    if source == "ZERO":
        div = sum(D[d] @ B[:, d] for d in range(3))
        want = want + np.outer(div, u)
    if not periodic:
        # bounded grid: inflow SATs act on boundary nodes only (b_boundary = 0 for uniform_flow)
        idx = np.arange(cfg.num_nodes)
        ix, iy, iz = idx % n[0], (idx // n[0]) % n[1], idx // (n[0] * n[1])
        interior = (ix > 0) & (ix < n[0] - 1) & (iy > 0) & (iy < n[1] - 1) & (iz > 0) & (iz < n[2] - 1)
        rhs, want = rhs[interior], want[interior]
    assert np.abs(rhs - want).max() < 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("ops", ["2ndOrder", "4thOrder", "6thOrder"])
def test_hall_term_is_minus_curl_of_emf(ops):
    n = (14, 13, 15)
    cfg = _cfg(ops, n, True, testcase="uniform_flow", hall=True)
    o = Oracle(cfg)
    t = T[OPS_NAME[ops]]
    Dx, Dy, Dz = kron3(assemble(t["diff"], n[0], True) / cfg.dx, assemble(t["diff"], n[1], True) / cfg.dy,
                       assemble(t["diff"], n[2], True) / cfg.dz)
    curl = lambda F: np.stack([Dy @ F[:, 2] - Dz @ F[:, 1], Dz @ F[:, 0] - Dx @ F[:, 2], Dx @ F[:, 1] - Dy @ F[:, 0]], 1)  # noqa
    B = _rand(cfg, 3)
    rho = 1.0 + 0.3 * np.random.default_rng(4).random(cfg.num_nodes)
    J = curl(B) / rho[:, None]
    E = np.cross(J, B)
    hJ = _aos(cfg, J)
    rhs = o.rhs(_aos(cfg, B), hJ, o.vec())[: cfg.num_nodes, :3]
    want = -curl(E)
    assert np.abs(rhs - want).max() < 1e-11 * np.abs(want).max()
    # and calc_curlB_rho produces J
    args = [_aos(cfg, B), o.vec(), o.vec(), o.vec(), o.vec(), np.concatenate([rho, np.ones(cfg.num_nodes_pad - cfg.num_nodes)]),
            o.time()]
    o.run("calc_curlB_rho_1_3args", args)
    assert np.abs(args[3][: cfg.num_nodes, :3] - J).max() < 1e-11 * np.abs(J).max()


@pytest.mark.parametrize("ops", list(OPS_NAME))
@pytest.mark.parametrize("periodic", [True, False])
def test_high_order_dissipation_matches_matrices(ops, periodic):
    n = (17, 16, 18)
    cfg = _cfg(ops, n, periodic, testcase="uniform_flow", dissipation="HIGH_ORDER", ho_factor=0.6)
    o = Oracle(cfg)
    t = T[OPS_NAME[ops]]
    p = ORDER[OPS_NAME[ops]]
    H = kron3(hod_matrix(t, n[0], p, periodic), hod_matrix(t, n[1], p, periodic), hod_matrix(t, n[2], p, periodic))
    B = _rand(cfg, 5)
    rhs = o.rhs(_aos(cfg, B), o.vec(), o.vec())[: cfg.num_nodes, :3]  # u = 0: convection and SATs vanish
    want = 0.6 * np.stack([(H[0] + H[1] + H[2]) @ B[:, c] for c in range(3)], 1)
    assert np.abs(rhs - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("integrator", INTEGRATORS)
def test_rk_order_on_a_fourier_mode(integrator):
    """Uniform advection of one periodic Fourier mode: B-hat is multiplied by R(lambda dt)
    per step; |R(z) - exp(z)| = O(z^(p+1)) with p = 3 (SSPRK33) or 4 (all others)."""
    n = (16, 8, 8)
    p = 3 if integrator == "SSPRK33" else 4
    errs = []
    for dt in (0.08, 0.04):
        cfg = _cfg("6thOrder", n, True, testcase="uniform_flow", ux=1.0, dt=dt)
        o = Oracle(cfg)
        x = np.arange(n[0]) * cfg.dx
        mode = np.exp(1j * 2 * x)
        B = np.zeros((cfg.num_nodes, 3))
        B[:, 1] = np.tile(mode.real, n[1] * n[2])
        B[:, 2] = np.tile(mode.imag, n[1] * n[2])
        names, nf = rk_schedule(integrator, False, False)
        fields = [_aos(cfg, B)] + [o.vec() for _ in range(nf - 1)]
        U = np.tile([1.0, 0, 0], (cfg.num_nodes, 1))
        tb = o.time()
        o.run(names, fields + [o.vec(), _aos(cfg, U), o.scal(1.0), tb])
        # semi-discrete eigenvalue: lambda = -i * symbol(2 dx) / dx
        th = 2 * cfg.dx
        c = np.array([0.75, -0.15, 1 / 60.0])
        lam = -1j * 2 * np.sum(c * np.sin(np.arange(1, 4) * th)) / cfg.dx
        got = fields[0][0, 1] + 1j * fields[0][0, 2]
        errs.append(abs(got - np.exp(lam * dt)))
        assert abs(tb[0] - dt) < 1e-15 and tb[1] == tb[0]
    rate = math.log(errs[0] / errs[1], 2)
    assert rate > p + 0.7, (integrator, errs, rate)


def _solution_error(ops, n, integrator, tc, hall, variable_u, T_end, cfl, periodic, lo, hi):
    cfg = ProgramConfig(opset=ops, nx=n, ny=n, nz=n, xmin=lo, xmax=hi, ymin=lo, ymax=hi, zmin=lo, zmax=hi, periodic=periodic,
                        hall=hall, testcase=tc, w_size=16)
    init = initialize(cfg, Oracle, cfl, T_end)
    cfg.dt = init["dt"]
    o = Oracle(cfg)
    res = compute_numerical_solution(cfg, o, integrator, init["b"], init["u"], init["rho"], init["num_steps"], T_end,
                                     variable_u=variable_u)
    return res["rel_err"], res


@pytest.mark.parametrize("ops,rate", [("2ndOrder", 1.7), ("4thOrder", 2.8), ("6thOrder", 3.6)])
def test_rotation_2D_converges(ops, rate):
    """rotation_2D (bounded grid, inflow SATs, analytic boundary data): error decreases at the order's rate."""
    e1, _ = _solution_error(ops, 16, "KennedyCarpenterLewis2R54C", "rotation_2D", False, False, 0.4, 0.5, False, -1.0, 1.0)
    e2, r2 = _solution_error(ops, 31, "KennedyCarpenterLewis2R54C", "rotation_2D", False, False, 0.4, 0.5, False, -1.0, 1.0)
    got = math.log(e1 / e2, 30 / 15)
    assert got > rate, (e1, e2, got)
    assert r2["energy"] > 0


@pytest.mark.parametrize("ops,rate", [("2ndOrder", 1.8), ("4thOrder", 3.7)])
def test_hall_periodic_converges(ops, rate):
    """hall_periodic: nonlinear Hall term + time-dependent velocity on a periodic grid, analytic solution."""
    L = 4 * math.pi / 3
    e1, _ = _solution_error(ops, 12, "CarpenterKennedy2N54", "hall_periodic", True, True, 0.3, 0.1, True, 0.0, L)
    e2, _ = _solution_error(ops, 24, "CarpenterKennedy2N54", "hall_periodic", True, True, 0.3, 0.05, True, 0.0, L)
    got = math.log(e1 / e2, 2)
    assert got > rate, (e1, e2, got)


# The reference runs CG with Euclidean dot products (kernel_dot_product.cl:56) although the
# -Laplacians are only M-self-adjoint; for the 4th/6th-order least-norm form the iteration
# (and the 6th-order Dirichlet form) does not converge on rough data (reproduced
# faithfully), so convergence is pinned where CG converges.
@pytest.mark.parametrize("ops,lap", [("4thOrder", "WIDE_STENCIL_DIRICHLET"), ("2ndOrder", "WIDE_STENCIL_DIRICHLET"),
                                     ("2ndOrder", "NARROW_STENCIL_DIRICHLET"), ("2ndOrder", "WIDE_STENCIL_LNS")])
def test_cg_matches_sparse_direct_solve(ops, lap):
    n = (12, 11, 13)
    cfg = _cfg(ops, n, False, testcase="rotation_2D", laplace=lap)
    o = Oracle(cfg)
    B = _rand(cfg, 9)
    b = _aos(cfg, B)
    t = T[OPS_NAME[ops]]
    key = {"WIDE_STENCIL_DIRICHLET": "lap_ws_d0", "NARROW_STENCIL_DIRICHLET": "lap_ns_d0", "WIDE_STENCIL_LNS": "lap_ws_ln"}[lap]
    L = kron3(assemble(t[key], n[0]) / cfg.dx**2, assemble(t[key], n[1]) / cfg.dy**2, assemble(t[key], n[2]) / cfg.dz**2)
    L = (L[0] + L[1] + L[2]).tocsr()
    D = kron3(assemble(t["diff"], n[0]) / cfg.dx, assemble(t["diff"], n[1]) / cfg.dy, assemble(t["diff"], n[2]) / cfg.dz)
    div = D[0] @ B[:, 0] + D[1] @ B[:, 1] + D[2] @ B[:, 2]
    bn = cfg.bnodes
    idx = np.arange(cfg.num_nodes)
    ix, iy, iz = idx % n[0], (idx // n[0]) % n[1], idx // (n[0] * n[1])
    inner = (ix >= bn) & (ix < n[0] - bn) & (iy >= bn) & (iy < n[1] - bn) & (iz >= bn) & (iz < n[2] - bn)
    res = clean_div_cg(cfg, o, b, 1e-11 * np.linalg.norm(div[inner]), 3000)
    got = res["phi"][: cfg.num_nodes]
    # the potential solves the assembled -Laplacian system on the interior
    resid = L[inner][:, inner] @ got[inner] - div[inner]
    assert np.linalg.norm(resid) < 1e-9 * np.linalg.norm(div[inner]), res["iterations"]
    if lap != "WIDE_STENCIL_LNS":  # non-singular (Dirichlet): unique, equals a sparse direct solve
        phi = spla.spsolve(L[inner][:, inner].tocsc(), div[inner])
        assert np.abs(got[inner] - phi).max() < 1e-8 * np.abs(phi).max()
        assert res["divergence_norm_final"] < res["divergence_norm_initial"]
    else:  # -Lap = D (M^-1 D^T M): the projection B - grad_adj(phi) is divergence free
        assert res["divergence_norm_final"] < 1e-8 * res["divergence_norm_initial"]


def test_cg_stops_on_the_absolute_threshold_like_the_reference():
    cfg = _cfg("4thOrder", (12, 12, 10), False, testcase="rotation_2D", laplace="WIDE_STENCIL_DIRICHLET")
    o = Oracle(cfg)
    b = _aos(cfg, _rand(cfg, 1))
    thr = 1e-1
    res = clean_div_cg(cfg, o, b, thr, 100)
    assert res["residual_norm"] <= thr and res["iterations"] > 0
    b2 = _aos(cfg, _rand(cfg, 1))
    res2 = clean_div_cg(cfg, o, b2, thr, res["iterations"] - 1)
    assert res2["residual_norm"] > thr  # one iteration fewer does not reach it


def _m_weights(cfg, key, periodic):
    t = T[key]
    d = [np.diag(norm_matrix(t["minv"], n, periodic)) * h for n, h in ((cfg.nx, cfg.dx), (cfg.ny, cfg.dy), (cfg.nz, cfg.dz))]
    return np.einsum("k,j,i->kji", d[2], d[1], d[0]).reshape(-1)


def _hall_energy_rate(ops, periodic, seed):
    n = (14, 13, 15)
    lo, hi = (0.0, 2 * math.pi) if periodic else (-1.0, 1.0)
    cfg = ProgramConfig(opset=ops, nx=n[0], ny=n[1], nz=n[2], xmin=lo, xmax=hi, ymin=lo, ymax=hi, zmin=lo, zmax=hi,
                        periodic=periodic, w_size=16, testcase="hall_travelling_wave", hall=True)
    o = Oracle(cfg)
    B = _rand(cfg, seed)
    hb = _aos(cfg, B)
    args = [hb, o.vec(), o.vec(), o.vec(), o.vec(), np.ones(cfg.num_nodes_pad), o.time()]
    o.run("calc_curlB_rho_1_3args", args)
    rhs = o.rhs(hb, args[3], o.vec())[: cfg.num_nodes, :3]
    w = _m_weights(cfg, OPS_NAME[ops], periodic)[:, None]
    return (w * B * rhs).sum(), (w * np.abs(B * rhs)).sum()


@pytest.mark.parametrize("ops", ["2ndOrder", "4thOrder", "6thOrder", "4thOrderExtended"])
def test_hall_term_conserves_energy_on_periodic_grids(ops):
    """B^T M RHS = -(curl B).(J x B) = 0 for the Hall term alone: skew-symmetric periodic SBP operators."""
    for seed in range(3):
        rate, scale = _hall_energy_rate(ops, True, seed)
        assert abs(rate) <= 1e-13 * scale


@pytest.mark.parametrize("ops", ["2ndOrder", "4thOrder", "6thOrder", "4thOrderExtended"])
def test_hall_outflow_sat_dissipates_energy(ops):
    """With walls, Hall_volume + the Hall outflow SAT (Hall_surface) give a non-positive M-energy rate."""
    for seed in range(3):
        rate, scale = _hall_energy_rate(ops, False, seed)
        assert rate < -1e-3 * scale


def _diss_only(ops, n, periodic, diss, B, U, **kw):
    """Dissipation term alone: RHS with it minus RHS without it (same convection code path)."""
    out = []
    for d in (diss, None):
        cfg = _cfg(ops, n, periodic, testcase="uniform_flow", dissipation=d, **kw)
        o = Oracle(cfg)
        out.append(o.rhs(_aos(cfg, B), o.vec(), _aos(cfg, U))[: cfg.num_nodes, :3])
    return out[0] - out[1], np.abs(out[1]).max()


def test_first_order_dissipation_is_rusanov_viscosity_for_constant_velocity():
    """2nd order, periodic, constant u: sum_d |u_c|/(2 dx_d) (B(+e_d) - 2B + B(-e_d)) per component (Lax-Friedrichs)."""
    n = (12, 11, 13)
    cfg = _cfg("2ndOrder", n, True)
    B = _rand(cfg, 8)
    u = np.array([0.7, -1.3, 0.4])
    got, conv = _diss_only("2ndOrder", n, True, "FIRST_ORDER", B, np.tile(u, (cfg.num_nodes, 1)))
    F = B.reshape(n[2], n[1], n[0], 3)
    want = np.zeros_like(F)
    for ax, h in ((2, cfg.dx), (1, cfg.dy), (0, cfg.dz)):
        want += np.abs(u) / (2 * h) * (np.roll(F, 1, ax) - 2 * F + np.roll(F, -1, ax))
    assert np.abs(got - want.reshape(-1, 3)).max() <= 1e-12 * max(conv, np.abs(want).max())


@pytest.mark.parametrize("ops", ["2ndOrder", "4thOrder", "6thOrder"])
def test_first_order_dissipation_dissipates_energy_with_variable_velocity(ops):
    """Weights |D_i| max(|u_k|,|u_n|) are symmetric in (n, k) on periodic grids: B.FO(B) <= 0, = 0 for constant B."""
    n = (12, 11, 13)
    cfg = _cfg(ops, n, True)
    B, U = _rand(cfg, 9), _rand(cfg, 10)
    got, conv = _diss_only(ops, n, True, "FIRST_ORDER", B, U)
    assert (B * got).sum() < -1e-3 * np.abs(B * got).sum()
    got_c, conv_c = _diss_only(ops, n, True, "FIRST_ORDER", np.ones_like(B) * [0.3, -2.0, 1.1], U)
    assert np.abs(got_c).max() <= 1e-12 * max(conv_c, 1.0)


@pytest.mark.parametrize("ops", ["2ndOrder", "4thOrder", "6thOrder"])
def test_adaptive_dissipation_limits(ops):
    """kappa = 0 everywhere (jumps below the MP2MIN*CMIN threshold) gives the high-order dissipation;
    kappa = 1 everywhere (jumps above MP2MAX*CMAX) gives the first-order dissipation."""
    n = (12, 11, 13)
    cfg = _cfg(ops, n, True)
    B, U = _rand(cfg, 11), _rand(cfg, 12)
    ad0, c0 = _diss_only(ops, n, True, "ADAPTIVE", B, U, ho_factor=0.7, mp2min=1e30, mp2max=2e30, cmin=1.0, cmax=1.0)
    ho, c1 = _diss_only(ops, n, True, "HIGH_ORDER", B, U, ho_factor=0.7)
    assert np.abs(ad0 - ho).max() <= 1e-12 * max(c0, np.abs(ho).max())
    ad1, c2 = _diss_only(ops, n, True, "ADAPTIVE", B, U, ho_factor=0.7, mp2min=0.0, mp2max=1e-300, cmin=1.0, cmax=1.0)
    fo, c3 = _diss_only(ops, n, True, "FIRST_ORDER", B, U)
    assert np.abs(ad1 - fo).max() <= 1e-12 * max(c2, np.abs(fo).max())
