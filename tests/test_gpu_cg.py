"""Device-side CG divergence cleaning (ieq_clean_div_cg) vs the oracle's reference loop.

Same inputs, same absolute stop test and iteration cap: the iteration counts agree and
the potential, the projected field and the divergence norms match to 1e-10 (FP64).
FP32 fields with FP64 dot products are compared against the FP64 oracle.
"""
import numpy as np
import pytest

from inductioneq_amd.harness import clean_div_cg
from inductioneq_amd.settings import ProgramConfig

pytestmark = pytest.mark.gpu


def _setup(ops, lap, n, real):
    from pyoracle import Oracle

    from inductioneq_amd.program import Program

    cfg = ProgramConfig(real=real, opset=ops, nx=n[0], ny=n[1], nz=n[2], laplace=lap, w_size=64, testcase="rotation_2D")
    ocfg = ProgramConfig(real="double", opset=ops, nx=n[0], ny=n[1], nz=n[2], laplace=lap, w_size=64, testcase="rotation_2D")
    rng = np.random.default_rng(21)
    B = rng.standard_normal((cfg.num_nodes, 3))
    aos = np.zeros((cfg.num_nodes_pad, 4))
    aos[: cfg.num_nodes, :3] = B
    return Program(cfg), Oracle(ocfg), cfg, aos


@pytest.mark.parametrize("ops,lap,n,maxit", [
    ("4thOrder", "WIDE_STENCIL_DIRICHLET", (40, 36, 30), 60),       # fused loop (half-width 4, 16-row tiles)
    ("2ndOrder", "NARROW_STENCIL_DIRICHLET", (33, 31, 35), 60),     # fused loop (half-width 1)
    ("2ndOrder", "WIDE_STENCIL_LNS", (24, 26, 22), 60),             # fused loop, BNODES = 0
    ("2ndOrderExtended", "WIDE_STENCIL_LNS", (26, 24, 28), 60),     # fused loop (half-width 3)
    ("4thOrder", "WIDE_STENCIL_LNS", (30, 32, 28), 12),             # fused loop (half-width 5); not M-symmetric, few iterations
    ("4thOrderExtended", "WIDE_STENCIL_DIRICHLET", (34, 30, 32), 60),  # fused loop (half-width 6, 8-row tiles)
    ("6thOrderExtended", "WIDE_STENCIL_DIRICHLET", (30, 28, 26), 60),  # half-width 9: reference kernel sequence
])
def test_cg_matches_oracle_fp64(ops, lap, n, maxit):
    import torch

    prog, orc, cfg, aos = _setup(ops, lap, n, "double")
    thr = 1e-6
    hb = aos.copy()
    want = clean_div_cg(orc.cfg, orc, hb, thr, maxit)
    db = prog.upload_real4(aos)
    got = prog.clean_div_cg(db, abs_threshold=thr, max_iterations=maxit)
    torch.cuda.synchronize()
    assert got["iterations"] == want["iterations"]
    phi_w = want["phi"][: cfg.num_nodes]
    phi_g = prog.download_real(got["phi"])[: cfg.num_nodes]
    assert np.abs(phi_g - phi_w).max() <= 1e-10 * np.abs(phi_w).max()
    b_g = prog.download_real4(db)[: cfg.num_nodes, :3]
    assert np.abs(b_g - hb[: cfg.num_nodes, :3]).max() <= 1e-10 * np.abs(hb[:, :3]).max()
    for k in ("divergence_norm_initial", "divergence_norm_final", "residual_norm"):
        assert abs(got[k] - want[k]) <= 1e-9 * max(abs(want[k]), 1e-300), k


def test_cg_mixed_precision_relative_stop():
    """FP32 stencils + FP64 dots vs all-FP64, both to a relative residual of 1e-5."""
    import torch

    n = (48, 48, 48)
    p64, orc, cfg64, aos = _setup("4thOrder", "WIDE_STENCIL_DIRICHLET", n, "double")
    p32, _, cfg32, _ = _setup("4thOrder", "WIDE_STENCIL_DIRICHLET", n, "float")
    b64, b32 = p64.upload_real4(aos), p32.upload_real4(aos.astype(np.float32))
    r64 = p64.clean_div_cg(b64, rel_tolerance=1e-5, max_iterations=500)
    r32 = p32.clean_div_cg(b32, rel_tolerance=1e-5, max_iterations=500)
    torch.cuda.synchronize()
    assert r64["residual_norm"] <= 1e-5 * r64["divergence_norm_initial"]
    assert r32["residual_norm"] <= 1e-5 * r32["divergence_norm_initial"] * 1.0001
    phi64 = p64.download_real(r64["phi"])[: cfg64.num_nodes]
    phi32 = p32.download_real(r32["phi"])[: cfg32.num_nodes].astype(np.float64)
    assert np.abs(phi32 - phi64).max() <= 1e-3 * np.abs(phi64).max()
    assert abs(r32["divergence_norm_final"] - r64["divergence_norm_final"]) <= 1e-2 * r64["divergence_norm_final"]
