"""The C ABI from a plain C program (tests/c_driver/ieq_driver.c): no Python, no torch.

CPU: the driver compiles and links against the in-tree library.  GPU: it runs 1000
CarpenterKennedy2N54 steps of the periodic Hall problem through ieq_step, through ieq_run
with the reference kernel list once per step, through the typed per-kernel entry points
with the same list, and through one ieq_run on the 1000-step list (16000 names, the
reference's repmat(RK_Step, 1, k)); all four results match the oracle to 1e-11, and the
list calls take the fused kernels while the typed single-kernel calls run the reference
kernels.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from inductioneq_amd.settings import ProgramConfig

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_driver")


def _build():
    subprocess.run(["make", "-s", "-C", HERE], check=True, capture_output=True)
    return os.path.join(HERE, "ieq_driver")


def test_c_driver_builds():
    exe = _build()
    assert os.access(exe, os.X_OK)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_c_driver_matches_oracle(tmp_path):
    from pyoracle import Oracle

    from inductioneq_amd.problems import hall_dt, random_modes_field, to_aos4
    from inductioneq_amd.schedules import rk_schedule

    exe = os.path.join(HERE, "ieq_driver")
    if not os.access(exe, os.X_OK):
        exe = _build()
    n, L = 16, 2 * math.pi
    cfg = ProgramConfig(opset="6thOrder", nx=n, ny=n, nz=n, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L, periodic=True,
                        hall=True, dissipation="HIGH_ORDER", testcase="uniform_flow", w_size=64)
    B = random_modes_field(n, seed=4)
    cfg.dt = hall_dt(B, (cfg.dx, cfg.dy, cfg.dz), 6, cfl=0.3)
    host = to_aos4(B, cfg.num_nodes_pad)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    host.astype(np.float64).tofile(inp)
    nsteps = 1000
    r = subprocess.run([exe, cfg.settings_string()] + cfg.modules() + ["--", "CarpenterKennedy2N54", str(nsteps),
                                                                         str(inp), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    paths = [l for l in r.stdout.splitlines() if l.startswith("ieq_driver: paths")][0].split()[2:]
    assert paths == ["1", "1", "0", "1"], r.stdout
    got = np.fromfile(out, dtype=np.float64).reshape(4, cfg.num_nodes_pad, 4)
    orc = Oracle(cfg)
    names, nf = rk_schedule("CarpenterKennedy2N54", True, False)
    hb = [host.copy(), orc.vec()]
    orc.run(names * nsteps, hb + [orc.vec(), orc.vec(), orc.scal(1.0), orc.time()])
    want = hb[0][: cfg.num_nodes, :3]
    for path in range(4):
        err = np.abs(got[path, : cfg.num_nodes, :3] - want).max() / np.abs(want).max()
        assert err <= 1e-11, (path, err)
