"""Per-step integrals (save_integrals_over_time) and the reference's result files, on the oracle executor.

compute_numerical_solution.m:69-89 records, before every step, the energy, the divergence norm
and the error against analytical_b, and completes the series with the final values (:218-250);
the sweep tables follow paper/convergence_hall_periodic.m:119-146, the time series
examples/energy_hall_outflow.m:150-161.
"""
import math
import os

import numpy as np

from inductioneq_amd.harness import compute_numerical_solution, initialize
from inductioneq_amd.outputs import (TABLE_HEADER, read_convergence_table, table_row, write_convergence_table,
                                     write_time_series)
from inductioneq_amd.settings import ProgramConfig


def _run(save, integrator="CarpenterKennedy2N54"):
    from pyoracle import Oracle

    L = 4 * math.pi / 3
    cfg = ProgramConfig(opset="4thOrder", nx=10, ny=10, nz=10, xmin=0, xmax=L, ymin=0, ymax=L, zmin=0, zmax=L,
                        periodic=True, hall=True, testcase="hall_periodic", w_size=16)
    init = initialize(cfg, Oracle, 0.2, 0.05)
    cfg.dt = init["dt"]
    res = compute_numerical_solution(cfg, Oracle(cfg), integrator, init["b"], init["u"], init["rho"], init["num_steps"],
                                     0.05, variable_u=True, save_integrals_over_time=save)
    return cfg, init, res


def test_integrals_over_time_series():
    for integrator in ("CarpenterKennedy2N54", "SSPRK33"):
        cfg, init, res = _run(True, integrator)
        _, _, plain = _run(False, integrator)
        n = init["num_steps"]
        for k in ("energy_over_time", "L2error_B_over_time", "L2error_divB_over_time", "time_series"):
            assert res[k].shape == (n + 1,), k
        # the analytic field written into the second register before each step does not change the run
        assert np.array_equal(res["field_b"], plain["field_b"])
        assert res["energy_over_time"][-1] == res["energy"] and res["L2error_B_over_time"][-1] == res["abs_err"]
        assert res["L2error_divB_over_time"][-1] == res["divergence_norm"]
        assert res["time_series"][0] == 0.0 and abs(res["time_series"][-1] - 0.05) < 1e-15
        # the initial field is the analytic solution at t = 0: zero error there, growing afterwards
        assert res["L2error_B_over_time"][0] <= 1e-14 * math.sqrt(res["energy_over_time"][0])
        assert res["L2error_B_over_time"][-1] > res["L2error_B_over_time"][1] > 0
        # energy (M-norm squared) of b0 equals the sum over the initial field with the M weights
        assert abs(res["energy_over_time"][0] - res["energy_over_time"][-1]) < 1e-3 * res["energy_over_time"][0]


def test_table_and_time_series_files(tmp_path):
    _, _, res = _run(True)
    forms = ("USE_UIBJ_PRODUCT", "USE_SOURCE_CENTRAL", "USE_UJBI_CENTRAL")
    path = tmp_path / "hall_periodic_convergence.txt"
    write_convergence_table(str(path), [table_row(10, 4, forms, res), table_row(20, 4, forms, res)], "gfx950")
    lines = open(path).read().splitlines(True)
    assert lines[0] == TABLE_HEADER and lines[1] == "#  gfx950 \n"
    assert lines[2].startswith("10, 4, USE_UIBJ_PRODUCT, USE_SOURCE_CENTRAL, USE_UJBI_CENTRAL, ")
    rows = read_convergence_table(str(path))
    assert [r[:2] for r in rows] == [(10, 4), (20, 4)]
    want = (res["energy"], res["abs_err"], res["divergence_norm"])  # printed with %.15e: 16 significant digits
    assert np.allclose(rows[0][6:], want, rtol=1e-15, atol=0)
    out = write_time_series(str(tmp_path / "hall_periodic_N_10_order_4.h5"), res)
    if out.endswith(".npz"):
        d = np.load(out)
        got = {k: d[k] for k in d.files}
    else:
        import h5py

        with h5py.File(out) as f:
            got = {k: f[k][()] for k in f.keys()}
    assert sorted(got) == ["L2error_B", "L2error_divB", "energy", "time"]
    assert np.array_equal(got["energy"], res["energy_over_time"]) and np.array_equal(got["time"], res["time_series"])


def test_convergence_sweep_table(tmp_path):
    """convergence_sweep (paper/convergence_hall_periodic.m:99-152) on the oracle executor: one table line per
    (N, order, forms), in the reference's loop order, and the hall_periodic rates of the operator orders."""
    from pyoracle import Oracle

    from inductioneq_amd.sweeps import FORMS_PAPER, convergence_sweep, observed_orders

    forms = [FORMS_PAPER[1], FORMS_PAPER[5]]
    path = tmp_path / "hall_periodic_convergence.txt"
    runs = convergence_sweep((12, 24), (2, 4), forms, final_time=0.05, make_executor=Oracle, path=str(path))
    rows = read_convergence_table(str(path))
    assert [(r[0], r[1], r[2]) for r in rows] == [(n, o, "USE_UIBJ_" + f[0]) for n in (12, 24) for o in (2, 4)
                                                  for f in forms]
    rates = observed_orders(runs)
    for (order, _), (rate,) in rates.items():
        assert rate > order - 0.2, (order, rate)


def test_divcleaning_sweep_table_and_series(tmp_path):
    """divcleaning_sweep (paper/divcleaning_confined_domain.m:130-212) on the oracle: per order and form triple one
    cleaned run per -Laplacian and one uncleaned run, each with its time series under the reference's file name
    and one table line in the 11-column format.  The confined_domain field is discretely divergence free to
    round-off, so no CG iteration is needed (r.r below the threshold) and the cleaned runs equal the uncleaned
    one."""
    from pyoracle import Oracle

    from inductioneq_amd.outputs import DC_TABLE_HEADER
    from inductioneq_amd.sweeps import FORMS_PAPER, divcleaning_sweep

    path = tmp_path / "confined_domain_divcleaning.txt"
    runs = divcleaning_sweep(10, (2,), [FORMS_PAPER[3]], final_time=0.1, make_executor=Oracle, path=str(path),
                             series_dir=str(tmp_path))
    assert [r["dc_form"] for r in runs] == ["WS_LN", "WS_D0", "NS_D0", ""]
    lines = open(path).read().splitlines()
    assert lines[0] + "\n" == DC_TABLE_HEADER and len(lines) == 2 + 4
    assert lines[2].split(", ")[9:] == ["1", "USE_LAPLACE_WIDE_STENCIL_LNS "]
    assert lines[5].split(", ")[9] == "0"
    names = sorted(os.path.basename(r["series"]) for r in runs)
    assert "confined_domain_N_10_order_2_product_central_product__DC_1_WS_LN" in names[0] or any(
        n.startswith("confined_domain_N_10_order_2_product_central_product__DC_1_WS_LN") for n in names)
    assert any(n.startswith("confined_domain_N_10_order_2_product_central_product__DC_0") for n in names)
    plain = runs[-1]["result"]
    for r in runs[:-1]:
        assert max(r["result"]["cg_iterations"]) == 0
        assert np.array_equal(r["result"]["field_b"], plain["field_b"])
    assert all(len(r["result"]["energy_over_time"]) == len(r["result"]["time_series"]) for r in runs)
