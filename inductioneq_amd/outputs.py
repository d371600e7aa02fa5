"""Result files in the reference's formats.

* ``write_convergence_table`` -- the sweep tables of paper/convergence_*.m (e.g.
  convergence_hall_periodic.m:119-146): a header line, the device line, then one line per run
  ``N, order, form_uiBj, form_source, form_ujBi, runtime, energy, error in B, error in div B``
  with the numbers printed as ``%.15e``.
* ``write_time_series`` -- the per-run series of examples/energy_hall_outflow.m:150-161 and
  paper/divcleaning_confined_domain.m:176-190: datasets ``/time``, ``/energy``, ``/L2error_B``,
  ``/L2error_divB`` (1-D, FP64).  Written as HDF5 when h5py is importable; otherwise as an
  ``.npz`` archive with the same dataset names (without the leading slash).
"""
from __future__ import annotations

import os
from typing import Dict, Iterable, Sequence

import numpy as np

TABLE_HEADER = "# N, order, form_uiBj, form_source, form_ujBi, runtime, energy, error in B, error in div B \n"
# paper/divcleaning_confined_domain.m:158 (and divcleaning_hall_outflow.m:134)
DC_TABLE_HEADER = ("# N, order, form_uiBj, form_source, form_ujBi, runtime, energy, error in B, error in div B, "
                   "divcleaning, form_divcleaning\n")


def table_row(n: int, order: int, forms: Sequence[str], res: Dict[str, object]) -> str:
    """One line of a sweep table; forms are the USE_* define names, res a compute_numerical_solution result."""
    return "%d, %d, %s, %s, %s, %.15e, %.15e, %.15e, %.15e \n" % (
        n, order, forms[0], forms[1], forms[2], res["runtime"], res["energy"], res["abs_err"], res["divergence_norm"])


def dc_table_row(n: int, order: int, forms: Sequence[str], res: Dict[str, object], dc_form: str = "") -> str:
    """One line of a divergence-cleaning sweep table (divcleaning_confined_domain.m:192-194): the table_row fields,
    then 1 and the USE_LAPLACE_* define when the run cleaned every step, 0 and a blank otherwise."""
    return "%d, %d, %s, %s, %s, %.15e, %.15e, %.15e, %.15e, %d, %s \n" % (
        n, order, forms[0], forms[1], forms[2], res["runtime"], res["energy"], res.get("abs_err", 0.0),
        res["divergence_norm"], 1 if dc_form else 0, dc_form if dc_form else " ")


def dc_series_name(prefix: str, n: int, order: int, forms: Sequence[str], dc_label: str = "") -> str:
    """File name of a run's time series (divcleaning_confined_domain.m:161-173): forms lowercased without the
    USE_UIBJ_ / USE_SOURCE_ / USE_UJBI_ prefix, __DC_1_<WS_LN|WS_D0|NS_D0> or __DC_0."""
    base = "%s_N_%d_order_%d_%s_%s_%s" % (prefix, n, order, forms[0][9:].lower(), forms[1][11:].lower(), forms[2][9:].lower())
    return base + ("__DC_1_%s.h5" % dc_label if dc_label else "__DC_0.h5")


def write_convergence_table(path: str, rows: Iterable[str], device: str, header: str = TABLE_HEADER) -> None:
    """Header, device line and the rows produced by table_row (or dc_table_row with DC_TABLE_HEADER)."""
    with open(path, "w") as f:
        f.write(header)
        f.write("#  %s \n" % device)
        for r in rows:
            f.write(r)


def read_convergence_table(path: str):
    """Rows of a sweep table as tuples (N, order, form_uiBj, form_source, form_ujBi, runtime, energy, err_B, err_divB)."""
    out = []
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        p = [v.strip() for v in line.split(",")]
        out.append((int(p[0]), int(p[1]), p[2], p[3], p[4]) + tuple(float(v) for v in p[5:9]))
    return out


def write_time_series(path: str, res: Dict[str, object]) -> str:
    """The /time, /energy, /L2error_B, /L2error_divB series of a run with save_integrals_over_time.
    Returns the path written (``.npz`` replaces ``.h5`` when h5py is not importable)."""
    data = {"time": res["time_series"], "energy": res["energy_over_time"], "L2error_divB": res["L2error_divB_over_time"]}
    if "L2error_B_over_time" in res:
        data["L2error_B"] = res["L2error_B_over_time"]
    data = {k: np.asarray(v, dtype=np.float64) for k, v in data.items()}
    try:
        import h5py  # noqa: F401
    except ImportError:
        h5py = None
    if path.endswith(".h5") and h5py is not None:
        if os.path.exists(path):
            os.remove(path)
        with h5py.File(path, "w") as f:
            for k, v in data.items():
                f.create_dataset("/" + k, data=v)
        return path
    out = path[:-3] + ".npz" if path.endswith(".h5") else path
    np.savez(out, **data)
    return out if out.endswith(".npz") else out + ".npz"
