"""Dense 1-D operator matrices assembled from the generated coefficient tables.

Assembly rule (the reference's per-node evaluation, include/utils.h:125-201): node i
of n uses table row ``row(i)`` (left rows 0..nb-1 for i < nb, right rows for
i >= n-nb, the central row otherwise; always central when periodic); column j of the
row weighs node ``i + j - (w-1)/2``, which wraps when periodic and otherwise falls
back to node i itself when it leaves the grid.
"""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLES_H = os.path.join(ROOT, "inductioneq_amd", "csrc", "ieq_tables.h")
OPSETS = ["2nd", "4th", "6th", "2nd_ext", "4th_ext", "6th_ext"]
ORDER = {"2nd": 2, "4th": 4, "6th": 6, "2nd_ext": 2, "4th_ext": 4, "6th_ext": 6}


def load_tables():
    text = open(TABLES_H).read()
    out = {}
    for m in re.finditer(r"double (ieq_(\w+?)_(diff_adj|diff|minv|dho|lap_ws_d0|lap_ns_d0|lap_ws_ln))\[(\d+)\] = \{(.*?)\};",
                         text, flags=re.S):
        name, ops, field, _n, body = m.groups()
        vals = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        comment = re.search(r"/\* %s: .*?, (\d+) rows x (\d+) \*/" % name, text)
        rows, w = int(comment.group(1)), int(comment.group(2))
        out.setdefault(ops, {})[field] = np.array(vals).reshape(rows, w)
    return out


def row_of(i, n, nb, periodic):
    if periodic:
        return nb
    b = 0
    for k in range(nb):
        b += (nb - k) * (int(i == n - (k + 1)) - int(i == k))
    return nb + b


def nbr(i, off, n, periodic):
    if periodic:
        return (i + off) % n
    j = i + off
    return j if 0 <= j < n else i


def assemble(table, n, periodic=False):
    """Matrix A with (A f)_i = sum_j table[row(i), j] f(nbr(i, j - c))."""
    rows, w = table.shape
    nb = (rows - 1) // 2
    c = (w - 1) // 2
    A = np.zeros((n, n))
    for i in range(n):
        r = row_of(i, n, nb, periodic)
        for j in range(w):
            A[i, nbr(i, j - c, n, periodic)] += table[r, j]
    return A


def norm_matrix(minv, n, periodic=False):
    nb = (minv.shape[0] - 1) // 2
    return np.diag([1.0 / minv[row_of(i, n, nb, periodic), 0] for i in range(n)])


def hod_matrix(t, n, order, periodic=False):
    """High-order dissipation of artificial_dissipation.h:58-101 (one direction, factor 1)."""
    dho, minv = t["dho"], t["minv"]
    rows, wh = dho.shape
    nb = (rows - 1) // 2
    sgn = (-1.0) ** (order // 2 + 1)
    A = np.zeros((n, n))
    for x in range(n):
        r = row_of(x, n, nb, periodic)
        for s in range(-wh + 1, 1):
            for i in range(wh):
                A[x, nbr(x, i + s, n, periodic)] += minv[r, 0] * dho[r, s + wh - 1] * dho[nb, i] * sgn
    return A
