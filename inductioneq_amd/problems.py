"""Synthetic inputs of the benchmark workloads and the time-step rules.

Fields are returned as numpy float64 arrays of shape ``(num_nodes, 3)`` in node order
ix + NX*(iy + NY*iz).

* :func:`random_modes_field` -- periodic ``[0, 2pi)^3``: a constant background plus
  divergence-free random Fourier modes ``0 < |k| <= kmax`` (half space, so each pair
  ``+-k`` appears once), amplitudes ``N(0, amp^2)`` projected orthogonal to ``k`` and
  scaled by ``1/|k|^2``; seeded with ``numpy.random.default_rng(seed)``.
* :func:`gradient_potential` -- the gradient of a random scalar potential
  ``psi = sum_{|k|<=kmax} c_k cos(pi k.(x+1)/2)`` on ``[-1, 1]^3`` (injects divergence).
* :func:`beltrami_hall_field` -- an exact periodic solution of the implemented equation
  with uniform velocity and density: B = B0 + eps*b(x - v t), b the ABC field of
  wavenumber m (curl b = m b), v = u + (m/rho) B0.
* :func:`hall_dt` -- the reference's CFL rule ``dt = cfl*min(dx)/max|u|``
  (``matlab/+induction_eq/initialize.m:93-96``) is undefined for ``u = 0``; for the
  Hall term the step is set from its spectral radius instead:
  ``dt = cfl / (c_H * max|B| * sum_d (s_max/dx_d)^2)`` with ``s_max`` the largest
  modulus of the central first-derivative symbol.
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np

# max_theta |sum_k c_k e^{i k theta}| of the central first-derivative rows (order 2, 4, 6)
_SYMBOL_MAX = {2: 1.0, 4: 1.3722, 6: 1.5860}


def half_space_modes(kmax: int):
    ks = []
    r = int(kmax)
    for kx in range(-r, r + 1):
        for ky in range(-r, r + 1):
            for kz in range(-r, r + 1):
                k = (kx, ky, kz)
                if k == (0, 0, 0) or kx * kx + ky * ky + kz * kz > kmax * kmax:
                    continue
                if k > (0, 0, 0):  # lexicographic half space
                    ks.append(k)
    return ks


def random_modes_field(n: Sequence[int], seed: int, kmax: int = 4, amp: float = 0.1,
                       background: Tuple[float, float, float] = (0.0, 0.0, 1.0), length: float = 2 * math.pi):
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * (length / nx)
    y = np.arange(ny) * (length / ny)
    z = np.arange(nz) * (length / nz)
    B = np.zeros((nz, ny, nx, 3))
    B[...] = background
    for k in half_space_modes(kmax):
        kv = np.array(k, float)
        k2 = kv @ kv
        a = rng.normal(0.0, amp, 3)
        b = rng.normal(0.0, amp, 3)
        a -= (a @ kv) / k2 * kv
        b -= (b @ kv) / k2 * kv
        a /= k2
        b /= k2
        ex = np.exp(1j * kv[0] * x)
        ey = np.exp(1j * kv[1] * y)
        ez = np.exp(1j * kv[2] * z)
        ph = ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
        c, s = ph.real, ph.imag
        B += c[..., None] * a + s[..., None] * b
    return B.reshape(-1, 3)


def beltrami_hall_field(n: Sequence[int], t: float = 0.0, m: int = 1, abc: Tuple[float, float, float] = (0.3, 0.2, 0.25),
                        eps: float = 1.0, background: Tuple[float, float, float] = (0.3, -0.2, 1.0),
                        u: Tuple[float, float, float] = (0.0, 0.0, 0.0), rho: float = 1.0,
                        length: float = 2 * math.pi):
    """Exact solution of dB/dt = curl(u x B) - curl((curl B)/rho x B) on the periodic cube for uniform u and rho.

    b = (A sin(m z) + C cos(m y), B sin(m x) + A cos(m z), C sin(m y) + B cos(m x)) satisfies curl b = m b
    and div b = 0, so with B = B0 + eps*b: J x B = (m/rho) eps b x B0, and both terms reduce to transport,
    dB/dt = -((u + (m/rho) B0) . grad) B, whose solution is B(x, t) = B0 + eps*b(x - v t), v = u + (m/rho) B0."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    v = np.asarray(u, float) + (m / rho) * np.asarray(background, float)
    x = np.arange(nx) * (length / nx) - v[0] * t
    y = np.arange(ny) * (length / ny) - v[1] * t
    z = np.arange(nz) * (length / nz) - v[2] * t
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    A, Bc, C = abc
    b = np.stack([A * np.sin(m * Z) + C * np.cos(m * Y), Bc * np.sin(m * X) + A * np.cos(m * Z),
                  C * np.sin(m * Y) + Bc * np.cos(m * X)], axis=-1)
    return (np.asarray(background, float) + eps * b).reshape(-1, 3)


def gradient_potential(n: Sequence[int], seed: int, kmax: int = 6, lo: float = -1.0, hi: float = 1.0):
    """grad psi, psi = sum_{0<|k|<=kmax} c_k cos(pi k.(x-lo)/L), c_k ~ N(0,1)/|k|^2 (non-periodic grid)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    L = hi - lo
    rng = np.random.default_rng(seed)
    xs = [lo + np.arange(m) * (L / (m - 1)) for m in (nx, ny, nz)]
    g = np.zeros((nz, ny, nx, 3))
    r = int(kmax)
    for kx in range(0, r + 1):
        for ky in range(0, r + 1):
            for kz in range(0, r + 1):
                k2 = kx * kx + ky * ky + kz * kz
                if k2 == 0 or k2 > kmax * kmax:
                    continue
                c = rng.normal() / k2
                w = math.pi / L
                ax = [w * kx * (xs[0] - lo), w * ky * (xs[1] - lo), w * kz * (xs[2] - lo)]
                # psi = c cos(ax + ay + az); grad = -c sin(.) * (w k)
                ph = ax[2][:, None, None] + ax[1][None, :, None] + ax[0][None, None, :]
                sn = -c * np.sin(ph)
                g[..., 0] += sn * w * kx
                g[..., 1] += sn * w * ky
                g[..., 2] += sn * w * kz
    return g.reshape(-1, 3)


def gaussian_rotation_field(n: Sequence[int], lo: float = -1.0, hi: float = 1.0):
    """curl (0, 0, exp(-20((x-1/2)^2 + y^2 + z^2))) on a bounded grid (divergence free)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    L = hi - lo
    x = lo + np.arange(nx) * (L / (nx - 1))
    y = lo + np.arange(ny) * (L / (ny - 1))
    z = lo + np.arange(nz) * (L / (nz - 1))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    fac = np.exp(-20 * ((X - 0.5) ** 2 + Y ** 2 + Z ** 2))
    B = np.stack([-40 * Y * fac, (40 * X - 20) * fac, np.zeros_like(X)], axis=-1)
    return B.reshape(-1, 3)


def hall_dt(B: np.ndarray, dx: Sequence[float], order: int, c_hall: float = 1.0, cfl: float = 0.5) -> float:
    bmax = float(np.sqrt((np.asarray(B, float) ** 2).sum(axis=-1)).max())
    s = _SYMBOL_MAX[order]
    rad = c_hall * bmax * sum((s / h) ** 2 for h in dx)
    return cfl / rad


def to_aos4(B3: np.ndarray, num_nodes_pad: int, dtype=np.float64) -> np.ndarray:
    """(num_nodes, 3) -> reference host layout (num_nodes_pad, 4) with zero w lane and padding."""
    out = np.zeros((num_nodes_pad, 4), dtype)
    out[: B3.shape[0], :3] = B3
    return out
