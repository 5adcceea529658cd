"""Closed forms of the distance outputs (render.py:248-276) on a segment of
constant density, used to pin the oracle and to check the GPU directly.

A density factor whose planes are all 1 and whose lines are all v gives every
sample of a ray the same feature sum 3 C v, so sigma = softplus(3 C v + shift)
along the whole box segment [t0, t0 + N dt] (map_coordinates of a constant
grid is that constant). With q = exp(-sigma dt):
  E_i = q^(i+1),  w_i = (1 - q) q^i,  t_i = t0 + (i + u_i) dt,
  median index i* = the first i with 1 - q^(i+1) > 1/2 = floor(ln 2 / (sigma dt)),
  mean = sum_i w_i t_i + q^N t_(N-1)
       = t0 (1 - q^N) + dt (1 - q) [S1 + sum_i q^i u_i] + q^N t_(N-1),
  S1 = sum_(i<N) i q^i = q (1 - N q^(N-1) + (N-1) q^N) / (1 - q)^2.
"""
import math

import numpy as np


def constant_density_factors(cd, grid_dim, sigma, shift):
    """(vector (3, cd, G), matrix (3, cd, G, G)) whose density is sigma everywhere."""
    x = math.log(math.expm1(sigma))  # softplus^-1
    v = (x - shift) / (3 * cd)
    return np.full((3, cd, grid_dim), v), np.ones((3, cd, grid_dim, grid_dim))


def box_segments(origins, directions, aabb):
    """Slab test in float64: (t0, t1) per ray (rays assumed to hit the box),
    with the reference's +1e-8 on the direction (render.py:412-415)."""
    lo = (aabb[0][None] - origins) / (directions + 1e-8)
    hi = (aabb[1][None] - origins) / (directions + 1e-8)
    t0 = np.maximum(np.minimum(lo, hi).max(axis=1), 0.0)
    t1 = np.maximum(lo, hi).min(axis=1)
    assert np.all(t1 > t0 + 1e-3)
    return t0, t1


def distance_closed_forms(t0, dt, sigma, n, u):
    """Per ray: (median select, median as written, mean, ln2/(sigma dt))."""
    u = np.zeros(n) if u is None else np.asarray(u, np.float64)
    a = sigma * dt
    q = np.exp(-a)
    x = math.log(2.0) / a
    istar = np.floor(x).astype(np.int64)
    t_at = t0 + (np.minimum(istar, n - 1) + u[np.minimum(istar, n - 1)]) * dt
    median = np.where(istar >= n, np.inf, np.where(istar == 0, 0.0, t_at))
    reached = 1.0 - q ** n > 0.5
    as_written = np.where(reached, np.nan, np.inf)
    s1 = q * (1.0 - n * q ** (n - 1) + (n - 1) * q ** n) / (1.0 - q) ** 2
    su = (q[:, None] ** np.arange(n)[None] * u[None]).sum(axis=1)
    t_last = t0 + (n - 1 + u[-1]) * dt
    mean = t0 * (1.0 - q ** n) + dt * (1.0 - q) * (s1 + su) + q ** n * t_last
    return median, as_written, mean, x
