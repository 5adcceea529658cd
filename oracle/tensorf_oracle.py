"""CPU oracle for the TensoRF render / train hot path (NumPy, FP64 or FP32).

TEST INFRASTRUCTURE. This module is the checker, never the product: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it. The product path (``tensorf_amd`` + the HIP library)
never calls into it and raises when the HIP library is missing.

It restates the reference algorithm op for op so that results can be compared
with the GPU kernels on the same inputs:

* every function cites the reference file:line it follows (paths relative to
  the reference repository root);
* all arithmetic runs in ``dtype`` (``np.float64`` by default; ``np.float32``
  reproduces the reference's FP32 rounding order up to transcendental ulps and
  XLA's reduction order);
* randomness is explicit: stratified jitter and Gumbel noise are inputs, so
  the JAX threefry PRNG (``jax.random.split/uniform/choice``) is not needed.

Pinning (what checks each function independently of this file's formulas) is
done in ``tests/test_oracle_*.py``: scipy ``map_coordinates`` and torch
``grid_sample`` for the VM lookup, closed forms and brute force for the
sampler and compositing, torch float64 ``Linear`` layers for the MLP, central
finite differences and torch float64 autograd for every backward, and
``torch.optim.Adam`` for the optimizer. Functions without such a check are
marked "parity unpinned" in their docstring.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, Optional, Tuple

import numpy as np
import scipy.sparse

# Mode permutations of TensorVM.interpolate: ijk, kij, jki (tensor_vm.py:50-52).
MODE_PERMS = ((0, 1, 2), (2, 0, 1), (1, 2, 0))

RGB, DIST_MEDIAN, DIST_MEAN = 0, 1, 2  # render.py:18-23


def _c(x, dtype):
    return np.asarray(x, dtype=dtype)


# --------------------------------------------------------------------------
# Ray sampling
# --------------------------------------------------------------------------
def ray_segment_from_bounding_box(origins, directions, aabb, min_segment_length=1e-3,
                                  dtype=np.float64):
    """Per-ray [t_min, t_max] inside the box. render.py:399-434 (vmapped).

    origins, directions: (R, 3); aabb: (2, 3). Returns (t_min (R,), t_max (R,)).
    """
    o = _c(origins, dtype)
    d = _c(directions, dtype)
    box = _c(aabb, dtype)
    offsets = box[None, :, :] - o[:, None, :]                        # render.py:412
    t_int = offsets / (d[:, None, :] + dtype(1e-8))                  # render.py:413-415
    t_min_axis = t_int.min(axis=1)                                   # render.py:418
    t_max_axis = t_int.max(axis=1)                                   # render.py:419
    t_min = np.maximum(dtype(0.0), t_min_axis.max(axis=1))           # render.py:423
    t_max = t_max_axis.min(axis=1)                                   # render.py:424
    t_max_clipped = np.maximum(t_max, t_min + dtype(min_segment_length))  # render.py:425
    valid = t_min < t_max                                            # render.py:429
    return (np.where(valid, t_min, dtype(0.0)).astype(dtype),        # render.py:432
            np.where(valid, t_max_clipped, dtype(min_segment_length)).astype(dtype))


def sample_points_along_ray_within_bbox(origins, directions, aabb, samples_per_ray,
                                        jitter=None, dtype=np.float64):
    """Stratified samples inside the box, render.py:350-390, vmapped as render.py:177-186.

    ``jitter`` (N,) is shared by every ray: the reference closes one PRNG key over
    the vmap (render.py:182), so all rays see the same uniform draws.
    Returns points (3, R, N), ts (R, N), step_sizes (R, N).
    """
    n = int(samples_per_ray)
    o = _c(origins, dtype)
    d = _c(directions, dtype)
    t_min, t_max = ray_segment_from_bounding_box(o, d, aabb, 1e-3, dtype)  # render.py:366
    step = (t_max - t_min) / dtype(n)                                # render.py:369
    ts = np.arange(n, dtype=dtype)                                   # render.py:372
    if jitter is not None:
        ts = ts + _c(jitter, dtype)                                  # render.py:375-379
    ts = ts[None, :] * step[:, None]                                 # render.py:380
    ts = t_min[:, None] + ts                                         # render.py:381
    points = o.T[:, :, None] + d.T[:, :, None] * ts[None, :, :]      # render.py:384-386
    steps = np.broadcast_to(step[:, None], ts.shape).astype(dtype)   # render.py:186
    return points.astype(dtype), ts.astype(dtype), steps


def _linspace32(start, stop, num):
    """FP32 linspace evaluated as start*(1-s)+stop*s, s=i/(num-1). Stands in for
    jnp.linspace(float32) at render.py:130; the exact FP32 rounding of XLA's
    linspace is parity unpinned (ulp level)."""
    if num == 1:
        return np.array([start], np.float32)
    s = (np.arange(num, dtype=np.float32) / np.float32(num - 1)).astype(np.float32)
    return (np.float32(start) * (np.float32(1.0) - s) + np.float32(stop) * s).astype(np.float32)


def contraction_schedule(samples_per_ray, near, far):
    """Un-jittered sample distances and step sizes of the contracted sampler,
    render.py:127-155. The far half is evaluated in float64 NumPy and cast to
    float32 exactly like the reference (render.py:141 "onp here is important").
    Returns (ts (N,) float32, step_sizes (N,) float32)."""
    n = int(samples_per_ray)
    nc = n // 2                                                      # render.py:127
    nf = n - nc                                                      # render.py:128
    close = _linspace32(near, near + 1.0, nc)                        # render.py:130
    far_start = near + 1.0 + 1.0 / nc                                # render.py:135
    k = 10.0                                                         # render.py:136
    far_deltas = (1.0 / (1.0 - np.linspace(0.0, 1.0 - 1 / ((far - far_start) / k + 1), nf))
                  - 1.0) * np.linspace(1.0, k, nf)                   # render.py:137-148
    far_ts = far_start + far_deltas                                  # render.py:149
    ts = np.concatenate([close, far_ts.astype(np.float32)]).astype(np.float32)  # render.py:151
    steps = (np.roll(ts, -1) - ts).astype(np.float32)                # render.py:154
    steps[-1] = steps[-2]                                            # render.py:155
    return ts, steps


def uniform_contracted_schedule(samples_per_ray, near, far):
    """Sample distances spaced uniformly in contracted distance
    g(t) = t (t <= 1), 2 - 1/t (t > 1) -- the contraction's radial map for a
    ray leaving the scene centre -- between near and far: t_i = g^-1(s_i),
    s_i evenly spaced from g(near) to g(far) in float64, cast to float32; step
    sizes as in the reference schedule (differences, the last one repeated,
    render.py:154-155). No counterpart in the reference (its schedule is
    contraction_schedule); the alternative spacing of the unbounded
    configuration. Returns (ts (N,) float32, step_sizes (N,) float32)."""
    n = int(samples_per_ray)
    g = lambda t: t if t <= 1.0 else 2.0 - 1.0 / t
    s = np.linspace(g(near), g(far), n)
    ts = np.where(s <= 1.0, s, 1.0 / (2.0 - np.minimum(s, 2.0 - 1e-300))).astype(np.float32)
    steps = (np.roll(ts, -1) - ts).astype(np.float32)
    steps[-1] = steps[-2] if n > 1 else np.float32(far - near)
    return ts, steps


SCHEDULES = {"reference": contraction_schedule, "uniform_contracted": uniform_contracted_schedule}


def sample_contracted(origins, directions, samples_per_ray, near, far, jitter,
                      dtype=np.float64, spacing="reference"):
    """Contracted-scene sampler + L-infinity contraction, render.py:122-173.

    ``jitter`` (R, N) is per ray and per sample (render.py:158-160).
    ``spacing`` picks the un-jittered schedule: the reference's
    (contraction_schedule) or uniform_contracted_schedule.
    Returns points (3, R, N) in contracted space, ts (R, N), step_sizes (R, N).
    """
    base, steps1 = SCHEDULES[spacing](samples_per_ray, near, far)
    o = _c(origins, dtype)
    d = _c(directions, dtype)
    r = o.shape[0]
    ts = np.tile(_c(base, dtype)[None, :], (r, 1))                   # render.py:151
    steps = np.tile(_c(steps1, dtype)[None, :], (r, 1))              # render.py:154-155
    ts = ts + steps * _c(jitter, dtype)                              # render.py:161
    pts = o[:, None, :] + ts[:, :, None] * d[:, None, :]             # render.py:164-167
    norm = np.abs(pts).max(axis=-1, keepdims=True)                   # render.py:170
    with np.errstate(divide="ignore", invalid="ignore"):
        contracted = (dtype(2.0) - dtype(1.0) / norm) * pts / norm
    pts = np.where(norm <= 1.0, pts, contracted)                     # render.py:171
    return np.moveaxis(pts, -1, 0).astype(dtype), ts.astype(dtype), steps.astype(dtype)


def contract_points(pts, dtype=np.float64):
    """L-infinity contraction p -> (2 - 1/|p|_inf) p/|p|_inf outside the unit cube,
    render.py:168-173. pts (..., 3)."""
    pts = _c(pts, dtype)
    norm = np.abs(pts).max(axis=-1, keepdims=True)                   # render.py:170
    with np.errstate(divide="ignore", invalid="ignore"):
        contracted = (dtype(2.0) - dtype(1.0) / norm) * pts / norm
    return np.where(norm <= 1.0, pts, contracted).astype(dtype)      # render.py:171


def sample_from_edges(origins, directions, edges, jitter, contraction, dtype=np.float64):
    """Samples of per-ray intervals (no reference counterpart; the interface a
    proposal sampler hands to the field): t_i = e_i + u_i (e_{i+1} - e_i), step
    size e_{i+1} - e_i, u = jitter (R, N) or 0.5. The point is contracted like
    render.py:168-173 when ``contraction``.
    Returns points (3, R, N), ts (R, N), step_sizes (R, N)."""
    o = _c(origins, dtype)
    d = _c(directions, dtype)
    e = _c(edges, dtype)
    steps = e[:, 1:] - e[:, :-1]
    u = dtype(0.5) if jitter is None else _c(jitter, dtype)
    ts = e[:, :-1] + u * steps
    pts = o[:, None, :] + ts[:, :, None] * d[:, None, :]
    if contraction:
        pts = contract_points(pts, dtype)
    return np.moveaxis(pts, -1, 0).astype(dtype), ts.astype(dtype), steps.astype(dtype)


def normalize_points(points, aabb, dtype=np.float64):
    """Map world (or contracted) points to [-1, 1] with the aabb, render.py:192-196."""
    p = _c(points, dtype)
    box = _c(aabb, dtype)
    shape = (3,) + (1,) * (p.ndim - 1)
    return (((p - box[0].reshape(shape)) / (box[1] - box[0]).reshape(shape)
             - dtype(0.5)) * dtype(2.0)).astype(dtype)


# --------------------------------------------------------------------------
# VM decomposition lookup
# --------------------------------------------------------------------------
def linear_taps(coord, size):
    """Order-1 'nearest' interpolation taps of map_coordinates as called at
    tensor_vm.py:241-248: i0 = floor(u), w1 = u - i0, w0 = 1 - w1, and both
    indices clamped to [0, size-1]."""
    lower = np.floor(coord)
    w1 = coord - lower
    w0 = np.asarray(1.0, coord.dtype) - w1
    i = np.clip(lower, -2, size + 1).astype(np.int64)
    return np.clip(i, 0, size - 1), np.clip(i + 1, 0, size - 1), w0, w1


def _mode_taps(ijk, grid_dim, m, dtype):
    coords = ijk[list(MODE_PERMS[m])]                                # tensor_vm.py:50-52
    u = (coords + dtype(1.0)) / dtype(2.0)                           # tensor_vm.py:150
    u = u * dtype(grid_dim - 1.0)                                    # tensor_vm.py:153
    return (linear_taps(u[0], grid_dim), linear_taps(u[1], grid_dim),
            linear_taps(u[2], grid_dim))


def vm_interpolate(vector, matrix, ijk, dtype=np.float64):
    """TensorVM.interpolate, tensor_vm.py:42-89 with TensorVMSingle.interpolate
    (tensor_vm.py:140-167) and map_coordinates(order=1, mode='nearest')
    (tensor_vm.py:226-250).

    vector (3, C, G), matrix (3, C, G, G) in the reference layout; ijk (3, M) in
    [-1, 1]. Returns features (3C, M) with channel index m*C + c."""
    vec = _c(vector, dtype)
    mat = _c(matrix, dtype)
    p = _c(ijk, dtype)
    g = mat.shape[-1]
    outs = []
    for m in range(3):
        line, plane = _line_plane(vec[m], mat[m], _mode_taps(p, g, m, dtype))
        outs.append(line * plane)                                    # tensor_vm.py:167
    return np.concatenate(outs, axis=0).astype(dtype)               # tensor_vm.py:86-88


def _line_plane(vec, mat, taps):
    """Line and plane values (C, M) of one mode at its taps. JAX's
    map_coordinates sums the 2-D taps in product order (a0,b0),(a0,b1),
    (a1,b0),(a1,b1), each as (wa*wb)*value. The texels are read as rows of
    the transposed tables (the same values, gathered row-wise)."""
    (v0, v1, wv0, wv1), (a0, a1, wa0, wa1), (b0, b1, wb0, wb1) = taps
    c, g = vec.shape
    vt = np.ascontiguousarray(vec.T)                                 # (G, C)
    mt = np.ascontiguousarray(mat.reshape(c, g * g).T)               # (G*G, C)
    line = (wv0[:, None] * vt[v0] + wv1[:, None] * vt[v1]).T
    plane = ((((wa0 * wb0)[:, None] * mt[a0 * g + b0] + (wa0 * wb1)[:, None] * mt[a0 * g + b1])
              + (wa1 * wb0)[:, None] * mt[a1 * g + b0]) + (wa1 * wb1)[:, None] * mt[a1 * g + b1]).T
    return line, plane


def vm_interpolate_vjp(vector, matrix, ijk, grad_features, dtype=np.float64):
    """Vector-Jacobian product of vm_interpolate w.r.t. (vector, matrix): the
    transposed gather (scatter-add). Returns (grad_vector, grad_matrix)."""
    vec = _c(vector, dtype)
    mat = _c(matrix, dtype)
    p = _c(ijk, dtype)
    gf = _c(grad_features, dtype)
    c = vec.shape[1]
    g = mat.shape[-1]
    gvec = np.zeros_like(vec)
    gmat = np.zeros_like(mat)
    for m in range(3):
        taps = _mode_taps(p, g, m, dtype)
        (v0, v1, wv0, wv1), (a0, a1, wa0, wa1), (b0, b1, wb0, wb1) = taps
        line, plane = _line_plane(vec[m], mat[m], taps)
        gm = gf[m * c:(m + 1) * c]
        gline = gm * plane
        gplane = gm * line
        # The scatter-add is the transposed tap matrix times the per-sample
        # gradients: T[texel, sample] = tap weight, duplicate taps (clamped at
        # the grid edge) summed.
        n = p.shape[1]
        cols = np.arange(n)
        tv = scipy.sparse.csr_matrix((np.concatenate([wv0, wv1]), (
            np.concatenate([v0, v1]), np.concatenate([cols, cols]))), shape=(g, n))
        gvec[m] += (tv @ gline.T).T
        rows_p, w_p = [], []
        for ia, wa in ((a0, wa0), (a1, wa1)):
            for ib, wb in ((b0, wb0), (b1, wb1)):
                rows_p.append(ia * g + ib)
                w_p.append(wa * wb)
        tp = scipy.sparse.csr_matrix((np.concatenate(w_p), (
            np.concatenate(rows_p), np.tile(cols, 4))), shape=(g * g, n))
        gmat[m] += (tp @ gplane.T).T.reshape(c, g, g)
    return gvec, gmat


def vm_resize(vector, matrix, grid_dim, dtype=np.float64):
    """TensorVM.resize with align-corners linear resampling, tensor_vm.py:91-100,
    183-223. For upsampling, jax.image.scale_and_translate(linear) with
    scale=(G'-1)/(G-1), translation=0.5-scale/2 samples input position
    o*(G-1)/(G'-1) with the triangle kernel, i.e. plain separable linear
    interpolation with align_corners=True. Downsampling (antialiased triangle
    kernel) is not implemented. Returns (vector', matrix')."""
    vec = _c(vector, dtype)
    mat = _c(matrix, dtype)
    g_in = mat.shape[-1]
    g_out = int(grid_dim)
    if g_out < g_in:
        raise NotImplementedError("vm_resize: only upsampling is restated")

    def interp_matrix(n_in, n_out):
        w = np.zeros((n_out, n_in), dtype)
        if n_out == 1 or n_in == 1:
            w[:, 0] = 1.0
            return w
        x = np.arange(n_out, dtype=np.float64) * (n_in - 1) / (n_out - 1)
        for o_idx, xv in enumerate(x):
            i0 = min(int(np.floor(xv)), n_in - 1)
            f = xv - i0
            w[o_idx, i0] += 1.0 - f
            if f > 0:
                w[o_idx, i0 + 1] += f
        return w

    w = interp_matrix(g_in, g_out)
    new_vec = np.einsum("oi,mci->mco", w, vec)
    new_mat = np.einsum("oi,mcij->mcoj", w, mat)
    new_mat = np.einsum("pj,mcoj->mcop", w, new_mat)
    return new_vec.astype(dtype), new_mat.astype(dtype)


# --------------------------------------------------------------------------
# Density and compositing
# --------------------------------------------------------------------------
def softplus(x):
    """jax.nn.softplus = logaddexp(x, 0) (render.py:207)."""
    return np.maximum(x, 0) + np.log1p(np.exp(-np.abs(x)))


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def density_from_features(features, dtype=np.float64):
    """sigma = softplus(sum over the 3C channels + 10), render.py:207.
    features (3C, ...) -> (...)."""
    f = _c(features, dtype)
    return softplus(f.sum(axis=0) + dtype(10.0)).astype(dtype)


def compute_segment_probabilities(sigmas, step_sizes, dtype=np.float64):
    """render.py:300-347: E = exp(cumsum(-sigma*delta)) (log-space transmittance),
    w = (1 - exp(-sigma*delta)) * [1, E_0 .. E_{N-2}]. Returns (p_exits, p_terminates)."""
    s = _c(sigmas, dtype)
    st = _c(step_sizes, dtype)
    neg = -s * st                                                    # render.py:322
    p_exits = np.exp(np.cumsum(neg, axis=-1))                        # render.py:323
    given = dtype(1.0) - np.exp(neg)                                 # render.py:327
    prev = np.concatenate([np.ones(s.shape[:-1] + (1,), dtype), p_exits[..., :-1]], axis=-1)
    return p_exits.astype(dtype), (given * prev).astype(dtype)       # render.py:331-341


def topk_indices(p_terminates, k, gumbel=None):
    """Per-ray appearance-sample selection, render.py:459-470.

    The reference draws k indices without replacement with probability p_terminates
    (jax.random.choice, one key shared by all rays). JAX implements that draw as a
    Gumbel top-k of log(p) + g with g shared by every ray; with g = 0 it is the
    deterministic top-k used here: order by weight descending, ties broken by the
    lower sample index first (lax.top_k's tie rule). Returned indices are sorted
    ascending (the compositing sum does not depend on their order).
    Parity unpinned for the JAX threefry noise itself (it is an explicit input)."""
    w = np.asarray(p_terminates)
    if gumbel is None:
        key = w
    else:
        with np.errstate(divide="ignore"):
            key = np.log(w.astype(np.float32)) + np.asarray(gumbel, np.float32)[None, :]
    order = np.argsort(-key, axis=-1, kind="stable")[..., :k]
    return np.sort(order, axis=-1).astype(np.int32)


# --------------------------------------------------------------------------
# Feature MLP
# --------------------------------------------------------------------------
def fourier_encode(coords, n_freqs, dtype=np.float64):
    """networks.py:13-35: sin([2^q x, 2^q x + pi/2]) per input dim -> (M, D*2n)."""
    x = _c(coords, dtype)
    m, dd = x.shape
    coeffs = (2.0 ** np.arange(n_freqs)).astype(dtype)               # networks.py:24
    inputs = x[..., None] * coeffs                                   # networks.py:25
    out = np.sin(np.concatenate([inputs, inputs + dtype(0.5 * np.pi)], axis=-1))  # networks.py:28-33
    return out.reshape(m, dd * 2 * n_freqs).astype(dtype)           # networks.py:35


@dataclasses.dataclass
class MlpShape:
    """FeatureMlp hyper-parameters, networks.py:38-43."""
    feat_dim: int = 144
    squash_dim: int = 27
    units: int = 128
    feature_n_freqs: int = 2
    viewdir_n_freqs: int = 2
    num_cameras: Optional[int] = None

    @property
    def encoded_dim(self):
        # networks.py:77-82
        return (self.squash_dim + 3 + 2 * self.feature_n_freqs * self.squash_dim
                + 2 * self.viewdir_n_freqs * 3)


def init_mlp_params(shape: MlpShape, rng: np.random.Generator):
    """Random parameters with the reference's shapes (networks.py:57-117): kernels
    are (in, out). Distributions follow the Flax initialisers named at
    networks.py:9-10 (truncated-normal variance scaling); the draws themselves
    are not the reference's (parity unpinned: Flax PRNG)."""
    def trunc(fan_in, fan_out, scale):
        std = math.sqrt(scale / fan_in) / 0.87962566103423978
        x = rng.standard_normal((fan_in, fan_out))
        while True:
            bad = np.abs(x) > 2.0
            if not bad.any():
                break
            x[bad] = rng.standard_normal(bad.sum())
        return x * std

    e = shape.encoded_dim
    u = shape.units
    p = {
        "w0": trunc(shape.feat_dim, shape.squash_dim, 1.0),
        "w1": trunc(e, u, 2.0), "b1": np.zeros(u),
        "w2": trunc(u, u, 2.0), "b2": np.zeros(u),
        "w3": trunc(u, 3, 1.0), "b3": np.zeros(3),
    }
    if shape.num_cameras is not None:
        p["embed"] = rng.standard_normal((shape.num_cameras, u)) / math.sqrt(u)
    return p


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64.
    Emulates the BF16 MFMA operands of the GPU MLP (FP32 accumulation)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, rnd):
    return (rnd(a) @ rnd(b)) if rnd is not None else a @ b


def feature_mlp_forward(params, features, viewdirs, camera_indices, shape: MlpShape,
                        dtype=np.float64, operand_round=None):
    """FeatureMlp.__call__, networks.py:46-121. Returns (rgb (M,3), cache).
    operand_round (e.g. bf16_round) is applied to both operands of every
    matrix product, to model reduced-precision MFMA inputs."""
    rnd = operand_round
    w0 = _c(params["w0"], dtype)
    f = _c(features, dtype)
    d = _c(viewdirs, dtype)
    h0 = _mm(f, w0, rnd)                                             # networks.py:57-61
    x = np.concatenate([h0, d, fourier_encode(h0, shape.feature_n_freqs, dtype),
                        fourier_encode(d, shape.viewdir_n_freqs, dtype)], axis=-1)  # networks.py:68-76
    z1 = _mm(x, _c(params["w1"], dtype), rnd) + _c(params["b1"], dtype)   # networks.py:86-89
    h1 = np.maximum(z1, 0)                                           # networks.py:90
    z2 = _mm(h1, _c(params["w2"], dtype), rnd) + _c(params["b2"], dtype)  # networks.py:94-97
    h2 = np.maximum(z2, 0)                                           # networks.py:98
    h2f = h2
    emb = None
    if shape.num_cameras is not None:                                # networks.py:103-111
        emb = _c(params["embed"], dtype)[np.asarray(camera_indices, np.int64)]
        half = shape.units // 2
        h2f = h2.copy()
        h2f[:, half:] = emb[:, :half] * h2[:, half:] + emb[:, half:]
    z3 = _mm(h2f, _c(params["w3"], dtype), rnd) + _c(params["b3"], dtype)  # networks.py:114-117
    rgb = sigmoid(z3)                                                # networks.py:120
    cache = dict(f=f, d=d, h0=h0, x=x, z1=z1, h1=h1, z2=z2, h2=h2, h2f=h2f, emb=emb,
                 rgb=rgb, cams=np.asarray(camera_indices, np.int64), rnd=rnd)
    return rgb.astype(dtype), cache


def feature_mlp_backward(params, cache, grad_rgb, shape: MlpShape, dtype=np.float64):
    """Hand-derived VJP of FeatureMlp w.r.t. its parameters and input features.
    Returns (param_grads dict, grad_features (M, feat_dim))."""
    g = _c(grad_rgb, dtype)
    rnd = cache.get("rnd")
    rgb = cache["rgb"]
    dz3 = g * rgb * (1.0 - rgb)
    grads = {"w3": _mm(cache["h2f"].T, dz3, rnd), "b3": dz3.sum(0)}
    dh2f = _mm(dz3, _c(params["w3"], dtype).T, rnd)
    dh2 = dh2f
    if shape.num_cameras is not None:
        half = shape.units // 2
        emb = cache["emb"]
        h2 = cache["h2"]
        dh2 = dh2f.copy()
        dh2[:, half:] = dh2f[:, half:] * emb[:, :half]
        demb_rows = np.concatenate([dh2f[:, half:] * h2[:, half:], dh2f[:, half:]], axis=-1)
        gemb = np.zeros_like(_c(params["embed"], dtype))
        np.add.at(gemb, cache["cams"], demb_rows)
        grads["embed"] = gemb
    dz2 = dh2 * (cache["z2"] > 0)
    grads["w2"] = _mm(cache["h1"].T, dz2, rnd)
    grads["b2"] = dz2.sum(0)
    dh1 = _mm(dz2, _c(params["w2"], dtype).T, rnd)
    dz1 = dh1 * (cache["z1"] > 0)
    grads["w1"] = _mm(cache["x"].T, dz1, rnd)
    grads["b1"] = dz1.sum(0)
    dx = _mm(dz1, _c(params["w1"], dtype).T, rnd)
    sq = shape.squash_dim
    nf = shape.feature_n_freqs
    h0 = cache["h0"]
    dh0 = dx[:, :sq].copy()
    fe = dx[:, sq + 3: sq + 3 + 2 * nf * sq].reshape(-1, sq, 2 * nf)
    for q in range(nf):
        s = 2.0 ** q
        dh0 += fe[:, :, q] * s * np.cos(s * h0)
        dh0 += fe[:, :, nf + q] * s * np.cos(s * h0 + 0.5 * np.pi)
    grads["w0"] = _mm(cache["f"].T, dh0, rnd)
    gfeat = _mm(dh0, _c(params["w0"], dtype).T, rnd)
    return {k: v.astype(dtype) for k, v in grads.items()}, gfeat.astype(dtype)


# --------------------------------------------------------------------------
# Full render and its backward
# --------------------------------------------------------------------------
@dataclasses.dataclass
class RenderConfig:
    """render.RenderConfig (render.py:26-36) plus the scene flag of
    LearnableParams (render.py:46)."""
    near: float = 0.05
    far: float = 200.0
    mode: int = RGB
    density_samples_per_ray: int = 64
    appearance_samples_per_ray: int = 64
    scene_contraction: bool = False
    sampling: int = 0  # 0: the reference's samplers, 1: per-ray interval edges
    density_shift: float = 10.0  # render.py:207
    median_as_written: bool = False  # DIST_MEDIAN: evaluate render.py:250-266 literally
    contracted_spacing: str = "reference"  # or "uniform_contracted"


SAMPLING_SCHEDULE, SAMPLING_EDGES = 0, 1


def render_rays(params, aabb, origins, directions, camera_indices, cfg: RenderConfig,
                shape: MlpShape, jitter, gumbel=None, dtype=np.float64, edges=None,
                mlp_operand_round=None):
    """render.render_rays, render.py:105-279 with _rgb_from_points (render.py:437-549).

    params: dict with 'density_vector' (3,Cd,G), 'density_matrix' (3,Cd,G,G),
    'appearance_vector', 'appearance_matrix' and the MLP entries of
    init_mlp_params. jitter: (N,) bbox mode or (R, N) contraction mode.
    With cfg.sampling == SAMPLING_EDGES the samples come from ``edges`` (R, N+1)
    (sample_from_edges) and jitter is (R, N) or None.
    mlp_operand_round (e.g. bf16_round) rounds the operands of the MLP's matrix
    products like a reduced-precision MFMA (feature_mlp_forward); the backward
    applies the same rounding.
    Returns (output, cache): output (R,3) for RGB, (R,) for the depth modes."""
    o = _c(origins, dtype)
    d = _c(directions, dtype)
    r = o.shape[0]
    n = cfg.density_samples_per_ray
    if cfg.sampling == SAMPLING_EDGES:
        pts, ts, steps = sample_from_edges(o, d, edges, jitter, cfg.scene_contraction, dtype)
    elif cfg.scene_contraction:
        pts, ts, steps = sample_contracted(o, d, n, cfg.near, cfg.far, jitter, dtype,
                                           cfg.contracted_spacing)
    else:
        pts, ts, steps = sample_points_along_ray_within_bbox(o, d, aabb, n, jitter, dtype)
    xhat = normalize_points(pts, aabb, dtype)                        # render.py:193-195
    flat = xhat.reshape(3, r * n)
    dfeat = vm_interpolate(params["density_vector"], params["density_matrix"], flat, dtype)
    s = dfeat.sum(axis=0).reshape(r, n) + dtype(cfg.density_shift)  # render.py:207
    sigma = softplus(s).astype(dtype)
    p_exits, w = compute_segment_probabilities(sigma, steps, dtype)  # render.py:211
    cache = dict(xhat=xhat, ts=ts, steps=steps, s=s, sigma=sigma, p_exits=p_exits, w=w, d=d,
                 cams=np.asarray(camera_indices, np.int64))
    if cfg.mode == DIST_MEDIAN and cfg.median_as_written:
        # render.py:250-266 literally: pad the distances with +inf and the
        # "not alive" probabilities with 1, xor neighbouring mask entries, and
        # sum mask * distance under IEEE rules (0 * inf = NaN).
        dist = np.concatenate([ts, np.full((r, 1), np.inf, dtype)], axis=1)
        alive_not = np.concatenate([1.0 - p_exits, np.ones((r, 1), dtype)], axis=1)
        mask = alive_not > 0.5
        sel = np.zeros_like(mask)
        sel[:, 1:] = np.logical_xor(mask[:, :-1], mask[:, 1:])
        with np.errstate(invalid="ignore"):
            out = (sel.astype(dtype) * dist).sum(axis=1)
        return out.astype(dtype), cache
    if cfg.mode == DIST_MEDIAN:
        # render.py:248-267: the first sample whose 1-E exceeds 0.5; 0 when it is
        # sample 0 (xor quirk), +inf when never reached. The reference multiplies
        # a 0/1 mask by a distance row that ends in +inf; this is the selection
        # the finite columns of that sum make (median_as_written evaluates the
        # sum itself).
        above = (1.0 - p_exits) > 0.5
        first = np.where(above.any(axis=1), above.argmax(axis=1), n)
        out = np.where(first == n, np.inf, np.where(first == 0, 0.0,
                       ts[np.arange(r), np.minimum(first, n - 1)]))
        return out.astype(dtype), cache
    if cfg.mode == DIST_MEAN:
        # render.py:269-276
        return ((w * ts).sum(axis=1) + p_exits[:, -1] * ts[:, -1]).astype(dtype), cache
    k = cfg.appearance_samples_per_ray
    idx = topk_indices(w, k, gumbel)                                  # render.py:461-469
    rows = np.arange(r)[:, None]
    sel = xhat[:, rows, idx].reshape(3, r * k)                       # render.py:472
    afeat = vm_interpolate(params["appearance_vector"], params["appearance_matrix"], sel,
                           dtype).T                                  # render.py:476-486
    vd = np.repeat(d, k, axis=0)                                     # render.py:487-490
    cams = np.repeat(np.asarray(camera_indices, np.int64), k)        # render.py:492-496
    rgb, mlp_cache = feature_mlp_forward(params, afeat, vd, cams, shape, dtype,
                                         mlp_operand_round)          # render.py:499-509
    rgb = rgb.reshape(r, k, 3)
    wsel = w[rows, idx]                                              # render.py:529-531
    e_last = p_exits[:, -1]
    ssum = wsel.sum(axis=1)
    unbias = (dtype(1.0) - e_last + dtype(1e-8)) / (ssum + dtype(1e-8))  # render.py:537-546
    expected = (rgb * wsel[:, :, None]).sum(axis=1) * unbias[:, None]  # render.py:233-236
    out = expected + e_last[:, None]                                 # render.py:241-244
    cache.update(idx=idx, sel=sel, afeat=afeat, mlp=mlp_cache, rgb=rgb, wsel=wsel,
                 unbias=unbias, ssum=ssum, e_last=e_last)
    return out.astype(dtype), cache


def render_rays_backward(params, cache, grad_out, cfg: RenderConfig, shape: MlpShape,
                         dtype=np.float64):
    """Analytic VJP of render_rays (RGB mode) w.r.t. the learnable parameters.

    With A = sum_K w_k c_k, S = sum_K w_k, E = E_{N-1}, U = (1-E+eps)/(S+eps),
    a_i = sigma_i * delta_i and g = dL/dC:
      dw_k/da_i = -w_k (i<k), E_k (i=k), 0 (i>k);  dE/da_i = -E
      dA/da_i = [i in K] E_i c_i - sum_{k in K, k>i} w_k c_k
      dS/da_i = [i in K] E_i - sum_{k in K, k>i} w_k
      dU/da_i = E/(S+eps) - U (dS/da_i)/(S+eps)
      dL/da_i = g . (dU/da_i A + U dA/da_i - E 1)
      dL/dc_k = U w_k g,  dsigma/ds = sigmoid(s)   (s includes the +10 shift).
    Returns a dict of gradients keyed like params."""
    if cfg.mode != RGB:
        raise ValueError("render_rays_backward: RGB mode only (the loss is on RGB)")
    g = _c(grad_out, dtype)
    w = cache["w"]
    e = cache["p_exits"]
    idx = cache["idx"]
    rgb = cache["rgb"]
    r, n = w.shape
    k = idx.shape[1]
    rows = np.arange(r)[:, None]
    eps = dtype(1e-8)
    e_last = cache["e_last"]
    u = cache["unbias"]
    ssum = cache["ssum"]
    # Dense per-sample selected colour (zero where not selected).
    cdense = np.zeros((r, n, 3), dtype)
    cdense[rows, idx] = rgb
    seld = np.zeros((r, n), dtype)
    seld[rows, idx] = 1.0
    wc = w[:, :, None] * cdense
    a_vec = wc.sum(axis=1)                                           # (R,3)
    # Exclusive suffix sums over k > i.
    suf_wc = np.flip(np.cumsum(np.flip(wc, 1), 1), 1) - wc
    suf_w = np.flip(np.cumsum(np.flip(w * seld, 1), 1), 1) - w * seld
    da = seld[:, :, None] * e[:, :, None] * cdense - suf_wc          # dA/da_i
    ds = seld * e - suf_w                                            # dS/da_i
    du = e_last[:, None] / (ssum[:, None] + eps) - u[:, None] * ds / (ssum[:, None] + eps)
    dc_da = du[:, :, None] * a_vec[:, None, :] + u[:, None, None] * da - e_last[:, None, None]
    dl_da = (dc_da * g[:, None, :]).sum(axis=-1)                     # (R,N)
    dl_dsigma = dl_da * cache["steps"]
    dl_ds = dl_dsigma * sigmoid(cache["s"])
    # Density gather transpose: every channel receives dL/ds.
    cd = np.asarray(params["density_vector"]).shape[1]
    gfeat_d = np.broadcast_to(dl_ds.reshape(1, r * n), (3 * cd, r * n))
    gdv, gdm = vm_interpolate_vjp(params["density_vector"], params["density_matrix"],
                                  cache["xhat"].reshape(3, r * n), gfeat_d, dtype)
    # Appearance branch.
    wsel = cache["wsel"]
    dl_dc = (u[:, None, None] * wsel[:, :, None] * g[:, None, :]).reshape(r * k, 3)
    mlp_grads, gafeat = feature_mlp_backward(params, cache["mlp"], dl_dc, shape, dtype)
    gav, gam = vm_interpolate_vjp(params["appearance_vector"], params["appearance_matrix"],
                                  cache["sel"], gafeat.T, dtype)
    out = dict(mlp_grads)
    out.update(density_vector=gdv, density_matrix=gdm, appearance_vector=gav,
               appearance_matrix=gam)
    return out


# --------------------------------------------------------------------------
# Optimizer and training step
# --------------------------------------------------------------------------
MLP_KEYS = ("w0", "w1", "b1", "w2", "b2", "w3", "b3", "embed")
TENSOR_KEYS = ("density_vector", "density_matrix", "appearance_vector", "appearance_matrix")


@dataclasses.dataclass
class OptimizerConfig:
    """train_config.OptimizerConfig, train_config.py:6-12."""
    lr_init_tensor: float = 0.02
    lr_init_mlp: float = 1e-3
    lr_decay_iters: Optional[int] = None
    lr_decay_target_ratio: float = 0.1
    lr_upsample_reset: bool = True


def lr_decay_coeff(step, n_iters, upsamp_iters, opt: OptimizerConfig):
    """training.py:161-181: steps since the last upsample (or since 0), then
    optax.exponential_decay(1.0, decay_iters, ratio, end_value=ratio)."""
    if opt.lr_upsample_reset:
        deltas = [step - u for u in (0,) + tuple(upsamp_iters) if step - u >= 0]
        s = min(deltas)
    else:
        s = step
    decay_iters = opt.lr_decay_iters if opt.lr_decay_iters is not None else n_iters
    v = np.float32(opt.lr_decay_target_ratio) ** (np.float32(s) / np.float32(decay_iters))
    return float(max(v, np.float32(opt.lr_decay_target_ratio)))


def adam_update(p, g, m, v, count, lr, decay, b1=0.9, b2=0.99, eps=1e-8, dtype=np.float64):
    """optax.scale_by_adam(b1=.9, b2=.99, eps=1e-8, eps_root=0) -> scale(-lr) ->
    x decay -> apply_updates (training.py:184-205, 213-243). ``count`` is the
    Adam step count after the increment (1 on the first update).
    Returns (p', m', v')."""
    p = _c(p, dtype)
    g = _c(g, dtype)
    m = (1 - b1) * g + b1 * _c(m, dtype)
    v = (1 - b2) * g * g + b2 * _c(v, dtype)
    m_hat = m / (1 - b1 ** count)
    v_hat = v / (1 - b2 ** count)
    upd = m_hat / (np.sqrt(v_hat) + eps)
    upd = upd * (-lr) * decay
    return (p + upd).astype(dtype), m.astype(dtype), v.astype(dtype)


def psnr_from_mse(mse):
    """utils.py:4-9."""
    return -10.0 * np.log10(max(float(mse), 1e-10))


def train_step(params, opt_state, step, aabb, origins, directions, camera_indices, colors,
               cfg: RenderConfig, shape: MlpShape, jitter, n_iters, upsamp_iters=(),
               opt: OptimizerConfig = OptimizerConfig(), gumbel=None, dtype=np.float64):
    """TrainState.training_step, training.py:101-205 (fwd + analytic bwd + Adam).

    opt_state: {'m': {...}, 'v': {...}, 'count': int}. Returns
    (params', opt_state', step+1, logs)."""
    out, cache = render_rays(params, aabb, origins, directions, camera_indices, cfg, shape,
                             jitter, gumbel, dtype)
    diff = out - _c(colors, dtype)
    mse = float(np.mean(diff * diff))                                # training.py:140
    gout = 2.0 * diff / diff.size
    grads = render_rays_backward(params, cache, gout, cfg, shape, dtype)
    decay = lr_decay_coeff(step, n_iters, upsamp_iters, opt)
    count = opt_state["count"] + 1
    new_p, new_m, new_v = {}, {}, {}
    for key, val in params.items():
        if key not in grads:
            new_p[key] = val
            continue
        lr = opt.lr_init_mlp if key in MLP_KEYS else opt.lr_init_tensor
        new_p[key], new_m[key], new_v[key] = adam_update(
            val, grads[key], opt_state["m"][key], opt_state["v"][key], count, lr, decay,
            dtype=dtype)
    gnorm = math.sqrt(sum(float((gv.astype(np.float64) ** 2).sum()) for gv in grads.values()))
    logs = {"mse": mse, "psnr": psnr_from_mse(mse), "lr_tensor": decay * opt.lr_init_tensor,
            "lr_mlp": decay * opt.lr_init_mlp, "grad_norm": gnorm}
    return new_p, {"m": new_m, "v": new_v, "count": count}, step + 1, logs, grads, out


def init_opt_state(params):
    return {"m": {k: np.zeros_like(v) for k, v in params.items()},
            "v": {k: np.zeros_like(v) for k, v in params.items()}, "count": 0}


# --------------------------------------------------------------------------
# Synthetic inputs (the shapes of the benchmark configurations)
# --------------------------------------------------------------------------
def make_params(grid_dim, cd, ca, shape: MlpShape, seed=0, std=0.1):
    """Planes and lines ~ N(0, std^2) (training.py:44) in the reference layout
    (tensor_vm.py:132-133) plus MLP parameters."""
    rng = np.random.default_rng(seed)
    p = {
        "density_vector": rng.normal(0, std, (3, cd, grid_dim)),
        "density_matrix": rng.normal(0, std, (3, cd, grid_dim, grid_dim)),
        "appearance_vector": rng.normal(0, std, (3, ca, grid_dim)),
        "appearance_matrix": rng.normal(0, std, (3, ca, grid_dim, grid_dim)),
    }
    p.update(init_mlp_params(shape, rng))
    return p


def make_rays(n_rays, seed=1, radius=4.0, target_half=0.5):
    """Origins uniform on a sphere of the given radius, each aimed at a uniform point
    of [-target_half, target_half]^3; unit directions."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n_rays, 3))
    o = radius * v / np.linalg.norm(v, axis=1, keepdims=True)
    tgt = rng.uniform(-target_half, target_half, size=(n_rays, 3))
    d = tgt - o
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    return o, d
