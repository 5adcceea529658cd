"""CPU oracle of the learned proposal sampler (NumPy, FP64 by default).

TEST INFRASTRUCTURE, like tensorf_oracle: only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may use it.

The reference has no proposal sampler: it stops at a comment ("An occupancy
grid or proposal network would really help us here!", render.py:126). This is
the build's own definition, restated here op for op as the GPU kernels run it:

* a VM-decomposed proposal field (reference layout (3, C, G) / (3, C, G, G),
  looked up with tensorf_oracle.vm_interpolate, i.e. tensor_vm.py:42-89) whose
  3C features feed an MLP 3C -> 64 -> 64 -> 1 with ReLU and sigma = softplus(z);
* rounds of samples: round 0 splits the box segment (render.py:399-434) into
  n0 equal intervals, or takes the contracted schedule (render.py:127-155) as
  interval starts; each round's weights are compute_segment_probabilities
  (render.py:300-347) over its intervals; the next round's edges are the
  quantiles m/n' of the piecewise-linear CDF of (w^power + pad/n);
* mip-NeRF 360's interval loss (Barron et al. 2022, eq. 13) of every round
  against the field's weights, differentiated w.r.t. the proposal only.

Pinning (tests/test_oracle_proposal.py): resample against numpy.interp (the
inverse of a piecewise-linear CDF), interval_loss against a brute-force
double loop and the bound >= weight property of a shared density, the
segment VJP against torch float64 autograd, the MLP against torch float64
Linear layers, and proposal_loss_backward against central finite
differences.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional

import numpy as np

from . import tensorf_oracle as T

HIDDEN = 64
LOSS_EPS = 1e-7
PROP_KEYS = ("prop_vector", "prop_matrix", "pw1", "pb1", "pw2", "pb2", "pw3", "pb3")
PROP_MLP_KEYS = ("pw1", "pb1", "pw2", "pb2", "pw3", "pb3")


@dataclasses.dataclass
class ProposalConfig:
    samples_round0: int = 128
    samples_round1: int = 64
    samples_out: int = 64
    scene_contraction: bool = False
    histogram_pad: float = 0.64
    loss_weight: float = 1.0
    weight_power: float = 1.0


def init_prop_params(grid_dim, channels, rng: np.random.Generator, std=0.3):
    """Proposal factors ~ N(0, std^2); MLP kernels (in, out) with the Flax
    lecun-normal scale, zero biases."""
    f = 3 * channels
    return {
        "prop_vector": rng.normal(0, std, (3, channels, grid_dim)),
        "prop_matrix": rng.normal(0, std, (3, channels, grid_dim, grid_dim)),
        "pw1": rng.normal(0, 1 / np.sqrt(f), (f, HIDDEN)),
        "pb1": np.zeros(HIDDEN),
        "pw2": rng.normal(0, 1 / np.sqrt(HIDDEN), (HIDDEN, HIDDEN)),
        "pb2": np.zeros(HIDDEN),
        "pw3": rng.normal(0, 1 / np.sqrt(HIDDEN), (HIDDEN, 1)),
        "pb3": np.zeros(1),
    }


# --------------------------------------------------------------------------
# MLP
# --------------------------------------------------------------------------
def prop_mlp_forward(params, feats, dtype=np.float64):
    """z = relu(relu(f W1 + b1) W2 + b2) W3 + b3, sigma = softplus(z).
    feats (M, F). Returns (z (M,), sigma (M,), cache)."""
    f = T._c(feats, dtype)
    a1 = f @ T._c(params["pw1"], dtype) + T._c(params["pb1"], dtype)
    h1 = np.maximum(a1, 0)
    a2 = h1 @ T._c(params["pw2"], dtype) + T._c(params["pb2"], dtype)
    h2 = np.maximum(a2, 0)
    z = (h2 @ T._c(params["pw3"], dtype) + T._c(params["pb3"], dtype))[:, 0]
    return z, T.softplus(z).astype(dtype), dict(f=f, h1=h1, h2=h2, z=z)


def prop_mlp_backward(params, cache, dz, dtype=np.float64):
    """VJP of prop_mlp_forward w.r.t. z. Returns (grads {pw1..pb3}, dfeats (M, F))."""
    dz = T._c(dz, dtype)[:, None]
    f, h1, h2 = cache["f"], cache["h1"], cache["h2"]
    g = {"pw3": h2.T @ dz, "pb3": dz.sum(axis=0)}
    d2 = (dz @ T._c(params["pw3"], dtype).T) * (h2 > 0)
    g["pw2"] = h1.T @ d2
    g["pb2"] = d2.sum(axis=0)
    d1 = (d2 @ T._c(params["pw2"], dtype).T) * (h1 > 0)
    g["pw1"] = f.T @ d1
    g["pb1"] = d1.sum(axis=0)
    return g, d1 @ T._c(params["pw1"], dtype).T


# --------------------------------------------------------------------------
# Rounds
# --------------------------------------------------------------------------
def round0_edges(origins, directions, aabb, n, contraction=False, base_ts=None,
                 base_steps=None, dtype=np.float64):
    """(R, n+1) edges of round 0: the box segment of render.py:399-434 split into
    n equal intervals, or the contracted schedule's starts (render.py:127-155)
    closed by its last step."""
    r = np.asarray(origins).shape[0]
    if contraction:
        bt = T._c(base_ts, dtype)
        bs = T._c(base_steps, dtype)
        e = np.concatenate([bt, bt[-1:] + bs[-1:]])
        return np.tile(e[None, :], (r, 1)).astype(dtype)
    t0, t1 = T.ray_segment_from_bounding_box(origins, directions, aabb, 1e-3, dtype)
    step = (t1 - t0) / dtype(n)
    j = np.arange(n + 1, dtype=dtype)
    return (t0[:, None] + j[None, :] * step[:, None]).astype(dtype)


def round_points(origins, directions, edges, jitter, aabb, contraction, dtype=np.float64):
    """Normalised sample points (3, R, n) of a round (tensorf_oracle.sample_from_edges
    then render.py:192-196)."""
    pts, _, steps = T.sample_from_edges(origins, directions, edges, jitter, contraction, dtype)
    return T.normalize_points(pts, aabb, dtype), steps


def resample(edges, w, n_out, pad, dtype=np.float64, power=1.0):
    """Inverse-CDF edges (R, n_out+1): h = w^power + pad/n, F = [0, cumsum(h)], out[m] =
    F^{-1}(F_n m / n_out) piecewise linearly between the edges, out[0] = e_0,
    out[n_out] = e_n."""
    e = T._c(edges, dtype)
    w = T._c(w, dtype)
    r, n = w.shape
    h = (w if power == 1.0 else np.power(w, power)) + dtype(pad) / dtype(n)
    cdf = np.concatenate([np.zeros((r, 1), dtype), np.cumsum(h, axis=1)], axis=1)
    out = np.empty((r, n_out + 1), dtype)
    out[:, 0] = e[:, 0]
    out[:, -1] = e[:, -1]
    m = np.arange(1, n_out, dtype=dtype) / dtype(n_out)
    for ri in range(r):
        q = cdf[ri, n] * m
        j = np.clip(np.searchsorted(cdf[ri, :n], q, side="right") - 1, 0, n - 1)
        lo = cdf[ri, j]
        den = cdf[ri, j + 1] - lo
        with np.errstate(divide="ignore", invalid="ignore"):
            frac = np.where(den > 0, np.clip((q - lo) / den, 0, 1), 0)
        e0, e1 = e[ri, j], e[ri, j + 1]
        out[ri, 1:n_out] = np.clip(e0 + frac * (e1 - e0), e0, e1)
    return out


def interval_loss(fedges, fw, pedges, pw, scale=1.0, dtype=np.float64):
    """Per-ray interval loss scale * sum_i max(0, w_i - bound_i)^2 / (w_i + eps),
    bound_i = sum of pw_j over the proposal intervals [pe_j, pe_{j+1}) that
    overlap the field interval [fe_i, fe_{i+1}) (pe_{j+1} > fe_i and pe_j <
    fe_{i+1}). Returns (loss (R,), dloss/dpw (R, n_p))."""
    fe, fw, pe, pw = (T._c(x, dtype) for x in (fedges, fw, pedges, pw))
    r, nf = fw.shape
    n = pw.shape[1]
    loss = np.zeros(r, dtype)
    grad = np.zeros((r, n), dtype)
    for ri in range(r):
        W = np.concatenate([[0.0], np.cumsum(pw[ri])]).astype(dtype)
        a, b = fe[ri, :-1], fe[ri, 1:]
        jlo = np.searchsorted(pe[ri, 1:], a, side="right")
        jend = np.searchsorted(pe[ri, :-1], b, side="left")
        bound = np.where(jend > jlo, W[jend] - W[np.minimum(jlo, n)], 0.0)
        dlt = np.maximum(fw[ri] - bound, 0.0)
        inv = 1.0 / (fw[ri] + LOSS_EPS)
        loss[ri] = scale * (dlt * dlt * inv).sum()
        g = -2.0 * dlt * inv * scale
        P = np.concatenate([[0.0], np.cumsum(g)])
        c_lo = np.searchsorted(a, pe[ri, 1:], side="left")
        c_hi = np.searchsorted(b, pe[ri, :-1], side="right")
        grad[ri] = np.where(c_lo > c_hi, P[c_lo] - P[np.minimum(c_hi, nf)], 0.0)
    return loss, grad


def segment_probabilities_vjp(gw, ge, w, e, dtype=np.float64):
    """VJP of compute_segment_probabilities (render.py:300-347) w.r.t. a = sigma
    delta: dL/da_i = gw_i E_i - sum_{j>i} gw_j w_j - sum_{j>=i} ge_j E_j."""
    w = T._c(w, dtype)
    e = T._c(e, dtype)
    gw = np.zeros_like(w) if gw is None else T._c(gw, dtype)
    ge = np.zeros_like(w) if ge is None else T._c(ge, dtype)
    gww = gw * w
    suffix_excl = np.flip(np.cumsum(np.flip(gww, -1), -1), -1) - gww
    suffix_incl_e = np.flip(np.cumsum(np.flip(ge * e, -1), -1), -1)
    return gw * e - suffix_excl - suffix_incl_e


def proposal_rounds(params, cfg: ProposalConfig, aabb, origins, directions, jitter=None,
                    base_ts=None, base_steps=None, fixed_edges: Optional[List] = None,
                    dtype=np.float64):
    """Forward of the proposal rounds. jitter: (R, n0 + n1) or None (midpoints).
    fixed_edges: optional per-round edges to use instead of the resampled ones
    (round 0 included), e.g. the GPU's, so that a gradient check sees the same
    intervals. Returns (edges_out (R, samples_out + 1), rounds), rounds[k] a dict
    with edges, xhat, steps, feats, z, sigma, w, e, mlp cache."""
    o = T._c(origins, dtype)
    d = T._c(directions, dtype)
    ns = [cfg.samples_round0] + ([cfg.samples_round1] if cfg.samples_round1 > 0 else [])
    rounds = []
    joff = 0
    edges = round0_edges(o, d, aabb, ns[0], cfg.scene_contraction, base_ts, base_steps, dtype)
    for k, n in enumerate(ns):
        if fixed_edges is not None:
            edges = T._c(fixed_edges[k], dtype)
        jit = None if jitter is None else T._c(jitter, dtype)[:, joff:joff + n]
        xhat, steps = round_points(o, d, edges, jit, aabb, cfg.scene_contraction, dtype)
        feats = T.vm_interpolate(params["prop_vector"], params["prop_matrix"],
                                 xhat.reshape(3, -1), dtype).T
        z, sigma, cache = prop_mlp_forward(params, feats, dtype)
        sigma = sigma.reshape(o.shape[0], n)
        e, w = T.compute_segment_probabilities(sigma, steps, dtype)
        rounds.append(dict(edges=edges, xhat=xhat, steps=steps, feats=feats, z=z, sigma=sigma,
                           w=w, e=e, mlp=cache))
        n_next = ns[k + 1] if k + 1 < len(ns) else cfg.samples_out
        edges = resample(edges, w, n_next, cfg.histogram_pad, dtype, cfg.weight_power)
        joff += n
    return edges, rounds


def proposal_loss(params, rounds_edges, cfg, aabb, origins, directions, fedges, fw,
                  jitter=None, base_ts=None, base_steps=None, dtype=np.float64):
    """Total interval loss (scalar) with every round's edges fixed (for finite
    differences)."""
    _, rounds = proposal_rounds(params, cfg, aabb, origins, directions, jitter, base_ts,
                                base_steps, rounds_edges, dtype)
    r = np.asarray(origins).shape[0]
    scale = cfg.loss_weight / r
    return sum(float(interval_loss(fedges, fw, rd["edges"], rd["w"], scale, dtype)[0].sum())
               for rd in rounds)


def proposal_loss_backward(params, rounds, cfg: ProposalConfig, fedges, fw, dtype=np.float64):
    """Interval loss of every round against the field's weights fw (R, n) on
    fedges (R, n+1), and its gradient w.r.t. the proposal parameters (the
    field's weights and every round's edges are constants).
    Returns (loss, grads {PROP_KEYS})."""
    r = np.asarray(fw).shape[0]
    scale = cfg.loss_weight / r
    grads = {k: np.zeros_like(np.asarray(params[k], dtype)) for k in PROP_KEYS}
    total = 0.0
    for rd in rounds:
        loss, gpw = interval_loss(fedges, fw, rd["edges"], rd["w"], scale, dtype)
        total += float(loss.sum())
        da = segment_probabilities_vjp(gpw, None, rd["w"], rd["e"], dtype)
        dsig = da * rd["steps"]
        dz = dsig.reshape(-1) * T.sigmoid(rd["z"])
        g, dfeat = prop_mlp_backward(params, rd["mlp"], dz, dtype)
        for k2, v in g.items():
            grads[k2] += v
        gvec, gmat = T.vm_interpolate_vjp(params["prop_vector"], params["prop_matrix"],
                                          rd["xhat"].reshape(3, -1), dfeat.T, dtype)
        grads["prop_vector"] += gvec
        grads["prop_matrix"] += gmat
    return total, grads


def train_step_proposal(params, prop_params, opt_state, prop_opt_state, step, aabb, origins,
                        directions, camera_indices, colors, cfg: "T.RenderConfig",
                        pcfg: ProposalConfig, shape: "T.MlpShape", prop_jitter, field_jitter,
                        n_iters, upsamp_iters=(), opt: T.OptimizerConfig = T.OptimizerConfig(),
                        gumbel=None, base_ts=None, base_steps=None, dtype=np.float64):
    """Training step with the proposal sampler: proposal rounds -> field render
    on the proposal's intervals -> MSE backward + interval loss backward ->
    Adam on both parameter sets (proposal factors at the tensor rate, its MLP
    at the MLP rate). Returns (params', prop_params', opt', prop_opt', logs, out)."""
    edges, rounds = proposal_rounds(prop_params, pcfg, aabb, origins, directions, prop_jitter,
                                    base_ts, base_steps, None, dtype)
    out, cache = T.render_rays(params, aabb, origins, directions, camera_indices, cfg, shape,
                               field_jitter, gumbel, dtype, edges=edges)
    diff = out - T._c(colors, dtype)
    mse = float(np.mean(diff * diff))
    grads = T.render_rays_backward(params, cache, 2.0 * diff / diff.size, cfg, shape, dtype)
    ploss, pgrads = proposal_loss_backward(prop_params, rounds, pcfg, edges, cache["w"], dtype)
    decay = T.lr_decay_coeff(step, n_iters, upsamp_iters, opt)

    def adam_all(ps, gs, st, mlp_keys):
        count = st["count"] + 1
        np_, nm, nv = {}, {}, {}
        for key, val in ps.items():
            if key not in gs:
                np_[key] = val
                continue
            lr = opt.lr_init_mlp if key in mlp_keys else opt.lr_init_tensor
            np_[key], nm[key], nv[key] = T.adam_update(val, gs[key], st["m"][key], st["v"][key],
                                                       count, lr, decay, dtype=dtype)
        return np_, {"m": nm, "v": nv, "count": count}

    new_p, new_st = adam_all(params, grads, opt_state, T.MLP_KEYS)
    new_pp, new_pst = adam_all(prop_params, pgrads, prop_opt_state, PROP_MLP_KEYS)
    logs = {"mse": mse, "prop_loss": ploss}
    return new_p, new_pp, new_st, new_pst, logs, out, edges
