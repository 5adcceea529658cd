"""Pin the proposal-sampler oracle (no reference counterpart: render.py:126 is a
comment) against numpy.interp, brute force, the bound property of a shared
density, torch float64 autograd and central finite differences."""
import numpy as np
import pytest
import torch

from oracle import proposal_oracle as po
from oracle import tensorf_oracle as orc

BOX = np.array([[-1.5, -1.5, -1.5], [1.5, 1.5, 1.5]])


def _sorted_edges(rng, r, n, lo=1.0, hi=5.0):
    e = np.sort(rng.uniform(lo, hi, (r, n + 1)), axis=1)
    e[:, 0], e[:, -1] = lo, hi
    return e


@pytest.mark.parametrize("power", [1.0, 0.5, 0.0])
def test_resample_matches_interp(power):
    """The inverse of a piecewise-linear CDF is numpy.interp with the axes swapped."""
    rng = np.random.default_rng(0)
    r, n, n_out, pad = 20, 37, 23, 0.01
    e = _sorted_edges(rng, r, n)
    w = rng.exponential(size=(r, n)) * (rng.uniform(size=(r, n)) < 0.3)
    w /= w.sum(1, keepdims=True) * 1.2
    out = po.resample(e, w, n_out, pad, power=power)
    h = w ** power + pad / n
    for ri in range(r):
        cdf = np.concatenate([[0.0], np.cumsum(h[ri])])
        q = cdf[-1] * np.arange(n_out + 1) / n_out
        np.testing.assert_allclose(out[ri], np.interp(q, cdf, e[ri]), rtol=0, atol=1e-12)
    assert np.all(np.diff(out, axis=1) >= 0)
    np.testing.assert_array_equal(out[:, 0], e[:, 0])
    np.testing.assert_array_equal(out[:, -1], e[:, -1])


def test_resample_places_samples_by_mass():
    """With zero pad, all new edges fall inside the intervals that hold mass."""
    rng = np.random.default_rng(1)
    e = _sorted_edges(rng, 5, 16)
    w = np.zeros((5, 16))
    w[:, 7] = 0.9
    out = po.resample(e, w, 10, 0.0)
    inner = out[:, 1:-1]
    assert np.all(inner >= e[:, 7:8]) and np.all(inner <= e[:, 8:9])
    # evenly spaced inside the interval (linear CDF there)
    np.testing.assert_allclose(np.diff(out[:, 1:-1], axis=1),
                               np.broadcast_to(((e[:, 8] - e[:, 7]) / 10)[:, None], (5, 8)),
                               rtol=1e-9)


def _brute_bounds(fe, pw, pe):
    nf, n = len(fe) - 1, len(pe) - 1
    bound = np.zeros(nf)
    for i in range(nf):
        for j in range(n):
            if pe[j + 1] > fe[i] and pe[j] < fe[i + 1]:
                bound[i] += pw[j]
    return bound


def test_interval_loss_brute_force_and_fd():
    rng = np.random.default_rng(2)
    r, nf, n = 6, 19, 11
    fe = _sorted_edges(rng, r, nf)
    pe = _sorted_edges(rng, r, n)
    # exact shared edges: half-open intervals do not overlap there
    fe[0, 5], fe[0, 9] = pe[0, 3], pe[0, 4]
    fe[0] = np.sort(fe[0])
    fw = rng.uniform(0, 0.2, (r, nf))
    pw = rng.uniform(0, 0.1, (r, n))
    scale = 0.37
    loss, grad = po.interval_loss(fe, fw, pe, pw, scale)
    for ri in range(r):
        b = _brute_bounds(fe[ri], pw[ri], pe[ri])
        want = scale * (np.maximum(fw[ri] - b, 0) ** 2 / (fw[ri] + po.LOSS_EPS)).sum()
        assert abs(loss[ri] - want) < 1e-12
        for j in range(n):
            h = 1e-6
            pp, pm = pw[ri].copy(), pw[ri].copy()
            pp[j] += h
            pm[j] -= h
            lp = po.interval_loss(fe[ri:ri + 1], fw[ri:ri + 1], pe[ri:ri + 1], pp[None], scale)[0]
            lm = po.interval_loss(fe[ri:ri + 1], fw[ri:ri + 1], pe[ri:ri + 1], pm[None], scale)[0]
            assert abs((lp - lm)[0] / (2 * h) - grad[ri, j]) < 1e-6


def test_bound_dominates_for_a_shared_density():
    """Field and proposal see one piecewise-constant density: the exact weight
    of every field interval is at most its bound, so the loss vanishes."""
    rng = np.random.default_rng(3)
    r, n, nf = 8, 24, 40
    pe = _sorted_edges(rng, r, n)
    sig = rng.exponential(2.0, (r, n)) * (rng.uniform(size=(r, n)) < 0.5)
    _, pw = orc.compute_segment_probabilities(sig, np.diff(pe, axis=1))
    fe = _sorted_edges(rng, r, nf)  # arbitrary, not aligned with the proposal edges

    def tau(ri, t):  # optical depth from the ray start to t
        seg = np.clip(t - pe[ri, :-1], 0, np.diff(pe[ri]))
        return (sig[ri] * seg).sum()

    fw = np.zeros((r, nf))
    for ri in range(r):
        T = np.exp(-np.array([tau(ri, t) for t in fe[ri]]))
        fw[ri] = T[:-1] - T[1:]
        b = _brute_bounds(fe[ri], pw[ri], pe[ri])
        assert np.all(b >= fw[ri] - 1e-12)
    loss, _ = po.interval_loss(fe, fw, pe, pw)
    assert np.all(loss < 1e-18)


def test_segment_vjp_matches_autograd():
    rng = np.random.default_rng(4)
    sig = rng.exponential(1.5, (5, 33))
    st = rng.uniform(0.01, 0.3, (5, 33))
    gw = rng.normal(size=(5, 33))
    ge = rng.normal(size=(5, 33))
    a = torch.tensor(sig * st, dtype=torch.float64, requires_grad=True)
    e = torch.exp(torch.cumsum(-a, -1))
    prev = torch.cat([torch.ones(5, 1, dtype=torch.float64), e[:, :-1]], -1)
    w = (1 - torch.exp(-a)) * prev
    (w * torch.tensor(gw) + e * torch.tensor(ge)).sum().backward()
    E, W = orc.compute_segment_probabilities(sig, st)
    np.testing.assert_allclose(po.segment_probabilities_vjp(gw, ge, W, E), a.grad.numpy(),
                               rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(po.segment_probabilities_vjp(gw, None, W, E),
                               po.segment_probabilities_vjp(gw, np.zeros_like(ge), W, E))


def test_prop_mlp_matches_torch():
    rng = np.random.default_rng(5)
    p = po.init_prop_params(4, 8, rng)
    p["pb1"] = rng.normal(0, 0.1, 64)
    p["pb2"] = rng.normal(0, 0.1, 64)
    p["pb3"] = np.array([0.3])
    f = rng.normal(0, 0.5, (50, 24))
    dz = rng.normal(size=50)
    z, sig, cache = po.prop_mlp_forward(p, f)
    tp = {k: torch.tensor(p[k], requires_grad=True) for k in po.PROP_MLP_KEYS}
    tf = torch.tensor(f, requires_grad=True)
    h = torch.relu(tf @ tp["pw1"] + tp["pb1"])
    h = torch.relu(h @ tp["pw2"] + tp["pb2"])
    tz = (h @ tp["pw3"] + tp["pb3"])[:, 0]
    np.testing.assert_allclose(z, tz.detach().numpy(), rtol=1e-12)
    np.testing.assert_allclose(sig, torch.nn.functional.softplus(tz).detach().numpy(), rtol=1e-12)
    (tz * torch.tensor(dz)).sum().backward()
    g, df = po.prop_mlp_backward(p, cache, dz)
    for k in po.PROP_MLP_KEYS:
        np.testing.assert_allclose(g[k], tp[k].grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(df, tf.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_edges_sampling_matches_stratified_sampler():
    """Equal intervals of the box segment with fractions u reproduce the
    reference's stratified sampler (render.py:350-390) with jitter u."""
    o, d = orc.make_rays(30, seed=6)
    n = 12
    u = np.random.default_rng(6).uniform(size=n)
    pts, ts, steps = orc.sample_points_along_ray_within_bbox(o, d, BOX, n, u)
    edges = po.round0_edges(o, d, BOX, n)
    pe, te, se = orc.sample_from_edges(o, d, edges, np.broadcast_to(u, (30, n)), False)
    np.testing.assert_allclose(te, ts, rtol=1e-13)
    np.testing.assert_allclose(se, steps, rtol=1e-12)
    np.testing.assert_allclose(pe, pts, rtol=1e-12, atol=1e-13)


def _small_case(seed=7, contraction=False):
    rng = np.random.default_rng(seed)
    cfg = po.ProposalConfig(samples_round0=10, samples_round1=7, samples_out=6,
                            scene_contraction=contraction, histogram_pad=0.05, loss_weight=0.8)
    p = po.init_prop_params(5, 2, rng, std=0.6)
    p["pb1"] = rng.normal(0, 0.2, 64)
    p["pb2"] = rng.normal(0, 0.2, 64)
    p["pb3"] = np.array([0.5])
    o, d = orc.make_rays(4, seed=seed)
    jit = rng.uniform(size=(4, 17))
    return rng, cfg, p, o, d, jit


def test_proposal_loss_backward_central_fd():
    rng, cfg, p, o, d, jit = _small_case()
    edges, rounds = po.proposal_rounds(p, cfg, BOX, o, d, jit)
    fw = rng.uniform(0, 0.5, (4, cfg.samples_out)) * (rng.uniform(size=(4, 6)) < 0.7)
    loss, grads = po.proposal_loss_backward(p, rounds, cfg, edges, fw)
    fixed = [rd["edges"] for rd in rounds]
    assert abs(po.proposal_loss(p, fixed, cfg, BOX, o, d, edges, fw, jit) - loss) < 1e-14
    assert loss > 0
    h = 1e-6
    for key in po.PROP_KEYS:
        flat = np.asarray(p[key]).reshape(-1)
        # the entries with the largest gradients plus a few random ones
        idx = set(np.argsort(-np.abs(grads[key].reshape(-1)))[:4].tolist())
        idx |= set(rng.choice(flat.size, size=min(3, flat.size), replace=False).tolist())
        for i in idx:
            pp = {k: np.array(v, copy=True) for k, v in p.items()}
            pm = {k: np.array(v, copy=True) for k, v in p.items()}
            pp[key].reshape(-1)[i] += h
            pm[key].reshape(-1)[i] -= h
            fd = (po.proposal_loss(pp, fixed, cfg, BOX, o, d, edges, fw, jit)
                  - po.proposal_loss(pm, fixed, cfg, BOX, o, d, edges, fw, jit)) / (2 * h)
            an = grads[key].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 + 1e-5 * abs(an), (key, i, fd, an)


def test_proposal_rounds_contracted_schedule():
    rng, cfg, p, o, d, jit = _small_case(seed=8, contraction=True)
    base_ts, base_steps = orc.contraction_schedule(cfg.samples_round0, 0.05, 200.0)
    box2 = np.array([[-2.0] * 3, [2.0] * 3])
    edges, rounds = po.proposal_rounds(p, cfg, box2, o, d, jit, base_ts, base_steps)
    e0 = rounds[0]["edges"]
    np.testing.assert_allclose(e0[:, :-1], np.broadcast_to(base_ts, (4, 10)))
    assert np.all(np.abs(rounds[0]["xhat"]) <= 1.0 + 1e-12)  # contracted into [-2,2]^3
    for rd in rounds:
        assert np.all(np.diff(rd["edges"], axis=1) >= 0)
        np.testing.assert_allclose(rd["w"].sum(1) + rd["e"][:, -1], 1.0, rtol=1e-12)
    assert np.all(np.diff(edges, axis=1) >= 0)
    np.testing.assert_array_equal(edges[:, 0], e0[:, 0])
    np.testing.assert_array_equal(edges[:, -1], e0[:, -1])
