"""The XLA custom-call entry points (tfh_xla_render_fwd / _bwd /
tfh_xla_train_step, include/tensorf_hip.h), called exactly as XLA's
custom-call thunk calls them -- stream, void** buffers (operands then results),
the ASCII descriptor and its length -- against the plain entry points on the
same inputs: the forward bit for bit (same kernels, no atomics), the VJP and
the training step within the rounding of their float atomics (summation order
changes from run to run): gradients within 1e-5 of the largest, the MSE within
1e-6 relative, Adam updates off by more than 1e-3 lr for at most 1e-3 of the
parameters (a near-zero gradient can flip sign)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import tensorf_oracle as orc
from tests.gpu_common import dev, f32, log, mlp_module, oracle_params, to_device_params

pytestmark = pytest.mark.gpu


def _env(r, n, k, precision, ncam, contraction, prng=5):
    from tensorf_amd import cameras, render
    shape = orc.MlpShape(num_cameras=ncam)
    p = oracle_params(24, 16, 48, shape, seed=2, std=0.3, bias_std=0.1)
    o, d = orc.make_rays(r, seed=3, radius=3.0 if contraction else 4.0, target_half=0.6)
    den, app, mp = to_device_params(p, shape)
    lp = render.LearnableParams(mp, app, den, contraction)
    cams = np.arange(r) % (ncam or 1)
    rays = cameras.Rays3D(dev(f32(o)), dev(f32(d)), dev(cams, torch.int32))
    box = 2.0 if contraction else 1.5
    aabb = dev(np.array([[-box] * 3, [box] * 3]))
    cfg = render.RenderConfig(0.05, 200.0, render.RenderMode.RGB, n, k)
    return render._RenderEnv(mlp_module(shape, precision), lp, aabb, rays, prng, cfg), shape


def _dummy():
    return torch.zeros(1, dtype=torch.float32, device="cuda")


def _operands(env, keep):
    """Operands 0-15 of the render entry points."""
    from tensorf_amd import _lib
    lv = env.leaves()
    emb = lv[11] if len(lv) > 11 else None
    if emb is None:
        emb = _dummy()
        keep.append(emb)
    return [_lib.ptr(t) for t in lv[:11]] + [_lib.ptr(emb), _lib.ptr(env.aabb), _lib.ptr(env.o),
                                             _lib.ptr(env.d), _lib.ptr(env.cams)]


def _desc(env, ws_bytes, **kw):
    from tensorf_amd import xla_abi
    return xla_abi.descriptor(env.cfg, env.dvm, env.avm, env.cmlp, env.r, ws_bytes,
                              jitter=env.jitter is not None, schedule=env.base_ts is not None,
                              gumbel=env.gumbel is not None, **kw)


def _noise(env, keep):
    from tensorf_amd import _lib
    out = []
    for t in (env.jitter, env.base_ts, env.base_steps, env.gumbel):
        if t is None:
            t = _dummy()
            keep.append(t)
        out.append(_lib.ptr(t))
    return out


@pytest.mark.parametrize("n,k,precision,ncam,contraction,alias", [
    (64, 16, "fp32", None, False, True), (64, 64, "bf16", 3, True, False),
    (96, 24, "bf16", None, False, False)])
def test_xla_render_fwd_bwd_match_plain_entry_points(n, k, precision, ncam, contraction, alias):
    log(f"case: xla render fwd/bwd N={n} k={k} {precision} ncam={ncam} contraction={contraction} "
        f"workspace aliased={alias}")
    from tensorf_amd import _lib, xla_abi
    r = 384
    env, shape = _env(r, n, k, precision, ncam, contraction)
    want = env.forward()
    g = dev(f32(np.random.default_rng(4).normal(size=(r, 3))))
    want_grads = env.backward(g)
    torch.cuda.synchronize()
    keep = []
    ws_bytes = env.ws.numel()
    out = torch.full((r, 3), float("nan"), dtype=torch.float32, device="cuda")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    ops = _operands(env, keep)
    opaque = _desc(env, ws_bytes)
    log(f"  descriptor: {opaque.decode()}")
    xla_abi.call("tfh_xla_render_fwd", _lib.stream(), ops + _noise(env, keep) +
                 [_lib.ptr(out), _lib.ptr(ws)], opaque)
    torch.cuda.synchronize()
    assert torch.equal(out, want), (out - want).abs().max().item()
    grads = [torch.full_like(t, float("nan")) for t in env.leaves()]
    if len(grads) == 11:
        grads.append(_dummy())
    ws_out = ws if alias else torch.empty_like(ws)
    xla_abi.call("tfh_xla_render_bwd", _lib.stream(),
                 ops + [_lib.ptr(g), _lib.ptr(ws)] + [_lib.ptr(t) for t in grads] +
                 [_lib.ptr(ws_out)], opaque)
    torch.cuda.synchronize()
    names = ["density plane", "density line", "appearance plane", "appearance line", "w0", "w1",
             "b1", "w2", "b2", "w3", "b3", "embed"]
    for name, a, b in zip(names, want_grads, grads):
        e = (a - b).abs().max().item()
        scale = a.abs().max().item()
        log(f"  grad {name}: max|difference| {e:.3e} of {scale:.3e}")
        assert torch.isfinite(b).all()
        assert e <= 1e-5 * scale + 1e-9, name


def test_xla_train_step_matches_train_step():
    log("case: xla train step vs tfh_train_step")
    from tensorf_amd import _lib, cameras, render, training, xla_abi
    g, r, n, k = 32, 1024, 64, 64
    cfg = training.TensorfConfig(minibatch_size=r, n_iters=100, initial_aabb_min=(-1.5,) * 3,
                                 initial_aabb_max=(1.5,) * 3, appearance_feat_dim=48,
                                 density_feat_dim=16, feature_n_freqs=2, viewdir_n_freqs=2,
                                 grid_dim_init=g, grid_dim_final=g, upsamp_iters=(),
                                 density_samples_per_ray=n, appearance_samples_per_ray=k)
    a = training.TrainState.initialize(cfg, g, prng_key=6, num_cameras=1)
    b = training.TrainState.initialize(cfg, g, prng_key=6, num_cameras=1)
    assert torch.equal(a.flat.params, b.flat.params)
    o, d = orc.make_rays(r, seed=7, radius=4.0, target_half=0.5)
    colors = dev(f32(np.random.default_rng(8).uniform(size=(r, 3))))
    rays = cameras.Rays3D(dev(f32(o)), dev(f32(d)), torch.zeros(r, dtype=torch.int32,
                                                                device="cuda"))
    # the plain fused step
    env_key = render._seed(a.prng_key, 0)
    a, logs = a.training_step(cameras.RenderedRays(colors, rays))
    torch.cuda.synchronize()
    mse_a = float(logs["mse"].item())
    # the XLA form: same key, same flat buffers, the Adam scalars from the caller
    lp = b.learnable_params
    env = render._RenderEnv(b.appearance_mlp, lp, b.aabb, rays, env_key, b.render_config(),
                            alloc_workspace=False)
    import ctypes as ct
    ws_bytes = int(_lib.lib.tfh_train_step_workspace_size(ct.byref(env.cfg), ct.byref(env.cmlp),
                                                          48, r))
    decay = training.lr_decay_coeff(0, cfg)
    lr_t, lr_m = cfg.optimizer.lr_init_tensor, cfg.optimizer.lr_init_mlp
    f = np.float32
    hyper = np.array([-f(lr_t) * f(decay), -f(lr_m) * f(decay), f(1) - f(0.9) ** f(1),
                      f(1) - f(0.99) ** f(1)], np.float32)
    fl = b.flat
    offs = list(fl.offsets) + [0] * (12 - len(fl.offsets))
    opaque = xla_abi.descriptor(env.cfg, env.dvm, env.avm, env.cmlp, r, ws_bytes,
                                jitter=env.jitter is not None, schedule=env.base_ts is not None,
                                gumbel=False, offsets=offs, total=fl.total,
                                tensor_count=fl.tensor_count)
    keep = []
    noise = _noise(env, keep)
    mse = torch.full((1,), float("nan"), dtype=torch.float32, device="cuda")
    grads = torch.full_like(fl.params, float("nan"))  # the call clears its scratch
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    hyp = dev(hyper)
    params_in = fl.params.clone()
    xla_abi.call("tfh_xla_train_step", _lib.stream(),
                 [_lib.ptr(params_in), _lib.ptr(fl.exp_avg), _lib.ptr(fl.exp_avg_sq),
                  _lib.ptr(env.aabb), _lib.ptr(env.o), _lib.ptr(env.d), _lib.ptr(env.cams),
                  _lib.ptr(colors)] + noise + [_lib.ptr(hyp), _lib.ptr(fl.params),
                                               _lib.ptr(fl.exp_avg), _lib.ptr(fl.exp_avg_sq),
                                               _lib.ptr(mse), _lib.ptr(grads), _lib.ptr(ws)],
                 opaque)
    torch.cuda.synchronize()
    mse_b = float(mse.item())
    log(f"  mse plain={mse_a:.9e} xla={mse_b:.9e}")
    assert abs(mse_a - mse_b) <= 1e-6 * mse_a
    assert torch.count_nonzero(grads).item() == 0  # left zero, like tfh_train_step
    for name, sl, lr in (("factors", slice(0, fl.tensor_count), lr_t),
                         ("mlp", slice(fl.tensor_count, None), lr_m)):
        upd_a = (a.flat.params - params_in)[sl]
        upd_b = (fl.params - params_in)[sl]
        e = (upd_a - upd_b).abs()
        frac = (e > 1e-3 * lr).float().mean().item()
        log(f"  {name}: max|update difference| {e.max().item():.3e}, fraction off by >1e-3 lr "
            f"{frac:.2e}")
        assert frac <= 1e-3, name
    for t_a, t_b in ((a.flat.exp_avg, fl.exp_avg), (a.flat.exp_avg_sq, fl.exp_avg_sq)):
        assert (t_a - t_b).abs().max().item() <= 1e-5 * t_a.abs().max().item() + 1e-12
