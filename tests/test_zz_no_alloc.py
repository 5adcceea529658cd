"""tfh_render_fwd_noalloc (the render_rays forward, render.py:105-279, with its
residuals in the library's fixed device array instead of a workspace) against
tfh_render_fwd: identical outputs bit for bit (the same kernels run on the same
inputs; the forward has no atomics), on two calls in a row where the second
renders fewer rays, in RGB (FP32 and BF16 MLP, top-k and all samples shaded)
and distance modes; a call whose residuals do not fit fails before any device
work."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import tensorf_oracle as orc
from tests.gpu_common import dev, f32, log, mlp_module, oracle_params, to_device_params

pytestmark = pytest.mark.gpu


def _env(r, n, k, precision, mode, contraction=False, g=48, alloc=True):
    from tensorf_amd import cameras, render
    shape = orc.MlpShape()
    p = oracle_params(g, 16, 48, shape, seed=3, std=0.3, bias_std=0.1)
    o, d = orc.make_rays(r, seed=4, radius=3.0 if contraction else 4.0, target_half=0.6)
    den, app, mp = to_device_params(p, shape)
    lp = render.LearnableParams(mp, app, den, contraction)
    rays = cameras.Rays3D(dev(f32(o)), dev(f32(d)), torch.zeros(r, dtype=torch.int32,
                                                               device="cuda"))
    box = 2.0 if contraction else 1.5
    aabb = dev(np.array([[-box] * 3, [box] * 3]))
    cfg = render.RenderConfig(0.05, 200.0, mode, n, k)
    return render._RenderEnv(mlp_module(shape, precision), lp, aabb, rays, 7, cfg,
                             alloc_workspace=alloc)


def _noalloc(env, r, out):
    from tensorf_amd import _lib
    return _lib.lib.tfh_render_fwd_noalloc(
        ctypes.byref(env.dvm), ctypes.byref(env.avm), env._mlp_ref(), ctypes.byref(env.cfg),
        _lib.ptr(env.aabb), _lib.ptr(env.o), _lib.ptr(env.d), _lib.ptr(env.cams), r,
        _lib.ptr(env.jitter), _lib.ptr(env.base_ts), _lib.ptr(env.base_steps),
        _lib.ptr(env.gumbel), _lib.ptr(out), _lib.stream())


@pytest.mark.parametrize("n,k,precision,mode,contraction", [
    (256, 32, "bf16", "rgb", False),
    (128, 128, "fp32", "rgb", False),
    (192, 48, "fp32", "rgb", True),
    (256, 256, "fp32", "median", False),
    (256, 256, "fp32", "mean", False),
])
def test_noalloc_forward_matches_workspace_forward(n, k, precision, mode, contraction):
    from tensorf_amd import _lib, render
    m = {"rgb": render.RenderMode.RGB, "median": render.RenderMode.DIST_MEDIAN,
         "mean": render.RenderMode.DIST_MEAN}[mode]
    for call, r in enumerate((1024, 517)):
        log(f"case: noalloc forward call {call} R={r} N={n} k={k} {precision} {mode} "
            f"contraction={contraction}")
        env = _env(r, n, k, precision, m, contraction)
        want = env.forward()
        shape = (r, 3) if mode == "rgb" else (r,)
        got = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
        st = _noalloc(env, r, got)
        assert st == 0, _lib.last_error()
        torch.cuda.synchronize()
        a, b = want.cpu().numpy(), got.cpu().numpy()
        log(f"  max |difference| {np.nanmax(np.abs(np.where(np.isinf(a), 0, a - b))):.3e}")
        assert np.array_equal(a, b, equal_nan=True)


def test_noalloc_rejects_what_does_not_fit():
    from tensorf_amd import _lib, render
    log("case: noalloc forward, too many rays")
    env = _env(4096, 512, 512, "fp32", render.RenderMode.RGB, alloc=False)
    # 512 shaded samples of 144 features per ray: ~310 KB a ray, so 4096 rays
    # (1.3 GB of residuals) cannot fit the fixed array
    out = torch.empty((4096, 3), dtype=torch.float32, device="cuda")
    st = _noalloc(env, 4096, out)
    assert st == -5, st
    msg = _lib.last_error()
    log(f"  {msg}")
    assert "scratch" in msg and "4096 rays" in msg
