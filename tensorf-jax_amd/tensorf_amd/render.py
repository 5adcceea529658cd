"""render.py of the reference (render.py:18-549) over libtensorf_hip.

Randomness is explicit: ``prng_key`` (an int) seeds the library's counter-based
generator (tfh_uniform) for the stratified jitter (render.py:158-160, 375-379)
and, with ``appearance_sampling="gumbel"``, for the Gumbel noise of the
weighted appearance-sample draw (render.py:461-469). The default appearance
selection is the deterministic per-ray top-k by weight (ties to the lower
sample index), which is the reference's draw with zero noise.
"""
from __future__ import annotations

import ctypes
import dataclasses
import enum
import math
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from . import _lib, cameras, networks, proposal, tensor_vm


class RenderMode(enum.Enum):
    """render.py:18-23."""
    RGB = _lib.MODE_RGB
    DIST_MEDIAN = _lib.MODE_DIST_MEDIAN
    DIST_MEAN = _lib.MODE_DIST_MEAN


@dataclasses.dataclass(frozen=True)
class RenderConfig:
    """render.py:26-36, plus the appearance selection rule (see module doc)."""
    near: float
    far: float
    mode: RenderMode
    density_samples_per_ray: int
    appearance_samples_per_ray: int
    appearance_sampling: str = "topk"  # "topk" (deterministic) or "gumbel"
    # sigma = softplus(sum of density features + density_shift): +10 in the
    # reference (render.py:207); TensoRF's original code starts transparent with -10
    density_shift: float = 10.0
    # Learned proposal sampler (no reference counterpart, render.py:126): when
    # set, density_samples_per_ray must equal proposal.samples_out and
    # LearnableParams.proposal_field holds its parameters.
    proposal: Optional[proposal.ProposalConfig] = None
    # DIST_MEDIAN only. False: the first sample whose 1 - E exceeds 0.5 (0 for
    # sample 0, +inf when never reached), what render.py:248-267 selects.
    # True: render.py:248-267 evaluated as written, where 0 * inf in the
    # padded last column makes every ray NaN (opacity above 0.5) or +inf.
    median_as_written: bool = False
    # Contracted scenes only: the un-jittered sample distances. "reference":
    # the reference's schedule (render.py:127-155); "uniform_contracted":
    # spaced uniformly in contracted distance g(t) = t (t <= 1), 2 - 1/t.
    contracted_spacing: str = "reference"


@dataclasses.dataclass
class LearnableParams:
    """render.py:39-46."""
    appearance_mlp_params: Dict
    appearance_tensor: tensor_vm.TensorVM
    density_tensor: tensor_vm.TensorVM
    scene_contraction: bool
    proposal_field: Optional[proposal.ProposalField] = None


@dataclasses.dataclass
class SegmentProbabilities:
    """render.py:282-297."""
    p_exits: torch.Tensor
    p_terminates: torch.Tensor

    def get_batch_axes(self) -> Tuple[int, ...]:
        return tuple(self.p_exits.shape)


# ---- explicit randomness -----------------------------------------------------------
def _seed(prng_key, salt: int) -> int:
    k = int(prng_key) & 0xFFFFFFFFFFFFFFFF
    return (k * 0x9E3779B97F4A7C15 + salt * 0xBF58476D1CE4E5B9 + 1) & 0xFFFFFFFFFFFFFFFF


def uniform(seed: int, n: int, device="cuda", offset: int = 0) -> torch.Tensor:
    """Counter-based U[0,1) floats from the HIP library."""
    out = torch.empty(n, dtype=torch.float32, device=device)
    if n:
        _lib.call("tfh_uniform", seed, offset, n, _lib.ptr(out), _lib.stream())
    return out


def _noise(prng_key, ray_count: int, config: RenderConfig, contraction: bool, device):
    n = config.density_samples_per_ray
    if prng_key is None:
        jitter = None
    elif contraction or config.proposal is not None:
        jitter = uniform(_seed(prng_key, 1), ray_count * n, device).view(ray_count, n)
    else:
        jitter = uniform(_seed(prng_key, 1), n, device)
    gumbel = None
    if config.appearance_sampling == "gumbel" and prng_key is not None:
        u = uniform(_seed(prng_key, 2), n, device).clamp_(1e-12, 1.0 - 1e-7)
        gumbel = -torch.log(-torch.log(u))
    return jitter, gumbel


def contraction_schedule(samples_per_ray: int, near: float, far: float):
    """Un-jittered distances and step sizes of the contracted sampler,
    render.py:127-155: half linear in [near, near+1], half spaced out to far,
    the far half evaluated in float64 and cast to float32 like the reference.
    Host-side table setup (N floats), not per-ray work."""
    n = int(samples_per_ray)
    nc, nf = n // 2, n - n // 2
    s = np.arange(nc, dtype=np.float32) / np.float32(max(nc - 1, 1))
    close = (np.float32(near) * (np.float32(1) - s) + np.float32(near + 1.0) * s).astype(np.float32)
    far_start = near + 1.0 + 1.0 / nc
    k = 10.0
    far_deltas = (1.0 / (1.0 - np.linspace(0.0, 1.0 - 1 / ((far - far_start) / k + 1), nf))
                  - 1.0) * np.linspace(1.0, k, nf)
    ts = np.concatenate([close, (far_start + far_deltas).astype(np.float32)]).astype(np.float32)
    steps = (np.roll(ts, -1) - ts).astype(np.float32)
    steps[-1] = steps[-2]
    return ts, steps


def uniform_contracted_schedule(samples_per_ray: int, near: float, far: float):
    """Un-jittered distances spaced uniformly in contracted distance
    g(t) = t (t <= 1), 2 - 1/t (t > 1) between near and far (evaluated in
    float64, cast to float32); step sizes are the differences with the last
    one repeated, as in the reference schedule (render.py:154-155). Host-side
    table setup (N floats)."""
    n = int(samples_per_ray)
    g = lambda t: t if t <= 1.0 else 2.0 - 1.0 / t
    s = np.linspace(g(near), g(far), n)
    ts = np.where(s <= 1.0, s, 1.0 / (2.0 - np.minimum(s, 2.0 - 1e-300))).astype(np.float32)
    steps = (np.roll(ts, -1) - ts).astype(np.float32)
    steps[-1] = steps[-2] if n > 1 else np.float32(far - near)
    return ts, steps


SCHEDULES = {"reference": contraction_schedule, "uniform_contracted": uniform_contracted_schedule}

_SCHEDULE_CACHE: Dict = {}


def _schedule_tensors(n, near, far, device, spacing="reference"):
    key = (n, float(near), float(far), str(device), spacing)
    if key not in _SCHEDULE_CACHE:
        if spacing not in SCHEDULES:
            raise ValueError(f"unknown contracted spacing {spacing!r}")
        ts, st = SCHEDULES[spacing](n, near, far)
        _SCHEDULE_CACHE[key] = (torch.tensor(ts, device=device), torch.tensor(st, device=device))
    return _SCHEDULE_CACHE[key]


# ---- C structs ------------------------------------------------------------------------
def c_config(config: RenderConfig, contraction: bool, precision: int) -> _lib.RenderConfig:
    sampling = _lib.SAMPLING_EDGES if config.proposal is not None else _lib.SAMPLING_SCHEDULE
    mode = config.mode.value
    if config.mode is RenderMode.DIST_MEDIAN and config.median_as_written:
        mode = _lib.MODE_DIST_MEDIAN_AS_WRITTEN
    return _lib.RenderConfig(float(config.near), float(config.far), mode,
                             int(config.density_samples_per_ray),
                             int(config.appearance_samples_per_ray), int(bool(contraction)),
                             int(precision), sampling, float(config.density_shift))


def proposal_noise(prng_key, ray_count: int, pc: proposal.ProposalConfig, device):
    """(R, n0 + n1) fractions placing the proposal rounds' samples, or None."""
    if prng_key is None:
        return None
    m = pc.jitter_count()
    return uniform(_seed(prng_key, 3), ray_count * m, device).view(ray_count, m)


class _RenderFn(torch.autograd.Function):
    """render_rays forward + hand-written VJP (the custom_vjp of the
    reference's jax.value_and_grad, training.py:153-156)."""

    @staticmethod
    def forward(ctx, env, *leaves):
        out = env.forward()
        ctx.env = env
        return out

    @staticmethod
    def backward(ctx, gout):
        return (None, *ctx.env.backward(gout))


class _RenderEnv:
    """Holds the structs, noise and workspace of one render_rays call."""

    def __init__(self, mlp: networks.FeatureMlp, lp: LearnableParams, aabb, rays, prng_key,
                 config: RenderConfig, workspace: Optional[torch.Tensor] = None,
                 sample_proposal: bool = True, alloc_workspace: bool = True):
        self.mlp, self.lp, self.config = mlp, lp, config
        self.device = rays.origins.device
        self.aabb = aabb.to(device=self.device, dtype=torch.float32).contiguous()
        self.o = rays.origins.to(torch.float32).contiguous()
        self.d = rays.directions.to(torch.float32).contiguous()
        self.cams = rays.camera_indices.to(torch.int32).contiguous()
        self.r = self.o.shape[0]
        self.contraction = bool(lp.scene_contraction)
        self.jitter, self.gumbel = _noise(prng_key, self.r, config, self.contraction, self.device)
        n = config.density_samples_per_ray
        self.base_ts = self.base_steps = None
        self.sampler = None
        if config.proposal is not None:
            pc = config.proposal
            if lp.proposal_field is None:
                raise ValueError("config.proposal is set but learnable_params has no proposal_field")
            if pc.samples_out != n:
                raise ValueError("density_samples_per_ray must equal proposal.samples_out")
            self.prop_base_ts = self.prop_base_steps = None
            if self.contraction:
                self.prop_base_ts, self.prop_base_steps = _schedule_tensors(
                    pc.samples_round0, config.near, config.far, self.device,
                    config.contracted_spacing)
            self.prop_jitter = proposal_noise(prng_key, self.r, pc, self.device)
            if sample_proposal:
                self.sampler = proposal.Sampler(lp.proposal_field, pc, self.aabb, self.o, self.d,
                                                self.prop_jitter, self.prop_base_ts,
                                                self.prop_base_steps, self.contraction)
                self.base_ts = self.sampler.edges
        elif self.contraction:
            self.base_ts, self.base_steps = _schedule_tensors(n, config.near, config.far,
                                                              self.device,
                                                              config.contracted_spacing)
        self.cfg = c_config(config, self.contraction, mlp.precision_code if mlp else 0)
        self.dvm = lp.density_tensor.c_struct()
        self.avm = lp.appearance_tensor.c_struct()
        ca = lp.appearance_tensor.stacked_single_vm.channel_dim()
        self.cmlp = mlp.c_struct(lp.appearance_mlp_params, 3 * ca) if mlp else None
        ws_bytes = _lib.lib.tfh_render_workspace_size(
            ctypes.byref(self.cfg), ctypes.byref(self.cmlp) if self.cmlp else None, ca, self.r)
        if not alloc_workspace:
            self.ws = None  # the caller runs a fused entry point with its own workspace
        elif workspace is not None and workspace.numel() >= int(ws_bytes):
            self.ws = workspace
        else:
            self.ws = torch.empty(max(int(ws_bytes), 1), dtype=torch.uint8, device=self.device)

    def _mlp_ref(self):
        return ctypes.byref(self.cmlp) if self.cmlp else None

    def forward(self) -> torch.Tensor:
        rgb = self.config.mode is RenderMode.RGB
        out = torch.empty((self.r, 3) if rgb else (self.r,), dtype=torch.float32,
                          device=self.device)
        _lib.call("tfh_render_fwd", ctypes.byref(self.dvm), ctypes.byref(self.avm), self._mlp_ref(),
                  ctypes.byref(self.cfg), _lib.ptr(self.aabb), _lib.ptr(self.o), _lib.ptr(self.d),
                  _lib.ptr(self.cams), self.r, _lib.ptr(self.jitter), _lib.ptr(self.base_ts),
                  _lib.ptr(self.base_steps), _lib.ptr(self.gumbel), _lib.ptr(out),
                  _lib.ptr(self.ws), self.ws.numel(), _lib.stream())
        return out

    def leaves(self):
        ds, as_ = self.lp.density_tensor.stacked_single_vm, self.lp.appearance_tensor.stacked_single_vm
        return [ds.plane, ds.line, as_.plane, as_.line,
                *networks.FeatureMlp.flat_leaves(self.lp.appearance_mlp_params)]

    def backward(self, gout):
        gout = gout.contiguous()
        self._gout = gout  # keep the buffer alive across the asynchronous call
        leaves = self.leaves()
        grads = [torch.zeros_like(t) for t in leaves]
        dg = _lib.VmGrad(_lib.ptr(grads[0]), _lib.ptr(grads[1]))
        ag = _lib.VmGrad(_lib.ptr(grads[2]), _lib.ptr(grads[3]))
        mg = _lib.MlpGrad(*[_lib.ptr(t) for t in grads[4:]], *([None] * (12 - len(grads))))
        _lib.call("tfh_render_bwd", ctypes.byref(self.dvm), ctypes.byref(self.avm),
                  self._mlp_ref(), ctypes.byref(self.cfg), _lib.ptr(self.aabb), _lib.ptr(self.o),
                  _lib.ptr(self.d), _lib.ptr(self.cams), self.r, _lib.ptr(gout),
                  ctypes.byref(dg), ctypes.byref(ag), ctypes.byref(mg), _lib.ptr(self.ws),
                  self.ws.numel(), _lib.stream())
        return grads


def render_rays(appearance_mlp: networks.FeatureMlp, learnable_params: LearnableParams,
                aabb: torch.Tensor, rays_wrt_world: cameras.Rays3D, prng_key: Optional[int],
                config: RenderConfig) -> torch.Tensor:
    """render.py:105-279: (R,3) RGB or (R,) distances. Differentiable w.r.t. the
    learnable parameters in RGB mode (hand-written backward in the library)."""
    env = _RenderEnv(appearance_mlp, learnable_params, aabb, rays_wrt_world, prng_key, config)
    if config.mode is RenderMode.RGB and torch.is_grad_enabled() and any(
            t.requires_grad for t in env.leaves()):
        return _RenderFn.apply(env, *env.leaves())
    return env.forward()


def render_rays_batched(appearance_mlp: networks.FeatureMlp, learnable_params: LearnableParams,
                        aabb: torch.Tensor, rays_wrt_world: cameras.Rays3D,
                        prng_key: Optional[int], config: RenderConfig, *,
                        batch_size: int = 4096, use_tqdm: bool = False) -> np.ndarray:
    """render.py:49-102: any leading batch axes, rendered in chunks; returns a
    host array shaped batch_axes (+ (3,) for RGB). Each chunk uses the same
    prng_key, like the reference (render.py:95)."""
    batch_axes = rays_wrt_world.get_batch_axes()
    flat = rays_wrt_world.reshape_flat()
    total = flat.origins.shape[0]
    outs = []
    it = range(math.ceil(total / batch_size))
    if use_tqdm:
        from tqdm.auto import tqdm
        it = tqdm(it)
    with torch.no_grad():
        for i in it:
            chunk = flat.slice(i * batch_size, min(total, (i + 1) * batch_size))
            outs.append(render_rays(appearance_mlp, learnable_params, aabb, chunk, prng_key,
                                    config).cpu().numpy())
    out = np.concatenate(outs, axis=0)
    return out.reshape(tuple(batch_axes) + out.shape[1:])


def compute_segment_probabilities(sigmas: torch.Tensor, step_sizes: torch.Tensor
                                  ) -> SegmentProbabilities:
    """render.py:300-347 (log-space transmittance scan per ray)."""
    shape = sigmas.shape
    n = shape[-1]
    s = sigmas.reshape(-1, n).to(torch.float32).contiguous()
    st = step_sizes.reshape(-1, n).to(torch.float32).contiguous()
    e = torch.empty_like(s)
    w = torch.empty_like(s)
    _lib.call("tfh_segment_probabilities", _lib.ptr(s), _lib.ptr(st), s.shape[0], n, _lib.ptr(e),
              _lib.ptr(w), _lib.stream())
    return SegmentProbabilities(e.view(shape), w.view(shape))


def sample_points_along_ray_within_bbox(rays_wrt_world: cameras.Rays3D, aabb: torch.Tensor,
                                        samples_per_ray: int, prng_key: Optional[int]):
    """render.py:350-390 vmapped over rays (render.py:177-186): points (3, R, N),
    ts (R, N), step sizes (R, N). One jitter row is shared by all rays."""
    flat = rays_wrt_world.reshape_flat()
    r, n = flat.origins.shape[0], samples_per_ray
    dev = flat.origins.device
    jitter = None if prng_key is None else uniform(_seed(prng_key, 1), n, dev)
    box = aabb.to(dev, torch.float32).contiguous()
    pts = torch.empty((3, r, n), dtype=torch.float32, device=dev)
    ts = torch.empty((r, n), dtype=torch.float32, device=dev)
    st = torch.empty((r, n), dtype=torch.float32, device=dev)
    _lib.call("tfh_sample_bbox", _lib.ptr(flat.origins), _lib.ptr(flat.directions), r,
              _lib.ptr(box), n, _lib.ptr(jitter),
              _lib.ptr(pts), _lib.ptr(ts), _lib.ptr(st), _lib.stream())
    return pts, ts, st


def ray_segment_from_bounding_box(rays_wrt_world: cameras.Rays3D, aabb: torch.Tensor,
                                  min_segment_length: float = 1e-3):
    """render.py:399-434 for a batch of rays: (t_min, t_max), each (R,). Uses
    one un-jittered sample: t_min = ts[0] and t_max = t_min + step."""
    assert abs(min_segment_length - 1e-3) < 1e-12, "the library fixes min_segment_length=1e-3"
    _, ts, st = sample_points_along_ray_within_bbox(rays_wrt_world, aabb, 1, None)
    return ts[:, 0], ts[:, 0] + st[:, 0]


def select_appearance_samples(p_terminates: torch.Tensor, k: int,
                              gumbel: Optional[torch.Tensor] = None) -> torch.Tensor:
    """render.py:459-470: (R, k) int32 indices, ascending."""
    r, n = p_terminates.shape
    w = p_terminates.to(torch.float32).contiguous()
    idx = torch.empty((r, k), dtype=torch.int32, device=w.device)
    _lib.call("tfh_topk_select", _lib.ptr(w), r, n, k, _lib.ptr(gumbel), _lib.ptr(idx),
              _lib.stream())
    return idx
