"""Learned proposal sampler over libtensorf_hip.

The reference has no proposal sampler; it stops at a comment (render.py:126:
"An occupancy grid or proposal network would really help us here!"). This
module supplies one: a small VM-decomposed density field (64^3, 8 channels
per mode by default) with a 3C -> 64 -> 64 -> 1 MLP (sigma = softplus),
evaluated in rounds (128 box-segment strata, then 64 resampled intervals) whose
composited weights place the field's samples by inverse-CDF resampling. The
field is trained with mip-NeRF 360's interval loss against the field's own
weights. Everything runs in the HIP library (tfh_proposal_sample,
tfh_proposal_loss_bwd, tfh_train_step_proposal); this module only owns
parameters, noise and workspaces.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Dict, List, Optional

import torch

from . import _lib, tensor_vm

HIDDEN = 64
MLP_KEYS = ("w1", "b1", "w2", "b2", "w3", "b3")


@dataclasses.dataclass(frozen=True)
class ProposalConfig:
    samples_round0: int = 128
    samples_round1: int = 64
    samples_out: int = 64
    # mass added evenly over the bins of each resampled histogram (h = w^p +
    # pad/n; 0.64 = 0.01 per bin of a 64-bin round): a quarter to a third of
    # the samples keep covering the whole ray, so views the proposal was not
    # trained on still find surfaces it under-bounds. On the 8-blob scene
    # (100 views, 10,000 steps, 64 field samples) 0.01 tests at 42.7 dB and
    # 0.64 at 57.6 dB.
    histogram_pad: float = 0.64
    loss_weight: float = 1.0
    grid_dim: int = 64
    channels: int = 8
    init_std: float = 0.3
    # initial output bias: softplus(-3) ~ 0.05 per unit length keeps the
    # untrained proposal nearly transparent, so its first resamplings spread
    # the field's samples along the whole ray (an opaque start concentrates
    # them at the ray entry, where an empty field never produces the weight
    # the interval loss needs to move them)
    init_bias: float = -3.0
    # resampling uses w^power; training ramps power linearly from 0 (even
    # resampling) to 1 over the first anneal_steps steps so that the field sees
    # samples along the whole ray while the proposal is still untrained
    anneal_steps: int = 1000

    def weight_power(self, step: Optional[int]) -> float:
        if step is None or self.anneal_steps <= 0:
            return 1.0
        return min(1.0, step / self.anneal_steps)

    def c_struct(self, contraction: bool, power: float = 1.0) -> _lib.ProposalConfig:
        return _lib.ProposalConfig(int(self.samples_round0), int(self.samples_round1),
                                   int(self.samples_out), int(bool(contraction)),
                                   float(self.histogram_pad), float(self.loss_weight),
                                   float(power))

    def jitter_count(self) -> int:
        return self.samples_round0 + self.samples_round1


def init_mlp(feat_dim: int, generator: torch.Generator, device) -> Dict[str, torch.Tensor]:
    """Kernels (in, out) ~ N(0, 1/fan_in) (lecun normal), zero biases."""
    def normal(shape, fan_in):
        return (torch.randn(shape, generator=generator) / math.sqrt(fan_in)).to(device)
    return {"w1": normal((feat_dim, HIDDEN), feat_dim), "b1": torch.zeros(HIDDEN, device=device),
            "w2": normal((HIDDEN, HIDDEN), HIDDEN), "b2": torch.zeros(HIDDEN, device=device),
            "w3": normal((HIDDEN, 1), HIDDEN), "b3": torch.zeros(1, device=device)}


def mlp_struct(mlp: Dict[str, torch.Tensor]) -> _lib.PropMlp:
    return _lib.PropMlp(*[_lib.ptr(mlp[k]) for k in MLP_KEYS], int(mlp["w1"].shape[0]))


def mlp_grad_struct(grads: Dict[str, torch.Tensor]) -> _lib.PropMlpGrad:
    return _lib.PropMlpGrad(*[_lib.ptr(grads[k]) for k in MLP_KEYS])


@dataclasses.dataclass
class ProposalField:
    """Proposal factors (channel-last, like tensor_vm.TensorVM) and MLP."""
    tensor: tensor_vm.TensorVM
    mlp: Dict[str, torch.Tensor]

    @staticmethod
    def initialize(config: ProposalConfig, prng_key=0, device="cuda") -> "ProposalField":
        gen = prng_key if isinstance(prng_key, torch.Generator) else \
            torch.Generator().manual_seed(int(prng_key))
        vm = tensor_vm.TensorVM.initialize(config.grid_dim, config.channels,
                                           init=tensor_vm.normal_init(config.init_std),
                                           prng_key=gen, device=device)
        mlp = init_mlp(3 * config.channels, gen, device)
        mlp["b3"].fill_(config.init_bias)
        return ProposalField(vm, mlp)

    def leaves(self) -> List[torch.Tensor]:
        s = self.tensor.stacked_single_vm
        return [s.plane, s.line] + [self.mlp[k] for k in MLP_KEYS]

    def c_vm(self) -> _lib.Vm:
        return self.tensor.c_struct()

    def c_mlp(self) -> _lib.PropMlp:
        return mlp_struct(self.mlp)


def workspace_bytes(config: ProposalConfig, field: ProposalField, num_rays: int,
                    contraction: bool = False) -> int:
    pc = config.c_struct(contraction)
    vm = field.c_vm()
    return int(_lib.lib.tfh_proposal_workspace_size(ctypes.byref(pc), ctypes.byref(vm), num_rays))


class Sampler:
    """One tfh_proposal_sample call's structs and workspace; keeps them for
    tfh_proposal_loss_bwd and tfh_proposal_round."""

    def __init__(self, field: ProposalField, config: ProposalConfig, aabb, origins, directions,
                 jitter=None, base_ts=None, base_steps=None, contraction: bool = False,
                 workspace: Optional[torch.Tensor] = None, power: float = 1.0):
        self.field, self.config = field, config
        self.r = origins.shape[0]
        self.pc = config.c_struct(contraction, power)
        self.vm = field.c_vm()
        self.cmlp = field.c_mlp()
        self.keep = (aabb, origins, directions, jitter, base_ts, base_steps)
        need = workspace_bytes(config, field, self.r, contraction)
        if workspace is not None and workspace.numel() >= need:
            self.ws = workspace
        else:
            self.ws = torch.empty(max(need, 1), dtype=torch.uint8, device=origins.device)
        self.edges = torch.empty((self.r, config.samples_out + 1), dtype=torch.float32,
                                 device=origins.device)
        if self.r:
            _lib.call("tfh_proposal_sample", ctypes.byref(self.vm), ctypes.byref(self.cmlp),
                      ctypes.byref(self.pc), _lib.ptr(aabb), _lib.ptr(origins),
                      _lib.ptr(directions), self.r, _lib.ptr(jitter), _lib.ptr(base_ts),
                      _lib.ptr(base_steps), _lib.ptr(self.edges), _lib.ptr(self.ws),
                      self.ws.numel(), _lib.stream())

    def round(self, k: int):
        """(edges (R, n_k + 1), weights (R, n_k)) of round k, as views of the workspace."""
        e, w = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.call("tfh_proposal_round", ctypes.byref(self.pc), ctypes.byref(self.vm), self.r, k,
                  _lib.ptr(self.ws), self.ws.numel(), ctypes.byref(e), ctypes.byref(w))
        n = self.config.samples_round0 if k == 0 else self.config.samples_round1
        base = self.ws.data_ptr()
        edges = self.ws[e.value - base:e.value - base + 4 * self.r * (n + 1)]
        weights = self.ws[w.value - base:w.value - base + 4 * self.r * n]
        return edges.view(torch.float32).view(self.r, n + 1), \
            weights.view(torch.float32).view(self.r, n)

    def loss_backward(self, field_edges: torch.Tensor, field_weights: torch.Tensor,
                      grads: ProposalField) -> torch.Tensor:
        """Interval loss against the field's weights; accumulates into grads
        (a ProposalField of gradient buffers). Returns the loss (device scalar)."""
        loss = torch.zeros(1, dtype=torch.float32, device=self.edges.device)
        gs = grads.tensor.stacked_single_vm
        vg = _lib.VmGrad(_lib.ptr(gs.plane), _lib.ptr(gs.line))
        mg = mlp_grad_struct(grads.mlp)
        fe = field_edges.contiguous()
        fw = field_weights.contiguous()
        self._keep_bwd = (fe, fw)
        _lib.call("tfh_proposal_loss_bwd", ctypes.byref(self.vm), ctypes.byref(self.cmlp),
                  ctypes.byref(self.pc), self.r, _lib.ptr(fe), _lib.ptr(fw), ctypes.byref(vg),
                  ctypes.byref(mg), _lib.ptr(loss), _lib.ptr(self.ws), self.ws.numel(),
                  _lib.stream())
        return loss


def resample(edges: torch.Tensor, weights: torch.Tensor, num_out: int, pad: float,
             power: float = 1.0) -> torch.Tensor:
    """Inverse-CDF resampling of per-ray histograms: (R, n+1), (R, n) -> (R, num_out+1)."""
    r, n = weights.shape
    e = edges.to(torch.float32).contiguous()
    w = weights.to(torch.float32).contiguous()
    out = torch.empty((r, num_out + 1), dtype=torch.float32, device=e.device)
    _lib.call("tfh_proposal_resample", _lib.ptr(e), _lib.ptr(w), r, n, num_out, float(pad),
              float(power), _lib.ptr(out), _lib.stream())
    return out


def interval_loss(edges, weights, prop_edges, prop_weights, scale: float = 1.0):
    """Per-ray interval loss and its gradient w.r.t. prop_weights."""
    r, n = weights.shape
    npr = prop_weights.shape[1]
    args = [t.to(torch.float32).contiguous() for t in (edges, weights, prop_edges, prop_weights)]
    loss = torch.empty(r, dtype=torch.float32, device=args[0].device)
    grad = torch.empty((r, npr), dtype=torch.float32, device=args[0].device)
    _lib.call("tfh_interval_loss", _lib.ptr(args[0]), _lib.ptr(args[1]), n, _lib.ptr(args[2]),
              _lib.ptr(args[3]), npr, r, float(scale), _lib.ptr(loss), _lib.ptr(grad),
              _lib.stream())
    return loss, grad


def mlp_forward(mlp: Dict[str, torch.Tensor], features: torch.Tensor):
    """The proposal MLP alone: (M, F) -> (z (M,), sigma (M,))."""
    f = features.to(torch.float32).contiguous()
    m = f.shape[0]
    z = torch.empty(m, dtype=torch.float32, device=f.device)
    sig = torch.empty_like(z)
    cm = mlp_struct(mlp)
    _lib.call("tfh_prop_mlp_fwd", ctypes.byref(cm), _lib.ptr(f), m, _lib.ptr(z), _lib.ptr(sig),
              _lib.stream())
    return z, sig


def mlp_backward(mlp: Dict[str, torch.Tensor], features: torch.Tensor, grad_z: torch.Tensor):
    """VJP of mlp_forward w.r.t. z: (grads dict, grad_features (M, F))."""
    f = features.to(torch.float32).contiguous()
    gz = grad_z.to(torch.float32).contiguous()
    m = f.shape[0]
    cm = mlp_struct(mlp)
    grads = {k: torch.zeros_like(v) for k, v in mlp.items()}
    gf = torch.empty_like(f)
    need = int(_lib.lib.tfh_prop_mlp_bwd_workspace_size(ctypes.byref(cm), m))
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device=f.device)
    gm = mlp_grad_struct(grads)
    _lib.call("tfh_prop_mlp_bwd", ctypes.byref(cm), _lib.ptr(f), m, _lib.ptr(gz), _lib.ptr(gf),
              ctypes.byref(gm), _lib.ptr(ws), ws.numel(), _lib.stream())
    torch.cuda.current_stream().synchronize()
    return grads, gf
