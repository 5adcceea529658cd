"""TrainState of the reference (training.py:19-276) over libtensorf_hip.

All learnable parameters live in ONE flat device buffer (density plane, line,
appearance plane, line, then the MLP leaves), with flat gradient and Adam
moment buffers of the same layout, so a training step is a single C call
(tfh_train_step: render -> MSE -> backward -> Adam on every group). The
reference's functional update is replaced by an in-place one: training_step
returns the same (mutated) object.
"""
from __future__ import annotations

import ctypes
import dataclasses
import json
import math
import pathlib
from typing import Dict, List, Optional, Tuple

import torch

from . import _lib, cameras, networks, proposal, render, tensor_vm, utils


@dataclasses.dataclass(frozen=True)
class OptimizerConfig:
    """train_config.py:6-12."""
    lr_init_tensor: float = 0.02
    lr_init_mlp: float = 1e-3
    lr_decay_iters: Optional[int] = None
    lr_decay_target_ratio: float = 0.1
    lr_upsample_reset: bool = True


@dataclasses.dataclass(frozen=True)
class TensorfConfig:
    """train_config.py:15-71 (dataset/run fields kept for signature parity)."""
    run_dir: Optional[pathlib.Path] = None
    dataset_path: Optional[pathlib.Path] = None
    dataset_type: str = "blender"
    minibatch_size: int = 4096
    n_iters: int = 30000
    optimizer: OptimizerConfig = dataclasses.field(default_factory=OptimizerConfig)
    initial_aabb_min: Tuple[float, float, float] = (-1.0, -1.0, -1.0)
    initial_aabb_max: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    appearance_feat_dim: int = 24
    density_feat_dim: int = 8
    feature_n_freqs: int = 6
    viewdir_n_freqs: int = 6
    grid_dim_init: int = 128
    grid_dim_final: int = 300
    upsamp_iters: Tuple[int, ...] = (2000, 3000, 4000, 5500, 7000)
    scene_contraction: bool = False
    scene_scale: float = 1.0
    camera_embeddings: bool = False
    render_near: float = 0.05
    render_far: float = 200.0
    train_ray_sample_multiplier: float = 1.0
    # Explicit sample counts and MLP arithmetic (the reference derives the counts
    # from the grid, training.py:115-118; None keeps that heuristic).
    density_samples_per_ray: Optional[int] = None
    appearance_samples_per_ray: Optional[int] = None
    mlp_precision: str = "fp32"
    # Learned proposal sampler (no reference counterpart): when set, the field
    # takes proposal.samples_out samples per ray on the proposal's intervals.
    proposal: Optional[proposal.ProposalConfig] = None
    # sigma = softplus(features + density_shift); +10 is the reference's
    # (render.py:207)
    density_shift: float = 10.0
    # contracted scenes: "reference" (render.py:127-155) or "uniform_contracted"
    contracted_spacing: str = "reference"


def lr_decay_coeff(step: int, config: TensorfConfig) -> float:
    """training.py:161-181."""
    opt = config.optimizer
    if opt.lr_upsample_reset:
        s = min(step - u for u in (0,) + tuple(config.upsamp_iters) if step - u >= 0)
    else:
        s = step
    decay_iters = opt.lr_decay_iters if opt.lr_decay_iters is not None else config.n_iters
    v = opt.lr_decay_target_ratio ** (s / decay_iters)
    return max(v, opt.lr_decay_target_ratio)


def sample_counts(config: TensorfConfig, grid_dim: int) -> Tuple[int, int]:
    """training.py:115-118, unless the config pins the counts (or a proposal
    sampler fixes N)."""
    if config.proposal is not None:
        n = config.proposal.samples_out
        k = config.appearance_samples_per_ray or n
        return n, max(1, min(k, n))
    n = config.density_samples_per_ray or int(
        math.sqrt(3 * grid_dim ** 2) * config.train_ray_sample_multiplier)
    k = config.appearance_samples_per_ray or int(0.15 * n)
    return n, max(1, min(k, n))


class FlatParams:
    """One flat buffer holding every learnable parameter, plus views."""

    def __init__(self, shapes: List[Tuple[str, Tuple[int, ...]]], tensor_groups: int, device):
        self.names = [n for n, _ in shapes]
        self.shapes = [s for _, s in shapes]
        sizes = [math.prod(s) for s in self.shapes]
        # 64-float alignment keeps every view 256-byte aligned
        self.offsets = []
        off = 0
        for sz in sizes:
            self.offsets.append(off)
            off += (sz + 63) // 64 * 64
        self.total = off
        self.tensor_count = self.offsets[tensor_groups]
        self.params = torch.zeros(self.total, dtype=torch.float32, device=device)
        self.grads = torch.zeros_like(self.params)
        self.exp_avg = torch.zeros_like(self.params)
        self.exp_avg_sq = torch.zeros_like(self.params)

    def view(self, buf: torch.Tensor, i: int) -> torch.Tensor:
        o = self.offsets[i]
        return buf[o:o + math.prod(self.shapes[i])].view(self.shapes[i])

    def c_struct(self) -> _lib.TrainBuffers:
        return _lib.TrainBuffers(_lib.ptr(self.params), _lib.ptr(self.grads),
                                 _lib.ptr(self.exp_avg), _lib.ptr(self.exp_avg_sq), self.total,
                                 self.tensor_count)


class TrainState:
    """training.py:19-276."""

    def __init__(self, config: TensorfConfig, appearance_mlp: networks.FeatureMlp,
                 flat: FlatParams, aabb: torch.Tensor, prng_key: int, step: int = 0,
                 adam_count: int = 0):
        self.config = config
        self.appearance_mlp = appearance_mlp
        self.flat = flat
        self.aabb = aabb
        self.prng_key = prng_key
        self.step = step
        self.adam_count = adam_count
        self._ws = None
        self._mse = torch.zeros(1, dtype=torch.float32, device=aabb.device)
        self._prop_loss = torch.zeros(1, dtype=torch.float32, device=aabb.device)
        self.prop_flat: Optional[FlatParams] = None

    # ---- construction -----------------------------------------------------------
    @staticmethod
    def initialize(config: TensorfConfig, grid_dim: int, prng_key: int, num_cameras: int,
                   device="cuda") -> "TrainState":
        """training.py:35-99."""
        gen = torch.Generator().manual_seed(int(prng_key))
        mlp = networks.FeatureMlp(feature_n_freqs=config.feature_n_freqs,
                                  viewdir_n_freqs=config.viewdir_n_freqs,
                                  num_cameras=num_cameras if config.camera_embeddings else None,
                                  precision=config.mlp_precision)
        ca, cd = config.appearance_feat_dim, config.density_feat_dim
        mparams = mlp.init(gen, torch.zeros((1, 3 * ca), device=device))
        app = tensor_vm.TensorVM.initialize(grid_dim, ca, prng_key=gen, device=device)
        den = tensor_vm.TensorVM.initialize(grid_dim, cd, prng_key=gen, device=device)
        aabb = torch.tensor([config.initial_aabb_min, config.initial_aabb_max],
                            dtype=torch.float32, device=device)
        state = TrainState(config, mlp, None, aabb, int(prng_key))
        state._load(den, app, mparams)
        if config.proposal is not None:
            state._load_proposal(proposal.ProposalField.initialize(config.proposal, gen, device))
        return state

    def _load_proposal(self, field: proposal.ProposalField):
        tensors = field.leaves()
        names = ["prop_plane", "prop_line", *proposal.MLP_KEYS]
        flat = FlatParams([(n, tuple(t.shape)) for n, t in zip(names, tensors)], 2,
                          self.aabb.device)
        for i, t in enumerate(tensors):
            flat.view(flat.params, i).copy_(t)
        self.prop_flat = flat

    def _load(self, den: tensor_vm.TensorVM, app: tensor_vm.TensorVM, mparams: Dict,
              moments=None):
        leaves = networks.FeatureMlp.flat_leaves(mparams)
        names = ["density_plane", "density_line", "appearance_plane", "appearance_line",
                 "w0", "w1", "b1", "w2", "b2", "w3", "b3", "embed"][:4 + len(leaves)]
        tensors = [den.stacked_single_vm.plane, den.stacked_single_vm.line,
                   app.stacked_single_vm.plane, app.stacked_single_vm.line, *leaves]
        flat = FlatParams([(n, tuple(t.shape)) for n, t in zip(names, tensors)], 4,
                          self.aabb.device)
        for i, t in enumerate(tensors):
            flat.view(flat.params, i).copy_(t)
        if moments is not None:
            for i, (m, v) in enumerate(moments):
                flat.view(flat.exp_avg, i).copy_(m)
                flat.view(flat.exp_avg_sq, i).copy_(v)
        self.flat = flat
        self._ws = None

    # ---- views in the reference's structure -----------------------------------------
    def _vm(self, buf, i) -> tensor_vm.TensorVM:
        return tensor_vm.TensorVM(tensor_vm.TensorVMSingle(self.flat.view(buf, i),
                                                           self.flat.view(buf, i + 1)))

    def _mlp_params(self, buf) -> Dict:
        f = self.flat
        p = {"Dense_0": {"kernel": f.view(buf, 4)},
             "Dense_1": {"kernel": f.view(buf, 5), "bias": f.view(buf, 6)},
             "Dense_2": {"kernel": f.view(buf, 7), "bias": f.view(buf, 8)},
             "Dense_3": {"kernel": f.view(buf, 9), "bias": f.view(buf, 10)}}
        if len(f.names) > 11:
            p["Embed_0"] = {"embedding": f.view(buf, 11)}
        return {"params": p}

    def _proposal_field(self, buf) -> Optional[proposal.ProposalField]:
        f = self.prop_flat
        if f is None:
            return None
        vm = tensor_vm.TensorVM(tensor_vm.TensorVMSingle(f.view(buf, 0), f.view(buf, 1)))
        return proposal.ProposalField(vm, {k: f.view(buf, 2 + i)
                                           for i, k in enumerate(proposal.MLP_KEYS)})

    @property
    def learnable_params(self) -> render.LearnableParams:
        return render.LearnableParams(
            self._mlp_params(self.flat.params), self._vm(self.flat.params, 2),
            self._vm(self.flat.params, 0), self.config.scene_contraction,
            self._proposal_field(self.prop_flat.params) if self.prop_flat else None)

    def render_config(self, mode=render.RenderMode.RGB) -> render.RenderConfig:
        n, k = sample_counts(self.config, self.flat.shapes[0][1])
        return render.RenderConfig(self.config.render_near, self.config.render_far, mode, n, k,
                                   proposal=self.config.proposal,
                                   density_shift=self.config.density_shift,
                                   contracted_spacing=self.config.contracted_spacing)

    # ---- training ------------------------------------------------------------------
    def training_step(self, minibatch: cameras.RenderedRays) -> Tuple["TrainState", Dict]:
        """training.py:101-205 as one library call. Logs hold device scalars."""
        if self.config.proposal is not None:
            return self._training_step_proposal(minibatch)
        rays = minibatch.rays_wrt_world.reshape_flat()
        colors = minibatch.colors.reshape(-1, 3).to(torch.float32).contiguous()
        lp = self.learnable_params
        cfg = self.render_config()
        env = render._RenderEnv(self.appearance_mlp, lp, self.aabb, rays,
                                render._seed(self.prng_key, self.step), cfg, self._ws,
                                alloc_workspace=False)
        # the training step's workspace also keeps the MLP activations between
        # its forward and backward when the shape allows it; reused by later
        # steps of the same shape
        ca = lp.appearance_tensor.stacked_single_vm.channel_dim()
        need = int(_lib.lib.tfh_train_step_workspace_size(
            ctypes.byref(env.cfg), ctypes.byref(env.cmlp), ca, env.r))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(max(need, 1), dtype=torch.uint8, device=env.device)
        grads_lp = (self._vm(self.flat.grads, 0), self._vm(self.flat.grads, 2),
                    self._mlp_params(self.flat.grads))
        dg = _lib.VmGrad(_lib.ptr(grads_lp[0].stacked_single_vm.plane),
                         _lib.ptr(grads_lp[0].stacked_single_vm.line))
        ag = _lib.VmGrad(_lib.ptr(grads_lp[1].stacked_single_vm.plane),
                         _lib.ptr(grads_lp[1].stacked_single_vm.line))
        leaves = networks.FeatureMlp.flat_leaves(grads_lp[2])
        mg = _lib.MlpGrad(*[_lib.ptr(t) for t in leaves], *([None] * (8 - len(leaves))))
        bufs = self.flat.c_struct()
        decay = lr_decay_coeff(self.step, self.config)
        self.adam_count += 1
        _lib.call("tfh_train_step", ctypes.byref(env.dvm), ctypes.byref(env.avm),
                  ctypes.byref(env.cmlp), ctypes.byref(env.cfg), _lib.ptr(env.aabb),
                  _lib.ptr(env.o), _lib.ptr(env.d), _lib.ptr(env.cams), _lib.ptr(colors), env.r,
                  _lib.ptr(env.jitter), _lib.ptr(env.base_ts), _lib.ptr(env.base_steps),
                  _lib.ptr(env.gumbel), ctypes.byref(dg), ctypes.byref(ag), ctypes.byref(mg),
                  ctypes.byref(bufs), self.config.optimizer.lr_init_tensor,
                  self.config.optimizer.lr_init_mlp, decay, self.adam_count, _lib.ptr(self._mse),
                  _lib.ptr(self._ws), self._ws.numel(), _lib.stream())
        self.step += 1
        logs = {"mse": self._mse, "lr_tensor": decay * self.config.optimizer.lr_init_tensor,
                "lr_mlp": decay * self.config.optimizer.lr_init_mlp}
        return self, logs

    def _grad_structs(self):
        grads_lp = (self._vm(self.flat.grads, 0), self._vm(self.flat.grads, 2),
                    self._mlp_params(self.flat.grads))
        dg = _lib.VmGrad(_lib.ptr(grads_lp[0].stacked_single_vm.plane),
                         _lib.ptr(grads_lp[0].stacked_single_vm.line))
        ag = _lib.VmGrad(_lib.ptr(grads_lp[1].stacked_single_vm.plane),
                         _lib.ptr(grads_lp[1].stacked_single_vm.line))
        leaves = networks.FeatureMlp.flat_leaves(grads_lp[2])
        mg = _lib.MlpGrad(*[_lib.ptr(t) for t in leaves], *([None] * (8 - len(leaves))))
        return dg, ag, mg

    def _training_step_proposal(self, minibatch: cameras.RenderedRays):
        """Training step with the proposal sampler as one library call
        (tfh_train_step_proposal): proposal rounds -> render on the proposal's
        intervals -> MSE and interval-loss backward -> Adam on both sets."""
        rays = minibatch.rays_wrt_world.reshape_flat()
        colors = minibatch.colors.reshape(-1, 3).to(torch.float32).contiguous()
        lp = self.learnable_params
        cfg = self.render_config()
        pc = self.config.proposal
        env = render._RenderEnv(self.appearance_mlp, lp, self.aabb, rays, None, cfg,
                                sample_proposal=False, alloc_workspace=False)
        r, n = env.r, cfg.density_samples_per_ray
        seed = render._seed(self.prng_key, self.step)
        jitter = render.uniform(seed, r * (pc.jitter_count() + n), env.device)
        field = lp.proposal_field
        pvm, pmlp = field.c_vm(), field.c_mlp()
        cpc = pc.c_struct(env.contraction, pc.weight_power(self.step))
        ca = lp.appearance_tensor.stacked_single_vm.channel_dim()
        need = int(_lib.lib.tfh_train_step_proposal_workspace_size(
            ctypes.byref(env.cfg), ctypes.byref(env.cmlp), ca, ctypes.byref(cpc),
            ctypes.byref(pvm), r))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=env.device)
        dg, ag, mg = self._grad_structs()
        pgrad = self._proposal_field(self.prop_flat.grads)
        pg_vm = _lib.VmGrad(_lib.ptr(pgrad.tensor.stacked_single_vm.plane),
                            _lib.ptr(pgrad.tensor.stacked_single_vm.line))
        pg_mlp = proposal.mlp_grad_struct(pgrad.mlp)
        bufs = self.flat.c_struct()
        pbufs = self.prop_flat.c_struct()
        decay = lr_decay_coeff(self.step, self.config)
        self.adam_count += 1
        self._keep = (colors, jitter, env)
        _lib.call("tfh_train_step_proposal", ctypes.byref(env.dvm), ctypes.byref(env.avm),
                  ctypes.byref(env.cmlp), ctypes.byref(env.cfg), ctypes.byref(pvm),
                  ctypes.byref(pmlp), ctypes.byref(cpc), _lib.ptr(env.aabb), _lib.ptr(env.o),
                  _lib.ptr(env.d), _lib.ptr(env.cams), _lib.ptr(colors), r, _lib.ptr(jitter),
                  _lib.ptr(env.prop_base_ts), _lib.ptr(env.prop_base_steps), None,
                  ctypes.byref(dg), ctypes.byref(ag), ctypes.byref(mg), ctypes.byref(pg_vm),
                  ctypes.byref(pg_mlp), ctypes.byref(bufs), ctypes.byref(pbufs),
                  self.config.optimizer.lr_init_tensor, self.config.optimizer.lr_init_mlp, decay,
                  self.adam_count, _lib.ptr(self._mse), _lib.ptr(self._prop_loss),
                  _lib.ptr(self._ws), self._ws.numel(), _lib.stream())
        self.step += 1
        logs = {"mse": self._mse, "prop_loss": self._prop_loss,
                "lr_tensor": decay * self.config.optimizer.lr_init_tensor,
                "lr_mlp": decay * self.config.optimizer.lr_init_mlp}
        return self, logs

    def resize_grid(self, new_grid_dim: int) -> "TrainState":
        """training.py:245-276: resample the factors AND the Adam moments."""
        f = self.flat
        out = {}
        for name, buf in (("p", f.params), ("m", f.exp_avg), ("v", f.exp_avg_sq)):
            out[name] = (self._vm(buf, 0).resize(new_grid_dim),
                         self._vm(buf, 2).resize(new_grid_dim))
        mparams = {k: {n: t.clone() for n, t in v.items()}
                   for k, v in self._mlp_params(f.params)["params"].items()}
        mlp_m = [t.clone() for t in networks.FeatureMlp.flat_leaves(self._mlp_params(f.exp_avg))]
        mlp_v = [t.clone() for t in networks.FeatureMlp.flat_leaves(self._mlp_params(f.exp_avg_sq))]
        moments = []
        for j in (0, 1):
            for vm_m, vm_v in ((out["m"][j], out["v"][j]),):
                moments.append((vm_m.stacked_single_vm.plane, vm_v.stacked_single_vm.plane))
                moments.append((vm_m.stacked_single_vm.line, vm_v.stacked_single_vm.line))
        moments += list(zip(mlp_m, mlp_v))
        self._load(out["p"][0], out["p"][1], {"params": mparams}, moments)
        return self

    def psnr(self) -> float:
        return utils.psnr_from_mse(self._mse.item())


def _config_to_json(config: TensorfConfig) -> str:
    d = dataclasses.asdict(config)
    for key in ("run_dir", "dataset_path"):
        if d[key] is not None:
            d[key] = str(d[key])
    return json.dumps(d)


def _config_from_json(text: str) -> TensorfConfig:
    d = json.loads(text)
    d["optimizer"] = OptimizerConfig(**d["optimizer"])
    if d.get("proposal") is not None:
        d["proposal"] = proposal.ProposalConfig(**d["proposal"])
    for key in ("initial_aabb_min", "initial_aabb_max", "upsamp_iters"):
        d[key] = tuple(d[key])
    for key in ("run_dir", "dataset_path"):
        if d[key] is not None:
            d[key] = pathlib.Path(d[key])
    return TensorfConfig(**d)


def save_checkpoint(state: TrainState, path) -> None:
    """Training state as one .npz in the reference's layouts (the fifteen
    checkpoint of training.py:279-368 / render_360.py:66-74 is not available):
    TensorVM factors as (3, C, G) vectors and (3, C, G, G) matrices
    (tensor_vm.py:129-133), Flax-named MLP leaves, Adam moments in the same
    layouts, step counters, aabb and the config as JSON. Loaded by
    load_checkpoint; no pickled objects."""
    import numpy as np
    out = {"config": np.array(_config_to_json(state.config)),
           "step": np.array(state.step), "adam_count": np.array(state.adam_count),
           "prng_key": np.array(state.prng_key), "aabb": state.aabb.cpu().numpy()}
    f = state.flat
    for tag, buf in (("params", f.params), ("exp_avg", f.exp_avg), ("exp_avg_sq", f.exp_avg_sq)):
        for name, i in (("density_tensor", 0), ("appearance_tensor", 2)):
            vec, mat = state._vm(buf, i).to_reference_layout()
            out[f"{tag}/{name}/vector"] = vec.cpu().numpy()
            out[f"{tag}/{name}/matrix"] = mat.cpu().numpy()
        for layer, leaves in state._mlp_params(buf)["params"].items():
            for leaf, t in leaves.items():
                out[f"{tag}/appearance_mlp_params/params/{layer}/{leaf}"] = t.cpu().numpy()
        if state.prop_flat is not None:
            field = state._proposal_field({"params": state.prop_flat.params,
                                           "exp_avg": state.prop_flat.exp_avg,
                                           "exp_avg_sq": state.prop_flat.exp_avg_sq}[tag])
            vec, mat = field.tensor.to_reference_layout()
            out[f"{tag}/proposal_field/vector"] = vec.cpu().numpy()
            out[f"{tag}/proposal_field/matrix"] = mat.cpu().numpy()
            for key, t in field.mlp.items():
                out[f"{tag}/proposal_field/mlp/{key}"] = t.cpu().numpy()
    np.savez(path, **out)


def load_checkpoint(path, device="cuda") -> TrainState:
    """Inverse of save_checkpoint (numpy.load without pickles)."""
    import numpy as np
    z = np.load(path, allow_pickle=False)
    config = _config_from_json(str(z["config"]))

    def t(key):
        return torch.tensor(z[key], device=device)

    mlp_names = (("Dense_0", "kernel"), ("Dense_1", "kernel"), ("Dense_1", "bias"),
                 ("Dense_2", "kernel"), ("Dense_2", "bias"), ("Dense_3", "kernel"),
                 ("Dense_3", "bias"), ("Embed_0", "embedding"))

    def pieces(tag):
        den = tensor_vm.TensorVM.from_reference_layout(t(f"{tag}/density_tensor/vector"),
                                                       t(f"{tag}/density_tensor/matrix"))
        app = tensor_vm.TensorVM.from_reference_layout(t(f"{tag}/appearance_tensor/vector"),
                                                       t(f"{tag}/appearance_tensor/matrix"))
        mp = {}
        for layer, leaf in mlp_names:
            key = f"{tag}/appearance_mlp_params/params/{layer}/{leaf}"
            if key in z:
                mp.setdefault(layer, {})[leaf] = t(key)
        return den, app, {"params": mp}

    num_cameras = z["params/appearance_mlp_params/params/Embed_0/embedding"].shape[0] \
        if "params/appearance_mlp_params/params/Embed_0/embedding" in z else None
    mlp = networks.FeatureMlp(feature_n_freqs=config.feature_n_freqs,
                              viewdir_n_freqs=config.viewdir_n_freqs,
                              num_cameras=num_cameras, precision=config.mlp_precision)
    aabb = torch.tensor(z["aabb"], dtype=torch.float32, device=device)
    state = TrainState(config, mlp, None, aabb, int(z["prng_key"]), int(z["step"]),
                       int(z["adam_count"]))
    den, app, mp = pieces("params")
    mom = []
    for tag in ("exp_avg", "exp_avg_sq"):
        d2, a2, m2 = pieces(tag)
        mom.append([d2.stacked_single_vm.plane, d2.stacked_single_vm.line,
                    a2.stacked_single_vm.plane, a2.stacked_single_vm.line,
                    *networks.FeatureMlp.flat_leaves(m2)])
    state._load(den, app, mp, list(zip(mom[0], mom[1])))
    if "params/proposal_field/vector" in z:
        fields = {}
        for tag in ("params", "exp_avg", "exp_avg_sq"):
            vm = tensor_vm.TensorVM.from_reference_layout(t(f"{tag}/proposal_field/vector"),
                                                          t(f"{tag}/proposal_field/matrix"))
            fields[tag] = proposal.ProposalField(
                vm, {k: t(f"{tag}/proposal_field/mlp/{k}") for k in proposal.MLP_KEYS})
        state._load_proposal(fields["params"])
        for i, (m, v) in enumerate(zip(fields["exp_avg"].leaves(), fields["exp_avg_sq"].leaves())):
            state.prop_flat.view(state.prop_flat.exp_avg, i).copy_(m)
            state.prop_flat.view(state.prop_flat.exp_avg_sq, i).copy_(v)
    return state


def upsampled_grid_dim(config: TensorfConfig, upsamp_index: int) -> int:
    """training.py:360-368: linear interpolation from grid_dim_init to grid_dim_final."""
    return int(config.grid_dim_init + (config.grid_dim_final - config.grid_dim_init)
               * ((upsamp_index + 1) / len(config.upsamp_iters)))


def run_training_loop(config: TensorfConfig, dataset, n_iters: Optional[int] = None,
                      log_every: int = 100, prng_key: int = 94709, log_fn=None) -> TrainState:
    """training.py:279-368 without the experiment harness: cycled shuffled
    minibatches, one training step per iteration, grid upsampling (with the
    Adam moments) at config.upsamp_iters. Logs are device scalars read every
    `log_every` steps (the reference syncs on the step counter every step)."""
    state = TrainState.initialize(config, config.grid_dim_init, prng_key,
                                  num_cameras=dataset.n_views)
    batches = dataset.minibatches(config.minibatch_size, seed=0)
    total = n_iters if n_iters is not None else config.n_iters
    for it in range(total):
        state, logs = state.training_step(next(batches))
        step = state.step
        if log_fn is not None and (step % log_every == 0 or step == total):
            logs["grid_dim"] = state.flat.shapes[0][1]  # the grid this step trained on
            log_fn(step, float(logs["mse"].item()), logs)
        if step in config.upsamp_iters:
            state = state.resize_grid(upsampled_grid_dim(config, config.upsamp_iters.index(step)))
    return state


def evaluate_psnr(state: TrainState, dataset, views, mode=None, chunk: int = 65536) -> float:
    """Mean PSNR of rendered views against the dataset's ground truth
    (render_rays_batched over each view, render.py:49-102)."""
    cfg = state.render_config()
    vals = []
    for v in views:
        rr = dataset.view(v)
        out = render.render_rays_batched(state.appearance_mlp, state.learnable_params, state.aabb,
                                         rr.rays_wrt_world, 0, cfg, batch_size=chunk)
        pred = torch.tensor(out, device=rr.colors.device)
        mse = float(((pred - rr.colors) ** 2).mean())
        vals.append(utils.psnr_from_mse(mse))
    return sum(vals) / len(vals)
