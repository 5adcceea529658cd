"""FeatureMlp of the reference (networks.py:38-121) over libtensorf_hip.

Parameters use the reference's (Flax) structure:
``{'params': {'Dense_0': {'kernel': (F, S)}, 'Dense_1': {'kernel', 'bias'}, ...,
'Embed_0': {'embedding': (num_cameras, units)}}}`` with (in, out) kernels.
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, Optional

import torch

from . import _lib

_DENSE = ("Dense_0", "Dense_1", "Dense_2", "Dense_3")


def _truncated_normal(gen, shape, std):
    """variance_scaling(..., 'truncated_normal') draw: N(0,1) truncated to
    [-2, 2], rescaled so the variance is std^2 (networks.py:9-10)."""
    x = torch.randn(shape, generator=gen, dtype=torch.float64)
    bad = x.abs() > 2.0
    while bool(bad.any()):
        x[bad] = torch.randn(int(bad.sum()), generator=gen, dtype=torch.float64)
        bad = x.abs() > 2.0
    return (x * (std / 0.87962566103423978)).to(torch.float32)


class FeatureMlp:
    """networks.FeatureMlp (networks.py:38-121).

    ``precision`` selects the MFMA arithmetic: "fp32" or "bf16" (BF16 operands,
    FP32 accumulation)."""

    def __init__(self, feature_squash_dim: int = 27, units: int = 128, feature_n_freqs: int = 6,
                 viewdir_n_freqs: int = 6, num_cameras: Optional[int] = None,
                 precision: str = "fp32"):
        self.feature_squash_dim = feature_squash_dim
        self.units = units
        self.feature_n_freqs = feature_n_freqs
        self.viewdir_n_freqs = viewdir_n_freqs
        self.num_cameras = num_cameras
        self.precision = precision

    @property
    def precision_code(self) -> int:
        return {"fp32": _lib.PREC_FP32, "bf16": _lib.PREC_BF16}[self.precision]

    def encoded_dim(self) -> int:
        """networks.py:77-82."""
        s = self.feature_squash_dim
        return s + 3 + 2 * self.feature_n_freqs * s + 2 * self.viewdir_n_freqs * 3

    def init(self, prng_key, features: torch.Tensor, viewdirs=None, camera_indices=None) -> Dict:
        """Parameters shaped like Flax's init of networks.py:46-121 (kaiming /
        lecun truncated normal kernels, zero biases, Flax's default embedding
        scale). Draws are torch's, not Flax's."""
        gen = prng_key if isinstance(prng_key, torch.Generator) else \
            torch.Generator().manual_seed(int(prng_key))
        f = features.shape[-1]
        s, u, e = self.feature_squash_dim, self.units, self.encoded_dim()
        dev = features.device
        p = {
            "Dense_0": {"kernel": _truncated_normal(gen, (f, s), math.sqrt(1.0 / f))},
            "Dense_1": {"kernel": _truncated_normal(gen, (e, u), math.sqrt(2.0 / e)),
                        "bias": torch.zeros(u)},
            "Dense_2": {"kernel": _truncated_normal(gen, (u, u), math.sqrt(2.0 / u)),
                        "bias": torch.zeros(u)},
            "Dense_3": {"kernel": _truncated_normal(gen, (u, 3), math.sqrt(1.0 / u)),
                        "bias": torch.zeros(3)},
        }
        if self.num_cameras is not None:
            p["Embed_0"] = {"embedding": torch.randn((self.num_cameras, u), generator=gen) /
                            math.sqrt(u)}
        return {"params": {k: {n: t.to(dev).contiguous() for n, t in v.items()}
                           for k, v in p.items()}}

    # ---- C structs -----------------------------------------------------------
    def c_struct(self, params: Dict, feat_dim: int) -> _lib.Mlp:
        p = params["params"]
        emb = p.get("Embed_0", {}).get("embedding") if self.num_cameras is not None else None
        return _lib.Mlp(
            _lib.ptr(p["Dense_0"]["kernel"]), _lib.ptr(p["Dense_1"]["kernel"]),
            _lib.ptr(p["Dense_1"]["bias"]), _lib.ptr(p["Dense_2"]["kernel"]),
            _lib.ptr(p["Dense_2"]["bias"]), _lib.ptr(p["Dense_3"]["kernel"]),
            _lib.ptr(p["Dense_3"]["bias"]), _lib.ptr(emb), feat_dim, self.feature_squash_dim,
            self.units, self.feature_n_freqs, self.viewdir_n_freqs,
            self.num_cameras if emb is not None else 0)

    @staticmethod
    def flat_leaves(params: Dict):
        """Leaves in the order (w0, w1, b1, w2, b2, w3, b3[, embed])."""
        p = params["params"]
        out = [p["Dense_0"]["kernel"], p["Dense_1"]["kernel"], p["Dense_1"]["bias"],
               p["Dense_2"]["kernel"], p["Dense_2"]["bias"], p["Dense_3"]["kernel"],
               p["Dense_3"]["bias"]]
        if "Embed_0" in p:
            out.append(p["Embed_0"]["embedding"])
        return out

    def apply(self, params: Dict, features: torch.Tensor, viewdirs: torch.Tensor,
              camera_indices: Optional[torch.Tensor] = None) -> torch.Tensor:
        """networks.py:46-121: features (*, F), viewdirs (*, 3) -> rgb (*, 3)."""
        batch = features.shape[:-1]
        f = features.reshape(-1, features.shape[-1]).to(torch.float32).contiguous()
        d = viewdirs.reshape(-1, 3).to(torch.float32).contiguous()
        cams = None
        if camera_indices is not None:
            cams = camera_indices.reshape(-1).to(torch.int32).contiguous()
        leaves = self.flat_leaves(params)
        out = _MlpFn.apply(self, params, f, d, cams, *leaves)
        return out.reshape(tuple(batch) + (3,))

    __call__ = apply


class _MlpFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mlp, params, f, d, cams, *leaves):
        m = f.shape[0]
        rgb = torch.empty((m, 3), dtype=torch.float32, device=f.device)
        st = mlp.c_struct(params, f.shape[1])
        ws_bytes = int(_lib.lib.tfh_feature_mlp_fwd_workspace_size(ctypes.byref(st),
                                                                   mlp.precision_code))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=f.device)
        _lib.call("tfh_feature_mlp_fwd", ctypes.byref(st), _lib.ptr(f), _lib.ptr(d),
                  _lib.ptr(cams), m, _lib.ptr(rgb), mlp.precision_code, _lib.ptr(ws), ws.numel(),
                  _lib.stream())
        ctx.mlp, ctx.params = mlp, params
        ctx.save_for_backward(f, d, cams if cams is not None else torch.empty(0))
        ctx.has_cams = cams is not None
        return rgb

    @staticmethod
    def backward(ctx, grgb):
        f, d, cams = ctx.saved_tensors
        cams = cams if ctx.has_cams else None
        mlp, params = ctx.mlp, ctx.params
        leaves = FeatureMlp.flat_leaves(params)
        grads = [torch.zeros_like(t) for t in leaves]
        gf = torch.empty_like(f)
        st = mlp.c_struct(params, f.shape[1])
        g = _lib.MlpGrad(*[_lib.ptr(t) for t in grads], *([None] * (8 - len(grads))))
        m = f.shape[0]
        ws_bytes = _lib.lib.tfh_feature_mlp_bwd_workspace_size(ctypes.byref(st), m,
                                                               mlp.precision_code)
        ws = torch.empty(max(int(ws_bytes), 1), dtype=torch.uint8, device=f.device)
        grgb = grgb.contiguous()  # keep the buffer alive across the asynchronous call
        _lib.call("tfh_feature_mlp_bwd", ctypes.byref(st), _lib.ptr(f), _lib.ptr(d),
                  _lib.ptr(cams), m, _lib.ptr(grgb), _lib.ptr(gf), ctypes.byref(g),
                  mlp.precision_code, _lib.ptr(ws), ws.numel(), _lib.stream())
        return (None, None, gf, None, None, *grads)
