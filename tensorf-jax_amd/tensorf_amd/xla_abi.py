"""Descriptors and buffer lists of the XLA custom-call entry points
(tfh_xla_render_fwd / tfh_xla_render_bwd / tfh_xla_train_step in
include/tensorf_hip.h), the form a JAX binding registers with
jax.ffi.register_ffi_target(..., platform="ROCM", api_version=0).

The descriptor is the ASCII backend config XLA hands the call as `opaque`;
`call` invokes an entry point exactly as XLA does (stream, void** buffers,
opaque, length), which is how the tests drive it without JAX.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

from . import _lib


def _f(x: float) -> str:
    return repr(float(x))  # a float32 value printed so that it round-trips


def descriptor(cfg: _lib.RenderConfig, density: _lib.Vm, appearance: _lib.Vm, mlp: _lib.Mlp,
               num_rays: int, workspace_bytes: int, *, jitter: bool, schedule: bool,
               gumbel: bool, offsets: Optional[Sequence[int]] = None,
               total: Optional[int] = None, tensor_count: Optional[int] = None) -> bytes:
    """The opaque descriptor of one call (see include/tensorf_hip.h)."""
    fields = [
        "tfhx=1", f"near={_f(cfg.near_plane)}", f"far={_f(cfg.far_plane)}", f"mode={cfg.mode}",
        f"N={cfg.density_samples_per_ray}", f"k={cfg.appearance_samples_per_ray}",
        f"contraction={cfg.scene_contraction}", f"precision={cfg.mlp_precision}",
        f"sampling={cfg.sampling}", f"shift={_f(cfg.density_shift)}", f"R={int(num_rays)}",
        f"dG={density.grid_dim}", f"dC={density.channels}", f"aG={appearance.grid_dim}",
        f"aC={appearance.channels}", f"squash={mlp.squash_dim}", f"units={mlp.units}",
        f"nf={mlp.feature_n_freqs}", f"nv={mlp.viewdir_n_freqs}", f"ncam={mlp.num_cameras}",
        f"jitter={int(bool(jitter))}", f"schedule={int(bool(schedule))}",
        f"gumbel={int(bool(gumbel))}", f"ws={int(workspace_bytes)}",
    ]
    if offsets is not None:
        assert len(offsets) == 12 and total is not None and tensor_count is not None
        fields += ["offsets=" + ",".join(str(int(o)) for o in offsets), f"total={int(total)}",
                   f"tcount={int(tensor_count)}"]
    return " ".join(fields).encode("ascii")


def call(name: str, stream: int, buffers: Sequence[Optional[int]], opaque: bytes) -> None:
    """Calls an XLA entry point the way XLA's custom-call thunk does."""
    arr = (ctypes.c_void_p * len(buffers))(*[ctypes.c_void_p(b) for b in buffers])
    getattr(_lib.lib, name)(ctypes.c_void_p(stream), arr, opaque, len(opaque))
