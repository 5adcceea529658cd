"""ctypes binding of libtensorf_hip.so (declared in include/tensorf_hip.h).

The library is the only compute path of this package: importing it when the
shared object is missing raises, and every call that returns a non-zero
status raises TfhError with the library's message. Device buffers are torch
tensors on the current HIP device; calls run on torch's current stream.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtensorf_hip.so")

TFH_OK = 0
MODE_RGB, MODE_DIST_MEDIAN, MODE_DIST_MEAN, MODE_DIST_MEDIAN_AS_WRITTEN = 0, 1, 2, 3
PREC_FP32, PREC_BF16 = 0, 1
SAMPLING_SCHEDULE, SAMPLING_EDGES = 0, 1
RENDER_NOALLOC_SCRATCH_BYTES = 128 << 20  # TFH_RENDER_NOALLOC_SCRATCH_BYTES


class TfhError(RuntimeError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python tensorf-jax_amd/build_lib.py` "
        "(there is no CPU fallback)")

# torch first: the library then binds to the HIP runtime torch has loaded (one
# runtime per process; loading the library first can leave two, and the second
# sees no device)
import torch  # noqa: E402,F401

lib = ctypes.CDLL(LIB_PATH)

c_p = ctypes.c_void_p
c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_f = ctypes.c_float
c_sz = ctypes.c_size_t


class Vm(ctypes.Structure):
    _fields_ = [("plane", c_p), ("line", c_p), ("grid_dim", c_i64), ("channels", c_i64)]


class VmGrad(ctypes.Structure):
    _fields_ = [("plane", c_p), ("line", c_p)]


class Mlp(ctypes.Structure):
    _fields_ = [("w0", c_p), ("w1", c_p), ("b1", c_p), ("w2", c_p), ("b2", c_p), ("w3", c_p),
                ("b3", c_p), ("embed", c_p), ("feat_dim", c_i64), ("squash_dim", c_i64),
                ("units", c_i64), ("feature_n_freqs", c_i64), ("viewdir_n_freqs", c_i64),
                ("num_cameras", c_i64)]


class MlpGrad(ctypes.Structure):
    _fields_ = [(n, c_p) for n in ("w0", "w1", "b1", "w2", "b2", "w3", "b3", "embed")]


class RenderConfig(ctypes.Structure):
    _fields_ = [("near_plane", c_f), ("far_plane", c_f), ("mode", c_i32),
                ("density_samples_per_ray", c_i32), ("appearance_samples_per_ray", c_i32),
                ("scene_contraction", c_i32), ("mlp_precision", c_i32), ("sampling", c_i32),
                ("density_shift", c_f)]


class PropMlp(ctypes.Structure):
    _fields_ = [(n, c_p) for n in ("w1", "b1", "w2", "b2", "w3", "b3")] + [("feat_dim", c_i64)]


class PropMlpGrad(ctypes.Structure):
    _fields_ = [(n, c_p) for n in ("w1", "b1", "w2", "b2", "w3", "b3")]


class ProposalConfig(ctypes.Structure):
    _fields_ = [("samples_round0", c_i32), ("samples_round1", c_i32), ("samples_out", c_i32),
                ("scene_contraction", c_i32), ("histogram_pad", c_f), ("loss_weight", c_f),
                ("weight_power", c_f)]


class TrainBuffers(ctypes.Structure):
    _fields_ = [("params", c_p), ("grads", c_p), ("exp_avg", c_p), ("exp_avg_sq", c_p),
                ("total", c_i64), ("tensor_count", c_i64)]


P = ctypes.POINTER
_SIGS = {
    "tfh_abi_version": (c_i32, []),
    "tfh_last_error": (ctypes.c_char_p, []),
    "tfh_device_count": (c_i32, [P(c_i32)]),
    "tfh_vm_to_channel_last": (c_i32, [c_p, c_p, c_i64, c_i64, c_p, c_p, c_p]),
    "tfh_vm_from_channel_last": (c_i32, [c_p, c_p, c_i64, c_i64, c_p, c_p, c_p]),
    "tfh_vm_interpolate": (c_i32, [P(Vm), c_p, c_i64, c_p, c_p]),
    "tfh_vm_interpolate_bwd": (c_i32, [P(Vm), c_p, c_i64, c_p, P(VmGrad), c_p]),
    "tfh_vm_resize": (c_i32, [P(Vm), c_i64, c_p, c_p, c_p]),
    "tfh_sample_bbox": (c_i32, [c_p, c_p, c_i64, c_p, c_i64, c_p, c_p, c_p, c_p, c_p]),
    "tfh_sample_contracted": (c_i32, [c_p, c_p, c_i64, c_i64, c_p, c_p, c_p, c_p, c_p, c_p,
                                      c_p]),
    "tfh_segment_probabilities": (c_i32, [c_p, c_p, c_i64, c_i64, c_p, c_p, c_p]),
    "tfh_segment_probabilities_bwd": (c_i32, [c_p, c_p, c_i64, c_i64, c_p, c_p, c_p, c_p, c_p,
                                              c_p]),
    "tfh_topk_select": (c_i32, [c_p, c_i64, c_i64, c_i64, c_p, c_p, c_p]),
    "tfh_feature_mlp_fwd_workspace_size": (c_sz, [P(Mlp), c_i32]),
    "tfh_feature_mlp_fwd": (c_i32, [P(Mlp), c_p, c_p, c_p, c_i64, c_p, c_i32, c_p, c_sz, c_p]),
    "tfh_feature_mlp_bwd_workspace_size": (c_sz, [P(Mlp), c_i64, c_i32]),
    "tfh_feature_mlp_bwd": (c_i32, [P(Mlp), c_p, c_p, c_p, c_i64, c_p, c_p, P(MlpGrad), c_i32,
                                    c_p, c_sz, c_p]),
    "tfh_render_workspace_size": (c_sz, [P(RenderConfig), P(Mlp), c_i64, c_i64]),
    "tfh_train_step_workspace_size": (c_sz, [P(RenderConfig), P(Mlp), c_i64, c_i64]),
    "tfh_render_fwd": (c_i32, [P(Vm), P(Vm), P(Mlp), P(RenderConfig), c_p, c_p, c_p, c_p, c_i64,
                               c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    "tfh_render_bwd": (c_i32, [P(Vm), P(Vm), P(Mlp), P(RenderConfig), c_p, c_p, c_p, c_p, c_i64,
                               c_p, P(VmGrad), P(VmGrad), P(MlpGrad), c_p, c_sz, c_p]),
    "tfh_render_fwd_noalloc": (c_i32, [P(Vm), P(Vm), P(Mlp), P(RenderConfig), c_p, c_p, c_p,
                                       c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p]),
    "tfh_adam_step": (c_i32, [c_p, c_p, c_p, c_p, c_i64, c_f, c_f, c_f, c_f, c_f, c_i64, c_i32,
                              c_p]),
    "tfh_train_step": (c_i32, [P(Vm), P(Vm), P(Mlp), P(RenderConfig), c_p, c_p, c_p, c_p, c_p,
                               c_i64, c_p, c_p, c_p, c_p, P(VmGrad), P(VmGrad), P(MlpGrad),
                               P(TrainBuffers), c_f, c_f, c_f, c_i64, c_p, c_p, c_sz, c_p]),
    "tfh_prop_mlp_fwd": (c_i32, [P(PropMlp), c_p, c_i64, c_p, c_p, c_p]),
    "tfh_prop_mlp_bwd_workspace_size": (c_sz, [P(PropMlp), c_i64]),
    "tfh_prop_mlp_bwd": (c_i32, [P(PropMlp), c_p, c_i64, c_p, c_p, P(PropMlpGrad), c_p, c_sz,
                                 c_p]),
    "tfh_proposal_resample": (c_i32, [c_p, c_p, c_i64, c_i64, c_i64, c_f, c_f, c_p, c_p]),
    "tfh_interval_loss": (c_i32, [c_p, c_p, c_i64, c_p, c_p, c_i64, c_i64, c_f, c_p, c_p, c_p]),
    "tfh_proposal_workspace_size": (c_sz, [P(ProposalConfig), P(Vm), c_i64]),
    "tfh_proposal_sample": (c_i32, [P(Vm), P(PropMlp), P(ProposalConfig), c_p, c_p, c_p, c_i64,
                                    c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    "tfh_proposal_round": (c_i32, [P(ProposalConfig), P(Vm), c_i64, c_i32, c_p, c_sz, P(c_p),
                                   P(c_p)]),
    "tfh_proposal_loss_bwd": (c_i32, [P(Vm), P(PropMlp), P(ProposalConfig), c_i64, c_p, c_p,
                                      P(VmGrad), P(PropMlpGrad), c_p, c_p, c_sz, c_p]),
    "tfh_train_step_proposal_workspace_size": (c_sz, [P(RenderConfig), P(Mlp), c_i64,
                                                      P(ProposalConfig), P(Vm), c_i64]),
    "tfh_train_step_proposal": (c_i32, [P(Vm), P(Vm), P(Mlp), P(RenderConfig), P(Vm),
                                        P(PropMlp), P(ProposalConfig), c_p, c_p, c_p, c_p, c_p,
                                        c_i64, c_p, c_p, c_p, c_p, P(VmGrad), P(VmGrad),
                                        P(MlpGrad), P(VmGrad), P(PropMlpGrad), P(TrainBuffers),
                                        P(TrainBuffers), c_f, c_f, c_f, c_i64, c_p, c_p, c_p,
                                        c_sz, c_p]),
    "tfh_xla_render_fwd": (None, [c_p, P(c_p), ctypes.c_char_p, c_sz]),
    "tfh_xla_render_bwd": (None, [c_p, P(c_p), ctypes.c_char_p, c_sz]),
    "tfh_xla_train_step": (None, [c_p, P(c_p), ctypes.c_char_p, c_sz]),
    "tfh_uniform": (c_i32, [ctypes.c_uint64, ctypes.c_uint64, c_i64, c_p, c_p]),
    "tfh_blob_scene_render": (c_i32, [c_p, c_i32, c_p, c_p, c_i64, ctypes.c_double,
                                      ctypes.c_double, c_i32, c_i32, c_p, c_p, c_p, c_p]),
    "tfh_profile_enable": (c_i32, [c_i32]),
    "tfh_profile_read": (c_i32, [c_i32, c_p, P(ctypes.c_double), P(c_i64), P(c_i32)]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


def last_error() -> str:
    msg = lib.tfh_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> None:
    if status != TFH_OK:
        raise TfhError(f"{what or 'tfh call'} failed with status {status}: {last_error()}")


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)


# ---- torch helpers -------------------------------------------------------------
def ptr(t) -> Optional[int]:
    """Device pointer of a contiguous CUDA/HIP tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise TfhError("tensor must live on the GPU")
    if not t.is_contiguous():
        raise TfhError("tensor must be contiguous")
    return t.data_ptr()


def stream() -> int:
    import torch
    return torch.cuda.current_stream().cuda_stream


def device_count() -> int:
    n = c_i32(0)
    check(lib.tfh_device_count(ctypes.byref(n)), "tfh_device_count")
    return n.value


def profile_enable(on: bool) -> None:
    check(lib.tfh_profile_enable(1 if on else 0), "tfh_profile_enable")


def profile_read(max_stages: int = 64):
    """{stage name: (total ms, launches)} of the recorded stages."""
    names = ctypes.create_string_buffer(48 * max_stages)
    ms = (ctypes.c_double * max_stages)()
    cnt = (c_i64 * max_stages)()
    n = c_i32(0)
    check(lib.tfh_profile_read(max_stages, ctypes.cast(names, c_p), ms, cnt, ctypes.byref(n)),
          "tfh_profile_read")
    out = {}
    for j in range(n.value):
        name = names.raw[48 * j:48 * (j + 1)].split(b"\0", 1)[0].decode()
        out[name] = (ms[j], cnt[j])
    return out
