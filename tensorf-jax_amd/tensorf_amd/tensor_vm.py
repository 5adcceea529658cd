"""TensorVM of the reference (tensor_vm.py:14-250) over libtensorf_hip.

The factors are stored channel-last (plane (3, G, G, C), line (3, G, C)) so a
tap of the gather reads C contiguous floats; ``vector`` / ``matrix`` expose
the reference's (3, C, G) / (3, C, G, G) layout as views of the same memory.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Callable, Optional, Union

import torch

from . import _lib


@dataclasses.dataclass
class TensorVMSingle:
    """Three stacked vector-matrix pairs (tensor_vm.py:109-181)."""

    plane: torch.Tensor  # (3, G, G, C)
    line: torch.Tensor   # (3, G, C)

    @property
    def vector(self) -> torch.Tensor:
        """(3, C, G) view in the reference layout (tensor_vm.py:132)."""
        return self.line.permute(0, 2, 1)

    @property
    def matrix(self) -> torch.Tensor:
        """(3, C, G, G) view in the reference layout (tensor_vm.py:133)."""
        return self.plane.permute(0, 3, 1, 2)

    def grid_dim(self) -> int:
        """tensor_vm.py:169-174."""
        g0, g1 = self.plane.shape[1:3]
        assert g0 == g1 == self.line.shape[1]
        return int(g0)

    def channel_dim(self) -> int:
        """tensor_vm.py:176-181."""
        assert self.plane.shape[-1] == self.line.shape[-1]
        return int(self.plane.shape[-1])


def _vm_struct(plane: torch.Tensor, line: torch.Tensor) -> _lib.Vm:
    return _lib.Vm(_lib.ptr(plane), _lib.ptr(line), plane.shape[1], plane.shape[-1])


class _Interpolate(torch.autograd.Function):
    """TensorVM.interpolate with its VJP w.r.t. the factors (the points are
    not differentiated: the reference only differentiates LearnableParams,
    training.py:153-156)."""

    @staticmethod
    def forward(ctx, ijk, plane, line):
        ijk = ijk.contiguous()
        m = ijk.shape[1]
        c = plane.shape[-1]
        out = torch.empty((3 * c, m), dtype=torch.float32, device=plane.device)
        vm = _vm_struct(plane, line)
        _lib.call("tfh_vm_interpolate", ctypes.byref(vm), _lib.ptr(ijk), m, _lib.ptr(out),
                  _lib.stream())
        ctx.save_for_backward(ijk, plane, line)
        return out

    @staticmethod
    def backward(ctx, gout):
        ijk, plane, line = ctx.saved_tensors
        gplane = torch.zeros_like(plane)
        gline = torch.zeros_like(line)
        vm = _vm_struct(plane, line)
        gout = gout.contiguous()
        grad = _lib.VmGrad(_lib.ptr(gplane), _lib.ptr(gline))
        _lib.call("tfh_vm_interpolate_bwd", ctypes.byref(vm), _lib.ptr(ijk), ijk.shape[1],
                  _lib.ptr(gout), ctypes.byref(grad), _lib.stream())
        return None, gplane, gline


def normal_init(stddev: float = 0.1):
    """jax.nn.initializers.normal(stddev) stand-in (training.py:44)."""
    def init(generator: torch.Generator, shape, dtype):
        return torch.randn(shape, generator=generator, dtype=dtype) * stddev
    return init


@dataclasses.dataclass
class TensorVM:
    """A tensor decomposition of three vector-matrix pairs (tensor_vm.py:14-106)."""

    stacked_single_vm: TensorVMSingle

    @staticmethod
    def initialize(grid_dim: int, per_axis_channel_dim: int, init: Optional[Callable] = None,
                   prng_key: Union[int, torch.Generator] = 0, dtype=torch.float32,
                   device="cuda") -> "TensorVM":
        """tensor_vm.py:21-40. ``init(generator, shape, dtype)`` draws in the
        reference layout; the default is N(0, 0.1^2) (training.py:44)."""
        init = init or normal_init(0.1)
        gen = prng_key if isinstance(prng_key, torch.Generator) else \
            torch.Generator().manual_seed(int(prng_key))
        c, g = per_axis_channel_dim, grid_dim
        vector = init(gen, (3, c, g), dtype)
        matrix = init(gen, (3, c, g, g), dtype)
        return TensorVM.from_reference_layout(vector.to(device), matrix.to(device))

    @staticmethod
    def from_reference_layout(vector: torch.Tensor, matrix: torch.Tensor) -> "TensorVM":
        """Builds the channel-last factors from (3,C,G) / (3,C,G,G) arrays."""
        vector = vector.to(torch.float32).contiguous()
        matrix = matrix.to(torch.float32).contiguous()
        _, c, g = vector.shape
        plane = torch.empty((3, g, g, c), dtype=torch.float32, device=matrix.device)
        line = torch.empty((3, g, c), dtype=torch.float32, device=matrix.device)
        _lib.call("tfh_vm_to_channel_last", _lib.ptr(matrix), _lib.ptr(vector), c, g,
                  _lib.ptr(plane), _lib.ptr(line), _lib.stream())
        return TensorVM(TensorVMSingle(plane, line))

    def to_reference_layout(self):
        """(vector (3,C,G), matrix (3,C,G,G)) contiguous copies."""
        s = self.stacked_single_vm
        c, g = s.channel_dim(), s.grid_dim()
        vector = torch.empty((3, c, g), dtype=torch.float32, device=s.plane.device)
        matrix = torch.empty((3, c, g, g), dtype=torch.float32, device=s.plane.device)
        _lib.call("tfh_vm_from_channel_last", _lib.ptr(s.plane), _lib.ptr(s.line), c, g,
                  _lib.ptr(matrix), _lib.ptr(vector), _lib.stream())
        return vector, matrix

    def interpolate(self, ijk: torch.Tensor) -> torch.Tensor:
        """tensor_vm.py:42-89: ijk (3, *) in [-1, 1] -> features (3C, *)."""
        batch = ijk.shape[1:]
        assert ijk.shape[0] == 3
        flat = ijk.reshape(3, -1).to(torch.float32)
        s = self.stacked_single_vm
        out = _Interpolate.apply(flat, s.plane, s.line)
        return out.reshape((3 * s.channel_dim(),) + tuple(batch))

    def resize(self, grid_dim: int) -> "TensorVM":
        """tensor_vm.py:91-100 (align-corners linear upsampling)."""
        s = self.stacked_single_vm
        c = s.channel_dim()
        plane = torch.empty((3, grid_dim, grid_dim, c), dtype=torch.float32, device=s.plane.device)
        line = torch.empty((3, grid_dim, c), dtype=torch.float32, device=s.plane.device)
        vm = _vm_struct(s.plane.detach(), s.line.detach())
        _lib.call("tfh_vm_resize", ctypes.byref(vm), grid_dim, _lib.ptr(plane), _lib.ptr(line),
                  _lib.stream())
        return TensorVM(TensorVMSingle(plane, line))

    def grid_dim(self) -> int:
        return self.stacked_single_vm.grid_dim()

    def channel_dim(self) -> int:
        """3 * per-mode channels (tensor_vm.py:105-106)."""
        return self.stacked_single_vm.channel_dim() * 3

    def c_struct(self) -> _lib.Vm:
        s = self.stacked_single_vm
        return _vm_struct(s.plane, s.line)
