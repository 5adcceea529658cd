"""MI355X-native TensoRF render / train hot path.

Host-side mirror of the reference's operator API (tensorf/tensor_vm.py,
tensorf/render.py, tensorf/networks.py, tensorf/training.py) over the C ABI of
libtensorf_hip.so (include/tensorf_hip.h). Device memory and streams come
from PyTorch-ROCm; every computation on the hot path runs in the HIP library.
"""
from . import _lib  # noqa: F401  (raises ImportError when the HIP library is missing)
from . import cameras, networks, render, tensor_vm, training, utils  # noqa: F401

__all__ = ["cameras", "networks", "render", "tensor_vm", "training", "utils"]
