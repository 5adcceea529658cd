"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/tensorf_hip.h declares, the header is valid C, and calls fail
with status codes (never crash) when arguments are bad or no GPU is visible."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tensorf_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tfh_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from tensorf_amd import _lib
    return _lib


def test_header_compiles_as_c():
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib.lib, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert set(names) <= set(lib.EXPORTED), set(names) - set(lib.EXPORTED)


def test_version_and_errors_without_device(lib):
    import torch
    assert lib.lib.tfh_abi_version() == 2
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert lib.device_count() == 0
    # compute entry points report "no device" instead of crashing
    st = lib.lib.tfh_uniform(1, 0, 16, ctypes.c_void_p(1), None)
    assert st == -3 and "no HIP device" in lib.last_error()
    # argument validation happens before any device work
    st = lib.lib.tfh_topk_select(None, 4, 10, 11, None, None, None)
    assert st == -1 and "k 11 outside" in lib.last_error()
    vm = lib.Vm(None, None, 8, 4)
    st = lib.lib.tfh_vm_interpolate(ctypes.byref(vm), None, 3, None, None)
    assert st == -1 and "null plane" in lib.last_error()
    cfg = lib.RenderConfig(0.05, 200.0, 0, 2000, 4, 0, 0, 0)
    assert lib.lib.tfh_render_workspace_size(ctypes.byref(cfg), None, 48, 16) == 0


def test_workspace_sizes_are_monotone(lib):
    cfg = lib.RenderConfig(0.05, 200.0, 1, 512, 512, 0, 0, 0)  # distance mode: no MLP needed
    a = lib.lib.tfh_render_workspace_size(ctypes.byref(cfg), None, 48, 1024)
    b = lib.lib.tfh_render_workspace_size(ctypes.byref(cfg), None, 48, 4096)
    assert 0 < a < b


def test_xla_entry_points_reject_bad_descriptors(lib):
    """The XLA custom calls have no status channel: a bad descriptor is
    reported through tfh_last_error (and stderr) before any buffer is read."""
    from tensorf_amd import xla_abi
    bufs = [None] * 31
    for name in ("tfh_xla_render_fwd", "tfh_xla_render_bwd", "tfh_xla_train_step"):
        xla_abi.call(name, 0, bufs, b"tfhx=2 near=1")
        assert "tfhx=1" in lib.last_error()
        xla_abi.call(name, 0, bufs, b"tfhx=1 near=0.05 far=200")
        assert "lacks 'mode'" in lib.last_error()
        xla_abi.call(name, 0, bufs, b"tfhx=1 junk")
        assert "not key=value" in lib.last_error()
    cfg = lib.RenderConfig(0.05, 200.0, 0, 64, 16, 0, 0, 0, 10.0)
    vm = lib.Vm(None, None, 1, 16)  # grid 1: rejected
    mlp = lib.Mlp(*([None] * 8), 144, 27, 128, 2, 2, 0)
    opaque = xla_abi.descriptor(cfg, vm, vm, mlp, 8, 0, jitter=False, schedule=False,
                                gumbel=False)
    assert opaque.startswith(b"tfhx=1 near=0.05000000074505806 far=200.0 mode=0 N=64 k=16")
    xla_abi.call("tfh_xla_render_fwd", 0, bufs, opaque)
    assert "bad sizes" in lib.last_error()
    xla_abi.call("tfh_xla_train_step", 0, bufs, opaque.replace(b"dG=1", b"dG=8").replace(
        b"aG=1", b"aG=8"))
    assert "needs offsets" in lib.last_error()
