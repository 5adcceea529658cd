"""bench.py on the CPU: its single-GPU guard, and the byte models of the
fused MLP forward and the compositing forward at the metric workload
(R=4096 rays x 512 samples, all shaded, VM 300^3 with 16/48 channels)."""
import os
import subprocess
import sys

from tests.conftest import ROOT

import bench

R, N, K, CD, CA, G = 4096, 512, 512, 16, 48, 300
M = R * K
# the appearance factors, read once: 3 planes (G, G, 48) and 3 lines (G, 48)
AFAC = 4 * (3 * G * G * CA + 3 * G * CA)


def test_more_than_one_gpu_is_refused():
    """--gpus 2 exits with status 2 and says why, before any GPU work."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_fused_forward_compulsory_bytes():
    """The gathering forward's unavoidable traffic: 12-byte points in, one
    12-byte view direction per ray, the appearance factors once, the
    576-byte feature row, the kept activations (1 KB FP32 / 512 B BF16) and
    the 12-byte colour out per row."""
    fp32 = bench.compulsory_bytes("mlp_fwd_kernel", R, N, K, CD, CA, G, 1, "fp32")
    bf16 = bench.compulsory_bytes("mlp_fwd_kernel", R, N, K, CD, CA, G, 1, "bf16")
    assert fp32 == M * 12 + R * 12 + AFAC + M * (576 + 1024 + 12)
    assert fp32 - bf16 == M * 512
    # 3.46 GB: at 8 TB/s the kernel could not beat 0.43 ms on bytes alone
    assert 3.4e9 < fp32 < 3.5e9


def test_compositing_compulsory_bytes():
    """The compositing forward reads each shaded sample's weight and colour
    once (16 B, plus a 4-byte index under top-k) and per ray 44 bytes."""
    assert bench.compulsory_bytes("composite_rgb", R, N, K, CD, CA, G, 1) == M * 16 + R * 44
    topk = bench.compulsory_bytes("composite_rgb", R, N, 64, CD, CA, G, 1)
    assert topk == R * 64 * 20 + R * 44
