"""Helpers shared by the GPU parity tests: build the same seeded inputs for the
HIP library (float32 on the device) and for the oracle (float64 on the host)."""
import sys

import numpy as np
import torch

from oracle import tensorf_oracle as orc


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def f32(a):
    """Round to float32 and back to float64 so both sides see identical inputs."""
    return np.asarray(a, np.float32).astype(np.float64)


def dev(a, dtype=torch.float32):
    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda").contiguous()


def oracle_params(g, cd, ca, shape, seed=0, std=0.1, bias_std=0.0):
    p = orc.make_params(g, cd, ca, shape, seed=seed, std=std)
    rng = np.random.default_rng(seed + 100)
    for b in ("b1", "b2", "b3"):
        if bias_std:
            p[b] = rng.normal(0, bias_std, p[b].shape)
    return {k: f32(v) for k, v in p.items()}


def to_device_params(p, shape):
    """Oracle parameter dict -> (density TensorVM, appearance TensorVM, mlp params dict)."""
    from tensorf_amd import tensor_vm
    den = tensor_vm.TensorVM.from_reference_layout(dev(p["density_vector"]), dev(p["density_matrix"]))
    app = tensor_vm.TensorVM.from_reference_layout(dev(p["appearance_vector"]),
                                                   dev(p["appearance_matrix"]))
    mp = {"Dense_0": {"kernel": dev(p["w0"])},
          "Dense_1": {"kernel": dev(p["w1"]), "bias": dev(p["b1"])},
          "Dense_2": {"kernel": dev(p["w2"]), "bias": dev(p["b2"])},
          "Dense_3": {"kernel": dev(p["w3"]), "bias": dev(p["b3"])}}
    if shape.num_cameras is not None:
        mp["Embed_0"] = {"embedding": dev(p["embed"])}
    return den, app, {"params": mp}


def mlp_module(shape, precision="fp32"):
    from tensorf_amd import networks
    return networks.FeatureMlp(feature_squash_dim=shape.squash_dim, units=shape.units,
                               feature_n_freqs=shape.feature_n_freqs,
                               viewdir_n_freqs=shape.viewdir_n_freqs,
                               num_cameras=shape.num_cameras, precision=precision)


def report(name, got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    err = np.abs(got - want)
    rel = err / np.maximum(np.abs(want), 1e-30)
    log(f"  {name}: max|err|={err.max():.3e} max rel={rel[np.abs(want) > 1e-6].max() if (np.abs(want) > 1e-6).any() else 0:.3e} "
        f"scale={np.abs(want).max():.3e}")
    return err
