"""Independent torch float64 restatement of the reference forward, built from
torch library routines (grid_sample for the VM lookup, cumsum, sort, linear)
rather than from the oracle's own index arithmetic. Used only by the tests to
pin the oracle's forward values and, through torch autograd, its hand-derived
backward.
"""
import math

import numpy as np
import torch
import torch.nn.functional as F

MODE_PERMS = ((0, 1, 2), (2, 0, 1), (1, 2, 0))


def vm_interpolate_torch(vector, matrix, ijk):
    """TensorVM.interpolate (tensor_vm.py:42-89) via grid_sample(bilinear, border,
    align_corners=True): vector (3,C,G), matrix (3,C,G,G), ijk (3,M) -> (3C,M)."""
    outs = []
    m_count = ijk.shape[1]
    for m, perm in enumerate(MODE_PERMS):
        c = ijk[list(perm)]
        # line: input (1,C,1,G), sample x = coordinate v.
        gl = torch.stack([c[0], torch.zeros_like(c[0])], -1).view(1, 1, m_count, 2)
        line = F.grid_sample(vector[m][None, :, None, :], gl, mode="bilinear",
                             padding_mode="border", align_corners=True)[0, :, 0, :]
        # plane: rows follow coordinate a, columns coordinate b.
        gp = torch.stack([c[2], c[1]], -1).view(1, 1, m_count, 2)
        plane = F.grid_sample(matrix[m][None], gp, mode="bilinear", padding_mode="border",
                              align_corners=True)[0, :, 0, :]
        outs.append(line * plane)
    return torch.cat(outs, 0)


def fourier_torch(x, n):
    coeffs = 2.0 ** torch.arange(n, dtype=x.dtype)
    inp = x[..., None] * coeffs
    return torch.cat([torch.sin(inp), torch.cos(inp)], -1).reshape(x.shape[0], -1)


def mlp_torch(p, feats, dirs, cams, shape):
    h0 = F.linear(feats, p["w0"].T)
    x = torch.cat([h0, dirs, fourier_torch(h0, shape.feature_n_freqs),
                   fourier_torch(dirs, shape.viewdir_n_freqs)], -1)
    h1 = torch.relu(F.linear(x, p["w1"].T, p["b1"]))
    h2 = torch.relu(F.linear(h1, p["w2"].T, p["b2"]))
    if shape.num_cameras is not None:
        e = p["embed"][cams]
        half = shape.units // 2
        h2 = torch.cat([h2[:, :half], e[:, :half] * h2[:, half:] + e[:, half:]], -1)
    return torch.sigmoid(F.linear(h2, p["w3"].T, p["b3"]))


def segment_torch(sigma, steps):
    a = sigma * steps
    # transmittance as an explicit product of per-segment survival probabilities
    surv = torch.exp(-a)
    t_incl = torch.cumprod(surv, -1)
    t_excl = torch.cat([torch.ones_like(t_incl[:, :1]), t_incl[:, :-1]], -1)
    return t_incl, (1 - surv) * t_excl


def render_rgb_torch(p, xhat, steps, dirs, cams, k, shape):
    """RGB render (render.py:193-246, 437-549) from normalised sample points
    xhat (3,R,N). Selection: weight descending, lower index first."""
    _, r, n = xhat.shape
    dfeat = vm_interpolate_torch(p["density_vector"], p["density_matrix"], xhat.reshape(3, -1))
    sigma = F.softplus(dfeat.sum(0).view(r, n) + 10.0)
    e, w = segment_torch(sigma, steps)
    order = torch.sort(w.detach(), dim=1, descending=True, stable=True).indices[:, :k]
    idx = torch.sort(order, dim=1).values
    rows = torch.arange(r)[:, None]
    sel = xhat[:, rows, idx].reshape(3, -1)
    af = vm_interpolate_torch(p["appearance_vector"], p["appearance_matrix"], sel).T
    rgb = mlp_torch(p, af, dirs.repeat_interleave(k, 0), cams.repeat_interleave(k, 0),
                    shape).view(r, k, 3)
    wsel = w[rows, idx]
    u = (1.0 - e[:, -1] + 1e-8) / (wsel.sum(1) + 1e-8)
    return (rgb * wsel[..., None]).sum(1) * u[:, None] + e[:, -1:], idx


def to_torch(params):
    return {k: torch.tensor(np.asarray(v), dtype=torch.float64, requires_grad=True)
            for k, v in params.items()}


def grid_sample_ok():
    return math.isfinite(1.0)
