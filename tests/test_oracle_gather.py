"""Pin the oracle's VM lookup (tensor_vm.py:42-89,140-167,226-250) and its VJP
against scipy.ndimage.map_coordinates, torch grid_sample, closed forms and
finite differences."""
import numpy as np
import pytest
import scipy.ndimage as ndi
import torch

from oracle import tensorf_oracle as orc
from tests.torch_ref import vm_interpolate_torch

RNG = np.random.default_rng(123)


def _grid(c, g):
    return RNG.normal(size=(3, c, g)), RNG.normal(size=(3, c, g, g))


def _pts(m, lo=-1.0, hi=1.0):
    return RNG.uniform(lo, hi, size=(3, m))


@pytest.mark.parametrize("g", [2, 5, 16])
def test_gather_vs_scipy_map_coordinates(g):
    c = 3
    vec, mat = _grid(c, g)
    ijk = _pts(200, -1.3, 1.3)  # includes out-of-range points (clamp-to-edge)
    out = orc.vm_interpolate(vec, mat, ijk)
    for m, perm in enumerate(orc.MODE_PERMS):
        u = (ijk[list(perm)] + 1.0) / 2.0 * (g - 1)
        for ch in range(c):
            line = ndi.map_coordinates(vec[m, ch], u[0:1], order=1, mode="nearest")
            plane = ndi.map_coordinates(mat[m, ch], u[1:3], order=1, mode="nearest")
            np.testing.assert_allclose(out[m * c + ch], line * plane, rtol=1e-12, atol=1e-12)


def test_gather_vs_torch_grid_sample():
    c, g = 4, 9
    vec, mat = _grid(c, g)
    ijk = _pts(300, -1.2, 1.2)
    ref = vm_interpolate_torch(torch.tensor(vec), torch.tensor(mat), torch.tensor(ijk)).numpy()
    np.testing.assert_allclose(orc.vm_interpolate(vec, mat, ijk), ref, rtol=1e-12, atol=1e-12)


def test_gather_special_cases():
    g, c = 7, 2
    # constant grids -> constant output
    vec = np.full((3, c, g), 2.0)
    mat = np.full((3, c, g, g), 3.0)
    out = orc.vm_interpolate(vec, mat, _pts(50, -2, 2))
    np.testing.assert_allclose(out, 6.0, rtol=0, atol=1e-13)
    # linear grid along the interpolated axis is reproduced exactly
    lin = np.arange(g, dtype=np.float64)
    vec = np.broadcast_to(lin, (3, c, g)).copy()
    mat = np.ones((3, c, g, g))
    ijk = _pts(50)
    out = orc.vm_interpolate(vec, mat, ijk)
    for m, perm in enumerate(orc.MODE_PERMS):
        u = (ijk[perm[0]] + 1) / 2 * (g - 1)
        np.testing.assert_allclose(out[m * c], u, atol=1e-12)
    # integer grid coordinates return texel values exactly
    vec, mat = _grid(c, g)
    idx = RNG.integers(0, g, size=(3, 40))
    ijk = idx / (g - 1) * 2.0 - 1.0
    out = orc.vm_interpolate(vec, mat, ijk)
    for m, perm in enumerate(orc.MODE_PERMS):
        v, a, b = idx[perm[0]], idx[perm[1]], idx[perm[2]]
        np.testing.assert_allclose(out[m * c:(m + 1) * c], vec[m][:, v] * mat[m][:, a, b],
                                   atol=1e-12)


def test_gather_vjp_adjoint_and_fd():
    c, g = 3, 6
    vec, mat = _grid(c, g)
    ijk = _pts(40, -1.1, 1.1)
    gout = RNG.normal(size=(3 * c, 40))
    gvec, gmat = orc.vm_interpolate_vjp(vec, mat, ijk, gout)
    # torch autograd of the grid_sample restatement
    tv = torch.tensor(vec, requires_grad=True)
    tm = torch.tensor(mat, requires_grad=True)
    (vm_interpolate_torch(tv, tm, torch.tensor(ijk)) * torch.tensor(gout)).sum().backward()
    np.testing.assert_allclose(gvec, tv.grad.numpy(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(gmat, tm.grad.numpy(), rtol=1e-11, atol=1e-12)
    # central finite differences on a few entries
    f = lambda v, mm: float((orc.vm_interpolate(v, mm, ijk) * gout).sum())
    h = 1e-6
    for _ in range(5):
        i = tuple(RNG.integers(0, s) for s in vec.shape)
        vp, vm_ = vec.copy(), vec.copy()
        vp[i] += h
        vm_[i] -= h
        assert abs((f(vp, mat) - f(vm_, mat)) / (2 * h) - gvec[i]) < 1e-6
        j = tuple(RNG.integers(0, s) for s in mat.shape)
        mp, mm_ = mat.copy(), mat.copy()
        mp[j] += h
        mm_[j] -= h
        assert abs((f(vec, mp) - f(vec, mm_)) / (2 * h) - gmat[j]) < 1e-6


def test_resize_vs_torch_interpolate():
    c, g, g2 = 3, 5, 12
    vec, mat = _grid(c, g)
    nv, nm = orc.vm_resize(vec, mat, g2)
    tv = torch.nn.functional.interpolate(torch.tensor(vec), size=g2, mode="linear",
                                         align_corners=True).numpy()
    tm = torch.nn.functional.interpolate(torch.tensor(mat), size=(g2, g2), mode="bilinear",
                                         align_corners=True).numpy()
    np.testing.assert_allclose(nv, tv, atol=1e-12)
    np.testing.assert_allclose(nm, tm, atol=1e-12)
