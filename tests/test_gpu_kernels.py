"""GPU parity of the individual C-ABI entry points against the oracle.

Tolerances (FP32 kernels vs the float64 oracle on identical float32 inputs):
gather / sampling / compositing 1e-5 relative (+ a small absolute floor for
values near zero), top-k indices bit-exact on identical weights, FP32 MLP
2e-5 absolute on sigmoid outputs.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import tensorf_oracle as orc
from tests.gpu_common import dev, f32, log, mlp_module, oracle_params, report, to_device_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from tensorf_amd import _lib
    return _lib


def test_layout_roundtrip(lib):
    log("case: layout roundtrip C=16 G=33")
    from tensorf_amd import tensor_vm
    rng = np.random.default_rng(0)
    vec = rng.normal(size=(3, 16, 33)).astype(np.float32)
    mat = rng.normal(size=(3, 16, 33, 33)).astype(np.float32)
    vm = tensor_vm.TensorVM.from_reference_layout(dev(vec), dev(mat))
    s = vm.stacked_single_vm
    assert np.array_equal(s.vector.cpu().numpy(), vec)
    assert np.array_equal(s.matrix.cpu().numpy(), mat)
    v2, m2 = vm.to_reference_layout()
    assert np.array_equal(v2.cpu().numpy(), vec) and np.array_equal(m2.cpu().numpy(), mat)


@pytest.mark.parametrize("c,g", [(16, 32), (48, 17), (8, 64), (5, 9)])
def test_vm_interpolate_fwd_bwd(lib, c, g):
    log(f"case: vm_interpolate C={c} G={g}")
    from tensorf_amd import tensor_vm
    rng = np.random.default_rng(c * 100 + g)
    vec = f32(rng.normal(size=(3, c, g)))
    mat = f32(rng.normal(size=(3, c, g, g)))
    m = 3000
    ijk = f32(rng.uniform(-1.15, 1.15, size=(3, m)))
    ijk[:, :8] = f32(np.array([[-1, 1, 0, 1, -1, 0.5, -0.25, 1]] * 3))  # grid edges exactly
    vm = tensor_vm.TensorVM.from_reference_layout(dev(vec), dev(mat))
    s = vm.stacked_single_vm
    plane = s.plane.clone().requires_grad_(True)
    line = s.line.clone().requires_grad_(True)
    vm2 = tensor_vm.TensorVM(tensor_vm.TensorVMSingle(plane, line))
    out = vm2.interpolate(dev(ijk))
    # FP32 oracle: same FP32 coordinate rounding as the reference's FP32 path
    want32 = orc.vm_interpolate(vec, mat, ijk, dtype=np.float32).astype(np.float64)
    want = orc.vm_interpolate(vec, mat, ijk)
    scale = np.abs(want).max()
    got = out.detach().cpu().numpy()
    err32 = report("features vs fp32 oracle", got, want32)
    err = report("features vs fp64 oracle", got, want)
    # a 1-ulp difference of u (~G * 6e-8 cells) moves a value by up to ~G * 6e-8 * scale
    assert np.all(err32 <= 1e-5 * np.abs(want32) + 1e-6 * max(1.0, g / 16) * scale)
    # vs float64 the FP32 texel coordinate u=(x+1)/2*(G-1) carries ~G*eps cells of error
    assert err.max() <= 4e-7 * max(g, 8) * scale
    gout = f32(rng.normal(size=want.shape))
    (out * dev(gout)).sum().backward()
    gvec, gmat = orc.vm_interpolate_vjp(vec, mat, ijk, gout, dtype=np.float32)
    got_vec = line.grad.permute(0, 2, 1).cpu().numpy()
    got_mat = plane.grad.permute(0, 3, 1, 2).cpu().numpy()
    e1 = report("grad vector", got_vec, gvec)
    e2 = report("grad matrix", got_mat, gmat)
    assert np.all(e1 <= 2e-5 * np.abs(gvec).max() + 1e-5 * np.abs(gvec))
    assert np.all(e2 <= 2e-5 * np.abs(gmat).max() + 1e-5 * np.abs(gmat))


def test_resize(lib):
    log("case: resize 7 -> 19")
    from tensorf_amd import tensor_vm
    rng = np.random.default_rng(5)
    vec = f32(rng.normal(size=(3, 6, 7)))
    mat = f32(rng.normal(size=(3, 6, 7, 7)))
    vm = tensor_vm.TensorVM.from_reference_layout(dev(vec), dev(mat)).resize(19)
    nv, nm = orc.vm_resize(vec, mat, 19)
    v2, m2 = vm.to_reference_layout()
    report("resized vector", v2.cpu().numpy(), nv)
    np.testing.assert_allclose(v2.cpu().numpy(), nv, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(m2.cpu().numpy(), nm, rtol=1e-5, atol=1e-6)


def test_sample_bbox(lib):
    log("case: sample_bbox R=500 N=64")
    from tensorf_amd import cameras, render
    o, d = orc.make_rays(500, seed=3, radius=4.0, target_half=2.0)  # includes misses
    o, d = f32(o), f32(d)
    box = f32(np.array([[-1.5, -1.2, -1.0], [1.5, 1.3, 1.1]]))
    n = 64
    rays = cameras.Rays3D(dev(o), dev(d), torch.zeros(500, dtype=torch.int32, device="cuda"))
    pts, ts, st = render.sample_points_along_ray_within_bbox(rays, dev(box), n, prng_key=7)
    jit = render.uniform(render._seed(7, 1), n).cpu().numpy().astype(np.float64)
    wp, wt, ws = orc.sample_points_along_ray_within_bbox(o, d, box, n, jit)
    report("ts", ts.cpu().numpy(), wt)
    np.testing.assert_allclose(ts.cpu().numpy(), wt, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(st.cpu().numpy(), ws, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(pts.cpu().numpy(), wp, rtol=1e-5, atol=1e-5)


def test_sample_contracted(lib):
    log("case: sample_contracted R=300 N=64")
    o, d = orc.make_rays(300, seed=4, radius=3.0)
    o, d = f32(o), f32(d)
    n = 64
    rng = np.random.default_rng(9)
    jit = f32(rng.uniform(size=(300, n)))
    base, steps = orc.contraction_schedule(n, 0.05, 200.0)
    pts = torch.empty((3, 300, n), device="cuda")
    ts = torch.empty((300, n), device="cuda")
    st = torch.empty((300, n), device="cuda")
    # every input stays referenced until the asynchronous call has completed
    inputs = [dev(o), dev(d), dev(base), dev(steps), dev(jit)]
    lib.call("tfh_sample_contracted", lib.ptr(inputs[0]), lib.ptr(inputs[1]), 300, n,
             lib.ptr(inputs[2]), lib.ptr(inputs[3]), lib.ptr(inputs[4]), lib.ptr(pts),
             lib.ptr(ts), lib.ptr(st), lib.stream())
    torch.cuda.synchronize()
    wp, wt, ws = orc.sample_contracted(o, d, n, 0.05, 200.0, jit)
    report("contracted points", pts.cpu().numpy(), wp)
    np.testing.assert_allclose(ts.cpu().numpy(), wt, rtol=1e-5)
    np.testing.assert_allclose(pts.cpu().numpy(), wp, rtol=1e-5, atol=2e-6)
    assert np.all(np.abs(pts.cpu().numpy()) <= 2.0)


@pytest.mark.parametrize("n", [1, 37, 64, 512, 1000])
def test_segment_probabilities(lib, n):
    log(f"case: segment_probabilities N={n}")
    from tensorf_amd import render
    rng = np.random.default_rng(n)
    sig = f32(rng.uniform(0, 40, size=(50, n)))
    st = f32(rng.uniform(0.001, 0.02, size=(50, n)))
    out = render.compute_segment_probabilities(dev(sig), dev(st))
    e, w = orc.compute_segment_probabilities(sig, st)
    ee = report("p_exits", out.p_exits.cpu().numpy(), e)
    ew = report("p_terminates", out.p_terminates.cpu().numpy(), w)
    # relative error of E equals the absolute error of the FP32 prefix sum
    assert np.all(ee <= 3e-5 * e + 1e-7)
    assert np.all(ew <= 3e-5 * w + 1e-7)


@pytest.mark.parametrize("n,k", [(64, 16), (512, 64), (100, 100), (1000, 1), (37, 36)])
def test_topk_bit_exact(lib, n, k):
    log(f"case: topk N={n} k={k}")
    from tensorf_amd import render
    rng = np.random.default_rng(n + k)
    w = rng.uniform(size=(200, n)).astype(np.float32)
    w[:, ::3] = np.round(w[:, ::3], 1)  # ties
    w[:20] = 0.25                        # a row of all-equal weights
    w[20:40, : n // 2] = 0.0
    idx = render.select_appearance_samples(dev(w), k).cpu().numpy()
    assert np.array_equal(idx, orc.topk_indices(w, k))
    g = rng.gumbel(size=n).astype(np.float32)
    w2 = rng.uniform(0.01, 1, size=(50, n)).astype(np.float32)
    idx2 = render.select_appearance_samples(dev(w2), k, dev(g)).cpu().numpy()
    want2 = orc.topk_indices(w2, k, g)
    # log(w) + g is formed in FP32 on both sides (logf and numpy's log may
    # differ by an ulp): a row may differ only by indices whose keys sit within
    # FP32 rounding of the row's selection threshold
    keys = np.log(w2.astype(np.float64)) + g.astype(np.float64)[None, :]
    bad = np.nonzero(np.any(idx2 != want2, axis=1))[0]
    for r in bad:
        thr = np.sort(keys[r])[::-1][k - 1]
        diff = np.setxor1d(idx2[r], want2[r])
        assert np.all(np.abs(keys[r][diff] - thr) <= 4e-7 * max(abs(thr), 1.0)), (r, diff)
    log(f"  gumbel top-k: {len(bad)} of {len(w2)} rows differ, all at key ties")


def _camera_layout(rng, ncam, m, layout):
    """Per-row camera indices: independent per row, or runs of 70 rows cycling
    through the cameras (32-row tiles with one camera and tiles that straddle
    a change; a persistent workgroup meets several cameras in turn)."""
    if layout == "runs":
        return (np.arange(m) // 70) % ncam
    return rng.integers(0, ncam or 1, size=m)


# (cameras, camera layout, rows, appearance channels, Fourier frequencies):
# 48 channels with 2+2 frequencies run the kernels specialised for that shape,
# 24 channels with 6+6 (the reference's defaults) the generic ones, 5
# channels (15 features, not a multiple of 4) the scalar feature loads
MLP_CASES = [(None, "random", 1000, 48, 2), (5, "random", 1000, 48, 2),
             (5, "runs", 20000, 48, 2), (None, "random", 1000, 24, 6), (3, "runs", 5000, 24, 6),
             (None, "random", 700, 5, 2)]


@pytest.mark.parametrize("ncam,layout,m,ca,nf", MLP_CASES)
def test_mlp_fwd_bwd(lib, ncam, layout, m, ca, nf):
    log(f"case: mlp fwd/bwd ncam={ncam} layout={layout} m={m} ca={ca} nf={nf}")
    shape = orc.MlpShape(feat_dim=3 * ca, feature_n_freqs=nf, viewdir_n_freqs=nf,
                         num_cameras=ncam)
    p = oracle_params(4, 16, ca, shape, seed=3, bias_std=0.1)
    _, _, mp = to_device_params(p, shape)
    mlp = mlp_module(shape)
    rng = np.random.default_rng(11)
    f = f32(rng.normal(0, 0.3, size=(m, 3 * ca)))
    d = rng.normal(size=(m, 3))
    d = f32(d / np.linalg.norm(d, axis=1, keepdims=True))
    cams = _camera_layout(rng, ncam, m, layout)
    for leaf in mlp.flat_leaves(mp):
        leaf.requires_grad_(True)
    tf = dev(f).requires_grad_(True)
    rgb = mlp.apply(mp, tf, dev(d), dev(cams, torch.int32))
    want, cache = orc.feature_mlp_forward(p, f, d, cams, shape)
    err = report("rgb", rgb.detach().cpu().numpy(), want)
    assert err.max() < 2e-5
    g = f32(rng.normal(size=(m, 3)))
    (rgb * dev(g)).sum().backward()
    grads, gf = orc.feature_mlp_backward(p, cache, g, shape)
    ef = report("grad features", tf.grad.cpu().numpy(), gf)
    assert ef.max() < 1e-4 * np.abs(gf).max() + 1e-6
    names = ["w0", "w1", "b1", "w2", "b2", "w3", "b3"] + (["embed"] if ncam else [])
    for name, leaf in zip(names, mlp.flat_leaves(mp)):
        e = report(f"grad {name}", leaf.grad.cpu().numpy(), grads[name])
        assert e.max() < 1e-4 * np.abs(grads[name]).max() + 1e-6, name


def test_adam(lib):
    log("case: adam")
    rng = np.random.default_rng(1)
    n = 1001
    p0 = f32(rng.normal(size=n))
    p, m, v = dev(p0), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    wp, wm, wv = p0.copy(), np.zeros(n), np.zeros(n)
    for count in range(1, 4):
        g = f32(rng.normal(size=n))
        tg = dev(g)
        lib.call("tfh_adam_step", lib.ptr(p), lib.ptr(tg), lib.ptr(m), lib.ptr(v), n, 0.02, 0.7,
                 0.9, 0.99, 1e-8, count, 1, lib.stream())
        wp, wm, wv = orc.adam_update(wp, g, wm, wv, count, 0.02, 0.7)
        assert float(tg.abs().max()) == 0.0
    report("adam params", p.cpu().numpy(), wp)
    np.testing.assert_allclose(p.cpu().numpy(), wp, rtol=1e-6, atol=1e-7)


def test_abi_errors(lib):
    log("case: abi errors")
    with pytest.raises(lib.TfhError, match="k .* outside"):
        lib.call("tfh_topk_select", lib.ptr(torch.zeros(10, device="cuda")), 1, 10, 11, None,
                 lib.ptr(torch.zeros(11, dtype=torch.int32, device="cuda")), lib.stream())
    vm = lib.Vm(None, None, 10, 4)
    with pytest.raises(lib.TfhError, match="null plane"):
        lib.call("tfh_vm_interpolate", ctypes.byref(vm), None, 5, None, lib.stream())


@pytest.mark.parametrize("ncam,layout,m,ca,nf",
                         [(None, "random", 777, 48, 2), (5, "random", 777, 48, 2)] + MLP_CASES[2:])
def test_mlp_bf16_fwd_bwd(lib, ncam, layout, m, ca, nf):
    """BF16 operands on MFMA, FP32 accumulation. Checked against the oracle with
    the same BF16 operand rounding (tight: 2e-3 of the largest value, the FP32
    accumulation order and the rounding of near-ties differ), and reported
    against float64 (bound: 2e-2 absolute on RGB).

    Gradients: every element within 2e-2 of the largest, except isolated
    rounding-tie events. With 6 frequencies an encoded feature sin(32 h0) that
    sits on a BF16 rounding tie rounds differently under FP32 and FP64
    accumulation of h0, and the chain rule scales that by up to 32: the oracle
    itself, FP32 against FP64 accumulation, differs by 0.174 (5% of the
    largest feature gradient) at 2 of the 360,000 feature gradients of the
    24-channel case. So at most 1e-3 of the elements may exceed 2e-2 of the
    largest, and the relative L2 error must stay below 1e-2."""
    log(f"case: mlp bf16 ncam={ncam} layout={layout} m={m} ca={ca} nf={nf}")
    shape = orc.MlpShape(feat_dim=3 * ca, feature_n_freqs=nf, viewdir_n_freqs=nf,
                         num_cameras=ncam)
    p = oracle_params(4, 16, ca, shape, seed=3, bias_std=0.1)
    _, _, mp = to_device_params(p, shape)
    mlp = mlp_module(shape, "bf16")
    rng = np.random.default_rng(12)
    f = f32(rng.normal(0, 0.3, size=(m, 3 * ca)))
    d = rng.normal(size=(m, 3))
    d = f32(d / np.linalg.norm(d, axis=1, keepdims=True))
    cams = _camera_layout(rng, ncam, m, layout)
    for leaf in mlp.flat_leaves(mp):
        leaf.requires_grad_(True)
    tf = dev(f).requires_grad_(True)
    rgb = mlp.apply(mp, tf, dev(d), dev(cams, torch.int32))
    want64, _ = orc.feature_mlp_forward(p, f, d, cams, shape)
    want, cache = orc.feature_mlp_forward(p, f, d, cams, shape, operand_round=orc.bf16_round)
    got = rgb.detach().cpu().numpy()
    assert report("rgb (bf16) vs float64", got, want64).max() < 2e-2
    assert report("rgb (bf16) vs bf16-operand oracle", got, want).max() < 2e-3
    g = f32(rng.normal(size=(m, 3)))
    (rgb * dev(g)).sum().backward()
    grads, gf = orc.feature_mlp_backward(p, cache, g, shape)
    def check(name, got_g, want_g):
        e = report(name, got_g, want_g)
        bound = 2e-2 * np.abs(want_g).max() + 1e-6
        rel_l2 = np.linalg.norm(got_g - want_g) / max(np.linalg.norm(want_g), 1e-30)
        assert (e > bound).mean() <= 1e-3, (name, (e > bound).sum())
        assert rel_l2 < 1e-2, (name, rel_l2)

    check("grad features (bf16)", tf.grad.cpu().numpy().astype(np.float64), gf)
    names = ["w0", "w1", "b1", "w2", "b2", "w3", "b3"] + (["embed"] if ncam else [])
    for name, leaf in zip(names, mlp.flat_leaves(mp)):
        check(f"grad {name} (bf16)", leaf.grad.cpu().numpy().astype(np.float64), grads[name])
