"""GPU tests of the training-loop layer: grid upsampling of the factors and the
Adam moments (training.py:245-276), checkpoint round trip, the ground-truth
scene renderer, chunked full-frame rendering (render.py:49-102,
render_360.py:122-160), and what the proposal sampler gains over uniform
sampling at the same number of field samples."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import tensorf_oracle as orc
from tests.gpu_common import log, report

pytestmark = pytest.mark.gpu


def _config(g=24, proposal=None, n=64, steps=100, upsamp=(), density_shift=10.0):
    from tensorf_amd import training
    return training.TensorfConfig(density_shift=density_shift,
        minibatch_size=2048, n_iters=steps, initial_aabb_min=(-2.0,) * 3,
        initial_aabb_max=(2.0,) * 3, appearance_feat_dim=48, density_feat_dim=16,
        feature_n_freqs=2, viewdir_n_freqs=2, grid_dim_init=g, grid_dim_final=g + 8,
        upsamp_iters=upsamp, density_samples_per_ray=None if proposal else n,
        appearance_samples_per_ray=n if not proposal else None, proposal=proposal)


def _small_dataset(views=12, res=48, seed=0):
    from tensorf_amd import scenes
    scene = scenes.bounded_blob_scene(seed=seed)
    c2ws = scenes.hemisphere_cameras(views, radius=4.0, seed=1)
    return scene, c2ws, scenes.make_dataset(scene, c2ws, res, res,
                                            scenes.focal_800() * res / 800.0, gt_samples=2048)


def test_resize_grid_moves_moments():
    log("case: resize_grid 24 -> 32 with Adam moments")
    from tensorf_amd import training
    _, _, data = _small_dataset(views=4, res=32)
    st = training.TrainState.initialize(_config(), 24, prng_key=1, num_cameras=4)
    batches = data.minibatches(1024, seed=0)
    for _ in range(3):
        st.training_step(next(batches))
    torch.cuda.synchronize()
    before = {}
    for tag, buf in (("p", st.flat.params), ("m", st.flat.exp_avg), ("v", st.flat.exp_avg_sq)):
        vec, mat = st._vm(buf, 2).to_reference_layout()
        before[tag] = (vec.cpu().numpy().astype(np.float64), mat.cpu().numpy().astype(np.float64))
    st.resize_grid(32)
    for tag, buf in (("p", st.flat.params), ("m", st.flat.exp_avg), ("v", st.flat.exp_avg_sq)):
        vec, mat = st._vm(buf, 2).to_reference_layout()
        wv, wm = orc.vm_resize(before[tag][0], before[tag][1], 32)
        e = report(f"{tag} matrix", mat.cpu().numpy(), wm)
        assert e.max() <= 1e-6 * max(np.abs(wm).max(), 1e-30)
        e = report(f"{tag} vector", vec.cpu().numpy(), wv)
        assert e.max() <= 1e-6 * max(np.abs(wv).max(), 1e-30)
    st.training_step(next(batches))  # the step after an upsample runs on the new grid
    torch.cuda.synchronize()
    assert np.isfinite(float(st._mse.item()))


def test_checkpoint_roundtrip(tmp_path):
    log("case: checkpoint round trip with proposal")
    from tensorf_amd import proposal, training
    pc = proposal.ProposalConfig(samples_round0=32, samples_round1=16, samples_out=24,
                                 grid_dim=16, channels=4)
    _, _, data = _small_dataset(views=4, res=32)
    st = training.TrainState.initialize(_config(proposal=pc), 24, prng_key=2, num_cameras=4)
    batches = list(zip(range(4), data.minibatches(1024, seed=0)))
    for _, b in batches[:2]:
        st.training_step(b)
    path = tmp_path / "state.npz"
    training.save_checkpoint(st, path)
    st2 = training.load_checkpoint(path)
    for a, b in ((st.flat, st2.flat), (st.prop_flat, st2.prop_flat)):
        for buf in ("params", "exp_avg", "exp_avg_sq"):
            assert torch.equal(getattr(a, buf), getattr(b, buf)), buf
    assert (st2.step, st2.adam_count) == (st.step, st.adam_count)
    st.training_step(batches[2][1])
    st2.training_step(batches[2][1])
    torch.cuda.synchronize()
    # the loss and the factor gradients are summed with atomics (order-dependent
    # rounding): compare with tolerances
    assert abs(st._mse.item() - st2._mse.item()) <= 1e-5 * st._mse.item()
    for a, b in ((st.flat, st2.flat), (st.prop_flat, st2.prop_flat)):
        d = (a.params - b.params).abs()
        frac = (d > 1e-5).float().mean().item()
        log(f"  max |param difference| after one more step: {d.max().item():.3e}, "
            f"fraction above 1e-5: {frac:.2e}")
        assert frac <= 1e-3


def test_ground_truth_renderer_closed_form():
    """One blob seen head-on through its centre, evaluated against a float64
    NumPy quadrature of the same integral; empty rays are white."""
    log("case: blob scene renderer")
    from tensorf_amd import cameras, scenes
    scene = scenes.BlobScene(np.array([[0.0, 0.0, 0.0, 0.3, 20.0, 0.2, 0.5, 0.9]]))
    o = np.array([[0.0, 0.0, -4.0], [0.0, 3.0, -4.0], [0.1, 0.05, -4.0]])
    d = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    rays = cameras.Rays3D(torch.tensor(o, dtype=torch.float32, device="cuda"),
                          torch.tensor(d, dtype=torch.float32, device="cuda"),
                          torch.zeros(3, dtype=torch.int32, device="cuda"))
    rgb, opa, dep = scenes.render_ground_truth(scene, rays, 8192, with_depth=True)
    ts = np.linspace(1.0, 7.0, 200001)
    for r in range(3):
        pts = o[r] + ts[:, None] * d[r]
        sig = scene.density(pts)
        tau = np.concatenate([[0.0], np.cumsum(0.5 * (sig[1:] + sig[:-1]) * np.diff(ts))])
        T = np.exp(-tau)
        opa_ref = 1 - T[-1]
        col_ref = opa_ref * scene.blobs[0, 5:8] + T[-1]
        np.testing.assert_allclose(rgb[r].cpu().numpy(), col_ref, atol=2e-4)
        assert abs(opa[r].item() - opa_ref) < 2e-4
        if opa_ref > 0.5:
            assert abs(dep[r].item() - np.interp(0.5, 1 - T, ts)) < 2e-3
        else:
            assert not np.isfinite(dep[r].item())
    assert abs(opa[1].item()) < 1e-6 and np.allclose(rgb[1].cpu().numpy(), 1.0, atol=1e-6)


def test_proposal64_matches_uniform512_at_a_fraction_of_the_samples():
    """What the learned proposal gains over the method without it: 64 field
    samples placed by the proposal (rounds 128 -> 64, BF16 MLP, 16,384 rays a
    step) against the reference's stratified 512-sample sampler (FP32 MLP,
    4,096 rays a step), each at its benchmark batch, on the 8-blob scene (24
    views at 128x128, VM 128^3, 3,000 steps, the field starting transparent with
    density shift -10 as TensoRF's original code). Judged on three held-out
    views: the proposal's PSNR at most 0.1 dB below uniform-512's (or above)
    and its median-distance error no worse, at 5x the training rays/s or more.
    Measured on one MI355X: +3.08 dB held out (depth error 0.0205 against
    0.0337) at 6.91x; with the training seeds 4 and 5 instead of 3: +1.04 and
    +0.72 dB at 6.86x and 6.90x. The proposal's PSNR on its training batches is
    2.3-2.6 dB lower than uniform-512's in these runs: uniform-512 fits the 24
    training views more closely and generalises less."""
    log("case: proposal-64 vs uniform-512, 3000 steps")
    import time
    from tensorf_amd import cameras, proposal, render, scenes, training, utils
    views, res, g, steps = 24, 128, 128, 3000
    scene = scenes.bounded_blob_scene(seed=0)
    c2ws = scenes.hemisphere_cameras(views, radius=4.0, seed=1)
    focal = scenes.focal_800() * res / 800.0
    data = scenes.make_dataset(scene, c2ws, res, res, focal, gt_samples=2048)
    spiral = scenes.hemisphere_cameras(views + 3, radius=4.0, seed=7)
    held = [cameras.pinhole_rays(spiral[i], res, res, focal, 0) for i in (3, 9, 15)]
    truth = [scenes.render_ground_truth(scene, r, 2048, with_depth=True) for r in held]
    res_psnr, held_psnr, depth_err, rate = {}, {}, {}, {}
    for name, pc, n, prec, r in (
            ("uniform-512", None, 512, "fp32", 4096),
            ("proposal-64", proposal.ProposalConfig(samples_round0=128, samples_round1=64,
                                                    samples_out=64, anneal_steps=500), 64,
             "bf16", 16384)):
        cfg = dataclasses.replace(_config(g=g, proposal=pc, n=n, steps=steps,
                                          density_shift=-10.0), mlp_precision=prec,
                                  minibatch_size=r)
        st = training.TrainState.initialize(cfg, g, prng_key=3, num_cameras=views)
        batches = data.minibatches(r, seed=3)
        mses = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(steps):
            st, logs = st.training_step(next(batches))
            if it >= steps - 50:
                mses.append(logs["mse"])
        torch.cuda.synchronize()
        rate[name] = r * steps / (time.perf_counter() - t0)
        res_psnr[name] = utils.psnr_from_mse(float(torch.stack(mses).mean().item()))
        ps, des = [], []
        for rays, (gt, _, gt_depth) in zip(held, truth):
            pred = torch.tensor(render.render_rays_batched(
                st.appearance_mlp, st.learnable_params, st.aabb, rays, None,
                st.render_config(), batch_size=65536), device=gt.device)
            ps.append(utils.psnr_from_mse(float(((pred - gt) ** 2).mean())))
            dep = render.render_rays_batched(
                st.appearance_mlp, st.learnable_params, st.aabb, rays, None,
                st.render_config(render.RenderMode.DIST_MEDIAN), batch_size=65536)
            gd = gt_depth.cpu().numpy()
            both = np.isfinite(gd) & np.isfinite(dep)
            des.append(float(np.abs(dep[both] - gd[both]).mean()))
        held_psnr[name], depth_err[name] = float(np.mean(ps)), float(np.mean(des))
        log(f"  {name}: training psnr over the last 50 steps {res_psnr[name]:.2f} dB, held-out "
            f"{held_psnr[name]:.2f} dB ({', '.join(f'{x:.2f}' for x in ps)}), median-distance "
            f"error {depth_err[name]:.4f}, {rate[name]:.0f} train rays/s")
    speedup = rate["proposal-64"] / rate["uniform-512"]
    log(f"  proposal-64 - uniform-512: {held_psnr['proposal-64'] - held_psnr['uniform-512']:+.2f} "
        f"dB held out at {speedup:.2f}x the rays/s")
    assert held_psnr["proposal-64"] >= held_psnr["uniform-512"] - 0.1
    assert depth_err["proposal-64"] <= depth_err["uniform-512"]
    assert speedup >= 5.0


@pytest.mark.parametrize("precision,film", [("fp32", False), ("bf16", False), ("fp32", True),
                                            ("bf16", True)])
def test_train_step_kept_activations_match_recompute(monkeypatch, precision, film):
    """The fused training step keeps the MLP's hidden activations between its
    forward and backward when the workspace has the room
    (tfh_train_step_workspace_size) and recomputes them with a render-sized
    workspace; both give the same step: the same MSE bit for bit, and the same
    Adam update except where the order of the scatter atomics tips a
    near-zero gradient (at most 1e-3 of the parameters off by > 1e-3 lr)."""
    log(f"case: train step, kept activations vs recompute, {precision}, FiLM {film}")
    from tensorf_amd import _lib, training
    _, c2ws, data = _small_dataset(views=4, res=32)
    out = []
    for keep in (True, False):
        cfg = dataclasses.replace(_config(g=24, n=64), mlp_precision=precision,
                                  camera_embeddings=film)
        st = training.TrainState.initialize(cfg, 24, prng_key=5, num_cameras=4)
        p0 = st.flat.params.clone()
        if not keep:
            monkeypatch.setattr(_lib.lib, "tfh_train_step_workspace_size",
                                _lib.lib.tfh_render_workspace_size)
        st, logs = st.training_step(next(data.minibatches(1024, seed=2)))
        torch.cuda.synchronize()
        out.append((float(logs["mse"].item()), (st.flat.params - p0).cpu().numpy(),
                    st._ws.numel(), st.flat.c_struct().tensor_count))
        monkeypatch.undo()
    (mse_a, upd_a, ws_a, split), (mse_b, upd_b, ws_b, _) = out
    log(f"  workspace {ws_a} vs {ws_b} bytes, mse {mse_a:.9e} vs {mse_b:.9e}")
    assert ws_a > ws_b
    assert abs(mse_a - mse_b) <= 1e-5 * mse_b
    opt = training.OptimizerConfig()
    for name, sl, lr in (("factors", slice(0, split), opt.lr_init_tensor),
                         ("mlp", slice(split, None), opt.lr_init_mlp)):
        e = report(f"update {name}", upd_a[sl], upd_b[sl])
        frac_bad = float(np.mean(e > 1e-3 * lr))
        log(f"  fraction of {name} updates off by >1e-3 lr: {frac_bad:.2e}")
        assert frac_bad <= 1e-3, name


@pytest.mark.parametrize("mode", ["rgb", "median", "mean"])
def test_render_rays_batched_800x800_chunks(mode):
    """render_rays_batched (render.py:49-102) of an 800x800 view in 65,536-ray
    chunks (9 full chunks and a ragged one of 50,176 rays) equals one
    render_rays call over all 640,000 rays bit for bit (every ray is rendered
    independently; the stratified jitter is shared by all rays), keeps the
    (H, W[, 3]) batch axes, and the full-frame driver (render_360.py:122-160)
    turns it into the same 8-bit image."""
    log(f"case: 800x800 chunked render, {mode}")
    from tensorf_amd import cameras, render, render_360, scenes, training
    st = training.TrainState.initialize(_config(g=48, n=64), 48, prng_key=11, num_cameras=1)
    m = {"rgb": render.RenderMode.RGB, "median": render.RenderMode.DIST_MEDIAN,
         "mean": render.RenderMode.DIST_MEAN}[mode]
    cfg = render.RenderConfig(0.05, 200.0, m, 64, 16)
    c2w = scenes.hemisphere_cameras(5, radius=4.0, seed=3)[2]
    rays = cameras.pinhole_rays(c2w, 800, 800, scenes.focal_800(), 0)
    out = render.render_rays_batched(st.appearance_mlp, st.learnable_params, st.aabb, rays, 4,
                                     cfg, batch_size=65536)
    assert out.shape == ((800, 800, 3) if mode == "rgb" else (800, 800))
    with torch.no_grad():
        whole = render.render_rays(st.appearance_mlp, st.learnable_params, st.aabb,
                                   rays.reshape_flat(), 4, cfg).cpu().numpy()
    whole = whole.reshape(out.shape)
    log(f"  finite outputs: {np.isfinite(out).mean():.3f}")
    assert np.array_equal(out, whole, equal_nan=True)
    frames = list(render_360.render_frames(
        st, c2w, 800, 800, 2 * np.arctan(400 / scenes.focal_800()), 1, m,
        density_samples_per_ray=64, appearance_samples_per_ray=16, batch_size=65536,
        translation_scale=1.0, prng_key=4))
    want = render_360.rgb_image(out) if mode == "rgb" else render_360.DistanceImages()(out)
    assert frames[0].shape == (800, 800, 3) and frames[0].dtype == np.uint8
    # the driver derives the focal length from the field of view (float64
    # round trip through tan/atan), so a pixel may land on the other side of
    # an 8-bit rounding step
    diff = np.abs(frames[0].astype(np.int32) - want.astype(np.int32))
    log(f"  driver image: {np.mean(diff > 0):.2e} of the bytes differ, by at most {diff.max()}")
    assert diff.max() <= 1 and np.mean(diff > 0) < 1e-3
