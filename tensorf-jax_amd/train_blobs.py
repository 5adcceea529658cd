"""Training and evaluation driver on the procedural blob scenes.

The role of the reference's train_lego.py / render_360.py (training.py:279-368,
render_360.py:59-170) for the synthetic scenes this build trains on: ground
truth by FP64 quadrature (tfh_blob_scene_render), cycled shuffled minibatches,
TrainState.training_step with grid upsampling (factors and Adam moments), then
held-out views rendered in chunks (render_rays_batched) for PSNR and median
depth error. Prints progress lines and writes one JSON summary.

    python tensorf-jax_amd/train_blobs.py --sampler uniform --precision fp32 \\
        --samples 512 --steps 10000 --out result.json
    python tensorf-jax_amd/train_blobs.py --sampler proposal --precision bf16 \\
        --rays 16384 --steps 10000 --out result.json
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def log(msg):
    print(f"[train_blobs] {msg}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scene-seed", type=int, default=0)
    ap.add_argument("--sampler", choices=["uniform", "proposal"], default="uniform")
    ap.add_argument("--precision", choices=["fp32", "bf16"], default="fp32")
    ap.add_argument("--samples", type=int, default=512, help="uniform samples per ray")
    ap.add_argument("--appearance-samples", type=int, default=0, help="0 = all samples")
    ap.add_argument("--prop-rounds", default="128,64", help="proposal samples per round")
    ap.add_argument("--prop-out", type=int, default=64, help="field samples with the proposal")
    ap.add_argument("--prop-pad", type=float, default=0.64, help="histogram pad mass")
    ap.add_argument("--prop-anneal", type=int, default=1000, help="weight-power ramp steps")
    ap.add_argument("--prop-grid", type=int, default=64)
    ap.add_argument("--rays", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--views", type=int, default=100)
    ap.add_argument("--res", type=int, default=800)
    ap.add_argument("--test-views", type=int, default=5)
    ap.add_argument("--test-res", type=int, default=800)
    ap.add_argument("--grid-init", type=int, default=128)
    ap.add_argument("--grid-final", type=int, default=300)
    ap.add_argument("--upsamp", default="2000,3000,4000,5500,7000")
    ap.add_argument("--box", type=float, default=2.0, help="half extent of the scene box")
    ap.add_argument("--density-shift", type=float, default=10.0,
                    help="sigma = softplus(features + shift); +10 is the reference's")
    ap.add_argument("--gt-samples", type=int, default=4096)
    ap.add_argument("--log-every", type=int, default=500)
    ap.add_argument("--chunk", type=int, default=65536)
    ap.add_argument("--eval-jitter", type=int, default=1,
                    help="1: render the test views with jittered samples (prng key 0), "
                         "0: deterministic sample positions")
    ap.add_argument("--checkpoint", default="")
    ap.add_argument("--out", default="")
    args = ap.parse_args()

    import numpy as np
    import torch
    from tensorf_amd import proposal, render, scenes, training, utils

    torch.cuda.set_device(0)
    scene = scenes.bounded_blob_scene(seed=args.scene_seed)
    c2ws = scenes.hemisphere_cameras(args.views + args.test_views, radius=4.0, seed=1)
    # every (views / test_views)-th camera of the spiral is held out
    stride = max(1, (args.views + args.test_views) // max(args.test_views, 1))
    test_ids = [i * stride for i in range(args.test_views)]
    train_c2ws = [c for i, c in enumerate(c2ws) if i not in test_ids]
    focal = scenes.focal_800() * args.res / 800.0
    t0 = time.perf_counter()
    data = scenes.make_dataset(scene, train_c2ws, args.res, args.res, focal,
                               gt_samples=args.gt_samples)
    torch.cuda.synchronize()
    log(f"ground truth: {len(train_c2ws)} views at {args.res}^2 in "
        f"{time.perf_counter() - t0:.1f} s")

    upsamp = tuple(int(x) for x in args.upsamp.split(",") if x)
    pc = None
    if args.sampler == "proposal":
        rounds = [int(x) for x in args.prop_rounds.split(",")]
        pc = proposal.ProposalConfig(samples_round0=rounds[0],
                                     samples_round1=rounds[1] if len(rounds) > 1 else 0,
                                     samples_out=args.prop_out, histogram_pad=args.prop_pad,
                                     anneal_steps=args.prop_anneal, grid_dim=args.prop_grid)
    n = args.prop_out if pc else args.samples
    k = args.appearance_samples or n
    cfg = training.TensorfConfig(
        minibatch_size=args.rays, n_iters=args.steps,
        initial_aabb_min=(-args.box,) * 3, initial_aabb_max=(args.box,) * 3,
        appearance_feat_dim=48, density_feat_dim=16, feature_n_freqs=2, viewdir_n_freqs=2,
        grid_dim_init=args.grid_init, grid_dim_final=args.grid_final, upsamp_iters=upsamp,
        density_samples_per_ray=None if pc else n, appearance_samples_per_ray=k,
        mlp_precision=args.precision, proposal=pc, density_shift=args.density_shift)

    t_train = [0.0]
    last = [time.perf_counter()]

    def log_fn(step, mse, logs):
        now = time.perf_counter()
        extra = f" prop_loss={float(logs['prop_loss'].item()):.3e}" if "prop_loss" in logs else ""
        log(f"step {step}: mse={mse:.6f} psnr={utils.psnr_from_mse(mse):.2f}{extra} "
            f"({now - last[0]:.1f} s)")
        last[0] = now

    torch.cuda.synchronize()
    t1 = time.perf_counter()
    state = training.run_training_loop(cfg, data, n_iters=args.steps, log_every=args.log_every,
                                       log_fn=log_fn)
    torch.cuda.synchronize()
    t_train[0] = time.perf_counter() - t1
    log(f"trained {args.steps} steps in {t_train[0]:.1f} s "
        f"({args.rays * args.steps / t_train[0]:.0f} rays/s incl. upsampling and logging)")
    if args.checkpoint:
        training.save_checkpoint(state, args.checkpoint)
        log(f"checkpoint written to {args.checkpoint}")

    # held-out views
    from tensorf_amd import cameras
    rgb_cfg = state.render_config()
    depth_cfg = state.render_config(render.RenderMode.DIST_MEDIAN)
    psnrs, depth_err, depth_miss = [], [], []
    eval_key = 0 if args.eval_jitter else None
    t_render = 0.0
    for i in test_ids:
        tf = focal * args.test_res / args.res
        rays = cameras.pinhole_rays(c2ws[i], args.test_res, args.test_res, tf, 0)
        gt, opa, gd = scenes.render_ground_truth(scene, rays, args.gt_samples, with_depth=True)
        torch.cuda.synchronize()
        tr = time.perf_counter()
        pred = render.render_rays_batched(state.appearance_mlp, state.learnable_params,
                                          state.aabb, rays, eval_key, rgb_cfg,
                                          batch_size=args.chunk)
        t_render += time.perf_counter() - tr
        dep = render.render_rays_batched(state.appearance_mlp, state.learnable_params,
                                         state.aabb, rays, eval_key, depth_cfg,
                                         batch_size=args.chunk)
        pred = torch.tensor(pred, device=gt.device)
        psnrs.append(utils.psnr_from_mse(float(((pred - gt) ** 2).mean())))
        gd = gd.cpu().numpy()
        fg = (opa.cpu().numpy() > 0.5) & np.isfinite(gd)
        ok = fg & np.isfinite(dep)
        depth_err.append(float(np.abs(dep[ok] - gd[ok]).mean()) if ok.any() else float("nan"))
        depth_miss.append(float((fg & ~np.isfinite(dep)).sum() / max(fg.sum(), 1)))
        log(f"test view {i}: psnr={psnrs[-1]:.3f} median-depth |err|={depth_err[-1]:.4f} "
            f"missed={depth_miss[-1]:.4f}")
    # one training view, to show the gap between fitted and held-out views
    rays = cameras.pinhole_rays(train_c2ws[0], args.test_res, args.test_res,
                                focal * args.test_res / args.res, 0)
    gt, _ = scenes.render_ground_truth(scene, rays, args.gt_samples)
    pred = torch.tensor(render.render_rays_batched(state.appearance_mlp, state.learnable_params,
                                                   state.aabb, rays, eval_key, rgb_cfg,
                                                   batch_size=args.chunk), device=gt.device)
    train_view_psnr = utils.psnr_from_mse(float(((pred - gt) ** 2).mean()))
    log(f"training view 0: psnr={train_view_psnr:.3f}")
    res = {
        "sampler": args.sampler, "precision": args.precision, "rays_per_step": args.rays,
        "train_view_psnr": round(train_view_psnr, 4),
        "density_shift": args.density_shift, "eval_jitter": bool(args.eval_jitter),
        "samples_per_ray": n, "appearance_samples_per_ray": k,
        "proposal_rounds": args.prop_rounds if pc else None, "steps": args.steps,
        "train_views": len(train_c2ws), "res": args.res, "grid": [args.grid_init, args.grid_final],
        "upsamp_iters": list(upsamp), "train_seconds": round(t_train[0], 2),
        "train_rays_per_s_incl_overheads": round(args.rays * args.steps / t_train[0], 1),
        "test_views": test_ids, "test_res": args.test_res,
        "test_psnr": [round(p, 4) for p in psnrs], "mean_test_psnr": round(float(np.mean(psnrs)), 4),
        "median_depth_abs_err": [round(e, 5) for e in depth_err],
        "mean_median_depth_abs_err": round(float(np.nanmean(depth_err)), 5),
        "depth_missed_fraction": [round(m, 5) for m in depth_miss],
        "render_rays_per_s": round(args.test_views * args.test_res ** 2 / max(t_render, 1e-9), 1),
    }
    log(json.dumps(res))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
