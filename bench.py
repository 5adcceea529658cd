"""Benchmark of the TensoRF training step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the plain uniform-512 FP32 configuration the metric is quoted on):
one TrainState.training_step (training.py:101-205) = render 4096 rays x 512
stratified samples in the [-1.5,1.5]^3 box through a 300^3 VM grid (density
16, appearance 48 channels per mode), every sample shaded by the FP32 feature
MLP (144->27, 150->128->128->3, 2+2 Fourier frequencies), MSE loss,
hand-written backward, Adam on all 17.3 M parameters. Rays: origins on the
upper hemisphere of radius 4 aimed at uniform points of [-0.8,0.8]^3;
targets uniform in [0,1] (synthetic data, resident in HBM before timing).

Prints ONE JSON line with train rays/s (value), render rays/s, the live
roofline of the dominant kernel (HIP events on the library's launch stream)
and the CPU baseline (the oracle timed on host cores on a bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "tensorf-jax_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "train rays/s (fwd+bwd) & render rays/s; proposal vs uniform at <=0.1 dB PSNR"

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0}


def stage_model(stage, R, N, k, cd, ca, precision, n_params, prop=None):
    """Algorithmic work of one launch of a stage: (bound, amount, unit, peak).
    Bytes: VM taps 72*C per sample (3 modes x 6 taps x C x 4 B) plus the
    streamed per-sample I/O; FLOPs: 79,712 per feature-MLP row forward
    (nf=nv=2) and 11,392 per proposal-MLP row (2 (24*64 + 64*64 + 64)). The
    training step keeps the MLP's hidden activations, so its backward does
    the input and weight gradients of every layer (2 x 79,712 per row) plus
    the recompute of layer 0 (2 x 144 x 27).
    prop: (samples per ray over the proposal rounds, proposal channels)."""
    M = R * k
    mlp_row = 2 * (144 * 27 + 150 * 128 + 128 * 128 + 128 * 3)
    mfma_peak = PEAK_TFLOPS[precision]
    if stage == "density_pass":
        return "hbm", R * N * (72 * cd + 28), "GB/s", PEAK_HBM_GBS
    if stage == "appearance_gather":
        return "hbm", M * (72 * ca + 12 + 4 * 3 * ca), "GB/s", PEAK_HBM_GBS
    if stage == "mlp_fwd_kernel":
        return "mfma", M * mlp_row, "TFLOP/s", mfma_peak
    if stage == "mlp_bwd":
        return "mfma", M * (2 * mlp_row + 2 * 144 * 27), "TFLOP/s", mfma_peak
    if stage == "composite_bwd_density_scatter":
        return "hbm", R * N * (2 * 72 * cd + 36), "GB/s", PEAK_HBM_GBS
    if stage == "appearance_scatter":
        return "hbm", M * (2 * 72 * ca + 12 + 4 * 3 * ca), "GB/s", PEAK_HBM_GBS
    # every sample shaded on one grid: the compositing backward (+ the pair
    # sort) and one scatter of both fields
    if stage == "composite_bwd":
        return "hbm", R * N * 36, "GB/s", PEAK_HBM_GBS
    if stage == "factor_scatter":
        return ("hbm", R * N * 2 * 72 * cd + M * (2 * 72 * ca + 12 + 4 * 3 * ca), "GB/s",
                PEAK_HBM_GBS)
    if stage == "adam":
        return "hbm", n_params * 32, "GB/s", PEAK_HBM_GBS
    if prop is not None:
        ns, cp = prop
        prow = 2 * (3 * cp * 64 + 64 * 64 + 64)
        if stage == "prop_mlp_fwd":
            return "mfma", R * ns * prow, "TFLOP/s", PEAK_TFLOPS["fp32"]
        if stage == "prop_mlp_bwd":
            return "mfma", R * ns * 3 * prow, "TFLOP/s", PEAK_TFLOPS["fp32"]
    return None


def compulsory_bytes(stage, R, N, k, cd, ca, G, n_params, precision="fp32"):
    """HBM bytes a launch cannot avoid: every input read once, every output
    written once, the factors (or their gradients, read and written) once --
    no tap re-reads. Achieved / 8 TB/s on this model is at most 1 for any
    kernel, so it grades the memory-bound stages (the logical 72*C bytes per
    sample of the stage model counts cache hits and can exceed the peak)."""
    M = R * k
    sel = 4 if k < N else 0  # top-k indices
    dfac = 4 * (3 * G * G * cd + 3 * G * cd)
    afac = 4 * (3 * G * G * ca + 3 * G * ca)
    if stage == "density_pass":
        # rays in; points, distances, sigma, step, weight, transmittance out;
        # per-ray terms and the selection
        return R * 24 + dfac + R * N * 32 + R * 16 + M * sel
    if stage == "appearance_gather":
        return M * (12 + sel) + afac + M * 4 * 3 * ca
    if stage == "composite_bwd_density_scatter":
        # per sample: point, sigma, step, weight, transmittance; per shaded
        # sample: colour in, dL/dc out; per ray: terms and dL/dC; gradients RMW
        return R * N * 28 + M * (24 + sel) + R * 28 + 2 * dfac
    if stage == "appearance_scatter":
        return M * (12 + sel + 4 * 3 * ca) + afac + 2 * afac
    if stage == "composite_bwd":
        # the scans' inputs and dL/ds, dL/dc out; the sorted pair order out
        return R * N * (28 + 4 + 12) + M * 24 + R * 28
    if stage == "factor_scatter":
        # the pair order, points, dL/ds and the feature gradient in; both
        # fields' factors read and their gradients read and written once
        return R * N * (12 + 12 + 4) + M * 4 * 3 * ca + 3 * (dfac + afac)
    if stage == "adam":
        return n_params * 32
    if stage == "composite_rgb":
        # per shaded sample: its weight and colour (and the selection) in; per
        # ray: the transmittance and unbiasing terms and the target colour
        # in, the colour and dL/dC out
        return M * (16 + sel) + R * (8 + 12 + 12 + 12)
    if stage == "mlp_fwd_kernel":
        # the standard MLP forward gathers its own rows: points (and the
        # selection), view directions and the appearance factors in; the
        # feature rows (kept for the backward), the hidden activations kept
        # by a training step (1 KB a row FP32, 512 B BF16) and the colours out
        kept = 1024 if precision == "fp32" else 512
        return M * (12 + sel) + R * 12 + afac + M * (4 * 3 * ca + kept + 12)
    return None


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_batches(R, count, seed, device):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    out = []
    for _ in range(count):
        v = torch.randn((R, 3), generator=g, dtype=torch.float64)
        v[:, 2] = v[:, 2].abs()
        o = 4.0 * v / v.norm(dim=1, keepdim=True)
        tgt = (torch.rand((R, 3), generator=g, dtype=torch.float64) * 2 - 1) * 0.8
        d = tgt - o
        d = d / d.norm(dim=1, keepdim=True)
        colors = torch.rand((R, 3), generator=g)
        out.append((o.float().to(device), d.float().to(device), colors.to(device)))
    return out


_CPU = {}


def _cpu_worker_init(G, N, k, cd, ca, R_total):
    """One CPU-baseline worker: its own copy of the parameters and rays."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle import tensorf_oracle as orc
    _CPU["limits"] = threadpool_limits(1)
    shape = orc.MlpShape()
    p = orc.make_params(G, cd, ca, shape, seed=0)
    _CPU.update(orc=orc, shape=shape, R=R_total,
                p={key: v.astype(np.float32) for key, v in p.items()},
                rays=orc.make_rays(R_total, seed=1, radius=4.0, target_half=0.8),
                colors=np.random.default_rng(2).uniform(size=(R_total, 3)).astype(np.float32),
                cfg=orc.RenderConfig(0.05, 200.0, orc.RGB, N, k),
                jit=np.random.default_rng(3).uniform(size=N).astype(np.float32))


def _cpu_worker(span):
    """Forward and analytic backward of the rays [a, b) (render.py:105-279 and
    its VJP) in float32: the squared error sum and the parameter gradients of
    the batch-mean loss."""
    import numpy as np
    c = _CPU
    orc = c["orc"]
    a, b = span
    o, d = c["rays"]
    box = np.array([[-1.5] * 3, [1.5] * 3], np.float32)
    out, cache = orc.render_rays(c["p"], box, o[a:b], d[a:b], np.zeros(b - a), c["cfg"],
                                 c["shape"], c["jit"], dtype=np.float32)
    diff = out - c["colors"][a:b]
    gout = (2.0 * diff / (3 * c["R"])).astype(np.float32)
    grads = orc.render_rays_backward(c["p"], cache, gout, c["cfg"], c["shape"], np.float32)
    return float((diff.astype(np.float64) ** 2).sum()), grads


def cpu_baseline(R_sample, N, k, G, cd, ca, cores):
    """The oracle (NumPy float32 restatement of the reference) timed for one
    training step of R_sample rays of the workload on `cores` host cores: the
    rays are split into one contiguous part per core, each worker process runs
    the forward and analytic backward of its part single-threaded, and the
    parent sums the gradients and runs Adam on every parameter. Must run before
    the process touches the GPU (the workers are forked)."""
    import multiprocessing as mp
    import platform

    import numpy as np
    from oracle import tensorf_oracle as orc
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    bounds = np.linspace(0, R_sample, cores + 1).astype(int)
    spans = [(int(bounds[i]), int(bounds[i + 1])) for i in range(cores)]
    ctx = mp.get_context("fork")
    with ctx.Pool(cores, initializer=_cpu_worker_init,
                  initargs=(G, N, k, cd, ca, R_sample)) as pool:
        pool.map(_cpu_worker, [(0, 1)] * cores)  # every worker initialised and warm
        t0 = time.perf_counter()
        parts = pool.map(_cpu_worker, spans)
        sq = sum(pp[0] for pp in parts)
        grads = {key: sum(pp[1][key] for pp in parts) for key in parts[0][1]}
        t_grad = time.perf_counter()
        _cpu_worker_init(G, N, k, cd, ca, R_sample)  # the parent's parameters for Adam
        t_init = time.perf_counter()
        opt = orc.OptimizerConfig()
        params = _CPU["p"]
        for key, g in grads.items():
            lr = opt.lr_init_mlp if key in orc.MLP_KEYS else opt.lr_init_tensor
            orc.adam_update(params[key], g, np.zeros_like(g), np.zeros_like(g), 1, lr, 1.0,
                            dtype=np.float32)
        dt = (t_grad - t0) + (time.perf_counter() - t_init)
    mse = sq / (3 * R_sample)
    return {"value": R_sample / dt, "unit": "rays/s", "cores": cores, "kind": "port",
            "cpu_model": model,
            "sample": f"one oracle training step (fwd + analytic bwd + Adam over all "
                      f"{sum(v.size for v in grads.values())} params) on {R_sample} of the "
                      f"{N}-sample rays at G={G}, NumPy float32, the rays split over {cores} "
                      f"worker processes of one thread each; {dt:.1f} s, mse {mse:.5f}"}


def bench_proposal(args, cd, ca, G, n_out=64):
    """Proposal-sampled training and rendering: R rays, proposal rounds 128 ->
    64 on a 64^3 x 8-channel proposal field, n_out field samples per ray on the
    proposal's intervals, all shaded by the BF16 feature MLP, Adam on both
    parameter sets. Returns the numbers merged into the JSON line."""
    import torch
    from tensorf_amd import _lib, cameras, proposal, render, training
    R = args.prop_rays
    pc = proposal.ProposalConfig(samples_round0=128, samples_round1=64, samples_out=n_out)
    cfg = training.TensorfConfig(
        minibatch_size=R, n_iters=30000, initial_aabb_min=(-1.5,) * 3,
        initial_aabb_max=(1.5,) * 3, appearance_feat_dim=ca, density_feat_dim=cd,
        feature_n_freqs=2, viewdir_n_freqs=2, grid_dim_init=G, grid_dim_final=G, upsamp_iters=(),
        appearance_samples_per_ray=n_out, mlp_precision="bf16", proposal=pc)
    progress(f"proposal: init state G={G} R={R} 128->64->{n_out} bf16")
    state = training.TrainState.initialize(cfg, G, prng_key=0, num_cameras=1)
    batches = make_batches(R, 4, seed=5, device="cuda")
    cams = torch.zeros(R, dtype=torch.int32, device="cuda")
    mbs = [cameras.RenderedRays(c, cameras.Rays3D(o, d, cams)) for o, d, c in batches]
    for i in range(args.warmup):
        state.training_step(mbs[i % 4])
        torch.cuda.synchronize()
    progress("proposal: warmup done")
    t0 = time.perf_counter()
    for i in range(args.steps):
        state.training_step(mbs[i % 4])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    train = R * args.steps / dt
    progress(f"proposal: {1e3 * dt / args.steps:.3f} ms/step, {train:.1f} rays/s")
    lp = state.learnable_params
    rcfg = state.render_config()
    rays = [m.rays_wrt_world for m in mbs]
    ws = {}

    def render_once(i):
        env = render._RenderEnv(state.appearance_mlp, lp, state.aabb, rays[i % 4], 7, rcfg,
                                ws.get("w"))
        ws["w"] = env.ws
        env.forward()

    for i in range(2):
        render_once(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_render = max(3, args.steps // 2)
    for i in range(n_render):
        render_once(i)
    torch.cuda.synchronize()
    rend = R * n_render / (time.perf_counter() - t1)
    progress(f"proposal render: {rend:.1f} rays/s")
    stages = {}
    if not args.no_stages:
        _lib.profile_enable(True)
        for i in range(3):
            state.training_step(mbs[i % 4])
        torch.cuda.synchronize()
        rec = _lib.profile_read()
        _lib.profile_enable(False)
        stages = {name: round(ms / max(cnt, 1), 4) for name, (ms, cnt) in rec.items()}
        progress("proposal stages: " + ", ".join(f"{k}={v:.3f}ms" for k, v in stages.items()))
    return {"workload": f"proposal-sampled training step: R={R} rays, proposal 64^3 x 8 "
                        f"channels + 2x64 FP32 MLP, rounds 128 -> 64, {n_out} field samples on "
                        f"the proposal intervals all shaded by the BF16 MLP, VM {G}^3, Adam on "
                        f"both",
            "train_rays_per_s": round(train, 1), "ms_per_step": round(1e3 * dt / args.steps, 3),
            "render_rays_per_s": round(rend, 1), "stages_ms": stages,
            "final_mse": float(state._mse.item())}


def fisheye_batches(R, count, n_frames, seed, device):
    """Minibatches of R rays drawn from equidistant-fisheye frames (960x540,
    180 degree horizontal field) of cameras on a ring of radius 3 with +-0.3
    height jitter; camera index = frame index; uniform random targets."""
    import numpy as np
    import torch
    from tensorf_amd import cameras, scenes
    c2ws = scenes.ring_cameras(n_frames, radius=3.0, height_jitter=0.3)
    rng = np.random.default_rng(seed)
    w, h = 960, 540
    focal = (w / 2.0) / (np.pi / 2.0)
    out = []
    per = 4  # frames mixed per batch
    for b in range(count):
        os_, ds_, cs_ = [], [], []
        for f in rng.choice(n_frames, per, replace=False):
            rays = scenes.fisheye_rays(c2ws[f], w, h, focal, int(f), device)
            flat = rays.reshape_flat()
            sel = torch.tensor(rng.choice(w * h, R // per, replace=False), device=device)
            os_.append(flat.origins[sel])
            ds_.append(flat.directions[sel])
            cs_.append(flat.camera_indices[sel])
        colors = torch.tensor(rng.uniform(size=(R, 3)), dtype=torch.float32, device=device)
        out.append(cameras.RenderedRays(colors, cameras.Rays3D(torch.cat(os_), torch.cat(ds_),
                                                                torch.cat(cs_))))
    return out


def bench_leg(args, label, R, N, k, G, cd, ca, precision, contraction=False, n_frames=0,
              spacing="reference"):
    """Uniform-sampled training and rendering on one further workload; returns
    the numbers merged into the JSON line."""
    import torch
    from tensorf_amd import _lib, cameras, render, training
    box = 2.0 if contraction else 1.5
    cfg = training.TensorfConfig(
        minibatch_size=R, n_iters=30000, initial_aabb_min=(-box,) * 3,
        initial_aabb_max=(box,) * 3, appearance_feat_dim=ca, density_feat_dim=cd,
        feature_n_freqs=2, viewdir_n_freqs=2, grid_dim_init=G, grid_dim_final=G, upsamp_iters=(),
        scene_contraction=contraction, camera_embeddings=n_frames > 0,
        density_samples_per_ray=N, appearance_samples_per_ray=k, mlp_precision=precision,
        contracted_spacing=spacing)
    progress(f"{label}: init state G={G} R={R} N={N} k={k} {precision} "
             f"contraction={contraction} ({spacing} spacing) cameras={n_frames}")
    state = training.TrainState.initialize(cfg, G, prng_key=0, num_cameras=max(n_frames, 1))
    if contraction:
        mbs = fisheye_batches(R, 4, n_frames, seed=9, device="cuda")
    else:
        cams = torch.zeros(R, dtype=torch.int32, device="cuda")
        mbs = [cameras.RenderedRays(c, cameras.Rays3D(o, d, cams))
               for o, d, c in make_batches(R, 4, seed=7, device="cuda")]
    steps = max(3, min(args.steps, 10))
    for i in range(max(1, min(args.warmup, 3))):
        state.training_step(mbs[i % 4])
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        state.training_step(mbs[i % 4])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    train = R * steps / dt
    progress(f"{label}: {1e3 * dt / steps:.3f} ms/step, {train:.1f} rays/s")
    lp = state.learnable_params
    rcfg = state.render_config()
    ws = state._ws
    rays = [m.rays_wrt_world for m in mbs]

    def render_once(i):
        render._RenderEnv(state.appearance_mlp, lp, state.aabb, rays[i % 4], 7, rcfg,
                          ws).forward()

    render_once(0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        render_once(i)
    torch.cuda.synchronize()
    rend = R * steps / (time.perf_counter() - t1)
    progress(f"{label} render: {rend:.1f} rays/s")
    stages = {}
    if not args.no_stages:
        _lib.profile_enable(True)
        for i in range(3):
            state.training_step(mbs[i % 4])
        torch.cuda.synchronize()
        rec = _lib.profile_read()
        _lib.profile_enable(False)
        stages = {name: round(ms / max(cnt, 1), 4) for name, (ms, cnt) in rec.items()}
    return {"train_rays_per_s": round(train, 1), "ms_per_step": round(1e3 * dt / steps, 3),
            "render_rays_per_s": round(rend, 1), "stages_ms": stages,
            "final_mse": float(state._mse.item())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rays", type=int, default=4096)
    ap.add_argument("--samples", type=int, default=512)
    ap.add_argument("--appearance-samples", type=int, default=512)
    ap.add_argument("--grid", type=int, default=300)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--cpu-sample-rays", type=int, default=0,
                    help="rays of the CPU baseline's training step (default: 1024 per core)")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="host cores for the CPU baseline (default: the affinity set, at "
                         "most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-stages", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="bytes of one plain device copy run first (calibrates HBM counters)")
    ap.add_argument("--prop-rays", type=int, default=16384)
    ap.add_argument("--no-proposal", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the top-k 64 BF16 and contracted fisheye workloads")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the training step has no multi-GPU path "
              "(data parallelism over rays is a design note, not code). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    # the CPU baseline forks its workers, so it runs before anything touches the GPU
    cpu = None
    if not args.no_cpu_baseline:
        cores = args.cpu_cores or min(len(os.sched_getaffinity(0)), 16)
        r_cpu = args.cpu_sample_rays or 1024 * cores
        progress(f"cpu baseline: {r_cpu} rays on {cores} cores")
        cpu = cpu_baseline(r_cpu, args.samples, args.appearance_samples, args.grid, 16, 48,
                           cores)
        progress(f"cpu baseline: {cpu['value']:.1f} rays/s")

    import torch
    from tensorf_amd import _lib, cameras, render, training

    torch.cuda.set_device(0)
    if args.copy_probe:
        # one copy of known size: its FETCH_SIZE / WRITE_SIZE counters scale
        # the counters of the kernels profiled in the same run
        src = torch.ones(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    R, N, k, G = args.rays, args.samples, args.appearance_samples, args.grid
    cd, ca = 16, 48
    cfg = training.TensorfConfig(
        minibatch_size=R, n_iters=30000, initial_aabb_min=(-1.5,) * 3,
        initial_aabb_max=(1.5,) * 3, appearance_feat_dim=ca, density_feat_dim=cd,
        feature_n_freqs=2, viewdir_n_freqs=2, grid_dim_init=G, grid_dim_final=G, upsamp_iters=(),
        density_samples_per_ray=N, appearance_samples_per_ray=k, mlp_precision=args.precision)
    progress(f"init state G={G} R={R} N={N} k={k} {args.precision}")
    state = training.TrainState.initialize(cfg, G, prng_key=0, num_cameras=1)
    batches = make_batches(R, 4, seed=1, device="cuda")
    cams = torch.zeros(R, dtype=torch.int32, device="cuda")
    mbs = [cameras.RenderedRays(c, cameras.Rays3D(o, d, cams)) for o, d, c in batches]

    def step(i):
        state.training_step(mbs[i % len(mbs)])

    for i in range(args.warmup):
        step(i)
        torch.cuda.synchronize()
        progress(f"warmup step {i} done")
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = R * args.steps / dt
    progress(f"timed {args.steps} steps: {1e3 * dt / args.steps:.3f} ms/step")
    mse = float(state._mse.item())

    # render rays/s (RGB forward, no gradient) on the same batches
    lp = state.learnable_params
    rcfg = state.render_config()
    ws = state._ws
    rays = [m.rays_wrt_world for m in mbs]

    def render_once(i):
        env = render._RenderEnv(state.appearance_mlp, lp, state.aabb, rays[i % len(rays)], 7,
                                rcfg, ws)
        env.forward()

    for i in range(2):
        render_once(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_render = max(3, args.steps // 2)
    for i in range(n_render):
        render_once(i)
    torch.cuda.synchronize()
    render_value = R * n_render / (time.perf_counter() - t1)
    progress(f"render: {render_value:.1f} rays/s")

    # live per-stage timing (HIP events on the launch stream) of a few steps
    stages = {}
    roofline = None
    if not args.no_stages:
        _lib.profile_enable(True)
        n_prof = 3
        for i in range(n_prof):
            step(i)
        torch.cuda.synchronize()
        rec = _lib.profile_read()
        _lib.profile_enable(False)
        progress("stages: " + ", ".join(f"{k}={v[0] / max(v[1], 1):.3f}ms" for k, v in rec.items()))
        n_params = state.flat.total
        best = None
        for name, (ms, cnt) in rec.items():
            avg = ms / max(cnt, 1)
            stages[name] = round(avg, 4)
            model = stage_model(name, R, N, k, cd, ca, args.precision, n_params)
            if model is not None and (best is None or avg > best[1]):
                best = (name, avg, model)
        if best is not None:
            name, avg, (bound, amount, unit, peak) = best
            achieved = amount / (avg * 1e-3) / (1e9 if unit == "GB/s" else 1e12)
            traffic = None
            tpath = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
            if os.path.exists(tpath):
                with open(tpath) as f:
                    tr = json.load(f)
                if tr.get("workload") == f"R{R}_N{N}_k{k}_G{G}_{args.precision}":
                    traffic = tr.get("bytes_per_launch")
            roofline = {"kernel": name, "bound": bound, "achieved": round(achieved, 3),
                        "peak": peak, "unit": unit, "frac": round(achieved / peak, 4),
                        "traffic": traffic, "avg_ms": round(avg, 4),
                        "algorithmic_per_launch": amount,
                        "traffic_source": (f"profiles/traffic_{name}.json: rocprofv3 --pmc "
                                           f"FETCH_SIZE and WRITE_SIZE of this workload, "
                                           f"scaled by a copy of known size (not measured "
                                           f"in this run)") if traffic is not None else None}
        hbm = {}
        for name, (ms, cnt) in rec.items():
            b = compulsory_bytes(name, R, N, k, cd, ca, G, n_params, args.precision)
            if b is None:
                continue
            avg = ms / max(cnt, 1)
            gbs = b / (avg * 1e-3) / 1e9
            hbm[name] = {"compulsory_bytes": int(b), "avg_ms": round(avg, 4),
                         "achieved_GBs": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4)}
        if hbm:
            roofline["hbm_compulsory"] = hbm

    prop = prop128 = topk = unb = None
    del state, mbs, batches, rays, lp, ws
    torch.cuda.empty_cache()
    if not args.no_extra:
        topk = bench_leg(args, "topk64", 8192, 512, 64, G, cd, ca, "bf16")
        topk["workload"] = (f"uniform training step: R=8192 rays x 512 stratified samples, "
                            f"per-ray top-64 by weight shaded by the BF16 MLP, VM {G}^3")
        torch.cuda.empty_cache()
        unb = bench_leg(args, "unbounded", 16384, 512, 512, G, cd, ca, "bf16",
                        contraction=True, n_frames=200, spacing="uniform_contracted")
        unb["workload"] = (f"unbounded training step: L-inf scene contraction, R=16384 "
                           f"equidistant-fisheye rays (960x540 frames, 200 ring cameras, "
                           f"FiLM camera embedding) x 512 samples spaced uniformly in "
                           f"contracted distance g(t) = t (t <= 1), 2 - 1/t (t > 1), all "
                           f"shaded by the BF16 MLP, VM {G}^3")
        torch.cuda.empty_cache()
    if not args.no_proposal:
        prop = bench_proposal(args, cd, ca, G, 64)
        prop128 = bench_proposal(args, cd, ca, G, 128)
        for p in (prop, prop128):
            p["train_speedup_vs_this_config"] = round(p["train_rays_per_s"] / value, 2)
            p["render_speedup_vs_this_config"] = round(p["render_rays_per_s"] / render_value, 2)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": f"plain TensoRF training step: R={R} rays x N={N} stratified "
                               f"samples, k={k} shaded, VM {G}^3 (Cd={cd}, Ca={ca}), "
                               f"{args.precision} MLP, Adam",
                   "rays_per_step": R, "samples_per_ray": N, "appearance_samples_per_ray": k,
                   "grid": G, "parallelism": "single GPU"},
        "render_rays_per_s": round(render_value, 1),
        "final_mse": mse,
        "stages_ms": stages,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "proposal": prop,
        "proposal_128": prop128,
        "topk64_bf16": topk,
        "unbounded_fisheye_bf16": unb,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
