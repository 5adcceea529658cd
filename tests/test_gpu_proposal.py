"""GPU parity of the proposal sampler (tfh_prop_mlp_*, tfh_proposal_*,
tfh_interval_loss, tfh_segment_probabilities_bwd, EDGES sampling and
tfh_train_step_proposal) against the FP64 proposal oracle."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import proposal_oracle as po
from oracle import tensorf_oracle as orc
from tests.gpu_common import dev, f32, log, report

pytestmark = pytest.mark.gpu

BOX = np.array([[-1.5] * 3, [1.5] * 3])


def _prop_params(g=16, c=4, seed=0, std=0.5, bias_std=0.2):
    rng = np.random.default_rng(seed)
    p = po.init_prop_params(g, c, rng, std=std)
    p["pb1"] = rng.normal(0, bias_std, 64)
    p["pb2"] = rng.normal(0, bias_std, 64)
    p["pb3"] = np.array([0.4])
    return {k: f32(v) for k, v in p.items()}


def _device_field(pp):
    from tensorf_amd import proposal, tensor_vm
    vm = tensor_vm.TensorVM.from_reference_layout(dev(pp["prop_vector"]), dev(pp["prop_matrix"]))
    mlp = {k: dev(pp["p" + k]) for k in proposal.MLP_KEYS}
    return proposal.ProposalField(vm, mlp)


def _sorted_edges(rng, r, n, lo=1.0, hi=5.0):
    e = np.sort(rng.uniform(lo, hi, (r, n + 1)), axis=1)
    e[:, 0], e[:, -1] = lo, hi
    return f32(e)


def test_prop_mlp_fwd_bwd():
    log("case: proposal mlp fwd/bwd M=3001 F=24")
    from tensorf_amd import proposal
    pp = _prop_params(c=8)
    rng = np.random.default_rng(1)
    feats = f32(rng.normal(0, 0.5, (3001, 24)))
    dz = f32(rng.normal(size=3001))
    field = _device_field(pp)
    z, sig = proposal.mlp_forward(field.mlp, dev(feats))
    zr, sr, cache = po.prop_mlp_forward(pp, feats)
    ez = report("z", z.cpu().numpy(), zr)
    es = report("sigma", sig.cpu().numpy(), sr)
    assert ez.max() <= 2e-5 * max(1.0, np.abs(zr).max())
    assert es.max() <= 2e-5 * max(1.0, np.abs(sr).max())
    grads, gf = proposal.mlp_backward(field.mlp, dev(feats), dev(dz))
    gr, gfr = po.prop_mlp_backward(pp, cache, dz)
    for k in proposal.MLP_KEYS:
        e = report(f"d{k}", grads[k].cpu().numpy(), gr["p" + k])
        assert e.max() <= 1e-4 * np.abs(gr["p" + k]).max() + 1e-6
    e = report("dfeat", gf.cpu().numpy(), gfr)
    assert e.max() <= 1e-4 * np.abs(gfr).max()


@pytest.mark.parametrize("n,n_out,pad,power", [(37, 23, 0.01, 1.0), (128, 64, 0.0, 1.0),
                                               (200, 300, 0.05, 0.5), (50, 20, 0.01, 0.0),
                                               (64, 64, 0.64, 1.0)])
def test_resample(n, n_out, pad, power):
    log(f"case: resample n={n} n_out={n_out} pad={pad} power={power}")
    from tensorf_amd import proposal
    rng = np.random.default_rng(n)
    r = 97
    e = _sorted_edges(rng, r, n)
    w = rng.exponential(size=(r, n)) * (rng.uniform(size=(r, n)) < 0.3)
    w = f32(w / (w.sum(1, keepdims=True) * 1.1 + 1e-12))
    got = proposal.resample(dev(e), dev(w), n_out, pad, power).cpu().numpy().astype(np.float64)
    want = po.resample(e, w, n_out, pad, power=power)
    err = report("edges", got, want)
    # the FP32 CDF carries ~n ulps of its total; an edge moves by that error
    # over the local histogram height (>= pad / n) times the interval width
    assert err.max() <= 1e-4 * (e[:, -1] - e[:, 0]).max()
    assert np.all(np.diff(got, axis=1) >= 0)
    np.testing.assert_array_equal(got[:, 0], e[:, 0])
    np.testing.assert_array_equal(got[:, -1], e[:, -1])


def test_interval_loss():
    log("case: interval loss nf=64 n=128")
    from tensorf_amd import proposal
    rng = np.random.default_rng(3)
    r, nf, n = 128, 64, 128
    pe = _sorted_edges(rng, r, n)
    fe = _sorted_edges(rng, r, nf)
    fe[:8, 10:20] = pe[:8, 30:40]  # exact shared edges on some rays
    fe = np.sort(fe, axis=1)
    fw = f32(rng.uniform(0, 0.05, (r, nf)))
    pw = f32(rng.uniform(0, 0.02, (r, n)) * (rng.uniform(size=(r, n)) < 0.6))
    loss, grad = proposal.interval_loss(dev(fe), dev(fw), dev(pe), dev(pw), 0.5)
    lr, gr = po.interval_loss(fe, fw, pe, pw, 0.5)
    e = report("loss", loss.cpu().numpy(), lr)
    assert e.max() <= 1e-5 * np.abs(lr).max()
    e = report("grad", grad.cpu().numpy(), gr)
    assert e.max() <= 1e-5 * np.abs(gr).max()


def test_segment_probabilities_bwd():
    log("case: segment probabilities bwd R=64 N=300")
    from tensorf_amd import _lib, render
    rng = np.random.default_rng(4)
    r, n = 64, 300
    sig = f32(rng.exponential(3.0, (r, n)) * (rng.uniform(size=(r, n)) < 0.4))
    st = f32(rng.uniform(0.001, 0.02, (r, n)))
    gw, ge = f32(rng.normal(size=(r, n))), f32(rng.normal(size=(r, n)))
    sp = render.compute_segment_probabilities(dev(sig), dev(st))
    gsig = torch.empty((r, n), dtype=torch.float32, device="cuda")
    args = [dev(sig), dev(st), sp.p_exits, sp.p_terminates, dev(ge), dev(gw)]
    _lib.call("tfh_segment_probabilities_bwd", _lib.ptr(args[0]), _lib.ptr(args[1]), r, n,
              _lib.ptr(args[2]), _lib.ptr(args[3]), _lib.ptr(args[4]), _lib.ptr(args[5]),
              _lib.ptr(gsig), _lib.stream())
    torch.cuda.synchronize()
    E, W = orc.compute_segment_probabilities(sig, st)
    want = po.segment_probabilities_vjp(gw, ge, W, E) * st
    e = report("grad sigma", gsig.cpu().numpy(), want)
    assert e.max() <= 1e-4 * np.abs(want).max()


def _rays(r, seed):
    o, d = orc.make_rays(r, seed=seed, radius=4.0, target_half=0.6)
    return f32(o), f32(d)


def _sample(pp, cfg, o, d, jit, contraction=False, base_ts=None, base_steps=None, power=1.0):
    from tensorf_amd import proposal
    field = _device_field(pp)
    box = BOX if not contraction else np.array([[-2.0] * 3, [2.0] * 3])
    keep = [dev(box), dev(o), dev(d), dev(jit) if jit is not None else None,
            dev(base_ts) if base_ts is not None else None,
            dev(base_steps) if base_steps is not None else None]
    pcfg = proposal.ProposalConfig(cfg.samples_round0, cfg.samples_round1, cfg.samples_out,
                                   cfg.histogram_pad, cfg.loss_weight,
                                   pp["prop_vector"].shape[2], pp["prop_vector"].shape[1])
    s = proposal.Sampler(field, pcfg, *keep, contraction=contraction, power=power)
    torch.cuda.synchronize()
    return s, field, box, keep


@pytest.mark.parametrize("contraction,n0,n1,nout,power,pad", [
    (False, 32, 16, 16, 1.0, 0.02), (True, 32, 16, 16, 1.0, 0.02),
    (False, 128, 64, 64, 1.0, 0.02), (False, 128, 64, 64, 0.3, 0.02),
    (True, 256, 96, 48, 1.0, 0.02), (False, 128, 64, 64, 1.0, 0.64)])
def test_proposal_sample(contraction, n0, n1, nout, power, pad):
    log(f"case: proposal sample contraction={contraction} rounds {n0} -> {n1} -> {nout} "
        f"power={power} pad={pad}")
    cfg = po.ProposalConfig(samples_round0=n0, samples_round1=n1, samples_out=nout,
                            scene_contraction=contraction, histogram_pad=pad,
                            weight_power=power)
    pp = _prop_params()
    r = 96
    o, d = _rays(r, 5)
    jit = f32(np.random.default_rng(5).uniform(size=(r, n0 + n1)))
    bt = bs = None
    if contraction:
        bt, bs = orc.contraction_schedule(n0, 0.05, 200.0)
    s, field, box, keep = _sample(pp, cfg, o, d, jit, contraction, bt, bs, power)
    gpu_rounds = [s.round(k) for k in range(2)]
    gpu_edges = [gpu_rounds[k][0].cpu().numpy().astype(np.float64) for k in range(2)]
    # weights at the GPU's own intervals
    fixed_out, rounds = po.proposal_rounds(pp, cfg, box, o, d, jit, bt, bs, gpu_edges)
    for k in range(2):
        e = report(f"round {k} weights", gpu_rounds[k][1].cpu().numpy(), rounds[k]["w"])
        assert e.max() <= 1e-4
    span = (gpu_edges[0][:, -1] - gpu_edges[0][:, 0]).max()
    e = report("edges out (GPU round edges)", s.edges.cpu().numpy(), fixed_out)
    assert e.max() <= 1e-3 * span
    # end to end against the oracle's own rounds
    out, orounds = po.proposal_rounds(pp, cfg, box, o, d, jit, bt, bs)
    e = report("round 0 edges", gpu_edges[0], orounds[0]["edges"])
    assert e.max() <= 1e-5 * max(span, 1.0)
    e = report("edges out", s.edges.cpu().numpy(), out)
    assert np.quantile(e, 0.999) <= 1e-3 * span
    got = s.edges.cpu().numpy()
    assert np.all(np.diff(got, axis=1) >= 0)


# field weights: "half" of the intervals non-zero, "sparse" (1 in 64: most
# proposal rows then have a zero loss gradient and are compacted away) or
# "none" (no row left: zero loss and gradients)
@pytest.mark.parametrize("n0,n1,nout,r,fw_kind", [
    (32, 16, 16, 128, "half"), (128, 64, 64, 48, "half"), (40, 0, 24, 64, "half"),
    (128, 64, 64, 300, "sparse"), (32, 16, 16, 64, "none")])
def test_proposal_loss_bwd(n0, n1, nout, r, fw_kind):
    log(f"case: proposal loss bwd rounds {n0} -> {n1} -> {nout}, R={r}, field weights {fw_kind}")
    from tensorf_amd import proposal
    cfg = po.ProposalConfig(samples_round0=n0, samples_round1=n1, samples_out=nout,
                            histogram_pad=0.02, loss_weight=0.7)
    pp = _prop_params(seed=6)
    o, d = _rays(r, 6)
    jit = f32(np.random.default_rng(6).uniform(size=(r, n0 + n1)))
    s, field, box, keep = _sample(pp, cfg, o, d, jit)
    fe = s.edges
    rng = np.random.default_rng(7)
    frac = {"half": 0.5, "sparse": 1.0 / 64, "none": 0.0}[fw_kind]
    fw = f32(rng.uniform(0, 0.3, (r, nout)) * (rng.uniform(size=(r, nout)) < frac) / (nout / 16))
    grads = proposal.ProposalField(
        type(field.tensor)(type(field.tensor.stacked_single_vm)(
            torch.zeros_like(field.tensor.stacked_single_vm.plane),
            torch.zeros_like(field.tensor.stacked_single_vm.line))),
        {k: torch.zeros_like(v) for k, v in field.mlp.items()})
    fw_d = dev(fw)
    loss = s.loss_backward(fe, fw_d, grads)
    torch.cuda.synchronize()
    gpu_edges = [s.round(k)[0].cpu().numpy().astype(np.float64) for k in range(1 + (n1 > 0))]
    fe_np = fe.cpu().numpy().astype(np.float64)
    _, rounds = po.proposal_rounds(pp, cfg, box, o, d, jit, None, None, gpu_edges)
    lref, gref = po.proposal_loss_backward(pp, rounds, cfg, fe_np, fw)
    log(f"  loss gpu={loss.item():.8e} oracle={lref:.8e}")
    assert abs(loss.item() - lref) <= 1e-4 * abs(lref) + 1e-12
    gv, gm = grads.tensor.to_reference_layout()
    got = {"prop_vector": gv.cpu().numpy(), "prop_matrix": gm.cpu().numpy()}
    got.update({"p" + k: v.cpu().numpy() for k, v in grads.mlp.items()})
    for k in po.PROP_KEYS:
        e = report(f"grad {k}", got[k], gref[k])
        assert e.max() <= 2e-3 * np.abs(gref[k]).max() + 1e-9, k


def test_render_edges_mode():
    log("case: render on explicit interval edges N=48 k=48")
    from tensorf_amd import _lib, render
    from tests.test_gpu_render import _gpu_objects
    shape = orc.MlpShape()
    from tests.gpu_common import oracle_params
    p = oracle_params(16, 16, 48, shape, seed=3, std=0.2, bias_std=0.1)
    r, n = 128, 48
    o, d = _rays(r, 9)
    rng = np.random.default_rng(9)
    t0, t1 = orc.ray_segment_from_bounding_box(o, d, BOX)
    u = np.sort(rng.uniform(size=(r, n + 1)), axis=1)
    u[:, 0], u[:, -1] = 0.0, 1.0
    edges = f32(t0[:, None] + u * (t1 - t0)[:, None])
    jit = f32(rng.uniform(size=(r, n)))
    cams = np.zeros(r)
    lp, rays, mlp = _gpu_objects(shape, p, o, d, cams, False)
    cfg = render.RenderConfig(0.05, 200.0, render.RenderMode.RGB, n, n)
    ccfg = render.c_config(cfg, False, 0)
    ccfg.sampling = _lib.SAMPLING_EDGES
    dvm, avm = lp.density_tensor.c_struct(), lp.appearance_tensor.c_struct()
    cm = mlp.c_struct(lp.appearance_mlp_params, 144)
    need = int(_lib.lib.tfh_render_workspace_size(ctypes.byref(ccfg), ctypes.byref(cm), 48, r))
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    out = torch.empty((r, 3), dtype=torch.float32, device="cuda")
    keep = [dev(BOX), dev(edges), dev(jit), torch.zeros(r, dtype=torch.int32, device="cuda")]
    _lib.call("tfh_render_fwd", ctypes.byref(dvm), ctypes.byref(avm), ctypes.byref(cm),
              ctypes.byref(ccfg), _lib.ptr(keep[0]), _lib.ptr(rays.origins),
              _lib.ptr(rays.directions), _lib.ptr(keep[3]), r, _lib.ptr(keep[2]),
              _lib.ptr(keep[1]), None, None, _lib.ptr(out), _lib.ptr(ws), ws.numel(),
              _lib.stream())
    torch.cuda.synchronize()
    ocfg = orc.RenderConfig(0.05, 200.0, orc.RGB, n, n, sampling=orc.SAMPLING_EDGES)
    want, _ = orc.render_rays(p, BOX, o, d, cams, ocfg, shape, jit, edges=edges)
    e = report("rgb", out.cpu().numpy(), want)
    assert e.max() <= 2e-4


def test_train_step_proposal():
    log("case: train step with proposal")
    from tensorf_amd import cameras, proposal, render, training
    shape = orc.MlpShape()
    g, cd, ca, r = 16, 16, 48, 256
    pcfg = proposal.ProposalConfig(samples_round0=32, samples_round1=16, samples_out=24,
                                   histogram_pad=0.02, grid_dim=12, channels=4, anneal_steps=0)
    cfg = training.TensorfConfig(minibatch_size=r, n_iters=100, initial_aabb_min=(-1.5,) * 3,
                                 initial_aabb_max=(1.5,) * 3, appearance_feat_dim=ca,
                                 density_feat_dim=cd, feature_n_freqs=2, viewdir_n_freqs=2,
                                 grid_dim_init=g, grid_dim_final=g, upsamp_iters=(),
                                 appearance_samples_per_ray=24, proposal=pcfg)
    st = training.TrainState.initialize(cfg, g, prng_key=0, num_cameras=1)
    lp = st.learnable_params
    dv, dm = lp.density_tensor.to_reference_layout()
    av, am = lp.appearance_tensor.to_reference_layout()
    mp = lp.appearance_mlp_params["params"]
    p = {"density_vector": dv, "density_matrix": dm, "appearance_vector": av,
         "appearance_matrix": am, "w0": mp["Dense_0"]["kernel"], "w1": mp["Dense_1"]["kernel"],
         "b1": mp["Dense_1"]["bias"], "w2": mp["Dense_2"]["kernel"], "b2": mp["Dense_2"]["bias"],
         "w3": mp["Dense_3"]["kernel"], "b3": mp["Dense_3"]["bias"]}
    p = {k: v.detach().cpu().numpy().astype(np.float64) for k, v in p.items()}
    pv, pm = lp.proposal_field.tensor.to_reference_layout()
    pp = {"prop_vector": pv, "prop_matrix": pm}
    pp.update({"p" + k: v for k, v in lp.proposal_field.mlp.items()})
    pp = {k: (v.detach().cpu().numpy() if torch.is_tensor(v) else v).astype(np.float64)
          for k, v in pp.items()}
    o, d = _rays(r, 8)
    colors = f32(np.random.default_rng(8).uniform(size=(r, 3)))
    batch = cameras.RenderedRays(dev(colors), cameras.Rays3D(
        dev(o), dev(d), torch.zeros(r, dtype=torch.int32, device="cuda")))
    seed = render._seed(st.prng_key, 0)
    jit = render.uniform(seed, r * (48 + 24)).cpu().numpy().astype(np.float64)
    st, logs = st.training_step(batch)
    torch.cuda.synchronize()
    ocfg = orc.RenderConfig(0.05, 200.0, orc.RGB, 24, 24, sampling=orc.SAMPLING_EDGES)
    opc = po.ProposalConfig(32, 16, 24, False, 0.02, 1.0)
    newp, newpp, _, _, ologs, _, oedges = po.train_step_proposal(
        p, pp, orc.init_opt_state(p), orc.init_opt_state(pp), 0, BOX, o, d, np.zeros(r), colors,
        ocfg, opc, shape, jit[:r * 48].reshape(r, 48), jit[r * 48:].reshape(r, 24), n_iters=100)
    mse, pl = float(logs["mse"].item()), float(logs["prop_loss"].item())
    log(f"  mse gpu={mse:.8f} oracle={ologs['mse']:.8f}; prop loss gpu={pl:.6e} "
        f"oracle={ologs['prop_loss']:.6e}")
    assert abs(mse - ologs["mse"]) <= 1e-4 * ologs["mse"]
    assert abs(pl - ologs["prop_loss"]) <= 2e-2 * ologs["prop_loss"] + 1e-7
    lp = st.learnable_params
    dv, dm = lp.density_tensor.to_reference_layout()
    pv, pm = lp.proposal_field.tensor.to_reference_layout()
    got = {"density_matrix": (dm.cpu().numpy(), p, newp, 0.02),
           "prop_matrix": (pm.cpu().numpy(), pp, newpp, 0.02),
           "pw2": (lp.proposal_field.mlp["w2"].cpu().numpy(), pp, newpp, 1e-3)}
    for key, (val, before, after, lr) in got.items():
        e = report(f"update {key}", val - before[key], after[key] - before[key])
        frac_bad = np.mean(e > 1e-3 * lr)
        log(f"  fraction of {key} updates off by >1e-3 lr: {frac_bad:.2e}")
        assert frac_bad < 2e-2


def _params_from_state(st):
    lp = st.learnable_params
    dv, dm = lp.density_tensor.to_reference_layout()
    av, am = lp.appearance_tensor.to_reference_layout()
    mp = lp.appearance_mlp_params["params"]
    p = {"density_vector": dv, "density_matrix": dm, "appearance_vector": av,
         "appearance_matrix": am, "w0": mp["Dense_0"]["kernel"], "w1": mp["Dense_1"]["kernel"],
         "b1": mp["Dense_1"]["bias"], "w2": mp["Dense_2"]["kernel"], "b2": mp["Dense_2"]["bias"],
         "w3": mp["Dense_3"]["kernel"], "b3": mp["Dense_3"]["bias"]}
    p = {k: v.detach().cpu().numpy().astype(np.float64) for k, v in p.items()}
    pv, pm = lp.proposal_field.tensor.to_reference_layout()
    pp = {"prop_vector": pv, "prop_matrix": pm}
    pp.update({"p" + k: v for k, v in lp.proposal_field.mlp.items()})
    pp = {k: v.detach().cpu().numpy().astype(np.float64) for k, v in pp.items()}
    return p, pp


def test_train_steps_proposal_follow_oracle():
    """Five fused proposal training steps with the weight-power ramp and a
    transparent start (density shift -10) track the oracle's trajectory."""
    log("case: 5 proposal train steps vs oracle")
    from tensorf_amd import cameras, proposal, render, training
    shape = orc.MlpShape()
    g, r, steps = 16, 256, 5
    pcfg = proposal.ProposalConfig(samples_round0=32, samples_round1=16, samples_out=24,
                                   histogram_pad=0.02, grid_dim=12, channels=4, anneal_steps=3)
    cfg = training.TensorfConfig(minibatch_size=r, n_iters=100, initial_aabb_min=(-1.5,) * 3,
                                 initial_aabb_max=(1.5,) * 3, appearance_feat_dim=48,
                                 density_feat_dim=16, feature_n_freqs=2, viewdir_n_freqs=2,
                                 grid_dim_init=g, grid_dim_final=g, upsamp_iters=(),
                                 appearance_samples_per_ray=24, proposal=pcfg,
                                 density_shift=-10.0)
    st = training.TrainState.initialize(cfg, g, prng_key=4, num_cameras=1)
    p, pp = _params_from_state(st)
    ost, opst = orc.init_opt_state(p), orc.init_opt_state(pp)
    ocfg = orc.RenderConfig(0.05, 200.0, orc.RGB, 24, 24, sampling=orc.SAMPLING_EDGES,
                            density_shift=-10.0)
    opc = po.ProposalConfig(32, 16, 24, False, 0.02, 1.0)
    for step in range(steps):
        o, d = _rays(r, 20 + step)
        colors = f32(0.5 + 0.4 * d)
        batch = cameras.RenderedRays(dev(colors), cameras.Rays3D(
            dev(o), dev(d), torch.zeros(r, dtype=torch.int32, device="cuda")))
        jit = render.uniform(render._seed(st.prng_key, step), r * (48 + 24)).cpu().numpy()
        jit = jit.astype(np.float64)
        st, logs = st.training_step(batch)
        torch.cuda.synchronize()
        opc.weight_power = pcfg.weight_power(step)
        p, pp, ost, opst, ologs, _, _ = po.train_step_proposal(
            p, pp, ost, opst, step, BOX, o, d, np.zeros(r), colors, ocfg, opc, shape,
            jit[:r * 48].reshape(r, 48), jit[r * 48:].reshape(r, 24), n_iters=100)
        mse, pl = float(logs["mse"].item()), float(logs["prop_loss"].item())
        log(f"  step {step} power {opc.weight_power:.2f}: mse gpu={mse:.8f} "
            f"oracle={ologs['mse']:.8f}; prop loss gpu={pl:.6e} oracle={ologs['prop_loss']:.6e}")
        assert abs(mse - ologs["mse"]) <= 1e-3 * ologs["mse"]
        assert abs(pl - ologs["prop_loss"]) <= 5e-2 * ologs["prop_loss"] + 1e-7
