"""The learned BSDF surrogate: the CPU restatement of the MLP pinned by finite differences and by
torch's bfloat16 rounding; on the GPU, the matrix-core GEMM, forward pass and gradients against
that restatement, the exact training samples against the FP64 oracle's material evaluation, the
exact fallback, and what training gains."""
import ctypes as C

import numpy as np
import pytest

import mlp_ref as R
import oracle_lib as O


def test_mlp_restatement_gradient_matches_finite_differences():
    """Float64 back-propagation of the restated MLP equals central finite differences of its loss."""
    rng = np.random.default_rng(3)
    p = R.he_init(1)
    X = rng.uniform(-1, 1, (12, 22))
    T = rng.normal(0, 1, (12, 16))
    _, g = R.loss_and_grad(p, X, T)
    idx = rng.choice(p.size, 300, replace=False)
    h = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += h
        lp, _ = R.loss_and_grad(q, X, T)
        q[k] -= 2 * h
        lm, _ = R.loss_and_grad(q, X, T)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_bf16_rounding_matches_torch():
    import torch
    x = np.random.default_rng(0).normal(0, 10, 100000).astype(np.float32)
    x[:10] = [0.0, -0.0, 1.0, 1.00390625, 1.001953125, 3.0e-39, -7.5, 65504.0, 1e30, 2.0 ** -130]
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(R.bf16(x), ref)


def test_target_transform_round_trip():
    from pltfalcor_amd.surrogate import mueller_to_targets, targets_to_mueller
    M = np.random.default_rng(1).normal(0, 1, (50, 16))
    M[:, 0] = np.abs(M[:, 0]) + 0.1
    assert np.allclose(targets_to_mueller(mueller_to_targets(M)), M, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------------------------ GPU


def _ctx_with_drude():
    """The default context (the synthetic Drude conductor's n, k profiles at ids 0 and 1)."""
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.spectra import drude_profiles
    n_prof, k_prof = drude_profiles()
    return default_context(), bytes(n_prof) + bytes(k_prof)


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K", [(16, 128, 32), (100, 70, 64), (64, 64, 128), (300, 16, 256), (128, 128, 4096)])
def test_gpu_mfma_gemm_matches_numpy(gpu, M, N, K):
    """C = A Bt^T on the matrix cores equals the float64 product of the same bf16 operands (their
    products are exact in FP32; only the FP32 accumulation differs)."""
    import torch
    from pltfalcor_amd import _lib
    g = torch.Generator().manual_seed(M * 7 + N * 13 + K)
    A = torch.randn((M, K), generator=g).to(torch.bfloat16)
    B = torch.randn((N, K), generator=g).to(torch.bfloat16)
    Ad, Bd = A.to("cuda"), B.to("cuda")
    Cd = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().plt_mlp_gemm_nt(M, N, K, C.c_void_p(Ad.data_ptr()), C.c_void_p(Bd.data_ptr()),
                                           C.c_void_p(Cd.data_ptr()), None))
    torch.cuda.synchronize()
    ref = A.double().numpy() @ B.double().numpy().T
    got = Cd.cpu().numpy()
    assert np.all(np.abs(got - ref) <= 1e-5 * np.sqrt(K) * (np.abs(A.double().numpy()) @ np.abs(B.double().numpy()).T) + 1e-6)


@pytest.mark.gpu
def test_gpu_surrogate_forward_and_gradients_match_restatement(gpu):
    """Forward outputs, loss and every parameter's gradient of the matrix-core MLP equal the CPU
    restatement with bf16 rounding at the same places (up to FP32 accumulation order, which can
    flip an occasional bf16 rounding)."""
    import torch
    from pltfalcor_amd.surrogate import Surrogate
    ctx, _ = _ctx_with_drude()
    s = Surrogate(ctx, max_batch=512, seed=5)
    p = R.he_init(7).astype(np.float32)
    s.set_params(p)
    assert np.array_equal(s.get_params(), p)
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, (448, 22)).astype(np.float32)
    T = rng.normal(0, 1, (448, 16)).astype(np.float32)
    Xd, Td = torch.tensor(X, device="cuda"), torch.tensor(T, device="cuda")
    Y = s.forward(Xd).cpu().numpy()
    Yr, _ = R.forward(p.astype(np.float64), X.astype(np.float64), R.bf16)
    scale = np.abs(Yr).max()
    assert np.abs(Y - Yr).max() <= 1e-2 * scale
    assert np.abs(Y - Yr).mean() <= 1e-3 * scale
    g, loss = s.gradients(Xd, Td)
    lr, gr = R.loss_and_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64), R.bf16)
    assert abs(loss - lr) <= 1e-3 * lr
    o = 0
    for name, shape in R.SHAPES:
        k = int(np.prod(shape))
        a, b = g[o:o + k], gr[o:o + k]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"{name}: relative gradient error {err:.2e}")
        assert err <= 2e-2, name
        o += k
    # a train step moves the parameters along -gradient (Adam's first step is lr * sign(g))
    s.train_step(Xd, Td, 1e-3)
    d = s.get_params() - p
    big = np.abs(gr) > 1e-3 * np.abs(gr).max()
    assert np.mean(np.sign(d[big]) == -np.sign(gr[big])) > 0.99
    s.close()


@pytest.mark.gpu
def test_gpu_surrogate_samples_match_fp64_oracle(gpu):
    """The generator's materials, inputs and exact Mueller matrices: each column equals the FP64
    oracle's evaluation of the same binary16 material for that unit Stokes vector."""
    import torch
    from pltfalcor_amd.surrogate import ExactSamples
    from pltfalcor_amd import materials as PM
    ctx, prof = _ctx_with_drude()
    n = 2048
    S = ExactSamples(ctx, n, seed=3)
    mats = S.materials.cpu().numpy()
    geo = S.geometry.cpu().numpy()
    X = S.X.cpu().numpy()
    M = S.M.cpu().numpy()
    assert np.isfinite(X).all() and np.isfinite(M).all()
    fam = X[:, 10:13].argmax(1)
    assert np.all(X[:, 10:13].sum(1) == 1)
    types = mats[:, :4].copy().view(np.uint32)[:, 0]
    want = np.array([PM.COATED_OPAQUE_DIELECTRIC, PM.COATED_CONDUCTOR, PM.GRATED_CONDUCTOR])[fam]
    assert np.array_equal(types, want)
    assert all((fam == f).mean() > 0.28 for f in range(3))
    # the derived lobe input is the log of k l |wi_t + wo_t| of the sample's own geometry
    lam_um = geo[6] * 1e-3
    ht = np.hypot(geo[0] + geo[3], geo[1] + geo[4])
    l_um = X[:, 8] * .41 + .46
    assert np.allclose(X[:, 13], (np.log(2 * np.pi / lam_um * l_um * ht + 1e-3) - 1) / 3, atol=2e-5)
    bad = 0
    for i in range(0, n, 7):
        wi, wo = geo[0:3, i].astype(np.float64), geo[3:6, i].astype(np.float64)
        lobe = int(geo[7, i:i + 1].view(np.uint32)[0])
        lam = np.full(4, geo[6, i], np.float64)
        Mo = np.zeros((4, 4))
        for c in range(4):
            st = np.zeros(16)
            st[c] = 1.0
            out, _, _ = O.material_eval(bytes(mats[i]), prof, 2, wi, wo, lobe, lam, st, float(geo[8, i]), True)
            Mo[:, c] = out[0]
        tol = 2e-3 * max(abs(Mo[0, 0]), 1e-12) + 1e-9
        bad += np.abs(M[i].reshape(4, 4) - Mo).max() > tol
    assert bad <= 0.01 * len(range(0, n, 7)), bad


@pytest.mark.gpu
def test_gpu_surrogate_out_of_distribution_falls_back_to_exact(gpu):
    """Inputs outside the training box are evaluated exactly (bit-identical to the exact mode) and
    counted; inside the box the network is used."""
    import torch
    from pltfalcor_amd.surrogate import ExactSamples, Surrogate
    ctx, _ = _ctx_with_drude()
    S = ExactSamples(ctx, 1000, seed=9)
    s = Surrogate(ctx, max_batch=256, seed=1)
    rng = np.random.default_rng(2)
    Sin = torch.tensor(np.c_[np.ones(1000), rng.uniform(-.5, .5, (1000, 3))].astype(np.float32), device="cuda")
    ymu, ysd = np.zeros(16), np.ones(16)
    X = S.X.cpu().numpy()
    # box that excludes the samples with wi.z < 0.5
    lo, hi = X.min(0) - 1, X.max(0) + 1
    lo[2] = 0.5
    s.set_stats(ymu, ysd, lo, hi)
    out, fb = s.eval(S, Sin, mode=0)
    ex, fb1 = s.eval(S, Sin, mode=1)
    ood = X[:, 2] < 0.5
    assert fb == ood.sum() and fb1 == 0
    o, e = out.cpu().numpy(), ex.cpu().numpy()
    assert np.array_equal(o[ood], e[ood])
    assert not np.array_equal(o[~ood], e[~ood])
    s.close()


@pytest.mark.gpu
def test_gpu_surrogate_training_reduces_error(gpu):
    """A short training run on exact samples (matrix-core forward/backward + Adam) lowers the
    training loss and predicts held-out radiance (S0 for random partially polarised incident
    light) much better than the best input-independent predictor (each target's mean)."""
    import torch
    from pltfalcor_amd.surrogate import ExactSamples, Surrogate, train, mueller_to_targets
    ctx, _ = _ctx_with_drude()
    tr = ExactSamples(ctx, 1 << 18, seed=21)
    te = ExactSamples(ctx, 1 << 14, seed=22)
    s = Surrogate(ctx, max_batch=4096, seed=3)
    losses = train(s, tr, epochs=8, batch=4096, lr=2e-3)
    print("losses", losses)
    assert losses[-1] < 0.7 * losses[0]
    rng = np.random.default_rng(4)
    n = te.n
    Sin = torch.tensor(np.c_[np.ones(n), rng.uniform(-.5, .5, (n, 3))].astype(np.float32), device="cuda")
    out, fb = s.eval(te, Sin, mode=0)
    ex, _ = s.eval(te, Sin, mode=1)
    o, e = out.cpu().numpy(), ex.cpu().numpy()
    ok = e[:, 0] > 1e-6
    rel = np.abs(o[ok, 0] - e[ok, 0]) / e[ok, 0]
    Mtr = tr.M.cpu().numpy()
    good = Mtr[:, 0] > 1e-12
    Mb = np.sinh(mueller_to_targets(Mtr[good]).mean(0))
    Mb[0] = np.exp(mueller_to_targets(Mtr[good]).mean(0)[0])
    Mb[1:] *= Mb[0]
    base = np.einsum("j,ij->i", Mb[:4], Sin.cpu().numpy())
    rel_base = np.abs(base[ok] - e[ok, 0]) / e[ok, 0]
    print(f"held-out S0 relative error: surrogate median {np.median(rel):.4f} mean {rel.mean():.4f}; "
          f"mean-target baseline median {np.median(rel_base):.4f}; fallbacks {fb}")
    assert np.median(rel) < 0.5 * np.median(rel_base)
    assert fb < 0.01 * n
    s.close()


def _render_frames(ctx, scene, w, h, opts, surrogate=None, mode=0, frames=2, solve_split=True):
    """Colour of `frames` frames of the pass (split solve pass), with the surrogate in `mode` or
    exact; and the pass's surrogate counts of the last frame."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    p = PLTPT(ctx, opts)
    p.set_sample_pattern(PLTPT.HALTON, 16)
    if solve_split:
        p.set_launch_config(solve_schedule=PLTPT.SOLVE_SPLIT)
    p.set_scene(scene)
    if surrogate is not None:
        p.set_surrogate(surrogate, mode)
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    out = []
    for _ in range(frames):
        color.fill_(float("nan"))
        p.execute(w, h, color, albedo, normal)
        torch.cuda.synchronize()
        out.append(color.cpu().numpy().copy())
    stats = p.surrogate_stats() if surrogate is not None else None
    p.close()
    return np.stack(out), stats


def _coated_scene(ctx):
    from pltfalcor_amd.scene import Scene
    # eight material kinds: the coated opaque dielectric, grated conductor and coated conductor among them
    return Scene(ctx, seed=7, sphere_count=24, subdivision=3, material_kinds=8)


@pytest.mark.gpu
def test_gpu_surrogate_in_render_outside_box_equals_exact(gpu):
    """With a training box that no input lies in, every step the surrogate would stand for falls
    back to the exact lobe: the render (query pass, inference, apply pass) is bit-identical to the
    exact split pass, every such step appends one row per HWSS wavelength and is counted as exact."""
    from pltfalcor_amd.surrogate import Surrogate
    ctx, _ = _ctx_with_drude()
    s = Surrogate(ctx, max_batch=4096, seed=3)
    s.set_stats(np.zeros(16), np.ones(16), np.full(22, 1e30), np.full(22, -1e30))
    scene = _coated_scene(ctx)
    w, h, opts = 96, 64, {"maxBounces": 6, "HWSS": 4}
    ref, _ = _render_frames(ctx, scene, w, h, opts)
    got, st = _render_frames(ctx, scene, w, h, opts, surrogate=s)
    print("surrogate counts", st)
    assert st["exact_steps"] > 0 and st["rows"] == 4 * st["exact_steps"]
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("hwss", [4, 1])
def test_gpu_surrogate_teacher_render_matches_exact(gpu, hwss):
    """The in-render plumbing apart from the network: with the exact lobe's matrix taken in the
    network's frames (incident tangent perp_stark(-wi), exitant frame of the exact evaluation), its
    coherence area mapped from the beam's coherence and encoded as the network's outputs, in place
    of the network's outputs, the render equals the exact one up to rounding -- a wrong frame, area,
    family or lobe, or a slip in the decoding, shows as percent-level differences on the coated and
    grated spheres."""
    from pltfalcor_amd.surrogate import Surrogate
    ctx, _ = _ctx_with_drude()
    s = Surrogate(ctx, max_batch=4096, seed=3)
    s.set_stats(np.full(16, 0.25), np.full(16, 1.5), np.full(22, -1e30), np.full(22, 1e30))
    scene = _coated_scene(ctx)
    w, h, opts = 96, 64, {"maxBounces": 6, "HWSS": hwss}
    ref, _ = _render_frames(ctx, scene, w, h, opts, frames=3)
    got, st = _render_frames(ctx, scene, w, h, opts, surrogate=s, mode=1, frames=3)
    print("teacher counts", st)
    assert st["rows"] > 0 and st["exact_steps"] == 0
    a, b = ref[..., :3].astype(np.float64), got[..., :3].astype(np.float64)
    assert np.isfinite(b).all()
    floor = 1e-4 * np.abs(a).mean()
    rel = np.abs(a - b) / (np.abs(a) + floor)
    print(f"teacher vs exact: rel quantiles 50/99/99.9/max {np.quantile(rel, [.5, .99, .999])} {rel.max():.3e}; "
          f"mean {a.mean():.6f} {b.mean():.6f}")
    assert np.quantile(rel, 0.999) < 1e-3 and abs(a.mean() - b.mean()) <= 1e-4 * abs(a.mean())
    s.close()


@pytest.mark.gpu
def test_gpu_surrogate_network_in_render_against_exact(gpu):
    """A/B of the trained network against the exact lobes in the render: a surrogate trained on
    exact samples of the three families renders the same frames; the network answers the steps
    inside its training box (most of the eligible ones), the image stays finite, and its error
    against the exact render is of the size of the network's held-out error on those lobes
    (reported), not larger."""
    from pltfalcor_amd.surrogate import ExactSamples, Surrogate, train
    ctx, _ = _ctx_with_drude()
    tr = ExactSamples(ctx, 1 << 19, seed=21)
    s = Surrogate(ctx, max_batch=8192, seed=3)
    train(s, tr, epochs=4, batch=8192, lr=2e-3)
    scene = _coated_scene(ctx)
    w, h, opts = 96, 64, {"maxBounces": 6, "HWSS": 4}
    ref, _ = _render_frames(ctx, scene, w, h, opts, frames=4)
    got, st = _render_frames(ctx, scene, w, h, opts, surrogate=s, frames=4)
    n_steps = st["rows"] // 4
    print("network counts", st, f"network share of eligible steps {1 - st['exact_steps'] / max(1, n_steps):.3f}")
    assert np.isfinite(got).all() and (got[..., 3] == 1).all()
    assert st["exact_steps"] < 0.5 * n_steps
    a, b = ref[..., :3].mean(0), got[..., :3].mean(0)
    err = np.abs(a - b).mean() / np.abs(a).mean()
    print(f"network render: mean |difference| / mean {err:.4f}; image means exact {a.mean():.5f} network {b.mean():.5f}")
    assert err < 0.5
    s.close()


@pytest.mark.gpu
def test_gpu_surrogate_prepare_matches_host_pipeline(gpu):
    """plt_surrogate_prepare against the same pipeline on the host (float64): the valid rows (finite,
    M00 > 1e-12) in row order, the targets' mean and (population) standard deviation divided by the
    output weights, the valid inputs' box, and the standardised targets."""
    from pltfalcor_amd.surrogate import OUTPUT_SCALE, ExactSamples, Surrogate, mueller_to_targets, prepare
    ctx, _ = _ctx_with_drude()
    S = ExactSamples(ctx, 1 << 16, seed=31)
    s = Surrogate(ctx, max_batch=1024, seed=1)
    T, idx, nv = prepare(s, S)
    M = S.M.cpu().numpy().astype(np.float64)
    X = S.X.cpu().numpy()
    ok = np.isfinite(M).all(1) & (M[:, 0] > 1e-12)
    assert nv == ok.sum() and nv > 0.5 * len(ok)
    assert np.array_equal(idx.cpu().numpy()[:nv], np.nonzero(ok)[0])
    y = mueller_to_targets(M[ok])
    mean, sd = y.mean(0), y.std(0)
    sd = np.where(sd > 1e-6, sd, 1.0) / OUTPUT_SCALE
    ym, ys, lo, hi = s.stats
    assert np.allclose(ym, mean, rtol=1e-6, atol=1e-6) and np.allclose(ys, sd, rtol=1e-6)
    assert np.array_equal(lo, X[ok].min(0)) and np.array_equal(hi, X[ok].max(0))
    Tg = T.cpu().numpy()[ok]
    assert np.allclose(Tg, (y - mean) / sd, rtol=1e-5, atol=1e-5)
    s.close()


@pytest.mark.gpu
def test_gpu_surrogate_graph_training_equals_eager(gpu):
    """plt_surrogate_train with one step captured as a graph and replayed gives the parameters and
    losses of the same steps launched eagerly, bit for bit (every step-dependent value lives in
    device state), and the epoch order is a permutation that changes with the epoch."""
    import ctypes as C
    import torch
    from pltfalcor_amd import _lib
    from pltfalcor_amd.surrogate import ExactSamples, Surrogate, train
    ctx, _ = _ctx_with_drude()
    S = ExactSamples(ctx, 1 << 15, seed=32)
    runs = []
    for graph in (True, False):
        s = Surrogate(ctx, max_batch=2048, seed=5)
        losses = train(s, S, epochs=3, batch=2048, lr=2e-3, seed=7, graph=graph)
        runs.append((s.get_params(), losses))
        s.close()
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]
    assert runs[0][1][-1] < runs[0][1][0]
    lib = _lib.load()
    n = 30001
    orders = []
    for epoch in (0, 1):
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        _lib.check(lib.plt_surrogate_shuffle(n, 7, epoch, C.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        assert np.array_equal(np.sort(o), np.arange(n))
        orders.append(o)
    assert (orders[0] != orders[1]).mean() > 0.9
