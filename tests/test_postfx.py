"""AccumulatePass and ToneMapper: the numpy restatement pinned by closed forms and invariants,
the host-side option logic of the library (no GPU), and on the GPU the kernels against the
restatement -- bit-exact for accumulation, within a stated tolerance for tone mapping -- plus the
reference graph's chain PLTPT.color -> AccumulatePass (Double) -> ToneMapper."""
import math
import warnings

import numpy as np
import pytest

from oracle import postfx_oracle as P

F32 = np.float32


def _ulp(x):
    return np.spacing(np.abs(np.asarray(x, F32))).astype(np.float64)


# ------------------------------------------------------------------------------ oracle pins


def test_oracle_double_accumulation_is_the_mean_within_two_ulp():
    """Double mode returns the float32 rounding of the exact mean up to the float weight's own
    rounding (Accumulate.cs.slang:157): within 2 ulp of math.fsum(frames) / N for every N."""
    rng = np.random.default_rng(0)
    frames = rng.lognormal(0, 2, (300, 6, 5, 4)).astype(F32)
    acc = P.Accumulator(P.DOUBLE)
    for n in range(1, len(frames) + 1):
        out = acc.execute(frames[n - 1])
        if n in (1, 2, 3, 7, 10, 100, 299, 300):
            exact = np.apply_along_axis(math.fsum, 0, frames[:n].astype(np.float64)) / n
            assert np.all(np.abs(out - exact) <= 2 * _ulp(exact)), n
    first = P.Accumulator(P.DOUBLE).execute(frames[0])
    assert np.array_equal(first, frames[0])


def test_oracle_compensated_summation_stays_exact_where_naive_single_drifts():
    """Kahan summation (Accumulate.cs.slang:123-135) keeps the running mean within a few ulp over
    20000 frames; plain Single summation drifts by more than ten times as much."""
    rng = np.random.default_rng(1)
    frames = (1.0 + 1e-3 * rng.standard_normal((20000, 2, 2, 4))).astype(F32)
    comp, naive = P.Accumulator(P.SINGLE_COMPENSATED), P.Accumulator(P.SINGLE)
    for f in frames:
        oc, on = comp.execute(f), naive.execute(f)
    exact = np.apply_along_axis(math.fsum, 0, frames.astype(np.float64)) / len(frames)
    err_c = np.abs(oc - exact).max() / _ulp(exact).max()
    err_n = np.abs(on - exact).max() / _ulp(exact).max()
    assert err_c <= 2, err_c
    assert err_n >= 10 * max(err_c, 1), (err_n, err_c)


def test_oracle_moving_average_closed_form():
    """maxFrameCount M with overflow EMA: M frames of a, then b forever. After the step the
    output is b + (a - b)(1 - 1/(M+1))^k (the weight stays 1/(M+1), AccumulatePass.cpp:290-293)."""
    M, a, b = 8, 2.0, 5.0
    for prec, tol in ((P.DOUBLE, 2e-7), (P.SINGLE, 2e-6)):
        acc = P.Accumulator(prec, max_frames=M, overflow=P.EMA)
        for _ in range(M):
            out = acc.execute(np.full((1, 1, 4), a, F32))
        assert np.allclose(out, a, rtol=1e-7)
        for k in range(1, 40):
            out = acc.execute(np.full((1, 1, 4), b, F32))
            want = b + (a - b) * (1 - 1 / (M + 1)) ** k
            assert abs(float(out[0, 0, 0]) - want) <= tol * b, (prec, k)
        assert acc.frame_count == M


def test_oracle_frame_counting_and_overflow_modes():
    """Stop keeps the image and the count; Reset restarts; compensated ignores the moving average;
    refresh and a resolution change reset; a disabled pass blits float input without counting but
    counts integer input (AccumulatePass.cpp:144-301)."""
    f = lambda v, h=2, w=3: np.full((h, w, 4), v, F32)
    acc = P.Accumulator(P.SINGLE, max_frames=2, overflow=P.STOP)
    acc.execute(f(1)); acc.execute(f(3))
    assert acc.execute(f(100)) is None and acc.frame_count == 2
    acc = P.Accumulator(P.DOUBLE, max_frames=2, overflow=P.RESET)
    acc.execute(f(1)); acc.execute(f(3))
    assert np.array_equal(acc.execute(f(7)), f(7)) and acc.frame_count == 1
    acc = P.Accumulator(P.SINGLE_COMPENSATED, max_frames=2, overflow=P.EMA)
    for v in (1, 3, 5, 7):
        out = acc.execute(f(v))
    assert np.array_equal(out, f(4)) and acc.frame_count == 4
    acc = P.Accumulator(P.SINGLE)
    acc.execute(f(1)); acc.execute(f(3))
    assert np.array_equal(acc.execute(f(9), refresh=True), f(9))
    acc.execute(f(1))
    assert np.array_equal(acc.execute(f(6, 4, 4)), f(6, 4, 4)) and acc.frame_count == 1
    acc = P.Accumulator(P.SINGLE, auto_reset=False)
    acc.execute(f(1))
    assert np.array_equal(acc.execute(f(3), refresh=True), f(2))
    acc = P.Accumulator(P.SINGLE, enabled=False)
    acc.execute(f(2.5))
    assert acc.frame_count == 0
    acc.execute(np.full((2, 3, 4), 4, np.int32))
    assert acc.frame_count == 1


def test_oracle_luminance_samples_reproduce_a_linear_ramp():
    """Bilinear filtering reproduces a linear function exactly, so the luminance texels of a ramp
    image equal the ramp at the texel-centre sample positions (u = (i + 0.5) W / w - 0.5)."""
    H, W = 23, 37
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    L = 1.0 + 0.01 * x + 0.02 * y
    img = np.stack([L, L, L, np.ones_like(L)], -1)
    lum = P.luminance_image(img)
    assert lum.shape == (16, 32)
    u = (np.arange(32) + 0.5) * W / 32 - 0.5
    v = (np.arange(16) + 0.5) * H / 16 - 0.5
    assert np.allclose(2.0 ** lum, 1.0 + 0.01 * u[None, :] + 0.02 * v[:, None], rtol=1e-12)
    # power-of-two sources are sampled at their own texels
    img2 = np.random.default_rng(2).uniform(0.1, 4, (8, 16, 4))
    assert np.allclose(P.luminance_image(img2), np.log2(img2[..., :3] @ P.LUMA), rtol=1e-12)


@pytest.mark.parametrize("shape", [(1, 1), (1, 64), (32, 1), (16, 16), (64, 256), (512, 8)])
def test_oracle_mip_chain_top_is_the_mean(shape):
    a = np.random.default_rng(sum(shape)).normal(0, 3, shape)
    assert abs(P.mip_top(a) - a.mean()) <= 1e-12 * max(1.0, np.abs(a).max())


def test_oracle_tone_map_closed_forms():
    o = dict(P.TONEMAP_DEFAULTS)
    # auto exposure of a uniform image: the exposed luminance is the key 0.042 (Linear, no clamp)
    img = np.ones((4, 8, 4)) * np.array([0.7, 1.3, 0.2, 0.5])
    out = P.tone_map(img, dict(o, autoExposure=True, operator=0, clamp=False))
    assert np.allclose(out[..., :3] @ P.LUMA, 0.042, rtol=1e-12)
    assert np.array_equal(out[..., 3], img[..., 3])
    # Reinhard maps luminance L to L / (1 + L)
    c = np.random.default_rng(3).uniform(0, 5, (3, 3, 4))
    out = P.tone_map(c, dict(o, operator=1, clamp=False))
    L = c[..., :3] @ P.LUMA
    assert np.allclose(out[..., :3] @ P.LUMA, L / (1 + L), rtol=1e-12)
    # HableUc2 maps whiteScale / 2 (the 2x exposure bias) to exactly 1
    ws = 7.5
    out = P.tone_map(np.full((1, 1, 4), ws / 2), dict(o, whiteScale=ws, clamp=False))
    assert np.allclose(out[..., :3], 1.0, rtol=1e-12)
    # ACES: 0 -> 0, large -> 1 (2.51 / 2.43 > 1 saturates); HejiHableAlu: <= 0.004 -> 0, large -> 1
    out = P.tone_map(np.array([[[0, 1e6, 0, 1]]], float), dict(o, operator=5, clamp=False))
    assert out[0, 0, 0] == 0 and out[0, 0, 1] == 1
    out = P.tone_map(np.array([[[0.004, 1e9, 0, 1]]], float), dict(o, operator=3, clamp=False))
    assert out[0, 0, 0] == 0 and abs(out[0, 0, 1] - 1) < 1e-6
    # manual exposure: ISO 400, f/2, 1/50 s -> 0.01 * 400 / (50 * 4)
    assert np.allclose(P.color_transform(dict(o, filmSpeed=400, fNumber=2, shutter=50)), np.eye(3) * 0.02)


def test_oracle_white_balance_maps_the_source_white_to_the_target_white():
    """At 6500 K the transform is the identity; at T it carries the RGB of T's white to the RGB of
    the 6500 K white (von Kries: LMS of T scaled to LMS of 6500 K)."""
    assert np.allclose(P.white_balance(6500.0), np.eye(3), atol=1e-12)
    for T in (1905.0, 3000.0, 4500.0, 9000.0, 25000.0):
        src = P.XYZ_TO_RGB @ P.color_temperature_xyz(T)
        dst = P.XYZ_TO_RGB @ P.color_temperature_xyz(6500.0)
        assert np.allclose(P.white_balance(T) @ src, dst, rtol=1e-10), T
    # the locus fit: chromaticity of 6500 K near D65 (0.3127, 0.3290)
    xyz = P.color_temperature_xyz(6500.0)
    assert abs(xyz[0] / xyz.sum() - 0.3135) < 2e-3 and abs(xyz[1] / xyz.sum() - 0.3236) < 2e-3


# ------------------------------------------------------------------------------ host logic (no GPU)


def test_accumulate_options_follow_the_reference():
    from pltfalcor_amd.postfx import AccumulatePass, AccumulatePrecision
    from pltfalcor_amd import _lib
    a = AccumulatePass(None, {"enabled": True, "precisionMode": AccumulatePrecision.Double})
    assert a.get_scripting_dictionary() == {"enabled": 1.0, "outputSize": 0.0, "autoReset": 1.0, "precisionMode": 0.0,
                                            "maxFrameCount": 0.0, "overflowMode": 0.0, "outputFormat": 0.0}
    assert AccumulatePass(None).get_option("precisionMode") == AccumulatePrecision.Single
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert not a.set_option("enableAccumulationX", 1)
        assert "Unknown field" in str(w[0].message)
    with pytest.raises(_lib.PLTError):
        a.set_option("precisionMode", 3)
    with pytest.raises(_lib.PLTError):
        a.set_option("outputSize", 1)
    a.maxAccmulatedFrames = 16
    assert a.maxAccmulatedFrames == 16 and a.frame_count == 0
    a.enabled = False
    assert not a.enabled
    with pytest.raises(_lib.PLTError):  # no context: execute needs a device
        import ctypes as C
        _lib.check(a.lib.plt_accumulate_execute(a.handle, C.c_void_p(16), 0, 1, 1, C.c_void_p(16), 0, None))


def test_tonemapper_setters_clamp_and_couple_like_the_reference():
    from pltfalcor_amd.postfx import ToneMapper, ToneMapOp, ExposureMode
    t = ToneMapper(None, {"autoExposure": True, "exposureCompensation": 0.0})
    d = t.get_scripting_dictionary()
    assert d["operator"] == ToneMapOp.HableUc2 and d["autoExposure"] == 1 and d["whiteScale"] == pytest.approx(11.2)
    assert list(d) == ["outputSize", "useSceneMetadata", "exposureCompensation", "autoExposure", "filmSpeed",
                       "whiteBalance", "whitePoint", "operator", "clamp", "whiteMaxLuminance", "whiteScale", "fNumber",
                       "shutter", "exposureMode"]
    t.exposureCompensation = 20
    assert t.exposureCompensation == 12
    t.whitePoint = 1000
    assert t.whitePoint == 1905
    t.whiteScale = -1
    assert t.whiteScale == pytest.approx(0.001)
    t.fNumber = 2
    t.shutter = 50
    assert t.exposureValue == pytest.approx(math.log2(50 * 4), rel=1e-6)
    t.exposureValue = 10  # aperture priority: shutter = 2^EV / f^2
    assert t.shutter == pytest.approx(1024 / 4, rel=1e-6)
    t.exposureMode = ExposureMode.ShutterPriority
    t.exposureValue = 12  # f = sqrt(2^EV / shutter)
    assert t.fNumber == pytest.approx(math.sqrt(4096 / 256), rel=1e-6)
    t.set_scene_metadata(film_iso=800, shutter_speed=None)
    assert t.filmSpeed == 800
    t.set_option("useSceneMetadata", 0)
    t.set_scene_metadata(film_iso=200)
    assert t.filmSpeed == 800


@pytest.mark.parametrize("opts", [dict(), dict(whiteBalance=True, whitePoint=3200.0, exposureCompensation=1.5),
                                  dict(whiteBalance=True, whitePoint=11000.0, autoExposure=True),
                                  dict(filmSpeed=400.0, fNumber=2.8, shutter=60.0, exposureCompensation=-2.0)])
def test_tonemapper_color_transform_matches_oracle(opts):
    from pltfalcor_amd.postfx import ToneMapper
    t = ToneMapper(None, opts)
    o = dict(P.TONEMAP_DEFAULTS, **opts)
    assert np.allclose(t.color_transform(), P.color_transform(o), rtol=2e-6, atol=2e-7)


# ------------------------------------------------------------------------------ GPU


def _gpu_frames(rng, n, shape, kind):
    if kind == "float":
        x = rng.lognormal(0, 2, (n,) + shape + (4,)).astype(F32)
        x[..., 1] *= rng.choice([-1, 1], x[..., 1].shape)
        return x
    if kind == "uint":
        return rng.integers(0, 2 ** 32, (n,) + shape + (4,), dtype=np.uint64).astype(np.uint32)
    return rng.integers(-2 ** 31, 2 ** 31, (n,) + shape + (4,), dtype=np.int64).astype(np.int32)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,max_frames,overflow", [(0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 5, 2), (1, 5, 2),
                                                      (2, 5, 2), (0, 4, 1), (1, 3, 0)])
@pytest.mark.parametrize("kind,shape", [("float", (23, 37)), ("uint", (5, 7)), ("int", (1, 1)), ("float", (135, 240))])
def test_gpu_accumulate_bit_exact(gpu, prec, max_frames, overflow, kind, shape):
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.postfx import AccumulatePass
    rng = np.random.default_rng(prec * 100 + max_frames * 10 + overflow + shape[0])
    frames = _gpu_frames(rng, 12, shape, kind)
    ctx = default_context()
    a = AccumulatePass(ctx, {"precisionMode": prec, "maxFrameCount": max_frames, "overflowMode": overflow})
    o = P.Accumulator(prec, max_frames, overflow)
    dst = torch.full(shape + (4,), float("nan"), dtype=torch.float32, device="cuda")
    last = None
    for i, f in enumerate(frames):
        src = torch.from_numpy(f.view(np.int32) if kind == "uint" else f).to("cuda")
        a.execute(src, dst, refresh=(i == 9), unsigned=(kind == "uint"))
        want = o.execute(f, refresh=(i == 9))
        got = dst.cpu().numpy()
        if want is None:
            assert np.array_equal(got, last, equal_nan=True), i  # Stop: output not written
        else:
            assert np.array_equal(got.view(np.uint32), want.astype(F32).view(np.uint32)), (i, np.abs(got - want).max())
            last = got
        assert a.frame_count == o.frame_count
    a.close()


@pytest.mark.gpu
def test_gpu_accumulate_resize_disable_and_reset(gpu):
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.postfx import AccumulatePass
    rng = np.random.default_rng(7)
    ctx = default_context()
    a = AccumulatePass(ctx, {"precisionMode": 0})
    o = P.Accumulator(P.DOUBLE)
    seq = [((4, 4), None), ((4, 4), None), ((8, 3), None), ((8, 3), "disable"), ((8, 3), None), ((8, 3), "enable"),
           ((8, 3), None), ((8, 3), "reset"), ((8, 3), None)]
    for shape, action in seq:
        if action == "disable":
            a.enabled = False; o.set_enabled(False)
        elif action == "enable":
            a.enabled = True; o.set_enabled(True)
        elif action == "reset":
            a.reset(); o.frame_count = 0
        f = rng.lognormal(0, 1, shape + (4,)).astype(F32)
        got = a.execute(torch.from_numpy(f).cuda()).cpu().numpy()
        assert np.array_equal(got, o.execute(f)), (shape, action)
        assert a.frame_count == o.frame_count
    a.close()


def _tm_options(**kw):
    o = dict(P.TONEMAP_DEFAULTS)
    o.update(kw)
    return o


@pytest.mark.gpu
@pytest.mark.parametrize("op", range(6))
@pytest.mark.parametrize("auto,clamp", [(True, True), (False, False), (True, False)])
@pytest.mark.parametrize("shape", [(1080, 1920), (23, 37), (64, 16), (1, 1), (3, 1000)])
def test_gpu_tone_mapper_matches_oracle(gpu, op, auto, clamp, shape):
    """Within 2e-5 relative (+1e-6 absolute) of the float64 restatement given the library's float
    colour transform (the matrix, computed in float as the reference's float matrices are, is
    compared with the float64 one in its own test: its small off-diagonal entries carry absolute
    errors near 1e-7, which a texel with channels 1000x apart turns into a 1e-4 relative error of
    the smaller channel in both implementations). Reinhard of a black pixel is 0/0 in both (NaN,
    saturated to 0 when clamped)."""
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.postfx import ToneMapper
    rng = np.random.default_rng(op * 7 + shape[0])
    img = rng.lognormal(-1, 1.5, shape + (4,)).astype(F32)
    img[0, 0, :3] = 0.0
    opts = dict(autoExposure=auto, clamp=clamp, operator=op, whiteBalance=op % 2 == 1, whitePoint=4000.0 + 900 * op,
                exposureCompensation=0.5 * (op - 2), whiteMaxLuminance=1.5, whiteScale=9.0, filmSpeed=200.0)
    t = ToneMapper(default_context(), opts)
    ml = torch.zeros(1, dtype=torch.float32, device="cuda")
    got = t.execute(torch.from_numpy(img).cuda(), mean_log_luminance=ml).cpu().numpy().astype(np.float64)
    want_ml = P.mip_top(P.luminance_image(img)) if auto else None
    if auto:
        assert abs(float(ml.item()) - want_ml) <= 2e-5 * max(1.0, abs(want_ml)), (ml.item(), want_ml)
    want = P.tone_map(img, _tm_options(**opts), want_ml, transform=t.color_transform())
    ok = np.isclose(got, want, rtol=2e-5, atol=1e-6, equal_nan=True)
    assert ok.all(), (np.argwhere(~ok)[:5], got[~ok][:5], want[~ok][:5])
    t.close()


@pytest.mark.gpu
def test_gpu_reference_graph_chain(gpu):
    """PLTPT.color of several frames -> AccumulatePass (Double, as scripts/PLTPathTracer.py:19) ->
    ToneMapper (autoExposure, :26): the accumulation is bit-identical to the restatement fed the
    same frames, and the tone-mapped image matches the restatement."""
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.postfx import AccumulatePass, ToneMapper
    from pltfalcor_amd.scene import Scene
    ctx = default_context()
    scene = Scene(ctx, seed=2, sphere_count=6, subdivision=2)
    pt = PLTPT(ctx)  # the reference's defaults
    pt.set_scene(scene)
    W, H = 96, 64
    acc = AccumulatePass(ctx, {"enabled": True, "precisionMode": 0})
    alb = AccumulatePass(ctx, {"enabled": True, "precisionMode": 1})
    tm = ToneMapper(ctx, {"autoExposure": True, "exposureCompensation": 0.0})
    o, oa = P.Accumulator(P.DOUBLE), P.Accumulator(P.SINGLE)
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    out, out_a = torch.empty_like(color), torch.empty_like(color)
    for _ in range(6):
        pt.execute(W, H, color, albedo, normal)
        acc.execute(color, out)
        alb.execute(albedo, out_a)
        want = o.execute(color.cpu().numpy())
        want_a = oa.execute(albedo.cpu().numpy())
        assert np.array_equal(out.cpu().numpy(), want)
        assert np.array_equal(out_a.cpu().numpy(), want_a)
    assert np.isfinite(want).all() and want[..., :3].max() > 0
    shown = tm.execute(out).cpu().numpy()
    ref = P.tone_map(want, _tm_options(autoExposure=True))
    assert np.allclose(shown, ref, rtol=2e-5, atol=1e-6)
    for p in (acc, alb, tm, pt):
        p.close()
