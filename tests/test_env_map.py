"""Environment map light: the oracle's env-map sampling pinned by a furnace closed form and by the
agreement of two independent unbiased estimators, and GPU parity of the path tracer with an
environment map (alone and with emissive and analytic lights, with and without manifold NEE)."""
import numpy as np
import pytest

import oracle_lib as O
from parity import color_close
from test_pathtracer import _look_down, _topology_equal


def _ground_only():
    from pltfalcor_amd.scene import Scene
    proc = Scene(None, seed=2, sphere_count=0, subdivision=1, material_kinds=1)
    a = proc.numpy()
    return Scene.from_arrays(None, a["positions"][:4], a["normals"][:4], a["indices"][:2], a["tri_mesh"][:2],
                             a["mesh_material"][:1], a["materials"], a["profiles"], proc.arrays.camera)


def _patchy_sky(w=64, h=32):
    """A dim sky with a bright patch above the horizon towards -z (u = 0.5) and a coloured band."""
    img = np.full((h, w, 3), 0.2, np.float32)
    img[4:10, 26:38] = (8.0, 6.0, 4.0)
    img[12:16, :] = (0.1, 0.3, 0.6)
    return img


def _mean(osc, frames, w=16, h=16, **kw):
    acc = np.zeros((w * h, 4))
    for f in range(frames):
        acc += osc.render(w, h, frame=f, do_rr=0, threads=8, **kw)
    return acc / frames


# The two estimators are taken where the reference's path weighting leaves them unbiased: light
# sampling at a vertex that ends the path (maxBounces 0), since a vertex's record carries the pdf of
# its own continuation sample, which the solve pass applies to that vertex's light samples too;
# BSDF sampling through one bounce (maxBounces 1), the sky reached by the continuation ray.


def test_env_furnace_ground_is_half_the_sky():
    """Under a constant sky of radiance L a Lambertian ground of (white-basis) albedo 0.5 reflects
    0.5 L, by light sampling of the sky alone and by BSDF sampling alone."""
    s = _ground_only()
    s.set_env_map(np.full((16, 32, 3), 0.8, np.float32))
    a = s.numpy()
    down = _look_down(dict(a), at=(0.3, 0.2), height=4.0)
    up = dict(a)
    cam = down["camera"].copy()
    cam[0:3], cam[3:6], cam[6:9] = [0.0, 1.0, 0.0], [0.0, 5.0, 0.0], [0.0, 0.0, 1.0]
    up["camera"] = cam
    sky = _mean(O.OracleScene(up), 64, max_bounces=0)[:, :3].mean(0)
    assert np.all(sky > 0)
    od = O.OracleScene(down)
    for kw in (dict(max_bounces=0, do_mis=0), dict(max_bounces=1, do_nee=0)):
        ground = _mean(od, 128, **kw)[:, :3].mean(0)
        # 2.5 %: about three standard errors of the blue channel's spectral noise
        assert np.allclose(ground / sky, 0.5, rtol=0.025), (kw, ground / sky)


def test_env_importance_sampling_agrees_with_bsdf_sampling():
    """A sky with a bright patch (rotated 30 degrees): light sampling alone (importance-sampled
    directions weighted by their pdf) and BSDF sampling alone are both unbiased, so their images
    agree. Pins the sampler, its pdf and the lookup against each other; light sampling has the
    lower variance on this sky."""
    s = _ground_only()
    s.set_env_map(_patchy_sky(), rotation=(0.0, 30.0, 0.0))
    osc = O.OracleScene(_look_down(s.numpy(), at=(0.3, 0.2), height=4.0))
    nee = _mean(osc, 96, max_bounces=0, do_mis=0)[:, :3]
    bsdf = _mean(osc, 192, max_bounces=1, do_nee=0)[:, :3]
    rel = np.abs(bsdf.mean(0) - nee.mean(0)) / nee.mean(0)
    assert (rel < 0.03).all(), (bsdf.mean(0), nee.mean(0))
    f = lambda k, **kw: osc.render(16, 16, frame=k, do_rr=0, threads=8, **kw)[:, :3].sum(1)  # noqa: E731
    v_nee = np.var([f(k, max_bounces=0, do_mis=0).mean() for k in range(24)])
    v_bsdf = np.var([f(k, max_bounces=1, do_nee=0).mean() for k in range(24)])
    assert v_nee < v_bsdf


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(name="env-only", opts={"doMNEE": 0, "useEmissiveLights": 0}, lights=False),
    dict(name="env+emissive+analytic", opts={"doMNEE": 0}, lights=True),
    dict(name="env+all-mnee", opts={"doMNEE": 1, "maxBounces": 8}, lights=True),
    dict(name="env-noMIS", opts={"doMNEE": 0, "doMIS": 0}, lights=False),
], ids=lambda c: c["name"])
def test_gpu_env_map_matches_oracle(gpu, case):
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from test_lights import _all_lights
    from test_pathtracer import OPTION_KW
    ctx = default_context()
    proc = Scene(None, seed=31, sphere_count=16, subdivision=3, material_kinds=8)
    pa = proc.numpy()
    s = Scene.from_arrays(ctx, pa["positions"], pa["normals"], pa["indices"], pa["tri_mesh"], pa["mesh_material"],
                          pa["materials"], pa["profiles"], proc.arrays.camera)
    s.set_env_map(_patchy_sky(), rotation=(0.0, 45.0, 0.0), intensity=0.5)
    if case["lights"]:
        s.set_lights(_all_lights())
    a = s.numpy()
    w, h = 64, 48
    opts = {"maxBounces": 6, "HWSS": 4}
    opts.update(case["opts"])
    p = PLTPT(ctx, opts)
    p.set_scene(s)
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    p.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    g_rec = p.bounce_records(w, h)
    g_col = color.cpu().numpy().reshape(-1, 4)
    kw = {OPTION_KW[k]: int(v) for k, v in opts.items() if k in OPTION_KW}
    o_col, o_rec = O.render(a, w, h, frame=0, max_bounces=int(opts["maxBounces"]), records=True, **kw)
    assert np.array_equal(g_rec[0, :3], o_rec[0, :3])
    topo = _topology_equal(g_rec, o_rec)
    assert topo.mean() >= 0.99, topo.mean()
    close = color_close(g_col, o_col)
    assert close[topo].mean() >= 0.995, close[topo].mean()
    gm, om = g_col[:, :3].mean(), o_col[:, :3].mean()
    assert np.isfinite(g_col).all() and om > 0
    assert abs(gm - om) <= 0.02 * abs(om), (gm, om)
    p.close()
