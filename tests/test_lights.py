"""Analytic lights (point, directional, distant, rect, disc, sphere): the oracle's light samplers
pinned by closed forms of the irradiance they deliver to a Lambertian ground with NEE only, and
GPU parity of the path tracer with analytic lights (alone, mixed with emissive geometry, with
manifold NEE) against the oracle.

Radiance scale: a rect light and the emissive quad of the same place and spectral profile deliver
the same irradiance (the emissive quad's white RGB basis is 1 to 0.08 % over the visible range),
so every other shape is compared with the rect light's image, pixel block by pixel block."""
import math

import numpy as np
import pytest

import oracle_lib as O
from parity import color_close
from test_pathtracer import _camera_rays, _lambert_polygon_irradiance, _look_down, _topology_equal

EMISSION_PROFILE = 2  # the procedural scene's blackbody emission profile


def _ground_scene(with_quad_light: bool):
    """The procedural ground (and optionally its 3 x 3 emissive quad at y = 7) as a host scene."""
    from pltfalcor_amd.scene import Scene
    proc = Scene(None, seed=2, sphere_count=0, subdivision=1, material_kinds=1)
    a = proc.numpy()
    ntri = 4 if with_quad_light else 2
    s = Scene.from_arrays(None, a["positions"][: 8 if with_quad_light else 4], a["normals"][: 8 if with_quad_light else 4],
                          a["indices"][:ntri], a["tri_mesh"][:ntri], a["mesh_material"][: 2 if with_quad_light else 1],
                          a["materials"], a["profiles"], proc.arrays.camera)
    return s


def _rect_over_ground(half=1.5, height=7.0):
    from pltfalcor_amd.lights import RectLight
    r = RectLight()
    c, s_ = 0.0, 1.0  # rotate +90 degrees about x: local +z (the emitting side) -> world -y
    rot = np.array([[1, 0, 0, 0], [0, c, -s_, 0], [0, s_, c, 0], [0, 0, 0, 1]], np.float32)
    tr = np.eye(4, dtype=np.float32)
    tr[1, 3] = height
    r.transform = tr @ rot
    r.scaling = (half, half, 1.0)
    r.setIntensity(EMISSION_PROFILE)
    return r


def _render_mean(a, frames, w=16, h=16, **kw):
    acc = np.zeros((w * h, 4))
    for f in range(frames):
        acc += O.render(a, w, h, frame=f, max_bounces=0, do_rr=0, is_emitters=0, do_mis=0, threads=8, **kw)
    return acc / frames


def _blocks(img, w=16, h=16):
    return img[:, :3].sum(1).reshape(4, h // 4, 4, w // 4).sum((1, 3))


def _setup(lights, with_quad=False):
    s = _ground_scene(with_quad)
    if lights:
        s.set_lights(lights)
    return _look_down(s.numpy(), at=(0.3, 0.2), height=4.0)


def test_rect_light_equals_the_emissive_quad_of_the_same_place():
    frames = 48
    quad = _render_mean(_setup([], with_quad=True), frames)
    rect_light = _rect_over_ground()
    assert abs(rect_light.data.surface_area - 9.0) < 1e-5
    rect = _render_mean(_setup([rect_light]), frames)
    ratio = _blocks(rect) / _blocks(quad)
    assert np.abs(ratio - 1).max() < 0.04, ratio
    # and its spatial falloff is Lambert's polygon irradiance
    cam = _setup([rect_light])["camera"]
    o, d = _camera_rays(cam, 16, 16)
    pts = o + (-o[1] / d[:, 1])[:, None] * d
    poly = [np.array(p, float) for p in ([-1.5, 7, -1.5], [1.5, 7, -1.5], [1.5, 7, 1.5], [-1.5, 7, 1.5])]
    E = np.array([_lambert_polygon_irradiance(p, np.array([0, 1.0, 0]), poly) for p in pts])
    shape = _blocks(rect) / _blocks(np.c_[np.repeat(E[:, None], 3, 1), np.zeros(len(E))])
    assert np.abs(shape / shape.mean() - 1).max() < 0.04


def test_distant_disc_and_sphere_lights_follow_their_closed_forms():
    """Distant light at the zenith: the sample's radiance is not divided by its pdf (the reference
    treats a distant light's intensity as the irradiance it delivers, LightHelpers.slang:187-191),
    so E = L E[cos] = L (1 + c) / 2 over the cone of half-angle cosine c. Disc of radius 1.5 at y = 7: E by quadrature
    of cos cos / d^2. Sphere light: the reference's area formula over the true area, times
    Nusselt's pi L (R/d)^2 cos. Each against the rect light's image E_rect."""
    from pltfalcor_amd.lights import DiscLight, DistantLight, SphereLight
    frames = 48
    rect = _render_mean(_setup([_rect_over_ground()]), frames)
    cam = _setup([])["camera"]
    o, d = _camera_rays(cam, 16, 16)
    pts = o + (-o[1] / d[:, 1])[:, None] * d
    poly = [np.array(p, float) for p in ([-1.5, 7, -1.5], [1.5, 7, -1.5], [1.5, 7, 1.5], [-1.5, 7, 1.5])]
    E_rect = np.array([_lambert_polygon_irradiance(p, np.array([0, 1.0, 0]), poly) for p in pts])

    def check(img, E, tol):
        """Per 4 x 4-pixel block within `tol` (both images carry Monte Carlo noise), the image level
        within 2.5 %."""
        ratio = _blocks(img) / _blocks(rect)
        want = (E.reshape(4, 4, 4, 4).sum((1, 3))) / (E_rect.reshape(4, 4, 4, 4).sum((1, 3)))
        assert np.abs(ratio / want - 1).max() < tol, (ratio / want)
        level = _blocks(img).sum() / _blocks(rect).sum() / (E.sum() / E_rect.sum())
        assert abs(level - 1) < 0.025, level

    dl = DistantLight()
    dl.direction = (0.0, -1.0, 0.0)
    dl.setSourceSolidAngle(0.05)
    dl.setIntensity(EMISSION_PROFILE)
    c = dl.data.cos_subtended_angle
    check(_render_mean(_setup([dl]), frames), np.full(256, (1 + c) / 2), 0.04)

    disc = DiscLight()
    disc.transform = _rect_over_ground().transform
    disc.scaling = (1.5, 1.5, 1.0)
    disc.setIntensity(EMISSION_PROFILE)
    assert abs(disc.data.surface_area - math.pi * 2.25) < 1e-4
    m = 200
    rr, tt = np.meshgrid((np.arange(m) + 0.5) / m, (np.arange(m) + 0.5) / m * 2 * math.pi)
    qx, qz = 1.5 * np.sqrt(rr) * np.cos(tt), 1.5 * np.sqrt(rr) * np.sin(tt)
    E_disc = []
    for p in pts:
        dx, dz, dy = qx - p[0], qz - p[2], 7.0 - p[1]
        d2 = dx * dx + dy * dy + dz * dz
        E_disc.append(float(((dy / np.sqrt(d2)) ** 2 / d2).mean() * math.pi * 2.25))
    check(_render_mean(_setup([disc]), frames), np.array(E_disc), 0.04)

    sph = SphereLight()
    tr = np.eye(4, dtype=np.float32)
    tr[0, 3], tr[1, 3], tr[2, 3] = 0.0, 6.0, 0.0
    sph.transform = tr
    sph.scaling = (0.5, 0.5, 0.5)
    sph.setIntensity(EMISSION_PROFILE)
    A_true = 4 * math.pi * 0.25
    E_sph = []
    for p in pts:
        v = np.array([0.0, 6.0, 0.0]) - p
        dd = np.linalg.norm(v)
        E_sph.append(sph.data.surface_area / A_true * math.pi * (0.5 / dd) ** 2 * (v[1] / dd))
    # half of the sphere's samples face away and are rejected: more frames for the same noise
    check(_render_mean(_setup([sph]), 4 * frames), np.array(E_sph), 0.1)


def test_point_and_directional_lights_deliver_nothing():
    """Their samples have pdf 0 (a Dirac), which generateLightSample rejects (pdf > 1e-6): the
    reference's PLT path tracer renders them black."""
    from pltfalcor_amd.lights import DirectionalLight, PointLight
    p = PointLight()
    p.position = (0.0, 5.0, 0.0)
    p.setIntensity(EMISSION_PROFILE)
    dl = DirectionalLight()
    dl.direction = (0.1, -1.0, 0.0)
    dl.setIntensity(EMISSION_PROFILE)
    img = _render_mean(_setup([p, dl]), 4)
    assert np.all(img[:, :3] == 0)


def test_mixed_emissive_and_analytic_lights_add_up():
    """Emissive quad + a rect light elsewhere: each type is chosen with probability 1/2 and its
    sample weighted by 2, so the image is the sum of the single-light images."""
    frames = 64
    r = _rect_over_ground()
    tr = r.transform
    tr[0, 3] = 4.0  # move the rect light 4 units along +x
    r.transform = tr
    both = _render_mean(_setup([r], with_quad=True), frames)
    quad = _render_mean(_setup([], with_quad=True), frames)
    rect = _render_mean(_setup([r]), frames)
    ratio = _blocks(both) / (_blocks(quad) + _blocks(rect))
    assert np.abs(ratio - 1).max() < 0.05, ratio


# ------------------------------------------------------------------------------ GPU


def _gpu_scene(ctx, kinds, seed, lights):
    from pltfalcor_amd.scene import Scene
    proc = Scene(None, seed=seed, sphere_count=16, subdivision=3, material_kinds=kinds)
    a = proc.numpy()
    s = Scene.from_arrays(ctx, a["positions"], a["normals"], a["indices"], a["tri_mesh"], a["mesh_material"],
                          a["materials"], a["profiles"], proc.arrays.camera)
    s.set_lights(lights)
    return s


def _all_lights():
    from pltfalcor_amd.lights import DiscLight, DistantLight, PointLight, SphereLight
    r = _rect_over_ground(1.0, 6.0)
    tr = r.transform
    tr[0, 3] = -3.0
    r.transform = tr
    disc = DiscLight()
    t2 = _rect_over_ground().transform
    t2[0, 3], t2[2, 3] = 3.0, 2.0
    disc.transform = t2
    disc.setIntensity(EMISSION_PROFILE)
    sph = SphereLight()
    t3 = np.eye(4, dtype=np.float32)
    t3[:3, 3] = (0.0, 5.0, -3.0)
    sph.transform = t3
    sph.scaling = (0.4, 0.4, 0.4)
    sph.setIntensity(EMISSION_PROFILE)
    dl = DistantLight()
    dl.direction = (0.3, -1.0, 0.2)
    dl.setSourceSolidAngle(0.02)
    dl.setIntensity(EMISSION_PROFILE)
    pl = PointLight()
    pl.position = (0.0, 4.0, 0.0)
    pl.setIntensity(EMISSION_PROFILE)
    return [r, disc, sph, dl, pl]


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(name="analytic+emissive", kinds=4, seed=21, opts={"doMNEE": 0}),
    dict(name="analytic+emissive-mnee", kinds=8, seed=22, opts={"doMNEE": 1, "maxBounces": 8}),
    dict(name="analytic-only", kinds=4, seed=23, opts={"doMNEE": 0, "useEmissiveLights": 0}),
    dict(name="emissive-only-with-analytic-off", kinds=4, seed=24, opts={"doMNEE": 0, "useAnalyticLights": 0}),
], ids=lambda c: c["name"])
def test_gpu_analytic_lights_match_oracle(gpu, case):
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from test_pathtracer import OPTION_KW
    ctx = default_context()
    s = _gpu_scene(ctx, case["kinds"], case["seed"], _all_lights())
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
    kw["analytic"] = int(opts.get("useAnalyticLights", 1))
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
