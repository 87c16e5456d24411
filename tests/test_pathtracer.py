"""PLTPT render pass: option handling (CPU), oracle self-checks (CPU), and GPU parity of the
two-pass path tracer against the oracle's render of the same scene, frame and pixels.

GPU parity bar:
  * primary (bounce 0) records: hit triangle ids identical for every pixel, barycentrics
    bit-identical (camera rays and traversal are exact-IEEE on both sides);
  * whole paths: hit triangles, sampled lobes and flags identical at every bounce for >= 99 % of
    the paths (the remainder diverge after a sampled direction differs in its last bits: GPU
    transcendental functions are not correctly rounded, so a later ray can graze a different
    triangle); on those paths the float record fields agree to 1e-3 for >= 99.9 %;
  * colour: on topologically identical paths |GPU - oracle| <= 1e-3 |oracle| + 1e-4 x (the
    image's mean pixel value) per channel for >= 99.5 % of them (float32 operation-order
    differences through up to nine BSDF evaluations), and the image mean over all pixels within 2 %;
  * albedo and view-space normal outputs within 1e-5 on >= 99.9 % of identical paths.
"""
import ctypes as C
import json
import warnings
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from parity import color_close, color_floor

# Reference defaults: the PLTPT.h member initialisers behind the PLTPT.cpp:102-134 keys, read from the
# reference's source text by tests/golden/make_pltpt_defaults.py.
REFERENCE_DEFAULTS = json.loads((Path(__file__).parent / "golden" / "pltpt_defaults.json").read_text())

OPTION_KW = dict(doNEE="do_nee", doMIS="do_mis", doRussianRoulette="do_rr", HWSSDoMIS="hwss_mis",
                 doImportanceSampleEmitters="is_emitters", useDirectLights="direct", useEmissiveLights="emissive",
                 doMNEE="do_mnee", MNEEMaxOccluders="mnee_max_occluders", MNEEMaxIterations="mnee_max_iterations",
                 emissiveSampler="emissive_sampler", sampleGenerator="sample_generator", NEEUsePerTileSG="per_tile_sg",
                 tileSize="tile_size")


def test_pass_options_defaults_and_unknown_keys():
    from pltfalcor_amd.pltpt import PLTPT
    p = PLTPT(None)
    d = p.get_scripting_dictionary()
    assert set(d) == set(REFERENCE_DEFAULTS)
    for k, v in REFERENCE_DEFAULTS.items():
        assert d[k] == pytest.approx(v, rel=1e-7), k
    with pytest.warns(UserWarning, match="Unknown field 'noSuchKey'"):
        assert p.set_option("noSuchKey", 3) is False
    assert p.get_scripting_dictionary() == d
    # HWSS is clamped to [1, 4] (PLTPT.cpp:291-292)
    p.set_option("HWSS", 9)
    assert p.get_option("HWSS") == 4
    p.set_option("HWSS", 0)
    assert p.get_option("HWSS") == 1
    p.set_option("maxBounces", 6)
    assert p.get_option("maxBounces") == 6
    from pltfalcor_amd import _lib
    with pytest.raises(_lib.PLTError):
        p.set_option("maxBounces", float("nan"))
    p.close()


def test_host_only_pass_rejects_scene():
    from pltfalcor_amd import _lib
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    p = PLTPT(None)
    s = Scene(None, seed=1, sphere_count=2, subdivision=1)
    with pytest.raises(_lib.PLTError):
        p.set_scene(s)


def _scene(sphere_count=12, subdivision=3, kinds=4, seed=2):
    from pltfalcor_amd.scene import Scene
    return Scene(None, seed=seed, sphere_count=sphere_count, subdivision=subdivision, material_kinds=kinds).numpy()


def test_oracle_render_deterministic_and_finite():
    a = _scene()
    c1, r1 = O.render(a, 24, 24, frame=3, max_bounces=5, records=True, threads=4)
    c2, r2 = O.render(a, 24, 24, frame=3, max_bounces=5, records=True, threads=1)
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
    assert np.isfinite(c1).all() and (c1[:, :3] >= 0).all() and (c1[:, 3] == 1).all()
    # a sub-range render reproduces the same pixels
    c3 = O.render(a, 24, 24, frame=3, max_bounces=5, first=100, count=50)
    assert np.array_equal(c3, c1[100:150])
    # different frames decorrelate
    c4 = O.render(a, 24, 24, frame=4, max_bounces=5)
    assert not np.array_equal(c1, c4)


def _camera_rays(cam, w, h):
    """Pinhole camera rays in float64 from the camera description (Camera.cpp:187-193,
    Camera.slang:46-78), independent of the oracle."""
    pos, tgt, up = cam[0:3].astype(np.float64), cam[3:6].astype(np.float64), cam[6:9].astype(np.float64)
    fov, aspect, fd = float(cam[9]), float(cam[10]), float(cam[11])
    W = tgt - pos
    W = W / np.linalg.norm(W) * fd
    U = np.cross(W, up)
    U /= np.linalg.norm(U)
    V = np.cross(U, W)
    V /= np.linalg.norm(V)
    U *= fd * np.tan(fov / 2) * aspect
    V *= fd * np.tan(fov / 2)
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    sx = (px.ravel() + 0.5) / w - cam[12]
    sy = (py.ravel() + 0.5) / h + cam[13]
    d = (2 * sx - 1)[:, None] * U + (1 - 2 * sy)[:, None] * V + W
    return pos, d / np.linalg.norm(d, axis=1, keepdims=True)


def test_oracle_primary_hits_match_brute_force_camera():
    """Bounce-0 records equal brute-force intersection of independently built camera rays."""
    a = _scene(sphere_count=6, subdivision=2)
    w = h = 32
    _, rec = O.render(a, w, h, max_bounces=0, records=True)
    o, d = _camera_rays(a["camera"], w, h)
    n = w * h
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = o[:, None]
    rays[3:6] = d.T
    rays[7] = 3.4e38
    hit = O.trace(a["positions"], a["indices"], np.ascontiguousarray(rays), brute=True)
    tri = rec[0, 0]
    agree = tri == hit[0].view(np.uint32)
    # float64 camera rays vs float32: only grazing pixels may differ
    assert agree.mean() > 0.995


def _lambert_polygon_irradiance(x, n, poly):
    """Irradiance at x (normal n) from a unit-radiance Lambertian polygon (Lambert's formula)."""
    E = 0.0
    v = [p - x for p in poly]
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        g = np.cross(a, b)
        theta = np.arccos(np.clip(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)), -1, 1))
        E += theta * np.dot(n, g / np.linalg.norm(g))
    return abs(E) / 2.0


def _mis_weighted_irradiance(x, poly, m=160):
    """Irradiance at x (normal +y) from the unit-radiance horizontal quad `poly`, each light point
    weighted by the balance heuristic between light sampling (area pdf 1/A in solid angle) and
    cosine BSDF sampling (cos/pi), by midpoint quadrature."""
    lo, hi = poly[0], poly[2]
    u = (np.arange(m) + 0.5) / m
    gx, gz = np.meshgrid(lo[0] + u * (hi[0] - lo[0]), lo[2] + u * (hi[2] - lo[2]))
    area = (hi[0] - lo[0]) * (hi[2] - lo[2])
    d = np.stack([gx - x[0], np.full_like(gx, poly[0][1] - x[1]), gz - x[2]], -1)
    r2 = (d ** 2).sum(-1)
    cos = d[..., 1] / np.sqrt(r2)  # same cosine at the receiver (+y) and at the emitter (-y)
    p_light = r2 / (cos * area)
    p_bsdf = cos / np.pi
    return float((cos * cos / r2 * p_light / (p_light + p_bsdf)).mean() * area)


def test_oracle_direct_lighting_matches_lambert_closed_form():
    """maxBounces 0, NEE only: the ground (white-basis albedo 0.5, Lambertian) under the 3 x 3 quad
    light. Without MIS the per-pixel expectation is albedo/pi * E(x) * Le_rgb, with E Lambert's
    closed-form polygon irradiance and Le_rgb the light's own measured colour (the camera sees the
    light from below in a second render). With MIS the light-sampling weight
    p_light / (p_light + p_bsdf) has no complementary BSDF-sampled hit at maxBounces 0, so the
    expectation is the balance-weighted irradiance. Pins camera rays, light sampling (alias table,
    uniform triangle sampling, area-to-solid-angle pdf), shadow rays, the diffuse eval and pdf, the
    MIS weight and the spectral measure."""
    a = _scene(sphere_count=0, subdivision=1, kinds=1)
    frames = 96
    w = h = 16
    cam = a["camera"].copy()
    # look straight down at the ground patch below the light
    cam[0:3] = [0.3, 4.0, 0.2]
    cam[3:6] = [0.3, 0.0, 0.2]
    cam[6:9] = [0.0, 0.0, 1.0]
    cam[9] = np.radians(60.0)
    cam[10] = 1.0
    cam[12:14] = 0.0
    a["camera"] = cam
    o, d = _camera_rays(cam, w, h)
    t = -o[1] / d[:, 1]
    pts = o + t[:, None] * d
    poly = [np.array(p, float) for p in ([-1.5, 7, -1.5], [1.5, 7, -1.5], [1.5, 7, 1.5], [-1.5, 7, 1.5])]
    E = np.array([_lambert_polygon_irradiance(p, np.array([0, 1.0, 0]), poly) for p in pts])
    E_mis = np.array([_mis_weighted_irradiance(p, poly) for p in pts])
    assert np.all(E_mis < E) and np.all(E_mis > 0.9 * E)
    # light colour: look up at the light from below (emission is Lambertian: radiance independent of angle)
    b = dict(a)
    camL = cam.copy()
    camL[0:3] = [0.0, 1.0, 0.0]
    camL[3:6] = [0.0, 7.0, 0.0]
    camL[9] = np.radians(10.0)
    b["camera"] = camL
    Le = np.zeros(3)
    for f in range(frames):
        Le += O.render(b, 8, 8, frame=f, max_bounces=0, do_rr=0, is_emitters=0)[:, :3].mean(0)
    Le /= frames
    for do_mis, Ex in ((0, E), (1, E_mis)):
        acc = np.zeros((w * h, 4))
        for f in range(frames):
            acc += O.render(a, w, h, frame=f, max_bounces=0, do_rr=0, is_emitters=0, do_mis=do_mis, threads=8)
        acc /= frames
        expected = 0.5 / np.pi * Ex[:, None] * Le[None, :]
        got = acc[:, :3]
        rel = np.abs(got.mean(0) - expected.mean(0)) / expected.mean(0)
        # 1.5 %: about three standard errors of the blue channel's spectral noise at 96 frames
        assert (rel < 0.015).all(), (do_mis, got.mean(0), expected.mean(0))
        # the closed-form spatial falloff, over 4 x 4-pixel blocks (per-pixel spectral noise is ~10 %)
        blk = lambda v: v.reshape(4, 4, 4, 4).sum((1, 3))  # noqa: E731
        ratio = blk(got.sum(1).reshape(h, w)) / blk(expected.sum(1).reshape(h, w))
        assert np.abs(ratio - 1).max() < 0.04, (do_mis, np.abs(ratio - 1).max())


def _fan_light_scene():
    """The procedural ground and light materials with the 3 x 3 light polygon split into a fan of four
    triangles of unequal area around an off-centre point (uniform and power selection then differ)."""
    from pltfalcor_amd.scene import Scene
    proc = Scene(None, seed=2, sphere_count=0, subdivision=1, material_kinds=1)
    base = proc.numpy()
    corners = [[-1.5, 7, -1.5], [1.5, 7, -1.5], [1.5, 7, 1.5], [-1.5, 7, 1.5]]
    pos = [[-10, 0, -10], [10, 0, -10], [10, 0, 10], [-10, 0, 10]] + corners + [[0.9, 7, -0.6]]
    nrm = [[0, 1, 0]] * 4 + [[0, -1, 0]] * 5
    idx = [[0, 2, 1], [0, 3, 2]] + [[8, 4 + i, 4 + (i + 1) % 4] for i in range(4)]
    a = Scene.from_arrays(None, np.array(pos, np.float32), np.array(nrm, np.float32), np.array(idx, np.uint32),
                          np.array([0, 0, 1, 1, 1, 1], np.uint32), np.array([0, 1], np.uint32), base["materials"],
                          base["profiles"], proc.arrays.camera).numpy()
    e1, e2 = a["positions"][a["indices"][2:, 1]] - a["positions"][a["indices"][2:, 0]], \
        a["positions"][a["indices"][2:, 2]] - a["positions"][a["indices"][2:, 0]]
    n = np.cross(e1, e2)
    assert np.all(n[:, 1] < 0)  # the fan faces the ground
    areas = 0.5 * np.linalg.norm(n, axis=1)
    assert areas.max() > 1.5 * areas.min()
    return a, corners


def test_oracle_uniform_emissive_sampler_matches_lambert_closed_form():
    """maxBounces 0, NEE only, no MIS, under the fan-split light: both the uniform sampler
    (1/4 per triangle) and the power sampler (area/total) are unbiased for Lambert's polygon
    irradiance, while their images differ sample by sample."""
    a, corners = _fan_light_scene()
    frames, w, h = 64, 16, 16
    a = _look_down(a, at=(0.3, 0.2), height=4.0)
    cam = a["camera"]
    o, d = _camera_rays(cam, w, h)
    pts = o + (-o[1] / d[:, 1])[:, None] * d
    poly = [np.array(p, float) for p in corners]
    E = np.array([_lambert_polygon_irradiance(p, np.array([0, 1.0, 0]), poly) for p in pts])
    b = dict(a)
    camL = cam.copy()
    camL[0:3], camL[3:6], camL[9] = [0.0, 1.0, 0.0], [0.0, 7.0, 0.0], np.radians(10.0)
    b["camera"] = camL
    Le = sum(O.render(b, 8, 8, frame=f, max_bounces=0, do_rr=0, is_emitters=0)[:, :3].mean(0) for f in range(frames)) / frames
    imgs = {}
    for sampler in (0, 2):
        acc = np.zeros((w * h, 4))
        for f in range(frames):
            acc += O.render(a, w, h, frame=f, max_bounces=0, do_rr=0, is_emitters=0, do_mis=0, threads=8,
                            emissive_sampler=sampler)
        acc /= frames
        imgs[sampler] = acc
        expected = 0.5 / np.pi * E[:, None] * Le[None, :]
        rel = np.abs(acc[:, :3].mean(0) - expected.mean(0)) / expected.mean(0)
        assert (rel < 0.02).all(), (sampler, acc[:, :3].mean(0), expected.mean(0))
    assert not np.array_equal(imgs[0], imgs[2])


def test_oracle_alias_table_reproduces_power_pdf():
    """The power sampler's alias table selects each emissive triangle with probability flux/sum."""
    a = _scene(sphere_count=0, subdivision=1, kinds=1)
    thr, redirect, perm, pdf = O.alias_table(a)
    n = len(pdf)
    p = np.zeros(n)
    for i in range(n):
        p[perm[i]] += thr[i] / n
        p[redirect[i]] += (1 - thr[i]) / n
    assert np.allclose(p, pdf, atol=2e-3)  # thresholds are stored as binary16


# ------------------------------------------------------------------------------------ GPU


def _render_pair(gpu_scene, host_arrays, w, h, frame, opts, oracle_threads=8, first=0, count=None, oracle=None,
                 pass_=None):
    """The GPU pass's frame and the oracle's render of pixels [first, first + count) of it: colour,
    bounce records, albedo and view-space normal of both (the GPU's cut to the same pixels)."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.bsdf import default_context
    count = w * h - first if count is None else count
    p = pass_ if pass_ is not None else PLTPT(default_context(), dict(opts))
    p.set_scene(gpu_scene)
    p.set_frame(frame)
    dev = torch.device("cuda")
    color = torch.empty((h, w, 4), dtype=torch.float32, device=dev)
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    p.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    sl = slice(first, first + count)
    g_rec = p.bounce_records(w, h)[:, :, sl]
    g_col = color.cpu().numpy().reshape(-1, 4)[sl]
    g_aov = np.concatenate([albedo.cpu().numpy().reshape(-1, 4)[sl], normal.cpu().numpy().reshape(-1, 4)[sl]], axis=1)
    kw = {OPTION_KW[k]: int(v) for k, v in opts.items() if k in OPTION_KW}
    if "MNEESolverThreshold" in opts:
        kw["mnee_threshold"] = float(opts["MNEESolverThreshold"])
    renderer = oracle if oracle is not None else O.OracleScene(host_arrays)
    o_col, o_rec, o_aov = renderer.render(w, h, frame=frame, max_bounces=int(opts["maxBounces"]), records=True, aov=True,
                                          threads=oracle_threads, first=first, count=count, **kw)
    if oracle is None:
        renderer.close()
    if pass_ is None:
        p.close()
    return g_col, g_rec, o_col, o_rec, g_aov, o_aov


def _path_len(rec):
    """Number of valid records per path (up to and including the miss/terminated one)."""
    flags = rec[:, 6, :]
    end = (flags & (0x10 | 0x20)) != 0
    first_end = np.argmax(end, axis=0)
    return first_end + 1


def _topology_equal(g, o):
    """Paths whose integer record fields (hit triangle, sampled lobe, flags) agree at every bounce."""
    n = g.shape[2]
    L = _path_len(o)
    eq = np.ones(n, bool)
    for p in range(g.shape[0]):
        m = p < L
        for f in (0, 5, 6):
            eq &= ~m | (g[p, f] == o[p, f])
    return eq & (_path_len(g) == L)


_FLOAT_FIELDS = {1: "b1", 2: "b2", 3: "wi.x", 4: "wi.y", 7: "contrib", 8: "lastPdf"}


def _float_fields_close(g, o, topo, rel=1e-3, report=None, bounces=None):
    """Per path: barycentrics, octahedral wi, contribution and last pdf within `rel` (of max(1, |x|)),
    over the first `bounces` vertices (all by default). With `report` (a list), appends (bounce,
    field, paths off, paths compared) for every bounce and field that has any."""
    L = _path_len(o)
    ok = np.ones(g.shape[2], bool)
    for p in range(g.shape[0] if bounces is None else min(bounces, g.shape[0])):
        m = (p < L) & topo
        for f in _FLOAT_FIELDS:
            gv, ov = g[p, f].view(np.float32).astype(np.float64), o[p, f].view(np.float32).astype(np.float64)
            good = ~m | (np.abs(gv - ov) <= rel * np.maximum(1.0, np.abs(ov)))
            if report is not None and not good.all():
                report.append((p, _FLOAT_FIELDS[f], int((~good).sum()), int(m.sum())))
            ok &= good
    return ok


BASE_OPTS = {"maxBounces": 6, "doMNEE": 0, "HWSS": 4}

def _check_parity(name, g_col, g_rec, o_col, o_rec, g_aov, o_aov, topo_min=0.99, color_min=0.995, mean_tol=0.02,
                  deep_fields_min=0.999):
    """The parity bar of the GPU pass against the oracle on the same pixels. The float record fields
    of the first two vertices hold 1e-3 for >= 99.9 % of topologically identical paths; over all
    vertices for >= `deep_fields_min` (the workload scenes' subdivision-5 spheres have triangles a
    few mm across, on which the ulp differences of the device's and the host's sin / cos / sqrt in
    sampled directions, compounded over several bounces, move barycentrics by more than 1e-3)."""
    assert np.array_equal(g_rec[0, 0], o_rec[0, 0]), "primary hit triangle ids differ"
    hit0 = o_rec[0, 0] != 0xFFFFFFFF
    assert np.array_equal(g_rec[0, 1:3, hit0], o_rec[0, 1:3, hit0]), "primary barycentrics differ"
    topo = _topology_equal(g_rec, o_rec)
    print(f"[{name}] topologically identical paths {topo.mean():.5f}")
    assert topo.mean() >= topo_min
    off = []
    fclose = _float_fields_close(g_rec, o_rec, topo, report=off)
    fclose2 = _float_fields_close(g_rec, o_rec, topo, bounces=2)
    print(f"[{name}] float record fields close: first two vertices {fclose2[topo].mean():.5f}, all {fclose[topo].mean():.5f}; "
          f"off (bounce, field, paths, of): {off}")
    floor = color_floor(o_col)
    err = np.abs(g_col[:, :3] - o_col[:, :3])
    close = color_close(g_col, o_col)
    rel = np.max(err / (np.abs(o_col[:, :3]) + floor), axis=1)
    print(f"[{name}] colour within 1e-3 rel + floor {floor:.3g} on identical paths {close[topo].mean():.5f}; "
          f"rel err quantiles 50/99/99.9 % {np.quantile(rel[topo], [.5, .99, .999])}")
    assert close[topo].mean() >= color_min
    assert np.isfinite(g_col).all() and (g_col[:, 3] == 1).all()
    gm, om = g_col[:, :3].mean(0), o_col[:, :3].mean(0)
    print(f"[{name}] mean gpu {gm} oracle {om}")
    assert np.all(np.abs(gm - om) <= mean_tol * np.maximum(om, 1e-3) + 1e-4)
    # albedo (getProperties.albedo) and view-space normal of the first non-delta vertex
    a_ok = np.all(np.abs(g_aov[:, :4] - o_aov[:, :4]) <= 1e-5 * np.maximum(1.0, np.abs(o_aov[:, :4])), axis=1)
    n_ok = np.all(np.abs(g_aov[:, 4:] - o_aov[:, 4:]) <= 1e-5, axis=1)
    print(f"[{name}] albedo close {a_ok[topo].mean():.5f}, view-space normal close {n_ok[topo].mean():.5f}")
    assert a_ok[topo].mean() >= 0.999 and n_ok[topo].mean() >= 0.999
    assert fclose2[topo].mean() >= 0.999 and fclose[topo].mean() >= deep_fields_min
    return topo


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(name="mixed-d6", scene=dict(sphere_count=12, subdivision=3, kinds=4, seed=2), opts={}),
    dict(name="ground-only", scene=dict(sphere_count=0, subdivision=1, kinds=1, seed=2), opts={}),
    dict(name="dielectric-noRR", scene=dict(sphere_count=10, subdivision=3, kinds=1, seed=5), opts={"doRussianRoulette": 0}),
    dict(name="conductor-noMIS", scene=dict(sphere_count=10, subdivision=3, kinds=2, seed=6), opts={"doMIS": 0}),
    dict(name="noNEE", scene=dict(sphere_count=12, subdivision=3, kinds=4, seed=7), opts={"doNEE": 0}),
    dict(name="noIS-noHWSSMIS", scene=dict(sphere_count=12, subdivision=3, kinds=4, seed=8),
         opts={"doImportanceSampleEmitters": 0, "HWSSDoMIS": 0}),
    dict(name="8-materials-d8", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=3), opts={"maxBounces": 8}),
    dict(name="8-materials-noHWSSMIS", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=9),
         opts={"maxBounces": 8, "HWSSDoMIS": 0}),
    dict(name="mnee-8-materials-d8", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=3),
         opts={"maxBounces": 8, "doMNEE": 1}),
    dict(name="mnee-dielectric", scene=dict(sphere_count=12, subdivision=3, kinds=1, seed=5), opts={"doMNEE": 1}),
    dict(name="mnee-noMIS-4occ", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=11),
         opts={"doMNEE": 1, "doMIS": 0, "MNEEMaxOccluders": 4, "MNEEMaxIterations": 30}),
    dict(name="mnee-noHWSSMIS", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=12),
         opts={"doMNEE": 1, "HWSSDoMIS": 0, "MNEESolverThreshold": 1e-4}),
    dict(name="uniform-sampler", scene=dict(sphere_count=12, subdivision=3, kinds=4, seed=13),
         opts={"emissiveSampler": 0}),
    dict(name="uniform-sampler-mnee", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=14),
         opts={"emissiveSampler": 0, "doMNEE": 1}),
    dict(name="tiny-generator", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=15),
         opts={"sampleGenerator": 0, "maxBounces": 8, "doMNEE": 1}),
    dict(name="per-tile-nee-selector", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=16),
         opts={"NEEUsePerTileSG": 1, "tileSize": 16, "doMNEE": 1}),
    dict(name="per-tile-tiny", scene=dict(sphere_count=12, subdivision=3, kinds=4, seed=17),
         opts={"NEEUsePerTileSG": 1, "tileSize": 40, "sampleGenerator": 0}),
    dict(name="grated-opaque-dielectric", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=18),
         opts={"maxBounces": 8, "doMNEE": 1}, variant="grated-opaque"),
    dict(name="camera-polarizer", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=19),
         opts={"maxBounces": 6, "doMNEE": 1}, variant="polarizer"),
    dict(name="partial-polarizer-tiny", scene=dict(sphere_count=24, subdivision=3, kinds=8, seed=20),
         opts={"maxBounces": 6, "sampleGenerator": 0}, variant="half-polarizer"),
], ids=lambda c: c["name"])
def test_gpu_pltpt_matches_oracle(gpu, case):
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    sc = case["scene"]
    s = Scene(default_context(), seed=sc["seed"], sphere_count=sc["sphere_count"], subdivision=sc["subdivision"],
              material_kinds=sc["kinds"])
    a = s.numpy()
    opts = dict(BASE_OPTS, **case["opts"])
    w, h = 64, 48
    if case.get("variant"):
        s, a = _scene_variant(s, a, case["variant"])
    out = _render_pair(s, a, w, h, frame=1, opts=opts)
    _check_parity(case["name"], *out)


@pytest.mark.gpu
@pytest.mark.parametrize("size", [(53, 37), (8, 3), (1, 1)], ids=["53x37", "8x3", "1x1"])
def test_gpu_frame_sizes_off_the_8x8_blocks(gpu, size):
    """Paths run in 8 x 8 pixel blocks (8-row bands, column by column; a shorter last band): frames
    whose sides are not multiples of 8 still give every pixel exactly the oracle's path (colour,
    records returned in pixel order, albedo and normal), with the split solve pass and manifold
    NEE, and the wavefront schedule writes the megakernel's records."""
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    w, h = size
    s = Scene(default_context(), seed=3, sphere_count=24, subdivision=3, material_kinds=8, aspect=w / h)
    a = s.numpy()
    opts = dict(BASE_OPTS, maxBounces=6, doMNEE=1)
    p = PLTPT(default_context(), dict(opts))
    p.set_launch_config(solve_schedule=PLTPT.SOLVE_SPLIT)
    g_col, g_rec, o_col, o_rec, g_aov, o_aov = _render_pair(s, a, w, h, frame=1, opts=opts, pass_=p)
    if w * h >= 1000:  # enough paths for the statistical part of the bar
        _check_parity(f"frame {w}x{h}", g_col, g_rec, o_col, o_rec, g_aov, o_aov)
    else:  # every pixel's camera ray and first hit, exactly
        assert np.array_equal(g_rec[0, 0], o_rec[0, 0])
        hit0 = o_rec[0, 0] != 0xFFFFFFFF
        assert np.array_equal(g_rec[0, 1:3, hit0], o_rec[0, 1:3, hit0])
    recs = []
    for sched in (PLTPT.MEGAKERNEL, PLTPT.WAVEFRONT):
        p.set_schedule(sched)
        p.set_frame(1)
        color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        albedo, normal = torch.empty_like(color), torch.empty_like(color)
        p.execute(w, h, color, albedo, normal)
        torch.cuda.synchronize()
        recs.append((p.bounce_records(w, h), color.cpu().numpy()))
    p.close()
    (r0, c0), (r1, c1) = recs
    L = _path_len(r0)
    valid = np.arange(r0.shape[0])[:, None] < L[None, :]
    assert np.array_equal(L, _path_len(r1))
    assert np.array_equal(r0.transpose(0, 2, 1)[valid], r1.transpose(0, 2, 1)[valid])
    assert np.array_equal(c0, c1)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [dict(kinds=8, seed=3, opts={"maxBounces": 8}),
                                  dict(kinds=4, seed=2, opts={"maxBounces": 6, "doRussianRoulette": 0}),
                                  dict(kinds=8, seed=9, opts={"maxBounces": 3, "doMNEE": 1})],
                         ids=["8-materials-d8", "4-materials-noRR", "8-materials-d3-mnee"])
def test_gpu_wavefront_schedule_equals_megakernel(gpu, case):
    """The wavefront schedule of the sample pass (per-bounce traversal kernel, material-sorted
    queues, shading kernel) writes exactly the megakernel's bounce records, albedo, normal and
    colour: the paths draw the same random numbers in the same order whichever schedule runs."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    ctx = default_context()
    s = Scene(ctx, seed=case["seed"], sphere_count=24, subdivision=3, material_kinds=case["kinds"])
    w, h = 96, 64
    outs = []
    for sched in (PLTPT.MEGAKERNEL, PLTPT.WAVEFRONT):
        p = PLTPT(ctx, dict(BASE_OPTS, **case["opts"]))
        p.set_schedule(sched)
        p.set_scene(s)
        p.set_frame(2)
        color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        albedo, normal = torch.empty_like(color), torch.empty_like(color)
        p.execute(w, h, color, albedo, normal)
        torch.cuda.synchronize()
        outs.append((p.bounce_records(w, h), color.cpu().numpy(), albedo.cpu().numpy(), normal.cpu().numpy()))
        p.close()
    (r0, c0, a0, n0), (r1, c1, a1, n1) = outs
    L = _path_len(r0)
    assert np.array_equal(L, _path_len(r1))
    valid = np.arange(r0.shape[0])[:, None] < L[None, :]  # records past a path's end are not written
    assert np.array_equal(r0.transpose(0, 2, 1)[valid], r1.transpose(0, 2, 1)[valid])
    assert np.array_equal(a0, a1) and np.array_equal(n0, n1)
    assert np.array_equal(c0, c1)


@pytest.mark.gpu
def test_gpu_pltpt_frames_accumulate_and_options(gpu):
    """Frame counter advances per execute; set_scene resets it; unsupported options raise."""
    import torch
    from pltfalcor_amd import _lib
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    ctx = default_context()
    s = Scene(ctx, seed=2, sphere_count=4, subdivision=2)
    a = s.numpy()
    p = PLTPT(ctx, dict(BASE_OPTS))
    p.set_scene(s)
    w, h = 32, 32
    dev = torch.device("cuda")
    color = torch.empty((h, w, 4), dtype=torch.float32, device=dev)
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    frames = []
    for _ in range(3):
        p.execute(w, h, color, albedo, normal)
        frames.append(color.cpu().numpy().reshape(-1, 4).copy())
    for f in range(3):
        o = O.render(a, w, h, frame=f, max_bounces=6)[:, :3]
        close = color_close(frames[f], o)
        assert close.mean() >= 0.99, (f, close.mean())
    assert not np.array_equal(frames[0], frames[1])
    p.set_scene(s)  # resets the frame counter (PLTPT.cpp:510)
    p.execute(w, h, color, albedo, normal)
    assert np.array_equal(color.cpu().numpy().reshape(-1, 4), frames[0])
    p.set_option("MNEEMaxOccluders", 9)
    with pytest.raises(_lib.PLTError):
        p.execute(w, h, color, albedo, normal)
    p.close()


def _mnee_scene_pair(with_sheet, glass):
    """Ground, light and either a thin dielectric sheet (y = 4) or a glass slab (faces at y = 1 and
    1.5, camera below it at y = 0.6), as GPU scene and host arrays. The occluder trace re-offsets its
    origin at every crossing while only subtracting the hit distance from its extent
    (Manifolds.slangh:146-147), so the ray's end drifts towards the light by the crossings' offsets;
    faces at |y| < 2 keep that drift (2 x 9.2e-5) short of the light point's own offset (3.7e-4 at
    y = 7), where the light's triangle would become an occluder."""
    from pltfalcor_amd import materials as M
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    if glass:
        g = M.dielectric(1.5, 0.0)
        g.specular[0] = M.h(0.0)
        mats = [M.diffuse(.5, .5, .5), M.diffuse(.8, .8, .8, emission=2), g]
        quads = [(0.0, 10.0, 1.0, 0), (7.0, 1.5, -1.0, 1), (1.5, 10.0, 1.0, 2), (1.0, 10.0, -1.0, 2)]
    else:
        mats = [M.diffuse(.5, .5, .5), M.diffuse(.8, .8, .8, emission=2), M.thin_dielectric(1.5, 1e-3)]
        quads = [(0.0, 10.0, 1.0, 0), (7.0, 1.5, -1.0, 1)] + ([(4.0, 10.0, 1.0, 2)] if with_sheet else [])
    a = _look_down(_quad_scene(quads, mats), height=0.6 if glass else 2.0)
    s = Scene.from_arrays(default_context(), a["positions"], a["normals"], a["indices"], a["tri_mesh"], a["mesh_material"],
                          a["materials"], a["profiles"], _camera_struct(a["camera"]))
    return s, a


def _scene_variant(s, a, variant):
    """The procedural scene with its grated conductors replaced by grated opaque dielectrics
    ("grated-opaque"), or with a camera polarizer ("polarizer": full strength at 0.7 rad;
    "half-polarizer": strength 0.5 at 2.1 rad)."""
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd import _lib
    from pltfalcor_amd import materials as M
    mats = (_lib.Material * (len(a["materials"]) // C.sizeof(_lib.Material))).from_buffer_copy(a["materials"])
    cam = np.array(a["camera"], np.float32)
    cam = np.concatenate([cam, np.zeros(max(0, 16 - cam.size), np.float32)])
    if variant == "grated-opaque":
        swapped = 0
        for m in mats:
            if m.type == M.GRATED_CONDUCTOR:
                m.type = M.GRATED_OPAQUE_DIELECTRIC
                m.base_color[0], m.base_color[1], m.base_color[2] = M.h(.8), M.h(.6), M.h(.3)
                m.ior_n_profile = m.ior_k_profile = M.NO_PROFILE
                swapped += 1
        assert swapped > 0
    elif variant == "polarizer":
        cam[14], cam[15] = 1.0, 0.7
    elif variant == "half-polarizer":
        cam[14], cam[15] = 0.5, 2.1
    mb = C.string_at(C.addressof(mats), C.sizeof(mats))
    s2 = Scene.from_arrays(default_context(), a["positions"], a["normals"], a["indices"], a["tri_mesh"], a["mesh_material"],
                           mb, a["profiles"], _camera_struct(cam))
    return s2, s2.numpy()


def _camera_struct(cam):
    from pltfalcor_amd import _lib
    c = _lib.Camera()
    for i in range(3):
        c.pos[i], c.target[i], c.up[i] = float(cam[i]), float(cam[3 + i]), float(cam[6 + i])
    c.fov_y, c.aspect, c.focal_distance = float(cam[9]), float(cam[10]), float(cam[11])
    c.jitter[0], c.jitter[1] = float(cam[12]), float(cam[13])
    if len(cam) >= 16:
        c.polarizer, c.polarizer_theta = float(cam[14]), float(cam[15])
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("glass", [False, True], ids=["thin-sheet", "glass-slab"])
def test_gpu_mnee_flat_occluders_match_oracle(gpu, glass):
    """Manifold NEE through a thin sheet (straight connection) and through both faces of a glass slab
    (a two-vertex Newton walk): the GPU image equals the oracle's, including which connections the
    light's own triangle rejects, and manifold NEE reaches the ground (it is black without it)."""
    s, a = _mnee_scene_pair(True, glass)
    opts = dict(BASE_OPTS, maxBounces=2, doMNEE=1, doRussianRoulette=0)
    w, h = 48, 48
    g_col, g_rec, o_col, o_rec, _, _ = _render_pair(s, a, w, h, frame=3, opts=opts)
    assert np.array_equal(g_rec[0, 0], o_rec[0, 0])
    assert (o_col[:, :3] > 0).any(axis=1).mean() > 0.05
    lit_g, lit_o = (g_col[:, :3] > 0).any(axis=1), (o_col[:, :3] > 0).any(axis=1)
    print(f"[{'glass' if glass else 'sheet'}] lit pixels gpu {lit_g.mean():.4f} oracle {lit_o.mean():.4f} "
          f"agree {(lit_g == lit_o).mean():.5f}")
    assert (lit_g == lit_o).mean() >= 0.995
    close = color_close(g_col, o_col)
    print(f"[{'glass' if glass else 'sheet'}] colour within 1e-3 rel + floor {close.mean():.5f}")
    assert close.mean() >= 0.995



def _quad_scene(quads, materials):
    """Scene of horizontal squares: quads = [(y, half size, normal y (+1 / -1), material[, centre x,
    centre z])], the procedural scene's spectral profiles (profile 2 is the emission spectrum)."""
    from pltfalcor_amd.scene import Scene
    proc = Scene(None, seed=1, sphere_count=0, subdivision=1)
    pa = proc.numpy()
    pos, nrm, idx, tm, mm = [], [], [], [], []
    for m, q in enumerate(quads):
        (y, hs, ny, mat), (cx, cz) = q[:4], (q[4:6] if len(q) > 4 else (0.0, 0.0))
        v0 = len(pos)
        pos += [[cx - hs, y, cz - hs], [cx + hs, y, cz - hs], [cx + hs, y, cz + hs], [cx - hs, y, cz + hs]]
        nrm += [[0, ny, 0]] * 4
        # winding so that the face normal follows the vertex normal
        idx += [[v0, v0 + 2, v0 + 1], [v0, v0 + 3, v0 + 2]] if ny > 0 else [[v0, v0 + 1, v0 + 2], [v0, v0 + 2, v0 + 3]]
        tm += [m, m]
        mm.append(mat)
    from pltfalcor_amd import materials as M
    return Scene.from_arrays(None, np.array(pos, np.float32), np.array(nrm, np.float32), np.array(idx, np.uint32),
                             np.array(tm, np.uint32), np.array(mm, np.uint32), M.pack(materials), pa["profiles"],
                             proc.arrays.camera).numpy()


def _look_down(a, at=(0.3, 0.2), height=2.0):
    cam = a["camera"].copy()
    cam[0:3], cam[3:6], cam[6:9] = [at[0], height, at[1]], [at[0], 0.0, at[1]], [0.0, 0.0, 1.0]
    cam[9], cam[10], cam[12:14] = np.radians(60.0), 1.0, 0.0
    a["camera"] = cam
    return a


def _fresnel_unpolarised(c, eta):
    ct = np.sqrt(1 - (1 - c * c) / eta ** 2)
    rs = (c - eta * ct) / (c + eta * ct)
    rp = (eta * c - ct) / (eta * c + ct)
    return (rs * rs + rp * rp) / 2


def test_oracle_mnee_through_thin_slab_equals_unoccluded_connection():
    """Manifold NEE through a thin dielectric sheet lying between the ground and the light: the
    sheet does not deviate light, so the walk converges on the straight segment at once, the hinted
    lobe choice is forced to transmission (pdf 1), and the sheet's transmission is 1 - R per
    polarisation (T (1 - R)^-1 T). Each ground sample therefore equals the same sample with the sheet
    removed times 1 - R(theta) (unpolarised source), R between its normal-incidence value and its
    value at the widest angle of the configuration; without manifold NEE the sheet occludes the
    light. The reference's final visibility ray ends on the light's own surface
    (Manifolds.slangh:266-276 with ls.beam.posW), so some connections are rejected by the light's
    triangle; those samples are 0 and are not counted here."""
    from pltfalcor_amd import materials as M
    eta = float(np.float16(1.5))
    mats = [M.diffuse(.5, .5, .5), M.diffuse(.8, .8, .8, emission=2), M.thin_dielectric(1.5, 1e-3)]
    ground_light = [(0.0, 10.0, 1.0, 0), (7.0, 1.5, -1.0, 1)]
    free = O.OracleScene(_look_down(_quad_scene(ground_light, mats)))
    sheet = O.OracleScene(_look_down(_quad_scene(ground_light + [(4.0, 10.0, 1.0, 2)], mats)))
    kw = dict(max_bounces=1, do_rr=0, do_mis=0, is_emitters=0, threads=8)
    assert np.all(sheet.render(16, 16, frame=0, do_mnee=0, **kw)[:, :3] == 0)
    # widest angle: ground corner seen by the camera (~1.15 off-axis at most) to the far light corner
    cmax = 3.0 / np.hypot(3.0, np.hypot(1.5 + 1.5, 1.5 + 1.5))
    lo, hi = 1 - _fresnel_unpolarised(cmax, eta), 1 - _fresnel_unpolarised(1.0, eta)
    connected = 0
    for f in range(6):
        A = free.render(16, 16, frame=f, do_mnee=1, **kw)[:, :3]
        B = sheet.render(16, 16, frame=f, do_mnee=1, **kw)[:, :3]
        nz = (B != 0).any(1)
        assert np.all(A[nz] > 0)
        ratio = B[nz] / A[nz]
        assert np.all(ratio > lo - 1e-4) and np.all(ratio < hi + 1e-4), (ratio.min(), ratio.max(), lo, hi)
        assert np.all(np.ptp(ratio, axis=1) < 1e-4)  # the same factor in every colour channel
        connected += nz.sum()
    assert connected > 0.05 * 6 * 256


def test_oracle_mnee_walk_satisfies_snell_through_glass_slab():
    """The Newton walk (Manifolds.slangh:165-281) through both faces of a non-dispersive glass slab
    (eta 1.5, faces at y = 3.5 and 4) from a ground point to a light point converges on the refracted
    path: the vertices agree with the flat-slab solution found by bisection on the incidence angle,
    the exit direction is parallel to the entry direction, and Snell's law holds at both faces. With
    one occluder allowed the walk is refused."""
    from pltfalcor_amd import materials as M
    glass = M.dielectric(1.5, 0.0)
    glass.specular[0] = M.h(0.0)  # no dispersion
    mats = [M.diffuse(.5, .5, .5), M.diffuse(.8, .8, .8, emission=2), glass]
    a = _quad_scene([(0.0, 10.0, 1.0, 0), (7.0, 1.5, -1.0, 1), (4.0, 10.0, 1.0, 2), (3.5, 10.0, -1.0, 2)], mats)
    osc = O.OracleScene(a)
    eta = 1.5
    # end points just below the light's plane, so that the final visibility ray does not end on it
    cases = [([1.0, 0.01, -0.5], [0.0, 6.99, 0.0]), ([1.0, 0.01, -0.5], [1.2, 6.99, 1.2]), ([2.5, 0.01, 1.0], [1.2, 6.99, 1.2]),
             ([-1.5, 0.01, 0.7], [0.5, 6.99, -1.0]), ([0.2, 0.01, 0.3], [-1.0, 6.99, 1.4])]
    converged = 0
    for x, L in cases:
        x, L = np.array(x, np.float64), np.array(L, np.float64)
        n, v, lobes, it, pdf = osc.mnee_probe(x, L)
        assert osc.mnee_probe(x, L, max_occluders=1)[0] == -1
        if n < 0:
            continue  # iteration cap
        converged += 1
        assert n == 2 and it <= 60 and pdf == 1.0
        assert np.all(lobes == 0x40)  # delta transmission (not dispersive)
        # flat-slab solution: horizontal distance = (h_below + h_above) tan(t) + thickness tan(t_inside)
        hb, ha, th = 3.5 - x[1], L[1] - 4.0, 0.5
        D = np.hypot(*(L - x)[[0, 2]])
        lo_t, hi_t = 0.0, np.pi / 2 - 1e-9
        for _ in range(200):
            t = (lo_t + hi_t) / 2
            s2 = np.sin(t) / eta
            f = (hb + ha) * np.tan(t) + th * s2 / np.sqrt(1 - s2 * s2) - D
            lo_t, hi_t = (t, hi_t) if f < 0 else (lo_t, t)
        u = (L - x)[[0, 2]] / D
        r1 = hb * np.tan(t)
        r2 = r1 + th * np.tan(np.arcsin(np.sin(t) / eta))
        want = np.array([[x[0] + u[0] * r1, 3.5, x[2] + u[1] * r1], [x[0] + u[0] * r2, 4.0, x[2] + u[1] * r2]])
        assert np.abs(v - want).max() < 2e-3, (v, want)
        d0, d1, d2 = v[0] - x, v[1] - v[0], L - v[1]
        d0, d1, d2 = (d / np.linalg.norm(d) for d in (d0, d1, d2))
        assert np.linalg.norm(d0 - d2) < 1e-3
        sin0, sin1 = np.hypot(d0[0], d0[2]), np.hypot(d1[0], d1[2])
        assert abs(sin0 - eta * sin1) < 1e-3
    assert converged >= 4


def test_oracle_walk_trig_accuracy():
    """The Newton step's FP64 acos / atan2 (a fixed rational evaluation shared by host and device)
    against numpy's float64 libm: both within 2 ulp over [-1, 1] (acos) and all quadrants and the
    axes (atan2), acos(1) exactly 0. A wrong coefficient of the rational approximation moves either
    by orders of magnitude more."""
    n = 400001
    x = np.linspace(-1, 1, n)
    rng = np.random.default_rng(1)
    y = rng.uniform(-1, 1, n)
    y[::11] = 0.0
    x2 = x.copy()
    x2[::13] = 0.0
    x3 = np.concatenate([x2, rng.uniform(-1e3, 1e3, n), np.tan(rng.uniform(-1.57, 1.57, n))])
    y3 = np.concatenate([y, rng.uniform(-1e3, 1e3, n), rng.uniform(-1, 1, n)])
    ac, _ = O.walk_trig(x, np.zeros_like(x))
    _, at = O.walk_trig(x3, y3)
    ra = np.arccos(x)
    rt = np.arctan2(y3, x3)
    pos = ra > 0
    assert np.max(np.abs(ac - ra)[pos] / np.spacing(ra[pos])) <= 2.0
    assert np.all(ac[~pos] == 0)
    nz = rt != 0
    assert np.max(np.abs(at[nz] - rt[nz]) / np.spacing(np.abs(rt[nz]))) <= 2.0
    assert np.all(at[~nz] == 0)


def test_oracle_walk_derivatives_match_finite_differences():
    """The walk's specular direction derivatives along the surface parameters (differential shading
    geometry of a triangle with non-parallel vertex normals, then the material's partial
    derivatives) equal central finite differences of the direction itself, in FP64: mirror
    reflection (conductor, coated conductor), refraction from the normal's side (dielectric) and
    the thin sheet's straight transmission. (The grating's Jacobian is the reference's
    approximation, DiffractionGrating.slang:229-232, and refraction from behind the normal is a
    branch the walk never takes; neither is pinned here.)"""
    from pltfalcor_amd import materials as M
    from pltfalcor_amd.scene import Scene
    pa = Scene(None, seed=1, sphere_count=0, subdivision=1).numpy()
    prof, nprof = pa["profiles"], pa["n_profiles"]
    P = np.array([[0, 0, 0], [1.0, 0.1, 0], [0.2, -0.05, 1.0]])
    N = np.array([[0.1, 1, 0.05], [-0.2, 1, 0.1], [0.05, 1, -0.3]])
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    cases = [(M.conductor(0.0, 0, 1), 0x04), (M.coated_conductor(0.0, 1.5, 2.0, 0, 1), 0x04),
             (M.dielectric(1.5, 0.0), 0x40), (M.dielectric(1.8, 0.0), 0x04), (M.thin_dielectric(1.5, 1e-3), 0x40)]
    h = 1e-6
    for mat, lobe in cases:
        mb = M.pack([mat])
        for src in ([0.7, 2.0, 0.4], [-1.0, 0.8, 1.5], [2.0, 1.5, -0.5]):
            for b1, b2 in ((0.3, 0.25), (0.1, 0.7), (0.6, 0.1)):
                f = lambda u, v: O.walk_derivatives(mb, prof, nprof, P, N, u, v, src, lobe, 550.0)  # noqa: E731
                r = f(b1, b2)
                assert abs(np.linalg.norm(r[0]) - 1) < 1e-9
                du = (f(b1 + h, b2)[0] - f(b1 - h, b2)[0]) / (2 * h)
                dv = (f(b1, b2 + h)[0] - f(b1, b2 - h)[0]) / (2 * h)
                assert np.abs(du - r[1]).max() < 1e-6 * max(1.0, np.abs(r[1]).max()), (mat.type, src, b1, b2)
                assert np.abs(dv - r[2]).max() < 1e-6 * max(1.0, np.abs(r[2]).max()), (mat.type, src, b1, b2)


# ------------------------------------------------------------------------------ camera jitter patterns


def _radical_inverse(i, base):
    """Van der Corput radical inverse as an exact fraction."""
    from fractions import Fraction
    r, f = Fraction(0), Fraction(1)
    while i:
        f /= base
        r += f * (i % base)
        i //= base
    return r


def _pattern(kind, count, n):
    from pltfalcor_amd import _lib
    xy = (C.c_float * (2 * n))()
    _lib.check(_lib.load().plt_sample_pattern_generate(kind, count, n, xy))
    return np.array(list(xy), np.float64).reshape(n, 2)


def test_halton_jitter_pattern_is_the_shifted_radical_inverse():
    """HaltonSamplePattern::next: (h2(i), h3(i)) shifted by fract(h + 0.5) - 0.5, cyclic with the
    sample count; every value within float rounding of the exact radical inverse."""
    from fractions import Fraction
    got = _pattern(2, 16, 40)
    for k in range(40):
        i = k % 16
        for d, base in enumerate((2, 3)):
            h = _radical_inverse(i, base) + Fraction(1, 2)
            want = float(h - (h.numerator // h.denominator) - Fraction(1, 2))
            assert abs(got[k, d] - want) <= 2e-7, (k, d, got[k, d], want)
    assert np.array_equal(got[0], [0, 0])
    assert np.allclose(got[1], [-0.5, 1 / 3], atol=1e-6)  # fract(1/2 + 1/2) - 1/2, fract(1/3 + 1/2) - 1/2
    # count 0 never repeats
    assert not np.array_equal(_pattern(2, 0, 40)[16], got[0])


def test_directx_and_stratified_jitter_patterns():
    dx = _pattern(1, 5, 16)  # DxSamplePattern uses 8 whatever the count
    table = np.array([[1, -3], [-1, 3], [5, 1], [-3, -5], [-5, 5], [-7, -1], [3, 7], [7, -7]]) / 16
    assert np.array_equal(dx, np.vstack([table, table]))
    for count, (bx, by) in ((16, (4, 4)), (12, (3, 4)), (7, (7, 1))):  # bx from floor(sqrt(count)) up
        st = _pattern(3, count, 3 * count)
        assert np.all((st >= -0.5) & (st < 0.5))
        for r in range(3):  # each round visits every bin exactly once
            b = st[r * count:(r + 1) * count] + 0.5
            bins = np.floor(b[:, 0] * bx).astype(int) + bx * np.floor(b[:, 1] * by).astype(int)
            assert sorted(bins) == list(range(count)), (count, r)
    from pltfalcor_amd import _lib
    with pytest.raises(_lib.PLTError):
        _pattern(4, 16, 1)


@pytest.mark.gpu
def test_gpu_halton_jittered_frames_match_oracle(gpu):
    """The reference graph's jitter (Halton, 16 samples): frame k uses offset k scaled by the frame
    size; primary hits bit-identical to the oracle rendered with that jitter, colour to the path
    tracer's parity bar; set_scene restarts the pattern."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    ctx = default_context()
    s = Scene(ctx, seed=4, sphere_count=8, subdivision=3)
    a = s.numpy()
    w, h = 48, 40
    p = PLTPT(ctx, dict(BASE_OPTS))
    p.set_sample_pattern(PLTPT.HALTON, 16)
    p.set_scene(s)
    pat = _pattern(2, 16, 4).astype(np.float32)
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    jit = []
    for f in range(4):
        p.execute(w, h, color, albedo, normal)
        torch.cuda.synchronize()
        j = p.jitter()
        jit.append(j)
        assert j == (float(pat[f, 0] * np.float32(1 / w)), float(pat[f, 1] * np.float32(1 / h))), (f, j)
        cam = np.array(a["camera"], np.float32).copy()
        cam[12], cam[13] = j
        g_rec = p.bounce_records(w, h)
        g_col = color.cpu().numpy().reshape(-1, 4)
        o_col, o_rec = O.render(a, w, h, frame=f, max_bounces=6, records=True, camera=cam)
        assert np.array_equal(g_rec[0, :3], o_rec[0, :3]), f
        topo = _topology_equal(g_rec, o_rec)
        assert topo.mean() >= 0.99
        close = color_close(g_col, o_col)
        assert close[topo].mean() >= 0.995
    assert len(set(jit)) == 4
    p.set_scene(s)
    p.execute(w, h, color, albedo, normal)
    assert p.jitter() == jit[0] == (0.0, 0.0)
    p.close()


@pytest.mark.gpu
def test_gpu_executes_with_reference_defaults_dictionary(gpu):
    """A pass configured with every key of the reference's own defaults dictionary (the PLTPT.h
    initialisers: sampleGenerator 1, maxBounces 32, manifold NEE on, ...) renders a frame: no key
    value is refused, the colour is finite with alpha 1, and the records match the oracle's at
    those options on every primary hit."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    ctx = default_context()
    s = Scene(ctx, seed=4, sphere_count=16, subdivision=3, material_kinds=8)
    p = PLTPT(ctx, REFERENCE_DEFAULTS)
    assert p.get_scripting_dictionary() == pytest.approx(REFERENCE_DEFAULTS)
    w, h = 48, 32
    out = _render_pair(s, s.numpy(), w, h, frame=0, opts=REFERENCE_DEFAULTS, pass_=p)
    g_col = out[0]
    assert np.isfinite(g_col).all() and (g_col[:, 3] == 1).all()
    _check_parity("reference-defaults", *out)
    p.close()


@pytest.mark.gpu
def test_gpu_set_scene_upload_failure_leaves_pass_unbound(gpu):
    """A device upload that fails while a scene is bound (injected on the host) returns the error,
    leaves the pass without a scene (execute then reports PLT_ESTATE instead of launching kernels on
    partly uploaded tables), and a later bind succeeds and renders."""
    import torch
    from pltfalcor_amd import _lib
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    lib = _lib.load()
    ctx = default_context()
    s = Scene(ctx, seed=2, sphere_count=4, subdivision=2, material_kinds=4)
    p = PLTPT(ctx, {"maxBounces": 3})
    w, h = 16, 16
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    for n in (0, 2, 5):
        _lib.check(lib.plt_debug_fail_upload(n))
        with pytest.raises(_lib.PLTError) as e:
            p.set_scene(s)
        assert e.value.status == _lib.PLT_ENOMEM
        with pytest.raises(_lib.PLTError) as e:
            p.execute(w, h, color, albedo, normal)
        assert e.value.status == _lib.PLT_ESTATE
    _lib.check(lib.plt_debug_fail_upload(-1))
    p.set_scene(s)
    p.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    assert torch.isfinite(color).all()
    p.close()


def _band_rows(h, bands, rows):
    """Row ranges of `bands` bands of `rows` full rows spread over the frame (sky, horizon, objects, ground)."""
    starts = np.linspace(0, h - rows, bands).astype(int)
    return [(int(r0), rows) for r0 in starts]


@pytest.mark.gpu
@pytest.mark.parametrize("config", ["C1", "C2"])
def test_gpu_workload_configs_match_oracle(gpu, config):
    """The pass on the workload scenes at their full shapes, against the oracle's render of 64 full
    rows (four bands of 16 rows over the frame) of the same frame:
      C1: 512 x 512, 64 icospheres of subdivision 5 (1,310,724 triangles), 4 material kinds, depth 6,
          manifold NEE on, frame 0, no jitter;
      C2: 1920 x 1080, 256 icospheres (5,242,884 triangles), 8 material kinds, depth 8, manifold NEE
          on, frame 0 with the reference graph's Halton jitter (the bench's configuration).
    On those rows: the parity bar of the small cases (primary hits exact, topology, records, colour
    within 1e-3 relative + the image-scale floor, albedo and normal). Over the whole frame: finite,
    non-negative colour with alpha 1, and primary hit ids on the sampled rows bit-identical."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import PATH_1, PATH_2, Scene
    from pltfalcor_amd.bsdf import default_context
    ctx = default_context()
    if config == "C1":
        w, h, cfg, opts = 512, 512, PATH_1, dict(REFERENCE_DEFAULTS, maxBounces=6)
    else:
        w, h, cfg, opts = 1920, 1080, PATH_2, dict(REFERENCE_DEFAULTS, maxBounces=8)
    s = Scene(ctx, aspect=w / h, **cfg)
    a = s.numpy()
    assert a["indices"].shape[0] == cfg["sphere_count"] * 20 * 4 ** cfg["subdivision"] + 4  # ground and light quads
    p = PLTPT(ctx, opts)
    if config == "C2":
        p.set_sample_pattern(PLTPT.HALTON, 16)
    p.set_scene(s)
    p.set_frame(0)
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    p.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    g_rec_all = p.bounce_records(w, h)
    g_col_all = color.cpu().numpy().reshape(-1, 4)
    g_aov_all = np.concatenate([albedo.cpu().numpy().reshape(-1, 4), normal.cpu().numpy().reshape(-1, 4)], axis=1)
    assert np.isfinite(g_col_all).all() and (g_col_all[:, :3] >= 0).all() and (g_col_all[:, 3] == 1).all()
    cam = np.array(a["camera"], np.float32).copy()
    cam = np.concatenate([cam, np.zeros(max(0, 16 - cam.size), np.float32)])
    cam[12], cam[13] = p.jitter()
    osc = O.OracleScene(a)
    kw = {OPTION_KW[k]: int(v) for k, v in opts.items() if k in OPTION_KW}
    kw["mnee_threshold"] = float(opts["MNEESolverThreshold"])
    cols, recs, aovs, gcols, grecs, gaovs = [], [], [], [], [], []
    for r0, rows in _band_rows(h, 4, 16):
        first, count = r0 * w, rows * w
        o_col, o_rec, o_aov = osc.render(w, h, frame=0, max_bounces=int(opts["maxBounces"]), records=True, aov=True,
                                         first=first, count=count, camera=cam, threads=0, **kw)
        cols.append(o_col); recs.append(o_rec); aovs.append(o_aov)
        sl = slice(first, first + count)
        gcols.append(g_col_all[sl]); grecs.append(g_rec_all[:, :, sl]); gaovs.append(g_aov_all[sl])
    osc.close()
    cat = lambda xs, ax=0: np.concatenate(xs, axis=ax)  # noqa: E731
    _check_parity(config, cat(gcols), cat(grecs, 2), cat(cols), cat(recs, 2), cat(gaovs), cat(aovs), deep_fields_min=0.99)
    p.close()


def test_oracle_two_bounce_mirror_matches_lambert_closed_form():
    """Forward unwind beyond depth 0: a camera at (7, 1, 0) looks up at a downward-facing smooth
    opaque-dielectric mirror (albedo 0.8, roughness 0: a depolarising delta reflection of
    reflectance 0.8) at y = 3 and sees, in it, the Lambertian ground (albedo 0.5) lit by the 3 x 3
    quad light at y = 7 (the mirror casts no shadow on the seen patch). maxBounces 1, NEE without
    MIS: every pixel's light arrives by NEE at the ground vertex (bounce 1) and is carried through the
    mirror vertex (bounce 0) to the camera, so its expectation is 0.8 x 0.5 / pi x E(x') x Le_rgb,
    E Lambert's polygon irradiance at the mirrored ground point x'. A dropped unwind step, a wrong
    vertex order or a wrong contribution bookkeeping at bounce 1 changes it."""
    from pltfalcor_amd import materials as M
    mats = [M.diffuse(.5, .5, .5), M.diffuse(.8, .8, .8, emission=2), M.opaque_dielectric(.8, .8, .8, 0.0)]
    a = _quad_scene([(0.0, 12.0, 1.0, 0), (7.0, 1.5, -1.0, 1), (3.0, 1.5, -1.0, 2, 7.0, 0.0)], mats)
    w = h = 16
    cam = a["camera"].copy()
    cam[0:3], cam[3:6], cam[6:9] = [7.0, 1.0, 0.0], [7.0, 3.0, 0.0], [0.0, 0.0, 1.0]
    cam[9], cam[10], cam[12:14] = np.radians(30.0), 1.0, 0.0
    a["camera"] = cam
    o, d = _camera_rays(cam, w, h)
    t = (3.0 - o[1]) / d[:, 1]
    hit = o + t[:, None] * d
    assert np.all(np.abs(hit[:, 0] - 7.0) < 1.5) and np.all(np.abs(hit[:, 2]) < 1.5)  # every pixel sees the mirror
    dr = d * np.array([1.0, -1.0, 1.0])
    g = hit + (3.0 / -dr[:, 1])[:, None] * dr  # the mirrored ray reaches the ground y = 0
    poly = [np.array(p, float) for p in ([-1.5, 7, -1.5], [1.5, 7, -1.5], [1.5, 7, 1.5], [-1.5, 7, 1.5])]
    E = np.array([_lambert_polygon_irradiance(p, np.array([0, 1.0, 0]), poly) for p in g])
    # the light's own colour, seen from below
    b = dict(a)
    camL = cam.copy()
    camL[0:3], camL[3:6], camL[9] = [0.0, 1.0, 0.0], [0.0, 7.0, 0.0], np.radians(10.0)
    b["camera"] = camL
    frames = 96
    Le = np.zeros(3)
    for f in range(frames):
        Le += O.render(b, 8, 8, frame=f, max_bounces=0, do_rr=0, is_emitters=0)[:, :3].mean(0)
    Le /= frames
    acc = np.zeros((w * h, 4))
    for f in range(frames):
        acc += O.render(a, w, h, frame=f, max_bounces=1, do_rr=0, is_emitters=0, do_mis=0, threads=8)
    acc /= frames
    expected = 0.8 * 0.5 / np.pi * E[:, None] * Le[None, :]
    got = acc[:, :3]
    rel = np.abs(got.mean(0) - expected.mean(0)) / expected.mean(0)
    print("two-bounce mirror: got", got.mean(0), "expected", expected.mean(0))
    assert (rel < 0.015).all(), (got.mean(0), expected.mean(0))
    blk = lambda v: v.reshape(4, 4, 4, 4).sum((1, 3))  # noqa: E731
    for c in range(3):
        rb = np.abs(blk(got[:, c].reshape(16, 16)) - blk(expected[:, c].reshape(16, 16))) / blk(expected[:, c].reshape(16, 16))
        assert rb.max() < 0.04, (c, rb.max())


@pytest.mark.gpu
def test_gpu_launch_configs_and_graph_replay_are_identical(gpu):
    """Every launch configuration of the pass -- the megakernel at 64, 128 and 256 threads per block,
    the wavefront schedule at 4 and 32 blocks per CU, the split solve pass with ample and with
    overflowing event areas, each eager and replayed from captured HIP graphs -- writes the default
    configuration's bounce records, colour, albedo and normal bit for bit. The
    graphs are replayed for several frames with eager work and synchronisations between the replays,
    and every replay is checked against an eager render of the same frame."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    ctx = default_context()
    s = Scene(ctx, seed=3, sphere_count=24, subdivision=3, material_kinds=8)
    w, h = 96, 64
    opts = dict(BASE_OPTS, maxBounces=6, doMNEE=1)

    def render(cfg, frame):
        p = PLTPT(ctx, opts)
        p.set_sample_pattern(PLTPT.HALTON, 16)
        p.set_launch_config(**cfg)
        p.set_scene(s)
        out = []
        color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        albedo, normal = torch.empty_like(color), torch.empty_like(color)
        for f in range(frame + 1):
            color.fill_(float("nan"))
            p.execute(w, h, color, albedo, normal)
            torch.cuda.synchronize()
            junk = torch.randn(1 << 16, device="cuda").sum().item()  # eager work between replays
            assert np.isfinite(junk)
            out.append((p.bounce_records(w, h), color.cpu().numpy().copy(), albedo.cpu().numpy().copy(),
                        normal.cpu().numpy().copy()))
        p.close()
        return out

    frames = 3
    ref = render(dict(schedule=PLTPT.MEGAKERNEL, sample_block=64, graph=False), frames - 1)
    L = [_path_len(r[0]) for r in ref]
    for cfg in (dict(schedule=PLTPT.MEGAKERNEL, sample_block=128), dict(schedule=PLTPT.MEGAKERNEL, sample_block=256),
                dict(schedule=PLTPT.MEGAKERNEL, sample_block=64, graph=True),
                dict(schedule=PLTPT.MEGAKERNEL, sample_block=256, graph=True),
                dict(schedule=PLTPT.WAVEFRONT, wf_blocks_per_cu=4), dict(schedule=PLTPT.WAVEFRONT, wf_blocks_per_cu=32),
                dict(schedule=PLTPT.WAVEFRONT, wf_blocks_per_cu=16, graph=True),
                # split solve pass: ample event area, then areas so small that most regions overflow
                # into the megakernel fallback (1 event and no walk vertex per path), eager and replayed
                dict(solve_schedule=PLTPT.SOLVE_SPLIT, solve_events_per_path=8, solve_hits_per_path=4),
                dict(solve_schedule=PLTPT.SOLVE_SPLIT, solve_events_per_path=1, solve_hits_per_path=0),
                dict(solve_schedule=PLTPT.SOLVE_SPLIT, solve_events_per_path=2, solve_hits_per_path=1, graph=True),
                dict(schedule=PLTPT.WAVEFRONT, solve_schedule=PLTPT.SOLVE_SPLIT, solve_events_per_path=4,
                     solve_hits_per_path=2, graph=True)):
        got = render(dict(dict(sample_block=64, wf_blocks_per_cu=16, graph=False), **cfg), frames - 1)
        for f in range(frames):
            valid = np.arange(ref[f][0].shape[0])[:, None] < L[f][None, :]
            assert np.array_equal(ref[f][0].transpose(0, 2, 1)[valid], got[f][0].transpose(0, 2, 1)[valid]), (cfg, f)
            for k in (1, 2, 3):
                assert np.array_equal(ref[f][k].view(np.uint32), got[f][k].view(np.uint32)), (cfg, f, k)


@pytest.mark.gpu
def test_gpu_autotune_picks_a_config_and_restores_state(gpu):
    """The autotuner times the candidates, returns a valid configuration no slower than the starting
    one, sets it, and leaves the frame counter and the jitter pattern where they were: the next frame
    renders as it would have without the autotuning."""
    import torch
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    from pltfalcor_amd.bsdf import default_context
    ctx = default_context()
    s = Scene(ctx, seed=5, sphere_count=16, subdivision=3, material_kinds=8)
    w, h = 128, 96
    opts = dict(BASE_OPTS, maxBounces=6)
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    ref = PLTPT(ctx, opts)
    ref.set_sample_pattern(PLTPT.HALTON, 16)
    ref.set_scene(s)
    ref.set_frame(2)
    ref.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    want = color.cpu().numpy().copy()
    p = PLTPT(ctx, opts)
    p.set_sample_pattern(PLTPT.HALTON, 16)
    p.set_scene(s)
    p.set_frame(2)
    best, d_ms, b_ms = p.autotune(w, h, color, albedo, normal, frames=2)
    print("autotune:", best, f"default {d_ms:.3f} ms, best {b_ms:.3f} ms")
    assert best["schedule"] in (0, 1) and best["sample_block"] in (64, 128, 256) and 1 <= best["wf_blocks_per_cu"] <= 64
    assert b_ms <= d_ms * 1.0001 and p.launch_config() == best
    color.zero_()
    p.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    assert np.array_equal(color.cpu().numpy().view(np.uint32), want.view(np.uint32))
    p.close()
    ref.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(name="C2", workload=True),
    dict(name="env+analytic-mnee", lights=True, opts={"maxBounces": 8, "HWSS": 4, "doMNEE": 1}),
    dict(name="env+analytic-hwss1", lights=True, opts={"maxBounces": 6, "HWSS": 1, "doMNEE": 1}),
    dict(name="hwss1-noMIS", lights=False, opts={"maxBounces": 6, "HWSS": 1, "doMNEE": 1, "doMIS": 0}),
    dict(name="no-hwss-mis", lights=False, opts={"maxBounces": 6, "HWSS": 4, "doMNEE": 1, "HWSSDoMIS": 0}),
    dict(name="tiny-pertile", lights=True, opts={"maxBounces": 5, "HWSS": 4, "doMNEE": 1, "sampleGenerator": 0,
                                                  "NEEUsePerTileSG": 1, "tileSize": 16}),
    # walks that run out of iterations (their events turn void) and walks through three occluders
    dict(name="mnee-3occ-it4", lights=False, opts={"maxBounces": 7, "HWSS": 4, "doMNEE": 1, "MNEEMaxOccluders": 3,
                                                   "MNEEMaxIterations": 4}),
    dict(name="mnee-off", lights=True, opts={"maxBounces": 6, "HWSS": 4, "doMNEE": 0}),
], ids=lambda c: c["name"])
def test_gpu_split_solve_equals_megakernel(gpu, case):
    """The split solve pass (light-sampling kernel, deferred manifold walks, unwind kernel,
    megakernel fallback for overflowed regions) writes the megakernel's colour bit for bit: on the C2
    workload at 1920 x 1080 (manifold NEE, Halton jitter), and on scenes that take the analytic-light /
    environment-map kernels, one HWSS wavelength, MIS off, the tiny generator with per-tile NEE draws,
    walks cut short by a small iteration limit, manifold NEE off -- with the default event area and
    with one so small that the fallback solves most paths."""
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import PATH_2, Scene
    from test_env_map import _patchy_sky
    from test_lights import _all_lights
    ctx = default_context()
    if case.get("workload"):
        w, h = 1920, 1080
        s = Scene(ctx, aspect=w / h, **PATH_2)
        opts = dict(REFERENCE_DEFAULTS, maxBounces=8)
        sizes = [(8, 4)]
    else:
        w, h = 96, 64
        proc = Scene(None, seed=41, sphere_count=16, subdivision=3, material_kinds=8)
        pa = proc.numpy()
        s = Scene.from_arrays(ctx, pa["positions"], pa["normals"], pa["indices"], pa["tri_mesh"], pa["mesh_material"],
                              pa["materials"], pa["profiles"], proc.arrays.camera)
        if case["lights"]:
            s.set_env_map(_patchy_sky(), rotation=(0.0, 45.0, 0.0), intensity=0.5)
            s.set_lights(_all_lights())
        opts = case["opts"]
        sizes = [(8, 4), (1, 0)]

    def render(cfg, frames=2):
        p = PLTPT(ctx, opts)
        p.set_sample_pattern(PLTPT.HALTON, 16)
        p.set_launch_config(**cfg)
        p.set_scene(s)
        out = []
        color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        albedo, normal = torch.empty_like(color), torch.empty_like(color)
        stats = None
        for _ in range(frames):
            color.fill_(float("nan"))
            p.execute(w, h, color, albedo, normal)
            torch.cuda.synchronize()
            out.append(color.cpu().numpy().copy())
        if cfg["solve_schedule"] == PLTPT.SOLVE_SPLIT:
            stats = p.solve_stats()
        p.close()
        return out, stats

    ref, _ = render(dict(solve_schedule=PLTPT.SOLVE_MEGAKERNEL))
    assert all(np.isfinite(r).all() for r in ref) and ref[0][..., :3].mean() > 0
    for ev, hits in sizes:
        got, st = render(dict(solve_schedule=PLTPT.SOLVE_SPLIT, solve_events_per_path=ev, solve_hits_per_path=hits))
        print(case["name"], (ev, hits), st)
        assert st["events"] > 0 and st["event_vertices"] >= st["events"]
        assert st["failed_walks"] <= st["walks"] <= st["events"]
        if opts.get("doMNEE", 1) and hits > 0:
            assert st["walks"] > 0 and st["walk_vertices"] >= st["walks"]
        else:
            assert st["walks"] == 0 and st["walk_vertices"] == 0
        if case["name"] == "mnee-3occ-it4" and hits > 0:
            assert st["failed_walks"] > 0
        if ev == 1:
            assert st["overflowed_paths"] > 0  # the fallback took part
        for f, (a, b) in enumerate(zip(ref, got)):
            diff = a.view(np.uint32) != b.view(np.uint32)
            if diff.any():
                rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-30)
                print(f"{case['name']} {(ev, hits)} frame {f}: {int(diff.any(axis=2).sum())} pixels differ, max rel "
                      f"{rel.max():.3e}, quantiles 50/99 % of differing {np.quantile(rel[diff], [.5, .99])}")
            assert not diff.any(), (case["name"], ev, hits, f, int(diff.any(axis=2).sum()), np.argwhere(diff.any(axis=2))[:5])
