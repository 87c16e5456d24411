"""PBRT scene import (the PLT-material branch of the reference's importer): the grammar, the
material conversions field by field, transforms, shapes and face orientation, the camera, lights
and the octahedral environment map (CPU), and a render of an imported scene against the oracle
(GPU)."""
import math
import struct
import warnings

import numpy as np
import pytest

import oracle_lib as O
from parity import color_close

SCENE = """
# a pbrt-v4 scene exercising every supported directive kind
LookAt 0 3 8   0 0.6 0   0 1 0
Camera "perspective" "float fov" [ 40 ]
Film "rgb" "integer xresolution" [ 96 ] "integer yresolution" [ 64 ] "string filename" "out.exr"
Sampler "halton" "integer pixelsamples" 16
Integrator "volpath" "integer maxdepth" [ 8 ]
WorldBegin
LightSource "infinite" "rgb L" [ 0.05 0.07 0.1 ]
LightSource "distant" "point3 from" [ 0 10 0 ] "point3 to" [ 0.3 0 0.2 ] "rgb L" [ 1 1 1 ]
LightSource "point" "point3 from" [ 0 4 0 ]
MakeNamedMaterial "floor" "string type" "diffuse" "rgb reflectance" [ 0.6 0.55 0.5 ]
MakeNamedMaterial "gold coat"
    "string type" "coatedconductor" "rgb conductor.eta" [ 0.2 0.4 1.3 ] "rgb conductor.k" [ 3.9 2.4 1.8 ]
    "float interface.eta" 1.6
Texture "refl" "spectrum" "constant" "rgb value" [ 0.3 0.5 0.7 ]
AttributeBegin
  NamedMaterial "floor"
  Shape "trianglemesh" "point3 P" [ -6 0 -6  6 0 -6  6 0 6  -6 0 6 ] "integer indices" [ 0 2 1  0 3 2 ]
AttributeEnd
AttributeBegin
  Translate -1.5 0.6 0
  Material "conductor" "rgb eta" [ 0.2 0.9 1.1 ] "rgb k" [ 3.9 2.4 2.2 ] "float roughness" 0.01
  Shape "sphere" "float radius" 0.6
AttributeEnd
AttributeBegin
  Translate 1.5 0.6 0
  NamedMaterial "gold coat"
  Shape "sphere" "float radius" 0.6
AttributeEnd
AttributeBegin
  Translate 0 0.5 1.5
  Material "dielectric" "float eta" 1.5 "float roughness" 0.04
  Shape "sphere" "float radius" 0.5
AttributeEnd
AttributeBegin
  Material "coateddiffuse" "texture reflectance" "refl" "float roughness" 0.1 "bool remaproughness" false
  Translate 0 0.3 -1.5
  Shape "sphere" "float radius" 0.3
AttributeEnd
AttributeBegin
  Material "thindielectric" "float eta" 1.33
  Translate 2.5 0 1.5
  Shape "trianglemesh" "point3 P" [ 0 0 0  1 0 0  1 1.5 0  0 1.5 0 ] "integer indices" [ 0 1 2  0 2 3 ]
AttributeEnd
AttributeBegin
  Material "diffusetransmission" "rgb reflectance" [ 0.4 0.4 0.2 ] "rgb transmittance" [ 0.3 0.5 0.3 ]
  Translate -3.5 0 1.5
  Shape "trianglemesh" "point3 P" [ 0 0 0  1 0 0  1 1.5 0  0 1.5 0 ] "integer indices" [ 0 1 2  0 2 3 ]
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 8 8 8 ]
  Translate 0 100 0  # not applied to disks, as in the reference
  Shape "disk" "float radius" 1.5 "float height" -4
AttributeEnd
"""


def _f16(bits):
    return float(np.array(bits, np.uint16).view(np.float16))


def _h(x):
    return _f16(int(np.array(x, np.float16).view(np.uint16)))


def _parse(text=SCENE):
    from pltfalcor_amd.pbrt import Parser
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        prs = Parser(text).run()
    return prs, [str(x.message) for x in w]


def test_tokenizer_and_parameter_lists():
    from pltfalcor_amd.pbrt import _Stream, tokenize
    toks = list(tokenize('Shape "trianglemesh" # comment "x"\n "point3 P" [1 2 3] "bool b" true "string s" "a b"'))
    assert toks[:2] == [("w", "Shape"), ("s", "trianglemesh")] and ("s", "x") not in toks
    st = _Stream(toks[2:])
    p = st.params()
    assert p["P"] == ("point3", [1.0, 2.0, 3.0]) and p["b"] == ("bool", [True]) and p["s"] == ("string", ["a b"])


def test_materials_convert_like_the_reference():
    from pltfalcor_amd import materials as M
    prs, msgs = _parse()
    mats = prs.builder.materials
    assert [m.type for m in mats] == [M.DIFFUSE, M.CONDUCTOR, M.COATED_CONDUCTOR, M.DIELECTRIC,
                                      M.COATED_OPAQUE_DIELECTRIC, M.THIN_DIELECTRIC, M.DIELECTRIC, M.DIFFUSE]
    floor, cu, coat, glass, cdiff, thin, dtrans, lamp = mats
    assert [_f16(floor.base_color[i]) for i in range(3)] == [_h(0.6), _h(0.55), _h(0.5)]
    # conductor: constructor (gamma 3, ext IoR 1, specular 1), base colour = RGB eta, roughness sqrt(u + v)
    assert [_f16(cu.base_color[i]) for i in range(4)] == [_h(0.2), _h(0.9), _h(1.1), 1.0]
    assert _f16(cu.specular[0]) == 3.0 and _f16(cu.specular[2]) == 1.0 and _f16(cu.ior) == 1.0
    assert _f16(cu.specular[1]) == _h(math.sqrt(0.02))
    assert cu.ior_n_profile == M.NO_PROFILE and cu.ior_k_profile == M.NO_PROFILE
    # coated conductor: base colour = conductor.eta, coat IoR set, roughness left at the constructor's 0
    assert [_f16(coat.base_color[i]) for i in range(3)] == [_h(0.2), _h(0.4), _h(1.3)]
    assert _f16(coat.data1[3]) == _h(1.6) and _f16(coat.data1[0]) == 3.0 and _f16(coat.specular[1]) == 0.0
    # dielectric: Abbe 100 at the default IoR 1.5, then roughness and IoR
    assert _f16(glass.ior) == 1.5 and _f16(glass.specular[1]) == _h(math.sqrt(0.08))
    assert _f16(glass.specular[0]) == _h(.52 * 0.5 / 100) and _f16(glass.specular[3]) == 1.0
    # coated diffuse: constant texture reflectance, roughness without remapping = u + v
    assert [_f16(cdiff.base_color[i]) for i in range(3)] == [_h(0.3), _h(0.5), _h(0.7)]
    assert _f16(cdiff.specular[1]) == _h(0.2) and _f16(cdiff.data1[3]) == 1.5 and _f16(cdiff.ior) == 1.0
    # thin dielectric: constructor specular 0 and thickness 0.0003, then the IoR
    assert _f16(thin.ior) == _h(1.33) and _f16(thin.specular[3]) == _h(0.0003) and _f16(thin.specular[0]) == 0.0
    # diffuse transmission: a dielectric of roughness 0.6 with the two colours
    assert _f16(dtrans.specular[1]) == _h(0.6) and _f16(dtrans.ior) == 1.5
    assert [_f16(dtrans.transmission[i]) for i in range(3)] == [_h(0.3), _h(0.5), _h(0.3)]
    # the area light: base colour 0 with an emission profile
    assert lamp.emission_profile != M.NO_PROFILE and _f16(lamp.base_color[0]) == 0.0
    assert any("point" in m for m in msgs)  # point lights are skipped with a warning, as in the reference


def test_roughness_and_eta_rules():
    from pltfalcor_amd.pbrt import Params, Parser
    prs = Parser("WorldBegin")
    p = Params(uroughness=("float", [0.01]), vroughness=("float", [0.03]))
    assert abs(prs.roughness(p) - math.sqrt(np.float32(0.04))) < 1e-7  # both components are u + v
    p = Params(roughness=("float", [0.09]), remaproughness=("bool", [False]))
    assert abs(prs.roughness(p) - np.float32(0.18)) < 1e-7
    assert prs.scalar_eta(Params()) == 1.5
    assert abs(prs.scalar_eta(Params(eta=("rgb", [1.4, 1.5, 1.6]))) - 1.5) < 1e-6
    p = Params(reflectance=("rgb", [0.5, 0.5, 0.5]))
    assert prs.conductor_eta(p) == (1.0, 1.0, 1.0)
    with pytest.raises(NotImplementedError):
        prs.conductor_eta(Params())  # the reference default is the named metal-Cu spectrum
    with pytest.raises(NotImplementedError):
        prs.conductor_eta(Params(eta=("spectrum", ["metal-Au-eta"])))
    with pytest.raises(ValueError):
        prs.conductor_eta(Params(reflectance=("rgb", [0.5] * 3), eta=("rgb", [1.0] * 3)))


def test_geometry_transforms_and_camera():
    prs, _ = _parse()
    meshes = prs.builder.meshes
    floor, cu = meshes[0], meshes[1]
    assert np.allclose(floor.positions[:, 1], 0) and np.allclose(floor.normals, [0, 1, 0])
    # createSphere(0.6): 33 x 17 vertices, 32 x 16 x 2 triangles, all at the radius around the translation
    assert len(cu.positions) == 33 * 17 and len(cu.indices) == 32 * 16 * 2
    assert np.allclose(np.linalg.norm(cu.positions - [-1.5, 0.6, 0], axis=1), 0.6, atol=1e-6)
    disk = meshes[-1]
    # the disk: z = height, facing +z, and the Translate ignored
    assert np.allclose(disk.positions[:, 2], -4) and np.allclose(disk.normals, [0, 0, 1])
    assert np.allclose(np.linalg.norm(disk.positions[1:, :2], axis=1), 1.5, atol=1e-6)
    cam = prs.builder.camera
    assert np.allclose(list(cam.pos), [0, 3, 8]) and abs(cam.aspect - 1.5) < 1e-6
    fwd = np.asarray(cam.target) - np.asarray(cam.pos)
    want = np.array([0, 0.6, 0]) - [0, 3, 8]
    assert np.allclose(fwd / np.linalg.norm(fwd), want / np.linalg.norm(want), atol=1e-6)
    assert abs(cam.fov_y - math.radians(40)) < 1e-5  # fov is the vertical field of view (24 mm frame)
    assert (prs.width, prs.height) == (96, 64)
    env = prs.env
    assert env["image"].shape == (1, 1, 3) and np.allclose(env["image"].reshape(3), [0.05, 0.07, 0.1])
    (light,) = prs.builder.lights
    d = np.array([0.3, -10, 0.2])
    assert np.allclose(light.direction, d / np.linalg.norm(d), atol=1e-6)


def _face_normals(mesh):
    p, i = mesh.positions.astype(np.float64), mesh.indices
    n = np.cross(p[i[:, 1]] - p[i[:, 0]], p[i[:, 2]] - p[i[:, 0]])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


@pytest.mark.parametrize("prefix,reverse", [("", False), ("Scale -1 1 1", False), ("", True),
                                            ("Scale 1 -1 1", True), ("Rotate 30 1 1 0 Scale 2 1 -3", False)])
def test_face_orientation_follows_getFaceNormalW(prefix, reverse):
    """Scene::getFaceNormalW: the object-space cross product (negated for clockwise front faces)
    taken through the inverse-transpose of the transform, whatever the transform's handedness."""
    from pltfalcor_amd.pbrt import Parser
    text = (f"WorldBegin\nAttributeBegin\n{prefix}\n{'ReverseOrientation' if reverse else ''}\n"
            'Shape "trianglemesh" "point3 P" [0 0 0 1 0 0 0 1 0.2] "integer indices" [0 1 2]\nAttributeEnd\n')
    prs = Parser(text).run()
    from pltfalcor_amd.pbrt import _rotate
    m = np.eye(4)
    if prefix.startswith("Scale"):
        m = np.diag([float(x) for x in prefix.split()[1:]] + [1.0])
    elif prefix.startswith("Rotate"):
        m = _rotate(30, (1, 1, 0)) @ np.diag([2.0, 1.0, -3.0, 1.0])
    obj = np.cross([1.0, 0, 0], [0, 1.0, 0.2]) * (-1 if reverse else 1)
    want = np.linalg.inv(m[:3, :3]).T @ obj
    got = _face_normals(prs.builder.meshes[0])[0]
    assert np.allclose(got, want / np.linalg.norm(want), atol=1e-6)


def test_euler_extraction_round_trips_env_rotation():
    """extractEulerAngleXYZ inverts eulerAngleXYZ (= Rx Ry Rz, as EnvMap::setRotation builds it)."""
    from pltfalcor_amd.pbrt import _rotate, extract_euler_xyz
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = rng.uniform(-80, 80, 3)
        m = _rotate(a[0], (1, 0, 0)) @ _rotate(a[1], (0, 1, 0)) @ _rotate(a[2], (0, 0, 1))
        t = extract_euler_xyz(m)
        r = _rotate(math.degrees(t[0]), (1, 0, 0)) @ _rotate(math.degrees(t[1]), (0, 1, 0)) @ \
            _rotate(math.degrees(t[2]), (0, 0, 1))
        assert np.allclose(r, m, atol=1e-9)
    # a typical pbrt sky transform: z up rotated to y up
    t = extract_euler_xyz(_rotate(-90, (1, 0, 0)))
    assert np.allclose(np.degrees(t), [-90, 0, 0], atol=1e-9)


def _oct_to_dir(px, py):
    """Equal-area octahedral map point in [-1,1]^2 to a direction (Clarberg's mapping)."""
    d = 1 - (abs(px) + abs(py))
    r = 1 - abs(d)
    phi = ((abs(py) - abs(px)) / r + 1) * math.pi / 4 if r > 0 else 0.0
    f = r * math.sqrt(2 - r * r)
    return np.array([f * math.copysign(1, px) * math.cos(phi), f * math.copysign(1, py) * math.sin(phi),
                     math.copysign(1, d) * (1 - r * r)])


def test_octahedral_to_latlong_conversion():
    """Each lat-long texel reads the octahedral texel of its own direction: an octahedral image of a
    smooth function of direction converts to that function of the lat-long directions."""
    from pltfalcor_amd.pbrt import _latlong_to_world, oct_to_latlong
    n = 96
    f = lambda d: 1.0 + 0.5 * d[..., 0] + 0.3 * d[..., 1] - 0.2 * d[..., 2]  # noqa: E731
    oct_img = np.zeros((n, n, 3))
    for y in range(n):
        for x in range(n):
            d = _oct_to_dir((x + 0.5) / n * 2 - 1, (y + 0.5) / n * 2 - 1)
            oct_img[y, x] = f(d)
    ll = oct_to_latlong(oct_img)
    assert ll.shape == (n, 2 * n, 3)
    yy, xx = np.meshgrid(np.arange(n), np.arange(2 * n), indexing="ij")
    want = f(_latlong_to_world((xx + 0.5) / (2 * n), (yy + 0.5) / n))
    assert np.abs(ll[..., 0] - want).max() < 0.03
    assert np.allclose(oct_to_latlong(np.full((8, 8, 3), 2.5)), 2.5)
    # lat-long conventions: row 0 is +y, u = 0.5 looks down -z
    assert np.allclose(_latlong_to_world(0.5, 0.0), [0, 1, 0], atol=1e-12)
    assert np.allclose(_latlong_to_world(0.5, 0.5), [0, 0, -1], atol=1e-12)


def test_ply_reader(tmp_path):
    from pltfalcor_amd.pbrt import Parser, load_ply
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    asc = tmp_path / "q.ply"
    asc.write_text("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n"
                   "element face 1\nproperty list uchar int vertex_indices\nend_header\n"
                   + "".join(f"{a} {b} {c}\n" for a, b, c in pos) + "4 0 1 2 3\n")
    p, n, i = load_ply(asc)
    assert np.allclose(p, pos) and n is None and i.tolist() == [[0, 1, 2], [0, 2, 3]]
    binf = tmp_path / "q_bin.ply"
    hdr = ("ply\nformat binary_little_endian 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
           "property float z\nproperty float nx\nproperty float ny\nproperty float nz\n"
           "element face 2\nproperty list uchar int vertex_indices\nend_header\n").encode()
    body = b"".join(struct.pack("<6f", *v, 0, 0, 1) for v in pos)
    body += struct.pack("<B3i", 3, 0, 1, 2) + struct.pack("<B3i", 3, 0, 2, 3)
    binf.write_bytes(hdr + body)
    p, n, i = load_ply(binf)
    assert np.allclose(p, pos) and np.allclose(n, [0, 0, 1]) and i.tolist() == [[0, 1, 2], [0, 2, 3]]
    scene = f'WorldBegin\nTranslate 0 0 2\nShape "plymesh" "string filename" "{binf.name}"\n'
    (tmp_path / "s.pbrt").write_text(scene)
    prs = Parser(str(tmp_path / "s.pbrt")).run()
    assert np.allclose(prs.builder.meshes[0].positions, pos + [0, 0, 2])


def test_unsupported_input_is_reported():
    from pltfalcor_amd.pbrt import Parser
    with pytest.raises(NotImplementedError):
        Parser('WorldBegin\nMaterial "diffuse" "spectrum reflectance" "metal-Cu-eta"\nShape "sphere"\n').run()
    with pytest.raises(ValueError):
        Parser("WorldBegin\nFrobnicate 1 2 3\n").run()
    with pytest.warns(UserWarning):
        Parser('WorldBegin\nShape "cylinder"\nShape "sphere"\n').run()


@pytest.mark.gpu
def test_gpu_imported_pbrt_scene_matches_oracle(gpu):
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pbrt import load_scene
    from pltfalcor_amd.pltpt import PLTPT
    from test_pathtracer import _topology_equal
    ctx = default_context()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        s, w, h = load_scene(SCENE, ctx)
    a = s.numpy()
    p = PLTPT(ctx, {"maxBounces": 6})  # the reference's defaults otherwise (manifold NEE on)
    p.set_scene(s)
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    p.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    g_rec = p.bounce_records(w, h)
    g_col = color.cpu().numpy().reshape(-1, 4)
    o_col, o_rec = O.render(a, w, h, frame=0, max_bounces=6, records=True, do_mnee=1)
    assert np.array_equal(g_rec[0, :3], o_rec[0, :3])
    topo = _topology_equal(g_rec, o_rec)
    assert topo.mean() >= 0.99, topo.mean()
    close = color_close(g_col, o_col)
    assert close[topo].mean() >= 0.995, close[topo].mean()
    gm, om = g_col[:, :3].mean(), o_col[:, :3].mean()
    assert np.isfinite(g_col).all() and om > 0 and abs(gm - om) <= 0.02 * abs(om), (gm, om)
    p.close()
