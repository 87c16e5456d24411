"""Scene ingest: the `.spd` reader, the material-name and emitter-type profile helpers, the
standard-material → PLT conversion and the scene builder (CPU), and a render of a built scene
with tabulated Cu and Au and the FL4 emission spectrum against the oracle (GPU).

Fixtures: tests/golden/tables (copies of four optical-constant tables and one emission table of
the reference's Source/Tables). When the reference's tables are present, every file of them is
also read."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from parity import color_close

TABLES = Path(__file__).resolve().parent / "golden" / "tables"
REF_TABLES = Path("/root/reference/Source/Tables")


def _parse_spd(path):
    """The table format restated independently: trimmed lines, '#' and blank lines skipped, two
    leading numbers per line, the first line without them ends the table."""
    wl, val = [], []
    for line in Path(path).read_bytes().decode("latin-1").split("\n"):
        t = line.strip()
        if not t or t[0] == "#":
            continue
        parts = t.split()
        try:
            a, b = np.float32(parts[0]), np.float32(parts[1])
        except (ValueError, IndexError):
            break
        wl.append(a)
        val.append(b)
    return np.array(wl, np.float32), np.array(val, np.float32)


def _tables():
    files = sorted(TABLES.glob("*/*.spd"))
    if REF_TABLES.is_dir():
        files += sorted(REF_TABLES.glob("*/*.spd"))
    return files


def test_spd_reader_matches_restatement_on_every_table():
    from pltfalcor_amd.scene_builder import load_spd
    files = _tables()
    assert len(files) >= 5
    for f in files:
        wl, v = load_spd(f)
        rw, rv = _parse_spd(f)
        assert np.array_equal(wl, rw) and np.array_equal(v, rv), f
        assert len(wl) >= 2 and np.all(np.diff(wl) >= 0), f


def test_spd_known_entries():
    """Cu.eta.spd (luxpop): first entry 302.400421 nm -> n 1.38; 60 lines with a 5-line header."""
    from pltfalcor_amd.scene_builder import load_spd
    wl, n = load_spd(TABLES / "ior" / "Cu.eta.spd")
    assert (wl[0], n[0]) == (np.float32(302.400421), np.float32(1.38))
    assert len(wl) == 55
    wl2, k = load_spd(TABLES / "ior" / "Cu.k.spd")
    assert np.array_equal(wl, wl2)
    # copper's visible optical constants: n < 1.2 and k > 2 across 550-700 nm
    band = (wl > 550) & (wl < 700)
    assert np.all(n[band] < 1.2) and np.all(k[band] > 2.0)
    from pltfalcor_amd import _lib
    with pytest.raises(_lib.PLTError):
        load_spd(TABLES / "ior" / "Missing.eta.spd")


def test_material_profiles_are_the_table_profiles():
    """addSpectralProfileFromMaterial = the profile builder (bit-exact vs the oracle) on the two
    tables, without inverse CDF; "none" = the constant n = 1, k = 0 profiles (one bin, 380-720 nm)."""
    from pltfalcor_amd.scene_builder import SceneBuilder, load_spd
    b = SceneBuilder(str(TABLES))
    n_id, k_id = b.addSpectralProfileFromMaterial("Cu")
    for pid, suffix in ((n_id, "eta"), (k_id, "k")):
        wl, v = load_spd(TABLES / "ior" / f"Cu.{suffix}.spd")
        assert bytes(b.profiles[pid]) == O.build_profile(wl, v, False)
    n0, k0 = b.addSpectralProfileFromMaterial("none")
    for pid, val in ((n0, 1.0), (k0, 0.0)):
        p = b.profiles[pid]
        assert (p.bins, p.min_wavelength, p.max_wavelength, p.pdf[0], p.intensity) == (1, 380.0, 720.0, val, 1.0)
    from pltfalcor_amd import _lib
    with pytest.raises(_lib.PLTError):
        b.addSpectralProfileFromMaterial("Unobtainium")


def test_emitter_profile_scales_intensity_and_keeps_the_distribution():
    from pltfalcor_amd.scene_builder import SceneBuilder, load_spd
    b = SceneBuilder(str(TABLES))
    a = b.profiles[b.addSpectralProfileForEmitterType("FL4")]
    s = b.profiles[b.addSpectralProfileForEmitterType("FL4", 4.0)]
    wl, v = load_spd(TABLES / "emission" / "FL4.spd")
    assert bytes(a) == O.build_profile(wl, v, True)
    assert s.intensity == 4.0 * a.intensity  # a power-of-two scale is exact through the sums
    assert list(s.pdf) == list(a.pdf) and list(s.icdf) == list(a.icdf)
    assert abs(sum(a.pdf[:a.bins]) - 1.0) < 1e-5
    from pltfalcor_amd import _lib
    with pytest.raises(_lib.PLTError):
        b.addSpectralProfileForEmitterType("NoSuchLamp")


def _f16(bits):
    return float(np.array(bits, np.uint16).view(np.float16))


def test_standard_material_conversion():
    """genMaterial (StandardMaterial.cpp:219-276): metal -> conductor with the named n, k tables, ext
    IoR 1, gamma 3; isDiffuse -> diffuse; no transmission -> opaque dielectric with IoR >= 1.1;
    transmission -> dielectric with Abbe number 80 at its IoR, ext IoR 1, transmission colour."""
    from pltfalcor_amd import materials as M
    from pltfalcor_amd.scene_builder import SceneBuilder, StandardMaterialPLTWrapper as W
    b = SceneBuilder(str(TABLES))
    cu = b.materials[b.addMaterial(W("cu", metalName="Cu", specularParams=(0, 0.2, 1, 0), baseColor=(1, .5, .25, 1)))]
    assert cu.type == M.CONDUCTOR and (cu.ior_n_profile, cu.ior_k_profile) == (0, 1)
    assert [_f16(x) for x in cu.specular] == [3.0, _f16(M.h(0.2)), 1.0, 0.0]
    assert [_f16(x) for x in cu.base_color] == [1.0, 0.5, 0.25, 1.0] and _f16(cu.ior) == 1.0
    d = b.materials[b.addMaterial(W("wall", isDiffuse=True, baseColor=(.8, .7, .6, 1)))]
    assert d.type == M.DIFFUSE and _f16(d.base_color[1]) == _f16(M.h(.7))
    o = b.materials[b.addMaterial(W("plastic", indexOfRefraction=1.0, specularParams=(0, .3, 0, 0)))]
    assert o.type == M.OPAQUE_DIELECTRIC and _f16(o.ior) == _f16(M.h(1.1)) and _f16(o.specular[1]) == _f16(M.h(.3))
    g = b.materials[b.addMaterial(W("glass", indexOfRefraction=1.52, specularTransmission=0.9,
                                    transmissionColor=(1, .5, 1), specularParams=(0, .05, 0, 0)))]
    assert g.type == M.DIELECTRIC and _f16(g.ior) == _f16(M.h(1.52))
    B = _f16(g.specular[0])
    assert abs(.52 * (_f16(g.ior) - 1) / B - 80.0) < 0.2  # Abbe number back from the stored dispersion
    assert _f16(g.specular[3]) == 1.0 and _f16(g.specular[1]) == _f16(M.h(.05))
    assert [_f16(x) for x in g.transmission] == [_f16(M.h(.9)), _f16(M.h(.45)), _f16(M.h(.9))]


def _uv_sphere(center, radius, n_lat=24, n_lon=48):
    th = np.linspace(0, np.pi, n_lat + 1)
    ph = np.linspace(0, 2 * np.pi, n_lon + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    nrm = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pos = np.asarray(center) + radius * nrm
    idx = []
    for i in range(n_lat):
        for j in range(n_lon):
            a, b = i * (n_lon + 1) + j, i * (n_lon + 1) + j + 1
            c, d = a + n_lon + 1, b + n_lon + 1
            if i > 0:
                idx.append([a, b, c])
            if i < n_lat - 1:
                idx.append([b, d, c])
    return pos.astype(np.float32), nrm.astype(np.float32), np.array(idx, np.uint32)


def _quad(y, hs, ny, cx=0.0, cz=0.0):
    pos = np.array([[cx - hs, y, cz - hs], [cx + hs, y, cz - hs], [cx + hs, y, cz + hs], [cx - hs, y, cz + hs]], np.float32)
    nrm = np.tile(np.array([0, ny, 0], np.float32), (4, 1))
    idx = np.array([[0, 2, 1], [0, 3, 2]] if ny > 0 else [[0, 1, 2], [0, 2, 3]], np.uint32)
    return pos, nrm, idx


def _built_scene(ctx):
    from pltfalcor_amd import materials as M
    from pltfalcor_amd.scene_builder import SceneBuilder, StandardMaterialPLTWrapper as W
    b = SceneBuilder(str(TABLES))
    lamp = b.addSpectralProfileForEmitterType("FL4", 4.0)
    ground = b.addMaterial(W("ground", isDiffuse=True, baseColor=(.6, .6, .6, 1)))
    light = b.addMaterial(M.diffuse(0.0, 0.0, 0.0, emission=lamp))
    cu = b.addMaterial(W("copper", metalName="Cu", specularParams=(0, .15, 1, 0)))
    au = b.addMaterial(W("gold", metalName="Au", specularParams=(0, .05, 1, 0)))
    glass = b.addMaterial(W("glass", indexOfRefraction=1.5, specularTransmission=1.0, specularParams=(0, .02, 0, 0)))
    b.addTriangleMesh(*_quad(0.0, 4.0, 1), ground)
    b.addTriangleMesh(*_quad(4.0, 1.0, -1), light)
    for mat, cx in ((cu, -1.3), (au, 0.0), (glass, 1.3)):
        b.addTriangleMesh(*_uv_sphere((cx, 0.6, 0.0), 0.6), mat)
    b.addCamera((0.0, 2.2, 4.5), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), focalLength=35.0, aspect=1.5)
    return b.getScene(ctx)


def test_scene_builder_host_scene():
    s = _built_scene(None)
    a = s.numpy()
    assert a["n_profiles"] == 5 and len(a["mesh_material"]) == 5
    assert a["indices"].max() < len(a["positions"])
    cam = a["camera"]
    assert abs(cam[9] - 2 * math.atan(12 / 35)) < 1e-6  # 35 mm lens on the 24 mm frame


@pytest.mark.gpu
def test_gpu_built_scene_with_tabulated_metals_matches_oracle(gpu):
    """Cu and Au spheres with the tabulated optical constants, a glass sphere and an FL4 lamp, built
    through the scene builder: primary hits bit-identical, paths and colour to the path tracer's
    parity bar against the oracle rendering the same arrays."""
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from test_pathtracer import _topology_equal
    ctx = default_context()
    s = _built_scene(ctx)
    a = s.numpy()
    w, h = 96, 64
    p = PLTPT(ctx, {"maxBounces": 6, "doMNEE": 0})
    p.set_scene(s)
    color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    albedo, normal = torch.empty_like(color), torch.empty_like(color)
    p.execute(w, h, color, albedo, normal)
    torch.cuda.synchronize()
    g_rec = p.bounce_records(w, h)
    g_col = color.cpu().numpy().reshape(-1, 4)
    o_col, o_rec = O.render(a, w, h, frame=0, max_bounces=6, records=True)
    assert np.array_equal(g_rec[0, :3], o_rec[0, :3])
    hit = g_rec[0, 0] != 0xFFFFFFFF
    assert hit.mean() > 0.5
    topo = _topology_equal(g_rec, o_rec)
    assert topo.mean() >= 0.99, topo.mean()
    close = color_close(g_col, o_col)
    assert close[topo].mean() >= 0.995, close[topo].mean()
    assert np.isfinite(g_col).all() and g_col[:, :3].max() > 0
    gm, om = g_col[:, :3].mean(), o_col[:, :3].mean()
    assert abs(gm - om) <= 0.02 * abs(om), (gm, om)
    p.close()
