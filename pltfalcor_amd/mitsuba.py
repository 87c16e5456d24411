"""Mitsuba scene files for the PLT path: the PLT subset of the reference's Mitsuba importer
(Source/plugins/importers/MitsubaImporter/MitsubaImporter.cpp) on top of the scene builder.

Built: `<default>` substitution, `<ref>` to named BSDFs, transforms (translate, rotate, scale,
matrix, lookat); BSDFs diffuse, dielectric / roughdielectric, thindielectric, conductor /
roughconductor, coated_conductor / coated_roughconductor, grating (sinusoidal, rectangular and
linear profiles) and twosided, converted setter by setter as `buildBSDF` does
(MitsubaImporter.cpp:287-530; the named IoR table of Tables.h; `roughness` or
`convertGSSigma2(sigma2)`); shapes rectangle, disk, cube and sphere built by the reference's mesh
generators (TriangleMesh::createQuad / createDisk / createCube / createSphere with 128 x 32
segments), obj and ply files, `flip_normals`, with a nested area emitter (`emission_profile` +
`scale`, or `radiance` RGB); emitters constant (a 1 x 1 environment map), envmap from a `.pfm` or
`.npy` image, point (sourcing area `A`) and directional (a distant light with `Omega` or
`source_area` / `distance`); the perspective sensor (`fov` over the film width, 24 mm frame
height) and the film size. Other types raise NotImplementedError naming them, as do radial
gratings, which need texture coordinates.
"""
from __future__ import annotations

import math
import re
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

from . import materials as M
from . import meshes as G
from .lights import DistantLight, PointLight
from .scene_builder import SceneBuilder

F32 = np.float32
INF = float("inf")

# Named indices of refraction and Abbe numbers (MitsubaImporter/Tables.h:14-40).
IOR_TABLE = {
    "vacuum": (1.0, INF), "helium": (1.000036, INF), "hydrogen": (1.000132, INF), "air": (1.000277, INF),
    "carbon dioxide": (1.00045, INF), "water": (1.3330, 55.0), "acetone": (1.36, 54.0), "ethanol": (1.361, 59.0),
    "carbon tetrachloride": (1.461, 49.0), "glycerol": (1.4729, INF), "benzene": (1.501, 30.3),
    "silicone oil": (1.52045, INF), "bromine": (1.661, 31.7), "water ice": (1.31, 64.0), "fused quartz": (1.458, 67.8),
    "pyrex": (1.470, 60.0), "acrylic glass": (1.49, INF), "polypropylene": (1.49, INF), "bk7": (1.5046, 64.0),
    "sodium chloride": (1.544, 42.9), "amber": (1.55, INF), "pet": (1.5750, INF), "diamond": (2.419, 55.3),
}


def convert_gs_sigma2(sigma2: float) -> float:
    """Roughness from a Gaussian-surface sigma^2 (MitsubaImporter.cpp:182-184)."""
    return float(np.clip(F32(2.0) * (F32(1.0) - F32(max(0.0, sigma2 / 1000.0)) ** F32(1.0 / 32.0)), 0.0, 1.0))


# RGB basis spectra over 11 bins of 380-720 nm (Utils/Color/SpectrumUtils.h:154-215).
_BASIS = dict(
    white=[1.0000, 1.0000, 0.9999, 0.9993, 0.9992, 0.9998, 1.0000, 1.0000, 1.0000, 1.0000, 0.0],
    cyan=[0.9710, 0.9426, 1.0007, 1.0007, 1.0007, 1.0007, 0.1564, 0.0000, 0.0000, 0.0000, 0.0],
    magenta=[1.0000, 1.0000, 0.968, 0.22295, 0.0000, 0.0458, 0.8369, 1.0000, 1.0000, 0.9959, 0.0],
    yellow=[0.0001, 0.0000, 0.1088, 0.6651, 1.0000, 1.0000, 0.9996, 0.9586, 0.9685, 0.9840, 0.0],
    red=[0.1012, 0.0515, 0.0000, 0.0000, 0.0000, 0.0000, 0.8325, 1.0149, 1.0149, 1.014, 0.0],
    green=[0.0000, 0.0000, 0.0273, 0.7937, 1.0000, 0.9418, 0.1719, 0.0000, 0.0000, 0.0025, 0.0],
    blue=[1.0000, 1.0000, 0.8916, 0.3323, 0.0000, 0.0000, 0.0003, 0.0369, 0.0483, 0.0496, 0.0],
)


def rgb_to_spectrum(rgb, samples: int = 16) -> np.ndarray:
    """SpectrumUtils::rgbToSpectrum<float>: `samples` values over 380-720 nm (float arithmetic)."""
    r, g, b = (F32(x) for x in rgb)
    B = {k: np.asarray(v, F32) for k, v in _BASIS.items()}
    out = np.zeros(samples, F32)
    for i in range(samples):
        w = F32(380.0) + (F32(720.0) - F32(380.0)) / F32(samples - 1) * F32(i)
        k = min(10, max(0, int((w - F32(380.0)) / (F32(720.0) - F32(380.0)) * F32(10.0))))
        sd = F32(0.0)
        if r <= g and r <= b:
            sd += B["white"][k] * r
            if g <= b:
                sd += B["cyan"][k] * (g - r)
                sd += B["blue"][k] * (b - g)
            else:
                sd += B["cyan"][k] * (b - r)
                sd += B["green"][k] * (g - b)
        elif g <= r and g <= b:
            sd += B["white"][k] * g
            if r <= b:
                sd += B["magenta"][k] * (r - g)
                sd += B["blue"][k] * (b - r)
            else:
                sd += B["magenta"][k] * (b - g)
                sd += B["red"][k] * (r - b)
        else:
            sd += B["white"][k] * b
            if r <= g:
                sd += B["yellow"][k] * (r - b)
                sd += B["green"][k] * (g - r)
            else:
                sd += B["yellow"][k] * (g - b)
                sd += B["red"][k] * (r - g)
        out[i] = sd
    return out


# ------------------------------------------------------------------------------ XML properties


def _floats(text):
    return [float(x) for x in re.split(r"[,\s]+", text.strip()) if x]


def _rot(axis, angle_deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(math.radians(angle_deg)), math.sin(math.radians(angle_deg))
    x, y, z = a
    m = np.eye(4)
    m[:3, :3] = [[c + (1 - c) * x * x, (1 - c) * x * y - s * z, (1 - c) * x * z + s * y],
                 [(1 - c) * y * x + s * z, c + (1 - c) * y * y, (1 - c) * y * z - s * x],
                 [(1 - c) * z * x - s * y, (1 - c) * z * y + s * x, c + (1 - c) * z * z]]
    return m


def _lookat(origin, target, up):
    o, t, u = (np.asarray(v, np.float64) for v in (origin, target, up))
    d = (t - o) / np.linalg.norm(t - o)
    left = np.cross(u, d)
    left /= np.linalg.norm(left)
    nu = np.cross(d, left)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = left, nu, d, o
    return m


class Props(dict):
    """Named properties of one element, plus nested objects by tag."""

    def __init__(self):
        super().__init__()
        self.children = []


class Parser:
    def __init__(self, path_or_text, tables_dir: str | None = None):
        p = Path(str(path_or_text))
        if "<" not in str(path_or_text) and p.exists():
            self.base = p.parent
            root = ET.parse(p).getroot()
        else:
            self.base = Path(".")
            root = ET.fromstring(path_or_text)
        if root.tag != "scene":
            raise ValueError("not a Mitsuba scene (root element is not <scene>)")
        self.defaults = {}
        self.ids = {}
        self.builder = SceneBuilder(tables_dir)
        self.width, self.height = 768, 576
        self.env = None
        self.root = root

    def sub(self, s):
        return re.sub(r"\$(\w+)", lambda m: str(self.defaults.get(m.group(1), m.group(0))), s)

    def attr(self, e, name, default=None):
        v = e.get(name)
        return default if v is None else self.sub(v)

    def transform(self, e):
        m = np.eye(4)
        for c in e:
            if c.tag == "translate":
                v = _floats(self.attr(c, "value")) if c.get("value") else [float(self.attr(c, k, 0)) for k in "xyz"]
                t = np.eye(4)
                t[:3, 3] = v
                m = t @ m
            elif c.tag == "scale":
                if c.get("value") is not None:
                    v = _floats(self.attr(c, "value"))
                    v = v * 3 if len(v) == 1 else v
                else:
                    v = [float(self.attr(c, k, 1)) for k in "xyz"]
                m = np.diag(list(v) + [1.0]) @ m
            elif c.tag == "rotate":
                if c.get("value") is not None:
                    axis = _floats(self.attr(c, "value"))
                else:
                    axis = [float(self.attr(c, k, 0)) for k in "xyz"]
                m = _rot(axis, float(self.attr(c, "angle"))) @ m
            elif c.tag == "matrix":
                m = np.asarray(_floats(self.attr(c, "value")), np.float64).reshape(4, 4) @ m
            elif c.tag == "lookat":
                m = _lookat(_floats(self.attr(c, "origin")), _floats(self.attr(c, "target")),
                            _floats(self.attr(c, "up", "0,1,0"))) @ m
            else:
                raise NotImplementedError(f"transform element <{c.tag}>")
        return m

    def props(self, e) -> Props:
        p = Props()
        for c in e:
            name = self.attr(c, "name")
            if c.tag in ("float",):
                p[name] = float(self.attr(c, "value"))
            elif c.tag == "integer":
                p[name] = int(self.attr(c, "value"))
            elif c.tag == "boolean":
                p[name] = self.attr(c, "value").lower() == "true"
            elif c.tag == "string":
                p[name] = self.attr(c, "value")
            elif c.tag in ("rgb", "color"):
                v = _floats(self.attr(c, "value"))
                p[name] = tuple(v * 3 if len(v) == 1 else v)
            elif c.tag in ("point", "vector"):
                p[name] = tuple(_floats(self.attr(c, "value"))) if c.get("value") else \
                    tuple(float(self.attr(c, k, 0)) for k in "xyz")
            elif c.tag == "transform":
                p[name] = self.transform(c)
            elif c.tag == "ref":
                p.children.append(("ref", self.attr(c, "id"), name))
            elif c.tag == "spectrum":
                raise NotImplementedError("<spectrum> properties (give RGB, a material name or an emission_profile)")
            else:
                p.children.append((c.tag, c, name))
        return p

    # --- BSDFs (buildBSDF, MitsubaImporter.cpp:287-460)
    def ior(self, p, name, default):
        if name in p and isinstance(p[name], float):
            return p[name], INF
        key = p.get(name, default).lower()
        return IOR_TABLE.get(key, IOR_TABLE["air"])

    def roughness(self, p):
        return p["roughness"] if "roughness" in p else convert_gs_sigma2(p.get("sigma2", 100.0))

    def conductor_profiles(self, p, message):
        """n, k profile ids from a material name or eta / k values (RGB through rgbToSpectrum)."""
        if "material" in p:
            return self.builder.addSpectralProfileFromMaterial(p["material"])
        if "eta" in p and "k" in p:
            ids = []
            for key in ("eta", "k"):
                self.builder.profiles.append(_sampled_profile(rgb_to_spectrum(p[key]), False))
                ids.append(len(self.builder.profiles) - 1)
            return tuple(ids)
        raise ValueError(message)

    def bsdf(self, e):
        t = self.attr(e, "type")
        p = self.props(e)
        h = M.h
        ext = self.ior(p, "ext_ior", "air")[0]
        if t == "twosided":
            inner = [c for kind, c, _ in p.children if kind == "bsdf"] + \
                    [self.ids[c] for kind, c, _ in p.children if kind == "ref"]
            if len(inner) != 1:
                raise ValueError("'twosided' BSDF must contain a single nested BSDF")
            return inner[0] if not isinstance(inner[0], ET.Element) else self.bsdf(inner[0])
        if t == "diffuse":
            m = M.blank(M.DIFFUSE)
            c = p.get("reflectance", (0.5, 0.5, 0.5))
            for i in range(3):
                m.base_color[i] = h(c[i])
            return m
        if t in ("dielectric", "roughdielectric"):
            m = M.blank(M.DIELECTRIC)  # PLTDielectricMaterial defaults, then the importer's setters
            m.specular[3] = h(1.0)
            if "specular_reflectance" in p:
                for i in range(3):
                    m.base_color[i] = h(p["specular_reflectance"][i])
            if "specular_transmittance" in p:
                for i in range(3):
                    m.transmission[i] = h(p["specular_transmittance"][i])
            n, abbe = self.ior(p, "int_ior", "bk7")
            m.specular[3] = h(ext)
            m.ior = h(n)
            ior_h = F32(np.array(m.ior, np.uint16).view(np.float16))
            m.specular[0] = h(float(F32(.52) * (ior_h - F32(1.0)) / F32(abbe)))  # setAbbeNumber
            m.specular[1] = h(self.roughness(p))
            return m
        if t == "thindielectric":
            m = M.thin_dielectric(1.5, 0.0003)  # the constructor: specular 0, optic axis +z, birefringence 0
            m.specular[0] = h(0.0)
            m.ior = h(self.ior(p, "int_ior", "bk7")[0])
            m.specular[3] = h(p.get("thickness", .001))
            axis = np.asarray(p.get("optic_axis", (0.0, 0.0, 1.0)), np.float32)
            axis = axis / np.float32(np.sqrt(np.dot(axis, axis)))
            if axis[2] < 0:
                axis = -axis
            m.base_color[1], m.base_color[2] = h(float(axis[0])), h(float(axis[1]))
            m.base_color[0] = h(p.get("birefringence_scale", 1.0))
            b = p.get("birefringence", 0.0)
            m.specular[0] = h(b[0] if isinstance(b, tuple) else b)
            return m
        if t == "grating":
            kinds = {"sinusoidal": 0, "rectangular": 1, "linear": 2}
            kind = p.get("type", "sinusoidal")
            if kind.startswith("radial_"):
                raise NotImplementedError("radial gratings (they need texture coordinates)")
            if kind not in kinds:
                raise ValueError("Unsupported 'type' name.")
            n_id, k_id = self.conductor_profiles(p, "'grating' BSDF must specify either 'eta' and 'k' IOR values or "
                                                      "'material' name.")
            m = M.grated_conductor(p.get("pitch", 5.0), float(F32(F32(p.get("grating_direction", 0.0)) * math.pi / 180.0)),
                                   n_id, k_id, lobes=p.get("lobes", 3), grating_type=kinds[kind], height=1.0,
                                   height_scale=p.get("q", .5), roughness=0.0)
            m.ior = h(ext)
            m.specular[1] = h(p.get("amplify", 1.0))
            return m
        if t in ("conductor", "roughconductor"):
            m = M.blank(M.CONDUCTOR)
            for i in range(4):
                m.specular[i] = h(1.0)
            m.specular[0] = h(3.0)
            c = p.get("specular_reflectance", (1.0, 1.0, 1.0))
            for i in range(3):
                m.base_color[i] = h(c[i])
            m.ior_n_profile, m.ior_k_profile = self.conductor_profiles(
                p, "'conductor' BSDF must specify either 'eta' and 'k' or 'material'")
            m.ior = h(ext)
            m.specular[1] = h(self.roughness(p))
            return m
        if t in ("coated_conductor", "coated_roughconductor"):
            nested = [c for kind, c, _ in p.children if kind == "bsdf"] + \
                     [self.ids[c] for kind, c, _ in p.children if kind == "ref"]
            if len(nested) != 1:
                raise ValueError("'coated_conductor' BSDF must contain a single nested 'conductor' BSDF")
            cond = nested[0] if not isinstance(nested[0], ET.Element) else self.bsdf(nested[0])
            if cond.type != M.CONDUCTOR:
                raise ValueError("'coated_conductor' BSDF must contain only a nested 'conductor' BSDF")
            m = M.blank(M.COATED_CONDUCTOR)
            m.ior_n_profile, m.ior_k_profile = cond.ior_n_profile, cond.ior_k_profile
            m.ior = cond.ior
            for i in range(4):
                m.specular[i] = cond.specular[i]
            for i in range(4):
                m.base_color[i] = cond.base_color[i]
            m.data1[3] = h(self.ior(p, "int_ior", "bk7")[0])
            m.data1[0] = h(p.get("thickness", 50.0))
            return m
        raise NotImplementedError(f"BSDF type '{t}'")

    # --- shapes (buildShape, MitsubaImporter.cpp:575-745)
    def shape_mesh(self, t, p):
        """The reference's meshes for each shape type (TriangleMesh::createQuad / createDisk /
        createCube / createSphere, OBJ and PLY files), placed by `to_world`."""
        m = p.get("to_world", np.eye(4))
        if t == "rectangle":
            pos, nrm, idx, cw = G.create_quad((2.0, 2.0))
            m = m @ G.Y_TO_Z
        elif t == "disk":
            pos, nrm, idx, cw = G.create_disk(1.0)
            m = m @ G.Y_TO_Z
        elif t == "cube":
            pos, nrm, idx, cw = G.create_cube((2.0, 2.0, 2.0))
        elif t == "sphere":
            pos, nrm, idx, cw = G.create_sphere(p.get("radius", 1.0), 128, 32)
            m = m @ G.translate(p.get("center", (0.0, 0.0, 0.0)))
        elif t == "obj":
            pos, nrm, idx = _load_obj(self.base / p["filename"])
            cw = False
        elif t == "ply":
            pos, nrm, idx = G.load_ply(self.base / p["filename"])
            if p.get("face_normals", False):
                pos, nrm, idx = G.face_normal_mesh(pos, idx)
            elif nrm is None:
                nrm = G.smooth_normals(pos, idx)
            cw = False
        else:
            raise NotImplementedError(f"shape type '{t}'")
        if p.get("flip_normals", False) and t in ("rectangle", "disk", "cube", "sphere"):
            pos, nrm, idx, cw = G.flip_normals(pos, nrm, idx, cw)
        pos, nrm, cw = G.apply_transform(m, pos, nrm, cw)
        return G.bake(pos, nrm, idx, cw)

    def emission_profile(self, p):
        """getEmissionSpectrum (MitsubaImporter.cpp:160-168)."""
        if "emission_profile" in p:
            return self.builder.addSpectralProfileForEmitterType(p["emission_profile"], p.get("scale", 1.0))
        pr = _sampled_profile(rgb_to_spectrum(p["radiance"]), True)
        self.builder.profiles.append(pr)
        return len(self.builder.profiles) - 1

    def shape(self, e):
        t = self.attr(e, "type")
        p = self.props(e)
        mat = None
        for kind, c, _ in p.children:
            if kind == "bsdf":
                mat = self.bsdf(c)
            elif kind == "ref":
                mat = self.ids[c]
        if mat is None:
            mat = M.blank(M.DIFFUSE)  # the importer's default PLTDiffuse material
        for kind, c, _ in p.children:
            if kind == "emitter" and self.attr(c, "type") == "area":
                if mat.type != M.DIFFUSE:
                    raise NotImplementedError("area emitters on non-diffuse PLT materials")
                mat.emission_profile = self.emission_profile(self.props(c))
        pos, nrm, idx = self.shape_mesh(t, p)
        mid = self._material_id(mat)
        self.builder.addTriangleMesh(pos, nrm, idx, mid)

    def _material_id(self, mat):
        for i, m in enumerate(self.builder.materials):
            if m is mat:
                return i
        return self.builder.addMaterial(mat)

    # --- emitters (buildEmitter, MitsubaImporter.cpp:846-930)
    def emitter(self, e):
        t = self.attr(e, "type")
        p = self.props(e)
        if t == "constant":
            self.env = dict(image=np.asarray(p["radiance"], np.float32).reshape(1, 1, 3), scale=1.0)
        elif t == "envmap":
            self.env = dict(image=_load_image(self.base / p["filename"]), scale=p.get("scale", 1.0))
        elif t == "point":
            light = PointLight(p.get("name", ""))
            light.setIntensity(self.emission_profile_named(p, "intensity"))
            light.position = p.get("position", tuple(p.get("to_world", np.eye(4))[:3, 3]))
            if "A" not in p:
                raise ValueError("point emitter must specify the sourcing area 'A'")
            light.setLightArea(p["A"])
            self.builder.addLight(light)
        elif t == "directional":
            light = DistantLight(p.get("name", ""))
            light.setIntensity(self.emission_profile_named(p, "irradiance"))
            light.direction = p["direction"]
            if "Omega" in p:
                light.setSourceSolidAngle(p["Omega"])
            elif "source_area" in p and "distance" in p:
                light.setSourceSolidAngle(p["source_area"] / p["distance"] / p["distance"])
            else:
                raise ValueError("directional emitter must specify 'Omega' or 'source_area' and 'distance'")
            self.builder.addLight(light)
        else:
            raise NotImplementedError(f"emitter type '{t}'")

    def emission_profile_named(self, p, name):
        if "emission_profile" in p:
            return self.builder.addSpectralProfileForEmitterType(p["emission_profile"], p.get("scale", 1.0))
        pr = _sampled_profile(rgb_to_spectrum(p[name]), True)
        self.builder.profiles.append(pr)
        return len(self.builder.profiles) - 1

    # --- sensor (buildSensor, MitsubaImporter.cpp:744-845)
    def sensor(self, e):
        t = self.attr(e, "type")
        if t not in ("perspective", "thinlens"):
            raise NotImplementedError(f"sensor type '{t}'")
        p = self.props(e)
        for kind, c, _ in p.children:
            if kind == "film":
                fp = self.props(c)
                self.width, self.height = fp.get("width", 768), fp.get("height", 576)
        if "focal_length" in p and "fov" in p:
            raise ValueError("Cannot specify both 'focal_length' and 'fov'.")
        focal = p.get("focal_length", 50.0)
        if "fov" in p:
            film_w = 24.0 / self.height * self.width
            focal = film_w / (2.0 * math.tan(math.radians(p["fov"]) / 2.0))  # fovYToFocalLength over the film width
        m = p.get("to_world", np.eye(4))
        pos, fwd, up = m[:3, 3], m[:3, 2], m[:3, 1]
        self.builder.addCamera(tuple(pos), tuple(pos + fwd), tuple(up), focalLength=focal, frameHeight=24.0,
                               aspect=self.width / self.height, focalDistance=p.get("focus_distance", 1.0))

    def run(self):
        for e in self.root:
            if e.tag == "default":
                self.defaults[e.get("name")] = e.get("value")
        for e in self.root:
            if e.tag == "bsdf":
                self.ids[self.attr(e, "id")] = self.bsdf(e)
        for e in self.root:
            if e.tag == "shape":
                self.shape(e)
            elif e.tag == "emitter":
                self.emitter(e)
            elif e.tag == "sensor":
                self.sensor(e)
        return self


def _sampled_profile(values, compute_icdf):
    import ctypes as C
    from . import _lib
    v = np.ascontiguousarray(values, np.float32)
    out = _lib.SpectralProfile()
    _lib.check(_lib.load().plt_spectral_profile_build_sampled(len(v), 380.0, 720.0,
                                                             v.ctypes.data_as(C.POINTER(C.c_float)), int(compute_icdf),
                                                             C.byref(out)))
    return out


def _load_obj(path):
    vs, vns, faces = [], [], []
    for line in Path(path).read_text().splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            vs.append([float(x) for x in t[1:4]])
        elif t[0] == "vn":
            vns.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            corner = [c.split("/") for c in t[1:]]
            for k in range(1, len(corner) - 1):  # fan triangulation
                faces.append([corner[0], corner[k], corner[k + 1]])
    pos, nrm, idx, cache = [], [], [], {}
    vs = np.asarray(vs, np.float64)
    for f in faces:
        tri = []
        for c in f:
            vi = int(c[0]) - 1
            ni = int(c[2]) - 1 if len(c) > 2 and c[2] else None
            key = (vi, ni)
            if key not in cache:
                cache[key] = len(pos)
                pos.append(vs[vi])
                nrm.append(vns[ni] if ni is not None else None)
            tri.append(cache[key])
        idx.append(tri)
    idx = np.asarray(idx, np.uint32)
    pos = np.asarray(pos)
    if any(n is None for n in nrm):  # smooth normals from area-weighted face normals
        acc = np.zeros_like(pos)
        fn = np.cross(pos[idx[:, 1]] - pos[idx[:, 0]], pos[idx[:, 2]] - pos[idx[:, 0]])
        for k in range(3):
            np.add.at(acc, idx[:, k], fn)
        nrm = acc
    return pos, np.asarray(nrm, np.float64), idx


def _load_image(path):
    p = Path(path)
    if p.suffix == ".npy":
        return np.load(p).astype(np.float32)[..., :3]
    if p.suffix == ".pfm":
        with open(p, "rb") as f:
            kind = f.readline().strip()
            w, h = (int(x) for x in f.readline().split())
            scale = float(f.readline())
            data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4")
        ch = 3 if kind == b"PF" else 1
        img = data.reshape(h, w, ch)[::-1].astype(np.float32)
        return np.repeat(img, 3, 2) if ch == 1 else img
    raise NotImplementedError(f"environment image format '{p.suffix}' (use .pfm or .npy)")


def load_scene(path_or_text, ctx, tables_dir: str | None = None):
    """Reads a Mitsuba scene (file path or XML text) into a Scene on `ctx` (host-only when None).
    Returns (scene, width, height)."""
    prs = Parser(path_or_text, tables_dir).run()
    scene = prs.builder.getScene(ctx)
    if prs.env is not None:
        scene.set_env_map(prs.env["image"], intensity=prs.env["scale"])
    return scene, prs.width, prs.height


__all__ = ["load_scene", "Parser", "rgb_to_spectrum", "convert_gs_sigma2", "IOR_TABLE"]
