"""PBRT-v4 scene files for the PLT path: the PLT-material branch (`PBRTImporter:usePBRTMaterials`)
of the reference's PBRT importer (Source/plugins/importers/PBRTImporter/{Parser,Builder,
PBRTImporter}.cpp) on top of the scene builder.

Built:
- The directive grammar: comments, quoted strings, parameter lists `"type name" [values]`, Include
  and Import.
- The graphics state: Identity, Translate, Scale, Rotate, LookAt (left-handed), Transform and
  ConcatTransform (row-vector matrices, transposed), CoordinateSystem / CoordSysTransform,
  AttributeBegin / End, TransformBegin / End, ReverseOrientation, ObjectBegin / End /
  ObjectInstance.
- Materials, converted setter by setter as createMaterial does with PLT materials
  (PBRTImporter.cpp:969-1256), all double-sided as the reference sets them:
  - diffuse → PLT diffuse
  - coateddiffuse → coated opaque dielectric
  - conductor → conductor (base colour = RGB eta, as the reference sets it; no n, k profiles)
  - coatedconductor → coated conductor
  - dielectric → dielectric
  - thindielectric → thin dielectric
  - diffusetransmission → dielectric of roughness 0.6 with reflectance and transmittance
- Roughness as getRoughness computes it: the sum u + v, square-rooted when `remaproughness`.
- Scalar eta as getScalarEta computes it: the float, or the RGB mean.
- Shapes (PBRTImporter.cpp:1322-1505):
  - sphere: TriangleMesh::createSphere, 32 x 16 segments
  - disk: TriangleMesh::createDisk, 32 segments, lifted to `height` on the z axis; the shape's
    transform is not applied, as in the reference
  - trianglemesh
  - plymesh: ASCII and binary little-endian PLY with float positions and normals and list faces
  Each shape's face normal is as the reference computes it (getFaceNormalW): winding follows the
  mesh's front-face flag, toggled by ReverseOrientation and by mirroring transforms.
- Lights (createLight, PBRTImporter.cpp:635-744):
  - distant → a DirectionalLight along the transformed `to - from`; its intensity is left unset,
    as in the reference
  - infinite with `L` → a 1 x 1 environment map; its `scale` is ignored, as in the reference
  - infinite with `filename` → the equal-area octahedral image (.pfm or .npy), converted to a
    2N x N lat-long map with bilinear, clamped lookups (EnvMapConverter.cs.slang); intensity =
    `scale`; rotation = the XYZ Euler angles of the light's transform
    (glm::extractEulerAngleXYZ, then EnvMap::setRotation)
- The perspective camera: `fov` as the vertical field of view over a 24 mm frame; the node
  transform is world-from-camera x diag(1, 1, -1). The Film resolution is read too.

Where the reference has no PLT equivalent:
- A shape with no material gets the PLT diffuse material at its defaults. The reference's fallback
  is a StandardMaterial, which no PLT pass evaluates.
- A shape with a `diffuse` AreaLightSource gets a PLT diffuse material of base colour 0. Its
  emission profile is `L` x `scale` taken through rgbToSpectrum, as the Mitsuba importer converts
  RGB radiance. The reference gives such shapes a StandardMaterial and leaves their emission unset.
- A triangle mesh without normals gets area-weighted vertex normals. The reference passes zero
  normals on.

Spectra are accepted as `rgb` or `float` values only. Named, tabulated and blackbody spectra need
pbrt's CIE and named-spectrum tables, which are not part of this build, so they raise
NotImplementedError. So does a conductor with neither eta/k nor reflectance, whose reference default
is the named metal-Cu spectra. As in the reference, unsupported light and shape types are skipped
with a warning. Textures other than constants raise NotImplementedError.
"""
from __future__ import annotations

import math
import warnings
from pathlib import Path

import numpy as np

from . import materials as M
from .lights import DirectionalLight
from .meshes import Y_TO_Z, apply_transform, bake, create_disk, create_sphere, load_ply, smooth_normals, translate
from .mitsuba import _load_image, _sampled_profile, rgb_to_spectrum
from .scene_builder import SceneBuilder, constant_profile

F32 = np.float32
_K_INVERT_Z = np.diag([1.0, 1.0, -1.0, 1.0])


# --- tokens and parameter lists (Parser.cpp)
def tokenize(text: str):
    """Tokens of a pbrt file: ('s', string) for quoted strings, ('[',) / (']',), ('w', word)."""
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c.isspace():
            i += 1
        elif c == "#":
            while i < n and text[i] not in "\r\n":
                i += 1
        elif c == '"':
            j = text.find('"', i + 1)
            if j < 0:
                raise ValueError("unterminated string")
            yield ("s", text[i + 1:j])
            i = j + 1
        elif c in "[]":
            yield (c,)
            i += 1
        else:
            j = i
            while j < n and not text[j].isspace() and text[j] not in '[]"#':
                j += 1
            yield ("w", text[i:j])
            i = j


_NUMERIC = {"integer", "float", "point2", "vector2", "point3", "vector3", "normal3", "normal", "point", "vector",
            "rgb", "color", "blackbody"}


class Params(dict):
    """Named parameters: name -> (type, [values])."""

    def get1(self, name, default=None):
        return self[name][1][0] if name in self else default

    def floats(self, name):
        return [float(v) for v in self[name][1]] if name in self else []

    def type_of(self, name):
        return self[name][0] if name in self else None


def _value(tok, ptype):
    if tok[0] == "s":
        return tok[1]
    w = tok[1]
    if ptype == "bool":
        return w == "true"
    if ptype == "integer":
        return int(w)
    return float(w)


class _Stream:
    def __init__(self, tokens):
        self.t = list(tokens)
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def next(self):
        tok = self.peek()
        if tok is None:
            raise ValueError("unexpected end of file")
        self.i += 1
        return tok

    def string(self):
        tok = self.next()
        if tok[0] != "s":
            raise ValueError(f"expected a quoted string, got {tok}")
        return tok[1]

    def numbers(self, count):
        out = []
        bracket = self.peek() == ("[",)
        if bracket:
            self.next()
        while len(out) < count:
            out.append(float(self.next()[1]))
        if bracket and self.next() != ("]",):
            raise ValueError("expected ']'")
        return out

    def params(self) -> Params:
        p = Params()
        while True:
            tok = self.peek()
            if tok is None or tok[0] != "s" or len(tok[1].split()) != 2:
                return p
            self.next()
            ptype, name = tok[1].split()
            vals = []
            if self.peek() == ("[",):
                self.next()
                while self.peek() != ("]",):
                    vals.append(_value(self.next(), ptype))
                self.next()
            else:
                vals.append(_value(self.next(), ptype))
            if ptype == "bool":
                vals = [v if isinstance(v, bool) else v == "true" for v in vals]
            p[name] = (ptype, vals)


# --- transforms (Builder.cpp:228-517)
def _rotate(angle_deg, axis):
    """glm::rotate: rotation by the angle about the normalised axis."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(math.radians(angle_deg)), math.sin(math.radians(angle_deg))
    x, y, z = a
    m = np.eye(4)
    m[:3, :3] = [[c + (1 - c) * x * x, (1 - c) * x * y - s * z, (1 - c) * x * z + s * y],
                 [(1 - c) * y * x + s * z, c + (1 - c) * y * y, (1 - c) * y * z - s * x],
                 [(1 - c) * z * x - s * y, (1 - c) * z * y + s * x, c + (1 - c) * z * z]]
    return m


def look_at_lh(eye, center, up):
    """glm::lookAtLH: camera-from-world with +z towards `center`, x = up x forward."""
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(up, f)
    s /= np.linalg.norm(s)
    u = np.cross(f, s)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, f
    m[:3, 3] = [-s @ eye, -u @ eye, -f @ eye]
    return m


def extract_euler_xyz(m):
    """glm::extractEulerAngleXYZ of a row-major matrix (radians)."""
    g = np.asarray(m, np.float64)[:3, :3].T  # glm's M[column][row]
    t1 = math.atan2(g[2][1], g[2][2])
    c2 = math.sqrt(g[0][0] ** 2 + g[1][0] ** 2)
    t2 = math.atan2(-g[2][0], c2)
    s1, c1 = math.sin(t1), math.cos(t1)
    t3 = math.atan2(s1 * g[0][2] - c1 * g[0][1], c1 * g[1][1] - s1 * g[1][2])
    return -t1, -t2, -t3


# --- environment map (EnvMapConverter.cs.slang, Utils/Math/MathHelpers.slang)
def _latlong_to_world(u, v):
    phi = math.pi * (2.0 * np.clip(u, 0, 1) - 1.0)
    theta = math.pi * np.clip(v, 0, 1)
    return np.stack([np.sin(theta) * np.sin(phi), np.cos(theta), -np.sin(theta) * np.cos(phi)], -1)


def _ndir_to_oct_equal_area_unorm(n):
    r = np.sqrt(1.0 - np.abs(n[..., 2]))
    phi = np.arctan2(np.abs(n[..., 1]), np.abs(n[..., 0]))
    py = r * phi * (2.0 / math.pi)
    px = r - py
    neg = n[..., 2] < 0
    px, py = np.where(neg, 1.0 - py, px), np.where(neg, 1.0 - px, py)
    return np.stack([px * np.sign(n[..., 0]), py * np.sign(n[..., 1])], -1) * 0.5 + 0.5


def oct_to_latlong(img):
    """convertEqualAreaOctToLatLong: a square equal-area octahedral map [N, N, 3] to a lat-long map
    [N, 2N, 3], each texel centre looked up bilinearly with clamped addressing."""
    img = np.asarray(img, np.float64)
    n = img.shape[0]
    if img.shape[1] != n:
        raise ValueError("the octahedral environment map must be square")
    h, w = n, 2 * n
    y, x = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    d = _latlong_to_world((x + 0.5) / w, (y + 0.5) / h)
    oct_uv = _ndir_to_oct_equal_area_unorm(d)
    s, t = oct_uv[..., 0] * n - 0.5, oct_uv[..., 1] * n - 0.5
    x0, y0 = np.floor(s), np.floor(t)
    a, b = (s - x0)[..., None], (t - y0)[..., None]
    xi0, xi1 = np.clip(x0, 0, n - 1).astype(int), np.clip(x0 + 1, 0, n - 1).astype(int)
    yi0, yi1 = np.clip(y0, 0, n - 1).astype(int), np.clip(y0 + 1, 0, n - 1).astype(int)
    top = img[yi0, xi0] * (1 - a) + img[yi0, xi1] * a
    bot = img[yi1, xi0] * (1 - a) + img[yi1, xi1] * a
    return (top * (1 - b) + bot * b).astype(np.float32)


# --- the importer
class _State:
    def __init__(self):
        self.ctm = np.eye(4)
        self.reverse = False
        self.material = None
        self.area_light = None

    def copy(self):
        s = _State()
        s.ctm, s.reverse, s.material, s.area_light = self.ctm.copy(), self.reverse, self.material, self.area_light
        return s


class Parser:
    def __init__(self, path_or_text, tables_dir: str | None = None):
        p = Path(str(path_or_text))
        if "\n" not in str(path_or_text) and p.exists():
            self.base, text = p.parent, p.read_text()
        else:
            self.base, text = Path("."), str(path_or_text)
        self.text = text
        self.builder = SceneBuilder(tables_dir)
        self.width, self.height = 1280, 720  # the rgb film's default resolution
        self.camera = ("perspective", Params(), np.eye(4))
        self.materials, self.named_materials, self.textures = [], {}, {}
        self.shapes, self.lights, self.objects, self.instances = [], [], {}, []
        self.env = None
        self._state, self._stack, self._coords, self._object = _State(), [], {}, None

    # directives (Builder.cpp)
    def _parse(self, text):
        st = _Stream(tokenize(text))
        while st.peek() is not None:
            tok = st.next()
            if tok[0] != "w":
                raise ValueError(f"expected a directive, got {tok}")
            d = tok[1]
            g = self._state
            if d in ("AttributeBegin", "TransformBegin"):
                self._stack.append((d, g.copy()))
            elif d in ("AttributeEnd", "TransformEnd"):
                if not self._stack:
                    warnings.warn(f"unmatched {d} ignored")
                    continue
                kind, saved = self._stack.pop()
                if d == "TransformEnd":
                    g.ctm = saved.ctm
                else:
                    self._state = saved
            elif d == "Identity":
                g.ctm = np.eye(4)
            elif d == "Translate":
                g.ctm = g.ctm @ translate(st.numbers(3))
            elif d == "Scale":
                g.ctm = g.ctm @ np.diag(st.numbers(3) + [1.0])
            elif d == "Rotate":
                a, x, y, z = st.numbers(4)
                g.ctm = g.ctm @ _rotate(a, (x, y, z))
            elif d == "LookAt":
                v = st.numbers(9)
                g.ctm = g.ctm @ look_at_lh(v[0:3], v[3:6], v[6:9])
            elif d == "Transform":
                g.ctm = np.asarray(st.numbers(16)).reshape(4, 4).T
            elif d == "ConcatTransform":
                g.ctm = g.ctm @ np.asarray(st.numbers(16)).reshape(4, 4).T
            elif d == "CoordinateSystem":
                self._coords[st.string()] = g.ctm.copy()
            elif d == "CoordSysTransform":
                name = st.string()
                if name in self._coords:
                    g.ctm = self._coords[name].copy()
                else:
                    warnings.warn(f"unknown coordinate system '{name}'")
            elif d == "ReverseOrientation":
                g.reverse = not g.reverse
            elif d == "Camera":
                name, params = st.string(), st.params()
                world_from_camera = np.linalg.inv(g.ctm)
                self._coords["camera"] = world_from_camera
                self.camera = (name, params, world_from_camera)
            elif d == "WorldBegin":
                g.ctm = np.eye(4)
                self._coords["world"] = g.ctm.copy()
            elif d == "Film":
                st.string()
                p = st.params()
                self.width, self.height = int(p.get1("xresolution", 1280)), int(p.get1("yresolution", 720))
            elif d in ("Sampler", "Integrator", "PixelFilter", "Accelerator", "MakeNamedMedium", "Option"):
                st.string()
                st.params()
            elif d == "ColorSpace":
                st.string()
            elif d == "WorldEnd":
                pass
            elif d == "MediumInterface":
                st.string()
                if st.peek() is not None and st.peek()[0] == "s":
                    st.string()
            elif d == "Attribute":
                st.string()
                st.params()
            elif d == "ActiveTransform":
                st.next()
            elif d == "TransformTimes":
                st.numbers(2)
            elif d == "Texture":
                name, ttype, cls, params = st.string(), st.string(), st.string(), st.params()
                self.textures[name] = (ttype, cls, params)
            elif d == "Material":
                self.materials.append((st.string(), st.params()))
                g.material = ("index", len(self.materials) - 1)
            elif d == "MakeNamedMaterial":
                name, params = st.string(), st.params()
                self.named_materials[name] = (params.get1("type", ""), params)
            elif d == "NamedMaterial":
                g.material = ("name", st.string())
            elif d == "AreaLightSource":
                g.area_light = (st.string(), st.params())
            elif d == "LightSource":
                self.lights.append((st.string(), st.params(), g.ctm.copy()))
            elif d == "Shape":
                shape = (st.string(), st.params(), g.ctm.copy(), g.reverse, g.material,
                         None if self._object else g.area_light)
                (self.objects[self._object] if self._object else self.shapes).append(shape)
            elif d == "ObjectBegin":
                self._stack.append(("AttributeBegin", g.copy()))
                self._object = st.string()
                self.objects[self._object] = []
            elif d == "ObjectEnd":
                self._object = None
                if self._stack:
                    self._state = self._stack.pop()[1]
            elif d == "ObjectInstance":
                self.instances.append((st.string(), g.ctm.copy()))
            elif d in ("Include", "Import"):
                self._parse((self.base / st.string()).read_text())
            else:
                raise ValueError(f"unknown directive '{d}'")

    # spectra and textures (PBRTImporter.cpp:275-405)
    def _rgb(self, p: Params, name, default):
        if name not in p:
            return tuple(float(x) for x in default)
        t, v = p[name]
        if t in ("rgb", "color"):
            return tuple(float(x) for x in v[:3])
        if t == "texture":
            tex = self.textures.get(v[0])
            if tex and tex[1] == "constant":
                return self._rgb(tex[2], "value", default)
            raise NotImplementedError(f"texture '{v[0]}' for '{name}' (only constant textures)")
        raise NotImplementedError(f"'{t}' spectrum for '{name}' (give an rgb value)")

    def _float(self, p: Params, name, default):
        if name not in p:
            return None if default is None else float(default)
        t, v = p[name]
        if t == "float":
            return float(v[0])
        if t == "texture":
            tex = self.textures.get(v[0])
            if tex and tex[1] == "constant":
                return self._float(tex[2], "value", default)
            return None  # non-constant: the caller's fallback
        raise ValueError(f"'{name}' must be a float")

    def roughness(self, p: Params, base="roughness", u="uroughness", v="vroughness"):
        """getRoughness (PBRTImporter.cpp:417-452): both components are u + v, square-rooted when
        `remaproughness` (default true); 0.5 when a roughness is not a constant."""
        ru = self._float(p, u, None)
        rv = self._float(p, v, None)
        rb = self._float(p, base, 0.0)
        ru = rb if u not in p else ru
        rv = rb if v not in p else rv
        if ru is None or rv is None:
            warnings.warn("non-constant roughness: using 0.5")
            return 0.5
        r = F32(ru) + F32(rv)
        if p.get1("remaproughness", True):
            r = np.sqrt(r)
        return float(r)

    def scalar_eta(self, p: Params, name="eta"):
        """getScalarEta (PBRTImporter.cpp:473-486)."""
        if name not in p:
            return 1.5
        if p.type_of(name) == "float":
            return float(p.get1(name))
        r, g, b = self._rgb(p, name, (1.5, 1.5, 1.5))
        return float((F32(r) + F32(g) + F32(b)) / F32(3.0))

    def conductor_eta(self, p: Params, refl="reflectance", eta="eta", k="k"):
        """getConductorEtaK (PBRTImporter.cpp:509-565), the eta half (the PLT setters use only it)."""
        if refl in p and (eta in p or k in p):
            raise ValueError(f"Both '{refl}' and '{eta}' and '{k}' can't be provided.")
        if refl in p:
            return (1.0, 1.0, 1.0)
        if eta not in p:
            raise NotImplementedError("conductor without eta (its default is the named metal-Cu spectrum)")
        return self._rgb(p, eta, (1.0, 1.0, 1.0))

    # materials (createMaterial with PLT materials, PBRTImporter.cpp:969-1256)
    def material(self, mtype, p: Params):
        h = M.h
        if mtype in ("", "none", "interface", "mix"):
            if mtype == "mix":
                warnings.warn("material type 'mix' is not supported")
            return None
        if mtype == "diffuse":
            m = M.blank(M.DIFFUSE)
            for i, c in enumerate(self._rgb(p, "reflectance", (0.5, 0.5, 0.5))):
                m.base_color[i] = h(c)
            return m
        if mtype == "coateddiffuse":
            m = M.coated_opaque(1.0, 1.0, 1.0, 0.0, 1.5, 3.0)  # the constructor's defaults
            m.specular[1] = h(self.roughness(p))
            for i, c in enumerate(self._rgb(p, "reflectance", (0.5, 0.5, 0.5))):
                m.base_color[i] = h(c)
            return m
        if mtype == "conductor":
            eta = self.conductor_eta(p)
            rough = self.roughness(p)
            m = M.conductor(0.0, M.NO_PROFILE, M.NO_PROFILE)
            for i, c in enumerate(eta):
                m.base_color[i] = h(c)
            m.base_color[3] = h(1.0)
            m.specular[1] = h(rough)
            return m
        if mtype == "coatedconductor":
            self.roughness(p, "interface.roughness", "interface.uroughness", "interface.vroughness")
            coat = self.scalar_eta(p, "interface.eta")
            eta = self.conductor_eta(p, "conductor.reflectance", "conductor.eta", "conductor.k")
            self.roughness(p, "conductor.roughness", "conductor.uroughness", "conductor.vroughness")
            m = M.coated_conductor(0.0, 1.5, 3.0, M.NO_PROFILE, M.NO_PROFILE)  # the constructor's defaults
            for i, c in enumerate(eta):
                m.base_color[i] = h(c)
            m.base_color[3] = h(1.0)
            m.data1[3] = h(coat)
            return m
        if mtype == "dielectric":
            return M.dielectric(self.scalar_eta(p), self.roughness(p))
        if mtype == "thindielectric":
            m = M.thin_dielectric(1.5, 0.0003)
            m.specular[0] = h(0.0)  # the constructor's specular params are all 0
            m.ior = h(self.scalar_eta(p))
            return m
        if mtype == "diffusetransmission":
            m = M.dielectric(1.5, 0.6)
            for i, c in enumerate(self._rgb(p, "reflectance", (0.25, 0.25, 0.25))):
                m.base_color[i] = h(c)
            for i, c in enumerate(self._rgb(p, "transmittance", (0.25, 0.25, 0.25))):
                m.transmission[i] = h(c)
            return m
        if mtype in ("hair", "measured", "subsurface"):
            warnings.warn(f"material type '{mtype}' is not supported")
            return None
        raise ValueError(f"Unknown material type '{mtype}'.")

    # shapes (createShape, PBRTImporter.cpp:1322-1576)
    def shape_mesh(self, stype, p: Params, ctm):
        if stype == "sphere":
            pos, nrm, idx, cw = create_sphere(float(p.get1("radius", 1.0)))
            m = ctm
        elif stype == "disk":
            pos, nrm, idx, cw = create_disk(float(p.get1("radius", 1.0)))
            pos, nrm, cw = apply_transform(translate([0.0, 0.0, float(p.get1("height", 0.0))]) @ Y_TO_Z, pos, nrm, cw)
            m = np.eye(4)  # the reference does not apply the shape's transform to disks
        elif stype == "trianglemesh":
            idx = np.asarray(p[("indices")][1] if "indices" in p else [], np.int64)
            P = np.asarray(p.floats("P"), np.float64).reshape(-1, 3)
            if idx.size == 0:
                if len(P) != 3:
                    warnings.warn("Vertex indices 'indices' missing. Skipping.")
                    return None
                idx = np.arange(3)
            idx = idx[:len(idx) // 3 * 3].reshape(-1, 3)
            if len(P) == 0 or idx.min() < 0 or idx.max() >= len(P):
                warnings.warn("trianglemesh without positions or with indices out of range: skipped")
                return None
            N = np.asarray(p.floats("N"), np.float64).reshape(-1, 3)
            nrm = N if len(N) == len(P) else smooth_normals(P, idx)
            pos, cw, m = P, False, ctm
        elif stype == "plymesh":
            pos, nrm, idx = load_ply(self.base / p.get1("filename"))
            if nrm is None:
                nrm = smooth_normals(pos, idx)
            cw, m = False, ctm
        else:
            warnings.warn(f"shape type '{stype}' is not supported")
            return None
        pos, nrm, cw = apply_transform(m, pos, nrm, cw)
        return pos, nrm, idx, cw

    def _material_id(self, ref, cache):
        key = ref
        if key in cache:
            return cache[key]
        mat = None
        if ref is not None:
            kind, v = ref
            mtype, params = self.materials[v] if kind == "index" else self.named_materials[v]
            mat = self.material(mtype, params)
        if mat is None:
            if "default" not in cache:
                cache["default"] = self.builder.addMaterial(M.blank(M.DIFFUSE))
            cache[key] = cache["default"]
        else:
            cache[key] = self.builder.addMaterial(mat)
        return cache[key]

    def _emissive_material(self, area):
        atype, ap = area
        if atype != "diffuse":
            raise ValueError(f"Unknown area light type '{atype}'.")
        L = np.asarray(self._rgb(ap, "L", (1.0, 1.0, 1.0)), np.float64) * float(ap.get1("scale", 1.0))
        self.builder.profiles.append(_sampled_profile(rgb_to_spectrum(L), True))
        m = M.diffuse(0.0, 0.0, 0.0, len(self.builder.profiles) - 1)
        return self.builder.addMaterial(m)

    def _add_shape(self, shape, cache, instance=np.eye(4)):
        stype, p, ctm, reverse, mref, area = shape
        r = self.shape_mesh(stype, p, instance @ ctm)
        if r is None:
            return
        pos, nrm, idx, cw = r
        mid = self._emissive_material(area) if area is not None else self._material_id(mref, cache)
        self.builder.addTriangleMesh(*bake(pos, nrm, idx, cw ^ reverse), mid)

    # camera and lights (PBRTImporter.cpp:581-744, buildScene 1681-1793)
    def _camera(self):
        ctype, p, world_from_camera = self.camera
        if ctype != "perspective":
            warnings.warn(f"camera type '{ctype}' is not supported")
            return
        fov = F32(math.radians(float(p.get1("fov", 90.0))))
        focal = F32(24.0) / (F32(2.0) * F32(math.tan(F32(0.5) * fov)))  # fovYToFocalLength(fov, 24)
        node = world_from_camera @ _K_INVERT_Z
        pos, fwd, up = node[:3, 3], -node[:3, 2], node[:3, 1]  # Camera::updateFromAnimation
        self.builder.addCamera(tuple(pos), tuple(pos + fwd), tuple(up), focalLength=float(focal), frameHeight=24.0,
                               aspect=self.width / self.height, focalDistance=float(p.get1("focaldistance", 1e6)))

    def _light(self, ltype, p: Params, ctm):
        if ltype == "distant":
            frm = np.asarray(p.floats("from") or [0.0, 0.0, 0.0])
            to = np.asarray(p.floats("to") or [0.0, 0.0, 1.0])
            d = ctm[:3, :3] @ (to - frm)
            light = DirectionalLight()
            light.direction = d / np.linalg.norm(d)
            if not self.builder.profiles:
                self.builder.profiles.append(constant_profile(0.0))  # a valid id for the unset intensity
            self.builder.addLight(light)
        elif ltype == "infinite":
            has_l, fname = "L" in p, p.get1("filename", "")
            if has_l and fname:
                raise ValueError("Can't specify both emission 'L' and 'filename' for infinite light.")
            if self.env is not None:
                warnings.warn("No support for multiple infinite light. Discarding this light.")
                return
            if has_l:
                self.env = dict(image=np.asarray(self._rgb(p, "L", (1, 1, 1)), np.float32).reshape(1, 1, 3),
                                rotation=(0.0, 0.0, 0.0), scale=1.0)
            elif fname:
                img = oct_to_latlong(_load_image(self.base / fname))
                rot = tuple(math.degrees(a) for a in extract_euler_xyz(ctm))
                self.env = dict(image=img, rotation=rot, scale=float(p.get1("scale", 1.0)))
        elif ltype in ("point", "spot", "goniometric", "projection"):
            warnings.warn(f"light type '{ltype}' is not supported")
        else:
            raise ValueError(f"Unknown light type '{ltype}'.")

    def run(self):
        self._parse(self.text)
        self._camera()
        for ltype, p, ctm in self.lights:
            self._light(ltype, p, ctm)
        cache = {}
        for shape in self.shapes:
            self._add_shape(shape, cache)
        for name, inst in self.instances:
            if name not in self.objects:
                raise ValueError(f"Object instance '{name}' not defined.")
            for shape in self.objects[name]:
                self._add_shape(shape, cache, inst)
        return self


def load_scene(path_or_text, ctx, tables_dir: str | None = None):
    """Reads a pbrt-v4 scene (file path or text) into a Scene on `ctx` (host-only when None).
    Returns (scene, width, height)."""
    prs = Parser(path_or_text, tables_dir).run()
    scene = prs.builder.getScene(ctx)
    if prs.env is not None:
        scene.set_env_map(prs.env["image"], rotation=prs.env["rotation"], intensity=prs.env["scale"])
    return scene, prs.width, prs.height


__all__ = ["load_scene", "Parser", "tokenize", "look_at_lh", "extract_euler_xyz", "oct_to_latlong", "load_ply"]
