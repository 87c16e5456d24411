"""Scene ingest for the PLT path: the parts of the reference's SceneBuilder that a scene script or
importer uses to hand geometry, PLT materials, spectral profiles and a camera to the renderer
(Scene/SceneBuilder.{h,cpp}), and StandardMaterialPLTWrapper, the conversion of a standard
material description into a PLT material (Scene/Material/StandardMaterial.{h,cpp}:130-158,
215-276) that the reference's importers use.

Spectral tables are read from a directory laid out as the reference's `Source/Tables`
(`ior/<name>.eta.spd`, `ior/<name>.k.spd`, `emission/<name>.spd`).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import materials as M
from .scene import Scene


def focal_length_to_fov_y(focal_length: float, frame_height: float = 24.0) -> float:
    """Utils/Math/FalcorMath.h:146-149 (millimetres, 35 mm film by default)."""
    return float(2.0 * np.arctan(np.float32(0.5) * np.float32(frame_height) / np.float32(focal_length)))


@dataclass
class StandardMaterialPLTWrapper:
    """StandardMaterialPLTWrapper (StandardMaterial.h:130-158): a standard material description that
    genMaterial turns into a PLT conductor (metalName set), diffuse (isDiffuse), opaque dielectric
    (specularTransmission < 0.01) or dielectric. Textures are not part of this build."""
    name: str = ""
    indexOfRefraction: float = 1.5
    specularTransmission: float = 0.0
    baseColor: tuple = (1.0, 1.0, 1.0, 1.0)
    specularParams: tuple = (0.0, 0.0, 0.0, 0.0)  # unused, roughness, metallic, unused
    volumeScattering: tuple = (0.0, 0.0, 0.0)
    volumeAbsorption: tuple = (0.0, 0.0, 0.0)
    transmissionColor: tuple = (1.0, 1.0, 1.0)
    alphaThreshold: float = 0.5
    doubleSided: bool = False
    isDiffuse: bool = False
    metalName: str = ""

    def genMaterial(self, builder: "SceneBuilder") -> _lib.Material:
        """StandardMaterial.cpp:219-276, setter by setter (binary16 fields as the reference stores them)."""
        h = M.h
        sp = [float(v) for v in self.specularParams]
        if self.metalName:
            m = M.blank(M.CONDUCTOR)  # PLTConductorMaterial defaults, then the wrapper's setters
            for i in range(4):
                m.base_color[i] = h(self.baseColor[i])
                m.specular[i] = h(sp[i])
            n_id, k_id = builder.addSpectralProfileFromMaterial(self.metalName)
            m.ior_n_profile, m.ior_k_profile = n_id, k_id
            m.ior = h(1.0)             # setExtIndexOfRefraction(1)
            m.specular[0] = h(3.0)     # setGamma(3)
            return m
        if self.isDiffuse:
            m = M.blank(M.DIFFUSE)
            for i in range(4):
                m.base_color[i] = h(self.baseColor[i])
            return m
        ior = max(1.1, float(self.indexOfRefraction))
        if self.specularTransmission < .01:
            m = M.blank(M.OPAQUE_DIELECTRIC)
            for i in range(4):
                m.specular[i] = h(sp[i])
            m.ior = h(ior)
            return m
        m = M.blank(M.DIELECTRIC)
        for i in range(4):
            m.specular[i] = h(sp[i])
        t = np.float32(self.specularTransmission)
        for i in range(3):
            m.transmission[i] = h(float(t * np.float32(self.transmissionColor[i])))
        m.ior = h(ior)
        ior_h = np.float32(np.array(m.ior, np.uint16).view(np.float16))
        m.specular[0] = h(float(np.float32(.52) * (ior_h - np.float32(1.0)) / np.float32(80.0)))  # setAbbeNumber(80)
        m.specular[3] = h(1.0)         # setExtIndexOfRefraction(1)
        return m


@dataclass
class _Mesh:
    positions: np.ndarray
    normals: np.ndarray
    indices: np.ndarray
    material: int


class SceneBuilder:
    """Collects triangle meshes, PLT materials, spectral profiles and a camera, and builds the scene
    on a device (SceneBuilder.h: addSpectralProfile / addSpectralProfileFromMaterial /
    addSpectralProfileForEmitterType / addMaterial / addTriangleMesh / addCamera)."""

    def __init__(self, tables_dir: str | None = None):
        self.lib = _lib.load()
        self.tables_dir = tables_dir
        self.profiles: list[_lib.SpectralProfile] = []
        self.materials: list[_lib.Material] = []
        self.meshes: list[_Mesh] = []
        self.lights = []
        self.camera = _lib.Camera((0, 0, 1), (0, 0, 0), (0, 1, 0), math.radians(45.0), 1.0, 1.0, (0, 0))

    # --- spectral profiles (SceneBuilder.cpp:911-1000)
    def _push(self, p: _lib.SpectralProfile) -> int:
        self.profiles.append(p)
        return len(self.profiles) - 1

    def addSpectralProfile(self, wavelengths, values, computeICdf: bool = True) -> int:
        from .spectra import build_profile
        return self._push(build_profile(wavelengths, values, computeICdf))

    def addSpectralProfileFromMaterial(self, name: str) -> tuple[int, int]:
        n, k = _lib.SpectralProfile(), _lib.SpectralProfile()
        d = self.tables_dir.encode() if self.tables_dir else None
        _lib.check(self.lib.plt_spectral_profiles_from_material(d, name.encode(), C.byref(n), C.byref(k)))
        return self._push(n), self._push(k)

    def addSpectralProfileForEmitterType(self, name: str, scale: float = 1.0) -> int:
        p = _lib.SpectralProfile()
        d = self.tables_dir.encode() if self.tables_dir else None
        _lib.check(self.lib.plt_spectral_profile_from_emitter(d, name.encode(), float(scale), C.byref(p)))
        return self._push(p)

    # --- materials, geometry, camera
    def addMaterial(self, material) -> int:
        m = material.genMaterial(self) if isinstance(material, StandardMaterialPLTWrapper) else material
        self.materials.append(m)
        return len(self.materials) - 1

    def addTriangleMesh(self, positions, normals, indices, materialID: int) -> int:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        if nrm.shape != pos.shape:
            raise ValueError("one normal per vertex")
        if idx.size and idx.max() >= len(pos):
            raise ValueError("index out of range")
        if not 0 <= materialID < len(self.materials):
            raise ValueError("material id out of range")
        self.meshes.append(_Mesh(pos, nrm, idx, int(materialID)))
        return len(self.meshes) - 1

    def addLight(self, light) -> int:
        """An analytic light (pltfalcor_amd.lights): Scene::addLight."""
        self.lights.append(light)
        return len(self.lights) - 1

    def addCamera(self, position, target, up, fovY: float | None = None, focalLength: float | None = None,
                  frameHeight: float = 24.0, aspect: float = 1.0, focalDistance: float = 1.0,
                  polarizationIntensity: float = 0.0, polarizationAngle: float = 0.0):
        """Pinhole camera: vertical field of view in radians, or a focal length in mm
        (Camera::setFocalLength, with the 24 mm frame height). The polarizer is the camera's
        polarizationIntensity / polarizationAngle properties (Camera.cpp:315-323, 452-453)."""
        fov = fovY if fovY is not None else focal_length_to_fov_y(50.0 if focalLength is None else focalLength,
                                                                 frameHeight)
        self.camera = _lib.Camera(tuple(position), tuple(target), tuple(up), fov, aspect, focalDistance, (0, 0),
                                  float(polarizationIntensity), float(polarizationAngle))

    def getScene(self, ctx) -> Scene:
        """Uploads the collected scene (host-only when ctx is None)."""
        if not self.meshes:
            raise ValueError("no meshes")
        pos, nrm, idx, tri_mesh, off = [], [], [], [], 0
        for i, m in enumerate(self.meshes):
            pos.append(m.positions)
            nrm.append(m.normals)
            idx.append(m.indices + off)
            tri_mesh.append(np.full(len(m.indices), i, np.uint32))
            off += len(m.positions)
        mesh_material = np.array([m.material for m in self.meshes], np.uint32)
        mats = M.pack(self.materials)
        profs = b"".join(bytes(p) for p in self.profiles)
        scene = Scene.from_arrays(ctx, np.concatenate(pos), np.concatenate(nrm), np.concatenate(idx),
                                  np.concatenate(tri_mesh), mesh_material, mats, profs, self.camera)
        if self.lights:
            scene.set_lights(self.lights)
        return scene


def load_spd(path: str):
    """(wavelengths, values) of a `.spd` table, as float32 arrays (PiecewiseLinearSpectrum::fromFile)."""
    lib = _lib.load()
    n = C.c_uint32()
    _lib.check(lib.plt_spectrum_load_spd(str(path).encode(), 0, None, None, C.byref(n)))
    wl = np.zeros(n.value, np.float32)
    v = np.zeros(n.value, np.float32)
    _lib.check(lib.plt_spectrum_load_spd(str(path).encode(), n.value, wl.ctypes.data_as(C.POINTER(C.c_float)),
                                         v.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)))
    return wl, v


def constant_profile(value: float) -> _lib.SpectralProfile:
    p = _lib.SpectralProfile()
    _lib.check(_lib.load().plt_spectral_profile_constant(float(value), C.byref(p)))
    return p


__all__ = ["SceneBuilder", "StandardMaterialPLTWrapper", "load_spd", "constant_profile", "focal_length_to_fov_y"]
