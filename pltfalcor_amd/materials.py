"""PLT material parameter blocks (`plt_material`, the reference's binary16 BasicMaterialData
layout, Scene/Material/BasicMaterialData.slang:63-107), built the way each reference material's
constructor and setters fill them (Rendering/Materials/PLT/*Material.cpp). The field use per type
is documented in include/plt_abi.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Material

NO_PROFILE = 0xFFFFFFFF
# MaterialType values (Scene/Material/MaterialTypes.slang:36-53)
DIFFUSE, CONDUCTOR, DIELECTRIC, OPAQUE_DIELECTRIC, THIN_DIELECTRIC, COATED_CONDUCTOR, COATED_OPAQUE_DIELECTRIC, \
    GRATED_CONDUCTOR, GRATED_OPAQUE_DIELECTRIC, MULTILAYER_STACK = range(2, 12)


def h(x: float) -> int:
    """float -> binary16 bits (round to nearest even)."""
    return int(np.array(x, np.float16).view(np.uint16))


def blank(type_: int) -> Material:
    """BasicMaterialData defaults: base colour 1, transmission 1, IoR 1.5, no profiles."""
    m = Material()
    m.type = type_
    for i in range(4):
        m.base_color[i] = h(1.0)
    for i in range(3):
        m.transmission[i] = h(1.0)
    m.ior = h(1.5)
    m.emission_profile = m.ior_n_profile = m.ior_k_profile = NO_PROFILE
    return m


def dielectric(ior: float, roughness: float) -> Material:
    """PLTDielectricMaterial.cpp:46-101: ext IoR 1, Abbe-100 dispersion at the default IoR 1.5."""
    m = blank(DIELECTRIC)
    m.specular[3] = h(1.0)
    m.specular[0] = h(.52 * (float(np.float16(1.5)) - 1.0) / 100.0)
    m.ior = h(ior)
    m.specular[1] = h(roughness)
    return m


def conductor(roughness: float, n_profile: int, k_profile: int, gamma: float = 3.0) -> Material:
    """PLTConductorMaterial.cpp:45-85."""
    m = blank(CONDUCTOR)
    m.ior = h(1.0)
    for i in range(4):
        m.specular[i] = h(1.0)
    m.specular[1] = h(roughness)
    m.specular[0] = h(gamma)
    m.ior_n_profile, m.ior_k_profile = n_profile, k_profile
    return m


def diffuse(r: float, g: float, b: float, emission: int = NO_PROFILE) -> Material:
    m = blank(DIFFUSE)
    m.base_color[0], m.base_color[1], m.base_color[2] = h(r), h(g), h(b)
    m.emission_profile = emission
    return m


def opaque_dielectric(r: float, g: float, b: float, roughness: float) -> Material:
    m = blank(OPAQUE_DIELECTRIC)
    m.base_color[0], m.base_color[1], m.base_color[2] = h(r), h(g), h(b)
    m.ior = h(1.0)
    m.specular[1] = h(roughness)
    return m


def coated_opaque(r: float, g: float, b: float, roughness: float, coat_ior: float, thickness_um: float) -> Material:
    """PLTCoatedOpaqueDielectricMaterial.cpp:45-105."""
    m = blank(COATED_OPAQUE_DIELECTRIC)
    m.base_color[0], m.base_color[1], m.base_color[2] = h(r), h(g), h(b)
    m.ior = h(1.0)
    m.specular[1] = h(roughness)
    m.data1[3] = h(coat_ior)
    m.data1[0] = h(thickness_um)
    m.specular[2] = h(1.0)
    return m


def coated_conductor(roughness: float, coat_ior: float, thickness_um: float, n_profile: int, k_profile: int,
                     ext_ior: float = 1.0) -> Material:
    """PLTCoatedConductorMaterial.cpp:45-110."""
    m = blank(COATED_CONDUCTOR)
    m.ior = h(ext_ior)
    m.specular[1] = h(roughness)
    m.specular[2] = h(1.0)
    m.data1[3] = h(coat_ior)
    m.data1[0] = h(thickness_um)
    m.ior_n_profile, m.ior_k_profile = n_profile, k_profile
    return m


def grated_conductor(pitch_um: float, direction: float, n_profile: int, k_profile: int, lobes: int = 3,
                     grating_type: int = 0, height: float = 1.0, height_scale: float = .5, roughness: float = 0.0
                     ) -> Material:
    """PLTDiffractionGratedConductorMaterial.cpp:45-147 (1-D grating along `direction`)."""
    m = blank(GRATED_CONDUCTOR)
    m.ior = h(1.0)
    m.data1[0], m.data1[1], m.data1[2] = h(pitch_um), h(0.0), h(direction)
    m.data2[0], m.data2[1] = h(height), h(roughness)
    m.specular[0], m.specular[1] = h(height_scale), h(1.0)
    m.data1[3] = (grating_type << 8) | lobes
    m.ior_n_profile, m.ior_k_profile = n_profile, k_profile
    return m


def grated_opaque_dielectric(pitch_um: float, direction: float, albedo=(1.0, 1.0, 1.0), lobes: int = 3,
                             grating_type: int = 0, height: float = 1.0, height_scale: float = .5,
                             roughness: float = 0.0) -> Material:
    """PLTDiffractionGratedOpaqueDielectricMaterial.cpp:44-147 (1-D grating along `direction` over an
    opaque dielectric base of colour `albedo`; no exterior IoR or n, k profiles)."""
    m = blank(GRATED_OPAQUE_DIELECTRIC)
    m.base_color[0], m.base_color[1], m.base_color[2] = (h(c) for c in albedo)
    m.data1[0], m.data1[1], m.data1[2] = h(pitch_um), h(0.0), h(direction)
    m.data2[0], m.data2[1] = h(height), h(roughness)
    m.specular[0], m.specular[1] = h(height_scale), h(1.0)
    m.data1[3] = (grating_type << 8) | lobes
    return m


def thin_dielectric(ior: float, thickness_mm: float, birefringence: float = 0.0, axis=(0.0, 0.0, 1.0)) -> Material:
    """PLTThinDielectricMaterial.cpp:45-90; the birefringence scale (specular.x) is set to 1."""
    m = blank(THIN_DIELECTRIC)
    m.ior = h(ior)
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    if a[2] < 0:
        a = -a
    m.base_color[0], m.base_color[1], m.base_color[2], m.base_color[3] = h(birefringence), h(a[0]), h(a[1]), h(1.0)
    for i in range(4):
        m.specular[i] = h(0.0)
    m.specular[0] = h(1.0)
    m.specular[3] = h(thickness_mm)
    return m


def multilayer(ior1: float, ior2: float, thickness1_um: float, thickness2_um: float, layers: int,
               albedo=(1.0, 1.0, 1.0), ext_ior: float = 1.0) -> Material:
    """PLTMultiLayeredStackMaterial.cpp:45-115 (layer count stored as raw uint16 bits)."""
    m = blank(MULTILAYER_STACK)
    m.base_color[0], m.base_color[1], m.base_color[2] = (h(c) for c in albedo)
    m.ior = h(ext_ior)
    m.data2[0], m.data2[1] = h(ior1), h(ior2)
    m.data1[0], m.data1[1] = h(thickness1_um), h(thickness2_um)
    m.data1[3] = int(layers) & 0xFFFF
    return m


def pack(materials) -> bytes:
    """Contiguous plt_material array bytes."""
    arr = (Material * len(materials))(*materials)
    return C.string_at(C.addressof(arr), C.sizeof(arr))


__all__ = ["blank", "dielectric", "conductor", "diffuse", "opaque_dielectric", "coated_opaque", "coated_conductor",
           "grated_opaque_dielectric",
           "grated_conductor", "thin_dielectric", "multilayer", "pack", "NO_PROFILE"]
