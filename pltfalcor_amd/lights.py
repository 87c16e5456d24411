"""Analytic light sources: host mirrors of the reference's light classes (Scene/Lights/Light.{h,cpp})
that fill the LightData record (plt_light) the path tracer samples.

Each class keeps the reference's defaults and setter semantics (normalised directions, clamped spot
angles, the PLT beam parameters `setLightArea` / `setEmissionSolidAngle` / `setSourceSolidAngle`,
the area lights' transform, scaling and surface area), computed in float32 as the reference's
float members are. Intensities are spectral profiles of the scene (`setIntensity(profile id)`).
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import Light as _Light

POINT, DIRECTIONAL, DISTANT, RECT, DISC, SPHERE = range(6)
F32 = np.float32


def _normalize(v):
    v = np.asarray(v, F32)
    return (v / F32(np.sqrt(np.dot(v, v)))).astype(F32)


def _cos_from_solid_angle(omega: float) -> float:
    """sqrt(1 - Omega / 4 / pi), as `sqrt(1.f-Omega/4.f/M_PI)` evaluates (double, stored as float)."""
    return float(F32(math.sqrt(1.0 - float(F32(omega) / F32(4.0)) / math.pi)))


class AnalyticLight:
    """LightData defaults (Scene/Lights/LightData.slang:47-66)."""

    def __init__(self, type_: int, name: str = ""):
        self.name = name
        self.data = _Light()
        d = self.data
        d.type = type_
        d.dir[:] = (0.0, -1.0, 0.0)
        d.opening_angle = F32(math.pi)
        d.cos_opening_angle = -1.0
        d.cos_subtended_angle = 0.9999893
        d.penumbra_angle = 0.0
        d.intensity_profile = 0
        d.surface_area = 0.0

    def setIntensity(self, profile_id: int):
        """Light::setIntensity (Light.cpp:51-55): the spectral profile of the emitted intensity."""
        self.data.intensity_profile = int(profile_id)

    def _set_matrix(self, field: str, m):
        getattr(self.data, field)[:] = [float(x) for x in np.asarray(m, F32).reshape(16)]


class PointLight(AnalyticLight):
    """PointLight (Light.cpp:117-201): a (spot) point light; PLT beams from it have area 10000 and
    solid angle 1e-6 unless set."""

    def __init__(self, name: str = ""):
        super().__init__(POINT, name)
        self.setLightArea(10000.0)
        self.setEmissionSolidAngle(1e-6)

    @property
    def position(self):
        return tuple(self.data.pos)

    @position.setter
    def position(self, p):
        self.data.pos[:] = [float(x) for x in np.asarray(p, F32)]

    @property
    def direction(self):
        return tuple(self.data.dir)

    @direction.setter
    def direction(self, d):
        if not np.linalg.norm(np.asarray(d, np.float64)) > 0:
            return  # the reference ignores a zero-length direction
        self.data.dir[:] = [float(x) for x in _normalize(d)]

    @property
    def openingAngle(self):
        return self.data.opening_angle

    @openingAngle.setter
    def openingAngle(self, a):
        a = float(np.clip(F32(a), 0.0, F32(math.pi)))
        if a == self.data.opening_angle:
            return
        self.data.opening_angle = a
        self.data.penumbra_angle = min(self.data.penumbra_angle, a)
        self.data.cos_opening_angle = float(F32(math.cos(a)))

    @property
    def penumbraAngle(self):
        return self.data.penumbra_angle

    @penumbraAngle.setter
    def penumbraAngle(self, a):
        self.data.penumbra_angle = float(np.clip(F32(a), 0.0, F32(self.data.opening_angle)))

    def setLightArea(self, area: float):
        self.data.surface_area = float(F32(area))

    def setEmissionSolidAngle(self, omega: float):
        self.data.cos_subtended_angle = _cos_from_solid_angle(omega)


class DirectionalLight(AnalyticLight):
    """DirectionalLight (Light.cpp:203-250): parallel light along `direction` (a Dirac in direction)."""

    def __init__(self, name: str = ""):
        super().__init__(DIRECTIONAL, name)

    @property
    def direction(self):
        return tuple(self.data.dir)

    @direction.setter
    def direction(self, d):
        if not np.linalg.norm(np.asarray(d, np.float64)) > 0:
            return
        self.data.dir[:] = [float(x) for x in _normalize(d)]

    def setSourceSolidAngle(self, omega: float):
        self.data.cos_subtended_angle = _cos_from_solid_angle(omega)


def _rotation(angle: float, axis) -> np.ndarray:
    """glm::rotate(mat4(), angle, axis): rotation by `angle` about the normalised axis."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(angle), math.sin(angle)
    x, y, z = a
    r = np.array([[c + (1 - c) * x * x, (1 - c) * x * y - s * z, (1 - c) * x * z + s * y],
                  [(1 - c) * y * x + s * z, c + (1 - c) * y * y, (1 - c) * y * z - s * x],
                  [(1 - c) * z * x - s * y, (1 - c) * z * y + s * x, c + (1 - c) * z * z]])
    m = np.eye(4)
    m[:3, :3] = r
    return m.astype(F32)


class DistantLight(AnalyticLight):
    """DistantLight (Light.cpp:252-305): light from a cone of directions around -`direction`
    subtending `setSourceSolidAngle` (default: the sun seen from the earth); its matrix rotates +z
    onto the direction towards the light."""

    def __init__(self, name: str = ""):
        super().__init__(DISTANT, name)
        self.data.dir[:] = (0.0, -1.0, 0.0)
        self._update()

    @property
    def direction(self):
        return tuple(self.data.dir)

    @direction.setter
    def direction(self, d):
        if not np.linalg.norm(np.asarray(d, np.float64)) > 0:
            return
        self.data.dir[:] = [float(x) for x in _normalize(d)]
        self._update()

    def setSourceSolidAngle(self, omega: float):
        self.data.cos_subtended_angle = _cos_from_solid_angle(omega)
        self._update()

    def _update(self):
        up = np.array([0.0, 0.0, 1.0], F32)
        nd = -np.asarray(self.data.dir, F32)
        vec = np.cross(up, nd)
        sin_t = float(np.linalg.norm(vec))
        if sin_t > 0.0:
            m = _rotation(math.acos(float(np.clip(np.dot(up, nd), -1, 1))), vec)
        else:
            m = np.eye(4, dtype=F32)
        self._set_matrix("trans", m)
        self._set_matrix("trans_it", np.linalg.inv(m.astype(np.float64).T))


class _AreaLight(AnalyticLight):
    """AnalyticAreaLight (Light.cpp:307-340): shape in [-1,1]^2 (rect), the unit disc or the unit
    sphere, placed by `transform` x scale(`scaling`). Its LightData position stays where the
    reference leaves it (the origin), which is where PLT beams from it are sourced."""

    def __init__(self, type_: int, name: str = ""):
        super().__init__(type_, name)
        self.data.surface_area = 4.0
        self.setEmissionSolidAngle(1e-6)
        self._transform = np.eye(4, dtype=F32)
        self._scaling = np.ones(3, F32)
        self._update()

    def setEmissionSolidAngle(self, omega: float):
        self.data.cos_subtended_angle = _cos_from_solid_angle(omega)

    @property
    def transform(self):
        return self._transform.copy()

    @transform.setter
    def transform(self, m):
        self._transform = np.asarray(m, F32).reshape(4, 4)
        self._update()

    @property
    def scaling(self):
        return tuple(self._scaling)

    @scaling.setter
    def scaling(self, s):
        self._scaling = np.asarray(s, F32).reshape(3)
        self._update()

    def _update(self):
        m = (self._transform.astype(np.float64) @ np.diag(list(self._scaling.astype(np.float64)) + [1.0])).astype(F32)
        self._set_matrix("trans", m)
        self._set_matrix("trans_it", np.linalg.inv(m.astype(np.float64).T))
        self._m = m
        self._update_area()

    def _axis_lengths(self):
        return [float(np.linalg.norm(self._m[:, i].astype(np.float64))) for i in range(3)]

    def _update_area(self):
        pass


class RectLight(_AreaLight):
    """RectLight (Light.cpp:342-355): area 4 rx ry."""

    def __init__(self, name: str = ""):
        super().__init__(RECT, name)

    def _update_area(self):
        rx, ry, _ = self._axis_lengths()
        self.data.surface_area = float(F32(4.0) * F32(rx) * F32(ry))


class DiscLight(_AreaLight):
    """DiscLight (Light.cpp:357-370): area pi rx ry."""

    def __init__(self, name: str = ""):
        super().__init__(DISC, name)

    def _update_area(self):
        rx, ry, _ = self._axis_lengths()
        self.data.surface_area = float(F32(math.pi) * F32(rx) * F32(ry))


class SphereLight(_AreaLight):
    """SphereLight (Light.cpp:372-390): Knud Thomsen's ellipsoid area approximation as the
    reference writes it (only the last term divided by 3)."""

    def __init__(self, name: str = ""):
        super().__init__(SPHERE, name)

    def _update_area(self):
        rx, ry, rz = (F32(v) for v in self._axis_lengths())
        p = F32(1.6)
        inner = (rx * ry) ** p + (ry * rz) ** p + (rx * rz) ** p / F32(3.0)
        self.data.surface_area = float(F32(4.0) * F32(math.pi) * F32(inner ** (F32(1.0) / p)))


def pack(lights) -> bytes:
    """Contiguous plt_light records."""
    return b"".join(bytes(l.data) for l in lights)


__all__ = ["PointLight", "DirectionalLight", "DistantLight", "RectLight", "DiscLight", "SphereLight", "pack",
           "POINT", "DIRECTIONAL", "DISTANT", "RECT", "DISC", "SPHERE"]
