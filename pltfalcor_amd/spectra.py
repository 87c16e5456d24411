"""Spectral tables for the synthetic workloads, and host construction of spectral profiles.

Profiles are built by the library (plt_spectral_profile_build), which restates the reference's
PiecewiseLinearSpectrum -> SampledSpectrum -> genSpectralProfile pipeline
(Utils/Color/Spectrum.h:88-123, Utils/Color/SampledSpectrum.h:64-74, Scene/SceneBuilder.cpp:911-960).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

HC_EV_NM = 1239.841984  # h c in eV nm


def drude_table(n_points: int = 64, lo: float = 380.0, hi: float = 780.0,
                plasma_ev: float = 9.0, damping_ev: float = 0.05):
    """Synthetic Drude metal: eps = 1 - wp^2 / (w^2 + i g w), n + i k = sqrt(eps) (principal branch).

    Returns float32 arrays (wavelength nm, n, k) with n_points samples spaced uniformly on [lo, hi].
    Defaults give a silver-like metal (n ~ 0.05-0.1, k ~ 2.5-5 over the visible range).
    """
    wl = np.linspace(lo, hi, n_points, dtype=np.float64)
    w = HC_EV_NM / wl
    eps = 1.0 - plasma_ev ** 2 / (w * w + 1j * damping_ev * w)
    nk = np.sqrt(eps)
    return wl.astype(np.float32), nk.real.astype(np.float32), np.abs(nk.imag).astype(np.float32)


def blackbody_table(temperature: float = 6500.0, lo: float = 380.0, hi: float = 780.0, step: float = 5.0):
    """Planck radiance tabulated in `step` nm bins, as BlackbodySpectrum (Utils/Color/Spectrum.cpp:125-138)."""
    c, h, kb = 299792458.0, 6.62606957e-34, 1.3806488e-23
    wl = np.arange(lo, hi + 0.5 * step, step, dtype=np.float64)
    lm = wl * 1e-9
    le = (2.0 * h * c * c) / ((lm ** 5) * (np.exp((h * c) / (lm * kb * temperature)) - 1.0) * 1e9)
    return wl.astype(np.float32), le.astype(np.float32)


def build_profile(wavelengths, values, compute_icdf: bool) -> _lib.SpectralProfile:
    wl = np.ascontiguousarray(wavelengths, dtype=np.float32)
    v = np.ascontiguousarray(values, dtype=np.float32)
    if wl.shape != v.shape or wl.ndim != 1:
        raise ValueError("wavelengths and values must be 1-D arrays of equal length")
    out = _lib.SpectralProfile()
    lib = _lib.load()
    _lib.check(lib.plt_spectral_profile_build(len(wl), wl.ctypes.data_as(C.POINTER(C.c_float)),
                                               v.ctypes.data_as(C.POINTER(C.c_float)), int(compute_icdf),
                                               C.byref(out)))
    return out


def profile_to_numpy(p: _lib.SpectralProfile) -> np.ndarray:
    """Raw bytes of a profile (for bit-exact comparisons)."""
    return np.frombuffer(bytes(p), dtype=np.uint8).copy()


def drude_profiles():
    """(n, k) profiles of the synthetic Drude conductor (ids 0 and 1 of the BSDF batch workload)."""
    wl, n, k = drude_table()
    return build_profile(wl, n, False), build_profile(wl, k, False)
