"""ctypes access to the CPU oracle (oracle/liboracle.so) — the checker for the HIP path."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)

_SIGS = {
    "oracle_interleave_32bit": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "oracle_splitmix64_seq": (None, [C.c_uint64, C.c_uint32, _U64]),
    "oracle_xoshiro128ss_seq": (None, [_U32, C.c_uint32, _U32]),
    "oracle_usg_seq": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _U32]),
    "oracle_sg_streams": (None, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _U32]),
    "oracle_tea": (None, [C.c_uint32, C.c_uint32, _U32, C.c_uint32, _U32]),
    "oracle_f32_to_f16": (C.c_uint16, [C.c_float]),
    "oracle_f16_to_f32": (C.c_float, [C.c_uint16]),
    "oracle_spectral_profile_size": (C.c_uint32, []),
    "oracle_build_profile": (C.c_int, [C.c_uint32, _F, _F, C.c_int, C.c_void_p]),
    "oracle_eval_spectrum_f32": (C.c_float, [C.c_void_p, C.c_float]),
    "oracle_c0_generate": (None, [C.c_uint32, C.c_uint32, _F, C.c_int]),
    "oracle_c0_eval_f32": (None, [C.c_uint32, _F, C.c_void_p, C.c_uint32, _F, C.c_int]),
    "oracle_c0_eval_f64": (None, [C.c_uint32, _F, C.c_void_p, C.c_uint32, _D, C.c_int]),
    "oracle_fresnel_dielectric_f64": (None, [C.c_double, C.c_double, _D]),
    "oracle_fresnel_conductor_f64": (None, [C.c_double, C.c_double, C.c_double, _D]),
    "oracle_mueller_rotate_f64": (None, [_D, _D, _D, _D, _D]),
    "oracle_fractal_pgf_f64": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
    "oracle_gaussian_eval_f64": (C.c_double, [C.c_double, C.c_double, _D, _D, C.c_double, C.c_double]),
    "oracle_boxmueller_trunc_pdf_f64": (C.c_double, [C.c_double] * 5),
    "oracle_rgb_to_spectrum_f64": (C.c_double, [C.c_double] * 4),
    "oracle_wavelength_to_xyz_f64": (None, [C.c_double, _D]),
    "oracle_hwss_wavelengths_f32": (None, [C.c_float, C.c_int, _F]),
    "oracle_surface_sample_f64": (None, [C.c_int, C.c_double, C.c_double, C.c_double, _D, C.c_double, C.c_uint32,
                                         C.c_uint32, _D, _D]),
    "oracle_surface_pdf_f64": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double, _D, _D, C.c_double]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        l = C.CDLL(str(ORACLE_SO))
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def usg_seq(px: int, py: int, sn: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_usg_seq(px, py, sn, n, ptr(out, C.c_uint32))
    return out


def sg_streams(gen_type: int, x0: int, y0: int, w: int, n_streams: int, sample_number: int, draws: int) -> np.ndarray:
    """Raw draws [n_streams, draws] of generator type 0 (tiny) or 1 (uniform) for pixels
    (x0 + s % w, y0 + s // w)."""
    out = np.zeros((n_streams, draws), np.uint32)
    lib().oracle_sg_streams(gen_type, x0, y0, w, n_streams, sample_number, draws, ptr(out, C.c_uint32))
    return out


def tea(v0: int, v1: int, key, iterations: int):
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(2, np.uint32)
    lib().oracle_tea(v0, v1, ptr(k, C.c_uint32), iterations, ptr(out, C.c_uint32))
    return int(out[0]), int(out[1])


def build_profile(wl, val, compute_icdf: bool) -> bytes:
    size = lib().oracle_spectral_profile_size()
    buf = C.create_string_buffer(size)
    wl = np.ascontiguousarray(wl, dtype=np.float32)
    val = np.ascontiguousarray(val, dtype=np.float32)
    rc = lib().oracle_build_profile(len(wl), ptr(wl, C.c_float), ptr(val, C.c_float), int(compute_icdf), buf)
    assert rc == 0
    return buf.raw


def c0_generate(n: int, seed: int, threads: int = 0) -> np.ndarray:
    q = np.zeros((16, n), dtype=np.float32)
    lib().oracle_c0_generate(n, seed, ptr(q, C.c_float), threads)
    return q


def c0_eval(queries: np.ndarray, profiles: bytes, nprof: int, fp64: bool = False, threads: int = 0) -> np.ndarray:
    """queries: [16, n] float32 view (integer planes as raw bits). Returns [5, n] results."""
    q = np.ascontiguousarray(queries).view(np.float32)
    n = q.shape[1]
    pbuf = C.create_string_buffer(profiles, len(profiles))
    if fp64:
        out = np.zeros((5, n), dtype=np.float64)
        lib().oracle_c0_eval_f64(n, ptr(q, C.c_float), pbuf, nprof, ptr(out, C.c_double), threads)
    else:
        out = np.zeros((5, n), dtype=np.float32)
        lib().oracle_c0_eval_f32(n, ptr(q, C.c_float), pbuf, nprof, ptr(out, C.c_float), threads)
    return out


def dvec(*v) -> np.ndarray:
    return np.array(v, dtype=np.float64)


_SIGS_GEOM = {
    "oracle_trace": (C.c_int, [_F, C.c_uint32, _U32, C.c_uint32, _F, _U32, C.c_int, C.c_int, C.c_int]),
    "oracle_compute_ray_origin": (None, [_F, _F, _F]),
}


def _geom():
    l = lib()
    for name, (res, args) in _SIGS_GEOM.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    return l


def trace(positions: np.ndarray, indices: np.ndarray, rays: np.ndarray, any_hit=False, brute=False, threads=0):
    """positions [V,3] f32, indices [T,3] u32, rays [8,n] f32 -> [4,n] u32 or [n] u32."""
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    n = r.shape[1]
    out = np.zeros((1 if any_hit else 4, n), dtype=np.uint32)
    _geom().oracle_trace(ptr(pos, C.c_float), idx.shape[0], ptr(idx, C.c_uint32), n, ptr(r, C.c_float),
                         ptr(out, C.c_uint32), 1 if any_hit else 0, 1 if brute else 0, threads)
    return out[0] if any_hit else out


def compute_ray_origin(pos, n):
    p = np.ascontiguousarray(pos, dtype=np.float32)
    nn = np.ascontiguousarray(n, dtype=np.float32)
    out = np.zeros(3, np.float32)
    _geom().oracle_compute_ray_origin(ptr(p, C.c_float), ptr(nn, C.c_float), ptr(out, C.c_float))
    return out


REC_WORDS = 9


class OracleScene:
    """A scene built once in the oracle (BVH and light tables); keeps its arrays alive."""

    def __init__(self, scene: dict):
        l = lib()
        l.oracle_scene_create.restype = C.c_void_p
        l.oracle_scene_create.argtypes = [_F, _F, C.c_uint32, _U32, _U32, _U32, C.c_char_p, C.c_char_p, C.c_uint32]
        l.oracle_scene_destroy.restype = None
        l.oracle_scene_destroy.argtypes = [C.c_void_p]
        l.oracle_scene_render.restype = C.c_int
        l.oracle_scene_render.argtypes = [C.c_void_p, _F, C.POINTER(C.c_int32), _F, C.c_uint32, C.c_uint32, _F, _U32,
                                          _F, C.c_int]
        self.lib = l
        self.arrays = dict(
            pos=np.ascontiguousarray(scene["positions"], np.float32),
            nrm=np.ascontiguousarray(scene["normals"], np.float32),
            idx=np.ascontiguousarray(scene["indices"], np.uint32),
            tm=np.ascontiguousarray(scene["tri_mesh"], np.uint32),
            mm=np.ascontiguousarray(scene["mesh_material"], np.uint32),
            mats=scene["materials"], prof=scene["profiles"])
        a = self.arrays
        self.camera = np.ascontiguousarray(scene["camera"], np.float32)
        self.handle = l.oracle_scene_create(ptr(a["pos"], C.c_float), ptr(a["nrm"], C.c_float), a["idx"].shape[0],
                                            ptr(a["idx"], C.c_uint32), ptr(a["tm"], C.c_uint32), ptr(a["mm"], C.c_uint32),
                                            a["mats"], a["prof"], scene["n_profiles"])
        lights = scene.get("lights", b"")
        if lights:
            l.oracle_scene_set_lights.restype = C.c_int
            l.oracle_scene_set_lights.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32]
            self._lights = bytes(lights)
            l.oracle_scene_set_lights(self.handle, self._lights, len(self._lights) // 192)
        env = scene.get("env")
        if env is not None:
            l.oracle_scene_set_env.restype = C.c_int
            l.oracle_scene_set_env.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, _F, _F, _F]
            self._env = [np.ascontiguousarray(env[k], np.float32) for k in ("rgba", "transform", "scale")]
            l.oracle_scene_set_env(self.handle, env["width"], env["height"], ptr(self._env[0], C.c_float),
                                   ptr(self._env[1], C.c_float), ptr(self._env[2], C.c_float))

    def render(self, width, height, frame=0, max_bounces=6, first=0, count=None, do_nee=1, do_mis=1, do_rr=1,
               hwss_mis=1, is_emitters=1, direct=1, emissive=1, sourcing_area=1.0, max_omega=0.025, threads=0,
               records=False, camera=None, do_mnee=0, mnee_max_occluders=2, mnee_max_iterations=60,
               mnee_threshold=5e-5, emissive_sampler=2, analytic=1, env=1, sample_generator=1, per_tile_sg=0,
               tile_size=512, aov=False):
        """Renders pixels [first, first+count) (row-major). Returns color [count, 4], then optionally the
        bounce records [(max_bounces+1), 9, count] (uint32) and the albedo and view-space normal outputs
        [count, 8] (aov)."""
        if count is None:
            count = width * height - first
        cam = self.camera if camera is None else np.ascontiguousarray(camera, np.float32)
        if cam.size < 16:  # polarizer off
            cam = np.concatenate([cam, np.zeros(16 - cam.size, np.float32)])
        if emissive_sampler not in (0, 2):
            raise ValueError("emissive_sampler: 0 (Uniform) or 2 (Power)")
        iopt = np.array([width, height, frame, max_bounces, do_nee, do_mis, do_rr, hwss_mis, is_emitters, direct,
                         emissive, do_mnee, mnee_max_occluders, mnee_max_iterations, int(emissive_sampler == 0),
                         analytic, env, sample_generator, per_tile_sg, tile_size], np.int32)
        fopt = np.array([sourcing_area, max_omega, mnee_threshold], np.float32)
        color = np.zeros((count, 4), np.float32)
        rec = np.zeros(((max_bounces + 1) * REC_WORDS, count), np.uint32) if records else None
        aovs = np.zeros((count, 8), np.float32) if aov else None
        self.lib.oracle_scene_render(self.handle, ptr(cam, C.c_float), iopt.ctypes.data_as(C.POINTER(C.c_int32)),
                                     ptr(fopt, C.c_float), first, count, ptr(color, C.c_float),
                                     ptr(rec, C.c_uint32) if records else None, ptr(aovs, C.c_float) if aov else None,
                                     threads)
        out = [color]
        if records:
            out.append(rec.reshape(max_bounces + 1, REC_WORDS, count))
        if aov:
            out.append(aovs)
        return out[0] if len(out) == 1 else tuple(out)

    def mnee_probe(self, origin, light_pos, wavelength=550.0, max_ms_bounces=2, max_occluders=2, max_iterations=60,
                   threshold=5e-5, seed=0):
        """One manifold walk from `origin` to `light_pos`. Returns (vertex count or -1 on failure,
        vertices [count, 3], lobes [count], iterations, pdf)."""
        l = self.lib
        l.oracle_mnee_probe.restype = C.c_int
        l.oracle_mnee_probe.argtypes = [C.c_void_p, _F, _F, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint32, _F,
                                        _U32, C.POINTER(C.c_int), _F]
        o = np.ascontiguousarray(origin, np.float32)
        lp = np.ascontiguousarray(light_pos, np.float32)
        pos = np.zeros((8, 3), np.float32)
        lobes = np.zeros(8, np.uint32)
        it = C.c_int(0)
        pdf = np.zeros(1, np.float32)
        n = l.oracle_mnee_probe(self.handle, ptr(o, C.c_float), ptr(lp, C.c_float), wavelength, max_ms_bounces,
                                max_occluders, max_iterations, threshold, seed, ptr(pos, C.c_float), ptr(lobes, C.c_uint32),
                                C.byref(it), ptr(pdf, C.c_float))
        k = max(n, 0)
        return n, pos[:k].copy(), lobes[:k].copy(), it.value, float(pdf[0])

    def close(self):
        if getattr(self, "handle", None):
            self.lib.oracle_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render(scene: dict, width: int, height: int, **kw):
    """One-shot oracle render (builds the scene's BVH each call); see OracleScene.render."""
    s = OracleScene(scene)
    try:
        return s.render(width, height, **kw)
    finally:
        s.close()


def alias_table(scene: dict):
    """The oracle's emissive power alias table: (threshold, redirect, perm, pdf) per light."""
    l = lib()
    fn = l.oracle_alias_table
    fn.restype = C.c_int
    fn.argtypes = [_F, C.c_uint32, _U32, _U32, _U32, C.c_char_p, _F, _U32, _U32, _F, C.c_uint32]
    pos = np.ascontiguousarray(scene["positions"], np.float32)
    idx = np.ascontiguousarray(scene["indices"], np.uint32)
    tm = np.ascontiguousarray(scene["tri_mesh"], np.uint32)
    mm = np.ascontiguousarray(scene["mesh_material"], np.uint32)
    cap = idx.shape[0]
    thr = np.zeros(cap, np.float32)
    red = np.zeros(cap, np.uint32)
    perm = np.zeros(cap, np.uint32)
    pdf = np.zeros(cap, np.float32)
    n = fn(ptr(pos, C.c_float), idx.shape[0], ptr(idx, C.c_uint32), ptr(tm, C.c_uint32), ptr(mm, C.c_uint32),
           scene["materials"], ptr(thr, C.c_float), ptr(red, C.c_uint32), ptr(perm, C.c_uint32), ptr(pdf, C.c_float), cap)
    return thr[:n], red[:n], perm[:n], pdf[:n]


def material_eval(material: bytes, profiles: bytes, n_profiles: int, wi, wo, lobe: int, lambdas, stokes, area: float,
                  dbl: bool = True):
    """One material query in the local frame (N = +z) through the oracle's material interface.
    Returns (stokes_out [4, 4], pdf, lambda_pdf [4])."""
    l = lib()
    fn = l.oracle_material_eval
    fn.restype = C.c_int
    D = C.POINTER(C.c_double)
    fn.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, D, D, C.c_uint32, D, D, C.c_double, C.c_int, D, D, D]
    a = lambda v, n: np.ascontiguousarray(np.asarray(v, np.float64).reshape(n))  # noqa: E731
    wi_, wo_, lam, st = a(wi, 3), a(wo, 3), a(lambdas, 4), a(stokes, 16)
    out = np.zeros(16)
    pdf = np.zeros(1)
    lp = np.zeros(4)
    fn(material, profiles, n_profiles, wi_.ctypes.data_as(D), wo_.ctypes.data_as(D), lobe, lam.ctypes.data_as(D),
       st.ctypes.data_as(D), area, int(dbl), out.ctypes.data_as(D), pdf.ctypes.data_as(D), lp.ctypes.data_as(D))
    return out.reshape(4, 4), float(pdf[0]), lp


def material_sample(material: bytes, profiles: bytes, n_profiles: int, wi, wavelength, n: int = None, frame: int = 0):
    """FP32 samples, query i with direction wi[i] (or one shared wi) and generator (i, 0, frame):
    (wo [n, 3], pdf [n], weight [n], lobe [n], valid [n])."""
    l = lib()
    fn = l.oracle_material_sample
    fn.restype = C.c_int
    fn.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, _F, _F, C.c_uint32, C.c_uint32, _F, _F, _F, _U32, _U32]
    w = np.asarray(wi, np.float32)
    if w.ndim == 1:
        w = np.tile(w, (n, 1))
    w = np.ascontiguousarray(w.reshape(-1, 3))
    n = w.shape[0]
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(wavelength, np.float32), (n,)))
    wo = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    weight = np.zeros(n, np.float32)
    lobe = np.zeros(n, np.uint32)
    valid = np.zeros(n, np.uint32)
    fn(material, profiles, n_profiles, ptr(w, C.c_float), ptr(lam, C.c_float), n, frame, ptr(wo, C.c_float),
       ptr(pdf, C.c_float), ptr(weight, C.c_float), ptr(lobe, C.c_uint32), ptr(valid, C.c_uint32))
    return wo, pdf, weight, lobe, valid


def walk_trig(x, y):
    """The manifold walk's FP64 acos(x) and atan2(y, x) (its Newton step's spherical angles)."""
    l = lib()
    _D = C.POINTER(C.c_double)
    l.oracle_walk_trig.restype = None
    l.oracle_walk_trig.argtypes = [_D, _D, C.c_uint32, _D, _D]
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    a, t = np.zeros_like(x), np.zeros_like(x)
    l.oracle_walk_trig(ptr(x, C.c_double), ptr(y, C.c_double), x.size, ptr(a, C.c_double), ptr(t, C.c_double))
    return a, t


def walk_derivatives(material: bytes, profiles: bytes, nprof: int, P, N, b1, b2, src, lobe, wavelength):
    """FP64 specular direction and its (u, v) derivatives at a walk vertex: array [3 (v, ddu, ddv), 3]."""
    l = lib()
    _D = C.POINTER(C.c_double)
    l.oracle_walk_derivatives.restype = None
    l.oracle_walk_derivatives.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, _D, _D, C.c_double, C.c_double, _D,
                                          C.c_uint32, C.c_double, _D]
    P = np.ascontiguousarray(P, np.float64)
    N = np.ascontiguousarray(N, np.float64)
    src = np.ascontiguousarray(src, np.float64)
    out = np.zeros(9)
    l.oracle_walk_derivatives(material, profiles, nprof, ptr(P, C.c_double), ptr(N, C.c_double), b1, b2,
                              ptr(src, C.c_double), lobe, wavelength, ptr(out, C.c_double))
    return out.reshape(3, 3)
