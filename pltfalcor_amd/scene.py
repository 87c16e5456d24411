"""Scenes (geometry + BVH + materials + camera in HBM) and batched ray casting over the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


from ._lib import Camera, Material, ProceduralSceneDesc, SceneArrays  # noqa: F401


# Workload scenes: sphere count / subdivision / material kinds
PATH_1 = dict(seed=1, sphere_count=64, subdivision=5, material_kinds=4)   # 512x512, ~1.31M triangles
PATH_2 = dict(seed=1, sphere_count=256, subdivision=5, material_kinds=8)  # 1920x1080, ~5.24M triangles


class Scene:
    """A scene on one GPU (or host-only when ctx is None): the procedural workload scene, or
    `Scene.from_arrays` for caller geometry, materials and spectral profiles."""

    def __init__(self, ctx=None, seed=1, sphere_count=64, subdivision=5, material_kinds=4, ground_size=20.0,
                 aspect=1.0, _handle=None):
        self.lib = _lib.load()
        self.ctx = ctx
        if _handle is None:
            d = ProceduralSceneDesc(seed, sphere_count, subdivision, material_kinds, ground_size, aspect)
            _handle = C.c_void_p()
            _lib.check(self.lib.plt_scene_create_procedural(ctx.handle if ctx is not None else None, C.byref(d),
                                                            C.byref(_handle)))
        self.handle = _handle
        self.lights = b""
        self.env = None
        a = SceneArrays()
        _lib.check(self.lib.plt_scene_get_arrays(self.handle, C.byref(a)))
        self.arrays = a

    @classmethod
    def from_arrays(cls, ctx, positions, normals, indices, tri_mesh, mesh_material, materials: bytes,
                    profiles: bytes, camera):
        """materials / profiles: packed plt_material / plt_spectral_profile records; camera: Camera."""
        lib = _lib.load()
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        nrm = np.ascontiguousarray(normals, np.float32).reshape(-1)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        tm = np.ascontiguousarray(tri_mesh, np.uint32)
        mm = np.ascontiguousarray(mesh_material, np.uint32)
        nmat = len(materials) // C.sizeof(Material)
        nprof = len(profiles) // C.sizeof(_lib.SpectralProfile)
        mats = (Material * max(nmat, 1)).from_buffer_copy(materials) if nmat else None
        profs = (_lib.SpectralProfile * nprof).from_buffer_copy(profiles) if nprof else None
        d = _lib.SceneDesc(pos.size // 3, idx.size // 3, mm.size, nmat, nprof,
                           pos.ctypes.data_as(C.POINTER(C.c_float)), nrm.ctypes.data_as(C.POINTER(C.c_float)),
                           idx.ctypes.data_as(C.POINTER(C.c_uint32)), tm.ctypes.data_as(C.POINTER(C.c_uint32)),
                           mm.ctypes.data_as(C.POINTER(C.c_uint32)),
                           C.cast(mats, C.POINTER(Material)) if mats is not None else None,
                           C.cast(profs, C.POINTER(_lib.SpectralProfile)) if profs is not None else None, camera)
        h = C.c_void_p()
        _lib.check(lib.plt_scene_create(ctx.handle if ctx is not None else None, C.byref(d), C.byref(h)))
        return cls(ctx, _handle=h)

    def set_lights(self, lights):
        """Analytic lights (pltfalcor_amd.lights objects or packed plt_light bytes); bind the scene to
        a pass afterwards (the pass reads them then)."""
        from .lights import pack
        data = lights if isinstance(lights, (bytes, bytearray)) else pack(lights)
        n = len(data) // C.sizeof(_lib.Light)
        buf = C.create_string_buffer(bytes(data), len(data)) if n else None
        _lib.check(self.lib.plt_scene_set_lights(self.handle, buf, n))
        self.lights = bytes(data)

    def set_env_map(self, image, rotation=(0.0, 0.0, 0.0), intensity: float = 1.0, tint=(1.0, 1.0, 1.0)):
        """Environment map light (EnvMap): `image` a lat-long radiance image [height, width, 3 or 4]
        (row 0 at the +y pole, u = 0.5 towards -z), rotated by the reference's XYZ Euler angles in
        degrees (EnvMap::setRotation: eulerAngleXYZ = Rx Ry Rz), scaled by intensity x tint. None
        removes it."""
        if image is None:
            _lib.check(self.lib.plt_scene_set_env_map(self.handle, 0, 0, None, None, 1.0, None))
            self.env = None
            return
        img = np.asarray(image, np.float32)
        h, w = img.shape[:2]
        rgba = np.ones((h, w, 4), np.float32)
        rgba[..., :3] = img[..., :3]
        ax, ay, az = (np.radians(float(a)) for a in rotation)
        rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
        ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
        rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
        m = (rx @ ry @ rz).astype(np.float32)
        t = np.asarray(tint, np.float32)
        _lib.check(self.lib.plt_scene_set_env_map(self.handle, w, h, C.c_void_p(rgba.ctypes.data),
                                                  C.c_void_p(m.ctypes.data), float(intensity),
                                                  C.c_void_p(t.ctypes.data)))
        scale = (np.float32(intensity) * t).astype(np.float32)
        self.env = dict(width=w, height=h, rgba=rgba, transform=m.reshape(9), scale=scale)

    def numpy(self):
        """Host copies of the scene arrays."""
        a = self.arrays
        cp = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy()  # noqa: E731
        return dict(
            positions=cp(a.positions, 3 * a.n_vertices, np.float32).reshape(-1, 3),
            normals=cp(a.normals, 3 * a.n_vertices, np.float32).reshape(-1, 3),
            indices=cp(a.indices, 3 * a.n_triangles, np.uint32).reshape(-1, 3),
            tri_mesh=cp(a.tri_mesh, a.n_triangles, np.uint32),
            mesh_first_tri=cp(a.mesh_first_tri, a.n_meshes, np.uint32),
            mesh_material=cp(a.mesh_material, a.n_meshes, np.uint32),
            materials=C.string_at(a.materials, a.n_materials * C.sizeof(Material)),
            profiles=C.string_at(a.profiles, a.n_profiles * C.sizeof(_lib.SpectralProfile)),
            n_profiles=a.n_profiles,
            lights=self.lights,
            env=self.env,
            camera=np.array(list(a.camera.pos) + list(a.camera.target) + list(a.camera.up) +
                            [a.camera.fov_y, a.camera.aspect, a.camera.focal_distance] + list(a.camera.jitter) +
                            [a.camera.polarizer, a.camera.polarizer_theta],
                            dtype=np.float32),
        )

    def trace(self, rays: torch.Tensor, any_hit: bool = False, stream=None) -> torch.Tensor:
        """rays: [8, n] float32 on the GPU. Returns [4, n] int32 (id, b1, b2, t bits) or [n] int32 (0/1)."""
        if rays.dtype != torch.float32 or rays.dim() != 2 or rays.shape[0] != 8 or not rays.is_cuda:
            raise ValueError("rays must be a float32 CUDA tensor of shape [8, n]")
        rays = rays.contiguous()
        n = rays.shape[1]
        out = torch.empty((1 if any_hit else 4, n), dtype=torch.int32, device=rays.device)
        if stream is None:
            stream = torch.cuda.current_stream()
        _lib.check(self.lib.plt_scene_trace(self.handle, n, C.c_void_p(rays.data_ptr()), C.c_void_p(out.data_ptr()),
                                            1 if any_hit else 0, C.c_void_p(stream.cuda_stream)))
        return out[0] if any_hit else out

    def close(self):
        if getattr(self, "handle", None):
            self.lib.plt_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
