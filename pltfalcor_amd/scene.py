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


# The render with the learned surrogate (C4): the 1080p workload scene plus a rough diffraction
# grating on the ground under the light and, between them, a 5 mm glass slab (IoR 1.5, smooth), so
# that light reaches the grating through the slab's two faces by manifold NEE.
C4_GRATING = dict(size=3.0, height=0.01, pitch_um=1.0, direction=0.5, roughness=0.3)
C4_SLAB = dict(size=3.6, thickness=0.005, center_y=2.6, ior=1.5)


def c4_scene(ctx, aspect: float):
    """The 1080p workload's procedural scene (PATH_2) with the grating plate and the glass slab added."""
    from . import materials as M
    from .meshes import create_cube, create_quad
    base = Scene(None, aspect=aspect, **PATH_2)
    a = base.numpy()
    cam = base.arrays.camera
    base.close()
    pos, nrm, idx = [a["positions"]], [a["normals"]], [a["indices"].astype(np.int64)]
    tri_mesh, mesh_mat = [a["tri_mesh"]], [a["mesh_material"]]
    mats = bytearray(a["materials"])
    n_mat = len(mats) // C.sizeof(Material)
    n_mesh = int(a["mesh_material"].size)
    nv = int(a["positions"].shape[0])

    def add(p, n, i, material):
        nonlocal nv, n_mesh, n_mat, mats
        pos.append(np.asarray(p, np.float32))
        nrm.append(np.asarray(n, np.float32))
        idx.append(np.asarray(i, np.int64) + nv)
        tri_mesh.append(np.full(len(i), n_mesh, np.uint32))
        mesh_mat.append(np.array([n_mat], np.uint32))
        mats += M.pack([material])
        nv += len(p)
        n_mesh += 1
        n_mat += 1

    g = C4_GRATING
    qp, qn, qi, _ = create_quad((g["size"], g["size"]))
    add(qp + np.array([0.0, g["height"], 0.0]), qn, qi,
        M.grated_conductor(g["pitch_um"], g["direction"], 0, 1, roughness=g["roughness"]))
    sl = C4_SLAB
    cp, cn, ci, _ = create_cube((sl["size"], sl["thickness"], sl["size"]))
    add(cp + np.array([0.0, sl["center_y"], 0.0]), cn, ci, M.dielectric(sl["ior"], 0.0))
    return Scene.from_arrays(ctx, np.concatenate(pos), np.concatenate(nrm), np.concatenate(idx).astype(np.uint32),
                             np.concatenate(tri_mesh), np.concatenate(mesh_mat), bytes(mats), a["profiles"], cam)


def _builder(name: str) -> int:
    if name not in ("host", "device"):
        raise ValueError("BVH builder must be 'host' or 'device'")
    return _lib.BVH_HOST_SAH if name == "host" else _lib.BVH_DEVICE_PLOC


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
                    profiles: bytes, camera, bvh: str = "host"):
        """materials / profiles: packed plt_material / plt_spectral_profile records; camera: Camera.
        bvh: "host" (binned SAH on the CPU) or "device" (built on the GPU; needs ctx)."""
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
        _lib.check(lib.plt_scene_create_ex(ctx.handle if ctx is not None else None, C.byref(d), _builder(bvh),
                                           C.byref(h), None))
        return cls(ctx, _handle=h)

    def build_bvh(self, builder: str = "device", stream=None) -> dict:
        """Rebuilds the scene's BVH with the host ("host") or device ("device") builder and returns
        bvh_info(). A device tree deeper than the traversal stack holds raises (PLT_EINVAL) and the
        scene keeps its tree."""
        if stream is None and self.ctx is not None and torch.cuda.is_available():
            stream = torch.cuda.current_stream()
        info = _lib.BvhInfo()
        _lib.check(self.lib.plt_scene_build_bvh(self.handle, _builder(builder),
                                                C.c_void_p(stream.cuda_stream) if stream is not None else None,
                                                C.byref(info)))
        self._refresh_arrays()
        return self.bvh_info()

    def bvh_info(self) -> dict:
        info = _lib.BvhInfo()
        _lib.check(self.lib.plt_scene_get_bvh(self.handle, C.byref(info), None, 0, None, 0))
        return dict(builder="device" if info.builder == _lib.BVH_DEVICE_PLOC else "host", n_nodes=info.n_nodes,
                    max_depth=info.max_depth, iterations=info.iterations, build_ms=info.build_ms)

    def bvh_arrays(self):
        """Host copies of the tree: nodes as a [n_nodes, 32] float32 array (rows lo x, hi x, lo y, hi y,
        lo z, hi z of the four children, then the four child references and padding as int32 bits)
        and the leaf-order triangle records as [n_triangles, 3, 4] float32 (v0 with the triangle id's
        bits in w, v1, v2)."""
        info = _lib.BvhInfo()
        _lib.check(self.lib.plt_scene_get_bvh(self.handle, C.byref(info), None, 0, None, 0))
        nodes = np.zeros((info.n_nodes, 32), np.float32)
        tris = np.zeros((self.arrays.n_triangles, 3, 4), np.float32)
        _lib.check(self.lib.plt_scene_get_bvh(self.handle, None, C.c_void_p(nodes.ctypes.data), nodes.nbytes,
                                              C.c_void_p(tris.ctypes.data), tris.nbytes))
        return nodes, tris

    def _refresh_arrays(self):
        a = SceneArrays()
        _lib.check(self.lib.plt_scene_get_arrays(self.handle, C.byref(a)))
        self.arrays = a

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

    def set_camera(self, **fields):
        """Changes camera fields (pos, target, up, fov_y, aspect, focal_distance, jitter, polarizer,
        polarizer_theta) for the next execute of a bound pass."""
        c = self.arrays.camera
        for k, v in fields.items():
            if not hasattr(c, k):
                raise KeyError(k)
            cur = getattr(c, k)
            if isinstance(cur, C.Array):
                for i, x in enumerate(v):
                    cur[i] = float(x)
            else:
                setattr(c, k, float(v))
        _lib.check(self.lib.plt_scene_set_camera(self.handle, C.byref(c)))

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
