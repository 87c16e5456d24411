"""Triangle meshes for the scene importers: the reference's built-in shape generators and mesh
operations (Scene/TriangleMesh.cpp), a PLY reader, and the face-orientation rule the scene uses.

Every generator returns (positions, normals, indices, frontFaceCW) in float64. The reference keeps
a per-mesh `frontFaceCW` flag and computes a triangle's world face normal as the object-space
cross product (negated for clockwise meshes) taken through the inverse-transpose of the instance
transform (Scene.slang getFaceNormalW). Meshes here are baked into world space, so `bake` swaps
the winding of meshes whose flag ends up set, which gives the same face normals.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

F32 = np.float32
Y_TO_Z = np.array([[1.0, 0, 0, 0], [0, 0, 1.0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0]])  # swaps y and z


def translate(v):
    m = np.eye(4)
    m[:3, 3] = v
    return m


def create_quad(size=(2.0, 2.0)):
    """TriangleMesh::createQuad: a quad in the y = 0 plane facing +y."""
    hx, hz = 0.5 * size[0], 0.5 * size[1]
    pos = np.array([[-hx, 0, -hz], [hx, 0, -hz], [-hx, 0, hz], [hx, 0, hz]], np.float64)
    nrm = np.tile([0.0, 1.0, 0.0], (4, 1))
    return pos, nrm, np.array([[2, 1, 0], [1, 2, 3]], np.int64), size[0] * size[1] < 0


def create_disk(radius: float = 1.0, segments: int = 32):
    """TriangleMesh::createDisk: a disk in the y = 0 plane facing +y."""
    pos, nrm, idx = [[0.0, 0.0, 0.0]], [[0.0, 1.0, 0.0]], []
    for i in range(segments):
        phi = F32(i) / F32(segments) * F32(2.0) * F32(math.pi)
        c, s = np.cos(phi), -np.sin(phi)
        pos.append([float(c * F32(radius)), 0.0, float(s * F32(radius))])
        nrm.append([0.0, 1.0, 0.0])
        idx.append([0, i + 1, (i + 1) % segments + 1])
    return np.array(pos), np.array(nrm), np.array(idx, np.int64), False


_CUBE_POS = [
    [[-.5, -.5, -.5], [-.5, -.5, .5], [.5, -.5, .5], [.5, -.5, -.5]],
    [[-.5, .5, .5], [-.5, .5, -.5], [.5, .5, -.5], [.5, .5, .5]],
    [[-.5, .5, -.5], [-.5, -.5, -.5], [.5, -.5, -.5], [.5, .5, -.5]],
    [[.5, .5, .5], [.5, -.5, .5], [-.5, -.5, .5], [-.5, .5, .5]],
    [[-.5, .5, .5], [-.5, -.5, .5], [-.5, -.5, -.5], [-.5, .5, -.5]],
    [[.5, .5, -.5], [.5, -.5, -.5], [.5, -.5, .5], [.5, .5, .5]],
]
_CUBE_NRM = [[0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1], [-1, 0, 0], [1, 0, 0]]


def create_cube(size=(1.0, 1.0, 1.0)):
    """TriangleMesh::createCube: six faces of four vertices, two triangles each."""
    size = np.asarray(size, np.float64)
    sign = np.where(size < 0, -1.0, 1.0)
    pos, nrm, idx = [], [], []
    for f in range(6):
        b = len(pos)
        idx += [[b, b + 2, b + 1], [b, b + 3, b + 2]]
        for j in range(4):
            pos.append(np.asarray(_CUBE_POS[f][j]) * size)
            nrm.append(np.asarray(_CUBE_NRM[f], np.float64) * sign)
    return np.array(pos), np.array(nrm), np.array(idx, np.int64), bool(np.prod(size) < 0)


def create_sphere(radius: float = 0.5, segments_u: int = 32, segments_v: int = 16):
    """TriangleMesh::createSphere: a UV sphere (its pole triangles have zero area, as in the reference)."""
    u = np.arange(segments_u + 1, dtype=F32)
    v = np.arange(segments_v + 1, dtype=F32)
    theta = (u / F32(segments_u)) * F32(2.0) * F32(math.pi)
    phi = (v / F32(segments_v)) * F32(math.pi)
    T, P = np.meshgrid(theta, phi)  # rows: v, columns: u
    d = np.stack([np.cos(T) * np.sin(P), np.cos(P), np.sin(T) * np.sin(P)], -1).astype(F32).reshape(-1, 3)
    idx = []
    w = segments_u + 1
    for j in range(segments_v):
        for i in range(segments_u):
            i0, i1 = j * w + i, j * w + (i + 1) % w
            i2, i3 = (j + 1) * w + i, (j + 1) * w + (i + 1) % w
            idx += [[i0, i1, i2], [i2, i1, i3]]
    return (d * F32(radius)).astype(np.float64), d.astype(np.float64), np.array(idx, np.int64), False


def apply_transform(m, pos, nrm, cw):
    """TriangleMesh::applyTransform: positions, normals through the inverse-transpose (normalised),
    and the winding flag toggled by a mirroring transform."""
    m = np.asarray(m, np.float64)
    pos = (np.c_[pos, np.ones(len(pos))] @ m.T)[:, :3]
    nrm = nrm @ np.linalg.inv(m[:3, :3])
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(ln > 0, nrm / np.where(ln > 0, ln, 1.0), 0.0)
    return pos, nrm, cw ^ bool(np.linalg.det(m[:3, :3]) < 0)


def flip_normals(pos, nrm, idx, cw):
    """TriangleMesh::flipNormals."""
    return pos, -nrm, idx, not cw


def bake(pos, nrm, idx, cw):
    """World-space arrays whose winding gives the reference's face normals."""
    return pos.astype(F32), nrm.astype(F32), (idx[:, [0, 2, 1]] if cw else idx).astype(np.uint32)


def smooth_normals(pos, idx):
    """Area-weighted vertex normals."""
    acc = np.zeros_like(pos, dtype=np.float64)
    fn = np.cross(pos[idx[:, 1]] - pos[idx[:, 0]], pos[idx[:, 2]] - pos[idx[:, 0]])
    for k in range(3):
        np.add.at(acc, idx[:, k], fn)
    ln = np.linalg.norm(acc, axis=1, keepdims=True)
    return acc / np.where(ln > 0, ln, 1.0)


def face_normal_mesh(pos, idx):
    """Unshared vertices carrying their triangle's face normal (flat shading)."""
    p = pos[idx.reshape(-1)]
    fn = np.cross(pos[idx[:, 1]] - pos[idx[:, 0]], pos[idx[:, 2]] - pos[idx[:, 0]])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
    return p, np.repeat(fn, 3, 0), np.arange(len(p), dtype=np.int64).reshape(-1, 3)


_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "<i2", "int16": "<i2",
              "ushort": "<u2", "uint16": "<u2", "int": "<i4", "int32": "<i4", "uint": "<u4", "uint32": "<u4",
              "float": "<f4", "float32": "<f4", "double": "<f8", "float64": "<f8"}


def load_ply(path):
    """Positions, normals (None when absent) and fan-triangulated faces of an ASCII or binary
    little-endian PLY file."""
    data = Path(path).read_bytes()
    end = data.index(b"end_header") + len(b"end_header")
    end = data.index(b"\n", end) + 1
    fmt, elements = None, []
    for line in data[:end].decode("ascii").splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "format":
            fmt = t[1]
        elif t[0] == "element":
            elements.append((t[1], int(t[2]), []))
        elif t[0] == "property":
            elements[-1][2].append(t[1:])
    if fmt not in ("ascii", "binary_little_endian"):
        raise NotImplementedError(f"PLY format '{fmt}'")
    body = data[end:]
    words = body.split() if fmt == "ascii" else None
    off, wi = 0, 0
    verts, faces = {}, []
    for name, count, props in elements:
        if fmt == "binary_little_endian" and all(p[0] != "list" for p in props):
            dt = np.dtype([(p[1], _PLY_TYPES[p[0]]) for p in props])
            arr = np.frombuffer(body, dt, count, off)
            off += dt.itemsize * count
            if name == "vertex":
                verts = {p[1]: arr[p[1]].astype(np.float64) for p in props}
            continue
        rows = []
        for _ in range(count):
            row = {}
            for p in props:
                if p[0] == "list":
                    if fmt == "ascii":
                        k = int(words[wi])
                        row[p[3]] = [int(x) for x in words[wi + 1:wi + 1 + k]]
                        wi += 1 + k
                    else:
                        ct, it = np.dtype(_PLY_TYPES[p[1]]), np.dtype(_PLY_TYPES[p[2]])
                        k = int(np.frombuffer(body, ct, 1, off)[0])
                        off += ct.itemsize
                        row[p[3]] = np.frombuffer(body, it, k, off).astype(np.int64).tolist()
                        off += it.itemsize * k
                elif fmt == "ascii":
                    row[p[1]] = float(words[wi])
                    wi += 1
                else:
                    dt = np.dtype(_PLY_TYPES[p[0]])
                    row[p[1]] = float(np.frombuffer(body, dt, 1, off)[0])
                    off += dt.itemsize
            rows.append(row)
        if name == "vertex":
            verts = {p[1]: np.array([r[p[1]] for r in rows]) for p in props}
        elif name == "face":
            key = "vertex_indices" if rows and "vertex_indices" in rows[0] else "vertex_index"
            for r in rows:
                f = r[key]
                for k in range(1, len(f) - 1):
                    faces.append([f[0], f[k], f[k + 1]])
    pos = np.stack([verts["x"], verts["y"], verts["z"]], 1)
    nrm = np.stack([verts["nx"], verts["ny"], verts["nz"]], 1) if "nx" in verts else None
    idx = np.array(faces, np.int64).reshape(-1, 3)
    if idx.size and (idx.min() < 0 or idx.max() >= len(pos)):
        raise ValueError(f"{path}: face index out of range")
    return pos, nrm, idx


__all__ = ["create_quad", "create_disk", "create_cube", "create_sphere", "apply_transform", "flip_normals", "bake",
           "smooth_normals", "face_normal_mesh", "load_ply", "translate", "Y_TO_Z"]
