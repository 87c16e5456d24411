"""Scene construction (CPU) and GPU ray casting: hit triangle ids, barycentrics and distances
bit-exact against the oracle (brute force on small scenes, the oracle's own BVH at full size)."""
import numpy as np
import pytest

import oracle_lib as O


def random_rays(n, seed, lo=-10, hi=10, tmax=3.4e38):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(0.05, 8, n)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([o.T, d.T, np.zeros((1, n), np.float32), np.full((1, n), tmax, np.float32)]), dtype=np.float32)


def edge_rays(a):
    """Rays aimed exactly at mesh vertices and edge midpoints (watertightness cases)."""
    pos, idx = a["positions"], a["indices"]
    tgt = [pos[idx[:200, 0]], 0.5 * (pos[idx[:200, 0]] + pos[idx[:200, 1]])]
    # the ground quad's diagonal and border
    t = np.linspace(-9.5, 9.5, 64, dtype=np.float32)
    tgt.append(np.stack([t, np.zeros_like(t), t], 1))
    tgt = np.concatenate(tgt).astype(np.float32)
    o = tgt + np.array([0.3, 9.0, 0.7], np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = len(tgt)
    return np.ascontiguousarray(np.concatenate([o.T, d.T, np.zeros((1, n), np.float32), np.full((1, n), 3.4e38, np.float32)]), dtype=np.float32)


def test_procedural_scene_counts_and_bvh():
    from pltfalcor_amd.scene import Scene
    s = Scene(None, seed=1, sphere_count=8, subdivision=3)
    a = s.numpy()
    assert s.arrays.n_triangles == 2 + 8 * 20 * 4 ** 3 + 2
    assert s.arrays.n_meshes == 10 and s.arrays.n_materials == 10
    assert a["indices"].max() < s.arrays.n_vertices
    # vertex normals are unit within binary16 rounding and outward
    n = a["normals"]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=2e-3)
    # face winding agrees with the vertex normals (outward)
    p = a["positions"]
    i = a["indices"]
    fn = np.cross(p[i[:, 1]] - p[i[:, 0]], p[i[:, 2]] - p[i[:, 0]])
    assert np.all(np.einsum("ij,ij->i", fn, n[i[:, 0]]) > 0)
    # deterministic
    s2 = Scene(None, seed=1, sphere_count=8, subdivision=3)
    assert np.array_equal(s2.numpy()["positions"], a["positions"])


def test_oracle_bvh_matches_brute_force():
    from pltfalcor_amd.scene import Scene
    a = Scene(None, seed=3, sphere_count=6, subdivision=2).numpy()
    for rays in (random_rays(4000, 1), edge_rays(a)):
        hb = O.trace(a["positions"], a["indices"], rays, brute=True)
        hv = O.trace(a["positions"], a["indices"], rays)
        assert np.array_equal(hb, hv)
        assert np.array_equal(O.trace(a["positions"], a["indices"], rays, any_hit=True, brute=True),
                              O.trace(a["positions"], a["indices"], rays, any_hit=True))


def test_watertight_ground_diagonal():
    """Rays through the shared diagonal edge of the ground quad never leak."""
    from pltfalcor_amd.scene import Scene
    a = Scene(None, seed=3, sphere_count=0, subdivision=0).numpy()
    t = np.linspace(-9.9, 9.9, 2001, dtype=np.float32)
    n = len(t)
    o = np.stack([t, np.full(n, 5.0, np.float32), t], 1)
    d = np.tile(np.array([0, -1, 0], np.float32), (n, 1))
    rays = np.ascontiguousarray(np.concatenate([o.T, d.T, np.zeros((1, n), np.float32), np.full((1, n), 3.4e38, np.float32)]), dtype=np.float32)
    h = O.trace(a["positions"], a["indices"], rays, brute=True)
    assert np.all(h[0] < 2)  # one of the two ground triangles
    assert np.allclose(h[3].view(np.float32), 5.0)


def _gpu_scene(**kw):
    from pltfalcor_amd.bsdf import Context
    from pltfalcor_amd.scene import Scene
    ctx = Context(0)
    return ctx, Scene(ctx, **kw)


@pytest.mark.gpu
def test_gpu_trace_bit_exact_small_scene(gpu):
    import torch
    ctx, s = _gpu_scene(seed=3, sphere_count=6, subdivision=2)
    a = s.numpy()
    for rays in (random_rays(20000, 2), edge_rays(a), random_rays(5000, 3, tmax=4.0)):
        g = s.trace(torch.from_numpy(rays).cuda()).cpu().numpy().view(np.uint32)
        ref = O.trace(a["positions"], a["indices"], rays, brute=True)
        assert np.array_equal(g[0], ref[0]), f"{np.sum(g[0] != ref[0])} hit ids differ"
        assert np.array_equal(g, ref), "barycentrics / distances differ"
        ga = s.trace(torch.from_numpy(rays).cuda(), any_hit=True).cpu().numpy()
        ra = O.trace(a["positions"], a["indices"], rays, any_hit=True, brute=True)
        assert np.array_equal(ga.astype(np.uint32), ra)


@pytest.mark.gpu
def test_gpu_trace_bit_exact_path_scene(gpu):
    """The 64-sphere, 1.31M-triangle scene: 2^18 random rays + 2^16 camera-like rays."""
    import torch
    ctx, s = _gpu_scene(seed=1, sphere_count=64, subdivision=5)
    a = s.numpy()
    rays = random_rays(1 << 18, 4)
    rng = np.random.default_rng(5)
    n = 1 << 16
    cam = np.array([0, 5, 16], np.float32)
    tgt = np.stack([rng.uniform(-10, 10, n), rng.uniform(0, 2, n), rng.uniform(-10, 10, n)], 1).astype(np.float32)
    d = tgt - cam
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    crays = np.ascontiguousarray(np.concatenate([np.tile(cam, (n, 1)).T, d.T, np.zeros((1, n), np.float32),
                                                 np.full((1, n), 3.4e38, np.float32)]), dtype=np.float32)
    for rr in (rays, crays):
        g = s.trace(torch.from_numpy(rr).cuda()).cpu().numpy().view(np.uint32)
        ref = O.trace(a["positions"], a["indices"], rr)
        assert np.array_equal(g[0], ref[0]), f"{np.sum(g[0] != ref[0])} hit ids differ"
        assert np.array_equal(g, ref)
        assert (g[0] != 0xFFFFFFFF).mean() > 0.3


def test_scene_from_arrays_matches_procedural():
    """plt_scene_create from caller arrays reproduces the procedural scene (same arrays, and the
    oracle's intersections with rays agree for both), and rejects out-of-range ids."""
    from pltfalcor_amd import _lib
    from pltfalcor_amd.scene import Scene
    s = Scene(None, seed=4, sphere_count=5, subdivision=2)
    a = s.numpy()
    t = Scene.from_arrays(None, a["positions"], a["normals"], a["indices"], a["tri_mesh"], a["mesh_material"],
                          a["materials"], a["profiles"], s.arrays.camera)
    b = t.numpy()
    for k in ("positions", "normals", "indices", "tri_mesh", "mesh_first_tri", "mesh_material", "materials",
              "profiles", "camera"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])) if not isinstance(a[k], bytes) else a[k] == b[k], k
    bad = a["indices"].copy()
    bad[3, 1] = a["positions"].shape[0]
    with pytest.raises(_lib.PLTError):
        Scene.from_arrays(None, a["positions"], a["normals"], bad, a["tri_mesh"], a["mesh_material"], a["materials"],
                          a["profiles"], s.arrays.camera)
    badm = a["mesh_material"].copy()
    badm[0] = 99
    with pytest.raises(_lib.PLTError):
        Scene.from_arrays(None, a["positions"], a["normals"], a["indices"], a["tri_mesh"], badm, a["materials"],
                          a["profiles"], s.arrays.camera)


def _staircase(n=300):
    """Triangles whose sizes and positions grow geometrically (x1.1 each): a binned SAH peels off
    one triangle per level, so an unbounded binary tree would be about n levels deep."""
    from pltfalcor_amd.scene import Scene
    base = Scene(None, seed=3, sphere_count=1, subdivision=0)
    a = base.numpy()
    s = (1.1 ** np.arange(n)).astype(np.float32)
    pos = np.zeros((3 * n, 3), np.float32)
    pos[0::3, 0] = s
    pos[1::3] = np.stack([s, 0.5 * s, np.zeros_like(s)], 1)
    pos[2::3] = np.stack([s, np.zeros_like(s), 0.5 * s], 1)
    nrm = np.tile(np.array([1, 0, 0], np.float32), (3 * n, 1))
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return a, base, pos, nrm, idx


def test_bvh_depth_is_bounded_on_a_geometric_staircase():
    """The builder's depth bound: the staircase builds (no depth refusal) into a 4-wide BVH of at
    most 32 levels (the bound the traversal stack is sized for)."""
    from pltfalcor_amd.scene import Scene
    a, base, pos, nrm, idx = _staircase()
    t = Scene.from_arrays(None, pos, nrm, idx, np.zeros(len(idx), np.uint32), np.zeros(1, np.uint32), a["materials"],
                          a["profiles"], base.arrays.camera)
    d = t.arrays.bvh_max_depth
    assert 8 <= d <= 32, d


@pytest.mark.gpu
def test_gpu_trace_bit_exact_on_a_deep_tree(gpu):
    """The staircase's deep tree on the GPU: rays along the stair (deep stacks, past the LDS part of
    the traversal stack) and random rays hit bit-exactly what brute force hits."""
    import torch
    from pltfalcor_amd.bsdf import Context
    from pltfalcor_amd.scene import Scene
    a, base, pos, nrm, idx = _staircase()
    ctx = Context(0)
    t = Scene.from_arrays(ctx, pos, nrm, idx, np.zeros(len(idx), np.uint32), np.zeros(1, np.uint32), a["materials"],
                          a["profiles"], base.arrays.camera)
    rng = np.random.default_rng(11)
    n = 4096
    s = pos[0::3, 0]
    tgt = np.stack([rng.choice(s, n) * rng.uniform(1.0, 1.001, n), rng.uniform(0, 0.2, n) * s[-1],
                    rng.uniform(0, 0.2, n) * s[-1]], 1).astype(np.float32)
    o = np.stack([np.full(n, -10.0), tgt[:, 1] * 0.5, tgt[:, 2] * 0.5], 1).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.concatenate([o.T, d.T, np.zeros((1, n), np.float32), np.full((1, n), 3.4e38, np.float32)]),
                                dtype=np.float32)
    for rr in (rays, random_rays(4096, 12, lo=-1e3, hi=1e3)):
        g = t.trace(torch.from_numpy(rr).cuda()).cpu().numpy().view(np.uint32)
        ref = O.trace(pos, idx, rr, brute=True)
        assert np.array_equal(g, ref), f"{np.sum(g[0] != ref[0])} hit ids differ"
        ga = t.trace(torch.from_numpy(rr).cuda(), any_hit=True).cpu().numpy()
        assert np.array_equal(ga.astype(np.uint32), O.trace(pos, idx, rr, any_hit=True, brute=True))
    assert (g[0] != 0xFFFFFFFF).mean() > 0.05
