"""The two BVH builders (host binned SAH, device PLOC): the trees they leave in a scene are checked
for structure (every triangle in exactly one leaf, every box enclosing what lies below it, every
node reached once, the depth within the traversal stack), and the device builder's tree must give
bit-identical ray casts (brute force on small scenes, the oracle's BVH at full size) and a
bit-identical render to the host builder's tree. The structure checker is itself checked on the
host builder's tree and on planted mistakes (CPU)."""
import numpy as np
import pytest

import oracle_lib as O
from test_pathtracer import _path_len
from test_scene_trace import _staircase, edge_rays, random_rays

MAX_LEVELS = 33  # 3 (levels - 1) traversal stack entries <= 32 in LDS + 64 spilled


def check_tree(nodes, tris, pos, idx):
    """Structure of a 4-wide tree in the device layout (nodes [n, 32] float32, tris [nt, 3, 4]).
    Returns the number of levels."""
    nt = len(idx)
    assert tris.shape == (nt, 3, 4)
    ids = tris[:, 0, 3].view(np.uint32)
    assert np.array_equal(np.sort(ids), np.arange(nt, dtype=np.uint32)), "leaf order is not a permutation"
    assert np.array_equal(tris[:, :, :3], pos[idx[ids]]), "triangle records differ from the triangles"
    assert np.all(tris[:, 1:, 3].view(np.uint32) == 0)
    box = nodes[:, :24].reshape(-1, 6, 4)  # [node, (lo x, hi x, lo y, hi y, lo z, hi z), child]
    lo = box[:, [0, 2, 4], :]  # [node, axis, child]
    hi = box[:, [1, 3, 5], :]
    refs = nodes[:, 24:28].view(np.int32)
    visits = np.zeros(len(nodes), np.int64)
    places = np.zeros(nt, np.int64)
    frontier = np.array([0])
    levels = 0
    while frontier.size:
        levels += 1
        assert levels <= MAX_LEVELS, "deeper than the traversal stack holds"
        np.add.at(visits, frontier, 1)
        r = refs[frontier]
        used = ~np.isinf(lo[frontier, 0, :])
        assert np.all(r[~used] == 0), "unused slot with a child"
        assert np.all(np.isinf(lo[frontier][~np.broadcast_to(used[:, None, :], lo[frontier].shape)])), "partial empty box"
        pi, slot = np.nonzero(used & (r >= 0))
        child = r[pi, slot]
        assert np.all(child < len(nodes))
        plo, phi = lo[frontier[pi], :, slot], hi[frontier[pi], :, slot]
        cused = ~np.isinf(lo[child, 0, :])
        assert np.all(cused.any(1)), "inner node without children"
        clo = np.where(cused[:, None, :], lo[child], np.inf).min(2)
        chi = np.where(cused[:, None, :], hi[child], -np.inf).max(2)
        assert np.all(clo >= plo) and np.all(chi <= phi), "child boxes outside the parent's box"
        li, ls = np.nonzero(used & (r < 0))
        leaf = ~r[li, ls]
        start, cnt = leaf >> 3, (leaf & 7) + 1
        llo, lhi = lo[frontier[li], :, ls], hi[frontier[li], :, ls]
        for k in range(8):
            m = cnt > k
            p = start[m] + k
            assert np.all(p < nt)
            np.add.at(places, p, 1)
            v = tris[p, :, :3]  # [m, 3 vertices, 3]
            assert np.all(v >= llo[m][:, None, :]) and np.all(v <= lhi[m][:, None, :]), "triangle outside its leaf box"
        frontier = child
    assert np.all(visits == 1), "nodes unreachable or reached twice"
    assert np.all(places == 1), "leaf-order places not covered exactly once"
    return levels


def _host_tree(scene):
    nodes, tris = scene.bvh_arrays()
    a = scene.numpy()
    return nodes, tris, a["positions"], a["indices"]


def test_checker_accepts_the_host_tree_and_finds_planted_mistakes():
    from pltfalcor_amd.scene import Scene
    s = Scene(None, seed=3, sphere_count=6, subdivision=2)
    nodes, tris, pos, idx = _host_tree(s)
    info = s.bvh_info()
    assert info["builder"] == "host" and info["n_nodes"] == len(nodes) and info["iterations"] == 0
    assert check_tree(nodes, tris, pos, idx) == info["max_depth"]
    # a child box shrunk below its subtree
    bad = nodes.copy()
    inner = np.nonzero((bad[0, 24:28].view(np.int32) > 0))[0][0]
    bad[0, 1 * 4 + inner] = bad[0, 0 * 4 + inner]  # hi x := lo x
    with pytest.raises(AssertionError):
        check_tree(bad, tris, pos, idx)
    # two leaves sharing triangles (a leaf reference copied over another)
    bad = nodes.copy()
    refs = bad[:, 24:28].view(np.int32)
    n, c = np.argwhere(refs < 0)[:2].T
    refs[n[1], c[1]] = refs[n[0], c[0]]
    with pytest.raises(AssertionError):
        check_tree(bad, tris, pos, idx)
    # a triangle record with another triangle's vertices
    badt = tris.copy()
    badt[5, 1, :3] = badt[6, 1, :3]
    with pytest.raises(AssertionError):
        check_tree(nodes, badt, pos, idx)


def test_host_rebuild_is_deterministic():
    """plt_scene_build_bvh with the host builder on a host-only scene rebuilds the same tree (the
    binned SAH is deterministic, its parallel subtrees included) and reports it."""
    from pltfalcor_amd.scene import Scene
    s = Scene(None, seed=4, sphere_count=40, subdivision=3)  # > 32k triangles: parallel subtrees
    nodes0, tris0 = s.bvh_arrays()
    info = s.build_bvh("host")
    nodes1, tris1 = s.bvh_arrays()
    assert info["builder"] == "host" and info["n_nodes"] == len(nodes1) == len(nodes0)
    assert np.array_equal(nodes0.view(np.uint32), nodes1.view(np.uint32)) and np.array_equal(tris0, tris1)
    a = s.numpy()
    assert check_tree(nodes1, tris1, a["positions"], a["indices"]) == info["max_depth"]


def test_device_builder_needs_a_context():
    from pltfalcor_amd import _lib
    from pltfalcor_amd.scene import Scene
    s = Scene(None, seed=3, sphere_count=1, subdivision=0)
    a = s.numpy()
    with pytest.raises(_lib.PLTError):
        Scene.from_arrays(None, a["positions"], a["normals"], a["indices"], a["tri_mesh"], a["mesh_material"],
                          a["materials"], a["profiles"], s.arrays.camera, bvh="device")
    with pytest.raises(_lib.PLTError):
        s.build_bvh("device")
    with pytest.raises(ValueError):
        s.build_bvh("lbvh")


def _soup(n, seed, spread=10.0, size=0.5):
    """n random triangles in a box, plus the arrays a scene needs."""
    from pltfalcor_amd.scene import Scene
    base = Scene(None, seed=3, sphere_count=1, subdivision=0)
    a = base.numpy()
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3)).astype(np.float32)
    c[..., 1] = np.abs(c[..., 1])
    pos = (c + rng.uniform(-size, size, (n, 3, 3))).astype(np.float32).reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (3 * n, 1))
    return a, base, pos, nrm, idx


def _from(ctx, a, base, pos, nrm, idx, bvh):
    from pltfalcor_amd.scene import Scene
    return Scene.from_arrays(ctx, pos, nrm, idx, np.zeros(len(idx), np.uint32), np.zeros(1, np.uint32), a["materials"],
                             a["profiles"], base.arrays.camera, bvh=bvh)


def _trace_both(s, pos, idx, rays, brute=True):
    import torch
    g = s.trace(torch.from_numpy(rays).cuda()).cpu().numpy().view(np.uint32)
    ref = O.trace(pos, idx, rays, brute=brute)
    assert np.array_equal(g[0], ref[0]), f"{np.sum(g[0] != ref[0])} hit ids differ"
    assert np.array_equal(g, ref), "barycentrics / distances differ"
    ga = s.trace(torch.from_numpy(rays).cuda(), any_hit=True).cpu().numpy()
    assert np.array_equal(ga.astype(np.uint32), O.trace(pos, idx, rays, any_hit=True, brute=brute))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 3, 4, 5, 17, 1024, 1025, 3000])
def test_gpu_device_tree_small_soups(gpu, n):
    """Triangle counts around the leaf size and the host-built top (from 1024 clusters): the
    device tree is well formed and casts rays bit-exactly like brute force."""
    from pltfalcor_amd.bsdf import Context
    print(f"device BVH soup n={n}", flush=True)
    ctx = Context(0)
    a, base, pos, nrm, idx = _soup(n, seed=n)
    s = _from(ctx, a, base, pos, nrm, idx, "device")
    info = s.bvh_info()
    assert info["builder"] == "device"
    nodes, tris = s.bvh_arrays()
    assert check_tree(nodes, tris, pos, idx) == info["max_depth"]
    assert info["n_nodes"] == len(nodes)
    _trace_both(s, pos, idx, random_rays(4000, n, lo=-12, hi=12))


@pytest.mark.gpu
def test_gpu_device_tree_duplicate_and_degenerate_triangles(gpu):
    """Equal Morton keys (many copies of a few triangles, all centroids in one cell) and zero-area
    triangles: the clustering still terminates, the tree is well formed, and among equal distances
    the smallest triangle id still wins."""
    from pltfalcor_amd.bsdf import Context
    print("device BVH duplicates", flush=True)
    ctx = Context(0)
    a, base, pos, nrm, idx = _soup(5, seed=1)
    reps = 700
    pos = np.tile(pos.reshape(5, 3, 3), (reps, 1, 1)).reshape(-1, 3)
    pos[: 3 * 40] = pos[:1]  # 40 triangles collapsed to a point
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (len(pos), 1))
    s = _from(ctx, a, base, pos, nrm, idx, "device")
    nodes, tris = s.bvh_arrays()
    check_tree(nodes, tris, pos, idx)
    _trace_both(s, pos, idx, random_rays(3000, 2, lo=-12, hi=12))


@pytest.mark.gpu
def test_gpu_device_tree_procedural_scenes(gpu):
    """The device tree of a small procedural scene casts like brute force, and the 64-sphere 1.31M
    triangle scene's like the oracle's BVH; rebuilding gives the same leaf order and node count
    (the build is deterministic up to the numbering of nodes within a level)."""
    import torch
    from pltfalcor_amd.bsdf import Context
    from pltfalcor_amd.scene import Scene
    ctx = Context(0)
    print("device BVH procedural small", flush=True)
    s = Scene(ctx, seed=3, sphere_count=6, subdivision=2)
    a = s.numpy()
    host_hits = s.trace(torch.from_numpy(random_rays(20000, 2)).cuda()).cpu().numpy()
    info = s.build_bvh("device")
    assert info["builder"] == "device" and s.arrays.n_bvh_nodes == info["n_nodes"]
    assert s.arrays.bvh_max_depth == info["max_depth"]
    nodes, tris = s.bvh_arrays()
    assert check_tree(nodes, tris, a["positions"], a["indices"]) == info["max_depth"]
    assert np.array_equal(s.trace(torch.from_numpy(random_rays(20000, 2)).cuda()).cpu().numpy(), host_hits)
    for rays in (random_rays(20000, 2), edge_rays(a), random_rays(5000, 3, tmax=4.0)):
        _trace_both(s, a["positions"], a["indices"], rays)
    # back to the host tree
    assert s.build_bvh("host")["builder"] == "host"
    _trace_both(s, a["positions"], a["indices"], random_rays(5000, 4))

    print("device BVH procedural 1.31M", flush=True)
    big = Scene(ctx, seed=1, sphere_count=64, subdivision=5)
    b = big.numpy()
    info = big.build_bvh("device")
    nodes, tris = big.bvh_arrays()
    assert check_tree(nodes, tris, b["positions"], b["indices"]) == info["max_depth"]
    info2 = big.build_bvh("device")
    nodes2, tris2 = big.bvh_arrays()
    assert np.array_equal(tris, tris2) and info2["n_nodes"] == info["n_nodes"]
    assert info2["max_depth"] == info["max_depth"]
    _trace_both(big, b["positions"], b["indices"], random_rays(1 << 16, 4), brute=False)


@pytest.mark.gpu
def test_gpu_device_tree_on_the_staircase(gpu):
    """The geometric staircase (a chain for any greedy clustering): the device builder either
    builds a tree within the traversal stack or refuses (PLT_EINVAL) and the scene keeps its
    tree; scene creation with the device builder then falls back to the host builder. Casts stay
    bit-exact either way."""
    from pltfalcor_amd import _lib
    from pltfalcor_amd.bsdf import Context
    print("device BVH staircase", flush=True)
    ctx = Context(0)
    a, base, pos, nrm, idx = _staircase()
    s = _from(ctx, a, base, pos, nrm, idx, "host")
    try:
        info = s.build_bvh("device")
        nodes, tris = s.bvh_arrays()
        assert check_tree(nodes, tris, pos, idx) == info["max_depth"] <= MAX_LEVELS
    except _lib.PLTError as e:
        assert "levels" in str(e)
        assert s.bvh_info()["builder"] == "host"
    _trace_both(s, pos, idx, random_rays(4096, 12, lo=-1e3, hi=1e3))
    t = _from(ctx, a, base, pos, nrm, idx, "device")
    nodes, tris = t.bvh_arrays()
    assert check_tree(nodes, tris, pos, idx) == t.bvh_info()["max_depth"]
    _trace_both(t, pos, idx, random_rays(4096, 13, lo=-1e3, hi=1e3))


@pytest.mark.gpu
def test_gpu_render_identical_with_either_tree(gpu):
    """A PLTPT frame (bounce records, colour, albedo, normal) is bit-identical whichever builder
    made the scene's tree."""
    import torch
    from pltfalcor_amd.bsdf import default_context
    from pltfalcor_amd.pltpt import PLTPT
    from pltfalcor_amd.scene import Scene
    print("device BVH render", flush=True)
    ctx = default_context()
    s = Scene(ctx, seed=3, sphere_count=24, subdivision=3, material_kinds=8)
    w, h = 96, 64
    outs = []
    for builder in ("host", "device"):
        s.build_bvh(builder)
        p = PLTPT(ctx, {"maxBounces": 6, "doMNEE": 1})
        p.set_scene(s)
        p.set_frame(2)
        color = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        albedo, normal = torch.empty_like(color), torch.empty_like(color)
        p.execute(w, h, color, albedo, normal)
        torch.cuda.synchronize()
        outs.append((p.bounce_records(w, h), color.cpu().numpy(), albedo.cpu().numpy(), normal.cpu().numpy()))
        p.close()
    (r0, c0, a0, n0), (r1, c1, a1, n1) = outs
    L = _path_len(r0)
    assert np.array_equal(L, _path_len(r1))
    valid = np.arange(r0.shape[0])[:, None] < L[None, :]  # records past a path's end are not written
    assert np.array_equal(r0.transpose(0, 2, 1)[valid], r1.transpose(0, 2, 1)[valid])
    assert np.array_equal(c0, c1) and np.array_equal(a0, a1) and np.array_equal(n0, n1)
