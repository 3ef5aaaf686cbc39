"""Closest-hit parity: the HIP traversal (el_trace_closest) against the CPU
oracle on identical seeded ray batches.

Bar: the triangle index matches bit-exactly (minimum t, ties to the lower
index) and the hit distance is bit-identical (the same Moller-Trumbore
arithmetic, Tri.hpp:38-68, in FP32 without contraction).  The stated
tolerance for t is 1e-5 relative; identical bits are checked first and the
tolerance is asserted as the contract.
"""
import numpy as np
import pytest

import eleven
from eleven import scenegen
import oracle_py as orc
import raygen

T_RTOL = 1e-5


def _check(tri_g, t_g, tri_o, t_o, label):
    assert tri_g.shape == tri_o.shape
    bad = np.nonzero(tri_g != tri_o)[0]
    assert bad.size == 0, f"{label}: {bad.size} index mismatches, first {bad[:8]} gpu {tri_g[bad[:8]]} oracle {tri_o[bad[:8]]}"
    hit = tri_o >= 0
    assert np.all(np.isinf(t_g[~hit]))
    rel = np.abs(t_g[hit] - t_o[hit]) / np.maximum(np.abs(t_o[hit]), 1e-30)
    assert rel.max(initial=0) <= T_RTOL, f"{label}: t rel err {rel.max()}"
    return int(hit.sum()), int(np.sum(t_g[hit] != t_o[hit]))


@pytest.fixture(scope="module")
def field64(scene_root):
    d = scenegen.sphere_field(scene_root / "field64")
    return eleven.SceneData.load(d)


@pytest.mark.gpu
@pytest.mark.parametrize("width", [4, 8])
def test_closest_random_rays_vs_oracle(gpu_ctx, field64, width):
    print(f"[trace] sphere field, 2^20 random rays, BVH{width}", flush=True)
    gpu_ctx.upload(field64)
    gpu_ctx.render(0, bvh_width=width)          # switch width without rendering
    verts = field64.vertices()
    rays = raygen.random_rays(verts, 1 << 20, seed=1)
    tri_g, t_g, _ = gpu_ctx.trace_closest(rays)
    ref = orc.RefBVH(verts)
    tri_r, t_r = ref.closest(rays)
    # where the reference-faithful tree disagrees, brute force decides
    dis = np.nonzero(tri_r != tri_g)[0]
    if dis.size:
        tb, tt, _, _ = orc.closest_brute(verts, rays[dis])
        tri_r[dis], t_r[dis] = tb, tt
    nh, ndiff = _check(tri_g, t_g, tri_r, t_r, "random rays")
    assert nh > 1000
    assert ndiff == 0, f"{ndiff} hit distances not bit-identical"
    # exhaustive check on a slice against brute force over all triangles
    sub = rays[:8192]
    tb, tt, _, _ = orc.closest_brute(verts, sub)
    _check(tri_g[:8192], t_g[:8192], tb, tt, "brute-force slice")


@pytest.mark.gpu
def test_closest_camera_and_tie_rays(gpu_ctx, field64):
    print("[trace] sphere field camera rays (256x256x16) + 4096 tie rays", flush=True)
    gpu_ctx.upload(field64)
    verts = field64.vertices()
    cam = raygen.camera_rays(field64.camera, spp=16, seed=2)
    ties = raygen.tie_rays(verts, 4096, seed=3)
    rays = np.concatenate([cam, ties])
    tri_g, t_g, uv_g = gpu_ctx.trace_closest(rays)
    ref = orc.RefBVH(verts)
    tri_r, t_r = ref.closest(rays)
    dis = np.nonzero(tri_r != tri_g)[0]
    if dis.size:
        tb, tt, _, _ = orc.closest_brute(verts, rays[dis])
        tri_r[dis], t_r[dis] = tb, tt
    _check(tri_g, t_g, tri_r, t_r, "camera+tie rays")
    tb, tt, ub, vb = orc.closest_brute(verts, ties)
    _check(tri_g[-4096:], t_g[-4096:], tb, tt, "tie rays brute force")
    hit = tb >= 0
    assert np.array_equal(uv_g[-4096:][hit, 0], ub[hit]) and np.array_equal(uv_g[-4096:][hit, 1], vb[hit])


@pytest.mark.gpu
def test_any_hit_matches_closest(gpu_ctx, field64):
    print("[trace] any-hit vs closest", flush=True)
    gpu_ctx.upload(field64)
    verts = field64.vertices()
    rays = raygen.random_rays(verts, 1 << 16, seed=4)
    tri, t, _ = gpu_ctx.trace_closest(rays)
    rng = np.random.default_rng(5)
    tmax = rng.uniform(0, 40, size=rays.shape[0]).astype(np.float32)
    occ = gpu_ctx.trace_any(rays, tmax)
    expect = (tri >= 0) & (t < tmax)
    assert np.array_equal(occ, expect)


@pytest.mark.gpu
def test_random_triangle_soup_vs_brute_force(gpu_ctx):
    print("[trace] random triangle soups vs brute force", flush=True)
    rng = np.random.default_rng(11)
    for n, size in [(1, 1.0), (7, 0.5), (300, 0.3), (5000, 0.05)]:
        c = rng.uniform(-1, 1, size=(n, 1, 3))
        v = (c + rng.normal(scale=size, size=(n, 3, 3))).astype(np.float32).reshape(n, 9)
        sc = eleven.SceneData.from_arrays(v, camera=dict(x_res=4, y_res=4))
        gpu_ctx.upload(sc)
        rays = raygen.random_rays(v, 20000, seed=n)
        tri_g, t_g, _ = gpu_ctx.trace_closest(rays)
        tb, tt, _, _ = orc.closest_brute(v, rays)
        _check(tri_g, t_g, tb, tt, f"soup n={n}")


@pytest.mark.gpu
def test_duplicate_triangles_tie_to_lower_index(gpu_ctx):
    print("[trace] duplicated triangles", flush=True)
    rng = np.random.default_rng(12)
    base = rng.uniform(-1, 1, size=(50, 9)).astype(np.float32)
    v = np.concatenate([base, base, base])       # every triangle three times
    perm = rng.permutation(v.shape[0])
    v = v[perm]
    sc = eleven.SceneData.from_arrays(v, camera=dict(x_res=4, y_res=4))
    gpu_ctx.upload(sc)
    rays = raygen.random_rays(v, 20000, seed=13)
    tri_g, t_g, _ = gpu_ctx.trace_closest(rays)
    tb, tt, _, _ = orc.closest_brute(v, rays)
    _check(tri_g, t_g, tb, tt, "duplicates")
    hit = tb >= 0
    # the winner is the lowest index among the copies
    bid = perm % 50                              # base triangle of each entry
    for i in np.nonzero(hit)[0][:200]:
        copies = np.nonzero(bid == bid[tri_g[i]])[0]
        assert tri_g[i] == copies.min()


@pytest.mark.gpu
@pytest.mark.parametrize("width", [4, 8])
def test_tiny_distant_nodes_keep_every_subtree(gpu_ctx, width):
    """A cluster of tiny triangles with a larger triangle behind it, seen
    from 1000 units away on every axis (some 60 scene diagonals).  Node
    boxes are padded to 2^-15 of the diagonal, so the cluster's nodes are
    ~6e-4 across with grid steps ~4e-6: 255 steps are below the slab
    test's widening at that distance, which cannot reject an empty slot's
    inverted box there.  Such a slot must still never be entered, or a ray
    through the cluster stops before it reaches the triangle behind."""
    print(f"[trace] tiny cluster seen from far away, BVH{width}", flush=True)
    rng = np.random.default_rng(21)
    size = 3e-5
    n = 96
    c = rng.uniform(-size / 2, size / 2, size=(n, 1, 3))
    tiny = c + rng.normal(scale=size / 40, size=(n, 3, 3))
    # behind the cluster as seen from (-50, -50, -50): a triangle across (1, 1, 1)
    u = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    w = np.array([1.0, 1.0, -2.0]) / np.sqrt(6.0)
    back = np.array([10.0, 10.0, 10.0]) + 5.0 * np.stack([u, w, -(u + w)])
    v = np.concatenate([tiny.reshape(n, 9), back.reshape(1, 9)]).astype(np.float32)
    sc = eleven.SceneData.from_arrays(v, camera=dict(x_res=4, y_res=4))
    gpu_ctx.upload(sc)
    gpu_ctx.render(0, bvh_width=width)
    m = 1 << 15
    o = np.full((m, 3), -1000.0) + rng.uniform(-1e-2, 1e-2, size=(m, 3))
    target = rng.uniform(-size, size, size=(m, 3))
    rays = np.concatenate([o, target - o], -1).astype(np.float32)
    rays[:, 3:] = raygen._norm(rays[:, 3:])
    tri_g, t_g, _ = gpu_ctx.trace_closest(rays)
    tb, tt, _, _ = orc.closest_brute(v, rays)
    # the tiny triangles are below the intersection test's determinant
    # threshold: every ray passes the cluster and hits the triangle behind
    assert np.all(tb == n)
    _check(tri_g, t_g, tb, tt, f"tiny cluster BVH{width}")
    # any hit within 2000 (the triangle behind is ~1750 away): every ray
    occ = gpu_ctx.trace_any(rays, np.full(m, 2000.0, np.float32))
    assert np.all(tt < 2000.0) and np.all(occ)
