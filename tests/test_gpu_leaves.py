"""Full leaves, ties and the wave-cooperative leaf test.

Dense clusters of overlapping triangles, some of them duplicated exactly, in
trees whose leaves hold up to 63 triangles (lower size 64): a wave's lanes
then bring thousands of (ray, triangle) pairs per leaf step, more than one
window of the cooperative test, and duplicated triangles give equal t, u, v
in one leaf, where the triangle later in the leaf's bitmap must win (the
sequential loop accepts t <= th).  Bar: bit-exact against the oracle --
hit records and per-ray counters -- for both tree tracers, in input and in
coherent ray order.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_same_hits, assert_same_tree, dev, orc_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "these tests need the GPU"
    return T


def dense_scene(seed=7, clusters=24, per=60, dups=12, rays=150000):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.1, 0.9, (clusters, 3))
    tris = []
    for c in centers:
        v = c + rng.uniform(-0.02, 0.02, (per, 3, 3))
        tris.append(v)
        tris.append(v[rng.choice(per, dups, replace=False)])     # exact duplicates
    T = np.concatenate(tris).astype(np.float32)
    P = []
    for k in range(3):
        p = np.zeros((T.shape[0], 4), np.float32)
        p[:, :3] = T[:, k]
        P.append(p)
    # rays from a sphere around the scene towards jittered cluster centres
    o = rng.normal(size=(rays, 3))
    o = 0.5 + 2.0 * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = centers[rng.integers(0, clusters, rays)] + rng.uniform(-0.015, 0.015, (rays, 3))
    d = tgt - o
    org = np.zeros((rays, 4), np.float32); org[:, :3] = o
    dirs = np.zeros((rays, 4), np.float32); dirs[:, :3] = d
    return tuple(P), org, dirs


@pytest.mark.parametrize("preset", ["reference", "default"])
def test_full_leaves_and_ties_bit_exact(pm, orc, torch, preset):
    tris, org, dirs = dense_scene()
    bp = pm.build_params(preset, lower_size=64)
    P = [dev(torch, t) for t in tris]
    ctx = pm.Context(0)
    ctx.build(*P, bp)
    ref = orc.build(tris, orc_params(orc, bp))
    assert_same_tree(ctx.tree_arrays(), ref.arrays(), preset)
    st = ref.stats()
    assert st["max_leaf_size"] >= 40, st["max_leaf_size"]
    O, D = dev(torch, org), dev(torch, dirs)
    for algo in (pm.SHORT_STACK, pm.KD_RESTART):
        want, wc = orc.trace(ref, tris, org, dirs, algo, stack_size=8, counters=True, threads=8)
        assert (want["id"] >= 0).mean() > 0.9
        # the leaves these rays visit hold > 8 triangles on average: a wave
        # step of 64 lanes lists more pairs than one 512-pair window
        assert wc[:, 2].sum() / max(wc[:, 1].sum(), 1) > 8
        for order in (pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT):
            cnt = torch.zeros((org.shape[0], 4), dtype=torch.int32, device="cuda:0")
            h = pm.hits_to_numpy(ctx.trace(algo, O, D, pm.trace_params(stack_size=8, ray_order=order),
                                           counters=cnt))
            assert_same_hits(h, want, f"{preset} algo {algo} order {order}")
            assert np.array_equal(cnt.cpu().numpy(), wc)
    ctx.close()


def test_duplicate_triangles_later_slot_wins(pm, orc, torch):
    """two exact copies of one triangle: every ray that hits it reports
    the copy that comes later in its leaf (the oracle's sequential rule)"""
    tri = np.array([[[0.2, 0.2, 0.5], [0.8, 0.2, 0.5], [0.5, 0.8, 0.5]]], np.float32)
    T = np.concatenate([tri, tri, tri + np.float32(0.001)])       # copies 0, 1 and a near one
    P = []
    for k in range(3):
        p = np.zeros((T.shape[0], 4), np.float32)
        p[:, :3] = T[:, k]
        P.append(p)
    rng = np.random.default_rng(3)
    n = 200000
    org = np.zeros((n, 4), np.float32)
    org[:, :2] = rng.uniform(0.3, 0.7, (n, 2)); org[:, 2] = 2.0
    dirs = np.zeros((n, 4), np.float32)
    dirs[:, :2] = rng.uniform(-0.01, 0.01, (n, 2)); dirs[:, 2] = -1.0
    ctx = pm.Context(0)
    Pd = [dev(torch, t) for t in P]
    for preset in ("reference", "default"):
        bp = pm.build_params(preset)
        ctx.build(*Pd, bp)
        ref = orc.build(tuple(P), orc_params(orc, bp))
        want = orc.trace(ref, tuple(P), org, dirs, orc.SHORT_STACK, stack_size=8, threads=8)
        for order in (pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT):
            h = pm.hits_to_numpy(ctx.trace(pm.SHORT_STACK, dev(torch, org), dev(torch, dirs),
                                           pm.trace_params(stack_size=8, ray_order=order)))
            assert_same_hits(h, want, f"{preset} {order}")
        # copies 0 and 1 tie in t, u and v, so copy 0 is never reported; the
        # raised copy (z + 0.001, nearer the origins at z = 2) wins where it is hit
        assert set(np.unique(want["id"])) <= {-1, 1, 2} and 1 in set(want["id"])
    ctx.close()
