"""The oracle's forest (one kd-tree per mesh, built level by level together)
pinned by the single-mesh oracle: each mesh's tree, renumbered, is the tree a
build of that mesh alone makes, its traces are that tree's traces, and the
per-level totals are the sums over the meshes."""
import numpy as np
import pytest

import forest_util as fu
from forest_util import c4_small


def mesh_tris(tris, off, m):
    return tuple(t[off[m]:off[m + 1]] for t in tris)


@pytest.mark.parametrize("preset", ["reference", "bitmap", "config"])
def test_forest_trees_are_the_per_mesh_trees(orc, scenes, preset):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, empty=(0, 4, 11))
    p = orc.build_params(preset)
    forest = orc.build_batched(tris, off, p)
    a = forest.arrays()
    assert a["num_roots"] == len(off) - 1
    singles = [orc.build(mesh_tris(tris, off, m), p) for m in range(len(off) - 1)]
    for m, S in enumerate(singles):
        fu.assert_same_mesh_tree(fu.restrict(a, m, off), S.arrays(), f"{preset} mesh {m}")
    # nothing but the meshes' nodes, slots and upper leaves
    assert a["info"].shape[0] == sum(S.num_nodes for S in singles)
    stats = [S.stats() for S in singles]
    fs = forest.stats()
    assert fs["num_prim_indices"] == sum(s["num_prim_indices"] for s in stats)
    assert fs["num_upper_leaves"] == sum(s["num_upper_leaves"] for s in stats)
    assert fs["reachable_nodes"] == fs["num_nodes"]
    assert fs["upper_levels"] == max(s["upper_levels"] for s in stats)
    assert fs["lower_levels"] == max(s["lower_levels"] for s in stats)
    assert fs["max_depth"] == max(s["max_depth"] for s in stats)
    for lvl, refs in enumerate(fs["upper_level_refs"]):
        assert refs == sum(s["upper_level_refs"][lvl] for s in stats if lvl < s["upper_levels"])


@pytest.mark.parametrize("algo", ["KD_RESTART", "SHORT_STACK"])
def test_forest_traces_are_the_per_mesh_traces(orc, scenes, algo):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=6, empty=(2,))
    p = orc.build_params("reference")
    forest = orc.build_batched(tris, off, p)
    code = getattr(orc, algo)
    h = orc.trace_batched(forest, tris, org, dirs, ray_mesh, code, stack_size=8)
    for m in range(len(off) - 1):
        rs = np.nonzero(ray_mesh == m)[0]
        if m == 2:
            assert rs.size == 0
            continue
        sub = mesh_tris(tris, off, m)
        ref = orc.trace(orc.build(sub, p), sub, org[rs], dirs[rs], code, stack_size=8)
        got = h[rs].copy()
        got["id"][got["id"] >= 0] -= off[m]
        assert np.array_equal(got.view(np.int32), ref.view(np.int32)), m
        assert (ref["id"] >= 0).mean() > 0.95      # rays aim into the unit ball
    # a ray whose mesh index is not one of the forest's misses
    bad = np.array([-1, len(off) - 1, 10 ** 6], np.int32)
    hb = orc.trace_batched(forest, tris, org[:3], dirs[:3], bad, code)
    assert (hb["id"] == -1).all() and np.isinf(hb["t"]).all()


def test_single_mesh_forest_is_the_plain_build(orc, scenes):
    tris, off, org, dirs, _ = c4_small(scenes, num_meshes=1)
    for preset in ("reference", "bitmap"):
        p = orc.build_params(preset)
        a = orc.build_batched(tris, [0, tris[0].shape[0]], p).arrays()
        b = orc.build(tris, p).arrays()
        for k in fu.TREE_KEYS:
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (preset, k)


def test_forest_argument_errors(orc, scenes):
    tris, off, *_ = c4_small(scenes, num_meshes=3)
    n = tris[0].shape[0]
    for bad in ([0, 5, 3, n], [1, 5, 10, n], [0, 5, 10, n - 1], []):
        with pytest.raises(RuntimeError):
            orc.build_batched(tris, np.asarray(bad, np.int32), orc.build_params())
