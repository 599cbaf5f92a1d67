"""Per-mesh view of a forest (one kd-tree per mesh in one node array).

A batched build keeps each mesh's tree exactly as a build of that mesh
alone would make it, except for numbering: the forest interleaves the
meshes' nodes level by level and their leaf slots leaf by leaf.  Both
renumberings keep order, so a mesh's tree is recovered by ranking its
nodes (the nodes reachable from its root) and its slots (the slots whose
triangle belongs to the mesh).

One field is not recoverable this way and is left out of the per-mesh
comparison: the primitive_index of an inner node of the upper stage, which
is the node's offset into its level's reference array, an array the meshes
share.  The forest-level comparison (GPU forest against the oracle forest)
covers it bit for bit.
"""
import numpy as np

TREE_KEYS = ("info", "split_pos", "aabb_min", "aabb_max", "primitive_index", "primitive_bitmap",
             "children", "surface_area", "parent", "prim_indices", "leaf_ids")


def reachable(a: dict, roots) -> np.ndarray:
    """ascending ids of the nodes reachable from ``roots`` (an id or ids)"""
    info, ch = a["info"], a["children"]
    seen = np.zeros(info.shape[0], dtype=bool)
    frontier = np.unique(np.asarray(roots, dtype=np.int64).reshape(-1))
    while frontier.size:
        frontier = frontier[~seen[frontier]]
        seen[frontier] = True
        inner = frontier[(info[frontier] & 3) != 0]
        frontier = np.unique(ch[inner].reshape(-1))
    return np.nonzero(seen)[0]


def restrict(a: dict, m: int, tri_off) -> dict:
    """tree of mesh ``m`` out of forest arrays ``a`` (keys as TREE_KEYS),
    renumbered as a build of mesh m alone numbers it; triangle ids local to
    the mesh.  Adds ``upper_inner`` (bool per node) so callers can mask the
    one field that is not comparable."""
    lo, hi = int(tri_off[m]), int(tri_off[m + 1])
    nodes = reachable(a, m)
    n = a["info"].shape[0]
    local = np.full(n + 1, -1, dtype=np.int64)
    local[nodes] = np.arange(nodes.size)
    pi_all = a["prim_indices"]
    mine = (pi_all >= lo) & (pi_all < hi)
    before = np.concatenate([[0], np.cumsum(mine)])        # mesh slots before slot p
    info = a["info"][nodes]
    inner = (info & 3) != 0
    # upper leaves and everything below them belong to the lower stage
    upper_leaves = a["leaf_ids"][local[a["leaf_ids"]] >= 0]
    lower = np.zeros(n, dtype=bool)
    lower[reachable(a, upper_leaves)] = True
    upper_inner = inner & ~lower[nodes]

    out = {}
    out["info"] = info
    for k in ("split_pos", "aabb_min", "aabb_max", "primitive_bitmap", "surface_area"):
        out[k] = a[k][nodes]
    ch = a["children"][nodes].astype(np.int64)
    ch[inner] = local[ch[inner]]
    assert (ch[inner] >= 0).all(), "a child left the mesh's tree"
    out["children"] = ch.astype(np.int32)
    p = a["primitive_index"][nodes].astype(np.int64)
    p_local = before[np.clip(p, 0, pi_all.shape[0])]
    out["primitive_index"] = np.where(upper_inner, p, p_local).astype(np.int32)
    par = a["parent"][nodes].astype(np.int64)
    lp = local[par]
    # nodes whose parent field was never set (chain nodes) hold 0
    stray = lp < 0
    assert (par[stray] == 0).all(), "a parent outside the mesh's tree"
    lp[stray] = 0
    out["parent"] = lp.astype(np.int32)
    out["prim_indices"] = (pi_all[mine] - lo).astype(np.int32)
    out["leaf_ids"] = local[upper_leaves].astype(np.int32)
    out["upper_inner"] = upper_inner
    return out


def assert_same_mesh_tree(got: dict, ref: dict, label=""):
    """compare a restricted forest tree with a single-mesh tree"""
    mask = got["upper_inner"]
    for k in TREE_KEYS:
        g, r = np.asarray(got[k]), np.asarray(ref[k])
        assert g.shape == r.shape, (label, k, g.shape, r.shape)
        if g.dtype.kind == "f":
            g, r = g.view(np.int32), r.view(np.int32)
        if k == "primitive_index":
            g, r = g[~mask], r[~mask]
        if not np.array_equal(g, r):
            bad = np.nonzero((g != r).reshape(g.shape[0], -1).any(axis=1))[0]
            raise AssertionError(f"{label} {k}: {len(bad)} rows differ, first {bad[:5]}")


def c4_small(scenes, num_meshes=9, rays=1500, hi=1.5e4, frame=2.0, empty=()):
    c = scenes.config4(num_meshes=num_meshes, rays_per_mesh=rays, hi=hi)
    V = scenes.displace_sphere64(c["unit"], frame)
    F = c["faces"]
    tris = list(scenes.tris_from(V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]))
    # splice in empty meshes (no triangles) so that they end up at the given
    # mesh indices
    counts = list(np.diff(c["tri_off"]))
    label = list(range(len(counts)))
    for pos in sorted(empty):
        counts.insert(pos, 0)
        label.insert(pos, -1)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    new_index = np.array([i for i, l in enumerate(label) if l >= 0], np.int32)
    ray_mesh = new_index[c["ray_mesh"]]
    return tuple(tris), off, c["org"], c["dirs"], ray_mesh
