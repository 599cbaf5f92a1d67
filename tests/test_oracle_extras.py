"""Pins the oracle's optional paths: the DIVIDE split method (SAT step 3),
the Woop intersector, the balanced lower creator and the canvas shading,
each against mathematics that does not share its code: FP64 separating-axis
projections, FP64 Moller-Trumbore, exhaustive candidate enumeration in
numpy, hand-computed Phong colours, and hit equality with brute force."""
import numpy as np
import pytest

from test_oracle import check_against_brute, regions, tri_boxes


# ---------------------------------------------------------------- SAT (DIVIDE)
def sat_edge_axes64(v, bmin, bmax):
    """per (edge, axis) pair: (min, max) of the triangle's projection and the
    box radius on cross(axis, edge), in FP64, box-centred"""
    v = np.asarray(v, np.float64)
    h = (np.asarray(bmax, np.float64) - np.asarray(bmin, np.float64)) * 0.5
    c = np.asarray(bmin, np.float64) + h
    v = v - c
    out = []
    for j in range(3):
        a = np.zeros(3); a[j] = 1.0
        for k in range(3):
            e = v[(k + 1) % 3] - v[k]
            ax = np.cross(a, e)
            p = v @ ax
            r = np.abs(ax) @ h
            out.append((p.min(), p.max(), r))
    return out


def test_sat_step3_is_the_edge_axis_test(orc):
    rng = np.random.default_rng(3)
    inside = outside = 0
    for _ in range(4000):
        bmin = rng.uniform(-1, 0, 3)
        bmax = bmin + rng.uniform(0.05, 1.0, 3)
        c = rng.uniform(bmin - 0.4, bmax + 0.4)
        v = c + rng.normal(scale=rng.uniform(0.05, 0.8), size=(3, 3))
        v32 = v.astype(np.float32)
        got = orc.tri_box_step3(v32[0], v32[1], v32[2], bmin.astype(np.float32), bmax.astype(np.float32))
        axes = sat_edge_axes64(v32, bmin.astype(np.float32), bmax.astype(np.float32))
        if all(lo <= r and hi >= -r for lo, hi, r in axes):
            assert got, "no edge axis separates, yet the test says apart"
            inside += 1
        elif any(lo > (r + 1e-5) * 1.0001 + 1e-4 or hi < -((r + 1e-5) * 1.0001 + 1e-4) for lo, hi, r in axes):
            assert not got, "an edge axis separates by a clear margin, yet the test says overlap"
            outside += 1
    assert inside > 300 and outside > 300
    # a triangle inside the box, and one separated only along an edge axis
    assert orc.tri_box_step3([0.1, 0.1, 0.1], [0.3, 0.1, 0.1], [0.1, 0.3, 0.1], [0, 0, 0], [1, 1, 1])
    assert not orc.tri_box_step3([1.5, 0.6, 0.5], [0.6, 1.5, 0.5], [1.5, 1.5, 0.5], [0, 0, 0], [1, 1, 1])


def test_divide_tree_invariants_and_traces(orc, scenes):
    tris, (org, dirs) = scenes.config1(width=96, height=96)
    box = orc.build(tris, orc.build_params("reference")).stats()
    tree = orc.build(tris, orc.build_params("reference", split_method=orc.SPLIT_DIVIDE))
    a = tree.arrays()
    st = tree.stats()
    assert st["leaf_refs"] < box["leaf_refs"]          # straddlers only go where they reach
    tmin, tmax = tri_boxes(tris)
    seen = np.zeros(tris[0].shape[0], bool)
    for n, (lo, hi) in regions(a).items():
        if a["info"][n] & 3:
            continue
        bm = int(a["primitive_bitmap"][n]) & (2 ** 64 - 1)
        pi = a["primitive_index"][n]
        for i in range(64):
            if bm >> i & 1:
                t = a["prim_indices"][pi + i]
                seen[t] = True
                assert np.all(tmin[t] <= hi) and np.all(tmax[t] >= lo), (n, t)
    assert seen.all()
    brute = orc.trace(None, tris, org, dirs, orc.BRUTE)
    for algo in (orc.KD_RESTART, orc.SHORT_STACK):
        check_against_brute(a, orc.trace(tree, tris, org, dirs, algo), brute, org, dirs)
    with pytest.raises(RuntimeError):
        orc.build(tris, orc.build_params("reference", split_method=2))   # SPLIT: not implemented


# ---------------------------------------------------------------- Woop
def test_woop_matrix_closed_form(orc):
    m = orc.woop_matrix([1, 0, 0], [0, 1, 0], [0, 0, 0])
    assert np.array_equal(m, np.eye(3, 4, dtype=np.float32))
    # translated and scaled: the rows invert the edge matrix of (a-c, b-c, n')
    a, b, c = np.array([3.0, 1, 2]), np.array([1.0, 4, 2]), np.array([1.0, 1, 2])
    m = orc.woop_matrix(a, b, c).astype(np.float64)
    i0, i1 = a - c, b - c
    i2 = np.cross(i0, i1) - c
    Minv = np.linalg.inv(np.stack([i0, i1, i2], axis=1))
    assert np.allclose(m[:, :3], Minv, rtol=1e-6, atol=1e-7)
    assert np.allclose(m[:, 3], Minv @ c, rtol=1e-6, atol=1e-6)


def test_woop_hits_match_fp64_moller_trumbore(orc, scenes):
    rng = np.random.default_rng(11)
    n = 300
    tris = scenes.random_soup(n, seed=12, size=0.4)
    cent = (tris[0][:, :3] + tris[1][:, :3] + tris[2][:, :3]) / 3
    k = rng.integers(0, n, 2000)
    org = np.zeros((2000, 4), np.float32); org[:, :3] = rng.uniform(-1, 2, (2000, 3))
    d = cent[k] - org[:, :3]
    dirs = np.zeros((2000, 4), np.float32); dirs[:, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    w = orc.trace(None, tris, org, dirs, orc.BRUTE, intersection=orc.WOOP)
    m = orc.trace(None, tris, org, dirs, orc.BRUTE)
    both = (w["id"] >= 0) & (w["id"] == m["id"])
    assert both.mean() > 0.98
    for r in np.nonzero(both)[0][:400]:
        t = w["id"][r]
        ref = orc.mt64(tris[0][t, :3], tris[1][t, :3], tris[2][t, :3], org[r, :3], dirs[r, :3])
        assert abs(w["t"][r] - ref[0]) <= 1e-4 * max(1.0, abs(ref[0])), r
        assert abs(w["u"][r] - ref[1]) <= 1e-3 and abs(w["v"][r] - ref[2]) <= 1e-3, r


def test_woop_tree_traces_match_woop_brute(orc, scenes):
    tris, (org, dirs) = scenes.config1(width=64, height=64)
    tree = orc.build(tris, orc.build_params("reference"))
    a = tree.arrays()
    brute = orc.trace(None, tris, org, dirs, orc.BRUTE, intersection=orc.WOOP)
    for algo in (orc.KD_RESTART, orc.SHORT_STACK):
        h = orc.trace(tree, tris, org, dirs, algo, intersection=orc.WOOP)
        check_against_brute(a, h, brute, org, dirs)
        mt = orc.trace(tree, tris, org, dirs, algo)
        assert (h["id"] == mt["id"]).mean() > 0.99


# ---------------------------------------------------------------- balanced
def balanced_enumeration(bmin, bmax, bmp, ct, ls):
    """numpy restatement: candidates in (prim, axis, low/high) order, cost
    small^2 + large^2 of the two child counts, first minimum below ls^2"""
    idx = [i for i in range(64) if bmp >> i & 1]
    best, choice = ls * ls, None
    for i in idx:
        for ax in range(3):
            for plane in (bmin[i, ax], bmax[i, ax]):
                left = sum(1 << j for j in idx if bmin[j, ax] <= plane)
                right = sum(1 << j for j in idx if plane < bmax[j, ax])
                nl, nr = bin(left).count("1"), bin(right).count("1")
                cost = min(nl, nr) ** 2 + max(nl, nr) ** 2
                if cost < best:
                    best, choice = cost, (ax + 1, np.float32(plane), left, right)
    n = len(idx)
    split = n > 0 and np.float32(np.float32(ct) + np.float32(best)) / np.float32(n) < np.float32(n)
    return split, best, choice


@pytest.mark.parametrize("ct", [16.0, 24.0, 4.0])
def test_balanced_choice_is_the_enumerated_minimum(orc, ct):
    rng = np.random.default_rng(int(ct))
    splits = 0
    for trial in range(150):
        n = int(rng.integers(1, 20))
        bmin = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        bmax = (bmin + rng.uniform(0, 0.5, (n, 3))).astype(np.float32)
        if trial % 5 == 0:                       # coincident planes: ties
            bmin[:, 0] = np.round(bmin[:, 0] * 4) / 4
        bmp = (1 << n) - 1
        if trial % 3 == 0:
            bmp &= int(rng.integers(1, 1 << n))
        got = orc.balanced_node(bmin, bmax, bmp, ct, 32)
        split, best, choice = balanced_enumeration(bmin, bmax, bmp, ct, 32)
        assert got["split"] == split, trial
        if bmp and choice is not None:
            assert got["cost"] == best, trial
        if split:
            splits += 1
            ax, pos, left, right = choice
            assert (got["axis"], np.float32(got["pos"]), got["left"], got["right"]) == (ax, pos, left, right)
    assert splits > 20


def test_balanced_tree_invariants_and_traces(orc, scenes):
    tris, (org, dirs) = scenes.config1(width=96, height=96)
    tree = orc.build(tris, orc.build_params("balanced"))
    a = tree.arrays()
    assert tree.stats()["lower_levels"] >= 2
    reg = regions(a)
    assert len(reg) == a["info"].shape[0]
    # below each upper leaf: children hold subsets of their parent's prims, covering them
    splits = 0
    for leaf in a["leaf_ids"]:
        stack = [int(leaf)]
        while stack:
            n = stack.pop()
            if not (a["info"][n] & 3):
                continue
            l, r = (int(x) for x in a["children"][n])
            pb, lb, rb = (int(a["primitive_bitmap"][x]) & (2 ** 64 - 1) for x in (n, l, r))
            assert lb | rb == pb and lb & ~pb == 0 and rb & ~pb == 0, n
            assert a["primitive_index"][l] == a["primitive_index"][n] == a["primitive_index"][r]
            splits += 1
            stack += [l, r]
    assert splits > 100
    brute = orc.trace(None, tris, org, dirs, orc.BRUTE)
    for algo in (orc.KD_RESTART, orc.SHORT_STACK):
        check_against_brute(a, orc.trace(tree, tris, org, dirs, algo), brute, org, dirs)


# ---------------------------------------------------------------- canvas
def one_quad_scene(alpha=255):
    """two triangles forming the square [-1,1]^2 at z = 0, normals +z"""
    p0 = np.array([[-1, -1, 0, 1], [-1, -1, 0, 1]], np.float32)
    p1 = np.array([[1, -1, 0, 1], [1, 1, 0, 1]], np.float32)
    p2 = np.array([[1, 1, 0, 1], [-1, 1, 0, 1]], np.float32)
    nz = np.tile(np.array([[0, 0, 1, 0]], np.float32), (2, 1))
    col = np.array([[200, 100, 50, alpha]] * 2, np.uint8)
    return (p0, p1, p2), (nz, nz, nz), col


def test_canvas_phong_closed_form(orc):
    tris, normals, col = one_quad_scene()
    light = np.array([0, 0, 2, 0.1, 0.1, 0.1, 0.8, 0.8, 0.8, 1, 1, 1], np.float32)
    org = np.array([[0.25, 0.5, 1, 1], [5, 5, 1, 1]], np.float32)
    dirs = np.array([[0, 0, -1, 0], [0, 0, -1, 0]], np.float32)
    tree = orc.build(tris, orc.build_params("reference"))
    for algo, t in ((orc.BRUTE, None), (orc.KD_RESTART, tree), (orc.SHORT_STACK, tree)):
        c = orc.trace_canvas(t, tris, org, dirs, algo, normals, col, light)
        # opaque: no specular; diffuse = cos of the light angle at the hit
        p = np.array([0.25, 0.5, 0.0])
        ld = np.array([0, 0, 2.0]) - p
        diffuse = ld[2] / np.linalg.norm(ld)
        base = np.array([200, 100, 50]) / 255.0
        want = np.clip(base * (0.1 + 0.8 * diffuse), 0, 1)
        assert np.all(np.abs(c[0, :3].astype(int) - np.floor(want * 255)) <= 1), (algo, c[0])
        assert c[0, 3] == 255
        assert np.array_equal(c[1], [0, 0, 0, 0])          # a miss stays black


def test_canvas_reflection_blends(orc):
    """a half-transparent floor reflects the ray onto an opaque wall: the
    pixel is the floor's colour plus (1 - alpha) of the wall's"""
    floor, fn, fc = one_quad_scene(alpha=128)
    # the same square stood up: plane y = 1.5, x in [-1, 1], z in [0, 2], facing -y
    wall = tuple(np.stack([t[:, 0], np.full(2, 1.5, np.float32), t[:, 1] + 1.0, t[:, 3]], axis=1)
                 .astype(np.float32) for t in floor)
    tris = tuple(np.concatenate([a, b]) for a, b in zip(floor, wall))
    wn = np.tile(np.array([[0, -1, 0, 0]], np.float32), (2, 1))
    normals = tuple(np.concatenate([a, wn]) for a in fn)
    col = np.concatenate([fc, np.array([[0, 255, 0, 255]] * 2, np.uint8)])
    light = np.array([0, 0, 3, 0.2, 0.2, 0.2, 0.5, 0.5, 0.5, 0, 0, 0], np.float32)
    d = np.array([0, 1, -1.0]); d /= np.linalg.norm(d)
    org = np.array([[0, -1, 1, 1]], np.float32)
    dirs = np.array([[d[0], d[1], d[2], 0]], np.float32)
    tree = orc.build(tris, orc.build_params("reference"))
    c1 = orc.trace_canvas(tree, tris, org, dirs, orc.KD_RESTART, normals, col, light, max_bounces=1)
    c2 = orc.trace_canvas(tree, tris, org, dirs, orc.KD_RESTART, normals, col, light)
    assert c1[0, 3] < c2[0, 3] <= 255                 # the second bounce adds coverage
    assert c2[0, 1] > c1[0, 1]                         # and the wall's green
    cs = orc.trace_canvas(tree, tris, org, dirs, orc.SHORT_STACK, normals, col, light)
    assert np.array_equal(cs, c2)
