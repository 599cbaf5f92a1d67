"""Pins three oracle outputs that tree equality alone does not: the triangle
and node areas of the SAH lower stage (S1), the lower children's areas (S4)
and the empty-space threshold rule (B4).

Each check comes from a closed form or a constructed scene whose answer is
known in advance, so that plausible mistakes fail:
  - areas (CalcSurfaceArea, LowerTriangleMap.h:24-38): right triangles with
    power-of-two legs have area exactly a*b/2 (a dropped 0.5 fails); a
    collinear triangle has area 0 and leaves its SAH leaf bitmap
    (PreprocesLowerNodes, :56-81; keeping it fails); a node's area is the
    FP32 sum of its slots' areas in ascending slot order;
  - lower children (CalcAreaForSets :412-459, CreateLowerSAHChildren
    :557-600): each child's area is the ascending FP32 sum over its bitmap;
  - empty-space cuts (CalcEmptySpaceSplits, EmptySpaceSplitting.h:10-70;
    EmptySpaceSplitting2 :74-184; root exemption
    TriangleMapUpperCreator.cu:302): a gap of 0.36 of the PROPAGATED extent
    on one side emits exactly that side's cut, 0.34 emits none (using the
    reduced or the parent's extent fails), a gap of exactly 0.35 emits none
    (writing >= fails), several gaps are chained min-x, min-y, min-z, max-x,
    max-y, max-z (as an X cut at max.x), and the root never gets a chain.
"""
from fractions import Fraction

import numpy as np
import pytest


def pack(v):
    v = np.asarray(v, np.float64)
    return np.concatenate([v, np.ones((v.shape[0], 1))], axis=1).astype(np.float32)


def popcount(b):
    return bin(int(b) & (2 ** 64 - 1)).count("1")


def bits(b):
    b = int(b) & (2 ** 64 - 1)
    return [i for i in range(64) if b >> i & 1]


# ---------------------------------------------------------------- S1: areas
def right_triangle(a, b, plane, flip, origin):
    """legs a, b along two axes of ``plane`` (0: xy, 1: yz, 2: zx) from origin"""
    u, w = plane, (plane + 1) % 3
    o = np.array(origin, np.float64)
    p = o.copy(); p[u] += a
    q = o.copy(); q[w] += b
    return (o, q, p) if flip else (o, p, q)


def test_right_triangle_area_is_half_product(orc):
    checked = 0
    for i in range(-6, 7, 3):
        for j in range(-5, 6, 2):
            a, b = 2.0 ** i, 2.0 ** j
            for plane in range(3):
                for flip in (False, True):
                    for origin in ((0, 0, 0), (0.5, -0.25, 2.0), (-8.0, 4.0, 0.125)):
                        v0, v1, v2 = right_triangle(a, b, plane, flip, origin)
                        assert orc.tri_area(v0, v1, v2) == np.float32(0.5 * a * b), (a, b, plane, flip, origin)
                        checked += 1
    assert checked > 300


def test_collinear_triangle_has_zero_area(orc):
    for v0, v1, v2 in (((0, 0, 0), (1, 1, 1), (2, 2, 2)), ((0.5, 0, 0), (0.5, 0, 0), (0.5, 3, 1)),
                       ((1, 2, 3), (1, 2, 3), (1, 2, 3)), ((0, 0, 0), (0.25, 0, 0), (1, 0, 0))):
        assert orc.tri_area(v0, v1, v2) == 0.0


def dyadic_scene(n, seed):
    """n axis-aligned right triangles with power-of-two legs at dyadic
    positions (exact areas 2^(i+j-1)); every 7th one is made collinear"""
    rng = np.random.default_rng(seed)
    v0, v1, v2, exact = [], [], [], []
    for k in range(n):
        a, b = 2.0 ** int(rng.integers(-7, -3)), 2.0 ** int(rng.integers(-7, -3))
        origin = rng.integers(0, 64, 3) / 64.0
        t = right_triangle(a, b, int(rng.integers(0, 3)), bool(rng.integers(0, 2)), origin)
        if k % 7 == 3:
            t = (t[0], t[1], t[0] + 2.0 * (t[1] - t[0]))   # collinear: area 0
            exact.append(Fraction(0))
        else:
            exact.append(Fraction(a) * Fraction(b) / 2)
        v0.append(t[0]); v1.append(t[1]); v2.append(t[2])
    return (pack(v0), pack(v1), pack(v2)), exact


def test_upper_leaf_area_is_exact_sum_and_zero_area_prims_are_dropped(orc):
    """all areas are dyadic, so every FP32 sum is exact whatever its order:
    the leaf area must be the exact sum (pins the 0.5), and the SAH leaf's
    bitmap must be the bitmap creator's minus the collinear slots"""
    tris, exact = dyadic_scene(900, seed=41)
    never_split = dict(traversal_cost=1e9)
    sah = orc.build(tris, orc.build_params("reference", **never_split)).arrays()
    bmp = orc.build(tris, orc.build_params("bitmap")).arrays()
    assert np.array_equal(sah["leaf_ids"], bmp["leaf_ids"])
    assert np.array_equal(sah["prim_indices"], bmp["prim_indices"])
    dropped = 0
    for leaf in sah["leaf_ids"]:
        pi = int(sah["primitive_index"][leaf])
        count = popcount(bmp["primitive_bitmap"][leaf])
        assert bits(bmp["primitive_bitmap"][leaf]) == list(range(count))
        slots = [int(sah["prim_indices"][pi + i]) for i in range(count)]
        want = [i for i, t in enumerate(slots) if exact[t] > 0]
        assert bits(sah["primitive_bitmap"][leaf]) == want, leaf
        dropped += count - len(want)
        total = sum((exact[t] for t in slots), Fraction(0))
        assert Fraction(float(sah["surface_area"][leaf])) == total, leaf
    assert dropped > 50
    leaf_refs = lambda a: sum(popcount(a["primitive_bitmap"][n]) for n in range(a["info"].shape[0])
                              if a["info"][n] == 0)
    assert leaf_refs(sah) == leaf_refs(bmp) - dropped


def slot_areas(orc, tris, a):
    p0, p1, p2 = tris
    return np.array([orc.tri_area(p0[t, :3], p1[t, :3], p2[t, :3]) for t in a["prim_indices"]], np.float32)


def fp32_sum(values):
    s = np.float32(0.0)
    for v in values:
        s = np.float32(s + np.float32(v))
    return s


def test_upper_leaf_area_is_ascending_fp32_sum(orc, scenes):
    tris = scenes.cluster_soup(6000, n_clusters=3, sigma=0.03, seed=17)
    a = orc.build(tris, orc.build_params("reference", traversal_cost=1e9)).arrays()
    ar = slot_areas(orc, tris, a)
    bmp = orc.build(tris, orc.build_params("bitmap")).arrays()
    order_matters = 0
    for leaf in a["leaf_ids"]:
        pi = int(a["primitive_index"][leaf])
        vals = ar[pi:pi + popcount(bmp["primitive_bitmap"][leaf])]
        asc = fp32_sum(vals)
        assert a["surface_area"][leaf] == asc, leaf
        order_matters += fp32_sum(vals[::-1]) != asc
    assert order_matters > 0          # the scene tells ascending from descending order


# ---------------------------------------------------------------- S4: child areas
def test_lower_children_areas_are_ascending_sums_of_their_sets(orc, scenes):
    tris = scenes.cluster_soup(20000, n_clusters=4, sigma=0.02, seed=7)
    tree = orc.build(tris, orc.build_params("reference", traversal_cost=4.0))
    a = tree.arrays()
    assert tree.stats()["lower_levels"] >= 2
    ar = slot_areas(orc, tris, a)
    upper = set(a["leaf_ids"].tolist())
    lower_start = max(upper) + 1
    checked = order_matters = 0
    for n in range(lower_start, a["info"].shape[0]):
        par = int(a["parent"][n])
        assert a["info"][par] & 3 and n in a["children"][par]
        pi = int(a["primitive_index"][n])
        vals = [ar[pi + i] for i in bits(a["primitive_bitmap"][n])]
        assert a["surface_area"][n] == fp32_sum(vals), n
        order_matters += fp32_sum(vals[::-1]) != fp32_sum(vals)
        checked += 1
    assert checked > 100 and order_matters > 0


# ---------------------------------------------------------------- B4: empty space
def corner_tris(lo, hi, e=0.02):
    """8 small triangles, one in each corner of [lo, hi], pointing inwards:
    their bounding box is exactly [lo, hi]"""
    v0, v1, v2 = [], [], []
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                c = np.array([hi[0] if cx else lo[0], hi[1] if cy else lo[1], hi[2] if cz else lo[2]], np.float64)
                s = np.array([-1 if cx else 1, -1 if cy else 1, -1 if cz else 1], np.float64) * e
                v0.append(c); v1.append(c + [s[0], s[1], 0]); v2.append(c + [0, s[1], s[2]])
    return v0, v1, v2


def gap_scene(test_lo, test_hi, test_right, shift=0.0):
    """Root box x in [0, 4], y and z in [0, 1], split at x = 2.  The test
    node is the root's right child (region [2,4] x [0,1]^2) when test_right,
    else its left child ([0,2] x [0,1]^2); its triangles span exactly
    [test_lo, test_hi].  The other child's triangles span the rest of the
    root box without crossing x = 2."""
    if test_right:
        t = corner_tris(test_lo, test_hi)
        f = corner_tris((0.0, 0.0, 0.0), (1.5, 1.0, 1.0))
    else:
        t = corner_tris(test_lo, test_hi)
        f = corner_tris((2.5, 0.0, 0.0), (4.0, 1.0, 1.0))
    v = [np.array(t[k] + f[k], np.float64) + shift for k in range(3)]
    return pack(v[0]), pack(v[1]), pack(v[2])


def chain(a, slot):
    """the empty-space chain that replaced the root's child ``slot`` (0 left,
    1 right): [(axis, split, empty side)] and the node it ends at"""
    out = []
    n = int(a["children"][0, slot])
    while n >= 3:                       # the root's children are nodes 1 and 2
        l, r = (int(c) for c in a["children"][n])
        assert a["info"][n] & 3
        if r == n + 1 and a["info"][r] == 0 and a["primitive_bitmap"][r] == 0:
            out.append((int(a["info"][n]), float(a["split_pos"][n]), "right"))
            n = l
        else:
            assert l == n + 1 and a["info"][l] == 0 and a["primitive_bitmap"][l] == 0
            out.append((int(a["info"][n]), float(a["split_pos"][n]), "left"))
            n = r
    return out, n


X, Y, Z = 1, 2, 3
f32 = lambda x: float(np.float32(x))


def one_gap(side, g):
    """test node lo/hi and right-ness for a gap of fraction g of the
    propagated extent (x extent 2, y and z extent 1) on one side"""
    if side == "min_x":
        return (2.0 + 2.0 * g, 0.0, 0.0), (4.0, 1.0, 1.0), True
    lo, hi = [0.0, 0.0, 0.0], [2.0, 1.0, 1.0]
    if side == "max_x":
        hi[0] = 2.0 - 2.0 * g
    elif side.startswith("min"):
        lo["xyz".index(side[-1])] = g
    else:
        hi["xyz".index(side[-1])] = 1.0 - g
    return tuple(lo), tuple(hi), False


EXPECT = {  # the cut each side emits: (axis, position from the test box, empty side)
    "min_x": lambda lo, hi: (X, f32(lo[0]), "left"),
    "min_y": lambda lo, hi: (Y, f32(lo[1]), "left"),
    "min_z": lambda lo, hi: (Z, f32(lo[2]), "left"),
    "max_x": lambda lo, hi: (X, f32(hi[0]), "right"),
    "max_y": lambda lo, hi: (Y, f32(hi[1]), "right"),
    "max_z": lambda lo, hi: (X, f32(hi[0]), "right"),   # EmptySpaceSplitting.h:172-175
}


def build_bitmap(orc, tris):
    return orc.build(tris, orc.build_params("bitmap", lower_size=4)).arrays()


@pytest.mark.parametrize("side", list(EXPECT))
def test_empty_space_cut_per_side(orc, side):
    for g, cut in ((0.36, True), (0.34, False)):
        lo, hi, right = one_gap(side, g)
        a = build_bitmap(orc, gap_scene(lo, hi, right))
        assert (a["info"][0] & 3) == X and a["split_pos"][0] == 2.0
        got, end = chain(a, 1 if right else 0)
        assert end == (2 if right else 1)
        assert got == ([EXPECT[side](lo, hi)] if cut else []), (side, g)


def test_empty_space_threshold_is_strict(orc):
    """a gap of exactly 0.35 of the extent (0.35f - 0 > 0.35f * 1 is false)"""
    a = build_bitmap(orc, gap_scene((0.0, 0.35, 0.0), (2.0, 1.0, 1.0), False))
    assert chain(a, 0) == ([], 1)
    a = build_bitmap(orc, gap_scene((0.0, 0.3501, 0.0), (2.0, 1.0, 1.0), False))
    assert chain(a, 0) == ([(Y, f32(0.3501), "left")], 1)


def test_empty_space_chain_order(orc):
    lo, hi = (0.0, 0.4, 0.4), (1.2, 0.6, 0.6)          # left node: five gaps
    a = build_bitmap(orc, gap_scene(lo, hi, False))
    got, end = chain(a, 0)
    assert end == 1
    assert got == [(Y, f32(0.4), "left"), (Z, f32(0.4), "left"), (X, f32(1.2), "right"),
                   (Y, f32(0.6), "right"), (X, f32(1.2), "right")]
    lo, hi = (2.8, 0.45, 0.0), (4.0, 1.0, 0.55)       # right node: min-x, min-y, max-z
    a = build_bitmap(orc, gap_scene(lo, hi, True))
    got, end = chain(a, 1)
    assert end == 2
    assert got == [(X, f32(2.8), "left"), (Y, f32(0.45), "left"), (X, f32(4.0), "right")]


def test_root_is_exempt_from_empty_space_cuts(orc):
    """the root's propagated box is never set (zero): without the exemption a
    scene away from the origin would get cuts at the root"""
    a = build_bitmap(orc, gap_scene((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), False, shift=5.0))
    assert (a["info"][0] & 3) == X and a["split_pos"][0] == 7.0
    assert a["children"][0].tolist() == [1, 2]
    assert np.array_equal(a["aabb_min"][0, :3], [5, 5, 5]) and np.array_equal(a["aabb_max"][0, :3], [9, 6, 6])


def b4_scenes():
    """every constructed scene above, for the GPU parity test"""
    out = []
    for side in EXPECT:
        for g in (0.34, 0.36):
            lo, hi, right = one_gap(side, g)
            out.append(gap_scene(lo, hi, right))
    out.append(gap_scene((0.0, 0.35, 0.0), (2.0, 1.0, 1.0), False))
    out.append(gap_scene((0.0, 0.4, 0.4), (1.2, 0.6, 0.6), False))
    out.append(gap_scene((2.8, 0.45, 0.0), (4.0, 1.0, 0.55), True))
    out.append(gap_scene((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), False, shift=5.0))
    return out
