"""GPU parity: the HIP build and tracers against the CPU oracle.

Bar (integer/index work and every float value produced by the same IEEE
operation sequence): bit-exact.  The tree arrays (node numbering, split
positions, boxes, bitmaps, prim indices, leaf order) and every hit record
{id, t, u, v} must be identical bit for bit.  At full size (1M and 2M
triangles) the tree is compared whole and the traces on seeded ray
subsets; brute force is compared on a small subset.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TREE_KEYS = ("info", "split_pos", "aabb_min", "aabb_max", "primitive_index", "primitive_bitmap",
             "children", "surface_area", "parent", "prim_indices", "leaf_ids")


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "these tests need the GPU"
    return T


@pytest.fixture(scope="module")
def ctx(pm, torch):
    return pm.Context(0)


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def assert_same_tree(got, ref, label=""):
    for k in TREE_KEYS:
        g, r = got[k], ref[k]
        assert g.shape == r.shape, (label, k, g.shape, r.shape)
        if g.dtype.kind == "f":
            g, r = g.view(np.int32), r.view(np.int32)
        if not np.array_equal(g, r):
            bad = np.nonzero((g != r).reshape(g.shape[0], -1).any(axis=1))[0]
            raise AssertionError(f"{label} {k}: {len(bad)} rows differ, first {bad[:5]}")


def assert_same_hits(got, ref, label=""):
    g, r = got.view(np.int32).reshape(-1, 4), ref.view(np.int32).reshape(-1, 4)
    if not np.array_equal(g, r):
        bad = np.nonzero((g != r).any(axis=1))[0]
        raise AssertionError(f"{label}: {len(bad)} rays differ, first {bad[:5]}: "
                             f"gpu {got[bad[:2]]} oracle {ref[bad[:2]]}")


def gpu_build(pm, ctx, torch, tris, params):
    P = [dev(torch, t) for t in tris]
    ctx.build(*P, params)
    torch.cuda.synchronize()
    return P, ctx.tree_arrays()


def orc_params(orc, pm_params):
    """same fields, oracle struct"""
    p = orc.BuildParams()
    for name, _ in orc.BuildParams._fields_:
        setattr(p, name, getattr(pm_params, name))
    return p


PRESETS = {
    "reference": dict(preset="reference"),
    "default": dict(preset="default"),
    "config": dict(preset="config"),
    "no_empty_space": dict(preset="reference", empty_space_splitting=0),
    "no_propagate": dict(preset="reference", propagate_aabbs=0),
    "no_empty_no_propagate": dict(preset="default", empty_space_splitting=0, propagate_aabbs=0),
    "lower64": dict(preset="reference", lower_size=64),
    "lower8": dict(preset="default", lower_size=8),
    "ct4": dict(preset="reference", traversal_cost=4.0),
}


def make_params(pm, spec):
    spec = dict(spec)
    return pm.build_params(spec.pop("preset"), **spec)


@pytest.mark.parametrize("name", list(PRESETS))
def test_c1_tree_and_traces_bit_exact(pm, orc, scenes, torch, ctx, name):
    tris, (org, dirs) = scenes.config1(width=128, height=128)
    params = make_params(pm, PRESETS[name])
    P, got = gpu_build(pm, ctx, torch, tris, params)
    ref_tree = orc.build(tris, orc_params(orc, params))
    assert_same_tree(got, ref_tree.arrays(), name)
    st = ctx.stats()
    rs = ref_tree.stats()
    for k in ("num_nodes", "reachable_nodes", "inner_nodes", "leaves", "empty_leaves", "max_depth",
              "leaf_refs", "max_leaf_size", "upper_levels", "lower_levels"):
        assert st[k] == rs[k], (k, st[k], rs[k])
    O, D = dev(torch, org), dev(torch, dirs)
    skip = bool(params.propagate_aabbs)
    for algo, ss in ((pm.KD_RESTART, 4), (pm.SHORT_STACK, 4), (pm.SHORT_STACK, 1),
                     (pm.SHORT_STACK, 8), (pm.SHORT_STACK, 16)):
        h = pm.hits_to_numpy(ctx.trace(algo, O, D, pm.trace_params(stack_size=ss)))
        r = orc.trace(ref_tree, tris, org, dirs, algo, leaf_skipping=skip, stack_size=ss)
        assert_same_hits(h, r, f"{name} algo={algo} S={ss}")


def test_c1_config_256_camera_rays_bit_exact(pm, orc, scenes, torch, ctx):
    """the Cornell configuration as it is benchmarked: CONFIG preset, 256 x 256
    jittered camera rays made on the GPU; the tree, the rays and every hit of
    the short stack (depth 8), kd-restart and brute force against the oracle"""
    tris, _ = scenes.config1()
    params = pm.build_params("config")
    P, got = gpu_build(pm, ctx, torch, tris, params)
    ref_tree = orc.build(tris, orc_params(orc, params))
    assert_same_tree(got, ref_tree.arrays(), "C1 config")
    m, c = scenes.camera(scenes.CAMERA_C1, 256, 256)
    O, D = ctx.create_rays(m, c, 256, 256, jitter_seed=1)
    org, dirs = orc.create_rays(m, c, 256, 256, jitter_seed=1)
    assert np.array_equal(O.cpu().numpy().view(np.int32), org.view(np.int32))
    assert np.array_equal(D.cpu().numpy().view(np.int32), dirs.view(np.int32))
    for algo, ss in ((pm.SHORT_STACK, 8), (pm.KD_RESTART, 4)):
        h = pm.hits_to_numpy(ctx.trace(algo, O, D, pm.trace_params(stack_size=ss)))
        r = orc.trace(ref_tree, tris, org, dirs, algo, stack_size=ss)
        assert_same_hits(h, r, f"C1 config 256 algo={algo}")
        assert (h["id"] >= 0).mean() > 0.9
    hb = pm.hits_to_numpy(ctx.trace(pm.BRUTE, O, D))
    assert_same_hits(hb, orc.trace(None, tris, org, dirs, orc.BRUTE), "C1 config 256 brute")


def test_c1_brute_force_bit_exact(pm, orc, scenes, torch, ctx):
    tris, (org, dirs) = scenes.config1(width=128, height=128)
    P = [dev(torch, t) for t in tris]
    ctx.set_geometry(*P)
    O, D = dev(torch, org), dev(torch, dirs)
    h = pm.hits_to_numpy(ctx.trace(pm.BRUTE, O, D))
    assert_same_hits(h, orc.trace(None, tris, org, dirs, orc.BRUTE), "brute")
    h = pm.hits_to_numpy(ctx.trace(pm.BRUTE, O, D, pm.trace_params(brute_cap=4096)))
    assert_same_hits(h, orc.trace(None, tris, org, dirs, orc.BRUTE, brute_cap=4096), "brute cap")


@pytest.mark.parametrize("n", [0, 1, 2, 5, 31, 32, 33, 64, 100, 257, 1000, 20000])
def test_random_soups_bit_exact(pm, orc, scenes, torch, ctx, n):
    tris = scenes.random_soup(n, seed=1000 + n, size=0.15)
    org, dirs = scenes.random_rays(4096, seed=2000 + n)
    for name in ("reference", "default", "lower64"):
        params = make_params(pm, PRESETS[name])
        P, got = gpu_build(pm, ctx, torch, tris, params)
        ref_tree = orc.build(tris, orc_params(orc, params))
        assert_same_tree(got, ref_tree.arrays(), f"n={n} {name}")
        O, D = dev(torch, org), dev(torch, dirs)
        for algo in (pm.KD_RESTART, pm.SHORT_STACK):
            h = pm.hits_to_numpy(ctx.trace(algo, O, D))
            assert_same_hits(h, orc.trace(ref_tree, tris, org, dirs, algo), f"n={n} {name} {algo}")


SAH_SEARCH = {
    "ct1_lower64": dict(preset="reference", traversal_cost=1.0, lower_size=64),
    "ct2_forced16_bonus": dict(preset="config", traversal_cost=2.0, lower_size=64, sah_leaf_max=16,
                               sah_empty_bonus=0.5),
    "ct3_bonus_above_one": dict(preset="reference", traversal_cost=3.0, sah_empty_bonus=1.5),
}


@pytest.mark.parametrize("n,size", [(40, 0.3), (300, 0.3), (3000, 0.1)])
@pytest.mark.parametrize("name", list(SAH_SEARCH))
def test_sah_full_search_bit_exact(pm, orc, scenes, torch, ctx, n, size, name):
    """low traversal costs send most lower nodes past the early rejection
    into the full plane search, over several lower levels with sparse
    bitmaps (up to 64 prims per node)"""
    tris = scenes.random_soup(n, seed=3000 + n, size=size)
    params = make_params(pm, SAH_SEARCH[name])
    P, got = gpu_build(pm, ctx, torch, tris, params)
    ref_tree = orc.build(tris, orc_params(orc, params))
    assert_same_tree(got, ref_tree.arrays(), f"n={n} {name}")
    assert ctx.stats()["lower_levels"] >= 2, "no lower split: the full search was not exercised"
    org, dirs = scenes.random_rays(2048, seed=4000 + n)
    O, D = dev(torch, org), dev(torch, dirs)
    h = pm.hits_to_numpy(ctx.trace(pm.SHORT_STACK, O, D))
    assert_same_hits(h, orc.trace(ref_tree, tris, org, dirs, pm.SHORT_STACK), f"n={n} {name}")


def test_counters_match_oracle(pm, orc, scenes, torch, ctx):
    tris, (org, dirs) = scenes.config1(width=64, height=64)
    params = pm.build_params("reference")
    P, got = gpu_build(pm, ctx, torch, tris, params)
    ref_tree = orc.build(tris, orc_params(orc, params))
    O, D = dev(torch, org), dev(torch, dirs)
    for algo in (pm.KD_RESTART, pm.SHORT_STACK):
        cnt = torch.zeros((org.shape[0], 4), dtype=torch.int32, device="cuda:0")
        ctx.trace(algo, O, D, counters=cnt)
        _, rc = orc.trace(ref_tree, tris, org, dirs, algo, counters=True)
        assert np.array_equal(cnt.cpu().numpy(), rc), algo


def test_set_tree_traces_an_uploaded_tree(pm, orc, scenes, torch, ctx):
    tris, (org, dirs) = scenes.config1(width=64, height=64)
    ref_tree = orc.build(tris, orc.build_params("config"))
    P = [dev(torch, t) for t in tris]
    ctx.set_tree(ref_tree.arrays(), *P)
    O, D = dev(torch, org), dev(torch, dirs)
    for algo in (pm.KD_RESTART, pm.SHORT_STACK):
        h = pm.hits_to_numpy(ctx.trace(algo, O, D))
        assert_same_hits(h, orc.trace(ref_tree, tris, org, dirs, algo), f"uploaded {algo}")


def test_rebuild_in_same_context(pm, orc, scenes, torch, ctx):
    """scratch reuse across builds of different sizes (grow-only buffers)"""
    for n, seed in ((30000, 1), (500, 2), (60000, 3), (33, 4)):
        tris = scenes.random_soup(n, seed=seed, size=0.05)
        params = pm.build_params("reference")
        P, got = gpu_build(pm, ctx, torch, tris, params)
        assert_same_tree(got, orc.build(tris, orc_params(orc, params)).arrays(), f"rebuild n={n}")


@pytest.mark.skip(reason="capacity-planned builds are switched off in the library (PLANNED_BUILDS in "
                         "pm_build.hip): the 1M-triangle benchmark faulted the GPU with them on")
def test_planned_builds_match_level_by_level(pm, orc, scenes, torch):
    """a context that has built before builds with capacities from that build
    (no read-back per level); one that outgrows them falls back to sizing
    level by level; the tree is the oracle's either way"""
    c = pm.Context(0)
    params = pm.build_params("reference")
    steps = [("first", scenes.random_soup(3000, seed=41, size=0.1), 0),
             ("again", scenes.random_soup(3000, seed=41, size=0.1), 1),
             ("other", scenes.random_soup(2500, seed=43, size=0.12), 1),
             ("more refs", scenes.random_soup(3000, seed=42, size=0.3), 0),      # outgrows the capacities
             ("after growth", scenes.random_soup(3000, seed=42, size=0.3), 1),
             ("far smaller", scenes.random_soup(500, seed=44, size=0.1), 0),     # outside [n/2, 1.25n]
             ("cornell config", scenes.config1()[0], 0)]
    for name, tris, planned in steps:
        p = pm.build_params("config") if name == "cornell config" else params
        _, got = gpu_build(pm, c, torch, tris, p)
        assert c.stats()["planned_build"] == planned, name
        assert_same_tree(got, orc.build(tris, orc_params(orc, p)).arrays(), name)
    # the same scene under the config preset again: planned
    tris = scenes.config1()[0]
    p = pm.build_params("config")
    _, got = gpu_build(pm, c, torch, tris, p)
    assert c.stats()["planned_build"] == 1
    assert_same_tree(got, orc.build(tris, orc_params(orc, p)).arrays(), "cornell again")
    # a planned attempt that ends in a depth error raises it (5000 coincident triangles)
    q = [np.repeat(np.array([v], np.float32), 5000, axis=0) for v in ([0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1])]
    with pytest.raises(pm.PmError) as e:
        c.build(*[dev(torch, x) for x in q], params)
    assert e.value.code == 3                                   # PM_ERR_DEPTH
    c.close()


def test_error_codes(pm, torch, scenes):
    c = pm.Context(0)
    org, dirs = scenes.random_rays(16, seed=1)
    with pytest.raises(pm.PmError) as e:
        c.trace(pm.SHORT_STACK, dev(torch, org), dev(torch, dirs))
    assert e.value.code == 7                                   # PM_ERR_NO_TREE
    tris = [dev(torch, t) for t in scenes.random_soup(100, seed=1)]
    with pytest.raises(pm.PmError) as e:
        c.build(*tris, pm.build_params("reference", split_method=pm.SPLIT_SPLIT))
    assert e.value.code == 2                                   # PM_ERR_NOT_IMPLEMENTED
    with pytest.raises(pm.PmError) as e:
        c.build(*tris, pm.build_params("reference", lower_algorithm=7))
    assert e.value.code == 1                                   # PM_ERR_ARG
    p0 = np.repeat(np.array([[0, 0, 0, 1]], np.float32), 40, axis=0)
    p1 = np.repeat(np.array([[1, 0, 0, 1]], np.float32), 40, axis=0)
    p2 = np.repeat(np.array([[0, 1, 0, 1]], np.float32), 40, axis=0)
    with pytest.raises(pm.PmError) as e:
        c.build(dev(torch, p0), dev(torch, p1), dev(torch, p2), pm.build_params("reference"))
    assert e.value.code == 3                                   # PM_ERR_DEPTH
    with pytest.raises(pm.PmError) as e:
        c.trace(pm.SHORT_STACK, dev(torch, org), dev(torch, dirs), pm.trace_params(stack_size=17))
    assert e.value.code == 1
    # an uploaded tree whose indices leave their arrays is refused, not traced
    import pyoracle as orc
    tris_h = scenes.random_soup(100, seed=1)
    good = orc.build(tris_h).arrays()
    bad = {k: np.copy(v) for k, v in good.items()}
    inner = np.nonzero(bad["info"] & 3)[0][0]
    bad["children"][inner, 1] = bad["info"].shape[0] + 5
    with pytest.raises(pm.PmError) as e:
        c.set_tree(bad, *tris)
    assert e.value.code == 1
    bad = {k: np.copy(v) for k, v in good.items()}
    leaf = np.nonzero((bad["info"] == 0) & (bad["primitive_bitmap"] != 0))[0][0]
    bad["primitive_index"][leaf] = bad["prim_indices"].shape[0]
    with pytest.raises(pm.PmError) as e:
        c.set_tree(bad, *tris)
    assert e.value.code == 1
    bad = {k: np.copy(v) for k, v in good.items()}
    bad["prim_indices"][0] = 100          # triangle id outside p0/p1/p2
    with pytest.raises(pm.PmError) as e:
        c.set_tree(bad, *tris)
    assert e.value.code == 1
    c.close()


# ------------------------------------------------------------- full size
@pytest.fixture(scope="module")
def c3(scenes):
    tris = scenes.cluster_soup()                     # 1,000,002 triangles
    org, dirs = scenes.photon_emission(200_000)
    return tris, org, dirs


def test_c3_1m_tree_and_traces_bit_exact(pm, orc, torch, ctx, c3):
    tris, org, dirs = c3
    params = pm.build_params("reference")
    P, got = gpu_build(pm, ctx, torch, tris, params)
    ref_tree = orc.build(tris, orc_params(orc, params))
    assert_same_tree(got, ref_tree.arrays(), "C3")
    O, D = dev(torch, org), dev(torch, dirs)
    for algo, ss in ((pm.SHORT_STACK, 8), (pm.KD_RESTART, 4)):
        h = pm.hits_to_numpy(ctx.trace(algo, O, D, pm.trace_params(stack_size=ss)))
        r = orc.trace(ref_tree, tris, org, dirs, algo, stack_size=ss)
        assert_same_hits(h, r, f"C3 algo={algo}")
    # brute force on a subset: bit-exact with the oracle's brute force, and
    # the tree tracer's ids agree with it (slack-explained rays aside)
    n = 512
    ctx.set_geometry(*P)
    hb = pm.hits_to_numpy(ctx.trace(pm.BRUTE, O[:n].contiguous(), D[:n].contiguous()))
    rb = orc.trace(None, tris, org[:n], dirs[:n], orc.BRUTE)
    assert_same_hits(hb, rb, "C3 brute")
    hs = orc.trace(ref_tree, tris, org[:n], dirs[:n], orc.SHORT_STACK, stack_size=8)
    assert (hs["id"] != rb["id"]).sum() <= 2


def test_c2_heightfield_2m_tree_bit_exact(pm, orc, scenes, torch, ctx):
    tris, (org, dirs) = scenes.config2_primary(width=480, height=270)
    params = pm.build_params("reference")
    P, got = gpu_build(pm, ctx, torch, tris, params)
    ref_tree = orc.build(tris, orc_params(orc, params))
    assert_same_tree(got, ref_tree.arrays(), "C2")
    O, D = dev(torch, org), dev(torch, dirs)
    h = pm.hits_to_numpy(ctx.trace(pm.SHORT_STACK, O, D, pm.trace_params(stack_size=8)))
    assert_same_hits(h, orc.trace(ref_tree, tris, org, dirs, orc.SHORT_STACK, stack_size=8), "C2")
    # the benchmarked workload: 1920 x 1080 jittered camera rays made on the
    # GPU, then one cosine bounce ray per hit; both passes in both ray orders
    W, Hh = 1920, 1080
    m, c = scenes.camera(scenes.CAMERA_C2, W, Hh)
    O, D = ctx.create_rays(m, c, W, Hh, jitter_seed=2)
    org, dirs = orc.create_rays(m, c, W, Hh, jitter_seed=2)
    assert np.array_equal(O.cpu().numpy().view(np.int32), org.view(np.int32))
    assert np.array_equal(D.cpu().numpy().view(np.int32), dirs.view(np.int32))
    ntris = tris[0].shape[0]
    for order in (pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT):
        tp = pm.trace_params(stack_size=8, ray_order=order)
        H = ctx.trace(pm.SHORT_STACK, O, D, tp)
        h = pm.hits_to_numpy(H)
        if order == pm.RAY_ORDER_INPUT:
            assert_same_hits(h, orc.trace(ref_tree, tris, org, dirs, orc.SHORT_STACK, stack_size=8, threads=8),
                             "C2 camera")
            assert (h["id"] >= 0).mean() > 0.9
            bo, bd, _, nb = ctx.photon_bounce(O, D, H, ntris, seed=2002)
            assert nb == int((h["id"] >= 0).sum())
            bo, bd = bo[:nb].contiguous(), bd[:nb].contiguous()
            rb = orc.trace(ref_tree, tris, bo.cpu().numpy(), bd.cpu().numpy(), orc.SHORT_STACK, stack_size=8,
                           threads=8)
        else:
            assert_same_hits(h, prev, "C2 camera coherent")
        hb = pm.hits_to_numpy(ctx.trace(pm.SHORT_STACK, bo, bd, tp))
        assert_same_hits(hb, rb, f"C2 bounce order={order}")
        prev = h


@pytest.mark.parametrize("name", ["reference", "ct1_lower64"])
def test_build_work_counts(pm, orc, scenes, torch, ctx, name):
    """the build's work counts (inputs of bench.py's build roofline) against
    the oracle: refs read = the oracle's per-level refs summed; refs written =
    the later levels' refs + the final leaf refs; SAH lower prims = the bitmap
    sizes of every node below an upper leaf (each is examined once)"""
    spec = PRESETS.get(name) or SAH_SEARCH[name]
    tris = scenes.random_soup(3000, seed=3000 + 3000, size=0.1)
    params = make_params(pm, spec)
    P, got = gpu_build(pm, ctx, torch, tris, params)
    ref_tree = orc.build(tris, orc_params(orc, params))
    rs = ref_tree.stats()
    st = ctx.stats()
    lvl = rs["upper_level_refs"]
    assert st["upper_ref_reads"] == sum(lvl)
    assert st["upper_ref_writes"] == sum(lvl[1:]) + st["num_prim_indices"]
    bm = got["primitive_bitmap"].astype(np.uint64)
    ch = got["children"].reshape(-1, 2)
    info = got["info"]
    total, stack = 0, list(got["leaf_ids"])
    while stack:
        nd = stack.pop()
        total += bin(int(bm[nd])).count("1")
        if info[nd] & 3:
            stack.extend(ch[nd])
    assert st["lower_prims"] == total and total > 0


def test_random_configs_bit_exact(pm, orc, scenes, torch, ctx):
    """40 seeded random cases: soup size 1..120k (log-uniform), triangle size,
    every preset above and every chunk size; tree arrays and short-stack hits
    must equal the oracle's"""
    specs = dict(PRESETS)
    specs.update(SAH_SEARCH)
    names = sorted(specs)
    rng = np.random.default_rng(2024)
    checked = 0
    for case in range(40):
        n = int(np.exp(rng.uniform(0, np.log(120000))))
        size = min(float(np.exp(rng.uniform(np.log(0.005), np.log(0.3)))), 3.0 / n ** (1 / 3))
        name = names[rng.integers(len(names))]
        chunk = int(rng.choice([0, 128, 256, 512, 1024, 2048, 4096]))
        seed = int(rng.integers(1 << 30))
        spec = dict(specs[name])
        params = pm.build_params(spec.pop("preset"), chunk_size=chunk, **spec)
        tris = scenes.random_soup(n, seed=seed, size=size)
        label = f"case {case}: n={n} size={size:.4f} {name} chunk={chunk} seed={seed}"
        try:
            ref_tree = orc.build(tris, orc_params(orc, params))
        except RuntimeError:       # the oracle's depth/ref limits; the GPU build is checked elsewhere
            continue
        P, got = gpu_build(pm, ctx, torch, tris, params)
        assert_same_tree(got, ref_tree.arrays(), label)
        org, dirs = scenes.random_rays(2048, seed=seed + 1)
        h = pm.hits_to_numpy(ctx.trace(pm.SHORT_STACK, dev(torch, org), dev(torch, dirs)))
        r = orc.trace(ref_tree, tris, org, dirs, orc.SHORT_STACK, leaf_skipping=bool(params.propagate_aabbs))
        assert_same_hits(h, r, label)
        checked += 1
    assert checked >= 35


def test_pinned_area_and_empty_space_scenes_bit_exact(pm, orc, torch, ctx):
    """the constructed scenes that pin the oracle's areas (S1, S4) and its
    empty-space rule (B4) build the same trees on the GPU"""
    import test_oracle_pins as pins
    cases = [(t, dict(preset="default", lower_size=4)) for t in pins.b4_scenes()]
    dy, _ = pins.dyadic_scene(900, seed=41)
    cases += [(dy, dict(preset="reference", traversal_cost=1e9)), (dy, dict(preset="reference"))]
    for i, (tris, spec) in enumerate(cases):
        params = make_params(pm, spec)
        _, got = gpu_build(pm, ctx, torch, tris, params)
        ref = orc.build(tris, orc_params(orc, params)).arrays()
        assert_same_tree(got, ref, f"pinned scene {i}")
