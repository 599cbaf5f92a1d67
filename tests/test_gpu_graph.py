"""A batch's rebuild + trace recorded once and replayed as one hipGraph launch
per frame (pm_batch_graph_capture / _launch / _status).

Bar: every replay gives the tree and hits an eager batched build + trace of
the same geometry gives, bit for bit (and so the oracle forest's); replays
follow in-place geometry updates; replays interleaved with eager work on
the same context stay exact; a replay that outgrows its capacities reports
PM_ERR_CAPACITY and traces nothing.
"""
import numpy as np
import pytest

from forest_util import c4_small
from test_gpu_parity import assert_same_hits, assert_same_tree, dev, orc_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "these tests need the GPU"
    return T


def frame_tris(scenes, c, frame):
    """triangles of the C4-style batch ``c`` at ``frame`` (FP64 deformer)"""
    V = scenes.displace_sphere64(c["unit"], frame)
    F = c["faces"]
    return scenes.tris_from(V[F[:, 0]], V[F[:, 1]], V[F[:, 2]])


@pytest.mark.parametrize("preset", ["reference", "default"])
def test_replays_equal_eager_and_oracle(pm, orc, scenes, torch, preset):
    c = scenes.config4(num_meshes=10, rays_per_mesh=1000, hi=1.5e4)
    off = c["tri_off"]
    tris = frame_tris(scenes, c, 0.0)
    P = [dev(torch, t) for t in tris]
    O, D, RM = dev(torch, c["org"]), dev(torch, c["dirs"]), dev(torch, c["ray_mesh"])
    H = torch.empty((O.shape[0], 4), dtype=torch.float32, device="cuda:0")
    bp = pm.build_params(preset)
    tp = pm.trace_params(stack_size=8)
    g = pm.Context(0)
    caps = g.batch_graph_capture(*P, off, O, D, RM, H, pm.SHORT_STACK, bp, tp)
    assert caps["upper_levels"] > 0 and caps["nodes"] > 0
    eager = pm.Context(0)
    for frame in (0.0, 1.0, 5.0, 9.0):
        t = frame_tris(scenes, c, frame)
        for p, x in zip(P, t):
            p.copy_(torch.from_numpy(x))
        g.batch_graph_launch()
        assert g.batch_graph_status() == 0
        got_tree = g.tree_arrays()
        got_hits = pm.hits_to_numpy(H)
        # the same frame, eager
        eager.build_batched(*P, off, bp)
        ref_tree = eager.tree_arrays()
        assert_same_tree(got_tree, ref_tree, f"{preset} frame {frame}")
        h2 = pm.hits_to_numpy(eager.trace_batched(pm.SHORT_STACK, O, D, RM, tp))
        assert_same_hits(got_hits, h2, f"{preset} frame {frame}")
        # and the oracle forest of the same triangles
        ref = orc.build_batched(t, off, orc_params(orc, bp))
        assert_same_tree(got_tree, ref.arrays(), f"{preset} frame {frame} oracle")
        r = orc.trace_batched(ref, t, c["org"], c["dirs"], c["ray_mesh"], orc.SHORT_STACK, stack_size=8)
        assert_same_hits(got_hits, r, f"{preset} frame {frame} oracle hits")
        st = g.stats()
        assert st["num_nodes"] == ref_tree["info"].shape[0]


def test_replays_with_eager_work_in_between(pm, orc, scenes, torch):
    """several replays on one context, with eager builds/traces of other
    geometry on the same context between them (same stream); every replay
    is checked against an eager result"""
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=7, empty=(2,))
    P = [dev(torch, t) for t in tris]
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, ray_mesh)
    H = torch.empty((O.shape[0], 4), dtype=torch.float32, device="cuda:0")
    bp = pm.build_params("reference")
    ctx = pm.Context(0)
    ctx.batch_graph_capture(*P, off, O, D, RM, H, pm.KD_RESTART, bp)
    ref = orc.build_batched(tris, off, orc_params(orc, bp))
    r = orc.trace_batched(ref, tris, org, dirs, ray_mesh, orc.KD_RESTART)
    small = [dev(torch, t) for t in scenes.random_soup(3000, seed=21)]
    sorg, sdirs = scenes.random_rays(4000, seed=22)
    for i in range(4):
        H.fill_(0.0)
        ctx.batch_graph_launch()
        torch.cuda.synchronize()
        assert ctx.batch_graph_status() == 0
        assert_same_hits(pm.hits_to_numpy(H), r, f"replay {i}")
        assert_same_tree(ctx.tree_arrays(), ref.arrays(), f"replay {i} tree")
        # eager work on the same context (a smaller tree: no buffer regrows)
        ctx.build(*small, bp)
        ctx.trace(pm.SHORT_STACK, dev(torch, sorg), dev(torch, sdirs))
        torch.cuda.synchronize()


def test_capacity_errors(pm, scenes, torch):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=5)
    P = [dev(torch, t) for t in tris]
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, ray_mesh)
    H = torch.empty((O.shape[0], 4), dtype=torch.float32, device="cuda:0")
    ctx = pm.Context(0)
    # too few levels recorded, or too few nodes: the checked first replay fails
    for caps in ({"upper_levels": 2}, {"nodes": 1}):
        with pytest.raises(pm.PmError) as e:
            ctx.batch_graph_capture(*P, off, O, D, RM, H, pm.SHORT_STACK, caps=caps)
        assert e.value.code == 8, caps
        with pytest.raises(pm.PmError):
            ctx.batch_graph_launch()                  # nothing captured
    # a later call that regrows the captured buffers invalidates the graph
    ctx.batch_graph_capture(*P, off, O, D, RM, H, pm.SHORT_STACK)
    big = [dev(torch, t) for t in scenes.random_soup(200000, seed=23, size=0.01)]
    ctx.build(*big, pm.build_params("reference"))
    with pytest.raises(pm.PmError) as e:
        ctx.batch_graph_launch()
    assert e.value.code == 8
    with pytest.raises(pm.PmError):                   # BRUTE is not a batched tracer
        ctx.batch_graph_capture(*P, off, O, D, RM, H, pm.BRUTE)


def test_build_only_graph_then_eager_traces(pm, orc, scenes, torch):
    """a graph of the build alone (no rays): each replay rebuilds from the
    current triangles; eager traces after it see the replayed tree; a
    replay that outgrows its capacities makes the next trace fail cleanly"""
    tris = scenes.random_soup(20000, seed=41, size=0.02)
    org, dirs = scenes.random_rays(30000, seed=42)
    P = [dev(torch, t) for t in tris]
    O, D = dev(torch, org), dev(torch, dirs)
    bp = pm.build_params("reference")
    off = np.array([0, tris[0].shape[0]], np.int32)
    ctx = pm.Context(0)
    ctx.batch_graph_capture(*P, off, None, None, None, None, pm.SHORT_STACK, bp)
    for seed in (41, 43):
        t = scenes.random_soup(20000, seed=seed, size=0.02)
        for p, x in zip(P, t):
            p.copy_(torch.from_numpy(x))
        ctx.batch_graph_launch()
        h = pm.hits_to_numpy(ctx.trace(pm.SHORT_STACK, O, D, pm.trace_params(stack_size=8)))
        ref = orc.build(t, orc_params(orc, bp))
        assert_same_tree(ctx.tree_arrays(), ref.arrays(), f"replay seed {seed}")
        assert_same_hits(h, orc.trace(ref, t, org, dirs, orc.SHORT_STACK, stack_size=8), f"replay seed {seed}")
    # much larger triangles: far more references than the capacities allow
    big = scenes.random_soup(20000, seed=44, size=0.5)
    for p, x in zip(P, big):
        p.copy_(torch.from_numpy(x))
    ctx.batch_graph_launch()
    with pytest.raises(pm.PmError):
        ctx.trace(pm.SHORT_STACK, O, D)


def test_depth_error_in_a_replay(pm, orc, scenes, torch):
    """a replay whose geometry needs more upper levels than max_upper_levels
    stops with PM_ERR_DEPTH (the status call reports it; the captured levels
    after it do no work) and leaves the context usable"""
    n = 4000
    a = scenes.random_soup(n, seed=5, size=0.01)          # 8 upper levels
    b = [t.copy() for t in a]                               # 19: half of it shrunk 100x
    c = np.array([0.3, 0.3, 0.3], np.float32)
    for t in b:
        t[: n // 2, :3] = c + (t[: n // 2, :3] - c) * 0.01
    P = [dev(torch, t) for t in a]
    off = np.array([0, n], np.int32)
    ctx = pm.Context(0)
    bp = pm.build_params("reference", max_upper_levels=8)
    caps = {"upper_levels": 24, "lower_levels": 6, "nodes": 8192, "active": 4096, "segments": 8192,
            "level_refs": 65536, "leaf_refs": 65536, "leaf_ids": 8192, "lower_nodes": 16384}
    ctx.batch_graph_capture(*P, off, None, None, None, None, pm.SHORT_STACK, bp, caps=caps)
    for t, h in zip(P, b):
        t.copy_(torch.from_numpy(h))
    ctx.batch_graph_launch()
    assert ctx.batch_graph_status() == 3                  # PM_ERR_DEPTH
    # the same context then builds the deeper geometry eagerly, as the oracle does
    got = ctx.build(*P, pm.build_params("reference"))
    torch.cuda.synchronize()
    ref = orc.build(b, orc.build_params("reference"))
    arr, ra = ctx.tree_arrays(), ref.arrays()
    for k in ("info", "children", "primitive_bitmap", "prim_indices"):
        assert np.array_equal(arr[k], ra[k]), k
    ctx.close()


def test_random_configs_graph_equals_eager(pm, scenes, torch):
    """20 seeded random single-mesh builds (size, triangle size, preset, chunk
    size): a captured build-only graph's replay gives the eager build's tree
    bit for bit (the captured build runs fixed grids over its capacities)"""
    from test_gpu_parity import PRESETS, SAH_SEARCH, TREE_KEYS
    specs = dict(PRESETS)
    specs.update(SAH_SEARCH)
    names = sorted(specs)
    rng = np.random.default_rng(77)
    checked = 0
    for case in range(20):
        n = int(np.exp(rng.uniform(np.log(2), np.log(60000))))
        size = min(float(np.exp(rng.uniform(np.log(0.005), np.log(0.3)))), 3.0 / n ** (1 / 3))
        name = names[rng.integers(len(names))]
        chunk = int(rng.choice([0, 128, 256, 512, 1024, 2048, 4096]))
        seed = int(rng.integers(1 << 30))
        spec = dict(specs[name])
        bp = pm.build_params(spec.pop("preset"), chunk_size=chunk, **spec)
        tris = scenes.random_soup(n, seed=seed, size=size)
        label = f"case {case}: n={n} {name} chunk={chunk} seed={seed}"
        P = [dev(torch, t) for t in tris]
        off = np.array([0, n], np.int32)
        e = pm.Context(0)
        try:
            e.build_batched(*P, off, bp)
        except pm.PmError:          # the reference-count guard; not a capture case
            e.close()
            continue
        torch.cuda.synchronize()
        ref = {k: np.array(v, copy=True) for k, v in e.tree_arrays().items()}
        e.close()
        g = pm.Context(0)
        g.batch_graph_capture(*P, off, None, None, None, None, pm.SHORT_STACK, bp)
        g.batch_graph_launch()
        assert g.batch_graph_status() == 0, label
        assert_same_tree(g.tree_arrays(), ref, label)
        g.close()
        checked += 1
    assert checked >= 18


@pytest.mark.parametrize("meshes", [127, 130])
def test_many_roots_graph_equals_eager_and_oracle(pm, orc, scenes, torch, meshes):
    """forests on either side of 128 roots (fewer: the roots' and the first
    levels' box keys are kept as line-separated replicas that the readers
    fold; more: one copy): the eager build, a captured build's replay and the
    oracle give the same tree bit for bit"""
    c = scenes.config4(num_meshes=meshes, rays_per_mesh=1, hi=3000.0)
    off = c["tri_off"]
    tris = frame_tris(scenes, c, 2.0)
    P = [dev(torch, t) for t in tris]
    bp = pm.build_params("reference")
    e = pm.Context(0)
    e.build_batched(*P, off, bp)
    torch.cuda.synchronize()
    eager = {k: np.array(v, copy=True) for k, v in e.tree_arrays().items()}
    e.close()
    g = pm.Context(0)
    g.batch_graph_capture(*P, off, None, None, None, None, pm.SHORT_STACK, bp)
    g.batch_graph_launch()
    assert g.batch_graph_status() == 0
    assert_same_tree(g.tree_arrays(), eager, f"{meshes} meshes: replay vs eager")
    g.close()
    ref = orc.build_batched(tris, off, orc_params(orc, bp))
    assert_same_tree(eager, ref.arrays(), f"{meshes} meshes: eager vs oracle")


def test_copy_back_levels_equal_published_levels(pm, orc, scenes, torch, monkeypatch):
    """PMKD_HOST_MIRROR=0: an eager build that copies each level's state back
    (no coherent pinned memory) gives the tree and counts of one that reads the
    levels' published states, and the oracle's"""
    tris = scenes.cluster_soup(60_000, seed=11)
    P = [dev(torch, t) for t in tris]
    bp = pm.build_params("reference")
    pub = pm.Context(0)
    pub.build(*P, bp)
    monkeypatch.setenv("PMKD_HOST_MIRROR", "0")
    cpy = pm.Context(0)
    cpy.build(*P, bp)
    cpy.build(*P, bp)   # and again on the same context (grown buffers kept)
    monkeypatch.delenv("PMKD_HOST_MIRROR")
    assert_same_tree(cpy.tree_arrays(), pub.tree_arrays(), "copy-back vs published")
    sp, sc_ = pub.stats(), cpy.stats()
    for k in ("upper_levels", "lower_levels", "upper_ref_reads", "upper_ref_writes", "lower_prims"):
        assert sc_[k] == sp[k], k
    ref = orc.build(tris, orc_params(orc, bp)).arrays()
    assert_same_tree(cpy.tree_arrays(), ref, "copy-back vs oracle")
