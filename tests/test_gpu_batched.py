"""GPU batched build and trace (one kd-tree per mesh, built together as a
forest) and the per-frame deformers that feed it.

Bars: the GPU forest equals the oracle forest bit for bit in every array,
and batched traces equal the oracle's forest traces bit for bit (the oracle
forest itself is pinned mesh by mesh against single-mesh builds on the CPU).
Deformers: within 2e-6 absolute of an FP64 evaluation (FP32 sin, cos, acos
and atan2 are not correctly rounded, so bit equality is not the bar there).
"""
import numpy as np
import pytest

from forest_util import TREE_KEYS, c4_small
from test_gpu_parity import assert_same_hits, assert_same_tree, dev, make_params, orc_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "these tests need the GPU"
    return T


@pytest.fixture(scope="module")
def ctx(pm, torch):
    return pm.Context(0)


def gpu_forest(pm, ctx, torch, tris, off, params):
    P = [dev(torch, t) for t in tris]
    ctx.build_batched(*P, off, params)
    torch.cuda.synchronize()
    return P, ctx.tree_arrays()


STAT_KEYS = ("num_nodes", "reachable_nodes", "inner_nodes", "leaves", "empty_leaves", "max_depth",
             "leaf_refs", "max_leaf_size", "upper_levels", "lower_levels", "num_prim_indices",
             "num_upper_leaves")


@pytest.mark.parametrize("preset", ["reference", "default", "config"])
def test_forest_tree_and_traces_bit_exact(pm, orc, scenes, torch, ctx, preset):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, empty=(0, 4, 11))
    params = make_params(pm, {"preset": preset})
    P, got = gpu_forest(pm, ctx, torch, tris, off, params)
    ref = orc.build_batched(tris, off, orc_params(orc, params))
    assert got["num_roots"] == len(off) - 1
    assert_same_tree(got, ref.arrays(), f"forest {preset}")
    st, rs = ctx.stats(), ref.stats()
    for k in STAT_KEYS:
        assert st[k] == rs[k], (k, st[k], rs[k])
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, ray_mesh)
    for algo, stack in ((pm.KD_RESTART, 4), (pm.SHORT_STACK, 4), (pm.SHORT_STACK, 8)):
        h = pm.hits_to_numpy(ctx.trace_batched(algo, O, D, RM, pm.trace_params(stack_size=stack)))
        r = orc.trace_batched(ref, tris, org, dirs, ray_mesh, algo, stack_size=stack)
        assert_same_hits(h, r, f"{preset} algo {algo} stack {stack}")
    assert (h["id"] >= 0).mean() > 0.95


def test_forest_counters_and_missing_meshes(pm, orc, scenes, torch, ctx):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=5, empty=(3,))
    params = pm.build_params("reference")
    P, got = gpu_forest(pm, ctx, torch, tris, off, params)
    ref = orc.build_batched(tris, off, orc_params(orc, params))
    # a ray whose mesh index is not one of the forest's misses
    rm = ray_mesh.copy()
    rm[::97] = -1
    rm[1::89] = len(off) - 1 + 5
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, rm)
    cnt = torch.zeros((org.shape[0], 4), dtype=torch.int32, device="cuda:0")
    h = pm.hits_to_numpy(ctx.trace_batched(pm.SHORT_STACK, O, D, RM, pm.trace_params(stack_size=8),
                                           counters=cnt))
    r, rc = orc.trace_batched(ref, tris, org, dirs, rm, orc.SHORT_STACK, stack_size=8, counters=True)
    assert_same_hits(h, r, "counters run")
    assert np.array_equal(cnt.cpu().numpy(), rc)
    bad = (rm < 0) | (rm >= len(off) - 1)
    assert (h["id"][bad] == -1).all() and (cnt.cpu().numpy()[bad] == 0).all()
    with pytest.raises(pm.PmError):      # no brute force over a forest
        ctx.trace_batched(pm.BRUTE, O, D, RM)


def test_single_mesh_batch_is_the_plain_build(pm, orc, scenes, torch, ctx):
    tris, off, org, dirs, _ = c4_small(scenes, num_meshes=1)
    params = pm.build_params("reference")
    P = [dev(torch, t) for t in tris]
    ctx.build(*P, params)
    plain = ctx.tree_arrays()
    ctx.build_batched(*P, [0, tris[0].shape[0]], params)
    batched = ctx.tree_arrays()
    for k in TREE_KEYS:
        assert np.array_equal(plain[k].view(np.uint8), batched[k].view(np.uint8)), k
    # switching back and forth in one context reuses the scratch correctly
    tris2, off2, *_ = c4_small(scenes, num_meshes=7)
    P2, got = gpu_forest(pm, ctx, torch, tris2, off2, params)
    assert_same_tree(got, orc.build_batched(tris2, off2, orc_params(orc, params)).arrays(), "forest 2")
    ctx.build(*P, params)
    assert_same_tree(ctx.tree_arrays(), plain, "plain again")
    assert ctx.tree_arrays()["num_roots"] == 1


def test_forest_argument_errors(pm, scenes, torch, ctx):
    tris, off, *_ = c4_small(scenes, num_meshes=3)
    P = [dev(torch, t) for t in tris]
    n = tris[0].shape[0]
    for bad in ([0, 5, 3, n], [1, 5, 10, n], [0, 5, 10, n - 1]):
        with pytest.raises(pm.PmError):
            ctx.build_batched(*P, np.asarray(bad, np.int32))
    # a forest cannot be shaded into one canvas (every ray would start at root 0)
    ctx.build_batched(*P, off)
    O, D = dev(torch, np.zeros((4, 4), np.float32)), dev(torch, np.ones((4, 4), np.float32))
    N = [torch.zeros_like(P[0]) for _ in range(3)]
    col = torch.zeros((n, 4), dtype=torch.uint8, device="cuda:0")
    with pytest.raises(pm.PmError, match="forest"):
        ctx.trace_canvas(pm.SHORT_STACK, O, D, N, col, [0.0] * 12)
    # photons absorbed from an id past the geometry: rejected before any kernel runs
    H = torch.zeros((4, 4), dtype=torch.float32, device="cuda:0")
    with pytest.raises(pm.PmError, match="absorb_from"):
        ctx.photon_bounce(O, D, H, n + 1, seed=1)


def test_radial_deformer_matches_fp64(pm, scenes, torch, ctx):
    c = scenes.config4(num_meshes=6, rays_per_mesh=1, hi=2e4)
    U = dev(torch, c["unit"])
    for frame in (0.0, 3.0, 15.0):
        got = ctx.deform_radial(U, frame).cpu().numpy()
        ref = scenes.displace_sphere64(c["unit"], frame)
        assert np.abs(got[:, :3] - ref).max() <= 2e-6
        assert np.array_equal(got[:, 3], c["unit"][:, 3])


def test_osc_surface_matches_fp64(pm, scenes, torch, ctx):
    detail, height = 63, 0.5
    V, idx = scenes.osc_grid(4.0, 3.0, height, detail)
    Vd = dev(torch, V)
    Nd = torch.empty_like(Vd)
    for time in (0.0, 0.37, 2.5):
        ctx.osc_surface(Vd, Nd, detail, height, time)
        ref_v, ref_n = scenes.osc_surface64(V, detail, height, time)
        got_v, got_n = Vd.cpu().numpy(), Nd.cpu().numpy()
        assert np.array_equal(got_v[:, [0, 2, 3]], V[:, [0, 2, 3]])   # only y moves
        assert np.abs(got_v[:, 1] - ref_v[:, 1]).max() <= 2e-6
        assert np.abs(got_n - ref_n).max() <= 2e-6
    # the grid's index buffer stages through add_mesh like any mesh
    n = idx.shape[0] // 3
    p = [torch.empty((n, 4), dtype=torch.float32, device="cuda:0") for _ in range(3)]
    ctx.add_mesh(Vd, dev(torch, idx.astype(np.int32)), *p)
    got_v = Vd.cpu().numpy()
    tri = idx.reshape(-1, 3)
    assert np.array_equal(p[1].cpu().numpy(), got_v[tri[:, 1]])


def test_c4_frame_pipeline_bit_exact(pm, orc, scenes, torch, ctx):
    """one C4 frame at full size (256 meshes, ~10.3M triangles): GPU deform
    -> gather -> batched build -> batched trace, against the oracle forest
    built from the same (GPU-deformed) triangles; traces on 1,024 rays per
    mesh"""
    c = scenes.config4(num_meshes=256, rays_per_mesh=1024)
    U = dev(torch, c["unit"])
    faces = dev(torch, c["faces"].reshape(-1))
    n = int(c["tri_off"][-1])
    V = ctx.deform_radial(U, 7.0)
    P = [torch.empty((n, 4), dtype=torch.float32, device="cuda:0") for _ in range(3)]
    ctx.add_mesh(V, faces, *P)
    params = pm.build_params("reference")
    ctx.build_batched(*P, c["tri_off"], params)
    torch.cuda.synchronize()
    got = ctx.tree_arrays()
    tris = tuple(p.cpu().numpy() for p in P)
    ref = orc.build_batched(tris, c["tri_off"], orc_params(orc, params))
    assert_same_tree(got, ref.arrays(), "C4 frame")
    O, D, RM = dev(torch, c["org"]), dev(torch, c["dirs"]), dev(torch, c["ray_mesh"])
    h = pm.hits_to_numpy(ctx.trace_batched(pm.SHORT_STACK, O, D, RM, pm.trace_params(stack_size=8)))
    r = orc.trace_batched(ref, tris, c["org"], c["dirs"], c["ray_mesh"], orc.SHORT_STACK, stack_size=8)
    assert_same_hits(h, r, "C4 frame short-stack")
    assert (h["id"] >= 0).mean() > 0.95


@pytest.mark.parametrize("chunk", [128, 256, 512, 2048, 4096])
def test_chunk_size_changes_the_work_not_the_tree(pm, orc, scenes, torch, ctx, chunk):
    tris = scenes.random_soup(60000, seed=11)
    params = pm.build_params("reference", chunk_size=chunk)
    P = [dev(torch, t) for t in tris]
    ctx.build(*P, params)
    assert_same_tree(ctx.tree_arrays(), orc.build(tris, orc_params(orc, params)).arrays(), f"chunk {chunk}")
    btris, off, *_ = c4_small(scenes, num_meshes=6, empty=(1,))
    P2, got = gpu_forest(pm, ctx, torch, btris, off, params)
    assert_same_tree(got, orc.build_batched(btris, off, orc_params(orc, params)).arrays(), f"forest chunk {chunk}")
    with pytest.raises(pm.PmError):
        ctx.build(*P, pm.build_params("reference", chunk_size=300))


def test_autotune_picks_a_valid_setting(pm, orc, scenes, torch, ctx):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=8)
    P = [dev(torch, t) for t in tris]
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, ray_mesh)
    res = ctx.autotune(*P, O, D, mesh_offsets=off, ray_mesh=RM, base=pm.build_params("reference"),
                       reps=1)
    assert res["chunk_size"] in (256, 512, 1024, 2048)
    assert res["lower_size"] in (8, 16, 24, 32, 48, 64)
    assert res["stack_size"] in (4, 8, 16)
    assert res["ray_order"] == pm.RAY_ORDER_INPUT          # too few rays to reorder
    # leaf thresholds the meshes cannot reach are skipped, the rest are timed
    assert res["step_cap"] in (0, 4)
    # 4 chunk sizes, 1..6 leaf thresholds, 3 stack depths x 2 step caps
    assert 4 + 1 + 6 <= res["trials"] <= 4 + 6 + 6 and res["build_ms"] > 0 and res["trace_ms"] > 0
    # the context holds the chosen tree; tracing it with the chosen depth is
    # the oracle's trace of the oracle's tree with the same parameters
    bp, tp = res["build_params"], res["trace_params"]
    ref = orc.build_batched(tris, off, orc_params(orc, bp))
    assert_same_tree(ctx.tree_arrays(), ref.arrays(), "autotuned forest")
    h = pm.hits_to_numpy(ctx.trace_batched(pm.SHORT_STACK, O, D, RM, tp))
    r = orc.trace_batched(ref, tris, org, dirs, ray_mesh, orc.SHORT_STACK, stack_size=tp.stack_size)
    assert_same_hits(h, r, "autotuned trace")
    # one mesh, no mesh indices
    stris = scenes.random_soup(20000, seed=12)
    sorg, sdirs = scenes.random_rays(20000, seed=13)
    SP = [dev(torch, t) for t in stris]
    res1 = ctx.autotune(*SP, dev(torch, sorg), dev(torch, sdirs), reps=1)
    assert_same_tree(ctx.tree_arrays(), orc.build(stris, orc_params(orc, res1["build_params"])).arrays(),
                     "autotuned single mesh")
