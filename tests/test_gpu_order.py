"""Coherent ray order (trace_params.ray_order): the tree tracers may run a
large trace in the order of a spatial key; every ray's record and counters
must be the ones the input order gives, bit for bit, and the oracle's.

Covers: the photon workload's shape (clustered soup, emission rays and
diffuse bounce rays) with both tree tracers, degenerate rays (zero
directions, origins far outside the scene), forests with rays of empty and
out-of-range meshes, the captured batch, and the threshold below which the
input order is kept.
"""
import numpy as np
import pytest

from forest_util import c4_small
from test_gpu_parity import assert_same_hits, assert_same_tree, dev, orc_params

pytestmark = pytest.mark.gpu

BIG = 1 << 17      # traces of at least this many rays are reordered


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "these tests need the GPU"
    return T


def trace_both(pm, ctx, torch, algo, O, D, RM=None, stack=8):
    """hits and counters in input order and in coherent order"""
    out = []
    for order in (pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT):
        cnt = torch.zeros((O.shape[0], 4), dtype=torch.int32, device="cuda:0")
        tp = pm.trace_params(stack_size=stack, ray_order=order)
        if RM is None:
            h = ctx.trace(algo, O, D, tp, counters=cnt)
        else:
            h = ctx.trace_batched(algo, O, D, RM, tp, counters=cnt)
        out.append((pm.hits_to_numpy(h), cnt.cpu().numpy()))
    return out


def test_photon_passes_same_records(pm, orc, scenes, torch):
    tris = scenes.cluster_soup(200000)
    org, dirs = scenes.photon_emission(300000)
    # degenerate rays: zero directions, origins far outside the scene
    dirs[:64, :3] = 0.0
    org[64:128, :3] = 1e6
    P = [dev(torch, t) for t in tris]
    ctx = pm.Context(0)
    bp = pm.build_params("reference")
    ctx.build(*P, bp)
    ref = orc.build(tris, orc.build_params("reference"))
    O, D = dev(torch, org), dev(torch, dirs)
    ntris = tris[0].shape[0]
    for algo in (pm.SHORT_STACK, pm.KD_RESTART):
        (hi, ci), (hc, cc) = trace_both(pm, ctx, torch, algo, O, D)
        assert_same_hits(hc, hi, f"emission {algo}")
        assert np.array_equal(cc, ci)
        r, rc = orc.trace(ref, tris, org, dirs, algo, stack_size=8, counters=True, threads=8)
        assert_same_hits(hc, r, f"emission {algo} oracle")
        assert np.array_equal(cc, rc)
    assert (hc["id"] >= 0).mean() > 0.5
    # one diffuse bounce: rays spread over the clusters
    H = ctx.trace(pm.SHORT_STACK, O, D, pm.trace_params(stack_size=8))
    bo, bd, _, n = ctx.photon_bounce(O, D, H, ntris - 2, seed=11)
    assert n >= BIG
    bo, bd = bo[:n].contiguous(), bd[:n].contiguous()
    (hi, ci), (hc, cc) = trace_both(pm, ctx, torch, pm.SHORT_STACK, bo, bd)
    assert_same_hits(hc, hi, "bounce")
    assert np.array_equal(cc, ci)
    r = orc.trace(ref, tris, bo.cpu().numpy(), bd.cpu().numpy(), orc.SHORT_STACK, stack_size=8, threads=8)
    assert_same_hits(hc, r, "bounce oracle")
    ctx.close()


def test_forest_same_records(pm, orc, scenes, torch):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=8, rays=20000, empty=(3,))
    ray_mesh = ray_mesh.copy()
    ray_mesh[::97] = 9                   # no such mesh
    ray_mesh[1::89] = -1
    ray_mesh[2::83] = 3                  # the empty mesh
    assert org.shape[0] >= BIG
    P = [dev(torch, t) for t in tris]
    ctx = pm.Context(0)
    bp = pm.build_params("reference")
    ctx.build_batched(*P, off, bp)
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, ray_mesh)
    ref = orc.build_batched(tris, off, orc.build_params("reference"))
    for algo in (pm.SHORT_STACK, pm.KD_RESTART):
        (hi, ci), (hc, cc) = trace_both(pm, ctx, torch, algo, O, D, RM)
        assert_same_hits(hc, hi, f"forest {algo}")
        assert np.array_equal(cc, ci)
        r = orc.trace_batched(ref, tris, org, dirs, ray_mesh, algo, stack_size=8)
        assert_same_hits(hc, r, f"forest {algo} oracle")
    assert (hc["id"][ray_mesh == 9] == -1).all() and (hc["id"][ray_mesh == 3] == -1).all()
    ctx.close()


def test_graph_with_coherent_order(pm, scenes, torch):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=8, rays=20000)
    P = [dev(torch, t) for t in tris]
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, ray_mesh)
    H = torch.empty((O.shape[0], 4), dtype=torch.float32, device="cuda:0")
    bp = pm.build_params("reference")
    g = pm.Context(0)
    g.batch_graph_capture(*P, off, O, D, RM, H, pm.SHORT_STACK, bp,
                          pm.trace_params(stack_size=8, ray_order=pm.RAY_ORDER_COHERENT))
    eager = pm.Context(0)
    eager.build_batched(*P, off, bp)
    want = pm.hits_to_numpy(eager.trace_batched(pm.SHORT_STACK, O, D, RM,
                                                pm.trace_params(stack_size=8, ray_order=pm.RAY_ORDER_INPUT)))
    for i in range(3):
        H.fill_(0.0)
        g.batch_graph_launch()
        assert g.batch_graph_status() == 0
        assert_same_hits(pm.hits_to_numpy(H), want, f"replay {i}")
    g.close()
    eager.close()


def test_small_traces_and_bad_order_value(pm, scenes, torch):
    tris = scenes.random_soup(5000, seed=5)
    org, dirs = scenes.random_rays(BIG - 1, seed=6)
    P = [dev(torch, t) for t in tris]
    ctx = pm.Context(0)
    ctx.build(*P, pm.build_params("reference"))
    O, D = dev(torch, org), dev(torch, dirs)
    (hi, ci), (hc, cc) = trace_both(pm, ctx, torch, pm.SHORT_STACK, O, D)
    assert_same_hits(hc, hi, "below the threshold")
    with pytest.raises(pm.PmError) as e:
        ctx.trace(pm.SHORT_STACK, O, D, pm.trace_params(ray_order=2))
    assert e.value.code == 1
    ctx.close()


def test_autotune_times_both_orders(pm, orc, scenes, torch):
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=8, rays=20000)
    P = [dev(torch, t) for t in tris]
    O, D, RM = dev(torch, org), dev(torch, dirs), dev(torch, ray_mesh)
    ctx = pm.Context(0)
    res = ctx.autotune(*P, O, D, mesh_offsets=off, ray_mesh=RM, base=pm.build_params("reference"), reps=1)
    assert res["ray_order"] in (pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT)
    # 4 chunk sizes, 1..6 leaf thresholds, 3 stack depths x 2 orders x 2 step caps
    assert 4 + 1 + 12 <= res["trials"] <= 4 + 6 + 12
    bp, tp = res["build_params"], res["trace_params"]
    assert tp.ray_order == res["ray_order"] and tp.stack_size == res["stack_size"]
    assert tp.step_cap == res["step_cap"] and res["step_cap"] in (0, 4)
    ref = orc.build_batched(tris, off, orc_params(orc, bp))
    assert_same_tree(ctx.tree_arrays(), ref.arrays(), "autotuned forest")
    h = pm.hits_to_numpy(ctx.trace_batched(pm.SHORT_STACK, O, D, RM, tp))
    r = orc.trace_batched(ref, tris, org, dirs, ray_mesh, orc.SHORT_STACK, stack_size=tp.stack_size)
    assert_same_hits(h, r, "autotuned trace")
    ctx.close()


def test_step_cap_same_records(pm, orc, scenes, torch):
    """pm_trace_params.step_cap spreads a descent over rounds of the
    cooperative leaf test: records and counters equal the uncapped trace (and
    the oracle's) for every cap, both tree tracers, both ray orders"""
    tris = scenes.heightfield(256, seed=2)
    W, Hh = 512, 288                     # 147,456 rays: the coherent order applies
    m, c = scenes.camera(scenes.CAMERA_C2, W, Hh)
    ctx = pm.Context(0)
    P = [dev(torch, t) for t in tris]
    ctx.build(*P, pm.build_params("reference"))
    O, D = ctx.create_rays(m, c, W, Hh, jitter_seed=2)
    org, dirs = O.cpu().numpy(), D.cpu().numpy()
    ref = orc.build(tris, orc.build_params("reference"))
    for algo in (pm.SHORT_STACK, pm.KD_RESTART):
        want, wc = orc.trace(ref, tris, org, dirs, algo, stack_size=8, counters=True, threads=8)
        for order in (pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT):
            for cap in (0, 1, 2, 4, 16):
                cnt = torch.zeros((O.shape[0], 4), dtype=torch.int32, device="cuda:0")
                tp = pm.trace_params(stack_size=8, ray_order=order, step_cap=cap)
                h = pm.hits_to_numpy(ctx.trace(algo, O, D, tp, counters=cnt))
                assert_same_hits(h, want, f"algo={algo} order={order} cap={cap}")
                assert np.array_equal(cnt.cpu().numpy(), wc), (algo, order, cap)
    with pytest.raises(pm.PmError):
        ctx.trace(pm.SHORT_STACK, O, D, pm.trace_params(step_cap=-1))
    ctx.close()
