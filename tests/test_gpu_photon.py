"""Photon bounce driver (pm_photon_bounce) against the FP64 numpy
restatement in scenes.bounce_rays on the same hits.

Survivor selection and order are integer work: identical.  Positions and
directions are FP32 on the GPU (sqrt, sincos) against FP64 in numpy:
positions within 1e-5 absolute, directions within 1e-4 absolute, unit
length within 1e-5, and every direction on the incoming side of the
geometric normal (cosine-weighted hemisphere)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bounce_matches_numpy(pm, scenes):
    import torch
    tris = scenes.cluster_soup(50_000, seed=3)
    org, dirs = scenes.photon_emission(100_000, seed=3)
    dev = "cuda:0"
    P = [torch.from_numpy(t).to(dev) for t in tris]
    O, D = torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)
    ctx = pm.Context(0)
    ctx.build(*P, pm.build_params("reference"))
    H = ctx.trace(pm.SHORT_STACK, O, D, pm.trace_params(stack_size=8))
    hits = pm.hits_to_numpy(H)
    absorb = tris[0].shape[0] - 2
    seed = 3 + 1000
    o2, d2, src, n = ctx.photon_bounce(O, D, H, absorb, seed)
    torch.cuda.synchronize()
    keep = (hits["id"] >= 0) & (hits["id"] < absorb)
    assert n == int(keep.sum()) and n > 1000
    # numpy restatement (float64) on the same hits; it keeps every hit, so
    # select the non-absorbed ones (stream index = source ray index)
    ro, rd, alive = scenes.bounce_rays(tris, org, dirs, hits, seed=3, bounce=1)
    sel = np.isin(alive, np.nonzero(keep)[0])
    ro, rd, alive = ro[sel], rd[sel], alive[sel]
    assert np.array_equal(src[:n].cpu().numpy(), alive)
    go, gd = o2[:n].cpu().numpy(), d2[:n].cpu().numpy()
    assert np.abs(go[:, :3] - ro[:, :3]).max() < 1e-5
    assert np.abs(gd[:, :3] - rd[:, :3]).max() < 1e-4
    assert np.abs(np.linalg.norm(gd[:, :3], axis=1) - 1).max() < 1e-5
    # continuation leaves on the incoming side: d_in . d_out < 0 w.r.t. the surface
    tri = hits["id"][alive]
    p0, p1, p2 = (t[tri, :3].astype(np.float64) for t in tris)
    nrm = np.cross(p1 - p0, p2 - p0)
    side_in = np.sign(np.sum(nrm * dirs[alive, :3], axis=1))
    side_out = np.sign(np.sum(nrm * gd[:, :3], axis=1))
    assert np.mean(side_in == side_out) < 1e-3
    # second call with fewer rays reuses scratch and gives a prefix-consistent result
    o3, d3, src3, n3 = ctx.photon_bounce(O[:5000].contiguous(), D[:5000].contiguous(),
                                         H[:5000].contiguous(), absorb, seed)
    assert n3 == int(keep[:5000].sum())
    assert np.array_equal(src3[:n3].cpu().numpy(), src[:n3].cpu().numpy())
    assert torch.equal(o3[:n3], o2[:n3]) and torch.equal(d3[:n3], d2[:n3])
    ctx.close()


def test_bounce_compaction_beyond_1024_scan_blocks(pm, scenes):
    """6,000,000 rays: the survivor scan spans more than 1,024 blocks of
    4,096 flags, so it takes the recursive path (block scans, a scan of the
    block sums, an add).  Survivor count and order are integer work: exact."""
    import torch
    tris = scenes.cluster_soup(1000, seed=5)
    ntris = tris[0].shape[0]
    absorb = ntris - 2
    n = 6_000_000
    rng = np.random.default_rng(11)
    org = np.zeros((n, 4), np.float32)
    org[:, :3] = rng.uniform(-1, 1, (n, 3))
    dirs = np.zeros((n, 4), np.float32)
    d = rng.normal(size=(n, 3))
    dirs[:, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    ids = rng.integers(-1, ntris, n).astype(np.int32)
    hits = np.zeros((n, 4), np.float32)
    hits[:, 0] = ids.view(np.float32)
    hits[:, 1] = 0.5
    hits[:, 2] = 0.25
    hits[:, 3] = 0.25
    dev = "cuda:0"
    P = [torch.from_numpy(t).to(dev) for t in tris]
    ctx = pm.Context(0)
    ctx.set_geometry(*P)
    O, D, H = (torch.from_numpy(a).to(dev) for a in (org, dirs, hits))
    o2, d2, src, cnt = ctx.photon_bounce(O, D, H, absorb, 77)
    torch.cuda.synchronize()
    keep = np.nonzero((ids >= 0) & (ids < absorb))[0]
    assert cnt == keep.size
    assert np.array_equal(src[:cnt].cpu().numpy(), keep)
    assert torch.isfinite(o2[:cnt]).all() and torch.isfinite(d2[:cnt]).all()
    ctx.close()


def test_photon_driver_passes(pm, orc, scenes):
    """the photon driver (photons.PhotonTracer) runs emission + bounce passes:
    each pass's hits equal a direct pm_trace of that pass's rays, in either
    ray order (and the oracle's for the emission pass); each pass continues
    exactly the previous pass's hits below absorb_from; tuning picks one
    valid order per pass"""
    import importlib
    import torch
    photons = importlib.import_module("extensions-photonmapping_amd.photons")
    tris = scenes.cluster_soup(100_000, seed=3)
    dev = "cuda:0"
    P = [torch.from_numpy(t).to(dev) for t in tris]
    ntris = tris[0].shape[0]
    ctx = pm.Context(0)
    ctx.build(*P, pm.build_params("reference"))
    lights = [p[ntris - 2:].contiguous() for p in P]
    drv = photons.PhotonTracer(ctx, lights, ntris - 2, stack_size=8)
    n = 200_000          # above the coherent order's threshold
    ref = None
    for orders in ([pm.RAY_ORDER_COHERENT], [pm.RAY_ORDER_INPUT], [pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT]):
        passes = drv.trace(n, 3, seed=3, orders=orders)
        assert passes[0].count == n and len(passes) == 4
        got = [pm.hits_to_numpy(p.hits).copy() for p in passes]
        if ref is None:
            ref = got
            tree = orc.build(tris, orc.build_params("reference"))
            o0, d0 = passes[0].org.cpu().numpy(), passes[0].dirs.cpu().numpy()
            want = orc.trace(tree, tris, o0, d0, orc.SHORT_STACK, stack_size=8, threads=8)
            assert np.array_equal(got[0].view(np.int32), want.view(np.int32))
        for a, b in zip(got, ref):
            assert np.array_equal(a.view(np.int32), b.view(np.int32)), orders
        for b in range(1, len(passes)):
            prev = got[b - 1]
            keep = np.nonzero((prev["id"] >= 0) & (prev["id"] < ntris - 2))[0]
            assert passes[b].count == keep.shape[0]
            assert np.array_equal(passes[b].src.cpu().numpy(), keep)
            direct = ctx.trace(pm.SHORT_STACK, passes[b].org, passes[b].dirs, pm.trace_params(stack_size=8))
            assert torch.equal(direct.view(torch.int32), passes[b].hits.view(torch.int32))
    best = drv.tune_ray_orders(n, 3, seed=3, reps=1)
    assert len(best) == 4 and set(best) <= {pm.RAY_ORDER_INPUT, pm.RAY_ORDER_COHERENT}
    timed = drv.trace(n, 3, seed=3, timed=True)
    assert all(p.ms > 0 for p in timed) and drv.emit_ms > 0
    assert [p.order for p in timed] == best
    with pytest.raises(ValueError):
        drv.set_ray_orders([7])
    ctx.close()
