"""GPU parity of the optional paths: the DIVIDE split method, the balanced
lower creator, the Woop intersector, the canvas shading and the normal /
colour staging.  Bar: bit-exact against the oracle (trees, hit records,
counters, canvas bytes); the staging against numpy in the same FP32
operation order."""
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


@pytest.fixture(scope="module")
def ctx(pm, torch):
    return pm.Context(0)


BUILDS = {
    "divide_sah": dict(preset="reference", split_method=1),
    "divide_bitmap": dict(preset="default", split_method=1),
    "divide_no_empty": dict(preset="reference", split_method=1, empty_space_splitting=0),
    "balanced": dict(preset="balanced"),
    "balanced_ct4": dict(preset="balanced", traversal_cost=4.0),
    "balanced_divide": dict(preset="balanced", split_method=1),
}


def params_of(pm, spec):
    spec = dict(spec)
    return pm.build_params(spec.pop("preset"), **spec)


@pytest.mark.parametrize("name", list(BUILDS))
def test_builds_bit_exact(pm, orc, scenes, torch, ctx, name):
    params = params_of(pm, BUILDS[name])
    op = orc_params(orc, params)
    for tris, (org, dirs) in ((scenes.config1(width=96, height=96)),
                              (scenes.random_soup(20000, seed=31, size=0.05), scenes.random_rays(20000, seed=32))):
        P = [dev(torch, t) for t in tris]
        ctx.build(*P, params)
        ref = orc.build(tris, op)
        assert_same_tree(ctx.tree_arrays(), ref.arrays(), name)
        O, D = dev(torch, org), dev(torch, dirs)
        for algo in (pm.KD_RESTART, pm.SHORT_STACK):
            h = pm.hits_to_numpy(ctx.trace(algo, O, D, pm.trace_params(stack_size=8)))
            assert_same_hits(h, orc.trace(ref, tris, org, dirs, algo, stack_size=8), f"{name} {algo}")
    # a forest with the same options
    tris, off, org, dirs, ray_mesh = c4_small(scenes, num_meshes=6, empty=(2,))
    P = [dev(torch, t) for t in tris]
    ctx.build_batched(*P, off, params)
    ref = orc.build_batched(tris, off, op)
    assert_same_tree(ctx.tree_arrays(), ref.arrays(), f"{name} forest")


def test_split_method_split_is_not_implemented(pm, scenes, torch, ctx):
    P = [dev(torch, t) for t in scenes.random_soup(100, seed=1)]
    with pytest.raises(pm.PmError) as e:
        ctx.build(*P, pm.build_params("reference", split_method=pm.SPLIT_SPLIT))
    assert e.value.code == 2


@pytest.mark.parametrize("algo_name", ["BRUTE", "KD_RESTART", "SHORT_STACK"])
def test_woop_traces_bit_exact(pm, orc, scenes, torch, ctx, algo_name):
    tris, (org, dirs) = scenes.config1(width=96, height=96)
    algo = getattr(pm, algo_name)
    P = [dev(torch, t) for t in tris]
    ctx.build(*P, pm.build_params("reference"))
    ref = orc.build(tris, orc.build_params("reference"))
    O, D = dev(torch, org), dev(torch, dirs)
    n = org.shape[0] if algo != pm.BRUTE else 4096
    O, D = O[:n].contiguous(), D[:n].contiguous()
    cnt = torch.zeros((n, 4), dtype=torch.int32, device="cuda:0")
    for stack in ((4, 8) if algo == pm.SHORT_STACK else (4,)):
        h = pm.hits_to_numpy(ctx.trace(algo, O, D, pm.trace_params(stack_size=stack, intersection=pm.WOOP),
                                       counters=cnt))
        r, rc = orc.trace(ref if algo != pm.BRUTE else None, tris, org[:n], dirs[:n], algo, stack_size=stack,
                          intersection=orc.WOOP, counters=True)
        assert_same_hits(h, r, f"woop {algo_name} {stack}")
        assert np.array_equal(cnt.cpu().numpy(), rc)
    assert (h["id"] >= 0).mean() > 0.5
    # batched traces take the same option
    btris, off, borg, bdirs, ray_mesh = c4_small(scenes, num_meshes=4)
    if algo != pm.BRUTE:
        BP = [dev(torch, t) for t in btris]
        ctx.build_batched(*BP, off, pm.build_params("reference"))
        bref = orc.build_batched(btris, off, orc.build_params("reference"))
        hb = pm.hits_to_numpy(ctx.trace_batched(algo, dev(torch, borg), dev(torch, bdirs), dev(torch, ray_mesh),
                                                pm.trace_params(stack_size=8, intersection=pm.WOOP)))
        rb = orc.trace_batched(bref, btris, borg, bdirs, ray_mesh, algo, stack_size=8, intersection=orc.WOOP)
        assert_same_hits(hb, rb, f"woop batched {algo_name}")


def shading_scene(scenes):
    """C1 with face normals, opaque coloured walls and a half-transparent sphere"""
    tris, (org, dirs) = scenes.config1(width=64, height=64)
    n = tris[0].shape[0]
    e1 = tris[1][:, :3] - tris[0][:, :3]
    e2 = tris[2][:, :3] - tris[0][:, :3]
    fn = np.cross(e1, e2)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    N = np.zeros((n, 4), np.float32)
    N[:, :3] = fn
    rng = np.random.default_rng(7)
    col = rng.integers(40, 255, (n, 4)).astype(np.uint8)
    col[:24, 3] = 255                 # the room is opaque
    col[24:, 3] = 128                 # the sphere reflects
    light = np.array([0.0, 0.9, 0.3, 0.1, 0.1, 0.1, 0.7, 0.7, 0.7, 0.5, 0.5, 0.5], np.float32)
    return tris, (org, dirs), (N, N.copy(), N.copy()), col, light


@pytest.mark.parametrize("algo_name", ["BRUTE", "KD_RESTART", "SHORT_STACK"])
@pytest.mark.parametrize("woop", [False, True])
def test_canvas_bit_exact(pm, orc, scenes, torch, ctx, algo_name, woop):
    tris, (org, dirs), normals, col, light = shading_scene(scenes)
    algo = getattr(pm, algo_name)
    P = [dev(torch, t) for t in tris]
    ctx.build(*P, pm.build_params("reference"))
    ref = orc.build(tris, orc.build_params("reference"))
    n = org.shape[0] if algo != pm.BRUTE else 1024
    tp = pm.trace_params(stack_size=8, intersection=pm.WOOP if woop else pm.MOELLER)
    got = ctx.trace_canvas(algo, dev(torch, org[:n]), dev(torch, dirs[:n]), [dev(torch, x) for x in normals],
                           dev(torch, col), light, params=tp).cpu().numpy()
    want = orc.trace_canvas(ref if algo != pm.BRUTE else None, tris, org[:n], dirs[:n], algo, normals, col, light,
                            stack_size=8, intersection=orc.WOOP if woop else orc.MOELLER)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:5], got[bad[:3]], want[bad[:3]])
    assert (want[:, 3] > 0).mean() > 0.5 and len(np.unique(want[:, 1])) > 20


def test_mesh_positions_under_a_general_model_matrix(pm, torch, ctx):
    """AddMeshKernel's p = M v (GeometryList.cu:197-232) with a general 4x4
    row-major matrix: every gathered corner equals the FP32 row products
    (x, y, z, w summed left to right) bit for bit, and the FP64 product within
    FP32 rounding"""
    rng = np.random.default_rng(8)
    nv, nt = 700, 1500
    verts = np.concatenate([rng.uniform(-3, 3, (nv, 3)), rng.uniform(0.5, 2, (nv, 1))], 1).astype(np.float32)
    idx = rng.integers(0, nv, 3 * nt).astype(np.int32)
    model = rng.normal(size=(4, 4)).astype(np.float32)
    P = [torch.zeros((nt + 3, 4), dtype=torch.float32, device="cuda:0") for _ in range(3)]
    ctx.add_mesh(dev(torch, verts), dev(torch, idx), *P, offset=3, model=model)
    tri = idx.reshape(-1, 3)
    for k in range(3):
        v = verts[tri[:, k]]
        want = np.zeros((nt, 4), np.float32)
        for i in range(4):
            r = model[i]
            want[:, i] = ((v[:, 0] * r[0] + v[:, 1] * r[1]) + v[:, 2] * r[2]) + v[:, 3] * r[3]
        got = P[k].cpu().numpy()
        assert np.array_equal(got[3:].view(np.int32), want.view(np.int32))
        assert not got[:3].any()
        w64 = v.astype(np.float64) @ model.astype(np.float64).T
        assert np.abs(got[3:] - w64).max() <= 1e-5 * max(1.0, np.abs(w64).max())


def test_mesh_attributes_match_numpy(pm, torch, ctx):
    rng = np.random.default_rng(3)
    nv, nt = 500, 800
    normals = rng.normal(size=(nv, 4)).astype(np.float32)
    colors = rng.integers(0, 256, (nv, 4)).astype(np.uint8)
    idx = rng.integers(0, nv, 3 * nt).astype(np.int32)
    model = rng.normal(size=(4, 4)).astype(np.float32)
    out = [torch.zeros((nt + 5, 4), dtype=torch.float32, device="cuda:0") for _ in range(3)]
    cout = [torch.zeros((nt + 5, 4), dtype=torch.uint8, device="cuda:0") for _ in range(3)]
    ctx.add_mesh_attributes(dev(torch, normals), dev(torch, idx), *out, colors=dev(torch, colors),
                            c0=cout[0], c1=cout[1], c2=cout[2], offset=5, model=model)
    tri = idx.reshape(-1, 3)
    for k in range(3):
        v = normals[tri[:, k], :3]
        want = np.zeros((nt, 4), np.float32)
        for i in range(3):
            r = model[i, :3]
            want[:, i] = (v[:, 0] * r[0] + v[:, 1] * r[1]) + v[:, 2] * r[2]
        assert np.array_equal(out[k].cpu().numpy()[5:], want)
        assert np.array_equal(cout[k].cpu().numpy()[5:], colors[tri[:, k]])
        assert not out[k].cpu().numpy()[:5].any()
    bad = idx.copy()
    bad[7] = nv + 3
    with pytest.raises(pm.PmError):
        ctx.add_mesh_attributes(dev(torch, normals), dev(torch, bad), *out)
