"""Ray sources on the GPU: camera rays (pm_create_rays, CreateRays/Unproject
IRayTracer.cu:45-70) and photon emission (pm_photon_emit).

Bars: camera rays are the same IEEE binary32 operation sequence as the
oracle's: bit-exact.  Emission uses sqrt/sincos in FP32 against the FP64
restatement: the chosen light (integer) must be the same, origins within
1e-5 absolute, directions within 1e-4 absolute, unit length within 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available()
    return T


@pytest.fixture(scope="module")
def ctx(pm, torch):
    return pm.Context(0)


CASES = [("C1", 256, 256, None), ("C1", 256, 256, 1), ("C2", 1920, 1080, 2), ("C5", 512, 512, 5),
         ("C1", 3, 5, 7), ("C2", 1, 1, None), ("C5", 0, 0, None)]


@pytest.mark.parametrize("name,W,H,seed", CASES)
def test_camera_rays_bit_exact(pm, orc, scenes, torch, ctx, name, W, H, seed):
    cam = {"C1": scenes.CAMERA_C1, "C2": scenes.CAMERA_C2, "C5": scenes.CAMERA_C5}[name]
    m, c = scenes.camera(cam, max(W, 1), max(H, 1))
    O, D = ctx.create_rays(m, c, W, H, jitter_seed=seed)
    torch.cuda.synchronize()
    ro, rd = orc.create_rays(m, c, W, H, jitter_seed=seed)
    assert O.shape[0] == W * H
    assert np.array_equal(O.cpu().numpy().view(np.int32), ro.view(np.int32))
    assert np.array_equal(D.cpu().numpy().view(np.int32), rd.view(np.int32))


def test_camera_rays_trace_like_host_rays(pm, orc, scenes, torch, ctx):
    """C1 at 256 x 256: the device camera rays traced on the GPU give the
    oracle's hits on the oracle's rays, and brute force agrees"""
    tris, _ = scenes.config1()
    m, c = scenes.camera(scenes.CAMERA_C1, 256, 256)
    O, D = ctx.create_rays(m, c, 256, 256, jitter_seed=1)
    P = [torch.from_numpy(t).to("cuda:0") for t in tris]
    ctx.build(*P, pm.build_params("config"))
    tp = pm.trace_params(stack_size=8)
    got = pm.hits_to_numpy(ctx.trace(pm.SHORT_STACK, O, D, tp))
    ro, rd = orc.create_rays(m, c, 256, 256, jitter_seed=1)
    params = orc.build_params("config")
    tree = orc.build(tris, params)
    want = orc.trace(tree, tris, ro, rd, orc.SHORT_STACK, stack_size=8)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))
    assert (got["id"] >= 0).mean() > 0.9


def light_sets(scenes):
    c3 = scenes.cluster_soup(1000, seed=3)
    two = tuple(t[-2:].copy() for t in c3)               # the C3 area light
    rng = np.random.default_rng(12)
    v = rng.uniform(-1, 1, (3, 7, 3))
    v[1] = v[0] + rng.uniform(0.05, 1.5, (7, 3))          # uneven areas
    seven = tuple(np.concatenate([x, np.ones((7, 1))], 1).astype(np.float32) for x in v)
    return {"c3_light": two, "seven_lights": seven}


@pytest.mark.parametrize("which", ["c3_light", "seven_lights"])
def test_emission_matches_fp64(pm, orc, scenes, torch, ctx, which):
    lights = light_sets(scenes)[which]
    n, seed, off = 200_000, 3, 1e-4
    L = [torch.from_numpy(x).to("cuda:0") for x in lights]
    O, D = ctx.photon_emit(L, n, seed, off)
    torch.cuda.synchronize()
    go, gd = O.cpu().numpy(), D.cpu().numpy()
    ro, rd, k = orc.photon_emission64(lights, n, seed, off)
    assert np.all(go[:, 3] == 1) and np.all(gd[:, 3] == 0)
    assert np.abs(go[:, :3] - ro).max() < 1e-5
    assert np.abs(gd[:, :3] - rd).max() < 1e-4
    assert np.abs(np.linalg.norm(gd[:, :3], axis=1) - 1).max() < 1e-5
    # every photon leaves on the emitting side of its own light
    l0, l1, l2 = (x[:, :3].astype(np.float64) for x in lights)
    nrm = np.cross(l1 - l0, l2 - l0)[k]
    assert np.all(np.sum(nrm * gd[:, :3], axis=1) > 0)
    # and the light choice follows the areas
    area = 0.5 * np.linalg.norm(np.cross(l1 - l0, l2 - l0), axis=1)
    freq = np.bincount(k, minlength=len(area)) / n
    assert np.abs(freq - area / area.sum()).max() < 0.01
    # deterministic, and a prefix of a larger call
    O2, D2 = ctx.photon_emit(L, 5000, seed, off)
    assert torch.equal(O2, O[:5000]) and torch.equal(D2, D[:5000])


def test_emission_argument_errors(pm, torch, ctx):
    L = [torch.zeros((1, 4), dtype=torch.float32, device="cuda:0") for _ in range(3)]
    with pytest.raises(pm.PmError):
        ctx.photon_emit([x[:0] for x in L], 10, 1)
    big = [torch.zeros((pm.MAX_LIGHTS + 1, 4), dtype=torch.float32, device="cuda:0") for _ in range(3)]
    with pytest.raises(pm.PmError):
        ctx.photon_emit(big, 10, 1)
    O, D = ctx.photon_emit(L, 0, 1)
    assert O.shape[0] == 0
