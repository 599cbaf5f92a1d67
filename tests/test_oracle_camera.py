"""Pins the oracle's camera rays (CreateRays / Unproject, IRayTracer.cu:45-70)
and the FP64 restatement of the photon emission without the reference.

Camera: with the identity matrix and the camera at the origin the ray of
pixel (x, y) points along (s.x, s.y, 1), s = (x / W) * 2 - 1, so directions
have closed forms (the centre pixel is exactly +z, a corner is
(-1, -1, 1)/sqrt(3)); a translation in the last four matrix entries moves the
far point; a real perspective camera agrees with an independent FP64 look-at
construction (scenes.pinhole_rays), which fails if the matrix is read
transposed; the jitter is the PCG32 stream pinned by its known answer."""
import numpy as np


def ulp_close(a, b, ulps=2):
    a = np.asarray(a, np.float32)
    return np.all(np.abs(a.astype(np.float64) - b) <= ulps * np.spacing(np.abs(a)).astype(np.float64) + 1e-45)


def test_identity_camera_closed_forms(orc):
    W, H = 64, 32
    eye = np.eye(4, dtype=np.float32).reshape(16)
    org, d = orc.create_rays(eye, [0, 0, 0], W, H)
    assert org.shape == (W * H, 4) and np.all(org == [0, 0, 0, 1]) and np.all(d[:, 3] == 0)
    assert np.array_equal(d[(H // 2) * W + W // 2], np.float32([0, 0, 1, 0]))
    assert np.array_equal(d[0, :3], np.float32([-1, -1, 1]) * np.float32(1.0 / np.sqrt(np.float32(3.0))))
    x = np.arange(W * H) % W
    y = np.arange(W * H) // W
    s = np.stack([x / W * 2.0 - 1.0, y / H * 2.0 - 1.0, np.ones(W * H)], axis=1)
    want = s / np.linalg.norm(s, axis=1, keepdims=True)
    assert ulp_close(d[:, :3], want, 2)
    # x grows along +x, rows along +y, no half-pixel offset: s.x / s.z = x / W * 2 - 1 exactly
    assert np.all(np.abs(d[:, 0] / d[:, 2] - (x / W * 2 - 1)) < 1e-6)


def test_translation_and_w_divide(orc):
    m = np.eye(4, dtype=np.float64)
    m[:3, 3] = [0.5, -2.0, 4.0]          # column 3 = entries 12..14 in Unproject's order
    m[3, 3] = 2.0                        # (sx, sy, 1, 1) -> w = 2: the far point is divided by 2
    flat = np.ascontiguousarray(m.T.reshape(16), np.float32)
    cam = np.float32([0.25, -1.0, 2.0])
    org, d = orc.create_rays(flat, cam, 16, 16)
    x = np.arange(256) % 16
    y = np.arange(256) // 16
    end = np.stack([(x / 16 * 2 - 1 + 0.5) / 2, (y / 16 * 2 - 1 - 2.0) / 2, (1 + 4.0) / 2 * np.ones(256)], 1)
    v = end - cam.astype(np.float64)
    want = v / np.linalg.norm(v, axis=1, keepdims=True)
    assert np.abs(d[:, :3] - want).max() < 1e-6
    assert np.all(org[:, :3] == cam)


def test_perspective_camera_matches_fp64_look_at(orc, scenes):
    for cfg, (W, H) in ((scenes.CAMERA_C1, (256, 256)), (scenes.CAMERA_C2, (192, 108)), (scenes.CAMERA_C5, (64, 48))):
        m, cam = scenes.camera(cfg, W, H)
        org, d = orc.create_rays(m, cam, W, H)
        eye, target, fov = cfg
        _, ref = scenes.pinhole_rays(W, H, eye, target, fov_deg=fov, jitter=False)
        assert np.abs(d[:, :3] - ref[:, :3]).max() < 2e-6, cfg
        # the same formula in FP64 on the same float32 matrix
        x = np.arange(W * H) % W
        y = np.arange(W * H) // W
        s = np.stack([x / W * 2 - 1, y / H * 2 - 1, np.ones(W * H), np.ones(W * H)], 1)
        c = s @ m.astype(np.float64).reshape(4, 4)
        v = c[:, :3] / c[:, 3:4] - cam.astype(np.float64)
        assert np.abs(d[:, :3] - v / np.linalg.norm(v, axis=1, keepdims=True)).max() < 1e-6
        # a transposed read of the matrix is a different camera
        _, dt = orc.create_rays(m.reshape(4, 4).T.reshape(16), cam, W, H)
        assert np.abs(dt[:, :3] - ref[:, :3]).max() > 1e-2


def test_jitter_is_the_seeded_pcg_stream(orc):
    W, H, seed = 32, 16, 1234
    eye = np.eye(4, dtype=np.float32).reshape(16)
    _, d = orc.create_rays(eye, [0, 0, 0], W, H, jitter_seed=seed)
    u = orc._pcg_uniforms(seed, 1, W * H, 2)
    x = np.arange(W * H) % W
    y = np.arange(W * H) // W
    assert np.all((u >= 0) & (u < 1))
    sx = (x + u[0]) / W * 2 - 1
    sy = (y + u[1]) / H * 2 - 1
    assert np.abs(d[:, 0] / d[:, 2] - sx).max() < 1e-5 and np.abs(d[:, 1] / d[:, 2] - sy).max() < 1e-5
    _, d0 = orc.create_rays(eye, [0, 0, 0], W, H)
    assert not np.array_equal(d, d0)


def test_pcg_uniforms_match_the_scene_generator(orc, scenes):
    u = orc._pcg_uniforms(7, 5, 100, 3)
    r = scenes.Pcg32(7, 5, np.arange(100, dtype=np.uint64))
    for j in range(3):
        assert np.array_equal(u[j], r.uniform())


def test_emission_restatement_closed_forms(orc):
    """two lights of area 2 and 6 (1:3): choice frequencies, points inside
    their triangle, directions in the emitting hemisphere with cos-weighted
    mean cosine 2/3"""
    l0 = np.float32([[0, 0, 0, 1], [10, 0, 0, 1]])
    l1 = np.float32([[2, 0, 0, 1], [16, 0, 0, 1]])
    l2 = np.float32([[0, 0, 2, 1], [10, 0, 2, 1]])
    org, d, k = orc.photon_emission64((l0, l1, l2), 200_000, seed=9, offset=0.0)
    assert abs(np.mean(k == 1) - 0.75) < 0.005
    n = np.cross(l1[:, :3] - l0[:, :3], l2[:, :3] - l0[:, :3])[k].astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.allclose(n, [0, -1, 0])                         # e1 x e2 of both lights points -y
    cos = np.sum(d * n, axis=1)
    assert cos.min() >= 0 and abs(cos.mean() - 2 / 3) < 0.005
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-12
    # barycentric point inside the chosen triangle, on its plane y = 0
    assert np.all(org[:, 1] == 0)
    rel = org[:, [0, 2]] - l0[k][:, [0, 2]]
    size = np.where(k == 1, 6.0, 2.0)
    assert np.all(rel >= -1e-12) and np.all(rel.sum(axis=1) <= size + 1e-12)
