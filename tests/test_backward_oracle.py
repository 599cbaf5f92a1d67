"""Pins the intersection backward (oracle FP32 restatement and FP64 adjoint)
by the mathematics: FP64 central differences of the loss with the hit
triangles held fixed, torch.autograd.gradcheck on a double-precision
restatement, and FP32-vs-FP64 agreement."""
import numpy as np
import pytest


def mt64_t(V, faces, ids, org, dirs):
    """FP64 t of each ray against its (fixed) hit triangle"""
    tri = faces[ids]
    v0, v1, v2 = V[tri[:, 0]], V[tri[:, 1]], V[tri[:, 2]]
    o = org[:, :3].astype(np.float64)
    d = dirs[:, :3].astype(np.float64)
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = np.sum(p * e1, axis=1)
    q = np.cross(o - v0, e1)
    return np.sum(q * e2, axis=1) / det


def _t64_local(A, org, dirs):
    """FP64 t for rays against per-ray triangle corners A[k, 3 corners, 3]"""
    o = org[:, :3].astype(np.float64)
    d = dirs[:, :3].astype(np.float64)
    e1, e2 = A[:, 1] - A[:, 0], A[:, 2] - A[:, 0]
    p = np.cross(d, e2)
    q = np.cross(o - A[:, 0], e1)
    return np.sum(q * e2, axis=1) / np.sum(p * e1, axis=1)


def _fd_once(V, faces, ids, org, dirs, target, sample, touching, h):
    out = np.zeros((len(sample), 3))
    for si, j in enumerate(sample):
        rs = touching[si]
        if len(rs) == 0:
            continue
        tri = faces[ids[rs]]
        A = V[tri]
        hit = tri == j
        for c in range(3):
            Ap = A.copy(); Ap[hit, c] += h
            Am = A.copy(); Am[hit, c] -= h
            lp = np.sum((_t64_local(Ap, org[rs], dirs[rs]) - target[rs]) ** 2)
            lm = np.sum((_t64_local(Am, org[rs], dirs[rs]) - target[rs]) ** 2)
            out[si, c] = (lp - lm) / (2 * h)
    return out


def fd_vertex_grads(V, faces, ids, org, dirs, target, sample, h=1e-4, levels=3):
    """FP64 central differences of L = sum (t - target)^2 (hit triangles held
    fixed) with step h, Richardson-extrapolated over h, h/2, h/4: grazing rays
    make t strongly nonlinear in the vertices, so the plain O(h^2) difference
    at h = 1e-4 carries percent-level truncation error there."""
    touching = [np.nonzero(np.any(faces[ids] == j, axis=1))[0] for j in sample]
    D = [_fd_once(V, faces, ids, org, dirs, target, sample, touching, h / 2 ** k) for k in range(levels)]
    for k in range(1, levels):
        f = 4.0 ** k
        D = [(f * D[i + 1] - D[i]) / (f - 1) for i in range(len(D) - 1)]
    return D[0]


@pytest.fixture(scope="module")
def small_sphere(orc, scenes):
    v_ref, v_pert, faces, (org, dirs) = scenes.config5(width=64, height=64, subdiv=4)

    def tris(V):
        return scenes.tris_from(V[faces[:, 0]], V[faces[:, 1]], V[faces[:, 2]])

    t_ref, t_pert = tris(v_ref), tris(v_pert)
    h_ref = orc.trace(orc.build(t_ref), t_ref, org, dirs, orc.SHORT_STACK)
    h = orc.trace(orc.build(t_pert), t_pert, org, dirs, orc.SHORT_STACK)
    mask = (h["id"] >= 0) & (h_ref["id"] >= 0)
    return dict(v_pert=v_pert, faces=faces, org=org, dirs=dirs, tris=t_pert, hits=h,
                target=h_ref["t"], mask=mask)


def test_fp64_adjoint_matches_central_differences(orc, small_sphere):
    s = small_sphere
    hits = s["hits"].copy()
    hits["id"][~s["mask"]] = -1
    m = s["mask"]
    gt = np.zeros(m.shape[0], np.float32)
    gt[m] = 2.0 * (s["hits"]["t"][m].astype(np.float64) - s["target"][m])
    g64, _, _ = orc.trace_backward(s["tris"], s["org"], s["dirs"], hits, grad_t=gt, faces=s["faces"],
                                   num_vertices=s["v_pert"].shape[0], fp64=True)
    rng = np.random.default_rng(0)
    used = np.unique(s["faces"][hits["id"][m]])
    sample = rng.choice(used, size=min(64, len(used)), replace=False)
    V = s["v_pert"].astype(np.float64)
    fd = fd_vertex_grads(V, s["faces"], hits["id"][m], s["org"][m], s["dirs"][m],
                         s["target"][m].astype(np.float64), sample)
    # the FP32 t values inside gt differ from the FP64 ones by ~1e-7 relative
    err = np.linalg.norm(g64[sample] - fd) / np.linalg.norm(fd)
    assert err < 1e-4, err


def test_fp32_backward_matches_fp64(orc, small_sphere):
    s = small_sphere
    hits = s["hits"].copy()
    hits["id"][~s["mask"]] = -1
    rng = np.random.default_rng(1)
    gt, gu, gv = (rng.normal(size=hits.shape[0]).astype(np.float32) for _ in range(3))
    nv = s["v_pert"].shape[0]
    g32, o32, d32 = orc.trace_backward(s["tris"], s["org"], s["dirs"], hits, gt, gu, gv, s["faces"], nv)
    g64, o64, d64 = orc.trace_backward(s["tris"], s["org"], s["dirs"], hits, gt, gu, gv, s["faces"], nv,
                                       fp64=True)
    for a, b in ((g32, g64), (o32, o64), (d32, d64)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4
    # without faces: one gradient per triangle corner, summing to the same vertex gradients
    gc, _, _ = orc.trace_backward(s["tris"], s["org"], s["dirs"], hits, gt, gu, gv)
    summed = np.zeros_like(g64)
    np.add.at(summed, s["faces"].reshape(-1), gc.astype(np.float64))
    assert np.linalg.norm(summed - g64) / np.linalg.norm(g64) < 1e-4


def test_gradcheck_fp64_restatement(orc):
    """torch.autograd.gradcheck of (t, u, v) against the FP64 adjoint the
    backward kernel restates (single triangle, a few rays, double)."""
    import torch

    class MT64(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v0, v1, v2, o, d):
            ctx.save_for_backward(v0, v1, v2, o, d)
            r = orc.mt64(v0.numpy(), v1.numpy(), v2.numpy(), o.numpy(), d.numpy())
            return torch.tensor(r[:3], dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            v0, v1, v2, o, d = ctx.saved_tensors
            gs = orc.mt64_adjoint(v0.numpy(), v1.numpy(), v2.numpy(), o.numpy(), d.numpy(),
                                  float(g[0]), float(g[1]), float(g[2]))
            return tuple(torch.tensor(x, dtype=torch.float64) for x in gs)

    rng = np.random.default_rng(4)
    for _ in range(5):
        tri = [torch.tensor(rng.uniform(-1, 1, 3), dtype=torch.float64, requires_grad=True) for _ in range(3)]
        o = torch.tensor(rng.uniform(-3, 3, 3), dtype=torch.float64, requires_grad=True)
        d = (sum(t.detach() for t in tri) / 3 - o.detach())
        d = (d / d.norm()).clone().requires_grad_(True)
        assert torch.autograd.gradcheck(MT64.apply, (*tri, o, d), eps=1e-6, atol=1e-5, rtol=1e-4)
