"""Differentiable first-hit intersection as a torch.autograd.Function.

Forward: gather the indexed mesh into the triangle SoA (pm_add_mesh), rebuild
the kd-tree (pm_build) and trace with the short-stack tracer (pm_trace).
Backward: pm_trace_backward, the hand-written adjoint of Moller-Trumbore with
a deterministic sort + segmented-reduce scatter into per-vertex gradients.
The hit ids are held fixed through the backward (the intersection is
piecewise: a perturbation that changes which triangle is hit is not a
differentiable direction).
"""
from __future__ import annotations

import torch

from . import SHORT_STACK, build_params, hits_view, trace_params


def _f4(x: torch.Tensor, w: float) -> torch.Tensor:
    out = torch.empty((x.shape[0], 4), dtype=torch.float32, device=x.device)
    out[:, :3] = x
    out[:, 3] = w
    return out


class TraceFunction(torch.autograd.Function):
    """(t, u, v, ids) = TraceFunction.apply(vertices[V,3], faces[F,3] int32,
    origins[R,3], directions[R,3], context[, build_params, trace_params])."""

    @staticmethod
    def forward(fctx, vertices, faces, origins, directions, context, bparams=None, tparams=None):
        dev = vertices.device
        faces_flat = faces.to(torch.int32).reshape(-1).contiguous()
        nt = faces.shape[0]
        v4 = _f4(vertices.detach().float(), 1.0)
        p0, p1, p2 = (torch.empty((nt, 4), dtype=torch.float32, device=dev) for _ in range(3))
        context.add_mesh(v4, faces_flat, p0, p1, p2)
        context.build(p0, p1, p2, bparams or build_params("reference"))
        o4 = _f4(origins.detach().float(), 1.0)
        d4 = _f4(directions.detach().float(), 0.0)
        hits = context.trace(SHORT_STACK, o4, d4, tparams or trace_params(stack_size=8))
        ids, t, u, v = hits_view(hits)
        fctx.save_for_backward(o4, d4, hits, faces_flat, p0, p1, p2)
        fctx.context = context
        fctx.nv = vertices.shape[0]
        ids = ids.clone()
        fctx.mark_non_differentiable(ids)
        return t.clone(), u.clone(), v.clone(), ids

    @staticmethod
    def backward(fctx, gt, gu, gv, gid):
        o4, d4, hits, faces_flat, p0, p1, p2 = fctx.saved_tensors
        gt, gu, gv = (None if g is None else g.contiguous().float() for g in (gt, gu, gv))
        # the triangles this forward traced, passed explicitly: the context
        # may hold another mesh's tree and geometry by now, and is not changed
        gvert, go, gd = fctx.context.trace_backward(o4, d4, hits, gt, gu, gv, faces_flat, fctx.nv,
                                                    tris=(p0, p1, p2))
        return gvert, None, go, gd, None, None, None


def trace(vertices, faces, origins, directions, context, bparams=None, tparams=None):
    return TraceFunction.apply(vertices, faces, origins, directions, context, bparams, tparams)
