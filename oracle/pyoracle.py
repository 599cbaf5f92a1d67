"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

HIT_DTYPE = np.dtype([("id", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
BRUTE, KD_RESTART, SHORT_STACK = 0, 1, 2
LOWER_BITMAP, LOWER_BALANCED, LOWER_SAH = 0, 1, 2
SPLIT_BOX, SPLIT_DIVIDE = 0, 1
MOELLER, WOOP = 0, 1


class BuildParams(C.Structure):
    _fields_ = [("lower_algorithm", C.c_int), ("empty_space_splitting", C.c_int),
                ("empty_space_threshold", C.c_float), ("split_method", C.c_int),
                ("propagate_aabbs", C.c_int), ("traversal_cost", C.c_float),
                ("lower_size", C.c_int), ("max_upper_levels", C.c_int),
                ("sah_leaf_max", C.c_int), ("sah_empty_bonus", C.c_float),
                ("max_ref_factor", C.c_int)]


class TraceParams(C.Structure):
    _fields_ = [("leaf_skipping", C.c_int), ("stack_size", C.c_int), ("brute_cap", C.c_int),
                ("threads", C.c_int), ("intersection", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("num_nodes", C.c_int), ("reachable_nodes", C.c_int), ("inner_nodes", C.c_int),
                ("leaves", C.c_int), ("empty_leaves", C.c_int), ("max_depth", C.c_int),
                ("mean_leaf_depth", C.c_double), ("leaf_refs", C.c_longlong),
                ("max_leaf_size", C.c_int), ("upper_levels", C.c_int), ("lower_levels", C.c_int),
                ("num_prim_indices", C.c_int), ("num_upper_leaves", C.c_int)]


def build_params(preset: str = "reference", **kw) -> BuildParams:
    """'reference' = TriangleMap defaults with the SAH lower creator
    (TriangleMap.cu:27-49, TriangleMapUpperCreator.cu:28-32);
    'bitmap' = the reference default lower creator (BITMAP)."""
    p = BuildParams(lower_algorithm=LOWER_SAH, empty_space_splitting=1,
                    empty_space_threshold=0.35, split_method=0, propagate_aabbs=1,
                    traversal_cost=24.0, lower_size=32, max_upper_levels=0,
                    sah_leaf_max=0, sah_empty_bonus=1.0, max_ref_factor=0)
    if preset == "bitmap":
        p.lower_algorithm = LOWER_BITMAP
    elif preset == "balanced":
        p.lower_algorithm = LOWER_BALANCED
    elif preset == "config":
        p.sah_leaf_max = 8
        p.sah_empty_bonus = 0.8
    elif preset != "reference":
        raise ValueError(preset)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def ensure_built() -> None:
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        ensure_built()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_build.argtypes = [P, P, P, C.c_int, C.POINTER(BuildParams), C.POINTER(P)]
        L.orc_build.restype = C.c_int
        L.orc_build_batched.argtypes = [P, P, P, C.c_int, P, C.c_int, C.POINTER(BuildParams), C.POINTER(P)]
        L.orc_build_batched.restype = C.c_int
        L.orc_woop_matrix.argtypes = [P, P, P, P]
        L.orc_woop_matrix.restype = None
        L.orc_tri_box_step3.argtypes = [P, P, P, P, P]
        L.orc_tri_box_step3.restype = C.c_int
        L.orc_balanced_node.argtypes = [C.c_int, P, P, C.c_longlong, C.c_float, C.c_int, P, P, P, P, P]
        L.orc_balanced_node.restype = C.c_int
        L.orc_trace_canvas.argtypes = [P, P, P, P, C.c_int, C.c_int, P, P, C.c_int, P, P, P, P, P, C.c_int,
                                       C.POINTER(TraceParams), P]
        L.orc_trace_canvas.restype = C.c_int
        L.orc_free.argtypes = [P]
        L.orc_last_error.restype = C.c_char_p
        for name in ("orc_num_nodes", "orc_num_prim_indices", "orc_num_leaf_ids",
                     "orc_upper_levels", "orc_lower_levels", "orc_num_roots"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = C.c_int
        L.orc_upper_level_refs.argtypes = [P, C.c_int]
        L.orc_upper_level_refs.restype = C.c_longlong
        L.orc_lower_level_nodes.argtypes = [P, C.c_int]
        L.orc_lower_level_nodes.restype = C.c_longlong
        L.orc_copy_tree.argtypes = [P] * 12
        L.orc_stats_of.argtypes = [P, C.POINTER(Stats)]
        L.orc_stats_of.restype = C.c_int
        L.orc_tree_from_arrays.argtypes = [C.c_int, P, P, P, P, P, P, P, C.c_int, P, C.POINTER(P)]
        L.orc_tree_from_arrays.restype = C.c_int
        L.orc_trace.argtypes = [P, P, P, P, C.c_int, C.c_int, P, P, C.c_int,
                                C.POINTER(TraceParams), P, P]
        L.orc_trace_batched.argtypes = [P, P, P, P, C.c_int, C.c_int, P, P, P, C.c_int,
                                        C.POINTER(TraceParams), P, P]
        L.orc_trace_batched.restype = C.c_int
        L.orc_trace.restype = C.c_int
        L.orc_mt32.argtypes = [P, P, P, P, P, C.c_float, P]
        L.orc_mt32.restype = C.c_int
        L.orc_mt64.argtypes = [P] * 6
        L.orc_mt64_adjoint.argtypes = [P] * 5 + [C.c_double] * 3 + [P] * 5
        L.orc_sah_node.argtypes = [C.c_int, P, P, P, C.c_longlong, C.c_float, C.c_float, C.c_int,
                                   C.c_float, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                   C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_sah_node.restype = C.c_int
        L.orc_trace_backward32.argtypes = [P, P, P, C.c_int, P, P, P, C.c_int, P, P, P, P, C.c_int, P, P, P]
        L.orc_trace_backward32.restype = C.c_int
        L.orc_trace_backward64.argtypes = [P, P, P, C.c_int, P, P, P, C.c_int, P, P, P, P, C.c_int, P, P, P]
        L.orc_trace_backward64.restype = C.c_int
        L.orc_tri_area.argtypes = [P, P, P]
        L.orc_tri_area.restype = C.c_float
        L.orc_create_rays.argtypes = [P, P, C.c_int, C.c_int, C.c_ulonglong, C.c_int, P, P]
        L.orc_create_rays.restype = None
        _lib = L
    return _lib


def create_rays(inv_view_proj, cam_pos, width: int, height: int, jitter_seed=None):
    """camera rays (IRayTracer.cu:45-70) in FP32: (org, dir) float32 [W*H, 4]"""
    m = _c(np.asarray(inv_view_proj).reshape(16), np.float32)
    c = _c(np.asarray(cam_pos).reshape(3), np.float32)
    n = width * height
    org = np.zeros((n, 4), np.float32)
    dirs = np.zeros((n, 4), np.float32)
    lib().orc_create_rays(_ptr(m), _ptr(c), width, height, (jitter_seed or 0) & (2**64 - 1),
                          int(jitter_seed is not None), _ptr(org), _ptr(dirs))
    return org, dirs


def _pcg_uniforms(seed: int, purpose: int, n: int, k: int) -> np.ndarray:
    """k successive 24-bit PCG32 uniforms of streams (seed, purpose << 40 | i),
    i < n, as float64 [k, n] (each exactly a float32)"""
    mult = np.uint64(6364136223846793005)
    inc = ((np.uint64(purpose) << np.uint64(40)) | np.arange(n, dtype=np.uint64)) << np.uint64(1) | np.uint64(1)
    state = np.zeros(n, np.uint64) * mult + inc
    state = (state + np.uint64(seed & (2**64 - 1))) * mult + inc
    out = np.empty((k, n))
    for j in range(k):
        old = state
        state = old * mult + inc
        xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
        rot = old >> np.uint64(59)
        r = (xs >> rot) | ((xs << ((np.uint64(32) - rot) & np.uint64(31))) & np.uint64(0xFFFFFFFF))
        out[j] = (r >> np.uint64(8)).astype(np.float64) / 16777216.0
    return out


def photon_emission64(lights, n: int, seed: int, offset: float = 1e-4):
    """FP64 restatement of pm_photon_emit: light k chosen by the FP32 running
    area sum (the first k with u0 * total < cdf[k], FP32 as the device does
    it, so the choice is the same integer), a uniform point
    (1 - s) v0 + s (1 - u2) v1 + s u2 v2 with s = sqrt(u1), and a
    cosine-weighted direction (Malley) about n = normalize(e1 x e2) in the
    frame t = normalize(a x n), b = n x t; origin p + offset n.  Returns
    (org, dir, light index) with float64 rows."""
    l0, l1, l2 = (np.asarray(x, np.float32)[:, :3] for x in lights)
    m = l0.shape[0]
    cdf = np.zeros(m, np.float32)
    acc = np.float32(0.0)
    for k in range(m):
        acc = np.float32(acc + np.float32(tri_area(l0[k], l1[k], l2[k])))
        cdf[k] = acc
    u = _pcg_uniforms(seed, 5, n, 5)
    target = u[0].astype(np.float32) * cdf[-1]
    k = np.minimum(np.searchsorted(cdf, target, side="right"), m - 1)
    v0, v1, v2 = (x[k].astype(np.float64) for x in (l0, l1, l2))
    s = np.sqrt(u[1])
    p = (1 - s)[:, None] * v0 + (s * (1 - u[2]))[:, None] * v1 + (s * u[2])[:, None] * v2
    nrm = np.cross(v1 - v0, v2 - v0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    r = np.sqrt(u[3])
    phi = 2.0 * np.pi * u[4]
    x, y, z = r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u[3]))
    a = np.where(np.abs(nrm[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(a, nrm)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    d = x[:, None] * t + y[:, None] * b + z[:, None] * nrm
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return p + offset * nrm, d, k


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Tree:
    """Oracle tree; ``arrays()`` returns the reference SoA layout."""

    def __init__(self, handle, keep=()):
        self.h = handle
        self._keep = keep

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    @property
    def num_nodes(self) -> int:
        return lib().orc_num_nodes(self.h)

    def arrays(self) -> dict:
        L = lib()
        n = L.orc_num_nodes(self.h)
        npi = L.orc_num_prim_indices(self.h)
        nl = L.orc_num_leaf_ids(self.h)
        a = dict(info=np.zeros(n, np.int8), split_pos=np.zeros(n, np.float32),
                 aabb_min=np.zeros((n, 4), np.float32), aabb_max=np.zeros((n, 4), np.float32),
                 primitive_index=np.zeros(n, np.int32), primitive_bitmap=np.zeros(n, np.int64),
                 children=np.zeros((n, 2), np.int32), surface_area=np.zeros(n, np.float32),
                 parent=np.zeros(n, np.int32), prim_indices=np.zeros(npi, np.int32),
                 leaf_ids=np.zeros(nl, np.int32))
        L.orc_copy_tree(self.h, *[_ptr(a[k]) for k in (
            "info", "split_pos", "aabb_min", "aabb_max", "primitive_index", "primitive_bitmap",
            "children", "surface_area", "parent", "prim_indices", "leaf_ids")])
        a["num_roots"] = L.orc_num_roots(self.h)
        return a

    def stats(self) -> dict:
        s = Stats()
        if lib().orc_stats_of(self.h, C.byref(s)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        d = {k: getattr(s, k) for k, _ in Stats._fields_}
        L = lib()
        d["upper_level_refs"] = [L.orc_upper_level_refs(self.h, i) for i in range(s.upper_levels)]
        d["lower_level_nodes"] = [L.orc_lower_level_nodes(self.h, i) for i in range(s.lower_levels)]
        return d


def build(tris, params: BuildParams | None = None) -> Tree:
    p0, p1, p2 = (_c(t, np.float32) for t in tris)
    n = p0.shape[0]
    params = params or build_params()
    h = C.c_void_p()
    rc = lib().orc_build(_ptr(p0), _ptr(p1), _ptr(p2), n, C.byref(params), C.byref(h))
    if rc != 0:
        raise RuntimeError(f"orc_build failed ({rc}): {lib().orc_last_error().decode()}")
    return Tree(h)


def build_batched(tris, mesh_offsets, params: BuildParams | None = None) -> Tree:
    """forest: one tree per mesh, root m = node m (mesh m = triangles
    [mesh_offsets[m], mesh_offsets[m+1]))"""
    p0, p1, p2 = (_c(t, np.float32) for t in tris)
    off = _c(mesh_offsets, np.int32)
    params = params or build_params()
    h = C.c_void_p()
    rc = lib().orc_build_batched(_ptr(p0), _ptr(p1), _ptr(p2), p0.shape[0], _ptr(off),
                                 off.shape[0] - 1, C.byref(params), C.byref(h))
    if rc != 0:
        raise RuntimeError(f"orc_build_batched failed ({rc}): {lib().orc_last_error().decode()}")
    return Tree(h)


def trace_batched(tree: Tree, tris, org, dirs, ray_root, algo: int, leaf_skipping=True,
                  stack_size=4, threads=0, counters=False, intersection=MOELLER):
    """trace ray r from root ray_root[r] of a forest"""
    p0, p1, p2 = (_c(t, np.float32) for t in tris)
    org = _c(org, np.float32)
    dirs = _c(dirs, np.float32)
    root = _c(ray_root, np.int32)
    n = org.shape[0]
    out = np.zeros(n, HIT_DTYPE)
    cnt = np.zeros((n, 4), np.int32) if counters else None
    prm = TraceParams(int(bool(leaf_skipping)), stack_size, 0, threads, intersection)
    rc = lib().orc_trace_batched(tree.h, _ptr(p0), _ptr(p1), _ptr(p2), p0.shape[0], algo,
                                 _ptr(org), _ptr(dirs), _ptr(root), n, C.byref(prm), _ptr(out),
                                 _ptr(cnt))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return (out, cnt) if counters else out


def tree_from_arrays(a: dict) -> Tree:
    keep = {k: np.ascontiguousarray(v) for k, v in a.items()}
    h = C.c_void_p()
    rc = lib().orc_tree_from_arrays(
        keep["info"].shape[0], _ptr(_c(keep["info"], np.int8)), _ptr(_c(keep["split_pos"], np.float32)),
        _ptr(_c(keep["aabb_min"], np.float32)), _ptr(_c(keep["aabb_max"], np.float32)),
        _ptr(_c(keep["primitive_index"], np.int32)), _ptr(_c(keep["primitive_bitmap"], np.int64)),
        _ptr(_c(keep["children"], np.int32)), keep["prim_indices"].shape[0],
        _ptr(_c(keep["prim_indices"], np.int32)), C.byref(h))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return Tree(h)


def trace(tree: Tree | None, tris, org, dirs, algo: int, leaf_skipping=True, stack_size=4,
          brute_cap=0, threads=0, counters=False, intersection=MOELLER):
    p0, p1, p2 = (_c(t, np.float32) for t in tris)
    org = _c(org, np.float32)
    dirs = _c(dirs, np.float32)
    n = org.shape[0]
    out = np.zeros(n, HIT_DTYPE)
    cnt = np.zeros((n, 4), np.int32) if counters else None
    prm = TraceParams(int(bool(leaf_skipping)), stack_size, brute_cap, threads, intersection)
    rc = lib().orc_trace(tree.h if tree is not None else None, _ptr(p0), _ptr(p1), _ptr(p2),
                         p0.shape[0], algo, _ptr(org), _ptr(dirs), n, C.byref(prm),
                         _ptr(out), _ptr(cnt))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return (out, cnt) if counters else out


def mt32(v0, v1, v2, o, d, t_max=2139095040.0):
    a = [_c(x, np.float32) for x in (v0, v1, v2, o, d)]
    r = np.zeros(3, np.float32)
    hit = lib().orc_mt32(*[_ptr(x) for x in a], t_max, _ptr(r))
    return bool(hit), r


def mt64(v0, v1, v2, o, d):
    a = [_c(x, np.float64) for x in (v0, v1, v2, o, d)]
    r = np.zeros(4, np.float64)
    lib().orc_mt64(*[_ptr(x) for x in a], _ptr(r))
    return r


def mt64_adjoint(v0, v1, v2, o, d, gt, gu, gv):
    a = [_c(x, np.float64) for x in (v0, v1, v2, o, d)]
    outs = [np.zeros(3, np.float64) for _ in range(5)]
    lib().orc_mt64_adjoint(*[_ptr(x) for x in a], gt, gu, gv, *[_ptr(x) for x in outs])
    return outs


def sah_node(bmin, bmax, areas, bitmap, node_area, traversal_cost=24.0, leaf_max=0, bonus=1.0):
    bmin = _c(bmin, np.float32)
    bmax = _c(bmax, np.float32)
    areas = _c(areas, np.float32)
    n = bmin.shape[0]
    ax, pos = C.c_int(), C.c_float()
    l, r = C.c_longlong(), C.c_longlong()
    la, ra, cost = C.c_float(), C.c_float(), C.c_float()
    s = lib().orc_sah_node(n, _ptr(bmin), _ptr(bmax), _ptr(areas), int(bitmap), node_area,
                           traversal_cost, leaf_max, bonus, C.byref(ax), C.byref(pos), C.byref(l),
                           C.byref(r), C.byref(la), C.byref(ra), C.byref(cost))
    return dict(split=bool(s), axis=ax.value, pos=pos.value, left=l.value & (2**64 - 1),
                right=r.value & (2**64 - 1), left_area=la.value, right_area=ra.value,
                cost=cost.value)


def tri_area(v0, v1, v2) -> float:
    a = [_c(x, np.float32) for x in (v0, v1, v2)]
    return float(lib().orc_tri_area(*[_ptr(x) for x in a]))


def trace_backward(tris, org, dirs, hits, grad_t=None, grad_u=None, grad_v=None, faces=None,
                   num_vertices=None, fp64=False):
    """per-vertex (and per-ray) gradients of t/u/v for the given hit records"""
    p0, p1, p2 = (_c(t, np.float32) for t in tris)
    org = _c(org, np.float32)
    dirs = _c(dirs, np.float32)
    hits = np.ascontiguousarray(hits)
    n = org.shape[0]
    nt = p0.shape[0]
    if num_vertices is None:
        num_vertices = 3 * nt if faces is None else int(np.max(faces)) + 1
    dt = np.float64 if fp64 else np.float32
    gvtx = np.zeros((num_vertices, 3), dt)
    go = np.zeros((n, 3), dt)
    gd = np.zeros((n, 3), dt)
    gs = [None if g is None else _c(g, np.float32) for g in (grad_t, grad_u, grad_v)]
    fc = None if faces is None else _c(faces, np.int32)
    fn = lib().orc_trace_backward64 if fp64 else lib().orc_trace_backward32
    rc = fn(_ptr(p0), _ptr(p1), _ptr(p2), nt, _ptr(org), _ptr(dirs), _ptr(hits), n,
            *[_ptr(g) for g in gs], _ptr(fc), num_vertices, _ptr(gvtx), _ptr(go), _ptr(gd))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return gvtx, go, gd


def woop_matrix(a, b, c) -> np.ndarray:
    """WoopTransformationMatrix(a, b, c) -> float32 [3, 4] (m0, m1, m2)"""
    x = [_c(v, np.float32) for v in (a, b, c)]
    m = np.zeros((3, 4), np.float32)
    lib().orc_woop_matrix(*[_ptr(v) for v in x], _ptr(m))
    return m


def tri_box_step3(v0, v1, v2, bmin, bmax) -> bool:
    x = [_c(v, np.float32) for v in (v0, v1, v2, bmin, bmax)]
    return bool(lib().orc_tri_box_step3(*[_ptr(v) for v in x]))


def balanced_node(bmin, bmax, bitmap, traversal_cost=24.0, lower_size=32):
    """balanced creator's decision for one node -> dict(split, axis, pos, left, right, cost)"""
    bmin, bmax = _c(bmin, np.float32), _c(bmax, np.float32)
    ax, pos, l, r, cost = C.c_int(), C.c_float(), C.c_longlong(), C.c_longlong(), C.c_int()
    sp = lib().orc_balanced_node(bmin.shape[0], _ptr(bmin), _ptr(bmax), int(np.int64(np.uint64(bitmap))),
                                 traversal_cost, lower_size, C.byref(ax), C.byref(pos), C.byref(l),
                                 C.byref(r), C.byref(cost))
    return dict(split=bool(sp), axis=ax.value, pos=pos.value, left=l.value & (2 ** 64 - 1),
                right=r.value & (2 ** 64 - 1), cost=cost.value)


def trace_canvas(tree, tris, org, dirs, algo, normals, colors, light, max_bounces=64,
                 leaf_skipping=True, stack_size=4, brute_cap=0, threads=0, intersection=MOELLER):
    """shaded canvas (uchar4 per ray); normals = (n0, n1, n2) float32 [N, 4],
    colors = uint8 [N, 4], light = 12 floats (position, ambient, diffuse, specular)"""
    p0, p1, p2 = (_c(t, np.float32) for t in tris)
    org, dirs = _c(org, np.float32), _c(dirs, np.float32)
    n0, n1, n2 = (_c(x, np.float32) for x in normals)
    c0 = _c(colors, np.uint8)
    lt = _c(light, np.float32)
    n = org.shape[0]
    out = np.zeros((n, 4), np.uint8)
    prm = TraceParams(int(bool(leaf_skipping)), stack_size, brute_cap, threads, intersection)
    rc = lib().orc_trace_canvas(tree.h if tree is not None else None, _ptr(p0), _ptr(p1), _ptr(p2),
                                p0.shape[0], algo, _ptr(org), _ptr(dirs), n, _ptr(n0), _ptr(n1), _ptr(n2),
                                _ptr(c0), _ptr(lt), max_bounces, C.byref(prm), _ptr(out))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out
