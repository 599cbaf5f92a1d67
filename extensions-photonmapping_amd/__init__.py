"""pmkd -- MI355X (gfx950) kd-tree build and ray traversal.

Python binding of libpmkd.so (the C ABI in include/pmkd.h) over torch
device tensors.  The compute path is the HIP library only: if libpmkd.so is
missing or has no GPU to run on, these calls raise; there is no CPU fallback.

Layout conventions follow the reference's device SoA
(Utils/CUDA/GeometryList.h:58-68, Scene/KDNode.h:65-73):
  triangles  three float32 [N, 4] tensors p0, p1, p2 (w unused)
  rays       float32 [R, 4] origin and direction
  hits       int32/float32 [R, 4] records {id, t, u, v} (see ``hits_view``)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmkd.so")

PM_OK = 0
ERRORS = {1: "PM_ERR_ARG", 2: "PM_ERR_NOT_IMPLEMENTED", 3: "PM_ERR_DEPTH", 4: "PM_ERR_INTERNAL",
          5: "PM_ERR_UNSUPPORTED", 6: "PM_ERR_HIP", 7: "PM_ERR_NO_TREE", 8: "PM_ERR_CAPACITY"}
LOWER_BITMAP, LOWER_BALANCED, LOWER_SAH = 0, 1, 2
SPLIT_BOX, SPLIT_DIVIDE, SPLIT_SPLIT = 0, 1, 2
BRUTE, KD_RESTART, SHORT_STACK = 0, 1, 2
MOELLER, WOOP = 0, 1
RAY_ORDER_INPUT, RAY_ORDER_COHERENT = 0, 1
MAX_LIGHTS = 65536      # PM_MAX_LIGHTS: light triangles pm_photon_emit takes

# every symbol include/pmkd.h declares
EXPORTS = ("pm_build_params_default", "pm_trace_params_default", "pm_create", "pm_destroy",
           "pm_last_error", "pm_build", "pm_set_geometry", "pm_tree_view", "pm_set_tree",
           "pm_tree_stats", "pm_trace", "pm_photon_bounce", "pm_add_mesh", "pm_trace_backward",
           "pm_build_batched", "pm_trace_batched", "pm_osc_surface", "pm_deform_radial",
           "pm_autotune", "pm_batch_graph_capture", "pm_batch_graph_launch", "pm_batch_graph_status",
           "pm_trace_canvas", "pm_add_mesh_attributes", "pm_create_rays", "pm_photon_emit",
           "pm_trace_noalloc")


class PmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class BuildParams(C.Structure):
    _fields_ = [("lower_algorithm", C.c_int), ("empty_space_splitting", C.c_int),
                ("empty_space_threshold", C.c_float), ("split_method", C.c_int),
                ("propagate_aabbs", C.c_int), ("traversal_cost", C.c_float),
                ("lower_size", C.c_int), ("max_upper_levels", C.c_int),
                ("sah_leaf_max", C.c_int), ("sah_empty_bonus", C.c_float),
                ("max_ref_factor", C.c_int), ("chunk_size", C.c_int)]


class TraceParams(C.Structure):
    _fields_ = [("leaf_skipping", C.c_int), ("stack_size", C.c_int), ("brute_cap", C.c_int),
                ("intersection", C.c_int), ("ray_order", C.c_int), ("step_cap", C.c_int)]


class Tree(C.Structure):
    _fields_ = [("num_nodes", C.c_int), ("info", C.c_void_p), ("split_pos", C.c_void_p),
                ("aabb_min", C.c_void_p), ("aabb_max", C.c_void_p),
                ("primitive_index", C.c_void_p), ("primitive_bitmap", C.c_void_p),
                ("children", C.c_void_p), ("surface_area", C.c_void_p), ("parent", C.c_void_p),
                ("num_prim_indices", C.c_int), ("prim_indices", C.c_void_p),
                ("num_leaf_ids", C.c_int), ("leaf_ids", C.c_void_p), ("p0", C.c_void_p),
                ("p1", C.c_void_p), ("p2", C.c_void_p), ("num_tris", C.c_int),
                ("upper_levels", C.c_int), ("lower_levels", C.c_int), ("has_boxes", C.c_int),
                ("num_roots", C.c_int)]


class Light(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("ambient", C.c_float * 3), ("diffuse", C.c_float * 3),
                ("specular", C.c_float * 3)]


class GraphCaps(C.Structure):
    _fields_ = [("nodes", C.c_longlong), ("active", C.c_longlong), ("segments", C.c_longlong),
                ("level_refs", C.c_longlong), ("leaf_refs", C.c_longlong), ("leaf_ids", C.c_longlong),
                ("lower_nodes", C.c_longlong), ("upper_levels", C.c_int), ("lower_levels", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("chunk_size", C.c_int), ("lower_size", C.c_int), ("stack_size", C.c_int),
                ("build_ms", C.c_float), ("trace_ms", C.c_float), ("trials", C.c_int),
                ("ray_order", C.c_int), ("step_cap", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("num_nodes", C.c_int), ("reachable_nodes", C.c_int), ("inner_nodes", C.c_int),
                ("leaves", C.c_int), ("empty_leaves", C.c_int), ("max_depth", C.c_int),
                ("mean_leaf_depth", C.c_double), ("leaf_refs", C.c_longlong),
                ("max_leaf_size", C.c_int), ("upper_levels", C.c_int), ("lower_levels", C.c_int),
                ("num_prim_indices", C.c_int), ("num_upper_leaves", C.c_int),
                ("upper_ref_reads", C.c_longlong), ("upper_ref_writes", C.c_longlong),
                ("lower_prims", C.c_longlong), ("planned_build", C.c_int)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libpmkd.so and declare the C signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: run __graft_entry__.build() (make -C csrc)")
    L = C.CDLL(path)
    P, I = C.c_void_p, C.c_int
    L.pm_build_params_default.argtypes = [C.POINTER(BuildParams)]
    L.pm_build_params_default.restype = None
    L.pm_trace_params_default.argtypes = [C.POINTER(TraceParams)]
    L.pm_trace_params_default.restype = None
    L.pm_create.argtypes = [C.POINTER(P), I]
    L.pm_destroy.argtypes = [P]
    L.pm_last_error.argtypes = [P]
    L.pm_last_error.restype = C.c_char_p
    L.pm_build.argtypes = [P, P, P, P, I, C.POINTER(BuildParams), P]
    L.pm_set_geometry.argtypes = [P, P, P, P, I]
    L.pm_tree_view.argtypes = [P, C.POINTER(Tree)]
    L.pm_set_tree.argtypes = [P, C.POINTER(Tree), P]
    L.pm_tree_stats.argtypes = [P, C.POINTER(Stats), P]
    L.pm_trace.argtypes = [P, I, P, P, I, P, C.POINTER(TraceParams), P, P]
    L.pm_photon_bounce.argtypes = [P, P, P, P, I, I, C.c_ulonglong, C.c_float, P, P, P, P,
                                   C.POINTER(C.c_int), P]
    L.pm_add_mesh.argtypes = [P, P, I, P, I, P, P, P, P, I, P]
    L.pm_trace_backward.argtypes = [P, P, P, P, I, P, P, P, I, P, P, P, P, I, P, P, P, P]
    L.pm_build_batched.argtypes = [P, P, P, P, I, P, I, C.POINTER(BuildParams), P]
    L.pm_trace_batched.argtypes = [P, I, P, P, P, I, P, C.POINTER(TraceParams), P, P]
    F = C.c_float
    L.pm_osc_surface.argtypes = [P, P, P, I, F, F, F, F, P]
    L.pm_deform_radial.argtypes = [P, P, I, F, P, P, P]
    L.pm_autotune.argtypes = [P, P, P, P, I, P, I, P, P, P, I, C.POINTER(BuildParams), F, I,
                              C.POINTER(Tuning), P]
    L.pm_batch_graph_capture.argtypes = [P, P, P, P, I, P, I, P, P, P, I, P, I, C.POINTER(BuildParams),
                                         C.POINTER(TraceParams), C.POINTER(GraphCaps), P]
    L.pm_batch_graph_launch.argtypes = [P, P]
    L.pm_batch_graph_status.argtypes = [P, C.POINTER(C.c_int), P]
    L.pm_trace_canvas.argtypes = [P, I, P, P, I, P, P, P, P, C.POINTER(Light), I, P, C.POINTER(TraceParams), P]
    L.pm_add_mesh_attributes.argtypes = [P, P, P, I, P, I, P, P, P, P, P, P, P, I, P]
    L.pm_create_rays.argtypes = [P, P, P, I, I, C.c_ulonglong, I, P, P, P]
    L.pm_photon_emit.argtypes = [P, P, P, P, I, I, C.c_ulonglong, F, P, P, P]
    L.pm_trace_noalloc.argtypes = [P, I, P, P, I, P, C.POINTER(TraceParams), P, P]
    L.pm_trace_noalloc.restype = C.c_int
    for n in ("pm_create", "pm_destroy", "pm_build", "pm_set_geometry", "pm_tree_view",
              "pm_set_tree", "pm_tree_stats", "pm_trace", "pm_photon_bounce", "pm_add_mesh",
              "pm_trace_backward", "pm_build_batched", "pm_trace_batched", "pm_osc_surface",
              "pm_deform_radial", "pm_autotune", "pm_batch_graph_capture", "pm_batch_graph_launch",
              "pm_batch_graph_status", "pm_trace_canvas", "pm_add_mesh_attributes", "pm_create_rays",
              "pm_photon_emit"):
        getattr(L, n).restype = C.c_int
    _lib = L
    return L


def build_params(preset: str = "reference", **kw) -> BuildParams:
    """'default' = the reference TriangleMap defaults (BITMAP lower creator,
    TriangleMap.cu:27-49); 'reference' = the same with the SAH lower creator
    (TriangleMapSAHCreator, traversal cost 24); 'config' = 'reference' plus a
    forced split above 8 triangles and a 0.8 empty-side cost factor;
    'balanced' = the defaults with the balanced lower creator."""
    p = BuildParams()
    load_library().pm_build_params_default(C.byref(p))
    if preset in ("reference", "config"):
        p.lower_algorithm = LOWER_SAH
    if preset == "config":
        p.sah_leaf_max = 8
        p.sah_empty_bonus = 0.8
    elif preset == "balanced":
        p.lower_algorithm = LOWER_BALANCED
    elif preset not in ("default", "reference", "bitmap"):
        raise ValueError(preset)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def trace_params(**kw) -> TraceParams:
    p = TraceParams()
    load_library().pm_trace_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _torch():
    import torch
    return torch


def _dev_ptr(t) -> int:
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a device (cuda) tensor")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def _stream_ptr(stream) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def hits_view(hits):
    """split a [R, 4] hit tensor into (id int32, t, u, v float32) tensors"""
    torch = _torch()
    ids = hits.view(torch.int32)[:, 0]
    f = hits.view(torch.float32)
    return ids, f[:, 1], f[:, 2], f[:, 3]


def hits_to_numpy(hits) -> np.ndarray:
    """[R, 4] device hit tensor -> structured numpy array (id, t, u, v)"""
    a = hits.detach().cpu().contiguous().view(_torch().int32).numpy()
    return a.view(np.dtype([("id", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])).reshape(-1)


class Context:
    """A TriangleMap with its tracers (pm_ctx)."""

    def __init__(self, device: int = 0):
        self._L = load_library()
        h = C.c_void_p()
        rc = self._L.pm_create(C.byref(h), device)
        if rc != PM_OK:
            raise PmError(rc, "pm_create failed (no HIP device?)")
        self.h = h
        self.device = device
        self._geom = None

    def close(self):
        if getattr(self, "h", None):
            self._L.pm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != PM_OK:
            raise PmError(rc, self._L.pm_last_error(self.h).decode())

    # -- TriangleMap::Create
    def build(self, p0, p1, p2, params: BuildParams | None = None, stream=None):
        n = p0.shape[0]
        self._geom = (p0, p1, p2)
        params = params or build_params()
        self._check(self._L.pm_build(self.h, _dev_ptr(p0), _dev_ptr(p1), _dev_ptr(p2), n,
                                     C.byref(params), _stream_ptr(stream)))

    def build_batched(self, p0, p1, p2, mesh_offsets, params: BuildParams | None = None, stream=None):
        """one kd-tree per mesh (triangles of mesh m are rows
        [mesh_offsets[m], mesh_offsets[m+1]) of p0/p1/p2), built together as a
        forest whose root m is node m"""
        off = np.ascontiguousarray(mesh_offsets, dtype=np.int32)
        n = p0.shape[0]
        self._geom = (p0, p1, p2)
        params = params or build_params()
        self._check(self._L.pm_build_batched(self.h, _dev_ptr(p0), _dev_ptr(p1), _dev_ptr(p2), n,
                                             off.ctypes.data, off.shape[0] - 1, C.byref(params),
                                             _stream_ptr(stream)))

    def trace_batched(self, algo: int, org, dirs, ray_mesh, params: TraceParams | None = None,
                      out=None, counters=None, stream=None):
        """trace ray r in the tree of mesh ray_mesh[r] (device int32)"""
        torch = _torch()
        n = org.shape[0]
        if out is None:
            out = torch.empty((n, 4), dtype=torch.float32, device=org.device)
        params = params or trace_params()
        cptr = _dev_ptr(counters) if counters is not None else None
        self._check(self._L.pm_trace_batched(self.h, algo, _dev_ptr(org), _dev_ptr(dirs),
                                             _dev_ptr(ray_mesh), n, _dev_ptr(out) if n else None,
                                             C.byref(params), cptr, _stream_ptr(stream)))
        return out

    def autotune(self, p0, p1, p2, org, dirs, mesh_offsets=None, ray_mesh=None,
                 base: BuildParams | None = None, builds_per_trace: float = 1.0, reps: int = 3,
                 stream=None) -> dict:
        """time this workload over chunk size, leaf threshold, stack depth and
        ray order (pm_autotune); the context keeps the tree of the chosen setting.
        Returns the choice with its times and ready-made parameter structs."""
        base = base or build_params()
        off = None
        if mesh_offsets is not None:
            off = np.ascontiguousarray(mesh_offsets, dtype=np.int32)
        t = Tuning()
        self._geom = (p0, p1, p2)
        self._check(self._L.pm_autotune(
            self.h, _dev_ptr(p0), _dev_ptr(p1), _dev_ptr(p2), p0.shape[0],
            off.ctypes.data if off is not None else None, (off.shape[0] - 1) if off is not None else 1,
            _dev_ptr(org), _dev_ptr(dirs), _dev_ptr(ray_mesh) if ray_mesh is not None else None,
            org.shape[0], C.byref(base), builds_per_trace, reps, C.byref(t), _stream_ptr(stream)))
        res = {k: getattr(t, k) for k, _ in Tuning._fields_}
        bp = BuildParams.from_buffer_copy(base)
        bp.chunk_size, bp.lower_size = t.chunk_size, t.lower_size
        res["build_params"] = bp
        res["trace_params"] = trace_params(stack_size=t.stack_size, ray_order=t.ray_order, step_cap=t.step_cap)
        return res

    # -- one graph launch per frame
    def batch_graph_capture(self, p0, p1, p2, mesh_offsets, org, dirs, ray_mesh, out, algo=SHORT_STACK,
                            build: BuildParams | None = None, trace: TraceParams | None = None,
                            caps: dict | None = None, stream=None) -> dict:
        """record build + trace of this batch as one hipGraph (pm_batch_graph_capture);
        returns the capacities captured.  The tensors must stay alive while the
        graph is replayed.  org=None records the build alone (no rays)."""
        off = np.ascontiguousarray(mesh_offsets, dtype=np.int32)
        nrays = 0 if org is None else org.shape[0]
        ptr = (lambda t: _dev_ptr(t) if nrays else None)
        build = build or build_params()
        trace = trace or trace_params()
        c = GraphCaps()
        for k, v in (caps or {}).items():
            setattr(c, k, v)
        self._geom = (p0, p1, p2)
        self._graph_keep = (p0, p1, p2, org, dirs, ray_mesh, out, off)
        self._check(self._L.pm_batch_graph_capture(
            self.h, _dev_ptr(p0), _dev_ptr(p1), _dev_ptr(p2), p0.shape[0], off.ctypes.data, off.shape[0] - 1,
            ptr(org), ptr(dirs), ptr(ray_mesh), nrays, ptr(out), algo,
            C.byref(build), C.byref(trace), C.byref(c), _stream_ptr(stream)))
        return {k: getattr(c, k) for k, _ in GraphCaps._fields_}

    def batch_graph_launch(self, stream=None):
        self._check(self._L.pm_batch_graph_launch(self.h, _stream_ptr(stream)))

    def batch_graph_status(self, stream=None) -> int:
        st = C.c_int(0)
        self._check(self._L.pm_batch_graph_status(self.h, C.byref(st), _stream_ptr(stream)))
        return st.value

    def set_geometry(self, p0, p1, p2):
        self._geom = (p0, p1, p2)
        self._check(self._L.pm_set_geometry(self.h, _dev_ptr(p0), _dev_ptr(p1), _dev_ptr(p2),
                                            p0.shape[0]))

    def view(self) -> Tree:
        t = Tree()
        self._check(self._L.pm_tree_view(self.h, C.byref(t)))
        return t

    def tree_arrays(self) -> dict:
        """copy the tree to host numpy arrays (reference SoA layout)"""
        torch = _torch()
        torch.cuda.synchronize()
        t = self.view()
        n = t.num_nodes

        def fetch(ptr, count, dtype, shape=None):
            out = np.zeros(count if shape is None else shape, dtype=dtype)
            if count:
                _memcpy_d2h(out, ptr)
            return out

        return dict(info=fetch(t.info, n, np.int8), split_pos=fetch(t.split_pos, n, np.float32),
                    aabb_min=fetch(t.aabb_min, n, np.float32, (n, 4)),
                    aabb_max=fetch(t.aabb_max, n, np.float32, (n, 4)),
                    primitive_index=fetch(t.primitive_index, n, np.int32),
                    primitive_bitmap=fetch(t.primitive_bitmap, n, np.int64),
                    children=fetch(t.children, n, np.int32, (n, 2)),
                    surface_area=fetch(t.surface_area, n, np.float32),
                    parent=fetch(t.parent, n, np.int32),
                    prim_indices=fetch(t.prim_indices, t.num_prim_indices, np.int32),
                    leaf_ids=fetch(t.leaf_ids, t.num_leaf_ids, np.int32),
                    upper_levels=t.upper_levels, lower_levels=t.lower_levels,
                    num_roots=t.num_roots)

    def set_tree(self, arrays: dict, p0, p1, p2, has_boxes: bool = True, stream=None):
        """upload a host tree (reference SoA layout) to trace it"""
        keep = {k: np.ascontiguousarray(arrays[k]) for k in (
            "info", "split_pos", "aabb_min", "aabb_max", "primitive_index", "primitive_bitmap",
            "children", "prim_indices")}
        t = Tree()
        t.num_nodes = keep["info"].shape[0]
        for k in keep:
            if k != "prim_indices":
                setattr(t, k, keep[k].ctypes.data)
        t.num_prim_indices = keep["prim_indices"].shape[0]
        t.prim_indices = keep["prim_indices"].ctypes.data if t.num_prim_indices else None
        t.p0, t.p1, t.p2 = _dev_ptr(p0), _dev_ptr(p1), _dev_ptr(p2)
        t.num_tris = p0.shape[0]
        t.has_boxes = int(has_boxes)
        t.num_roots = int(arrays.get("num_roots", 1))
        self._geom = (p0, p1, p2)
        self._check(self._L.pm_set_tree(self.h, C.byref(t), _stream_ptr(stream)))

    def stats(self, stream=None) -> dict:
        s = Stats()
        self._check(self._L.pm_tree_stats(self.h, C.byref(s), _stream_ptr(stream)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    # -- IRayTracer::Trace
    def trace(self, algo: int, org, dirs, params: TraceParams | None = None, out=None,
              counters=None, stream=None):
        torch = _torch()
        n = org.shape[0]
        if out is None:
            out = torch.empty((n, 4), dtype=torch.float32, device=org.device)
        params = params or trace_params()
        cptr = _dev_ptr(counters) if counters is not None else None
        self._check(self._L.pm_trace(self.h, algo, _dev_ptr(org), _dev_ptr(dirs), n,
                                     _dev_ptr(out) if n else None, C.byref(params), cptr,
                                     _stream_ptr(stream)))
        return out

    def trace_canvas(self, algo: int, org, dirs, normals, colors, light, max_bounces: int = 64,
                     params: TraceParams | None = None, out=None, stream=None):
        """IRayTracer::Trace into a canvas: shaded uchar4 per ray.  normals =
        (n0, n1, n2) device float32 [N, 4]; colors device uint8 [N, 4]; light =
        12 floats (position, ambient, diffuse, specular)"""
        torch = _torch()
        n = org.shape[0]
        if out is None:
            out = torch.empty((max(n, 1), 4), dtype=torch.uint8, device=org.device)
        lt = Light()
        v = [float(x) for x in np.asarray(light, dtype=np.float32).reshape(12)]
        for i, f in enumerate(("position", "ambient", "diffuse", "specular")):
            getattr(lt, f)[:] = v[3 * i:3 * i + 3]
        params = params or trace_params()
        self._check(self._L.pm_trace_canvas(self.h, algo, _dev_ptr(org), _dev_ptr(dirs), n,
                                            *[_dev_ptr(x) for x in normals], _dev_ptr(colors), C.byref(lt),
                                            max_bounces, _dev_ptr(out), C.byref(params), _stream_ptr(stream)))
        return out[:n]

    def add_mesh_attributes(self, normals, indices, n0, n1, n2, colors=None, c0=None, c1=None, c2=None,
                            offset: int = 0, model=None, stream=None):
        """gather per-vertex normals (device float4) and colours (device
        uint8 [V, 4]) into per-corner triangle arrays rows [offset, ...)"""
        m = None
        if model is not None:
            m = np.ascontiguousarray(model, dtype=np.float32).reshape(16)
        cp = [None, None, None, None]
        if colors is not None:
            cp = [_dev_ptr(colors), _dev_ptr(c0), _dev_ptr(c1), _dev_ptr(c2)]
        self._check(self._L.pm_add_mesh_attributes(
            self.h, _dev_ptr(normals), cp[0], normals.shape[0], _dev_ptr(indices), indices.numel(),
            m.ctypes.data if m is not None else None, _dev_ptr(n0), _dev_ptr(n1), _dev_ptr(n2), cp[1], cp[2],
            cp[3], offset, _stream_ptr(stream)))

    def add_mesh(self, vertices, indices, p0, p1, p2, offset: int = 0, model=None, stream=None):
        """gather an indexed mesh (device float4 vertices, device int32/uint32
        indices) into p0/p1/p2 rows [offset, offset + len(indices)//3)"""
        m = None
        if model is not None:
            m = np.ascontiguousarray(model, dtype=np.float32).reshape(16)
        self._check(self._L.pm_add_mesh(self.h, _dev_ptr(vertices), vertices.shape[0],
                                        _dev_ptr(indices), indices.numel(),
                                        m.ctypes.data if m is not None else None,
                                        _dev_ptr(p0), _dev_ptr(p1), _dev_ptr(p2), offset,
                                        _stream_ptr(stream)))

    def osc_surface(self, vertices, normals, detail: int, height: float, time: float,
                    amplitude: float = 0.2, frequency: float = 0.3, stream=None):
        """one OscCUDAMeshNode animation step on a (detail+1)^2 vertex grid"""
        self._check(self._L.pm_osc_surface(self.h, _dev_ptr(vertices),
                                           _dev_ptr(normals) if normals is not None else None,
                                           detail, height, time, amplitude, frequency,
                                           _stream_ptr(stream)))

    def deform_radial(self, unit, frame: float, wave=(0.05, 8.0, 0.3, 6.0, 0.2), out=None,
                      stream=None):
        """unit-sphere vertices scaled by the radial wave at ``frame``"""
        torch = _torch()
        if out is None:
            out = torch.empty_like(unit)
        w = np.ascontiguousarray(wave, dtype=np.float32)
        self._check(self._L.pm_deform_radial(self.h, _dev_ptr(unit), unit.shape[0], frame,
                                             w.ctypes.data, _dev_ptr(out), _stream_ptr(stream)))
        return out

    def trace_backward(self, org, dirs, hits, grad_t=None, grad_u=None, grad_v=None, faces=None,
                       num_vertices=None, grad_vertices=None, grad_org=None, grad_dir=None,
                       tris=None, stream=None):
        """gradients of (t, u, v) w.r.t. vertices (deterministic per-vertex
        sums), ray origins and directions, for the hits of a previous trace
        over ``tris`` = (p0, p1, p2), the triangles that trace saw (default:
        the geometry last given to this context); returns (grad_vertices,
        grad_org, grad_dir)"""
        torch = _torch()
        n = org.shape[0]
        dev = org.device
        p0, p1, p2 = tris if tris is not None else self._geom
        nt = p0.shape[0]
        if num_vertices is None:
            num_vertices = (3 * nt) if faces is None else int(faces.max().item()) + 1
        if grad_vertices is None:
            grad_vertices = torch.empty((max(num_vertices, 1), 3), dtype=torch.float32, device=dev)
        if grad_org is None:
            grad_org = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
        if grad_dir is None:
            grad_dir = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
        gp = [None if g is None else _dev_ptr(g) for g in (grad_t, grad_u, grad_v)]
        tp = [_dev_ptr(x) if nt else None for x in (p0, p1, p2)]
        self._check(self._L.pm_trace_backward(
            self.h, *tp, nt, _dev_ptr(org) if n else None, _dev_ptr(dirs) if n else None,
            _dev_ptr(hits) if n else None, n, *gp,
            _dev_ptr(faces) if faces is not None else None, num_vertices, _dev_ptr(grad_vertices),
            _dev_ptr(grad_org), _dev_ptr(grad_dir), _stream_ptr(stream)))
        return grad_vertices[:num_vertices], grad_org[:n], grad_dir[:n]

    def create_rays(self, inv_view_proj, cam_pos, width: int, height: int, jitter_seed: int | None = None,
                    org=None, dirs=None, stream=None):
        """camera rays (pm_create_rays): one per pixel, id = y * width + x;
        inv_view_proj = 16 floats in Unproject's element order; jitter_seed
        None = no jitter (the reference), else a seeded sub-pixel offset.
        Returns (org, dirs) device float32 [width * height, 4]."""
        torch = _torch()
        n = width * height
        if org is None:
            org = torch.empty((max(n, 1), 4), dtype=torch.float32, device=f"cuda:{self.device}")
        if dirs is None:
            dirs = torch.empty_like(org)
        m = np.ascontiguousarray(inv_view_proj, dtype=np.float32).reshape(16)
        c = np.ascontiguousarray(cam_pos, dtype=np.float32).reshape(3)
        self._check(self._L.pm_create_rays(self.h, m.ctypes.data, c.ctypes.data, width, height,
                                           (jitter_seed or 0) & (2**64 - 1), int(jitter_seed is not None),
                                           _dev_ptr(org), _dev_ptr(dirs), _stream_ptr(stream)))
        return org[:n], dirs[:n]

    def photon_emit(self, lights, n: int, seed: int, offset: float = 1e-4, out_org=None, out_dir=None,
                    stream=None):
        """n photons from the light triangles ``lights`` = (l0, l1, l2) device
        float32 [L, 4] (pm_photon_emit); returns (org, dirs) [n, 4]"""
        torch = _torch()
        l0, l1, l2 = lights
        if out_org is None:
            out_org = torch.empty((max(n, 1), 4), dtype=torch.float32, device=l0.device)
        if out_dir is None:
            out_dir = torch.empty_like(out_org)
        self._check(self._L.pm_photon_emit(self.h, _dev_ptr(l0), _dev_ptr(l1), _dev_ptr(l2), l0.shape[0], n,
                                           seed & (2**64 - 1), offset, _dev_ptr(out_org), _dev_ptr(out_dir),
                                           _stream_ptr(stream)))
        return out_org[:n], out_dir[:n]

    def photon_bounce(self, org, dirs, hits, absorb_from: int, seed: int, offset: float = 1e-4,
                      out_org=None, out_dir=None, out_src=None, stream=None):
        """diffuse continuation rays of the surviving photons (compacted);
        returns (org, dir, src, count) with the first ``count`` rows valid"""
        torch = _torch()
        n = org.shape[0]
        if out_org is None:
            out_org = torch.empty((max(n, 1), 4), dtype=torch.float32, device=org.device)
        if out_dir is None:
            out_dir = torch.empty((max(n, 1), 4), dtype=torch.float32, device=org.device)
        if out_src is None:
            out_src = torch.empty(max(n, 1), dtype=torch.int32, device=org.device)
        cnt = C.c_int(0)
        self._check(self._L.pm_photon_bounce(
            self.h, _dev_ptr(org) if n else None, _dev_ptr(dirs) if n else None,
            _dev_ptr(hits) if n else None, n, absorb_from, seed & (2**64 - 1), offset,
            _dev_ptr(out_org), _dev_ptr(out_dir), _dev_ptr(out_src), None, C.byref(cnt),
            _stream_ptr(stream)))
        return out_org, out_dir, out_src, cnt.value


def _memcpy_d2h(out: np.ndarray, src: int):
    rc = _hip_runtime().hipMemcpy(out.ctypes.data, src, out.nbytes, 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise PmError(6, f"hipMemcpy failed ({rc})")


_hip = None


def _hip_runtime():
    global _hip
    if _hip is None:
        import ctypes
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise PmError(6, "libamdhip64.so not found")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    return _hip
