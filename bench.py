#!/usr/bin/env python3
"""Benchmark: kd-tree rebuild + photon tracing at 1M triangles on one MI355X.

Workload (the 1M-triangle configuration the headline metric is quoted on):
clustered triangle soup, 1,000,000 triangles in 64 Gaussian clusters plus a
2-triangle area light (1,000,002 triangles), 4M photons emitted from the
light on the GPU (area-weighted light choice, cosine-weighted directions),
up to 3 diffuse bounces.

One step = rebuild the SAH kd-tree (REFERENCE constants: lower size 32,
traversal cost 24, empty-space threshold 0.35) + emit the photons + trace
every photon segment with the short-stack tracer (stack depth 8), bounce
rays generated on the GPU between passes.  The triangles are resident in
HBM before the timed region.

Reported: value = trace throughput in Mrays/s (all rays of all passes over
the summed GPU time of the trace calls), the kd-build time in ms next to it,
the roofline of the dominant kernel (short-stack trace of the 4M emission
rays) from HIP events around that call, the per-configuration lines
(Cornell box, heightfield, batched meshes, differentiable intersection),
and the same build and trace algorithms run by the CPU oracle on the host
as the CPU baselines.
"""
from __future__ import annotations

import argparse
import hashlib
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E
SIMDS = 256 * 4                  # MI355X: 256 CUs x 4 SIMD units
CLOCK_GHZ = 2.4                  # peak engine clock
LANES_PER_SIMD = 16              # a wave64 VALU instruction occupies a SIMD for 4 cycles
# the library sources the tracer's instruction stream depends on: a counter
# summary is only used while they are unchanged
TRACE_SOURCES = ("pm_trace.hip", "pm_device.h", "pm_order.hip")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def source_sha() -> str:
    h = hashlib.sha256()
    for name in TRACE_SOURCES:
        with open(os.path.join(ROOT, "extensions-photonmapping_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def omp_threads() -> int:
    """threads the CPU baselines use: OMP_NUM_THREADS if set, else the cores
    this process may run on"""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(int(env), avail)) if env and env.isdigit() else avail


def ray_bytes(cnt) -> int:
    """algorithmic bytes of traced rays from per-ray counters
    (inner visits, leaf visits, triangle tests, leaf-box tests):
    32 ray + 16 hit + 13/inner node + 8/leaf + 52/triangle + 32/leaf box"""
    c = cnt.sum(dim=0).tolist()
    n = cnt.shape[0]
    return int(48 * n + 13 * c[0] + 8 * c[1] + 52 * c[2] + 32 * c[3])


def ray_counts(cnt) -> dict:
    """mean per-ray traversal counts"""
    c = cnt.float().mean(dim=0).tolist()
    return {"inner_nodes": round(c[0], 2), "leaves": round(c[1], 2), "triangle_tests": round(c[2], 2),
            "leaf_box_tests": round(c[3], 2)}


def timed(torch, fn, reps: int = 5) -> float:
    """median GPU ms of fn over reps runs after one warm-up (HIP events on
    the current stream, the one every call here launches on)"""
    import numpy as np
    stream = torch.cuda.current_stream()
    out = []
    for i in range(reps + 1):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        fn()
        ev[1].record(stream)
        torch.cuda.synchronize()
        if i:
            out.append(ev[0].elapsed_time(ev[1]))
    return float(np.median(out))


def wall(fn, reps: int = 1) -> float:
    """median host wall ms of fn (CPU baselines)"""
    import numpy as np
    out = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        out.append((time.perf_counter() - t) * 1e3)
    return float(np.median(out))


TREE_KEYS = ("num_nodes", "leaves", "empty_leaves", "max_depth", "leaf_refs", "upper_levels", "lower_levels")


def tree_stats(st: dict) -> dict:
    return {k: st[k] for k in TREE_KEYS}


# ------------------------------------------------------------------ C1
def bench_cornell(pm, scenes, torch, orc, reps: int = 5, stack: int = 8, cpu: bool = True) -> dict:
    """Cornell room (24 tris) + icosphere (5,120 tris), 256x256 jittered
    camera rays made on the GPU (pm_create_rays), CONFIG preset (forced SAH
    split above 8 triangles, 0.8 cost factor for an empty side); build, the
    three tracers, and how many hit ids equal brute force's"""
    import numpy as np
    dev = torch.device("cuda:0")
    W = H = 256
    tris = scenes.cornell_icosphere(4)
    P = [torch.from_numpy(t).to(dev) for t in tris]
    m, c = scenes.camera(scenes.CAMERA_C1, W, H)
    ctx = pm.Context(0)
    O, D = ctx.create_rays(m, c, W, H, jitter_seed=1)
    n = W * H
    rays_ms = timed(torch, lambda: ctx.create_rays(m, c, W, H, jitter_seed=1, org=O, dirs=D), reps)
    bp = pm.build_params("config")
    tp = pm.trace_params(stack_size=stack)
    build_ms = timed(torch, lambda: ctx.build(*P, bp), reps)
    out = {}
    hits = {}
    for name, algo in (("short_stack", pm.SHORT_STACK), ("kd_restart", pm.KD_RESTART), ("brute", pm.BRUTE)):
        Hh = torch.empty((n, 4), dtype=torch.float32, device=dev)
        ms = timed(torch, lambda: ctx.trace(algo, O, D, tp, out=Hh), reps)
        out[name] = {"ms": round(ms, 4), "mrays": round(n / (ms * 1e-3) / 1e6, 1)}
        hits[name] = pm.hits_to_numpy(Hh)
    same = lambda a, b: int(np.sum(hits[a]["id"] == hits[b]["id"]))
    rec = {"workload": f"Cornell room + icosphere ({tris[0].shape[0]} tris), {W}x{H} jittered camera rays "
                       f"made on the GPU, CONFIG preset (SAH, forced split above 8, empty-side factor 0.8), "
                       f"short stack depth {stack}",
           "tris": int(tris[0].shape[0]), "rays": n, "camera_rays_ms": round(rays_ms, 4),
           "kd_build_ms": round(build_ms, 4), "trace": out,
           "hit_ids_equal_brute": {"short_stack": same("short_stack", "brute"),
                                   "kd_restart": same("kd_restart", "brute"), "of": n},
           "hit_fraction": round(float(np.mean(hits["brute"]["id"] >= 0)), 4),
           "tree": tree_stats(ctx.stats())}
    if cpu:
        org, dirs = O.cpu().numpy(), D.cpu().numpy()
        th = omp_threads()
        ob = orc.build_params("config")
        tree = [None]
        b_ms = wall(lambda: tree.__setitem__(0, orc.build(tris, ob)), 3)
        res = {}
        for name, algo in (("short_stack", orc.SHORT_STACK), ("kd_restart", orc.KD_RESTART), ("brute", orc.BRUTE)):
            ms = wall(lambda: orc.trace(tree[0] if algo != orc.BRUTE else None, tris, org, dirs, algo,
                                        stack_size=stack, threads=th), 1)
            res[name] = round(n / (ms * 1e-3) / 1e6, 3)
        rec["cpu_baseline"] = {"kind": "port", "cores": th, "build_ms": round(b_ms, 2), "mrays": res,
                               "sample": "the oracle's build and all three traces of the same rays"}
    ctx.close()
    return rec


# ------------------------------------------------------------------ C2
def bench_heightfield(pm, scenes, torch, orc, reps: int = 5, stack: int = 8, cpu: bool = True) -> dict:
    """Heightfield terrain, 1024x1024 grid (2,093,058 triangles): SAH build,
    then 1920x1080 jittered camera rays made on the GPU and one
    cosine-distributed bounce ray per hit, short stack; build and the two
    trace passes timed separately (HIP events, median of ``reps`` after one
    warm-up)"""
    import numpy as np
    dev = torch.device("cuda:0")
    W, Hh = 1920, 1080
    tris = scenes.heightfield(1024, seed=2)
    ntris = int(tris[0].shape[0])
    P = [torch.from_numpy(x).to(dev) for x in tris]
    m, c = scenes.camera(scenes.CAMERA_C2, W, Hh)
    ctx = pm.Context(0)
    O, D = ctx.create_rays(m, c, W, Hh, jitter_seed=2)
    rays_ms = timed(torch, lambda: ctx.create_rays(m, c, W, Hh, jitter_seed=2, org=O, dirs=D), reps)
    n = int(O.shape[0])
    H = torch.empty((n, 4), dtype=torch.float32, device=dev)
    bp = pm.build_params("reference")
    t_in = pm.trace_params(stack_size=stack, ray_order=pm.RAY_ORDER_INPUT)
    t_co = pm.trace_params(stack_size=stack, ray_order=pm.RAY_ORDER_COHERENT)
    build_ms = timed(torch, lambda: ctx.build(*P, bp), reps)
    st = ctx.stats()
    primary_ms = timed(torch, lambda: ctx.trace(pm.SHORT_STACK, O, D, t_in, out=H), reps)
    # the same rays with the descent capped at 4 inner nodes per round of the
    # cooperative leaf test (pm_trace_params.step_cap; same records)
    t_in4 = pm.trace_params(stack_size=stack, ray_order=pm.RAY_ORDER_INPUT, step_cap=4)
    t_co4 = pm.trace_params(stack_size=stack, ray_order=pm.RAY_ORDER_COHERENT, step_cap=4)
    primary4_ms = timed(torch, lambda: ctx.trace(pm.SHORT_STACK, O, D, t_in4, out=H), reps)
    cnt = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    ctx.trace(pm.SHORT_STACK, O, D, t_in, out=H, counters=cnt)
    primary_counts = ray_counts(cnt)
    bo, bd, _, nb = ctx.photon_bounce(O, D, H, ntris, seed=2002)
    bo, bd = bo[:nb].contiguous(), bd[:nb].contiguous()
    HB = torch.empty((max(nb, 1), 4), dtype=torch.float32, device=dev)
    bounce_ms = timed(torch, lambda: ctx.trace(pm.SHORT_STACK, bo, bd, t_co, out=HB[:nb]), reps)
    bounce_in_ms = timed(torch, lambda: ctx.trace(pm.SHORT_STACK, bo, bd, t_in, out=HB[:nb]), reps)
    bounce4_ms = timed(torch, lambda: ctx.trace(pm.SHORT_STACK, bo, bd, t_co4, out=HB[:nb]), reps)
    bounce4_in_ms = timed(torch, lambda: ctx.trace(pm.SHORT_STACK, bo, bd, t_in4, out=HB[:nb]), reps)
    best_primary = min(primary_ms, primary4_ms)
    best_bounce = min(bounce_ms, bounce_in_ms, bounce4_ms, bounce4_in_ms)
    rec = {"workload": f"heightfield 1024x1024 ({ntris} tris), {W}x{Hh} jittered camera rays made on the GPU "
                       f"+ one cosine bounce ray per hit, short stack depth {stack}",
           "tris": ntris, "kd_build_ms": round(build_ms, 3), "tree": tree_stats(st),
           "camera_rays_ms": round(rays_ms, 4),
           "primary_rays": n, "primary_ms": round(primary_ms, 3), "primary_ms_step_cap4": round(primary4_ms, 3),
           "primary_mrays": round(n / (best_primary * 1e-3) / 1e6, 2), "primary_counts": primary_counts,
           "bounce_rays": int(nb), "bounce_ms_coherent": round(bounce_ms, 3),
           "bounce_ms_input_order": round(bounce_in_ms, 3),
           "bounce_ms_coherent_step_cap4": round(bounce4_ms, 3),
           "bounce_ms_input_order_step_cap4": round(bounce4_in_ms, 3),
           "trace_setting": "each pass at the fastest of its timed settings (ray order x step cap)",
           "trace_mrays": round((n + nb) / ((best_primary + best_bounce) * 1e-3) / 1e6, 2)}
    if cpu:
        th = omp_threads()
        tree = [None]
        b_ms = wall(lambda: tree.__setitem__(0, orc.build(tris, orc.build_params("reference"))), 1)
        rng = np.random.default_rng(2)
        org, dirs = O.cpu().numpy(), D.cpu().numpy()
        sel = np.sort(rng.choice(n, 131072, replace=False))
        ms = wall(lambda: orc.trace(tree[0], tris, org[sel], dirs[sel], orc.SHORT_STACK, stack_size=stack,
                                    threads=th), 1)
        bsel = sel[:1024]
        bms = wall(lambda: orc.trace(None, tris, org[bsel], dirs[bsel], orc.BRUTE, threads=th), 1)
        ost = tree[0].stats()
        rec["cpu_baseline"] = {"kind": "port", "cores": th, "build_ms": round(b_ms, 1),
                               "short_stack_mrays": round(sel.size / (ms * 1e-3) / 1e6, 3),
                               "brute_mrays": round(bsel.size / (bms * 1e-3) / 1e6, 6),
                               "sample": f"oracle build of all {ntris} tris; short stack on {sel.size} and brute "
                                         f"force on {bsel.size} seeded camera rays (scaled Mrays/s)"}
        rec["tree_oracle"] = tree_stats(ost)
        rec["tree_equal_oracle"] = tree_stats(ost) == tree_stats(st)
    ctx.close()
    return rec


# ------------------------------------------------------------------ C5
def bench_fwd_bwd(pm, scenes, torch, orc, reps: int = 5, cpu: bool = True) -> dict:
    """65,536-triangle perturbed sphere, 512x512 camera rays made on the GPU,
    loss sum (t - t_target)^2: plain forward trace vs forward + hand-written
    backward (same GPU)"""
    import numpy as np
    dev = torch.device("cuda:0")
    v_ref, v_pert, faces = scenes.perturbed_sphere(7, seed=5)
    W = 512
    m, c = scenes.camera(scenes.CAMERA_C5, W, W)
    ctx = pm.Context(0)
    O, D = ctx.create_rays(m, c, W, W, jitter_seed=5)
    F = torch.from_numpy(faces.reshape(-1)).to(dev)
    nt = faces.shape[0]

    def tris(V):
        P = [torch.empty((nt, 4), dtype=torch.float32, device=dev) for _ in range(3)]
        V4 = torch.ones((V.shape[0], 4), dtype=torch.float32, device=dev)
        V4[:, :3] = torch.from_numpy(V).to(dev)
        ctx.add_mesh(V4, F, *P)
        return P

    tp = pm.trace_params(stack_size=8)
    ctx.build(*tris(v_ref), pm.build_params("reference"))
    ids_ref, t_ref, _, _ = pm.hits_view(ctx.trace(pm.SHORT_STACK, O, D, tp))
    t_ref = t_ref.clone()
    ok_ref = ids_ref >= 0
    P = tris(v_pert)
    ctx.build(*P, pm.build_params("reference"))
    H = torch.empty((O.shape[0], 4), dtype=torch.float32, device=dev)
    gt = torch.empty(O.shape[0], dtype=torch.float32, device=dev)
    nv = v_pert.shape[0]
    gv = torch.empty((nv, 3), dtype=torch.float32, device=dev)
    go = torch.empty((O.shape[0], 3), dtype=torch.float32, device=dev)
    gd = torch.empty_like(go)
    stream = torch.cuda.current_stream()
    fwd, bwd = [], []
    for i in range(reps + 1):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record(stream)
        ctx.trace(pm.SHORT_STACK, O, D, tp, out=H)
        e[1].record(stream)
        ids, t, _, _ = pm.hits_view(H)
        mask = (ids >= 0) & ok_ref
        torch.where(mask, 2.0 * (t - t_ref), torch.zeros_like(t), out=gt)
        e[2].record(stream)
        ctx.trace_backward(O, D, H, gt, None, None, F, nv, gv, go, gd, tris=P)
        e[3].record(stream)
        torch.cuda.synchronize()
        if i:
            fwd.append(e[0].elapsed_time(e[1]))
            bwd.append(e[2].elapsed_time(e[3]))
    f, b = float(np.median(fwd)), float(np.median(bwd))
    rec = {"workload": f"{nt}-triangle perturbed sphere, {O.shape[0]} camera rays, loss sum (t - t_target)^2",
           "fwd_trace_ms": round(f, 4), "bwd_ms": round(b, 4), "fwd_bwd_over_fwd": round((f + b) / f, 3)}
    if cpu:
        th = omp_threads()
        host = [p.cpu().numpy() for p in P]
        org, dirs = O.cpu().numpy(), D.cpu().numpy()
        tree = orc.build(host, orc.build_params("reference"))
        hits = [None]
        f_ms = wall(lambda: hits.__setitem__(0, orc.trace(tree, host, org, dirs, orc.SHORT_STACK, stack_size=8,
                                                          threads=th)), 1)
        g = gt.cpu().numpy()
        b_ms = wall(lambda: orc.trace_backward(host, org, dirs, hits[0], g, None, None, faces, nv, fp64=True), 1)
        rec["cpu_baseline"] = {"kind": "port", "cores": th, "fwd_ms": round(f_ms, 1), "bwd_fp64_ms": round(b_ms, 1),
                               "fwd_bwd_over_fwd": round((f_ms + b_ms) / f_ms, 3),
                               "sample": "oracle FP32 short-stack forward (OpenMP) + FP64 adjoint (1 thread) "
                                         "of all rays"}
    ctx.close()
    return rec


# ------------------------------------------------------------------ C4
def bench_batched(pm, scenes, torch, orc, meshes: int = 256, rays_per_mesh: int = 65536,
                  frames: int = 4, per_mesh_frames: int = 2, stack: int = 8, cpu: bool = True) -> dict:
    """Animated meshes rebuilt every frame: ``meshes`` cube spheres of
    log-uniform size (1k-200k triangles), radial wave deformed on the GPU,
    gathered into one triangle SoA, then either ONE forest build + ONE
    batched trace (batched) or one build + one trace per mesh (per-mesh),
    ``rays_per_mesh`` rays each; wall time per frame, same GPU, same
    deform/gather step in both."""
    import numpy as np
    dev = torch.device("cuda:0")
    t0 = time.time()
    c = scenes.config4(num_meshes=meshes, rays_per_mesh=rays_per_mesh)
    log(f"[bench] batched scene generated in {time.time() - t0:.1f}s")
    ctx = pm.Context(0)
    U = torch.from_numpy(c["unit"]).to(dev)
    V = torch.empty_like(U)
    F = torch.from_numpy(c["faces"].reshape(-1)).to(dev)
    off = c["tri_off"]
    n = int(off[-1])
    P = [torch.empty((n, 4), dtype=torch.float32, device=dev) for _ in range(3)]
    O, D = torch.from_numpy(c["org"]).to(dev), torch.from_numpy(c["dirs"]).to(dev)
    RM = torch.from_numpy(c["ray_mesh"]).to(dev)
    H = torch.empty((O.shape[0], 4), dtype=torch.float32, device=dev)
    bp = pm.build_params("reference")
    tp = pm.trace_params(stack_size=stack)
    stream = torch.cuda.current_stream()
    R = rays_per_mesh

    def stage(f):
        ctx.deform_radial(U, float(f), out=V)
        ctx.add_mesh(V, F, *P)

    def batched(f, ev=None, bp=bp, tp=tp):
        stage(f)
        if ev:
            ev[0].record(stream)
        ctx.build_batched(*P, off, bp)
        if ev:
            ev[1].record(stream)
        ctx.trace_batched(pm.SHORT_STACK, O, D, RM, tp, out=H)
        if ev:
            ev[2].record(stream)

    def per_mesh(f):
        stage(f)
        for mm in range(meshes):
            a, b = int(off[mm]), int(off[mm + 1])
            ctx.build(P[0][a:b], P[1][a:b], P[2][a:b], bp)
            ctx.trace(pm.SHORT_STACK, O[mm * R:(mm + 1) * R], D[mm * R:(mm + 1) * R], tp,
                      out=H[mm * R:(mm + 1) * R])

    def run_batched(bp, tp):
        batched(0, bp=bp, tp=tp)
        bt, bb, btr = [], [], []
        for f in range(frames):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            torch.cuda.synchronize()
            t = time.perf_counter()
            batched(f + 1, ev, bp, tp)
            torch.cuda.synchronize()
            bt.append((time.perf_counter() - t) * 1e3)
            bb.append(ev[0].elapsed_time(ev[1]))
            btr.append(ev[1].elapsed_time(ev[2]))
        return float(np.median(bt)), float(np.median(bb)), float(np.median(btr))

    per_mesh(0)
    b, b_build, b_trace = run_batched(bp, tp)
    hit_frac = float((H.view(torch.int32)[:, 0] >= 0).float().mean())
    b_roof = build_roofline(ctx.stats(), n, b_build)
    # autotuned chunk size, leaf threshold and stack depth for this workload
    stage(0)
    tune = ctx.autotune(*P, O, D, mesh_offsets=off, ray_mesh=RM, base=bp, reps=2)
    tb, tb_build, tb_trace = run_batched(tune["build_params"], tune["trace_params"])
    # the whole rebuild + trace as one captured graph, replayed once per frame
    gctx = pm.Context(0)
    stage(0)
    caps = gctx.batch_graph_capture(*P, off, O, D, RM, H, pm.SHORT_STACK, bp, tp)
    gt, gk = [], []
    for f in range(frames):
        stage(f + 1)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        t = time.perf_counter()
        ev[0].record(stream)
        gctx.batch_graph_launch()
        ev[1].record(stream)
        torch.cuda.synchronize()
        gt.append((time.perf_counter() - t) * 1e3)
        gk.append(ev[0].elapsed_time(ev[1]))
    graph_status = gctx.batch_graph_status()
    gctx.close()
    pt = []
    for f in range(per_mesh_frames):
        torch.cuda.synchronize()
        t = time.perf_counter()
        per_mesh(f + 1)
        torch.cuda.synchronize()
        pt.append((time.perf_counter() - t) * 1e3)
    p = float(np.median(pt))
    rec = {"workload": f"{meshes} animated cube spheres, {n} tris, {O.shape[0]} rays per frame "
                       f"(radial wave deform + gather + SAH rebuild + short-stack depth {stack})",
           "batched_frame_ms": round(b, 3), "per_mesh_frame_ms": round(p, 3),
           "speedup": round(p / b, 2),
           "batched_build_ms": round(b_build, 3),
           "batched_build_roofline": b_roof,
           "batched_trace_ms": round(b_trace, 3),
           "batched_trace_mrays": round(O.shape[0] / (b_trace * 1e-3) / 1e6, 1),
           "hit_fraction": round(hit_frac, 4),
           "graph": {"launch_ms": round(float(np.median(gt)), 3),
                     "launch_gpu_ms": round(float(np.median(gk)), 3),
                     "speedup_vs_per_mesh": round(p / float(np.median(gt)), 2),
                     "status": graph_status, "upper_levels": caps["upper_levels"],
                     "lower_levels": caps["lower_levels"]},
           "autotuned": {"chunk_size": tune["chunk_size"], "lower_size": tune["lower_size"],
                         "stack_size": tune["stack_size"], "trials": tune["trials"],
                         "ray_order": ["input", "coherent"][tune["ray_order"]],
                         "frame_ms": round(tb, 3), "build_ms": round(tb_build, 3),
                         "trace_ms": round(tb_trace, 3), "speedup_vs_per_mesh": round(p / tb, 2)}}
    if cpu:
        th = omp_threads()
        stage(1)
        host = [x.cpu().numpy() for x in P]
        tree = [None]
        b_ms = wall(lambda: tree.__setitem__(0, orc.build_batched(host, off, orc.build_params("reference"))), 1)
        rng = np.random.default_rng(4)
        sel = np.sort(rng.choice(O.shape[0], 524288, replace=False))
        org, dirs, rm = c["org"][sel], c["dirs"][sel], c["ray_mesh"][sel]
        t_ms = wall(lambda: orc.trace_batched(tree[0], host, org, dirs, rm, orc.SHORT_STACK, stack_size=stack,
                                              threads=th), 1)
        frame = b_ms + t_ms * O.shape[0] / sel.size
        rec["cpu_baseline"] = {"kind": "port", "cores": th, "forest_build_ms": round(b_ms, 1),
                               "trace_mrays": round(sel.size / (t_ms * 1e-3) / 1e6, 3),
                               "frame_ms": round(frame, 1), "gpu_batched_speedup": round(frame / b, 1),
                               "sample": f"oracle forest build of the whole frame ({n} tris); short stack on "
                                         f"{sel.size} of the frame's rays, scaled to all {O.shape[0]}"}
    ctx.close()
    return rec


# ------------------------------------------------------------------ roofline
def trace_roofline(pmc_path: str, n0: int, ntris: int, stack: int, call_ms: float, alg_bytes: int,
                   counts: dict) -> dict:
    """The emission-batch trace call is bound by vector-instruction issue: the
    fraction is the VALU issue rate, SQ_INSTS_VALU x 4 cycles per wave64
    instruction over (SIMDs x the call's cycles at the peak clock).  The
    instruction count and the HBM traffic per call come from the rocprofv3
    counter runs summarised in profiles/ (used only while the tracer's
    sources hash to the value stamped there, and for the same workload); the
    call's duration is measured here."""
    rec = {"bound": "valu", "kernel": "trace_coop_kernel<short stack> emission call (ray sort + trace; "
                                      "HIP events bracket the call)",
           "launch_ms": round(call_ms, 4), "algorithmic_bytes": int(alg_bytes),
           "algorithmic_gbs": round(alg_bytes / (call_ms * 1e-3) / 1e9, 1),
           "ray_counts": counts,
           "peak": round(SIMDS * LANES_PER_SIMD * CLOCK_GHZ, 1), "unit": "G lane-instructions/s",
           "achieved": None, "frac": None, "traffic": None, "hbm_frac": None}
    sha = source_sha()
    if not os.path.exists(pmc_path):
        rec["pmc"] = "missing"
        return rec
    pmc = json.load(open(pmc_path))
    ok = (pmc.get("source_sha") == sha and pmc.get("rays") == n0 and pmc.get("tris") == ntris
          and pmc.get("stack") == stack and pmc.get("ray_order") == "coherent")
    rec["pmc"] = {"file": os.path.relpath(pmc_path, ROOT), "source_sha": pmc.get("source_sha"),
                  "current_sha": sha, "used": ok}
    if not ok:
        return rec
    valu = pmc["valu_insts"]
    cycles = call_ms * 1e-3 * CLOCK_GHZ * 1e9
    rec["achieved"] = round(valu * 64 / (call_ms * 1e-3) / 1e9, 1)
    rec["frac"] = round(valu * 4 / (SIMDS * cycles), 4)
    rec["lane_efficiency"] = pmc.get("lane_efficiency")
    rec["useful_lane_frac"] = round(rec["frac"] * pmc["lane_efficiency"], 4) if pmc.get("lane_efficiency") else None
    rec["valu_insts"] = valu
    rec["traffic"] = pmc["traffic_bytes"]
    rec["hbm_frac"] = round(pmc["traffic_bytes"] / (call_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return rec


def build_roofline(st: dict, ntris: int, build_ms: float, graph_ms: float = None) -> dict:
    """the build's algorithmic bytes (reference record sizes: 80 B per
    triangle for the box pass, 136 B per ref read and 32 B per ref written at
    every upper level, 136 B per final leaf ref for areas and planes, 52 B per
    prim of every SAH lower node) over the measured build time (and over the
    captured graph's replay time when given)"""
    b = (80 * ntris + 136 * st["upper_ref_reads"] + 32 * st["upper_ref_writes"]
         + 136 * st["num_prim_indices"] + 52 * st["lower_prims"])
    achieved = b / (build_ms * 1e-3) / 1e9
    rec = {"bound": "hbm", "algorithmic_bytes": int(b), "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)}
    if graph_ms:
        g = b / (graph_ms * 1e-3) / 1e9
        rec.update({"graph_achieved": round(g, 1), "graph_frac": round(g / HBM_PEAK_GBS, 4)})
    rec["counts"] = {"tris": ntris, "upper_ref_reads": st["upper_ref_reads"],
                     "upper_ref_writes": st["upper_ref_writes"], "leaf_refs": st["num_prim_indices"],
                     "lower_prims": st["lower_prims"]}
    return rec


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--photons", type=int, default=4_000_000)
    ap.add_argument("--bounces", type=int, default=3)
    ap.add_argument("--stack", type=int, default=8)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000,
                    help="rays the CPU baseline traces (0 disables every CPU baseline)")
    ap.add_argument("--meshes", type=int, default=256,
                    help="meshes of the batched-rebuild comparison (0 skips it)")
    ap.add_argument("--configs", default="c1,c2,c4,c5",
                    help="the other configurations to measure: any of c1 (Cornell box), c2 (heightfield), "
                         "c4 (batched meshes), c5 (differentiable intersection), or 'none'")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched-rebuild comparison and print its record (profiling runs)")
    ap.add_argument("--builds-only", nargs="?", const="c3", choices=("c3", "c4"), default=None,
                    help="profiling runs: rebuild the 1M-triangle scene (c3, the default) or the batched frame of "
                         "--meshes meshes (c4) --warmup + --steps times, print the build times, then copy 256 MiB "
                         "on the device (the HBM-counter calibration) and exit")
    ap.add_argument("--probe", nargs="?", const="c3", choices=("c3", "c2"), default=None,
                    help="counter runs: c3 (the default) = one build and one short-stack trace of the emission "
                         "batch; c2 = the heightfield build, its camera rays traced in input order, one cosine "
                         "bounce ray per hit traced in coherent order; then a 256 MiB device copy (the "
                         "HBM-counter calibration), then exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}): this build "
              "measures one MI355X; sharding over GPUs is a design note, not a run.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    pm = importlib.import_module("extensions-photonmapping_amd")
    scenes = importlib.import_module("extensions-photonmapping_amd.scenes")
    photons = importlib.import_module("extensions-photonmapping_amd.photons")
    cpu_on = args.cpu_sample > 0
    orc = None
    if cpu_on:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    configs = set() if args.configs == "none" else set(args.configs.split(","))

    def calibration_copy() -> int:
        src_buf = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst_buf = torch.empty_like(src_buf)
        dst_buf.copy_(src_buf)
        torch.cuda.synchronize()
        return src_buf.numel() * 4

    if args.builds_only == "c4":
        c = scenes.config4(num_meshes=args.meshes, rays_per_mesh=1)
        ctx = pm.Context(0)
        U = torch.from_numpy(c["unit"]).to(dev)
        V = torch.empty_like(U)
        F = torch.from_numpy(c["faces"].reshape(-1)).to(dev)
        off = c["tri_off"]
        n = int(off[-1])
        P = [torch.empty((n, 4), dtype=torch.float32, device=dev) for _ in range(3)]
        ctx.deform_radial(U, 1.0, out=V)
        ctx.add_mesh(V, F, *P)
        bp = pm.build_params("reference")
        ms = []
        for i in range(args.warmup + args.steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ctx.build_batched(*P, off, bp)
            e1.record()
            torch.cuda.synchronize()
            if i >= args.warmup:
                ms.append(e0.elapsed_time(e1))
            log(f"[bench] batched build {i + 1}/{args.warmup + args.steps}: {e0.elapsed_time(e1):.3f} ms")
        print(json.dumps({"builds_only": "c4", "meshes": args.meshes, "tris": n,
                          "kd_build_ms": [round(x, 4) for x in ms], "tree": tree_stats(ctx.stats()),
                          "copy_bytes": calibration_copy()}), flush=True)
        return 0

    if args.probe == "c2":
        W, Hh = 1920, 1080
        tris = scenes.heightfield(1024, seed=2)
        ntris = int(tris[0].shape[0])
        P = [torch.from_numpy(x).to(dev) for x in tris]
        m, c = scenes.camera(scenes.CAMERA_C2, W, Hh)
        ctx = pm.Context(0)
        ctx.build(*P, pm.build_params("reference"))
        O, D = ctx.create_rays(m, c, W, Hh, jitter_seed=2)
        H = ctx.trace(pm.SHORT_STACK, O, D, pm.trace_params(stack_size=args.stack, ray_order=pm.RAY_ORDER_INPUT))
        bo, bd, _, nb = ctx.photon_bounce(O, D, H, ntris, seed=2002)
        bo, bd = bo[:nb].contiguous(), bd[:nb].contiguous()
        ctx.trace(pm.SHORT_STACK, bo, bd, pm.trace_params(stack_size=args.stack, ray_order=pm.RAY_ORDER_COHERENT))
        torch.cuda.synchronize()
        print(json.dumps({"probe": "c2", "tris": ntris, "primary_rays": int(O.shape[0]), "bounce_rays": int(nb),
                          "copy_bytes": calibration_copy(), "source_sha": source_sha()}), flush=True)
        return 0

    if args.batched_only:
        print(json.dumps({"batched": bench_batched(pm, scenes, torch, orc, meshes=args.meshes, cpu=False)}),
              flush=True)
        return 0

    t0 = time.time()
    tris = scenes.cluster_soup(args.tris)
    log(f"[bench] scene {tris[0].shape[0]} tris generated in {time.time() - t0:.1f}s")
    P = [torch.from_numpy(t).to(dev) for t in tris]
    ntris = tris[0].shape[0]
    LIGHT = [p[ntris - 2:].contiguous() for p in P]   # the two area-light triangles
    absorb_from = ntris - 2                          # photons reaching the light are absorbed
    n0 = args.photons
    O = [torch.empty((n0, 4), dtype=torch.float32, device=dev) for _ in range(2)]
    D = [torch.empty((n0, 4), dtype=torch.float32, device=dev) for _ in range(2)]
    H = torch.empty((n0, 4), dtype=torch.float32, device=dev)
    src = torch.empty(n0, dtype=torch.int32, device=dev)
    ctx = pm.Context(0)
    bp = pm.build_params("reference")
    # photons and bounce rays are incoherent: the tracer sorts them by a
    # spatial key inside each trace call (included in the times)
    tp = pm.trace_params(stack_size=args.stack, ray_order=pm.RAY_ORDER_COHERENT)
    stream = torch.cuda.current_stream()
    EMIT_SEED = 3

    if args.builds_only:
        ms = []
        for i in range(args.warmup + args.steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            ctx.build(*P, bp)
            e1.record(stream)
            torch.cuda.synchronize()
            if i >= args.warmup:
                ms.append(e0.elapsed_time(e1))
            log(f"[bench] build {i + 1}/{args.warmup + args.steps}: {e0.elapsed_time(e1):.3f} ms")
        print(json.dumps({"builds_only": "c3", "tris": ntris, "kd_build_ms": [round(x, 4) for x in ms],
                          "tree": tree_stats(ctx.stats()), "copy_bytes": calibration_copy()}), flush=True)
        return 0

    if args.probe:
        ctx.build(*P, bp)
        ctx.photon_emit(LIGHT, n0, EMIT_SEED, out_org=O[0], out_dir=D[0])
        ctx.trace(pm.SHORT_STACK, O[0], D[0], tp, out=H)
        torch.cuda.synchronize()
        print(json.dumps({"probe": True, "copy_bytes": calibration_copy(), "rays": n0,
                          "tris": ntris, "source_sha": source_sha()}), flush=True)
        return 0

    # the photon driver: emission + passes, the ray order chosen per pass
    driver = photons.PhotonTracer(ctx, LIGHT, absorb_from, stack_size=args.stack, capacity=n0)

    def step(timing: dict | None, orders=None):
        """one frame: rebuild + emission + all photon passes; returns rays traced"""
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e[0].record(stream)
        tb = time.perf_counter()
        ctx.build(*P, bp)
        e[1].record(stream)
        build_wall = time.perf_counter() - tb
        passes = driver.trace(n0, args.bounces, EMIT_SEED, orders=orders, timed=timing is not None)
        rays = sum(p.count for p in passes)
        if timing is not None:
            timing["build_ms"].append(e[0].elapsed_time(e[1]))
            timing["emit_ms"].append(driver.emit_ms)
            timing["build_wall_ms"].append(build_wall * 1e3)
            timing["trace_ms"].append(sum(p.ms for p in passes))
            timing["pass0_ms"].append(passes[0].ms)
            timing["pass_ms"].append([p.ms for p in passes])
            timing["rays"].append(rays)
            timing["pass_rays"] = [p.count for p in passes]
        return rays

    keys = ("build_ms", "emit_ms", "build_wall_ms", "trace_ms", "pass0_ms", "pass_ms", "rays")
    ctx.build(*P, bp)
    pass_orders = driver.tune_ray_orders(n0, args.bounces, EMIT_SEED, reps=1)
    order_names = ["coherent" if o == pm.RAY_ORDER_COHERENT else "input" for o in pass_orders]
    log(f"[bench] ray order per pass (tuned): {order_names}")
    for i in range(args.warmup):
        step(None)
        log(f"[bench] warmup {i + 1}/{args.warmup}")
    timing = {k: [] for k in keys}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(timing)
        log(f"[bench] step {i + 1}/{args.steps}: build {timing['build_ms'][-1]:.2f} ms, emit "
            f"{timing['emit_ms'][-1]:.3f} ms, trace {timing['trace_ms'][-1]:.2f} ms, {timing['rays'][-1]} rays")
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t_start
    ms_per_step = wall_s * 1e3 / args.steps
    rays_total = sum(timing["rays"])
    trace_ms_total = sum(timing["trace_ms"])
    mrays = rays_total / (trace_ms_total * 1e-3) / 1e6
    build_ms = float(np.median(timing["build_ms"]))
    # the same passes traced in input order (no ray sort)
    t_in = {k: [] for k in keys}
    step(None, [pm.RAY_ORDER_INPUT])
    for i in range(2):
        step(t_in, [pm.RAY_ORDER_INPUT])
    input_mrays = sum(t_in["rays"]) / (sum(t_in["trace_ms"]) * 1e-3) / 1e6
    log(f"[bench] input ray order: {input_mrays:.1f} Mrays/s")

    # the same rebuild recorded once as a hipGraph (no host read-backs) and replayed
    gctx = pm.Context(0)
    off1 = np.array([0, ntris], np.int32)
    gctx.batch_graph_capture(*P, off1, None, None, None, None, pm.SHORT_STACK, bp)
    build_graph_ms = timed(torch, gctx.batch_graph_launch, args.steps)
    assert gctx.batch_graph_status() == 0
    gctx.close()

    # ---- roofline of the dominant kernel: short-stack trace of the emission batch
    ctx.build(*P, bp)
    ctx.photon_emit(LIGHT, n0, EMIT_SEED, out_org=O[0], out_dir=D[0])
    O0, D0 = O[0].clone(), D[0].clone()
    cnt = torch.zeros((n0, 4), dtype=torch.int32, device=dev)
    ctx.trace(pm.SHORT_STACK, O0, D0, tp, out=H, counters=cnt)
    torch.cuda.synchronize()
    bytes_pass0 = ray_bytes(cnt.to(torch.int64))
    counts0 = ray_counts(cnt)
    hits0 = H.view(torch.int32)[:, 0]
    hit_frac = float((hits0 >= 0).float().mean())
    # the call the counter summary describes: the emission batch in coherent
    # order (the photon driver may have timed its first pass in input order)
    roof_ms = timed(torch, lambda: ctx.trace(pm.SHORT_STACK, O0, D0, tp, out=H), max(args.steps, 3))
    roof = trace_roofline(os.path.join(ROOT, "profiles", "trace_c3_pmc.json"), n0, ntris, args.stack, roof_ms,
                          bytes_pass0, counts0)

    # ---- other tracers on the same tree (same GPU): kd-restart, brute force
    kd_ms = timed(torch, lambda: ctx.trace(pm.KD_RESTART, O0, D0, tp, out=H), 2)
    kd_mrays = n0 / (kd_ms * 1e-3) / 1e6
    nb = 16384
    ob, db = O0[:nb].contiguous(), D0[:nb].contiguous()
    brute_ms = timed(torch, lambda: ctx.trace(pm.BRUTE, ob, db, tp, out=H[:nb]), 1)
    brute_mrays = nb / (brute_ms * 1e-3) / 1e6
    st = ctx.stats()

    extra = {}
    if "c1" in configs:
        extra["cornell"] = bench_cornell(pm, scenes, torch, orc, stack=args.stack, cpu=cpu_on)
        log(f"[bench] cornell: {extra['cornell']}")
    if "c5" in configs:
        extra["fwd_bwd"] = bench_fwd_bwd(pm, scenes, torch, orc, cpu=cpu_on)
        log(f"[bench] fwd+bwd: {extra['fwd_bwd']}")
    if "c2" in configs:
        extra["heightfield"] = bench_heightfield(pm, scenes, torch, orc, stack=args.stack, cpu=cpu_on)
        log(f"[bench] heightfield: {extra['heightfield']}")
    if "c4" in configs and args.meshes > 0:
        extra["batched"] = bench_batched(pm, scenes, torch, orc, meshes=args.meshes, cpu=cpu_on)
        log(f"[bench] batched: {extra['batched']}")

    # ---- CPU baseline: the oracle's build and trace on the host cores
    cpu = None
    tree_oracle = None
    if cpu_on:
        threads = omp_threads()
        m = min(args.cpu_sample, n0)
        org0, dir0 = O0[:m].cpu().numpy(), D0[:m].cpu().numpy()
        tree = [None]
        cpu_build_ms = wall(lambda: tree.__setitem__(0, orc.build(tris, orc.build_params("reference"))), 1)
        tree_oracle = tree_stats(tree[0].stats())
        t0 = time.perf_counter()
        orc.trace(tree[0], tris, org0, dir0, orc.SHORT_STACK, stack_size=args.stack, threads=threads)
        cpu_trace_s = time.perf_counter() - t0
        cpu = {"value": round(m / cpu_trace_s / 1e6, 4), "unit": "Mrays/s", "cores": threads,
               "kind": "port", "build_ms": round(cpu_build_ms, 1),
               "sample": f"oracle SAH kd-tree build of the full {ntris}-triangle scene + short-stack "
                         f"(depth {args.stack}) trace of the first {m} emission photon rays, "
                         f"OpenMP {threads} threads (the cores this process may use)"}
        log(f"[bench] cpu baseline: build {cpu_build_ms:.0f} ms, {cpu['value']} Mrays/s")

    line = {
        "metric": "Mrays/s + kd-build ms @1M tris; batched & fwd+bwd vs plain, same GPU",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"photon tracing on a clustered triangle soup: {ntris} tris (64 Gaussian "
                        f"clusters + 2-tri light), {n0} photons emitted on the GPU, <= {args.bounces} diffuse "
                        f"bounces, SAH kd-tree rebuilt every step, short-stack depth {args.stack}",
            "tris": ntris,
            "rays_per_step": int(np.mean(timing["rays"])),
            "rays_per_pass": timing["pass_rays"],
            "kd_build_ms": round(build_ms, 3),
            "kd_build_wall_ms": round(float(np.median(timing["build_wall_ms"])), 3),
            "kd_build_graph_ms": round(build_graph_ms, 3),
            "emit_ms": round(float(np.median(timing["emit_ms"])), 4),
            "trace_ms_per_step": round(trace_ms_total / args.steps, 3),
            "ray_order": {"per_pass": order_names,
                          "chosen_by": "the photon driver's tuning run (both orders timed per pass, before the "
                                       "timed steps); coherent = spatial-key radix sort inside the trace call, "
                                       "included in the times"},
            "pass_ms": [round(float(np.median([t[b] for t in timing["pass_ms"]])), 3)
                        for b in range(len(timing["pass_rays"]))],
            "input_order_mrays": round(input_mrays, 2),
            "input_order_trace_ms_per_step": round(sum(t_in["trace_ms"]) / len(t_in["trace_ms"]), 3),
            "kd_restart_mrays": round(kd_mrays, 2),
            "brute_force_mrays": round(brute_mrays, 4),
            "short_stack_vs_brute": round(mrays / brute_mrays, 1),
            "emission_hit_fraction": round(hit_frac, 4),
            "tree": tree_stats(st),
            "tree_oracle": tree_oracle,
            "tree_equal_oracle": (tree_oracle == tree_stats(st)) if tree_oracle else None,
            "build_roofline": build_roofline(st, ntris, build_ms, build_graph_ms),
            **extra,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
