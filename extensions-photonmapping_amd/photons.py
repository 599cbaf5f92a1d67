"""Photon emission and bounce driver over a built context.

No counterpart in the reference: its tracers (Raytracer.cu, ShortStack.cu)
serve a photon-mapping renderer the extension does not contain.  The driver
emits photons from area lights on the GPU (pm_photon_emit), traces them with
a tree tracer (pm_trace), and continues every photon that hit a surface
below ``absorb_from`` with a cosine-weighted bounce ray (pm_photon_bounce),
up to ``bounces`` times.  Every pass's rays, hits and source indices stay on
the device.  The ray order (the coherent spatial sort or the input order) is
chosen per pass: given, or measured by ``tune_ray_orders``, which times both
orders of every pass and keeps the faster one (emission rays from a small
light gain from the sort; short, scattered bounce rays may not).

The C++ counterpart is ``pmkd::Utils::PhotonTracer`` (include/pmkd/PhotonTracer.hpp).
"""
from __future__ import annotations

import importlib

_pm = importlib.import_module(__package__)


class PhotonPass:
    """one pass: its rays (org, dirs [n, 4]), their first hits [n, 4]
    (pm_hit records), and for passes after the first the index of the ray of
    the previous pass each ray continues (src [n], int32); ms is the trace
    call's GPU time when the pass was timed"""

    def __init__(self, org, dirs, hits, src, count, order, ms=None):
        self.org, self.dirs, self.hits, self.src, self.count = org, dirs, hits, src, count
        self.order, self.ms = order, ms


class PhotonTracer:
    def __init__(self, ctx, lights, absorb_from: int, algo: int = None, stack_size: int = 8,
                 offset: float = 1e-4, capacity: int = 0):
        """ctx: a Context holding the tree (and its triangles); lights = (l0,
        l1, l2) device float32 [L, 4] light triangles; photons reaching a
        triangle id >= absorb_from end there (the light's own ids, or the
        triangle count for none)"""
        self.ctx = ctx
        self.lights = lights
        self.absorb_from = int(absorb_from)
        self.algo = _pm.SHORT_STACK if algo is None else algo
        self.stack_size = stack_size
        self.offset = offset
        self.orders: list[int] = [_pm.RAY_ORDER_COHERENT]
        self._cap = 0
        self._bufs = []
        if capacity:
            self._reserve(capacity, 1)

    def _reserve(self, n: int, passes: int):
        torch = _pm._torch()
        dev = self.lights[0].device
        if n > self._cap:
            self._bufs = []
            self._cap = n
        while len(self._bufs) < passes:
            self._bufs.append({
                "org": torch.empty((max(n, 1), 4), dtype=torch.float32, device=dev),
                "dirs": torch.empty((max(n, 1), 4), dtype=torch.float32, device=dev),
                "hits": torch.empty((max(n, 1), 4), dtype=torch.float32, device=dev),
                "src": torch.empty(max(n, 1), dtype=torch.int32, device=dev)})

    def order_of(self, b: int) -> int:
        """the ray order of pass b (the last given order for later passes)"""
        return self.orders[b] if b < len(self.orders) else self.orders[-1]

    def set_ray_orders(self, orders):
        if not orders:
            raise ValueError("at least one ray order")
        for o in orders:
            if o not in (_pm.RAY_ORDER_INPUT, _pm.RAY_ORDER_COHERENT):
                raise ValueError(f"unknown ray order {o}")
        self.orders = list(orders)

    def trace(self, photons: int, bounces: int, seed: int, orders=None, timed: bool = False,
              stream=None) -> list[PhotonPass]:
        """emit `photons` photons (PCG32 streams of `seed`), trace them, and
        bounce the survivors up to `bounces` times; pass b + 1 uses bounce
        seed seed + 1000 (b + 1).  Returns the passes (views of the driver's
        buffers, valid until its next trace)."""
        torch = _pm._torch()
        if photons < 0 or bounces < 0:
            raise ValueError("photons and bounces must be >= 0")
        self._reserve(photons, bounces + 1)
        ords = list(orders) if orders is not None else [self.order_of(b) for b in range(bounces + 1)]
        while len(ords) < bounces + 1:
            ords.append(ords[-1])
        b0 = self._bufs[0]
        s = stream if stream is not None else torch.cuda.current_stream()
        if timed:
            emit_ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            emit_ev[0].record(s)
        self.ctx.photon_emit(self.lights, photons, seed, self.offset, out_org=b0["org"], out_dir=b0["dirs"],
                             stream=stream)
        if timed:
            emit_ev[1].record(s)
        n, out, src = photons, [], None
        for b in range(bounces + 1):
            buf = self._bufs[b]
            tp = _pm.trace_params(stack_size=self.stack_size, ray_order=ords[b])
            ev = None
            if timed:
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record(s)
            hits = self.ctx.trace(self.algo, buf["org"][:n], buf["dirs"][:n], tp, out=buf["hits"][:n],
                                  stream=stream) if n else buf["hits"][:0]
            if timed:
                ev[1].record(s)
            out.append(PhotonPass(buf["org"][:n], buf["dirs"][:n], hits, src, n, ords[b], ev))
            if b == bounces or n == 0:
                break
            nxt = self._bufs[b + 1]
            _, _, _, m = self.ctx.photon_bounce(buf["org"][:n], buf["dirs"][:n], hits, self.absorb_from,
                                                seed=seed + 1000 * (b + 1), offset=self.offset,
                                                out_org=nxt["org"], out_dir=nxt["dirs"], out_src=nxt["src"],
                                                stream=stream)
            src = nxt["src"][:m]
            n = m
        if timed:
            torch.cuda.synchronize()
            for p in out:
                p.ms = p.ms[0].elapsed_time(p.ms[1])
            self.emit_ms = emit_ev[0].elapsed_time(emit_ev[1])
        return out

    def tune_ray_orders(self, photons: int, bounces: int, seed: int, reps: int = 2) -> list[int]:
        """time every pass in both ray orders (the median of `reps` runs each,
        after one untimed run) and keep the faster order per pass; the hits
        are the same either way"""
        best = []
        times = {}
        for order in (_pm.RAY_ORDER_COHERENT, _pm.RAY_ORDER_INPUT):
            self.trace(photons, bounces, seed, orders=[order])
            runs = [self.trace(photons, bounces, seed, orders=[order], timed=True) for _ in range(max(1, reps))]
            times[order] = [sorted(r[b].ms for r in runs)[len(runs) // 2] if b < len(runs[0]) else 0.0
                            for b in range(bounces + 1)]
        for b in range(bounces + 1):
            c, i = times[_pm.RAY_ORDER_COHERENT][b], times[_pm.RAY_ORDER_INPUT][b]
            best.append(_pm.RAY_ORDER_COHERENT if c <= i else _pm.RAY_ORDER_INPUT)
        self.orders = best
        self.tuned_ms = times
        return best
