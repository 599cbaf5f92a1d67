"""The allocation-free trace (pm_trace_noalloc) against pm_trace.

pm_trace_noalloc keeps its scratch (work counter, coherent-order sort
buffers) in a file-scope device array instead of context buffers.  Bar:
every hit record and counter equals pm_trace's bit for bit, on two calls in a
row (the second with fewer rays, the array keeping the first call's
contents), in both ray orders and for all three tracers; sizes beyond the
array, and the Woop intersection (its matrices need memory), are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available()
    return T


def noalloc(pm, ctx, algo, O, D, out, params, counters=None):
    L = pm.load_library()
    rc = L.pm_trace_noalloc(ctx.h, algo, O.data_ptr(), D.data_ptr(), O.shape[0], out.data_ptr(),
                            C.byref(params), counters.data_ptr() if counters is not None else None,
                            pm._stream_ptr(None))
    return rc


def test_noalloc_trace_equals_trace(pm, scenes, torch):
    dev = "cuda:0"
    tris = scenes.cluster_soup(200_000, seed=3)
    P = [torch.from_numpy(t).to(dev) for t in tris]
    ctx = pm.Context(0)
    ctx.build(*P, pm.build_params("reference"))
    L = [p[-2:].contiguous() for p in P]
    O, D = ctx.photon_emit(L, 400_000, seed=3)
    O, D = O.contiguous(), D.contiguous()
    coherent = pm.trace_params(stack_size=8, ray_order=pm.RAY_ORDER_COHERENT)
    plain = pm.trace_params(stack_size=8)
    cases = [(pm.SHORT_STACK, 400_000, coherent), (pm.SHORT_STACK, 150_000, coherent),
             (pm.KD_RESTART, 300_000, coherent), (pm.KD_RESTART, 200_000, plain),
             (pm.SHORT_STACK, 1000, plain), (pm.BRUTE, 512, plain)]
    for algo, n, prm in cases:
        o, d = O[:n].contiguous(), D[:n].contiguous()
        want = ctx.trace(algo, o, d, prm)
        cw = torch.zeros((n, 4), dtype=torch.int32, device=dev)
        ctx.trace(algo, o, d, prm, counters=cw)
        got = torch.full((n, 4), -7.0, dtype=torch.float32, device=dev)
        cg = torch.full((n, 4), -7, dtype=torch.int32, device=dev)
        assert noalloc(pm, ctx, algo, o, d, got, prm) == 0, ctx._L.pm_last_error(ctx.h)
        assert noalloc(pm, ctx, algo, o, d, got, prm, cg) == 0, ctx._L.pm_last_error(ctx.h)
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), (algo, n, prm.ray_order)
        assert torch.equal(cg, cw), (algo, n)
        assert (pm.hits_view(got)[0] >= 0).float().mean() > 0.3
    # refused: more rays than the array orders coherently, and Woop
    big = 1 << 20
    Ob = torch.zeros((big + 1, 4), dtype=torch.float32, device=dev)
    Hb = torch.empty_like(Ob)
    assert noalloc(pm, ctx, pm.SHORT_STACK, Ob, Ob, Hb, coherent) == 1
    assert b"orders at most" in ctx._L.pm_last_error(ctx.h)
    woop = pm.trace_params(stack_size=8, intersection=pm.WOOP)
    assert noalloc(pm, ctx, pm.SHORT_STACK, O[:64], D[:64], Hb[:64], woop) == 5
    ctx.close()
