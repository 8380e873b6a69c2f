"""vr_trace_device_noalloc_async: the trace entry point that allocates nothing
(its scratch is a static device array per GPU) gives, through the C ABI, the
same closest hits and any-hit bits as vr_trace_device_async, bit for bit, on
two calls in a row, the second with fewer rays; more rays than
VR_NOALLOC_MAX_RAYS are refused before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import announce
from voidray_amd import _native as N
from voidray_amd import workloads as W

pytestmark = pytest.mark.gpu
VR_ERR_INVALID = -1  # include/voidray_hip.h


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _trace(fn, acc, o, d, n, any_hit, engine):
    import torch
    prim = torch.full((n,), 7, dtype=torch.int32, device="cuda:0")
    t = torch.full((n,), -1.0, dtype=torch.float32, device="cuda:0")
    u = torch.full((n,), -1.0, dtype=torch.float32, device="cuda:0")
    v = torch.full((n,), -1.0, dtype=torch.float32, device="cuda:0")
    hit = torch.full((n,), 9, dtype=torch.uint8, device="cuda:0")
    desc = N.vr_trace_desc(n, 1e-5, N.VR_TRACE_ANY if any_hit else N.VR_TRACE_CLOSEST, engine)
    torch.cuda.synchronize()
    N.check(fn(acc.handle, C.byref(desc), o.data_ptr(), d.data_ptr(), prim.data_ptr(), t.data_ptr(),
               u.data_ptr(), v.data_ptr(), hit.data_ptr(), None))
    torch.cuda.synchronize()
    return {k: x.cpu().numpy() for k, x in dict(prim=prim, t=t, u=u, v=v, hit=hit).items()}


def _same(a, b, label, any_hit):
    # any hit defines only the hit bit (which occluder a query reports first
    # depends on the traversal's stack budget); closest hit defines everything
    for k in (("hit",) if any_hit else ("prim", "t", "u", "v", "hit")):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), f"{label}: {k} differs"


@pytest.mark.parametrize("case", ["soup", "cornell"])
def test_noalloc_trace_matches_the_allocating_entry_point(case, gpu_ctx):
    lib = N.lib()
    if case == "soup":
        scene = W.random_triangle_scene(50000, seed=9)
        o, d = W.microbench_rays(1 << 18, res=512, seed=4)
    else:
        scene = W.cornell_room()
        rng = np.random.default_rng(6)
        o = rng.uniform(0.05, 0.95, (1 << 17, 3)).astype(np.float32)
        d = W._normalize_rows(rng.normal(size=(1 << 17, 3)).astype(np.float32))
    acc = scene.build_acceleration(gpu_ctx)
    do, dd = _dev(o), _dev(d)
    for n in (len(o), len(o) // 8 + 13):  # a second, smaller call after the first
        for any_hit in (False, True):
            for engine in (N.VR_ENGINE_WIDE, N.VR_ENGINE_REFERENCE):
                if engine == N.VR_ENGINE_REFERENCE and n > 40000:
                    continue  # the exact engine is slow on big batches; the small call covers it
                announce(f"noalloc {case} n={n} any={any_hit} engine={engine}")
                ref = _trace(lib.vr_trace_device_async, acc, do, dd, n, any_hit, engine)
                got = _trace(lib.vr_trace_device_noalloc_async, acc, do, dd, n, any_hit, engine)
                _same(ref, got, f"{case} n={n} any={any_hit} engine={engine}", any_hit)
                if not any_hit:
                    assert np.any(got["prim"] != -1)


def test_noalloc_trace_refuses_oversized_batches(gpu_ctx):
    scene = W.random_triangle_scene(64, seed=3)
    acc = scene.build_acceleration(gpu_ctx)
    desc = N.vr_trace_desc(N.VR_NOALLOC_MAX_RAYS + 1, 1e-5, N.VR_TRACE_CLOSEST, N.VR_ENGINE_WIDE)
    # no buffers are touched: the size check comes first
    rc = N.lib().vr_trace_device_noalloc_async(acc.handle, C.byref(desc), 16, 16, 16, 16, 16, 16, 16, None)
    assert rc == VR_ERR_INVALID
    assert b"VR_NOALLOC_MAX_RAYS" in N.lib().vr_last_error()
