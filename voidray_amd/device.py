"""Device-side objects: context, uploaded scene, ray queries, rendering."""
import ctypes as C

import numpy as np

from . import _native as N


class Context:
    _default = None

    def __init__(self, device=0):
        h = C.c_void_p()
        N.check(N.lib().vr_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    @property
    def stream(self):
        return N.lib().vr_stream(self.handle)

    def close(self):
        if self.handle:
            N.lib().vr_destroy(self.handle)
            self.handle = None


class SceneAcceleration:
    """An uploaded scene (Scene::build_acceleration, core/scene.rs:163-179)."""

    def __init__(self, ctx, scene, build_flags=0):
        self.ctx = ctx
        self.scene = scene
        desc, keep = scene.to_desc(build_flags)
        h = C.c_void_p()
        N.check(N.lib().vr_scene_upload(ctx.handle, C.byref(desc), C.byref(h)))
        self.handle = h
        self.num_prims = scene.num_triangles

    def stats(self):
        s = N.vr_scene_stats()
        N.check(N.lib().vr_scene_get_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in N.vr_scene_stats._fields_}

    def release(self):
        if self.handle:
            N.lib().vr_scene_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    # -- ray queries: SceneAcceleration::hit (core/scene.rs:182-185) ---------
    def hit(self, origins, directions, t_min=1e-5, any_hit=False, engine=N.VR_ENGINE_WIDE):
        """Host arrays in, host arrays out.  Returns dict prim/t/u/v (closest) or hit (any)."""
        o = np.ascontiguousarray(origins, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, dtype=np.float32).reshape(-1, 3)
        n = len(o)
        desc = N.vr_trace_desc(n, t_min, N.VR_TRACE_ANY if any_hit else N.VR_TRACE_CLOSEST, engine)
        out = {}
        if any_hit:
            hit = np.zeros(n, np.uint8)
            N.check(N.lib().vr_trace(self.handle, C.byref(desc), o.ctypes.data, d.ctypes.data, None, None, None,
                                     None, hit.ctypes.data))
            out["hit"] = hit
        else:
            prim = np.zeros(n, np.uint32)
            t = np.zeros(n, np.float32)
            u = np.zeros(n, np.float32)
            v = np.zeros(n, np.float32)
            N.check(N.lib().vr_trace(self.handle, C.byref(desc), o.ctypes.data, d.ctypes.data, prim.ctypes.data,
                                     t.ctypes.data, u.ctypes.data, v.ctypes.data, None))
            out.update(prim=prim, t=t, u=u, v=v)
        return out

    def last_fallbacks(self):
        c = C.c_uint32()
        N.check(N.lib().vr_trace_last_fallbacks(self.handle, C.byref(c)))
        return c.value

    def last_wide_ms(self):
        """Device time of the wide-engine kernel of the last trace (HIP events;
        that trace must have been made with timing=True)."""
        ms = C.c_float()
        N.check(N.lib().vr_trace_last_timing(self.handle, C.byref(ms)))
        return ms.value

    def trace_device(self, d_origins, d_directions, n, any_hit=False, t_min=1e-5, engine=N.VR_ENGINE_WIDE,
                     d_prim=None, d_t=None, d_u=None, d_v=None, d_hit=None, stream=None, timing=False):
        """Device pointers (ints) in/out, enqueued on `stream` (None = context stream).
        timing: record the wide-engine kernel's device time for last_wide_ms()."""
        mode = (N.VR_TRACE_ANY if any_hit else N.VR_TRACE_CLOSEST) | (N.VR_TRACE_TIMING if timing else 0)
        desc = N.vr_trace_desc(n, t_min, mode, engine)
        N.check(N.lib().vr_trace_device_async(self.handle, C.byref(desc), d_origins, d_directions, d_prim, d_t,
                                              d_u, d_v, d_hit, stream))

    # -- rendering ---------------------------------------------------------------
    def render(self, rgba, desc):
        """rgba: (H, W, 4) float32 host array, accumulated in place."""
        assert rgba.dtype == np.float32 and rgba.flags.c_contiguous
        st = N.vr_render_stats()
        N.check(N.lib().vr_render(self.handle, C.byref(desc), rgba.ctypes.data, C.byref(st)))
        return {k: getattr(st, k) for k, _ in N.vr_render_stats._fields_}

    def render_device(self, d_rgba, desc, stream=None):
        st = N.vr_render_stats()
        N.check(N.lib().vr_render_device_async(self.handle, C.byref(desc), d_rgba, C.byref(st), stream))
        return st


class BatchRenderer:
    """Serves many render requests in one launch (vr_render_batch_async).

    The reference renders one frame per RenderThread::one_shot call
    (render/renderer.rs:67-107).  Here a request is (SceneAcceleration,
    vr_render_desc, device target pointer); requests on one scene share
    wavefront chunks, and with `graph=True` the whole batch is captured into a
    hipGraph that `relaunch` replays.  All calls go to one stream."""

    def __init__(self, ctx):
        self.ctx = ctx

    @staticmethod
    def _array(requests):
        arr = (N.vr_request * max(1, len(requests)))()
        for k, (accel, desc, d_rgba) in enumerate(requests):
            arr[k].scene = accel.handle.value
            arr[k].desc = desc
            arr[k].d_rgba = int(d_rgba)
        return arr

    def submit(self, requests, tuning=None, graph=True, timestamps=False, shortest_first=False, stream=None):
        flags = ((N.VR_BATCH_GRAPH if graph else 0) | (N.VR_BATCH_TIMESTAMPS if timestamps else 0)
                 | (N.VR_BATCH_SHORTEST_FIRST if shortest_first else 0))
        t = N.vr_tuning(*tuning) if tuning is not None else None
        N.check(N.lib().vr_render_batch_async(self.ctx.handle, self._array(requests), len(requests),
                                              C.byref(t) if t is not None else None, flags, stream))

    def relaunch(self, stream=None):
        N.check(N.lib().vr_batch_relaunch_async(self.ctx.handle, stream))

    def latencies(self, n):
        """Completion time of each request (ms after the batch started on the device)."""
        out = np.zeros(n, np.float64)
        N.check(N.lib().vr_batch_latencies(self.ctx.handle, out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out

    def stats(self):
        st = N.vr_render_stats()
        N.check(N.lib().vr_batch_last_stats(self.ctx.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in N.vr_render_stats._fields_}

    def autotune(self, requests, reps=2, shortest_first=False, stream=None):
        """Returns ((chunk_paths, block), ms) of the fastest candidate; the targets are overwritten."""
        best = N.vr_tuning()
        ms = C.c_float()
        flags = N.VR_BATCH_SHORTEST_FIRST if shortest_first else 0
        N.check(N.lib().vr_batch_autotune(self.ctx.handle, self._array(requests), len(requests), flags, reps,
                                          stream, C.byref(best), C.byref(ms)))
        return (best.chunk_paths, best.block), ms.value


class RadianceCache:
    """Learned radiance cache (vr_cache_*): multiresolution hash grid + MLP,
    BF16 on MFMA (precision=VR_CACHE_FP32: the all-FP32 debug path), trained
    online on path suffixes the tracer produces.  No counterpart in the
    reference, whose integrator traces every path to the end
    (core/tracer.rs:9-56)."""

    def __init__(self, ctx, bbox_lo, bbox_hi, log2_table=19, base_resolution=16, max_resolution=2048,
                 precision=N.VR_CACHE_BF16, lr_grid=1e-2, lr_mlp=1e-3, beta1=0.9, beta2=0.99, eps=1e-15, seed=0):
        d = N.vr_cache_desc()
        d.log2_table, d.base_resolution, d.max_resolution, d.precision = log2_table, base_resolution, \
            max_resolution, precision
        for a in range(3):
            d.bbox_lo[a], d.bbox_hi[a] = float(bbox_lo[a]), float(bbox_hi[a])
        d.lr_grid, d.lr_mlp, d.beta1, d.beta2, d.eps, d.seed = lr_grid, lr_mlp, beta1, beta2, eps, seed
        h = C.c_void_p()
        N.check(N.lib().vr_cache_create(ctx.handle, C.byref(d), C.byref(h)))
        self.ctx = ctx
        self.handle = h

    def info(self):
        i = N.vr_cache_info()
        N.check(N.lib().vr_cache_get_info(self.handle, C.byref(i)))
        return {"levels": i.levels, "table_size": i.table_size, "log2_table": int(i.table_size).bit_length() - 1,
                "precision": i.precision, "resolution": list(i.resolution), "origin": list(i.origin),
                "inv_extent": i.inv_extent, "grid_params": i.grid_params, "steps": i.steps}

    def get_params(self):
        info = self.info()
        mlp = np.zeros(N.VR_CACHE_MLP_PARAMS, np.float32)
        grid = np.zeros(info["grid_params"], np.float32)
        N.check(N.lib().vr_cache_get_params(self.handle, mlp.ctypes.data, grid.ctypes.data))
        return mlp, grid

    def set_params(self, mlp=None, grid=None):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        N.check(N.lib().vr_cache_set_params(self.handle, None if m is None else m.ctypes.data,
                                            None if g is None else g.ctypes.data))

    def infer(self, n, d_pos, d_dir, d_nrm, d_rough, d_out, stream=None):
        N.check(N.lib().vr_cache_infer_async(self.handle, n, d_pos, d_dir, d_nrm, d_rough, d_out, stream))

    def gradients(self, n, d_pos, d_dir, d_nrm, d_rough, d_target, d_mlp_grad=None, d_grid_grad=None,
                  d_loss=None, stream=None):
        N.check(N.lib().vr_cache_gradients_async(self.handle, n, d_pos, d_dir, d_nrm, d_rough, d_target,
                                                 d_mlp_grad, d_grid_grad, d_loss, stream))

    def train(self, n, d_pos, d_dir, d_nrm, d_rough, d_target, stream=None):
        N.check(N.lib().vr_cache_train_async(self.handle, n, d_pos, d_dir, d_nrm, d_rough, d_target, stream))

    def render_frame(self, accel, d_rgba, desc, train_paths=1 << 14, train_vertices=4, cache_depth=2,
                     train_steps=1, frame=0, stream=None, query_roughness=0.0):
        f = N.vr_cache_frame(train_paths, train_vertices, cache_depth, train_steps, frame, query_roughness)
        N.check(N.lib().vr_render_cached_async(accel.handle, self.handle, C.byref(desc), C.byref(f), d_rgba,
                                               stream))

    def stats(self):
        s = N.vr_cache_stats()
        N.check(N.lib().vr_cache_last_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in N.vr_cache_stats._fields_}

    def close(self):
        if self.handle:
            N.lib().vr_cache_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scene_bounds(scene):
    """Axis-aligned bounds of every mesh vertex of a Scene (for the cache's position normalisation);
    instances count with their placed vertices."""
    if getattr(scene, "instances", None):
        scene = scene.flattened()
    lo = np.full(3, np.inf, np.float64)
    hi = np.full(3, -np.inf, np.float64)
    for m in scene.meshes:
        if len(m.positions):
            lo = np.minimum(lo, m.positions.min(0))
            hi = np.maximum(hi, m.positions.max(0))
    return lo.astype(np.float32), hi.astype(np.float32)


def make_render_desc(width, height, samples, total_samples, settings=None, sample_offset=0, seed=0, request=0,
                     flags=N.VR_RENDER_ALPHA_PER_CALL, rr_depth=0, engine=N.VR_ENGINE_WIDE):
    from .scene import RenderSettings
    s = settings or RenderSettings()
    return N.vr_render_desc(width, height, samples, total_samples, sample_offset, s.max_bounces,
                            float(s.firefly_clamp), s.render_mode, flags, rr_depth, seed, request, engine)


def iterative_render(target, scene_accel, settings, samples, sample_offset=0, seed=0, request=0,
                     flags=N.VR_RENDER_ROWS_BY_HEIGHT | N.VR_RENDER_ALPHA_PER_CALL):
    """render/iterative.rs:11-55: adds (sum of `samples` traces) / total_samples into target
    ((H, W, 4) float32).  `sample_offset` selects the sample indices of the per-pixel stream."""
    h, w = target.shape[:2]
    desc = make_render_desc(w, h, samples, settings.total_samples, settings, sample_offset, seed, request, flags)
    return scene_accel.render(target, desc)


def tonemap(ctx, src, post, dst=None):
    """post_process.glsl over an (H, W, 4) float32 image."""
    h, w = src.shape[:2]
    src = np.ascontiguousarray(src, dtype=np.float32)
    dst = np.zeros_like(src) if dst is None else dst
    N.check(N.lib().vr_tonemap(ctx.handle, C.byref(post), src.ctypes.data, dst.ctypes.data))
    return dst


def trace_host(scene, origins, directions, t_min=1e-5, any_hit=False, engine=N.VR_ENGINE_WIDE, build_flags=0):
    """SceneAcceleration::hit answered on the host (no GPU): the scene is built
    as for the device and the rays run through the kernels' traversal routines
    on the CPU.  Returns prim/t/u/v (closest) or hit (any), plus the build
    statistics and the number of rays the wide engine handed to the exact one."""
    o = np.ascontiguousarray(origins, dtype=np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(directions, dtype=np.float32).reshape(-1, 3)
    n = len(o)
    desc, keep = scene.to_desc(build_flags)
    tdesc = N.vr_trace_desc(n, t_min, N.VR_TRACE_ANY if any_hit else N.VR_TRACE_CLOSEST, engine)
    prim = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    hit = np.zeros(n, np.uint8)
    stats = N.vr_scene_stats()
    fb = C.c_uint32(0)
    N.check(N.lib().vr_trace_host(C.byref(desc), C.byref(tdesc), o.ctypes.data, d.ctypes.data, prim.ctypes.data,
                                  t.ctypes.data, u.ctypes.data, v.ctypes.data, hit.ctypes.data, C.byref(stats),
                                  C.byref(fb)))
    out = {"hit": hit} if any_hit else {"prim": prim, "t": t, "u": u, "v": v}
    out["stats"] = {k: getattr(stats, k) for k, _ in N.vr_scene_stats._fields_}
    out["fallbacks"] = fb.value
    return out
