"""Batched serving: many requests (scenes, sizes, spp, seeds, estimators) in one
captured launch.  Bar: every request's image is bit-identical to its
standalone render (the render of that request alone, which the render tests
hold against the oracle); a relaunch of the captured graph adds the same frame
again; time stamps and statistics describe the batch.
"""
import numpy as np
import pytest

from tests.conftest import announce
from voidray_amd import _native as N
from voidray_amd import workloads as W
from voidray_amd.device import BatchRenderer, make_render_desc
from voidray_amd.scene import RenderMode, RenderSettings

pytestmark = pytest.mark.gpu

EXT = N.VR_RENDER_ENV_IS | N.VR_RENDER_RUSSIAN_ROULETTE | N.VR_RENDER_ALPHA_PER_CALL
REF = N.VR_RENDER_ROWS_BY_HEIGHT | N.VR_RENDER_ALPHA_PER_CALL


def _small_terrain():
    return W.terrain_spheres(grid=24, n_spheres=6, sphere_res=(16, 8), tex_size=64, env_size=(64, 32),
                             sun_peak=200.0)


@pytest.fixture(scope="module")
def scenes(gpu_ctx):
    built = [W.cornell_room(), W.material_spheres(), _small_terrain()]
    return [s.build_acceleration(gpu_ctx) for s in built]


def _requests():
    """(scene index, desc) pairs of uneven size, spp, seed and estimator."""
    inf = float("inf")
    S = RenderSettings
    out = [
        (0, make_render_desc(40, 30, 3, 3, S(total_samples=3, max_bounces=8), seed=1, request=0, flags=REF)),
        (2, make_render_desc(48, 27, 5, 5, S(total_samples=5, max_bounces=16, firefly_clamp=inf), seed=2,
                             request=1, flags=EXT, rr_depth=4)),
        (1, make_render_desc(32, 32, 2, 8, S(total_samples=8, max_bounces=10), seed=3, request=2,
                             flags=N.VR_RENDER_ALPHA_PER_CALL)),
        (0, make_render_desc(17, 23, 1, 1, S(total_samples=1, max_bounces=8), seed=4, request=3, flags=REF)),
        (2, make_render_desc(20, 20, 7, 7, S(total_samples=7, max_bounces=12, firefly_clamp=inf), seed=5,
                             request=4, flags=EXT, rr_depth=2)),
        (1, make_render_desc(24, 16, 4, 4, S(total_samples=4, render_mode=RenderMode.Normal), seed=6, request=5,
                             flags=N.VR_RENDER_ALPHA_PER_CALL)),
        (0, make_render_desc(8, 8, 2, 2, S(total_samples=2, max_bounces=0), seed=7, request=6, flags=REF)),
        (1, make_render_desc(16, 16, 0, 4, S(total_samples=4, max_bounces=6), seed=8, request=7,
                             flags=N.VR_RENDER_ALPHA_PER_CALL)),
        (1, make_render_desc(30, 20, 2, 8, S(total_samples=8, max_bounces=6), sample_offset=3, seed=9, request=8,
                             flags=N.VR_RENDER_ALPHA_PER_CALL)),
        (2, make_render_desc(33, 19, 3, 3, S(total_samples=3, max_bounces=6, firefly_clamp=inf), seed=10,
                             request=9, flags=EXT, rr_depth=3, engine=N.VR_ENGINE_REFERENCE)),
    ]
    return out


def _standalone(scenes, reqs):
    imgs = []
    for si, d in reqs:
        img = np.zeros((d.height, d.width, 4), np.float32)
        scenes[si].render(img, d)
        imgs.append(img)
    return imgs


def _targets(reqs):
    import torch
    t = [torch.zeros((d.height, d.width, 4), dtype=torch.float32, device="cuda") for _, d in reqs]
    torch.cuda.synchronize()
    return t


def _batch(scenes, reqs, targets):
    return [(scenes[si], d, t.data_ptr()) for (si, d), t in zip(reqs, targets)]


@pytest.mark.parametrize("graph,chunk", [(True, 1024), (False, 0), (True, 1 << 22)])
def test_batch_equals_standalone(gpu_ctx, scenes, graph, chunk):
    reqs = _requests()
    announce(f"batch of {len(reqs)} requests graph={graph} chunk_paths={chunk}")
    want = _standalone(scenes, reqs)
    targets = _targets(reqs)
    br = BatchRenderer(gpu_ctx)
    br.submit(_batch(scenes, reqs, targets), tuning=(chunk, 128), graph=graph, timestamps=True,
              shortest_first=True)
    st = br.stats()
    lat = br.latencies(len(reqs))
    for k, ((si, d), t) in enumerate(zip(reqs, targets)):
        got = t.cpu().numpy()
        assert np.array_equal(got, want[k]), f"request {k}: max diff {np.max(np.abs(got - want[k]))}"
    assert st["samples"] == sum(d.width * d.height * d.samples for _, d in reqs)
    assert np.all(np.isfinite(lat)) and np.all(lat >= 0.0), lat
    print(f"[batch] graph={graph} chunk={chunk}: segments {st['segments']}, latencies ms "
          f"{np.round(lat, 3).tolist()}", flush=True)


def test_graph_relaunch_adds_the_same_frame(gpu_ctx, scenes):
    reqs = _requests()[:5]
    announce("batch relaunch x2 (additive) and overwrite")
    targets = _targets(reqs)
    br = BatchRenderer(gpu_ctx)
    br.submit(_batch(scenes, reqs, targets), tuning=(2048, 256))
    br.stats()
    once = [t.cpu().numpy() for t in targets]
    br.relaunch()
    br.stats()
    for k, t in enumerate(targets):
        assert np.array_equal(t.cpu().numpy(), once[k] + once[k]), k
    # overwrite: a relaunch stores the same frame again
    reqs_o = []
    for si, d in reqs:
        d2 = N.vr_render_desc.from_buffer_copy(d)
        d2.flags |= N.VR_RENDER_OVERWRITE
        reqs_o.append((si, d2))
    targets = _targets(reqs_o)
    br.submit(_batch(scenes, reqs_o, targets), tuning=(2048, 256))
    br.stats()
    first = [t.cpu().numpy() for t in targets]
    br.relaunch()
    br.relaunch()
    br.stats()
    for k, t in enumerate(targets):
        got = t.cpu().numpy()
        assert np.array_equal(got, first[k]), k
        assert np.array_equal(got[..., :3], once[k][..., :3]), k


def test_shortest_first_orders_completion(gpu_ctx, scenes):
    reqs = _requests()
    targets = _targets(reqs)
    br = BatchRenderer(gpu_ctx)
    announce("batch time stamps, shortest group first")
    br.submit(_batch(scenes, reqs, targets), tuning=(1 << 22, 256), timestamps=True, shortest_first=True)
    lat = br.latencies(len(reqs))

    def size_class(p):  # paths in [4^c, 4^(c+1))
        c = 0
        while p >> (2 * (c + 1)):
            c += 1
        return c

    # groups: (scene, engine, size class), run in ascending order of their paths
    key = [(si, d.engine, size_class(d.width * d.height * d.samples)) for si, d in reqs]
    paths = {}
    for k, (si, d) in enumerate(reqs):
        paths[key[k]] = paths.get(key[k], 0) + d.width * d.height * d.samples
    order = sorted(paths, key=lambda g: paths[g])
    done = {g: max(lat[k] for k in range(len(reqs)) if key[k] == g) for g in paths}
    start = {g: min(lat[k] for k in range(len(reqs)) if key[k] == g) for g in paths}
    for a, b in zip(order, order[1:]):
        if paths[a] < paths[b]:
            assert done[a] <= start[b], (a, b, done, start)


def test_autotune_picks_a_candidate(gpu_ctx, scenes):
    reqs = _requests()[:4]
    targets = _targets(reqs)
    br = BatchRenderer(gpu_ctx)
    announce("batch autotune")
    (chunk, block), ms = br.autotune(_batch(scenes, reqs, targets), reps=1)
    assert chunk in (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27) and block in (128, 256, 512, 1024)
    assert ms > 0.0
    # the tuned configuration still renders every request exactly
    want = _standalone(scenes, reqs)
    targets = _targets(reqs)
    br.submit(_batch(scenes, reqs, targets), tuning=(chunk, block))
    br.stats()
    for k, t in enumerate(targets):
        assert np.array_equal(t.cpu().numpy(), want[k]), k


def test_batch_rejects_bad_requests(gpu_ctx, scenes):
    import torch
    br = BatchRenderer(gpu_ctx)
    t = torch.zeros((4, 4, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    d = make_render_desc(4, 4, 64, 64, RenderSettings(total_samples=64))
    with pytest.raises(N.VoidrayError, match="chunk"):
        br.submit([(scenes[0], d, t.data_ptr())], tuning=(32, 256))
    with pytest.raises(N.VoidrayError, match="block"):
        br.submit([(scenes[0], d, t.data_ptr())], tuning=(1 << 20, 100))
    with pytest.raises(N.VoidrayError, match="relaunch"):
        br.relaunch()
