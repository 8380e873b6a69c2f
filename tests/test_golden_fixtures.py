"""The mid-size traversal fixture (65,536 triangles, 262,144 rays: the
oracle's closest hits as per-block digests, tests/golden/c2_mid_digests.json)
and the tiny render fixture (the oracle's cornell room at 32x32, 4 spp,
tests/golden/c1_tiny.npz), both written by tests/golden/make_fixtures.py.
On the CPU: the generators and the oracle still give them, and the kernels'
traversal routines run on the host (vr_trace_host, wide engine) give the
same digests on two blocks.  The GPU side is in test_trace_gpu.py and
test_render_gpu.py."""
import json

import numpy as np

import oracle
from tests.golden.make_fixtures import (c1_tiny_desc, c2_full_digests, c2_mid_inputs, scene_digest)
from voidray_amd import _native as N
from voidray_amd import workloads as W
from voidray_amd.device import trace_host


def _mid():
    with open("tests/golden/c2_mid_digests.json") as f:
        return json.load(f)


def test_mid_fixture_oracle_reproduces_every_block():
    fx = _mid()
    scene, o, d = c2_mid_inputs()
    assert scene_digest(scene) == fx["scene"]
    r = oracle.OracleScene(scene).trace(o, d)
    assert c2_full_digests(r["prim"], r["t"], r["u"], r["v"], fx["block"]) == fx["blocks"]
    assert int(np.sum(r["prim"] != 0xFFFFFFFF)) == fx["hits"]


def test_mid_fixture_host_traversal_two_blocks():
    fx = _mid()
    scene, o, d = c2_mid_inputs()
    B = fx["block"]
    for blk in (1, len(fx["blocks"]) - 2):  # one coherent, one incoherent block
        sl = slice(blk * B, (blk + 1) * B)
        r = trace_host(scene, o[sl], d[sl], engine=N.VR_ENGINE_WIDE)
        assert c2_full_digests(r["prim"], r["t"], r["u"], r["v"], B) == [fx["blocks"][blk]], blk


def test_tiny_render_fixture_oracle_reproduces_it():
    fx = np.load("tests/golden/c1_tiny.npz")
    img, _, _ = oracle.OracleScene(W.cornell_room()).render(c1_tiny_desc())
    assert np.array_equal(img.astype(np.float32), fx["rgba"])
    assert fx["rgba"][..., :3].max() > 0  # the light is reached
