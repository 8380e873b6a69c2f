"""The full-size traversal fixture (tests/golden/c2_full_digests.json) is the
oracle's answer for all 4,194,304 microbenchmark rays, kept as per-block
digests.  Here the generators are checked to give the fixture's inputs, and
two blocks (one of coherent, one of incoherent rays) are traced again by the
oracle and must give the committed digests."""
import json

import numpy as np

import oracle
from tests.golden.make_fixtures import c2_full_digests, c2_full_inputs


def test_full_fixture_inputs_and_sample_blocks():
    with open("tests/golden/c2_full_digests.json") as f:
        fx = json.load(f)
    scene, o, d, ins = c2_full_inputs()
    assert ins == fx["inputs"]
    assert len(fx["blocks"]) == fx["num_rays"] // fx["block"]
    orc = oracle.OracleScene(scene)
    B = fx["block"]
    for blk in (3, len(fx["blocks"]) - 5):
        sl = slice(blk * B, (blk + 1) * B)
        r = orc.trace(o[sl], d[sl])
        assert c2_full_digests(r["prim"], r["t"], r["u"], r["v"], B) == [fx["blocks"][blk]], blk


def test_procedural_generators_reproduce_their_digests():
    """Every procedural workload (the benchmark workloads at full size: the cornell room,
    the 1080p terrain with 200 spheres, textures and HDR sky, the 8-scene
    serving pool and its 64 requests; and the scenes the tests render) gives
    the committed SHA-256 of everything it hands to the library."""
    from tests.golden.make_fixtures import generator_cases, scene_digest
    from voidray_amd import workloads as W
    with open("tests/golden/generator_digests.json") as f:
        fx = json.load(f)
    for name, make in generator_cases().items():
        assert scene_digest(make()) == fx[name], name
    assert [scene_digest(s) for s in W.serving_scene_pool()] == fx["serving_pool"]
    assert [list(r) for r in W.serving_requests()] == fx["serving_requests"]
