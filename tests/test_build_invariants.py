"""The wide BVH's stored encoding: every child's 8-bit bounds, decoded as the
kernels decode them (origin + q * 2^(e - 127)), contain the exact box of the
subtree or leaf they stand for (the build counts (child, axis) pairs and the
ones that do not contain; vr_scene_stats.qbox_*).  Checked on the microbench
soup, the terrain with textures, the edge-case fan, analytic surfaces beside
meshes, and instances (the object-space BVHs are checked too)."""
import numpy as np
import pytest

from tests.trace_cases import edge_case_scene, instanced_scene
from voidray_amd import workloads as W
from voidray_amd.device import trace_host


def _stats(scene):
    o = np.zeros((1, 3), np.float32)
    d = np.float32([[0.0, 0.0, 1.0]])
    return trace_host(scene, o, d)["stats"]


@pytest.mark.parametrize("make", [
    lambda: W.random_triangle_scene(50000, seed=9),
    lambda: W.terrain_spheres(grid=64, n_spheres=12, sphere_res=(16, 8), tex_size=16, env_size=(32, 16)),
    lambda: edge_case_scene()[0],
    lambda: W.analytic_and_meshes(),
    lambda: instanced_scene()[0],
], ids=["soup", "terrain", "edge_cases", "analytic_and_meshes", "instances"])
def test_quantised_bounds_contain_the_exact_boxes(make):
    st = _stats(make())
    assert st["qbox_checked"] >= 3 * st["num_qnodes"]  # every node has at least one child
    assert st["qbox_not_containing"] == 0, st
