"""voidray_amd — MI355X-native path-tracing core behind a C ABI.

The host mirror of voidray_renderer's scene/material/camera API lives in
``voidray_amd.scene``; uploaded scenes, ray queries and rendering in
``voidray_amd.device``; synthetic workloads in ``voidray_amd.workloads``.
"""
from ._native import VoidrayError, lib  # noqa: F401
from .scene import (Camera, ColorManagementSettings, Environments, Materials, Mesh, RenderMode,  # noqa: F401
                    RenderSettings, SampleType, Scene, Settings, Surfaces, Tonemap, degrees_to_radians)
from .device import Context, SceneAcceleration, iterative_render, make_render_desc, tonemap  # noqa: F401
