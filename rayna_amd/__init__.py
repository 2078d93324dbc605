"""rayna_amd: rayna_engine's render hot path (per-pixel path tracing) on AMD MI355X (gfx950).

Host API mirroring rayna_engine (Scene / objects / meshes / materials / Camera / RenderOpts / Renderer) over the
C ABI of librayna_amd.so (include/rayna_gpu.h), whose HIP kernels do all per-ray work on the GPU.
"""
from .scene import (AxisBoxMesh, BvhMesh, Camera, DielectricMaterial, IsotropicMaterial, LambertianMaterial,
                    LightMaterial, MetalMaterial, NoSkybox, ParallelogramMesh, RaymarchedIsosurfaceMesh, RenderMode,
                    RenderOpts, Scene, SdfMandelbulb, SdfSminSpheres, SdfSphere, SimpleObject, SimpleSkybox, SphereMesh,
                    Transform3, TriangleListMesh, TriangleMesh, VolumetricObject, WhiteSkybox)
from . import presets

__all__ = ["AxisBoxMesh", "BvhMesh", "Camera", "DielectricMaterial", "IsotropicMaterial", "LambertianMaterial", "LightMaterial",
           "MetalMaterial", "NoSkybox", "ParallelogramMesh", "RaymarchedIsosurfaceMesh", "RenderMode", "RenderOpts", "Scene",
           "SdfMandelbulb", "SdfSminSpheres", "SdfSphere", "SimpleObject", "SimpleSkybox", "SphereMesh", "Transform3",
           "TriangleListMesh", "TriangleMesh", "VolumetricObject", "WhiteSkybox", "presets"]
from .renderer import Render, Renderer, RenderStats, RayError, prepare_scene

__all__ += ["Render", "Renderer", "RenderStats", "RayError", "prepare_scene"]
