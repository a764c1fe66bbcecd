"""light_garden_amd — the light_garden trace + line accumulation path on MI355X (gfx950).

The compute runs in hand-written HIP kernels behind the C ABI of include/lg.h
(light_garden_amd/build/liblg.so). This package is the host-side mirror of the reference's
scene/object/light API and thin wrappers over that ABI; it has no CPU fallback.
"""
from . import abi
from . import geo
from .geo import CubicBezier, Circle, ConvexPolygon, Ellipse, LineSegment, Logic, Rect, lens, rotation
from .tracer import (Context, DirectionalLight, Image, Material, Object, PointLight, Segments, SpotLight,
                     Tracer, batch_render, lights_array, objects_array)
from .string_mod import ModRemColor, StringMod
from .server import Server
from . import configs

__all__ = [
    "abi", "geo", "configs", "CubicBezier", "Circle", "ConvexPolygon", "Ellipse", "LineSegment", "Logic",
    "Rect", "lens", "rotation", "Context", "DirectionalLight", "Image", "Material", "Object", "PointLight",
    "Segments", "SpotLight", "Tracer", "batch_render", "lights_array", "objects_array", "ModRemColor", "StringMod", "Server",
]
