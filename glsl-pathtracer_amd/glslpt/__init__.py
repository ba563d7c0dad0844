"""glslpt — MI355X path-tracing backend with the scene surface of
knightcrawler25/GLSL-PathTracer (Scene / Renderer), HIP kernels underneath."""
from . import _capi
from .scene import Scene, material, quad_light, sphere_light, distant_light, mat4

__all__ = ["Scene", "material", "quad_light", "sphere_light", "distant_light", "mat4", "_capi"]
