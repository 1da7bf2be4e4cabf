"""MI355X (gfx950) ray-tracing hot path for bling: BVH traversal, ray/shape
intersection and the spectral path integrator behind a C ABI (include/bgpu.h,
libbgpu.so).  See runtime.py for the host-side mirror of the reference's
accelerator and renderer interfaces."""
__all__ = ["abi", "runtime", "scene", "scenes", "transform"]
