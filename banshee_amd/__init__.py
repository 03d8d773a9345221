"""banshee_amd: MI355X (gfx950) global-illumination hot path with the Banshee
scene / primitive / tracer / sampler API.

The compute path is libbanshee_hip.so (HIP kernels + C ABI, include/banshee_hip.h);
``banshee_amd.hip`` binds it, ``banshee_amd.scenes`` generates the seeded
procedural scenes in the layout it takes.
"""
__all__ = ["hip", "scenes"]
