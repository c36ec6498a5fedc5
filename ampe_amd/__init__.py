"""MI355X (gfx950) implementation of AMPE's per-timestep hot path.

The compute path is the HIP library ampe_amd/libampe_hip.so (C ABI declared in
include/ampe_hip.h); ampe_amd.hip binds it with ctypes and mirrors the
reference's strategy/operator interfaces; ampe_amd.params holds the ABI
structs and the seeded workloads.
"""
