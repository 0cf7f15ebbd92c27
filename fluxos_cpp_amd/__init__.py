"""MI355X-native FLUXOS-OVERLAND time step (hydrodynamics_calc loop).

The compute path is libfluxos_hip.so (hand-written HIP for gfx950 behind a C
ABI, include/fluxos_hip.h); ``capi`` binds it, ``gen`` makes synthetic inputs.
"""
__all__ = ["capi", "gen"]
