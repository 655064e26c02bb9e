"""imd_amd -- MI355X-native neighbour-list pair/EAM path of IMD.

``imd_amd.imdg`` binds the HIP library (include/imdg.h); ``imd_amd.host``
holds the C host that keeps IMD's parameter files, table formats and
calc_forces/move_atoms entry points; ``imd_amd.workloads`` writes inputs in
IMD's formats.
"""
__all__ = ["imdg", "workloads", "build"]
