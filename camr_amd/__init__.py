"""camr_amd -- MI355X (gfx950) single-level hydro advance of CAMR.

The compute path is the HIP library ``libcamr_hydro.so`` (C-ABI in
include/camr_hydro.h); this package holds its host-side Python mirror
(``hydro``), the workload definitions and the build script.
"""
__all__ = ["hydro", "problems", "workloads"]
