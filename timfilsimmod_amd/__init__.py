"""MI355X-native (gfx950) RPY mobility hot path of TimFilSim.

Submodules: mobility (matvec back-end), flow (flow-field evaluator), solver (per-timestep saddle-point
solve and prescribed-beat driver), broyden (implicit-step Broyden engine), geometry and beat
(synthetic inputs), datafiles (the reference's data files and its file-driven flow-field run),
filament (the Cosserat filament model of the elastic timestep).
All compute runs in libtfs_rpy.so (include/tfs_rpy.h, tfs_solver.h, tfs_flow.h, tfs_broyden.h,
tfs_io.h, tfs_filament.h).
"""
from . import _lib, geometry, beat, mobility, solver, flow, ensemble, broyden, datafiles, filament  # noqa: F401
