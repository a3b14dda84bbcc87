"""fermiphasespace_amd -- MI355X (gfx950) fTWA trajectory engine.

A drop-in for the "fTWA" method of FermiPhaseSpace: the same .json system / parameters / results
format, behind an FHfTWASolver-shaped interface, computed by hand-written HIP kernels through the
C ABI of include/fps_ftwa.h.  Nothing here computes on the CPU.

    from fermiphasespace_amd import FHSystem, FHfTWASolver
    sys = FHSystem.from_json_file("sim.json")
    s = FHfTWASolver(sys)
    s.setMethodParameters({"dt": 0.01, "t_max": 1.0, "n_traj": 1000})
    s.initiateSolver(); s.runSimulation()
"""
from ._lib import (COULOMB_GENERAL, COULOMB_NULL, COULOMB_ONSITE, ENGINE_FP32_BATCHED, ENGINE_FP64_BATCHED,
                   ENGINE_FP64_PLAIN, Engine, FtwaError, device_info, ground_state, load, load_initial_state_json,
                   load_system_json, run_json_file)
from .solver import FHGPSRSolver, FHHFSolver, FHSystem, FHfTWASolver

__all__ = [
    "Engine", "FtwaError", "FHGPSRSolver", "FHHFSolver", "FHSystem", "FHfTWASolver", "device_info", "ground_state", "load",
    "load_initial_state_json", "load_system_json", "run_json_file",
    "COULOMB_NULL", "COULOMB_ONSITE", "COULOMB_GENERAL", "ENGINE_FP64_PLAIN", "ENGINE_FP64_BATCHED",
    "ENGINE_FP32_BATCHED",
]
