"""MI355X-native lattice self-consistent-field (Scheutjens-Fleer) hot path.

The compute path is the HIP library libnamics_scf.so (C ABI declared in
include/namics_scf.h); this package is its Python host binding.  There is no
CPU fallback: importing the binding without the built library raises.
"""
from .scf import ScfSystem, solve_batch, solve_resident, solve_resident_info, solve_jobs, Server, DeviceBuffer, device_info, synchronize, library_path, ScfError, autotune_forget, autotune_batch, run_input_file  # noqa: F401
