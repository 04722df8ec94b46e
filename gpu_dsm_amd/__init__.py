"""gpu_dsm_amd: Discrete Slip-Link Model ensemble KMC on AMD Instinct MI355X (gfx950).

The product is the C-ABI shared library built from gpu_dsm_amd/csrc (declared in
include/dsm.h) plus the gpu_DSM command-line driver.  `engine` is a thin ctypes
binding used by Python callers; it fails loudly if the library is not built.
"""
from .engine import (DsmError, Engine, load_library, mem_info, probe_com, probe_copy, probe_math, probe_philox, probe_scan,  # noqa: F401
                     run_main, serve, serve_plan, LIB_PATH, EXE_PATH, EXPORTS)

__all__ = ["DsmError", "Engine", "load_library", "mem_info", "probe_com", "probe_copy", "probe_math", "probe_philox", "probe_scan",
           "run_main", "serve", "serve_plan", "LIB_PATH", "EXE_PATH", "EXPORTS"]
