"""LS-SPH hot path on MI355X (gfx950): host-side package.

    config     config.h macros -> Param (init_param) -> lssph_params_t
    particles  lattice / wall / initial-condition setup (setup_closed_box ...)
    engine     ctypes binding of lib/liblssph_hip.so (the C ABI)
    io         reference output formats (write_data, write_param)
    workloads  the benchmark configurations
"""
from .config import Config, Param, Params, init_param, read_config_h  # noqa: F401
