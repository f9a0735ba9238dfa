"""MI355X-native per-time-step path of the 3D droplet solver
(farzan-t/3D_Droplet_simulation_2007): HIP kernels behind a C-ABI drop-in
(include/droplet_gpu.h) and this Python host side.

The package directory name starts with a digit; import it with
``importlib.import_module("3d_droplet_simulation_2007_amd")``.
"""
from .api import (FIELD_IDS, DropletError, DropletGPU, abcei, acfromac, addint, al3d, bc_scalar, fl3d,  # noqa: F401
                  filterc27, format_real, interp, lib, make_inputdata, poisson_solve, profile_start, profile_stop,
                  relax, resid, rstrct, set_lookahead, set_smoother, set_tile_smoother, set_tuning, slvsml)
