"""scp_amd — MI355X-native sfa/sim leapfrog step (real / complex / gauged).

Native pieces (built by `make -C scp_amd`):
  lib/libscp_amd.so  HIP kernels for gfx950 + the C-ABI of include/scp_amd.h
  lib/libscp_io.so   config keys, SFA snapshot container, seed generator (C)
  bin/scp_sim        the drop-in CLI of the reference sfa/sim/scp_sim binary
Python modules are thin ctypes views over those libraries.
"""
from .grid import (Grid, Params, ScpError, make_params, MODE_REAL, MODE_COMPLEX,  # noqa: F401
                   MODE_GAUGED, PREC_F64, PREC_F32C, DIAG_NAMES)
