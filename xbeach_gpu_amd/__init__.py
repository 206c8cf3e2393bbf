"""MI355X-native surf-beat step of XBeach_GPU: HIP kernels for gfx950 behind
the C-ABI of include/xbg.h, with a host-side mirror of the reference's step
interface (Domain) and main loop (driver)."""
from ._ffi import XbgParams, default_params, field_names, lib  # noqa: F401
from .domain import Domain, ModelCrashed, XbgError, next_dt, theta_tables  # noqa: F401
