"""ctypes binding of the C-ABI in include/xbg.h (libxbg.so).

The shared library is built in-tree by ``__graft_entry__.build()`` (hipcc for
gfx950).  Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxbg.so")

XBG_OK = 0
XBG_E_ARG = -1
XBG_E_HIP = -2
XBG_E_CRASH = -3
XBG_E_STATE = -4


class XbgParams(C.Structure):
    """Mirror of ``xbg_params`` (include/xbg.h), itself the numeric part of
    XBGPUParam (XBeachGPU.h:62-164)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("ntheta", C.c_int),
        ("dx", C.c_double), ("grdalpha", C.c_double),
        ("g", C.c_double), ("rho", C.c_double), ("eps", C.c_double), ("cf", C.c_double),
        ("cfsand", C.c_double), ("cfreef", C.c_double),
        ("nuh", C.c_double), ("nuhfac", C.c_double),
        ("usesmago", C.c_int),
        ("smag", C.c_double), ("lat", C.c_double), ("fc", C.c_double), ("Cd", C.c_double),
        ("wci", C.c_double), ("hwci", C.c_double),
        ("breakmodel", C.c_int),
        ("gammaa", C.c_double), ("n", C.c_double), ("alpha", C.c_double), ("gammax", C.c_double),
        ("beta", C.c_double), ("fw", C.c_double), ("fwsand", C.c_double), ("fwreef", C.c_double),
        ("roller", C.c_int),
        ("thetamin", C.c_double), ("thetamax", C.c_double), ("dtheta", C.c_double),
        ("D50", C.c_double), ("D90", C.c_double), ("rhosed", C.c_double), ("wws", C.c_double),
        ("drydzmax", C.c_double), ("wetdzmax", C.c_double), ("maxslpchg", C.c_double),
        ("por", C.c_double), ("morfac", C.c_double),
        ("sus", C.c_double), ("bed", C.c_double), ("facsk", C.c_double), ("facas", C.c_double),
        ("CFL", C.c_double), ("sedstart", C.c_double), ("endtime", C.c_double),
        ("outputtimestep", C.c_double),
        ("swave", C.c_int), ("flow", C.c_int), ("sedtrans", C.c_int), ("morphology", C.c_int),
        ("dtbc", C.c_double), ("rtlength", C.c_double), ("sprdthr", C.c_double),
        ("nmax", C.c_double), ("fcutoff", C.c_double),
        ("random", C.c_int), ("nspr", C.c_int),
        ("zb_accum", C.c_int),
    ]

    def copy(self) -> "XbgParams":
        q = XbgParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(self))
        return q

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.POINTER(XbgParams)
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_CTX = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/xbg.h
PROTOTYPES = {
    "xbg_default_params": (None, [_P]),
    "xbg_finalize_params": (None, [_P]),
    "xbg_theta_tables": (None, [_P, _F, _F, _F]),
    "xbg_next_dt": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_double, _D]),
    "xbg_interptime": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
    "xbg_create": (C.c_int, [_P, C.POINTER(_CTX)]),
    "xbg_destroy": (C.c_int, [_CTX]),
    "xbg_last_error": (C.c_char_p, [_CTX]),
    "xbg_set_field": (C.c_int, [_CTX, C.c_char_p, _F, C.c_size_t]),
    "xbg_get_field": (C.c_int, [_CTX, C.c_char_p, _F, C.c_size_t]),
    "xbg_field_size": (C.c_size_t, [_CTX, C.c_char_p]),
    "xbg_field_names": (C.c_void_p, []),
    "xbg_init_state": (C.c_int, [_CTX, C.c_double]),
    "xbg_set_wave_forcing": (C.c_int, [_CTX, _F, C.c_double]),
    "xbg_set_flow_forcing": (C.c_int, [_CTX, _F, _F, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double]),
    "xbg_set_bnd_series": (C.c_int, [_CTX, _D, _D, C.c_int]),
    "xbg_use_bnd_record": (C.c_int, [_CTX, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double]),
    "xbg_flow_bnd": (C.c_int, [_CTX, C.c_double, C.c_double]),
    "xbg_wave_step": (C.c_int, [_CTX, C.c_double, C.c_double]),
    "xbg_flow_step": (C.c_int, [_CTX, C.c_double]),
    "xbg_sed_step": (C.c_int, [_CTX, C.c_double]),
    "xbg_step": (C.c_int, [_CTX, C.c_double, C.c_double]),
    "xbg_cfl_dt": (C.c_int, [_CTX, C.c_int, _D, _D]),
    "xbg_advance_resident": (C.c_int, [_CTX, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _D, _D]),
    "xbg_accumulate_means": (C.c_int, [_CTX]),
    "xbg_reset_means": (C.c_int, [_CTX]),
    "xbg_divide_means": (C.c_int, [_CTX, C.c_double]),
    "xbg_mass_balance": (C.c_int, [_CTX, _D, _D]),
    "xbg_mass_balance_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_double, C.c_void_p, _D, _D]),
    "xbg_noalloc_last_error": (C.c_char_p, []),
    "xbg_sync": (C.c_int, [_CTX]),
    "xbg_stream": (C.c_void_p, [_CTX]),
    "xbg_profile": (C.c_int, [_CTX, C.c_int]),
    "xbg_profile_report": (C.c_int, [_CTX, C.c_char_p, C.c_size_t]),
    "xbg_jonswap_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "xbg_jonswap": (C.c_int, [C.c_void_p, _D, _D, _D]),
    "xbg_bnd_nrec": (C.c_int, [_P]),
    "xbg_wave_groups": (C.c_int, [_P, C.c_double, C.c_int, C.c_int, _D, _D, _D, C.c_uint, _D, _D, _D]),
    "xbg_constant_waves": (C.c_int, [_P, C.c_int, C.c_void_p, _F, _D, _D, _D]),
    "xbg_bnd_uniforms": (C.c_int, [C.c_uint, C.c_int, _D]),
    "xbg_hilbert": (C.c_int, [C.c_int, _D, _D, _D]),
    "xbg_bnd_synth": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), _D, _D, _D]),
    "xbg_nc_create": (C.c_void_p, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_double]),
    "xbg_nc_def_var": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, _F]),
    "xbg_nc_write_time": (C.c_int, [C.c_void_p, C.c_double]),
    "xbg_nc_write_var_step": (C.c_int, [C.c_void_p, C.c_char_p, _F]),
    "xbg_nc_close": (C.c_int, [C.c_void_p]),
    "xbg_nc_error": (C.c_char_p, []),
    "xbg_nc_grid_size": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "xbg_nc_read_zb": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _F]),
    "xbg_nc_spectrum_size": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "xbg_nc_spectrum": (C.c_int, [C.c_char_p, C.c_double, _D, _D, _D]),
    "xbg_xbbnd_read": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, _D, _D]),
    "xbg_bndnc_head": (C.c_int, [C.c_char_p, _P, C.POINTER(C.c_int)]),
    "xbg_bndnc_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, _D, _D]),
    "xbg_bndnc_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 9 + [_D, _D]),
    "xbg_set_points": (C.c_int, [_CTX, C.POINTER(C.c_int), C.c_int, C.c_int]),
    "xbg_record_points": (C.c_int, [_CTX]),
    "xbg_read_points": (C.c_int, [_CTX, _F, _F, C.c_int, C.POINTER(C.c_int)]),
    "xbg_default_inputs": (None, [C.c_void_p]),
    "xbg_param_line": (C.c_int, [C.c_char_p, _P, C.c_void_p]),
    "xbg_run_case": (C.c_int, [C.c_char_p, C.c_int]),
    "xbg_run_error": (C.c_char_p, []),
    "xbg_ens_create": (C.c_int, [C.c_int, C.POINTER(_CTX), C.c_void_p, C.POINTER(C.c_void_p)]),
    "xbg_ens_step": (C.c_int, [C.c_void_p, C.c_int]),
    "xbg_ens_sync": (C.c_int, [C.c_void_p]),
    "xbg_ens_status": (C.c_int, [C.c_void_p, _D, _D, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "xbg_ens_tune": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t]),
    "xbg_ens_kernel_times": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "xbg_ens_kernel_name": (C.c_char_p, [C.c_int]),
    "xbg_ens_last_error": (C.c_char_p, [C.c_void_p]),
    "xbg_ens_destroy": (C.c_int, [C.c_void_p]),
    "xbg_ens_window_enable": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "xbg_ens_window_add": (C.c_int, [C.c_void_p]),
    "xbg_ens_window_sample": (C.c_int, [C.c_void_p, C.c_int, _F, C.c_void_p]),
    "xbg_sur_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint, C.POINTER(C.c_void_p)]),
    "xbg_sur_destroy": (C.c_int, [C.c_void_p]),
    "xbg_sur_last_error": (C.c_char_p, [C.c_void_p]),
    "xbg_sur_nparams": (C.c_size_t, [C.c_void_p]),
    "xbg_sur_get_params": (C.c_int, [C.c_void_p, _F]),
    "xbg_sur_set_params": (C.c_int, [C.c_void_p, _F]),
    "xbg_sur_get_grads": (C.c_int, [C.c_void_p, _F]),
    "xbg_sur_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "xbg_sur_loss": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _D]),
    "xbg_sur_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _D]),
    "xbg_sur_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, _D]),
    "xbg_sur_sync": (C.c_int, [C.c_void_p]),
    "xbg_sur_conv_fwd": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int] * 7),
    "xbg_sur_conv_dgrad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int] * 6),
    "xbg_sur_conv_wgrad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int] * 6),
    "xbg_sur_pack_inputs": (C.c_int, [C.c_void_p] * 4 + [C.c_size_t, C.c_double, C.c_double, C.c_double, _F,
                                                         C.c_void_p, C.c_void_p]),
    "xbg_field_device": (C.c_void_p, [_CTX, C.c_char_p]),
    "xbg_state_save": (C.c_int, [_CTX]),
    "xbg_state_restore": (C.c_int, [_CTX]),
    "xbg_set_wave_products": (C.c_int, [_CTX, C.c_void_p, _F, C.c_double]),
    "xbg_step_nowave": (C.c_int, [_CTX, C.c_double, C.c_double]),
    "xbg_advance_window": (C.c_int, [_CTX, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, _D, _D, C.POINTER(C.c_int)]),
}


class XbgEnsDomain(C.Structure):
    """Mirror of ``xbg_ens_domain`` (run setup of one ensemble domain)."""

    _fields_ = [("t_start", C.c_double), ("dt_prev", C.c_double), ("endtime", C.c_double),
                ("t0", C.c_double), ("dtbc", C.c_double), ("Trep", C.c_double)]


MAX_TS = 64  # XBG_MAX_TS


class XbgInputs(C.Structure):
    """Mirror of ``xbg_inputs`` (the file names and options of XBGPUParam,
    XBeachGPU.h:116-139)."""

    _fields_ = [("bathy", C.c_char * 512), ("outfile", C.c_char * 512), ("sedthk", C.c_char * 512),
                ("wavebndfile", C.c_char * 512), ("slbndfile", C.c_char * 512), ("windfile", C.c_char * 512),
                ("wavebndtype", C.c_int), ("dt", C.c_double), ("outvars", C.c_char * 2048),
                ("ntsofile", C.c_int), ("ntsnode", C.c_int), ("tsofile", (C.c_char * 256) * MAX_TS),
                ("tsnode_i", C.c_int * MAX_TS), ("tsnode_j", C.c_int * MAX_TS)]


class XbgWavebnd(C.Structure):
    """Mirror of ``xbg_wavebnd`` (Wavebndparam, XBeachGPU.h:198-210)."""

    _fields_ = [("time", C.c_double), ("Hs", C.c_double), ("Tp", C.c_double), ("Dp", C.c_double),
                ("s", C.c_double), ("gamma", C.c_double)]

_lib = None


def lib(path: str | None = None) -> C.CDLL:
    """Load libxbg.so once (the in-tree build, or the file named by the
    XBG_LIB environment variable).  Raises if the HIP library has not been
    built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("XBG_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libxbg.so not found at {p}: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    L = C.CDLL(p)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def field_names() -> list[str]:
    ptr = lib().xbg_field_names()
    out, off = [], 0
    while True:
        s = C.string_at(ptr + off)
        if not s:
            break
        out.append(s.decode())
        off += len(s) + 1
    return out


def default_params(**kw) -> XbgParams:
    """Reference defaults (XBeachGPU.h:62-164) with keyword overrides, then
    the derived values of checkparamsanity (read_input.cpp:1227-1444).
    Angles (thetamin/thetamax/dtheta) are given in degrees, as in the
    parameter file, and converted with the reference's pi."""
    p = XbgParams()
    lib().xbg_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("thetamin", "thetamax", "dtheta"):
            v = v * 3.14159265 / 180.0
        setattr(p, k, v)
    lib().xbg_finalize_params(C.byref(p))
    return p
