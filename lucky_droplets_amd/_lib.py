"""ctypes binding of libld.so (the C-ABI declared in include/ld.h).

The library is built in-tree by ``lucky_droplets_amd.build``; importing this
module never builds.  When the library is missing, ``load()`` raises: there is
no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libld.so"

LD_OK, LD_ERR_ARG, LD_ERR_HIP, LD_ERR_BLOWUP, LD_ERR_CAPACITY = 0, 1, 2, 3, 4
LD_F32, LD_F64 = 0, 1
KERNELS = {"droplets": 0, "x_rhs": 1, "x_uw": 2, "y_pass": 3, "z_cross": 4, "z_vel": 5, "x_kn": 6, "z_kn": 7,
           "deposit": 8}

_c_double_p = ctypes.POINTER(ctypes.c_double)


class LdParams(ctypes.Structure):
    """Mirror of ``ld_params`` in include/ld.h."""

    _fields_ = [
        ("N", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("L", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("lp", ctypes.c_int32),
        ("isexplicit", ctypes.c_double),
        ("nshells", ctypes.c_int32),
        ("shell_no", ctypes.c_int32 * 8),
        ("f0", ctypes.c_double),
        ("isforcing", ctypes.c_double),
        ("tps", ctypes.c_double),
        ("g", ctypes.c_double),
        ("st_s", ctypes.c_double),
        ("tau_eta", ctypes.c_double),
        ("factor", ctypes.c_double),
        ("growthfactor", ctypes.c_double),
        ("max_particles", ctypes.c_int64),
        ("batch", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
    ]


class LdCnnParams(ctypes.Structure):
    """Mirror of ``ld_cnn_params`` in include/ld.h."""

    _fields_ = [
        ("N", ctypes.c_int32),
        ("nlayers", ctypes.c_int32),
        ("channels", ctypes.c_int32 * 7),
    ]


class LdError(RuntimeError):
    """An LD_ERR_* status from libld.so."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libld error {code}: {msg}")
        self.code = code


class BlowUp(LdError):
    """max Re(u_hat) exceeded 100 N^3 (the reference aborts, forced-dns-sm.py:618-620)."""


# name -> (restype, argtypes); every ld_* function returns int except ld_version.
_VP = ctypes.c_void_p
_SIGS = {
    "ld_create": [ctypes.POINTER(LdParams), ctypes.POINTER(_VP)],
    "ld_destroy": [_VP],
    "ld_set_stream": [_VP, _VP],
    "ld_last_error": [ctypes.c_char_p, ctypes.c_size_t],
    "ld_rfft": [_VP, _VP, _VP, ctypes.c_int],
    "ld_irfft": [_VP, _VP, _VP, ctypes.c_int],
    "ld_forcing": [_VP, _VP, _VP, _VP],
    "ld_clip_zero": [_VP, _VP, _VP],
    "ld_full_rhs": [_VP, _VP, _VP, _VP, ctypes.c_int64, _VP, _VP, _VP, _VP, ctypes.c_double, ctypes.c_double,
                    _VP, _VP],
    "ld_rk4": [_VP, ctypes.c_double, _VP, _VP, _VP, ctypes.c_int64, _VP, _VP, _VP],
    "ld_step": [_VP, ctypes.c_double, _VP, _VP, _VP, ctypes.c_int64, ctypes.c_int32],
    "ld_state_set": [_VP, _VP, _VP, _VP, ctypes.c_int64],
    "ld_state_get": [_VP, _VP, _VP, _VP],
    "ld_advance": [_VP, ctypes.c_double, ctypes.c_int32],
    "ld_advance_plain": [_VP, ctypes.c_double, ctypes.c_int32],
    "ld_status": [_VP, _c_double_p, ctypes.POINTER(ctypes.c_int32)],
    "ld_profile": [_VP, ctypes.c_int32, ctypes.c_int32],
    "ld_profile_read": [_VP, _c_double_p, ctypes.POINTER(ctypes.c_int64)],
    "ld_interp_cosine": [_VP, _VP, ctypes.c_int64, _VP, ctypes.c_int32, _VP],
    "ld_exterp_cosine_scalar": [_VP, _VP, ctypes.c_int64, _VP, _VP],
    "ld_prhs": [_VP, _VP, ctypes.c_int64, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    "ld_spectrum": [_VP, _VP, ctypes.c_int32, ctypes.c_double, _VP],
    "ld_spectrum_noalloc": [_VP, _VP, ctypes.c_int32, ctypes.c_double, _VP],
    "ld_flux": [_VP, _VP, _VP, _VP],
    "ld_field_stats": [_VP, _VP, ctypes.c_int32, _VP],
    "ld_dissipation": [_VP, _VP, _VP],
    "ld_divergence": [_VP, _VP, _VP],
    "ld_droplet_stats": [_VP, _VP],
    "ld_collide_pairs": [_VP, _VP, _VP, ctypes.c_int64, ctypes.c_double, _VP, ctypes.c_int64,
                         ctypes.POINTER(ctypes.c_int64)],
    "ld_coalesce": [_VP, _VP, ctypes.c_int64, _VP, ctypes.c_int64, _VP, ctypes.POINTER(ctypes.c_int64)],
    "ld_pair_stats": [_VP, _VP, ctypes.c_int64, ctypes.c_double, ctypes.c_int32, _VP, _VP, _VP],
    "ld_status_all": [_VP, _VP, _VP],
    "ld_tuning": [_VP, _VP, _VP],
    "ld_tuning_all": [_VP, _VP, _VP],
    "ld_spectral_resample": [_VP, _VP, ctypes.c_int32, _VP, ctypes.c_int32],
    "ld_fields_to_rows": [_VP, _VP, ctypes.c_int32, _VP, ctypes.c_float],
    "ld_velocity_gradient_spec": [_VP, _VP, _VP, ctypes.c_double],
    "ld_subgrid_compose": [_VP, _VP, ctypes.c_int32, _VP, _VP, ctypes.c_float],
    "ld_subgrid_residual": [_VP, _VP, _VP, ctypes.c_int32, _VP, ctypes.c_float],
    "ld_droplets_frozen_rk4": [_VP, _VP, ctypes.c_int64, _VP, ctypes.c_double, ctypes.c_int32],
    "ld_cnn_create": [ctypes.POINTER(LdCnnParams), ctypes.POINTER(_VP)],
    "ld_cnn_destroy": [_VP],
    "ld_cnn_set_stream": [_VP, _VP],
    "ld_cnn_nparams": [_VP, ctypes.POINTER(ctypes.c_int64)],
    "ld_cnn_set_params": [_VP, _VP],
    "ld_cnn_get_params": [_VP, _VP],
    "ld_cnn_forward": [_VP, _VP, _VP],
    "ld_cnn_grad": [_VP, _VP, _VP, _VP, ctypes.POINTER(ctypes.c_double)],
    "ld_cnn_train_step": [_VP, _VP, _VP, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                          ctypes.POINTER(ctypes.c_double)],
}
EXPORTS = tuple(_SIGS) + ("ld_version",)

_lib = None


def load(path: os.PathLike | None = None) -> ctypes.CDLL:
    """Load libld.so once and declare every signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -m lucky_droplets_amd.build` "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(p))
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    lib.ld_version.restype = ctypes.c_char_p
    lib.ld_version.argtypes = []
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load().ld_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc == LD_OK:
        return
    msg = last_error()
    if rc == LD_ERR_BLOWUP:
        raise BlowUp(rc, msg)
    raise LdError(rc, msg)
