"""ctypes binding of librbchem.so (the C ABI declared in include/rbchem.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C reprobus_amd/csrc``).  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librbchem.so")

RB_OK, RB_EINVAL, RB_EHIP, RB_ENOTABLE, RB_ENONFINITE, RB_ENOMEM, RB_EIO = 0, -1, -2, -3, -4, -5, -6
RB_NOALLOC_MAX_BOXES = 32768     # include/rbchem.h
ERROR_NAMES = {RB_EINVAL: "RB_EINVAL", RB_EHIP: "RB_EHIP", RB_ENOTABLE: "RB_ENOTABLE",
               RB_ENONFINITE: "RB_ENONFINITE", RB_ENOMEM: "RB_ENOMEM", RB_EIO: "RB_EIO"}

# every symbol include/rbchem.h declares
EXPORTS = ["rb_init", "rb_load_jtable", "rb_set_stream", "rb_chem_step", "rb_conserv_floor",
           "rb_chem_step_dev", "rb_chem_step_noalloc_dev", "rb_conserv_floor_dev",
           "rb_chem_converge_dev",
           "rb_chem_newton_dev", "rb_chem_surrogate_dev", "rb_sediment",
           "rb_sediment_dev", "rb_zenith", "rb_zenith_dev", "rb_altitude", "rb_altitude_dev",
           "rb_run_day_dev", "rb_transport_backtra_dev", "rb_transport_backtra",
           "rb_transport_weights_dev",
           "rb_transport_interp_dev", "rb_transport",
           "rb_sync", "rb_get_stats", "rb_last_error", "rb_finalize"]
# and include/rbmlp.h (the learned initial guess)
MLP_EXPORTS = ["rb_mlp_create", "rb_mlp_destroy", "rb_mlp_set_stream", "rb_mlp_last_error",
               "rb_mlp_param_count", "rb_mlp_set_params", "rb_mlp_get_params", "rb_mlp_set_norm",
               "rb_mlp_forward_dev", "rb_mlp_train_step_dev", "rb_mlp_gradients_dev",
               "rb_mlp_features_dev", "rb_mlp_guess_dev"]
# and include/rbio.h (history and forcing records, host only)
IO_EXPORTS = ["rb_history_write", "rb_history_read", "rb_ecmwf_write_time", "rb_ecmwf_read_time",
              "rb_io_last_error"]


class Grid(C.Structure):
    _fields_ = [("nlon", C.c_int32), ("nlat", C.c_int32), ("niv", C.c_int32),
                ("nivbas", C.c_int32), ("nivhau", C.c_int32)]


class Forcing(C.Structure):
    """rb_forcing: one forcing record of the trajectory code (device pointers)."""
    _fields_ = [("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p), ("t", C.c_void_p),
                ("lnps", C.c_void_p), ("day", C.c_double)]


class Config(C.Structure):
    _fields_ = [("niter", C.c_int32), ("check_finite", C.c_int32), ("block", C.c_int32),
                ("profile", C.c_int32), ("conv_refill", C.c_int32), ("conv_aa_start", C.c_int32),
                ("conv_aa_budget", C.c_int32), ("newton_waves", C.c_int32),
                ("conv_order", C.c_int32)]


class DayConfig(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("use_graph", C.c_int32), ("converged", C.c_int32),
                ("maxiter", C.c_int32), ("day0", C.c_double), ("dt", C.c_double),
                ("tol", C.c_double), ("accel", C.c_int32), ("pad_", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("boxes", C.c_int64), ("nonfinite", C.c_int64), ("last_kernel_ms", C.c_double),
                ("h2d_ms", C.c_double), ("d2h_ms", C.c_double), ("chem_ms_total", C.c_double),
                ("chem_launches", C.c_int64)]


class RbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load librbchem.so and declare its signatures (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "or `make -C reprobus_amd/csrc`")
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7).  Loading
    # it first makes this library bind to that runtime, so the process holds one HIP
    # runtime and device pointers from torch tensors are valid in our launches.
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    d = C.c_double
    L.rb_init.argtypes = [C.POINTER(P), C.c_int, C.POINTER(Config)]
    L.rb_load_jtable.argtypes = [P, P, P]
    L.rb_set_stream.argtypes = [P, P]
    L.rb_chem_step.argtypes = [P, C.POINTER(Grid), d] + [P] * 10
    L.rb_conserv_floor.argtypes = [P, C.POINTER(Grid), P, P]
    L.rb_chem_step_dev.argtypes = [P, C.POINTER(Grid), d] + [P] * 10 + [C.c_int]
    L.rb_chem_step_noalloc_dev.argtypes = [P, C.POINTER(Grid), d] + [P] * 10 + [C.c_int]
    L.rb_conserv_floor_dev.argtypes = [P, C.POINTER(Grid), P, P]
    L.rb_chem_converge_dev.argtypes = [P, C.POINTER(Grid), d, d, C.c_int32, C.c_int32] + [P] * 13 + [C.c_int]
    L.rb_chem_newton_dev.argtypes = [P, C.POINTER(Grid), d, d, C.c_int32, C.c_int32, C.c_int32] + \
        [P] * 13 + [C.c_int]
    L.rb_chem_surrogate_dev.argtypes = [P, C.POINTER(Grid), d, d, d, C.c_int32, C.c_int32] + \
        [P] * 13 + [C.c_int]
    L.rb_sediment.argtypes = [P, C.POINTER(Grid), d, C.c_int32] + [P] * 6
    L.rb_sediment_dev.argtypes = [P, C.POINTER(Grid), d, C.c_int32] + [P] * 6
    L.rb_zenith.argtypes = [P, C.POINTER(Grid), d, P, P, P]
    L.rb_zenith_dev.argtypes = [P, C.POINTER(Grid), d, P, P, P]
    L.rb_altitude.argtypes = [P, C.POINTER(Grid), d, d] + [P] * 10
    L.rb_altitude_dev.argtypes = [P, C.POINTER(Grid), d, d] + [P] * 10
    L.rb_run_day_dev.argtypes = [P, C.POINTER(Grid), C.POINTER(DayConfig)] + [P] * 13
    L.rb_transport_backtra_dev.argtypes = [P, C.POINTER(Grid), P, P, C.POINTER(Forcing),
                                           C.POINTER(Forcing), d, d, C.c_int32, C.c_int32,
                                           C.c_int32] + [P] * 10
    L.rb_transport_backtra.argtypes = [P, C.POINTER(Grid), P, P, C.POINTER(Forcing),
                                       C.POINTER(Forcing), d, d, C.c_int32, C.c_int32,
                                       C.c_int32] + [P] * 8
    L.rb_transport_weights_dev.argtypes = [P, C.POINTER(Grid)] + [P] * 8
    L.rb_transport_interp_dev.argtypes = [P, C.POINTER(Grid), P, P, C.c_int32, P, P]
    L.rb_transport.argtypes = [P, C.POINTER(Grid)] + [P] * 6 + [C.c_int32, P]
    L.rb_sync.argtypes = [P]
    L.rb_get_stats.argtypes = [P, C.POINTER(Stats)]
    L.rb_last_error.argtypes = [P]
    L.rb_last_error.restype = C.c_char_p
    L.rb_finalize.argtypes = [P]
    for name in EXPORTS:
        if name != "rb_last_error":
            getattr(L, name).restype = C.c_int
    f = C.c_float
    i64 = C.c_int64
    for name in MLP_EXPORTS:
        getattr(L, name).restype = C.c_int
    L.rb_mlp_create.argtypes = [C.POINTER(P), C.c_int]
    L.rb_mlp_destroy.argtypes = [P]
    L.rb_mlp_destroy.restype = None
    L.rb_mlp_set_stream.argtypes = [P, P]
    L.rb_mlp_last_error.argtypes = [P]
    L.rb_mlp_last_error.restype = C.c_char_p
    L.rb_mlp_param_count.argtypes = []
    L.rb_mlp_param_count.restype = i64
    L.rb_mlp_set_params.argtypes = [P, P]
    L.rb_mlp_get_params.argtypes = [P, P]
    L.rb_mlp_set_norm.argtypes = [P, P, P, P, P]
    L.rb_mlp_forward_dev.argtypes = [P, P, i64, P]
    L.rb_mlp_train_step_dev.argtypes = [P, P, P, i64, f, C.POINTER(f)]
    L.rb_mlp_gradients_dev.argtypes = [P, P, P, i64, P, C.POINTER(f)]
    L.rb_mlp_features_dev.argtypes = [P, C.POINTER(Grid), d] + [P] * 6
    L.rb_mlp_guess_dev.argtypes = [P, C.POINTER(Grid), d] + [P] * 7
    for name in IO_EXPORTS:
        getattr(L, name).restype = C.c_int
    L.rb_io_last_error.argtypes = []
    L.rb_io_last_error.restype = C.c_char_p
    i32 = C.c_int32
    L.rb_history_write.argtypes = [C.c_char_p, i32, i32, i32] + [P] * 9
    L.rb_history_read.argtypes = [C.c_char_p, i32, i32, i32] + [P] * 9
    L.rb_ecmwf_write_time.argtypes = [C.c_char_p, i32, i32, i32, i32] + [P] * 8
    L.rb_ecmwf_read_time.argtypes = [C.c_char_p, i32, i32, i32, i32] + [P] * 7
    _lib = L
    return L
