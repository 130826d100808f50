"""ctypes binding of libcfd_amd.so (the C ABI declared in include/cfd_amd.h).

The library is built in-tree by ``cfd_amd/csrc/Makefile``; importing this
module without it raises, so no code path can silently fall back to a CPU
implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfd_amd.so")

INT_FIELDS = ("ngalx ngalz ny dsty dsty1 ppp ygrid geometry_type ntilex ntilez "
              "extras24 flag_ctpress precision device ngalx2 ngalz2").split()
DBL_FIELDS = "Re alp bet dyq posth posth1 ytanh_beta CFL mpgx r c_i c_k stag".split()


class CParams(ctypes.Structure):
    """Mirror of ``cfd_params``."""
    _fields_ = [(n, ctypes.c_int) for n in INT_FIELDS] + [(n, ctypes.c_double) for n in DBL_FIELDS]


class CFileHeader(ctypes.Structure):
    """Mirror of ``cfd_file_header`` (record 1 and 2 of a field file)."""
    _fields_ = [(n, ctypes.c_double) for n in ("t", "Re", "alp", "bet", "mpgx")] + \
               [("nband", ctypes.c_int), ("iter", ctypes.c_int), ("N", ctypes.c_int * 20)]


# symbol -> (restype, argtypes)
_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_H = ctypes.c_void_p
SIGNATURES = {
    "cfd_create": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(ctypes.c_void_p)]),
    "cfd_destroy": (ctypes.c_int, [_H]),
    "cfd_last_error": (ctypes.c_char_p, [_H]),
    "cfd_build_ygrid": (ctypes.c_int, [ctypes.POINTER(CParams)] + [_D] * 6),
    "cfd_build_iblist": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.c_int, ctypes.c_int, _I, _D, _I]),
    "cfd_dims": (ctypes.c_int, [_H, _I]),
    "cfd_get_ygrid": (ctypes.c_int, [_H] + [_D] * 6),
    "cfd_get_iblist": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _I, _D, _I]),
    "cfd_set_field": (ctypes.c_int, [_H, ctypes.c_int, _D]),
    "cfd_get_field": (ctypes.c_int, [_H, ctypes.c_int, _D]),
    "cfd_set_scalars": (ctypes.c_int, [_H, _D]),
    "cfd_get_scalars": (ctypes.c_int, [_H, _D]),
    "cfd_rhs0": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_nonlinear": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_solve_u": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int]),
    "cfd_solve_v": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_meanflow_ctU": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_solve_p": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_vcorr": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_substep": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_step": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_sync": (ctypes.c_int, [_H]),
    "cfd_device_bytes": (ctypes.c_int, [_H, _D]),
    "cfd_profile": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_profile_read": (ctypes.c_int, [_H, _D, _I]),
    "cfd_field_file_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CFileHeader), ctypes.c_int,
                                            ctypes.c_double, _D, _D] + [ctypes.c_int] * 6 + [_D]),
    "cfd_field_file_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CFileHeader)] + [ctypes.c_int] * 4
                            + [_D, _I]),
    "cfd_record_out": (ctypes.c_int, [_H, ctypes.c_char_p, ctypes.c_int]),
    "cfd_band_table": (ctypes.c_int, [ctypes.POINTER(CParams), _I]),
    "cfd_plane_partition": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.c_int, _I]),
    "cfd_four_to_phys": (ctypes.c_int, [_H, ctypes.c_int, _D]),
    "cfd_four_to_phys_planes": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "cfd_cnn_nparams": (ctypes.c_int, []),
    "cfd_cnn_create": (ctypes.c_int, [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_void_p)]),
    "cfd_cnn_destroy": (ctypes.c_int, [_H]),
    "cfd_cnn_last_error": (ctypes.c_char_p, [_H]),
    "cfd_cnn_set_params": (ctypes.c_int, [_H, _F]),
    "cfd_cnn_get_params": (ctypes.c_int, [_H, _F]),
    "cfd_cnn_get_grads": (ctypes.c_int, [_H, _F]),
    "cfd_cnn_set_input": (ctypes.c_int, [_H, _F]),
    "cfd_cnn_set_target": (ctypes.c_int, [_H, _F]),
    "cfd_cnn_forward": (ctypes.c_int, [_H]),
    "cfd_cnn_get_output": (ctypes.c_int, [_H, _F]),
    "cfd_cnn_backward": (ctypes.c_int, [_H, _D]),
    "cfd_cnn_adam": (ctypes.c_int, [_H] + [ctypes.c_float] * 4),
    "cfd_cnn_dataset": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_cnn_train_slot": (ctypes.c_int, [_H, ctypes.c_int] + [ctypes.c_float] * 4 + [_D]),
    "cfd_cnn_eval_slot": (ctypes.c_int, [_H, ctypes.c_int, _D]),
    "cfd_cnn_load_slot": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_cnn_device_bytes": (ctypes.c_int, [_H, _D]),
    "cfd_cnn_set_scales": (ctypes.c_int, [_H, _F]),
    "cfd_cnn_get_scales": (ctypes.c_int, [_H, _F]),
    "cfd_closure_snapshot": (ctypes.c_int, [_H, _H, _H, ctypes.c_int, _D]),
    "cfd_closure_attach": (ctypes.c_int, [_H, _H, ctypes.c_int]),
    "cfd_closure_phys": (ctypes.c_int, [_H, ctypes.c_int, _D]),
    "cfd_create_tuned": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "cfd_get_tuning": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_int)]),
    "cfd_set_tuning": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_int)]),
    "cfd_closure_clip": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_meanflow_ctP": (ctypes.c_int, [_H]),
    "cfd_hist": (ctypes.c_int, [_H, _D]),
    "cfd_set_mode_filter": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
    "cfd_set_compat_lastplane": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_ensemble_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "cfd_ensemble_create_batched": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_void_p)]),
    "cfd_ensemble_destroy": (ctypes.c_int, [_H]),
    "cfd_ensemble_last_error": (ctypes.c_char_p, [_H]),
    "cfd_ensemble_substep": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_ensemble_step": (ctypes.c_int, [_H, ctypes.c_int]),
    "cfd_ensemble_sync": (ctypes.c_int, [_H]),
    "cfd_stats": (ctypes.c_int, [_H]),
    "cfd_write_stats": (ctypes.c_int, [_H, ctypes.c_char_p, ctypes.c_int]),
    "cfd_stats_request": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "cfd_stats_dims": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_int)]),
    "cfd_read_in": (ctypes.c_int, [_H, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _I]),
}

_lib = None


def load():
    """Load libcfd_amd.so once and attach the signatures; raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libcfd_amd.so is not built (%s); run `make -C cfd_amd/csrc` "
                "or __graft_entry__.build()" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class CfdError(RuntimeError):
    """Nonzero status from the C ABI."""


def check(rc, handle=None):
    if rc != 0:
        msg = load().cfd_last_error(handle)
        raise CfdError("cfd_amd error %d: %s" % (rc, msg.decode() if msg else "?"))
