"""ctypes binding of libcsipm.so (the C ABI declared in include/csipm.h).

The library is built in-tree (compressedsensingipm.jl_amd/csipm/libcsipm.so) by
`csipm.build.build()`.  Loading fails loudly when it is missing: there is no
CPU fallback anywhere on the product path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsipm.so")

OK = 0
ERR_ARG, ERR_SHAPE, ERR_ALLOC, ERR_HIP, ERR_UNSUPPORTED, ERR_NOT_CONVERGED = 1, 2, 3, 4, 5, 6
F64, F32, MIXED = 0, 1, 2

# every exported symbol, with (restype, argtypes)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_i = ctypes.c_int
SIGNATURES = {
    "csipm_version": (ctypes.c_char_p, []),
    "csipm_last_error": (ctypes.c_char_p, []),
    "csipm_op_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(_i64), _i, _i,
                             ctypes.POINTER(_i64), _i64, _vp]),
    "csipm_op_destroy": (_i, [_vp]),
    "csipm_op_set_stream": (_i, [_vp, _vp]),
    "csipm_op_info": (_i, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i)]),
    "csipm_op_fft_shape": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                ctypes.POINTER(_i)]),
    "csipm_tune_set": (_i, [_i, _i, _i, _i, _i, _i]),
    "csipm_tune_clear": (_i, []),
    "csipm_tune_bluestein": (_i, [_i, _i]),
    "csipm_op_bytes": (_i, [_vp, ctypes.POINTER(_i64)]),
    "csipm_op_footprint": (_i, [_i, ctypes.POINTER(_i64), _i, _i64, ctypes.POINTER(_i64),
                                ctypes.POINTER(_i64)]),
    "csipm_beta_to_dft": (_i, [_vp, _vp, _vp, _i]),
    "csipm_dft_to_beta": (_i, [_vp, _vp, _vp, _i]),
    "csipm_m_perp_beta": (_i, [_vp, _vp, _vp]),
    "csipm_m_perpt_z": (_i, [_vp, _vp, _vp]),
    "csipm_reconstruct": (_i, [_vp, _vp, _vp]),
    "csipm_gondzio_obj": (_i, [_vp, _vp, _d, ctypes.POINTER(_d)]),
    "csipm_gondzio_grad": (_i, [_vp, _vp, _d, _vp]),
    "csipm_gondzio_cons": (_i, [_vp, _vp, _vp, _vp]),
    "csipm_gondzio_jtprod": (_i, [_vp, _vp, _vp]),
    "csipm_gondzio_kkt_apply": (_i, [_vp, _vp, _vp, _vp, _vp, _d, _d]),
    "csipm_gondzio_kkt_mul_full": (_i, [_vp, _vp, _vp, _d, _d]),
    "csipm_gondzio_precond_build": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "csipm_gondzio_condense_rhs": (_i, [_vp, _vp, _vp]),
    "csipm_gondzio_unpack_solution": (_i, [_vp, _vp, _vp]),
    "csipm_gondzio_cg_solve": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i64,
                                    ctypes.POINTER(_i64), ctypes.POINTER(_d),
                                    ctypes.POINTER(_i)]),
    "csipm_gondzio_solve": (_i, [_vp, _vp, _d, _d, _i64, _d, _d, _i64, _vp, _vp, _vp]),
    "csipm_mask_index_missing": (_i, [_vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "csipm_mask_index_missing_noalloc": (_i, [_vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "csipm_mtm": (_i, [_vp, _vp, _vp]),
    "csipm_kkt_apply": (_i, [_vp, _vp, _vp, _vp, _vp, _d, _d]),
    "csipm_precond_build": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "csipm_precond_apply": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _d]),
    "csipm_cg_solve": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i64,
                            ctypes.POINTER(_i64), ctypes.POINTER(_d), ctypes.POINTER(_i)]),
    "csipm_solve": (_i, [_vp, _vp, _d, _d, _i64, _d, _d, _i64, _vp, ctypes.POINTER(_d)]),
    "csipm_admm": (_i, [_vp, _vp, _d, _d, _i64, _d, _vp, ctypes.POINTER(_d)]),
    "csipm_batch_create": (_i, [ctypes.POINTER(_vp), _i, _i64, _i, _vp]),
    "csipm_batch_destroy": (_i, [_vp]),
    "csipm_batch_slots": (_i, [_vp, ctypes.POINTER(_i64)]),
    "csipm_batch_bytes": (_i, [_i, _i64, ctypes.POINTER(_i64)]),
    "csipm_batch_admit": (_i, [_vp, _vp, _vp, _d, _vp, _i64, ctypes.POINTER(_i)]),
    "csipm_batch_step": (_i, [_vp, _d, _i64, _d, _d, ctypes.POINTER(_i)]),
    "csipm_batch_poll": (_i, [_vp, ctypes.POINTER(_i), _i, ctypes.POINTER(_i)]),
    "csipm_batch_harvest": (_i, [_vp, _i, ctypes.POINTER(_d)]),
    "csipm_batch_admit_many": (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_d), ctypes.POINTER(_vp), _i64,
                                    ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "csipm_batch_harvest_many": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_d)]),
    "csipm_batch_apply_timed": (_i, [_vp, _i64, ctypes.POINTER(_d), ctypes.POINTER(_d), _i,
                                     ctypes.POINTER(_i)]),
    "csipm_batch_launch_name": (ctypes.c_char_p, [_vp, _i]),
    "csipm_batch_stats": (_i, [_vp, ctypes.POINTER(_d)]),
    "csipm_batch_solve": (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                               ctypes.POINTER(_d), _d, _i64, _d, _d, _i64, ctypes.POINTER(_vp),
                               ctypes.POINTER(_d)]),
    "csipm_nlp_obj": (_i, [_vp, _vp, _vp, _d, ctypes.POINTER(_d)]),
    "csipm_nlp_grad": (_i, [_vp, _vp, _vp, _d, _vp]),
    "csipm_nlp_cons": (_i, [_vp, _vp, _vp]),
    "csipm_kkt_mul_full": (_i, [_vp, _vp, _vp, _d, _d]),
    "csipm_jtprod": (_i, [_vp, _vp, _vp]),
    "csipm_condense_rhs": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "csipm_unpack_solution": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "csipm_kkt_apply_timed": (_i, [_vp, _vp, _vp, _vp, _vp, _i64, _i, ctypes.POINTER(_d),
                                   ctypes.POINTER(_d), ctypes.POINTER(_i)]),
    "csipm_kkt_pass_name": (ctypes.c_char_p, [_vp, _i, _i]),
}


class CsipmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"csipm error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded library (loads on first use; raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libcsipm.so not found at {LIB_PATH}: run "
                               "`python -c 'import __graft_entry__ as g; g.build()'` first")
        # torch first: libcsipm then binds to the HIP runtime torch has loaded
        # (two runtimes in one process do not share devices)
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # the autotuned launch shapes of the batched FFT passes (committed table)
        from . import tune
        tune.load(L)
        _lib = L
    return _lib


def check(code):
    if code != OK:
        msg = lib().csipm_last_error().decode(errors="replace")
        raise CsipmError(code, msg)
    return code


def call(name, *args):
    return check(getattr(lib(), name)(*args))
