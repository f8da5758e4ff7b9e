"""ctypes binding of libdage_ops.so (the C-ABI declared in include/dage_ops.h).

The library is built in-tree (dage_amd/libdage_ops.so) by ``build()``; there
is no fallback: loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdage_ops.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "dage_ops.h")

DAGE_OK = 0
DAGE_ERR_ARG = 1
DAGE_ERR_HIP = 2
DAGE_ERR_ALLOC = 3
DAGE_ERR_STATE = 4
DAGE_ERR_UNSUPPORTED = 5

DAGE_PREC_FP64 = 0
DAGE_PREC_BF16X6 = 1
DAGE_PREC_BF16X3 = 2
DAGE_OP_NO_CODA = 1

_c_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_int = ctypes.c_int
_dbl = ctypes.c_double

# name -> argtypes (all return int unless listed in _RESTYPES)
_SIGNATURES = {
    "dage_abi_version": [],
    "dage_last_error": [],
    "dage_ctx_create": [_int, ctypes.POINTER(_c_p)],
    "dage_ctx_destroy": [_c_p],
    "dage_ctx_set_stream": [_c_p, _c_p],
    "dage_ctx_get_stream": [_c_p, ctypes.POINTER(_c_p)],
    "dage_ctx_synchronize": [_c_p],
    "dage_ctx_profile": [_c_p, _int],
    "dage_ctx_profile_query": [_c_p, ctypes.c_char_p, _dp, _ip],
    "dage_ctx_time_begin": [_c_p],
    "dage_ctx_time_end": [_c_p, _dp],
    "dage_cube_create": [_c_p, _ip, ctypes.POINTER(_c_p)],
    "dage_cube_destroy": [_c_p],
    "dage_cube_upload": [_c_p, _dp],
    "dage_cube_download": [_c_p, _dp],
    "dage_cube_set_to_zero": [_c_p],
    "dage_cube_device_pointer": [_c_p, ctypes.POINTER(_c_p)],
    "dage_cube_copy": [_c_p, _c_p],
    "dage_gauss_quad": [_int, _int, _dbl, _dbl, _dbl, _dp, _dp],
    "dage_grid3d_create": [_c_p, _dp, _ip, _int, _dp, _dp, _dp, _int, ctypes.POINTER(_c_p)],
    "dage_grid3d_create_equidistant": [_c_p, _dp, _int, _dbl, _int, ctypes.POINTER(_c_p)],
    "dage_grid3d_destroy": [_c_p],
    "dage_grid3d_shape": [_c_p, _ip],
    "dage_grid3d_coord": [_c_p, _int, _dp],
    "dage_grid3d_weights": [_c_p, _int, _dp],
    "dage_coulomb3d_create": [_c_p, _c_p, _c_p, _dp, _dp, _int, _dbl, _int, ctypes.POINTER(_c_p)],
    "dage_operator_set_weights": [_c_p, _dp],
    "dage_operator_get_weights": [_c_p, _dp],
    "dage_operator_set_energy": [_c_p, _dbl],
    "dage_operator_set_matrices": [_c_p, _dp, _dp, _dp],
    "dage_operator_get_matrices": [_c_p, _dp, _dp, _dp],
    "dage_operator_apply": [_c_p, _dp, _dp, _int],
    "dage_operator_apply_device": [_c_p, _c_p, _c_p, _int],
    "dage_operator_apply_device_noalloc": [_c_p, _c_p, _c_p],
    "dage_operator_noalloc_limit": [ctypes.POINTER(ctypes.c_size_t)],
    "dage_operator_workspace_bytes": [_c_p, ctypes.POINTER(ctypes.c_size_t)],
    "dage_operator_get_band": [_c_p, ctypes.POINTER(ctypes.c_int)],
    "dage_operator_get_lowrank": [_c_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "dage_operator_get_tiling": [_c_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int)],
    "dage_operator_destroy": [_c_p],
    "dage_cube_integrate_device": [_c_p, _c_p, _c_p, _dp],
    "dage_bubbles_create": [_c_p, _int, _int, _int, _int, ctypes.POINTER(_c_p)],
    "dage_bubbles_set": [_c_p, _int, _dp, _dbl, _int, _dp, _dp],
    "dage_bubbles_inject_device": [_c_p, _c_p, _int, _dbl, _c_p],
    "dage_bubbles_inject": [_c_p, _c_p, _int, _dbl, _dp],
    "dage_bubbles_destroy": [_c_p],
    "dage_bubbles_create_like": [_c_p, _int, _int, ctypes.POINTER(_c_p)],
    "dage_bubbles_info": [_c_p, _ip, _ip, _ip, _ip],
    "dage_bubbles_npoints": [_c_p, _int, _ip],
    "dage_bubbles_get": [_c_p, _int, _dp, _dp, _dp, _dp],
    "dage_bubbles_axpby": [_dbl, _c_p, _dbl, _c_p],
    "dage_bubbles_coulomb_transform": [_c_p, _c_p],
    "dage_bubbles_helmholtz_transform": [_c_p, _dbl, _c_p],
    "dage_bubbles_copy_lmax": [_c_p, _int, ctypes.POINTER(_c_p)],
    "dage_function3d_inject_extra_bubbles": [_c_p, _c_p, _c_p, _int, ctypes.POINTER(_c_p)],
    "dage_operator_operate_on": [_c_p, _c_p, _c_p, _c_p, _c_p, _int],
    "dage_bubbles_multiply": [_c_p, _c_p, _c_p],
    "dage_bubbles_integrate": [_c_p, _dp],
    "dage_function3d_multiply": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p],
    "dage_function3d_contaminants": [_c_p, _c_p, _c_p, _int, _dp],
    "dage_operator_apply_device_strided": [_c_p, _c_p, _int, _int, _c_p, _int, _int, _int],
    "dage_bubbles_inject_device_strided": [_c_p, _c_p, _int, _dbl, _c_p, _int, _int],
    "dage_function3d_taylor_bubbles": [_c_p, _c_p, _c_p, _int, ctypes.POINTER(_c_p)],
    "dage_function3d_multiply_taylor": [_c_p, _c_p, _c_p, _c_p, _c_p, _int, _c_p, _c_p],
    "dage_fun_write": [ctypes.c_char_p, ctypes.c_char_p, _int, _ip, _dp, _int, _int, _int, _dp,
                       _dp, _ip, ctypes.POINTER(_dp)],
    "dage_fun_read_info": [ctypes.c_char_p, ctypes.c_char_p, _ip, _ip, _ip, _ip, _ip],
    "dage_fun_read": [ctypes.c_char_p, _dp, _dp, _dp, _ip, ctypes.POINTER(_dp)],
    "dage_g3d_write": [ctypes.c_char_p, _int, _int, _ip, _dp, ctypes.POINTER(_dp)],
    "dage_g3d_read": [ctypes.c_char_p, _ip, _ip, _ip, _dp, ctypes.POINTER(_dp)],
    "dage_radial_grid_steps": [_dbl, _int, _int, _dbl, _dp],
    "dage_gbfmm_create": [_c_p, _c_p, _int, _int, _dp, _dp, _int, ctypes.POINTER(_c_p)],
    "dage_gbfmm_apply_device": [_c_p, _c_p, _c_p, _int],
    "dage_gbfmm_apply": [_c_p, _dp, _dp, _int],
    "dage_gbfmm_set_energy": [_c_p, _dbl],
    "dage_gbfmm_box_count": [_c_p, _ip],
    "dage_gbfmm_get_band": [_c_p, _ip],
    "dage_gbfmm_get_lowrank": [_c_p, _ip, _ip],
    "dage_gbfmm_get_multipoles": [_c_p, _dp],
    "dage_gbfmm_get_local_expansions": [_c_p, _dp],
    "dage_gbfmm_destroy": [_c_p],
    "dage_fmm_interaction_matrix": [_int, _dp, _dp],
    "dage_fmm_translation_matrix": [_int, _dp, _dp],
    "dage_bessel_modified_spherical": [_int, _dbl, _dp, _dp],
    "dage_fmm_yukawa_interaction_matrix": [_int, _dbl, _dp, _dbl, _dp],
    "dage_fmm_yukawa_translation_matrix": [_int, _dbl, _dp, _dbl, _dp],
    "dage_helmholtz_batch_create": [_c_p, _int, ctypes.POINTER(_c_p), _dp, _dp, _dp, _int, _dbl,
                                    _int, _int, ctypes.POINTER(_c_p)],
    "dage_batch_input": [_c_p, _int, ctypes.POINTER(_c_p)],
    "dage_batch_output": [_c_p, _int, ctypes.POINTER(_c_p)],
    "dage_batch_upload": [_c_p, _int, _dp],
    "dage_batch_download": [_c_p, _int, _dp],
    "dage_batch_set_energies": [_c_p, _dp],
    "dage_batch_apply": [_c_p],
    "dage_batch_apply_eager": [_c_p],
    "dage_batch_tune": [_c_p, _int, _ip],
    "dage_batch_group_count": [_c_p, _ip],
    "dage_batch_destroy": [_c_p],
    "dage_vec_axpby": [_c_p, ctypes.c_long, _dbl, _c_p, _dbl, _c_p],
    "dage_vec_mul": [_c_p, ctypes.c_long, _c_p, _c_p, _c_p],
    "dage_vec_square_acc": [_c_p, ctypes.c_long, _dbl, _c_p, _c_p],
    "dage_vec_lda_x": [_c_p, ctypes.c_long, _c_p, _c_p, _c_p],
    "dage_vec_lincomb": [_c_p, ctypes.c_long, _int, ctypes.POINTER(_c_p), _dp, _int,
                         ctypes.POINTER(_c_p)],
    "dage_cube_gram": [_c_p, _int, ctypes.POINTER(_c_p), _int, ctypes.POINTER(_c_p), _dp],
    "dage_cube_gradient": [_c_p, _int, _c_p, _c_p],
    "dage_cube_stiffness": [_c_p, _int, _c_p, _c_p],
    "dage_conv3d_forward": [_c_p, _int, _int, _int, _int, _int, _c_p, _c_p, _c_p, _int, _c_p],
    "dage_conv3d_backward_data": [_c_p, _int, _int, _int, _int, _int, _c_p, _c_p, _c_p],
    "dage_conv3d_backward_filter": [_c_p, _int, _int, _int, _int, _int, _c_p, _c_p, _c_p, _c_p],
    "dage_relu_backward": [_c_p, ctypes.c_long, _c_p, _c_p],
    "dage_adam_step": [_c_p, _int, _c_p, _c_p, _c_p, _c_p, ctypes.c_float, ctypes.c_float,
                       ctypes.c_float, ctypes.c_float, _int, _c_p],
    "dage_weights_to_bf16": [_c_p, _int, _c_p, _c_p],
    "dage_mse_loss": [_c_p, ctypes.c_long, _c_p, _c_p, _c_p, _dp],
    "dage_cvt_f64_f32": [_c_p, ctypes.c_long, _c_p, _dbl, _c_p],
    "dage_axpy_f32_f64": [_c_p, ctypes.c_long, _dbl, _c_p, _c_p],
    "dage_device_alloc": [_c_p, ctypes.c_size_t, ctypes.POINTER(_c_p)],
    "dage_device_free": [_c_p, _c_p],
    "dage_memcpy": [_c_p, _c_p, _c_p, ctypes.c_size_t, _int],
}
_RESTYPES = {"dage_last_error": ctypes.c_char_p}

_lib = None


class DageError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, func: str, code: int, message: str):
        super().__init__(f"{func} failed with status {code}: {message}")
        self.code = code


def build() -> None:
    """Compile libdage_ops.so for gfx950 with the committed Makefile."""
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(_HERE, "csrc")], check=True)


def header_symbols() -> list[str]:
    """Function names declared in include/dage_ops.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dage_\w+)\s*\(", text, re.M)))


def load():
    """Load libdage_ops.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, ctypes.c_int)
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise DageError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != DAGE_OK:
        msg = lib.dage_last_error()
        raise DageError(name, rc, msg.decode() if msg else "")
