"""ctypes binding of libmom25d.so, the gfx950 kernels of the solve.

The library is built in-tree (`mom25d_amd/libmom25d.so`, see csrc/Makefile);
there is no fallback: every entry point of the solve goes through it and a
missing library raises ImportError-like `RuntimeError` at first use.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmom25d.so")

_p = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_f = ctypes.c_float

# name -> (restype, argtypes); the C declarations are in include/mom25d.h
SIGNATURES = {
    "mom_last_error": (ctypes.c_char_p, []),
    "mom_abi_version": (_i, []),
    "mom_sommerfeld": (_i, [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p, _p]),
    "mom_j0": (_i, [_i, _p, _p, _i, _p]),
    "mom_fill_tier": (_i, [_i, _i, _p, _p]),
    "mom_near_statics": (_i, [_i, _p, _p, _p, _p]),
    "mom_fill_args_size": (_i, []),
    "mom_pair_tab_size": (_i, []),
    "mom_plan_struct_sizes": (_i, [_p]),
    "mom_plan_create": (_i, [_p, _i, _p, _p]),
    "mom_plan_destroy": (_i, [_p]),
    "mom_plan_info": (_i, [_p, _p]),
    "mom_plan_host_array": (_i, [_p, ctypes.c_char_p, _p, _p]),
    "mom_plan_set_pair": (_i, [_p, _i, _i, _i, _p, _p, _d, _d, _d, _d, _d, _i]),
    "mom_plan_fill": (_i, [_p, _p, _p]),
    "mom_plan_fill_host": (_i, [_p, _p]),
    "mom_gram_add": (_i, [_i, _p, _p, _d, _d, _p, _i64, _p]),
    "mom_via_blocks": (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _d, _d, _p, _i64, _p]),
    "mom_symmetrize": (_i, [_i, _p, _i64, _p, _i64, _p, _p]),
    "mom_diag_scale": (_i, [_i, _p, _i64, _p, _p]),
    "mom_getrf_work_bytes": (_i64, [_i, _i, _i]),
    "mom_getrf": (_i, [_i, _i, _p, _i64, _p, _p, _p, _i64, _i, _i, _p]),
    "mom_getrf_nopiv": (_i, [_i, _i, _p, _i64, _p, _p, _p, _i64, _i, _p]),
    "mom_sytrf_nopiv": (_i, [_i, _i, _p, _i64, _p, _p, _p, _i64, _i, _p]),
    "mom_build_perm": (_i, [_i, _p, _p, _p]),
    "mom_set_gemm_tile": (_i, [_i, _i, _i]),
    "mom_set_lu_block": (_i, [_i, _i]),
    "mom_set_spin_limit": (_i, [ctypes.c_uint]),
    "mom_get_lu_block": (_i, [_i]),
    "mom_get_gemm_tile": (_i, [_i, _p, _p]),
    "mom_getrs_work_bytes": (_i64, [_i, _i]),
    "mom_getrs": (_i, [_i, _i, _i, _p, _i64, _p, _p, _p, _p, _p]),
    "mom_getrs_noalloc": (_i, [_i, _i, _i, _p, _i64, _p, _p, _p, _p, _p]),
    "mom_getrs_inv_work_bytes": (_i64, [_i, _i]),
    "mom_trtri_blocks": (_i, [_i, _i, _p, _i64, _p, _p]),
    "mom_getrs_inv": (_i, [_i, _i, _i, _p, _i64, _p, _p, _p, _p, _p]),
    "mom_getrs_inv_z": (_i, [_i, _i, _i, _p, _i64, _p, _p, _p, _p, _p]),
    "mom_scaled_matvec": (_i, [_i, _i, _p, _i64, _p, _p, _p, _p]),
    "mom_scaled_symv_work_bytes": (_i64, [_i, _i]),
    "mom_scaled_symv": (_i, [_i, _i, _p, _i64, _p, _p, _p, _p, _i64, _p]),
    "mom_gmres_work_elems": (_i64, [_i, _i]),
    "mom_gmres_arnoldi": (_i, [_i, _i, _i, _i, _i, _p, _i64, _p, _p, _p, _i64, _p, _p, _i64, _p, _p, _p, _p, _p]),
    "mom_sbsolve_work_bytes": (_i64, [_i, _i, _i]),
    "mom_sbsolve_setup": (_i, [_i, _i, _p, _i64, _i, _p, _i64, _p]),
    "mom_sbsolve_z": (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p]),
    "mom_gmres_arnoldi_sb": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _p, _i64, _p, _p, _i64, _p, _p, _p, _p, _p]),
    "mom_set_solve_rows": (_i, [_i]),
    "mom_get_solve_rows": (_i, []),
    "mom_set_solve_chain": (_i, [_i]),
    "mom_get_solve_chain": (_i, []),
    "mom_interp_anchor": (_i, [_i, _i, _i, _p, _i64, _p, _d, _d, _p, _i64, _p]),
    "mom_interp_combine": (_i, [_i, _i, _i, _i, _p, _p, _i64, _p, _d, _d, _p, _i64, _p]),
    "mom_cgemm_sub": (_i, [_i, _i, _i, _i, _p, _i64, _p, _i64, _p, _i64, _p]),
    "mom_residual": (_i, [_i, _i, _p, _i64, _p, _p, _p, _p, _p, _p]),
    "mom_cvt_z_to_c": (_i, [_i64, _p, _p, _p]),
    "mom_refine_update": (_i, [_i, _i, _p, _p, _p, _p, _p]),
    "mom_timing_begin": (_i, [_i, _i]),
    "mom_timing_end": (_i, [_i, _p]),
    "mom_basis_dot": (_i, [_i, _i, _i, _p, _p, _p, _p]),
    "mom_basis_sub": (_i, [_i, _i, _i, _p, _p, _p, _p]),
    "mom_line_voltage": (_i, [_i, _p, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _d, _p, _p]),
    "mom_graph_begin": (_i, [_p]),
    "mom_graph_end": (_i, [_p, _p]),
    "mom_graph_launch": (_i, [_p, _p]),
    "mom_graph_destroy": (_i, [_p]),
    "mom_mlp_param_count": (_i, []),
    "mom_mlp_partial_stride": (_i, []),
    "mom_mlp_pack": (_i, [_p, _p, _p]),
    "mom_mlp_forward": (_i, [_i, _i, _p, _p, _p, _p, _p]),
    "mom_mlp_train_step": (_i, [_i, _p, _p, _p, _f, _p, _p, _p, _i, _p]),
    "mom_mlp_adam": (_i, [_i, _p, _p, _p, _p, _p, _p, _p, _f, _f, _f, _f, _p]),
}

_lib = None


def lib():
    """The loaded library (loaded once); RuntimeError if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                               "g.build()' or make -C mom25d_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call an entry point; a non-zero status raises RuntimeError with the library's message."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().mom_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t):
    """Device (or host) address of a tensor, or None for None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
