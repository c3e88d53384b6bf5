"""ctypes loader for libbksp_amd.so (the HIP kernels, C-ABI and C host layer).

The library is built in-tree by ``make -C bksp_amd/csrc``. Importing this module
raises if the shared object is missing: there is no Python or CPU fallback for
any solver or kernel.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbksp_amd.so")

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_ip = ctypes.POINTER(ctypes.c_int)
c_dp = ctypes.POINTER(ctypes.c_double)
c_u64 = ctypes.c_uint64


class DSpMat(ctypes.Structure):
    """The reference's ``dspmat`` (include/bksp.h:12-21): CSR with int* size fields."""
    _fields_ = [("row", c_ip), ("col", c_ip), ("nnz", c_ip), ("row_size", c_ip),
                ("col_size", c_ip), ("value", c_dp), ("format", ctypes.c_char_p),
                ("type", ctypes.c_char_p)]


class DMat(ctypes.Structure):
    """The reference's ``dmat`` (include/bksp.h:23-27): column-major dense block."""
    _fields_ = [("row_size", c_int), ("col_size", c_int), ("value", c_dp)]


class ZSpMat(ctypes.Structure):
    """The reference's ``zspmat`` (include/bksp.h:113-122): CSR with double _Complex values."""
    _fields_ = [("row", c_ip), ("col", c_ip), ("nnz", c_ip), ("row_size", c_ip),
                ("col_size", c_ip), ("value", c_vp), ("format", ctypes.c_char_p),
                ("type", ctypes.c_char_p)]


class ZMat(ctypes.Structure):
    """The reference's ``zmat`` (include/bksp.h:124-128): column-major complex block."""
    _fields_ = [("row_size", c_int), ("col_size", c_int), ("value", c_vp)]


class BkhOpts(ctypes.Structure):
    _fields_ = [("precision", c_int), ("graph_iters", c_int), ("fix_blcg_p0", c_int),
                ("max_inner", c_int), ("inner_tol", c_dbl), ("path", c_int)]


class BkhMatrixTuning(ctypes.Structure):
    _fields_ = [("spmv_variant", c_int), ("rows_per_wave", c_int), ("nnz_per_wave", c_int),
                ("candidates", c_int), ("spmv_ms", c_dbl), ("spmv_ms_default", c_dbl),
                ("spmm_width", c_int * 17), ("spmm_ms", c_dbl * 17), ("spmm_ms_full", c_dbl * 17)]


class BkhInfo(ctypes.Structure):
    _fields_ = [("ret", c_int), ("status", c_int), ("iters", c_int), ("inner_iters", c_int),
                ("passes", c_int), ("rec_res", c_dbl), ("true_res", c_dbl), ("time_ms", c_dbl)]


class CsrHost(ctypes.Structure):
    _fields_ = [("n", c_int), ("nnz", c_int), ("row_ptr", c_ip), ("col", c_ip), ("val", c_dp)]


_SIGS = {
    # runtime / memory
    "bkh_init": (c_int, []),
    "bkh_last_error": (ctypes.c_char_p, []),
    "bkh_device_name": (c_int, [ctypes.c_char_p, c_int]),
    "bkh_sync": (c_int, []),
    "bkh_malloc": (c_vp, [ctypes.c_size_t]),
    "bkh_free": (None, [c_vp]),
    "bkh_h2d": (c_int, [c_vp, c_vp, ctypes.c_size_t]),
    "bkh_d2h": (c_int, [c_vp, c_vp, ctypes.c_size_t]),
    "bkh_d2d": (c_int, [c_vp, c_vp, ctypes.c_size_t]),
    "bkh_memset": (c_int, [c_vp, c_int, ctypes.c_size_t]),
    # matrices and kernels
    "bkh_matrix_create": (c_vp, [c_int, c_vp, c_vp, c_vp, c_int, c_int, c_int]),
    "bkh_matrix_destroy": (None, [c_vp]),
    "bkh_matrix_nnz": (c_ll, [c_vp]),
    "bkh_matrix_dia_entries": (c_ll, [c_vp]),
    "bkh_matrix_rows": (c_int, [c_vp]),
    "bkh_matrix_blocks": (c_int, [c_vp]),
    "bkh_matrix_set_spmv": (c_int, [c_vp, c_int]),
    "bkh_matrix_tune": (c_int, [c_vp, c_int, c_vp]),
    "bkh_spmv": (c_int, [c_vp, ctypes.c_char, c_dbl, c_vp, c_dbl, c_vp]),
    "bkh_spmm": (c_int, [c_vp, c_int, c_dbl, c_vp, c_dbl, c_vp]),
    "bkh_spmm_gram": (c_int, [c_vp, c_int, c_vp, c_vp, c_dp, ctypes.POINTER(c_int)]),
    "bkh_zmatrix_create": (c_vp, [c_int, c_vp, c_vp, c_vp]),
    "bkh_zmatrix_destroy": (None, [c_vp]),
    "bkh_zmatrix_dia_entries": (c_ll, [c_vp]),
    "bkh_zmatrix_split_rows": (c_ll, [c_vp]),
    "bkh_zspmv": (c_int, [c_vp, ctypes.c_char, c_vp, c_vp]),
    "bkh_zspmm1": (c_int, [c_vp, ctypes.c_char, c_vp, c_int, c_vp, c_vp, c_vp]),
    "bkh_zbicg": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_int, c_vp]),
    "bkh_zbicgstabl": (c_int, [c_vp, c_vp, c_int, c_vp, c_dbl, c_int, c_vp]),
    "bkh_zshbicgstabl": (c_int, [c_vp, c_vp, c_int, c_dp, c_int, c_vp, c_dbl, c_int, c_vp, c_dp]),
    "bkh_zspmv_bench": (c_int, [c_vp, ctypes.c_char, c_int, c_dp]),
    "bkh_zblbicgstabl": (c_int, [c_vp, c_int, c_vp, c_int, c_vp, c_dbl, c_int, c_vp]),
    "bkh_zqr": (c_int, [c_int, c_int, c_vp, c_vp]),
    "bkh_spmv32": (c_int, [c_vp, c_vp, c_vp]),
    "bkh_dot": (c_int, [c_ll, c_vp, c_vp, c_dp]),
    "bkh_qr": (c_int, [c_int, c_int, c_vp, c_dp]),
    "bkh_sp_norm_f": (c_int, [c_int, c_vp, c_vp, c_vp, c_int, c_dp]),
    "bkh_last_solve_ms": (c_dbl, []),
    "bkh_bench_spmv": (c_int, [c_vp, c_int, c_vp, c_vp, c_int, c_dp]),
    "bkh_bench_copy": (c_int, [c_vp, c_vp, ctypes.c_size_t, c_int, c_dp]),
    # solvers (device vectors)
    "bkh_cg_noalloc": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_int, ctypes.POINTER(BkhInfo)]),
    "bkh_cg_noalloc_max_rows": (c_ll, []),
    "bkh_cg": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts),
                       ctypes.POINTER(BkhInfo)]),
    "bkh_cr": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts),
                       ctypes.POINTER(BkhInfo)]),
    "bkh_bicg": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts),
                         ctypes.POINTER(BkhInfo)]),
    "bkh_bicgstabl": (c_int, [c_vp, c_vp, c_int, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts),
                              ctypes.POINTER(BkhInfo)]),
    "bkh_blcg": (c_int, [c_vp, c_int, c_vp, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts),
                         ctypes.POINTER(BkhInfo)]),
    "bkh_blbicgstabl": (c_int, [c_vp, c_int, c_vp, c_int, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts),
                                ctypes.POINTER(BkhInfo)]),
    "bkh_shbicgstabl": (c_int, [c_vp, c_vp, c_int, c_dp, c_int, c_vp, c_dbl, c_int,
                                ctypes.POINTER(BkhOpts), ctypes.POINTER(BkhInfo), c_dp]),
    # batches of independent systems
    "bkh_batch_create": (c_vp, [c_int, c_ip, ctypes.POINTER(c_ip), ctypes.POINTER(c_ip),
                                ctypes.POINTER(c_dp), c_int]),
    "bkh_batch_destroy": (None, [c_vp]),
    "bkh_batch_rows": (c_ll, [c_vp]),
    "bkh_batch_nnz": (c_ll, [c_vp]),
    "bkh_batch_offset": (c_ll, [c_vp, c_int]),
    "bkh_batch_layout": (None, [c_vp, c_ip, c_ip, c_ip]),
    "bkh_batch_tune": (c_int, [c_vp, c_int, c_dp]),
    "bkh_batch_set_spmv_kernel": (c_int, [c_vp, c_int]),
    "bkh_batch_spmv_kernel": (c_int, [c_vp]),
    "bkh_batch_spmv": (c_int, [c_vp, c_int, c_vp, c_vp]),
    "bkh_batch_last_path": (c_int, []),
    "bkh_debug_fail_alloc": (None, [c_int]),
    "bkh_debug_fail_alloc_pending": (c_int, []),
    "bkh_debug_fin_lds_form": (None, [c_int]),
    "bkh_batch_bicgstabl": (c_int, [c_vp, c_vp, c_int, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts), c_ip,
                                    c_dp, ctypes.POINTER(BkhInfo)]),
    "bkh_matrix_spmv_bytes": (c_ll, [c_vp, c_int]),
    "bkh_bench_cg_kernel": (c_int, [c_vp, c_int, c_int, c_dp]),
    # learned FSAI and PCG
    "bkh_fsai_nparam": (c_int, []),
    "bkh_fsai_model_create": (c_vp, [ctypes.POINTER(ctypes.c_float)]),
    "bkh_fsai_model_destroy": (None, [c_vp]),
    "bkh_fsai_model_get": (c_int, [c_vp, ctypes.POINTER(ctypes.c_float)]),
    "bkh_fsai_model_set": (c_int, [c_vp, ctypes.POINTER(ctypes.c_float)]),
    "bkh_fsai_field_create": (c_vp, [c_int, c_vp, c_vp, c_vp, c_int, c_vp]),
    "bkh_fsai_field_destroy": (None, [c_vp]),
    "bkh_fsai_field_matrix": (c_vp, [c_vp]),
    "bkh_fsai_factor": (c_int, [c_vp, c_vp, c_vp]),
    "bkh_fsai_loss_grad": (c_int, [c_vp, c_vp, c_dp, c_vp]),
    "bkh_fsai_train_step": (c_int, [c_vp, c_vp, c_dbl, c_dp]),
    "bkh_fsai_bench_forward": (c_int, [c_vp, c_vp, c_int, c_dp]),
    "bkh_fsai_pcg": (c_int, [c_vp, c_vp, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts), ctypes.POINTER(BkhInfo)]),
    "bkh_pcg": (c_int, [c_vp, c_int, c_vp, c_vp, c_dbl, c_int, ctypes.POINTER(BkhOpts), ctypes.POINTER(BkhInfo)]),
    # bksp.h drop-in (host arrays)
    "bksp_dcg": (c_int, [ctypes.POINTER(DSpMat), c_dp, c_dbl, c_int, c_dp]),
    "dcr": (c_int, [ctypes.POINTER(DSpMat), c_dp, c_dbl, c_int, c_dp]),
    "dbicg": (c_int, [ctypes.POINTER(DSpMat), c_dp, c_dbl, c_int, c_dp]),
    "dbicgstabl": (c_int, [ctypes.POINTER(DSpMat), c_dp, c_int, c_dbl, c_int, c_dp]),
    "dblcg": (c_int, [ctypes.POINTER(DSpMat), ctypes.POINTER(DMat), c_dbl, c_int, ctypes.POINTER(DMat)]),
    "dblbicgstabl": (c_int, [ctypes.POINTER(DSpMat), ctypes.POINTER(DMat), c_int, c_dbl, c_int,
                             ctypes.POINTER(DMat)]),
    "dbicgstab": (c_int, [ctypes.POINTER(DSpMat), c_dp, c_dbl, c_int, c_dp]),
    "dblbicgstab": (c_int, [ctypes.POINTER(DSpMat), ctypes.POINTER(DMat), c_dbl, c_int, ctypes.POINTER(DMat)]),
    "dshbicg": (c_int, [ctypes.POINTER(DSpMat), c_dp, c_dp, c_int, c_dbl, c_int, c_dp]),
    "dshbicgstabl": (c_int, [ctypes.POINTER(DSpMat), c_dp, c_int, c_dp, c_int, c_dbl, c_int, c_dp]),
    "zbicg": (c_int, [ctypes.POINTER(ZSpMat), c_vp, c_dbl, c_int, c_vp]),
    "zbicgstabl": (c_int, [ctypes.POINTER(ZSpMat), c_vp, c_int, c_dbl, c_int, c_vp]),
    "zshbicgstabl": (c_int, [ctypes.POINTER(ZSpMat), c_vp, c_int, c_vp, c_int, c_dbl, c_int, c_vp]),
    "zblbicgstabl": (c_int, [ctypes.POINTER(ZSpMat), ctypes.POINTER(ZMat), c_int, c_dbl, c_int,
                             ctypes.POINTER(ZMat)]),
    "read_matrix_market": (ctypes.POINTER(DSpMat), [ctypes.c_char_p]),
    "myqr": (c_int, [c_int, c_int, c_dp, c_dp]),
    "sp_norm_f": (c_dbl, [ctypes.POINTER(DSpMat)]),
    "dcsrgemm0": (c_int, [ctypes.c_char, c_int, c_int, c_dp, c_ip, c_ip, c_int, c_dp, c_dp]),
    "dcsrsymm0": (c_int, [ctypes.c_char, c_int, c_int, c_dp, c_ip, c_ip, c_int, c_dp, c_dp]),
    "dcsrgemm1": (c_int, [ctypes.c_char, c_dbl, c_int, c_int, c_dp, c_ip, c_ip, c_int, c_dp, c_dbl, c_dp]),
    "dcsrsymm1": (c_int, [ctypes.c_char, c_dbl, c_int, c_int, c_dp, c_ip, c_ip, c_int, c_dp, c_dbl, c_dp]),
    "dhorner": (c_int, [c_dp, c_int, c_dbl, c_dp]),
    "comb": (c_int, [c_int, c_int]),
    "dmat_create": (ctypes.POINTER(DMat), [c_int, c_int]),
    "dmat_free": (c_int, [ctypes.POINTER(DMat)]),
    "darray2dmat": (c_int, [c_dp, c_int, c_int, ctypes.POINTER(DMat)]),
    "dvec2mat": (c_int, [c_dp, c_int, ctypes.POINTER(DMat)]),
    # generators
    "bkh_csr_host_free": (None, [ctypes.POINTER(CsrHost)]),
    "bkh_gen_uniform": (None, [c_u64, c_ll, c_dbl, c_dbl, c_dp]),
    "bkh_gen_normal": (None, [c_u64, c_ll, c_dp]),
    "bkh_gen_convdiff2d": (c_int, [c_int, c_dbl, c_dbl, ctypes.POINTER(CsrHost)]),
    "bkh_gen_poisson3d": (c_int, [c_int, c_int, ctypes.POINTER(CsrHost)]),
    "bkh_gen_convdiff3d": (c_int, [c_int, c_dbl, c_dbl, c_dbl, ctypes.POINTER(CsrHost)]),
    "bkh_gen_random_rows": (c_int, [c_u64, c_int, ctypes.POINTER(CsrHost)]),
    "bkh_gen_batch_sizes": (None, [c_u64, c_int, c_dbl, c_dbl, c_ip]),
    "bkh_gen_grf": (c_int, [c_u64, c_int, c_dbl, c_dbl, c_dp]),
    "bkh_gen_fv_diffusion": (c_int, [c_int, c_dp, c_int, ctypes.POINTER(CsrHost)]),
}

_lib = None


def load():
    """Loads libbksp_amd.so and declares every exported prototype."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C bksp_amd/csrc` "
            "(bksp_amd has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names this module declares (used by the export test)."""
    return list(_SIGS)
