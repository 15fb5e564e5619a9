"""ctypes binding of libsinv_hip.so (the C ABI declared in include/sinv.h).

The library is loaded after torch so that both use the HIP runtime bundled
with torch (the library's run path points there as well).  There is no CPU
fallback: if the library is missing, every entry point raises.
"""

import ctypes
import os

import numpy as np

SINV_OK = 0
SINV_E_SINGULAR = 1
SINV_E_INVAL = 2
SINV_E_HIP = 3
SINV_INFO_OK = (1 << 64) - 1
SINV_INFO_NOT_CONVERGED = 1 << 62

# block inverse methods (include/sinv.h SINV_INV_*)
INV_METHODS = {"pivoted": 0, "newton": 1}

PREC = {"fp64": 0, "mixed": 1, "int8x": 2, "mixed_refined": 3}
OPS = {"N": 0, "T": 1, "C": 2}

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsinv_hip.so")

class C128(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)


_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_i = ctypes.c_int


class RgfDesc(ctypes.Structure):
    _fields_ = [
        ("nblocks", _i), ("blocksize", _i), ("bsizes", ctypes.POINTER(_i)), ("batch", _i),
        ("sym_mat", _i), ("save_off_diag", _i), ("precision", _i),
        ("diag", _vp), ("diag_stride", _ll),
        ("upper", _vp), ("upper_stride", _ll),
        ("lower", _vp), ("lower_stride", _ll),
        ("dscale", C128), ("oscale", C128),
        ("shift", _vp),
        ("g_diag", _vp), ("g_upper", _vp), ("g_lower", _vp),
        ("info", ctypes.POINTER(ctypes.c_ulonglong)),
    ]


class BcrDesc(ctypes.Structure):
    _fields_ = [
        ("nblocks", _i), ("blocksize", _i), ("batch", _i),
        ("diag", _vp), ("diag_stride", _ll),
        ("upper", _vp), ("upper_stride", _ll),
        ("lower", _vp), ("lower_stride", _ll),
        ("dscale", C128), ("oscale", C128),
        ("shift", _vp),
        ("g_diag", _vp), ("g_upper", _vp), ("g_lower", _vp),
        ("g_fill", _vp),
        ("info", ctypes.POINTER(ctypes.c_ulonglong)),
    ]


class PsrDesc(ctypes.Structure):
    _fields_ = [
        ("nblocks", _i), ("blocksize", _i), ("batch", _i), ("npart", _i), ("first_rows", _i),
        ("diag", _vp), ("diag_stride", _ll),
        ("upper", _vp), ("upper_stride", _ll),
        ("lower", _vp), ("lower_stride", _ll),
        ("dscale", C128), ("oscale", C128),
        ("shift", _vp),
        ("g_diag", _vp), ("g_upper", _vp), ("g_lower", _vp),
        ("info", ctypes.POINTER(ctypes.c_ulonglong)),
    ]


class PdivDesc(ctypes.Structure):
    _fields_ = [
        ("nblocks", _i), ("blocksize", _i), ("batch", _i), ("npart", _i),
        ("diag", _vp), ("diag_stride", _ll),
        ("upper", _vp), ("upper_stride", _ll),
        ("lower", _vp), ("lower_stride", _ll),
        ("g_diag", _vp), ("g_upper", _vp), ("g_lower", _vp),
        ("info", ctypes.POINTER(ctypes.c_ulonglong)),
    ]


class AdjDesc(ctypes.Structure):
    _fields_ = [
        ("nblocks", _i), ("blocksize", _i), ("bsizes", ctypes.POINTER(_i)), ("batch", _i),
        ("precision", _i),
        ("checkpoint_every", _i),
        ("diag", _vp), ("diag_stride", _ll),
        ("upper", _vp), ("upper_stride", _ll),
        ("lower", _vp), ("lower_stride", _ll),
        ("dscale", C128), ("oscale", C128),
        ("shift", _vp),
        ("gbar_diag", _vp), ("gbar_diag_stride", _ll),
        ("gbar_upper", _vp), ("gbar_upper_stride", _ll),
        ("gbar_lower", _vp), ("gbar_lower_stride", _ll),
        ("abar_diag", _vp), ("abar_upper", _vp), ("abar_lower", _vp),
        ("g_diag", _vp),
        ("info", ctypes.POINTER(ctypes.c_ulonglong)),
    ]


_lib = None


def lib():
    """Load libsinv_hip.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m sinv_amd.build` (there is no CPU fallback)")
    import torch  # noqa: F401  -- load torch's HIP runtime first
    L = ctypes.CDLL(LIB_PATH)
    L.sinv_version.restype = _i
    L.sinv_last_error.restype = ctypes.c_char_p
    L.sinv_rgf_workspace_bytes.restype = ctypes.c_size_t
    L.sinv_rgf_workspace_bytes.argtypes = [ctypes.POINTER(RgfDesc)]
    L.sinv_rgf_batched.restype = _i
    L.sinv_rgf_batched.argtypes = [ctypes.POINTER(RgfDesc), _vp, ctypes.c_size_t, _vp]
    L.sinv_rgf_batched_noalloc.restype = _i
    L.sinv_rgf_batched_noalloc.argtypes = [ctypes.POINTER(RgfDesc), _vp]
    L.sinv_zinv_workspace_bytes.restype = ctypes.c_size_t
    L.sinv_zinv_workspace_bytes.argtypes = [_i, _i]
    L.sinv_zinv_batched.restype = _i
    L.sinv_zinv_batched.argtypes = [_vp, _i, _i, _ll, _i, _vp, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_ulonglong), _vp]
    L.sinv_zgemm_batched.restype = _i
    L.sinv_zgemm_batched.argtypes = [_i, _i, _i, _i, _i, C128, _vp, _i, _ll, _vp, _i, _ll,
                                     C128, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _i, _vp]
    L.sinv_zgemm_int8x_workspace_bytes.restype = ctypes.c_size_t
    L.sinv_zgemm_int8x_workspace_bytes.argtypes = [_i, _i, _i, _ll, _ll, _i, _i]
    L.sinv_zgemm_int8x_batched.restype = _i
    L.sinv_zgemm_int8x_batched.argtypes = [_i, _i, _i, _i, _i, C128, _vp, _i, _ll, _vp, _i, _ll,
                                           C128, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _i, _vp,
                                           ctypes.c_size_t, _vp]
    L.sinv_int8x_force_variant.restype = _i
    L.sinv_int8x_force_variant.argtypes = [_i]
    L.sinv_prof_begin.restype = _i
    L.sinv_prof_begin.argtypes = [_i]
    L.sinv_prof_end.restype = _i
    L.sinv_prof_end.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_ll),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                _i]
    L.sinv_pdiv_workspace_bytes.restype = ctypes.c_size_t
    L.sinv_pdiv_workspace_bytes.argtypes = [ctypes.POINTER(PdivDesc)]
    L.sinv_pdiv_batched.restype = _i
    L.sinv_pdiv_batched.argtypes = [ctypes.POINTER(PdivDesc), _vp, ctypes.c_size_t, _vp]
    L.sinv_dense_to_blocks.restype = _i
    L.sinv_dense_to_blocks.argtypes = [_vp, _i, _i, _ll, _i, _i, _vp, _vp, _vp, _vp]
    L.sinv_blocks_to_dense.restype = _i
    L.sinv_blocks_to_dense.argtypes = [_vp, _vp, _vp, _i, _i, _i, _vp, _i, _ll, _vp]
    L.sinv_scatter_blocks.restype = _i
    L.sinv_scatter_blocks.argtypes = [_vp, _ll, _i, _vp, _vp, _i, _i, _i, _vp, _i, _ll, _vp]
    L.sinv_gemm_config_count.restype = _i
    L.sinv_gemm_config_count.argtypes = []
    L.sinv_gemm_config_tile.restype = _i
    L.sinv_gemm_config_tile.argtypes = [_i, ctypes.POINTER(_i), ctypes.POINTER(_i)]
    L.sinv_gemm_config_ksplit.restype = _i
    L.sinv_gemm_config_ksplit.argtypes = [_i]
    L.sinv_gemm_force_config.restype = _i
    L.sinv_gemm_force_config.argtypes = [_i]
    L.sinv_leaf_force_variant.restype = _i
    L.sinv_leaf_force_variant.argtypes = [_i]
    L.sinv_leaf_spin_limit.restype = _i
    L.sinv_leaf_spin_limit.argtypes = [_i]
    L.sinv_gj_force_variant.restype = _i
    L.sinv_gj_force_variant.argtypes = [_i]
    L.sinv_gemm_tune_set.restype = _i
    L.sinv_gemm_tune_set.argtypes = [_i, _i, _i, _ll, _i, _i, _i, _i, _i]
    L.sinv_gemm_tune_clear.restype = _i
    L.sinv_gemm_tune_clear.argtypes = []
    L.sinv_gemm_select.restype = _i
    L.sinv_gemm_select.argtypes = [_i, _i, _i, _ll, _i, _i, _i, _i]
    L.sinv_inverse_method.restype = _i
    L.sinv_inverse_method.argtypes = [_i, _i]
    L.sinv_column_dominance.restype = _i
    L.sinv_column_dominance.argtypes = [ctypes.POINTER(RgfDesc), _vp, _vp]
    L.sinv_prof_records.restype = _i
    L.sinv_prof_records.argtypes = [_i, ctypes.POINTER(_i), ctypes.POINTER(_ll),
                                    ctypes.POINTER(ctypes.c_float)]
    if hasattr(L, "sinv_bcr_batched"):
        L.sinv_bcr_workspace_bytes.restype = ctypes.c_size_t
        L.sinv_bcr_workspace_bytes.argtypes = [ctypes.POINTER(BcrDesc)]
        L.sinv_bcr_fill_blocks.restype = _ll
        L.sinv_bcr_fill_blocks.argtypes = [_i]
        L.sinv_bcr_batched.restype = _i
        L.sinv_bcr_batched.argtypes = [ctypes.POINTER(BcrDesc), _vp, ctypes.c_size_t, _vp]
        L.sinv_bcr_parallel_batched.restype = _i
        L.sinv_bcr_parallel_batched.argtypes = [ctypes.POINTER(BcrDesc), _i, _vp, ctypes.c_size_t,
                                                _vp]
    if hasattr(L, "sinv_psr_batched"):
        L.sinv_psr_workspace_bytes.restype = ctypes.c_size_t
        L.sinv_psr_workspace_bytes.argtypes = [ctypes.POINTER(PsrDesc)]
        L.sinv_psr_batched.restype = _i
        L.sinv_psr_batched.argtypes = [ctypes.POINTER(PsrDesc), _vp, ctypes.c_size_t, _vp]
    if hasattr(L, "sinv_rgf_adjoint_batched"):
        L.sinv_rgf_adjoint_workspace_bytes.restype = ctypes.c_size_t
        L.sinv_rgf_adjoint_workspace_bytes.argtypes = [ctypes.POINTER(AdjDesc)]
        L.sinv_rgf_adjoint_batched.restype = _i
        L.sinv_rgf_adjoint_batched.argtypes = [ctypes.POINTER(AdjDesc), _vp, ctypes.c_size_t, _vp]
    _lib = L
    return L


def check(rc, what):
    """Map a C return code to the reference's exception types."""
    if rc == SINV_OK:
        return
    msg = lib().sinv_last_error().decode(errors="replace")
    if rc == SINV_E_INVAL:
        raise ValueError(f"{what}: {msg}")
    if rc == SINV_E_SINGULAR:
        raise np.linalg.LinAlgError("Singular matrix")
    raise RuntimeError(f"{what}: HIP error: {msg}")


class NotConverged(RuntimeError):
    """A Newton-Schulz block inverse did not reach its acceptance bound (the
    call's outputs are undefined; repeat it with the pivoted elimination)."""


def check_info(info_tensor, what="inverse"):
    """Raise numpy's LinAlgError (as numpy.linalg.inv does) if a pivot was zero,
    or NotConverged if a Newton-Schulz inverse missed its bound.

    Synchronises the stream of ``info_tensor``."""
    v = int(info_tensor.cpu().numpy().view(np.uint64)[0])
    if v != SINV_INFO_OK:
        if v & SINV_INFO_NOT_CONVERGED:
            v &= ~SINV_INFO_NOT_CONVERGED
            err = NotConverged(f"{what}: Newton-Schulz inverse of block {v >> 32}, instance "
                               f"{v & 0xFFFFFFFF} did not converge")
        else:
            err = np.linalg.LinAlgError("Singular matrix")
        err.block, err.instance = v >> 32, v & 0xFFFFFFFF
        raise err


class inverse_method:
    """Context manager selecting the block-inverse method of library calls made
    on this thread: "pivoted" (Gauss-Jordan with partial pivoting, numpy's
    semantics) or "newton" (Newton-Schulz iteration for launches of at most 4
    blocks above 64, for column diagonally dominant blocks; see
    include/sinv.h).  Workspace queries made inside the block size for it."""

    def __init__(self, method, max_iter=0):
        if method not in INV_METHODS:
            raise ValueError(f"inverse method must be one of {sorted(INV_METHODS)}")
        self.method, self.max_iter = INV_METHODS[method], int(max_iter)

    def __enter__(self):
        prev = lib().sinv_inverse_method(self.method, self.max_iter)
        if prev < 0:
            raise ValueError(lib().sinv_last_error().decode(errors="replace"))
        self.prev = prev
        return self

    def __exit__(self, *exc):
        # a limit given here is put back to the library's default (6)
        lib().sinv_inverse_method(self.prev, 6 if self.max_iter else 0)
        return False


PROF_KINDS = ("gemm_f64", "gemm_f32", "inv_leaf", "other", "gemm_i8x", "inv_panel", "inv_newton",
              "i8x_split")

# leaf inverse kernels (include/sinv.h SINV_LEAF_*)
LEAF_VARIANTS = {"auto": 1, "lds": 0, "rows8": 2, "panel8": 3, "panel16": 4, "panel16f": 5,
                 "panel8f": 6}


def force_leaf_variant(name):
    """Select the leaf inverse kernel for later launches ("auto" restores the
    default choice)."""
    check(lib().sinv_leaf_force_variant(LEAF_VARIANTS[name]), "leaf_force_variant")


class LaunchProfiler:
    """Context manager around libsinv_hip's launch profiler.

    ``stats`` maps kernel kind -> (total ms, launches, algorithmic flops);
    ``bytes`` maps kernel kind -> algorithmic bytes."""

    def __init__(self, max_launches=200000):
        self.max_launches = max_launches
        self.stats = None

    def __enter__(self):
        check(lib().sinv_prof_begin(self.max_launches), "prof_begin")
        return self

    def __exit__(self, *exc):
        n = len(PROF_KINDS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        fl = (ctypes.c_double * n)()
        by = (ctypes.c_double * n)()
        check(lib().sinv_prof_end(ms, cnt, fl, by, n), "prof_end")
        self.stats = {k: (ms[i], cnt[i], fl[i]) for i, k in enumerate(PROF_KINDS)}
        self.bytes = {k: by[i] for i, k in enumerate(PROF_KINDS)}
        return False

    def records(self):
        """Per-launch (kind, (M, N, K, batch, flags), ms) of the profiled region."""
        L = lib()
        n = L.sinv_prof_records(0, None, None, None)
        kind = (ctypes.c_int * max(n, 1))()
        dims = (ctypes.c_longlong * max(5 * n, 1))()
        ms = (ctypes.c_float * max(n, 1))()
        L.sinv_prof_records(n, kind, dims, ms)
        return [(PROF_KINDS[kind[i]] if 0 <= kind[i] < len(PROF_KINDS) else "other",
                 tuple(dims[5 * i:5 * i + 5]), ms[i]) for i in range(n)]
