"""ctypes binding of libemerge_amd.so (the C ABI declared in include/emerge_amd.h).

The library is required: every GPU entry point raises if it is missing, there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemerge_amd.so")

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32
F64 = ctypes.c_double
F32 = ctypes.c_float

# name -> (restype, argtypes).  Keep in step with include/emerge_amd.h.
SIGNATURES: dict[str, tuple] = {
    "emg_last_error": (ctypes.c_char_p, []),
    "emg_pattern_build": (ctypes.c_int, [P, I64, I64, P, ctypes.POINTER(P)]),
    "emg_pattern_info": (ctypes.c_int, [P, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64)]),
    "emg_pattern_csr": (ctypes.c_int, [P, P, P, P]),
    "emg_pattern_destroy": (None, [P]),
    "emg_assemble_iso": (ctypes.c_int, [P, P, P, P, P, ctypes.c_int, P, P, P]),
    "emg_assemble_tensor": (ctypes.c_int, [P, P, P, P, P, ctypes.c_int, P, P, P]),
    "emg_solver_create": (ctypes.c_int, [I64, I64, P, P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]),
    "emg_solver_set_matrix": (ctypes.c_int, [P, P, P, P]),
    "emg_solver_set_gradient": (ctypes.c_int, [P, I64, P, P, P, P, P, P, P]),
    "emg_solver_solve": (ctypes.c_int, [P, P, P, F64, ctypes.c_int, ctypes.c_int, P, P]),
    "emg_solver_destroy": (None, [P]),
    "emg_solver_apply": (ctypes.c_int, [P, ctypes.c_int, P, P, P]),
    "emg_affine_combine": (ctypes.c_int, [I64, P, P, F64, P, P]),
    "emg_csr_find": (ctypes.c_int, [I64, P, P, P, P, P, P]),
    "emg_scatter_add": (ctypes.c_int, [I64, P, P, F64, F64, P, P]),
    "emg_spmv": (ctypes.c_int, [I64, P, P, P, P, P, P, P]),
    "emg_spmv_blocked": (ctypes.c_int, [I64, I64, P, P, P, P, P, P, P]),
    "emg_mdotc": (ctypes.c_int, [I64, ctypes.c_int, P, P, P, P, P]),
    "emg_maxpy": (ctypes.c_int, [I64, ctypes.c_int, P, P, P, P, P, P]),
    "emg_mcomb": (ctypes.c_int, [I64, ctypes.c_int, P, P, P, P]),
    "emg_robin_blocks": (ctypes.c_int, [I64, P, P, P, ctypes.c_int, P, P, P, P]),
    "emg_interp_points": (ctypes.c_int, [I64, P, P, P, P, P, I64, ctypes.c_int, P, P, P]),
    "emg_interp_curl_points": (ctypes.c_int, [I64, P, P, P, P, P, I64, ctypes.c_int, P, P, P, P]),
    "emg_farfield": (ctypes.c_int, [I64, P, P, P, P, I64, P, P, F64, P, P, P]),
    "emg_locate_points": (ctypes.c_int, [I64, P, P, P, P, P, P, P, P, P]),
    "emg_solver_spmv_tuning": (ctypes.c_int, [P, P, P]),
    "emg_solver_smoother_tuning": (ctypes.c_int, [P, P, P]),
    "emg_pattern_assembly_tuning": (ctypes.c_int, [P, P, P]),
    "emg_assemble_iso_batched": (ctypes.c_int, [P, ctypes.c_int, I64, P, P, P, P, ctypes.c_int, P, P, P]),
    "emg_topology_build": (ctypes.c_int, [I64, I64, P, P, P, P]),
    "emg_topology_sizes": (ctypes.c_int, [P, P, P]),
    "emg_topology_get": (ctypes.c_int, [P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "emg_topology_destroy": (None, [P]),
    "emg_port_assemble": (ctypes.c_int, [I64, P, P, P, F64, ctypes.c_int, P, I64, P, P, P, P, P]),
    "emg_port_field": (ctypes.c_int, [I64, P, I64, P, I64, P, P, I64, P, F64, P, P, F64, F64, P, P, P]),
    "emg_port_factor": (ctypes.c_int, [I64, P, P, F64, F64, P, P, P]),
    "emg_port_inverse_step": (ctypes.c_int, [I64, P, P, P, I64, P, P, P]),
    "emg_port_operator": (ctypes.c_int, [I64, P, P, P, P, P]),
    "emg_port_krylov": (ctypes.c_int, [I64, P, ctypes.c_int, P, P, P]),
    "emg_zgemm": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, I64, I64, I64, F64, F64, P, I64, P, I64, F64, F64, P, I64,
                                 P]),
    "emg_solver_set_weights": (ctypes.c_int, [P, P, P]),
    "emg_solver_smoother_parts": (ctypes.c_int, [P, P, P, P, P]),
    "emg_solver_set_poly": (ctypes.c_int, [P, ctypes.c_int, P]),
    "emg_solver_arnoldi_probe": (ctypes.c_int, [P, ctypes.c_int, P, P, P]),
    "emg_csr_real_apply": (ctypes.c_int, [I64, P, P, P, P, P, ctypes.c_int, P]),
    "emg_gnn_gemm": (ctypes.c_int, [I64, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_float, P, P]),
    "emg_gnn_gemm_nt": (ctypes.c_int, [I64, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_float, P, P]),
    "emg_gnn_gemm_tn": (ctypes.c_int, [I64, ctypes.c_int, ctypes.c_int, P, P, P, P, P]),
    "emg_gnn_agg": (ctypes.c_int, [I64, ctypes.c_int, P, P, P, P, P, P, P]),
    "emg_gnn_relu": (ctypes.c_int, [I64, P, P, P, P, P, P]),
    "emg_gnn_readout": (ctypes.c_int, [I64, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, P, P, P, P, P, P]),
    "emg_gnn_weights": (ctypes.c_int, [I64, ctypes.c_int, P, P, P, P, P]),
    "emg_gnn_weight_grad": (ctypes.c_int, [I64, ctypes.c_int, I64, I64, P, P, P, P, P]),
    "emg_sweep_residual": (ctypes.c_int, [I64, P, P, P, P, P, ctypes.c_int, P, P, P, P, P, P, P, P, P, P, P]),
    "emg_csr_permute": (ctypes.c_int, [I64, P, P, P, P, P, P, P, P]),
    "emg_csr_map": (ctypes.c_int, [I64, P, P, P, P, P, P, P]),
    "emg_gather_c128": (ctypes.c_int, [I64, P, I64, P, P, P]),
    "emg_scatter_c128": (ctypes.c_int, [I64, P, I64, P, P, P, P]),
    "emg_sweep_residual_noalloc": (ctypes.c_int, [I64, P, P, P, P, P, ctypes.c_int, P, P, P, P, P, P, P, P, P, P]),
    "emg_port_apply": (ctypes.c_int, [I64, ctypes.c_int, P, P, P, P, P, P, P, P]),
    "emg_basis_combine": (ctypes.c_int, [I64, ctypes.c_int, P, ctypes.c_int, P, P, P]),
}


class SolveInfo(ctypes.Structure):
    """emg_solve_info (include/emerge_amd.h)."""
    _fields_ = [("iterations", ctypes.c_int), ("refinements", ctypes.c_int), ("converged", ctypes.c_int),
                ("reserved", ctypes.c_int), ("residual", ctypes.c_double)]

_lib = None


class EmgError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load libemerge_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EmgError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python emerge_amd/build.py`; there is no CPU fallback")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().emg_last_error().decode(errors="replace")
        raise EmgError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> ctypes.c_void_p:
    """Device (or host) pointer of a torch tensor / numpy array.

    Lazily conjugated or negated torch views and non-contiguous tensors are
    refused: their memory does not hold the values the view shows in order."""
    if hasattr(t, "data_ptr"):
        if (hasattr(t, "is_conj") and t.is_conj()) or (hasattr(t, "is_neg") and t.is_neg()):
            raise EmgError("pointer of a lazily conjugated/negated tensor view; call resolve_conj() first")
        if not t.is_contiguous():
            raise EmgError("pointer of a non-contiguous tensor; call contiguous() first")
        return ctypes.c_void_p(t.data_ptr())
    return t.ctypes.data_as(ctypes.c_void_p)


def stream_ptr(stream=None) -> ctypes.c_void_p:
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)
