"""ctypes binding of ``libtnx.so`` (the C-ABI declared in ``include/tnx.h``).

Every wrapper takes torch tensors that live on the GPU, checks dtype,
contiguity and device on the host, passes raw device pointers and the
current HIP stream, and raises ``RuntimeError`` with the library's message
when a call returns a nonzero status. There is no CPU fallback: without the
built library or without a GPU these functions raise.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np
import torch

LIB_PATH = Path(__file__).resolve().parent / "libtnx.so"
# GEMM tile table chosen by the autotuner on MI355X for the shapes of the
# DMRG path; loaded with the library so no GEMM is timed inline
TILE_TABLE = Path(__file__).resolve().parent / "gemm_tiles_mi355x.txt"

TNX_F64, TNX_F32, TNX_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": TNX_F64, "fp32": TNX_F32, "bf16x3": TNX_BF16X3}

TNX_OK = 0
TNX_ERR_NOCONV = -3

_sz = C.c_size_t
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_p = C.c_void_p


class EigshOpts(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("max_basis", C.c_int),
        ("max_matvecs", C.c_int),
        ("precision", C.c_int),
        ("keep", C.c_int),
        ("steps_per_check", C.c_int),
    ]


class EigshInfo(C.Structure):
    _fields_ = [
        ("eval", C.c_double),
        ("resnorm", C.c_double),
        ("anorm", C.c_double),
        ("matvecs", C.c_int),
        ("matvecs_f64", C.c_int),
        ("restarts", C.c_int),
        ("converged", C.c_int),
    ]


# name -> (restype, argtypes); mirrors include/tnx.h
SIGNATURES = {
    "tnx_last_error": (C.c_char_p, []),
    "tnx_version": (_i, []),
    "tnx_set_autotune": (_i, [_i]),
    "tnx_tune_clear": (_i, []),
    "tnx_tune_load": (_i, [C.c_char_p]),
    "tnx_tune_save": (_i, [C.c_char_p]),
    "tnx_gemm_f64": (_i, [_i, _i, _i, _i, _d, _p, _i64, _i64, _i, _p, _i64, _i64, _i, _d, _p, _i64, _i64, _i, _p]),
    "tnx_gemm": (_i, [_i, _i, _i, _i, _d, _p, _i, _i64, _i64, _i, _p, _i, _i64, _i64, _i, _d, _p, _i, _i64, _i64, _i,
                      _p]),
    "tnx_heff1_matvec_ex": (_i, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_to_f32": (_i, [_p, _p, _i64, _p]),
    "tnx_heff1_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "tnx_heff1_matvec": (_i, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_noalloc_scratch_bytes": (_sz, []),
    "tnx_heff1_matvec_noalloc": (_i, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p]),
    "tnx_heff1_dense": (_i, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _p]),
    "tnx_env_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "tnx_env_left": (_i, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_env_right": (_i, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_syev_lowest_workspace_bytes": (_sz, [_i]),
    "tnx_syev_lowest": (_i, [_p, _i, _p, C.POINTER(_d), _p, _sz, _p]),
    "tnx_eigsh_heff1_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i]),
    "tnx_eigsh_heff1": (
        _i,
        [_p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, C.POINTER(EigshOpts), C.POINTER(EigshInfo), _p, _sz, _p],
    ),
    "tnx_heff1_batched_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i]),
    "tnx_heff1_matvec_batched": (_i, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_heff1_dense_batched": (_i, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p]),
    "tnx_env_batched_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i]),
    "tnx_env_left_batched": (_i, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_env_right_batched": (_i, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_syev_lowest_batched_workspace_bytes": (_sz, [_i, _i]),
    "tnx_syev_lowest_batched": (_i, [_p, _i, _i, _p, C.POINTER(_d), _p, _sz, _p]),
    "tnx_eigsh_heff1_batched_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i, _i]),
    "tnx_eigsh_heff1_batched": (
        _i,
        [_p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, C.POINTER(EigshOpts), C.POINTER(EigshInfo), _p, _sz, _p],
    ),
    "tnx_gauge_batched_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "tnx_gauge_right_batched": (_i, [_p, _p, _p, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_gauge_left_batched": (_i, [_p, _p, _p, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_heff2_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i]),
    "tnx_heff2_matvec": (_i, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_eigsh_heff2_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i, _i]),
    "tnx_eigsh_heff2": (
        _i,
        [_p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, C.POINTER(EigshOpts), C.POINTER(EigshInfo), _p, _sz, _p],
    ),
    "tnx_gauge_workspace_bytes": (_sz, [_i, _i, _i]),
    "tnx_gauge_right": (_i, [_p, _p, _p, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_gauge_left": (_i, [_p, _p, _p, _i, _i, _i, _i, _p, _sz, _p]),
    "tnx_svd_trunc_workspace_bytes": (_sz, [_i, _i, _i]),
    "tnx_svd_trunc_f64": (_i, [_p, _i, _i, _i, _p, _p, C.POINTER(_d), C.POINTER(_i), _p, _sz, _p]),
    "tnx_syev_workspace_bytes": (_sz, [_i]),
    "tnx_syev_f64": (_i, [_p, _i, C.POINTER(_d), _p, _p, _sz, _p]),
    "tnx_mpo_absorb": (_i, [_p, _i, _i64, _p, _i, _i, _p, _p]),
    "tnx_sygv_lowest_workspace_bytes": (_sz, [_i]),
    "tnx_sygv_lowest": (_i, [_p, _p, _i, _p, C.POINTER(_d), _p, _sz, _p]),
    "tnx_eigsh_gen_heff1_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i, _i, _i]),
    "tnx_eigsh_gen_heff1": (
        _i,
        [_p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, C.POINTER(EigshOpts), C.POINTER(EigshInfo),
         _p, _sz, _p],
    ),
    "tnx_mpo_apply_site": (_i, [_p, _p, _i, _i, _i, _i, _i, _p, _p]),
    "tnx_qr_workspace_bytes": (_sz, [_i, _i]),
    "tnx_qr": (_i, [_p, _i, _i, _p, _p, _sz, _p]),
    "tnx_axpy": (_i, [_i64, _d, _p, _p, _p]),
    "tnx_scal": (_i, [_i64, _d, _p, _p]),
    "tnx_dot_workspace_bytes": (_sz, [_i64]),
    "tnx_dot": (_i, [_i64, _p, _p, C.POINTER(_d), _p, _sz, _p]),
}

_LIB: Optional[C.CDLL] = None


def load_library(path: Optional[os.PathLike] = None) -> C.CDLL:
    """Load libtnx.so and bind every exported entry point. Raises if the
    library has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -m tnpy_amd.build)"
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
        if TILE_TABLE.exists():
            lib.tnx_tune_load(str(TILE_TABLE).encode())
    return lib


def lib() -> C.CDLL:
    return load_library()


def check(status: int, what: str) -> int:
    if status != TNX_OK:
        msg = lib().tnx_last_error()
        raise RuntimeError(f"{what} failed (status {status}): {msg.decode() if msg else ''}")
    return status


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("tnpy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(device: Optional[torch.device] = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor, name: str = "tensor") -> int:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if t.dtype != torch.float64:
        raise ValueError(f"{name} must be float64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def svd_trunc(X: torch.Tensor, k: int, workspace: Optional["Workspace"] = None):
    """Top-k left singular vectors U (a x k), U^T X (k x b) and the k leading
    singular values (numpy, descending) of the device matrix X (a x b), by
    tnx_svd_trunc_f64."""
    if X.dim() != 2:
        raise ValueError("svd_trunc needs a matrix")
    a, b = int(X.shape[0]), int(X.shape[1])
    k = int(k)
    X = X.contiguous()
    ws = workspace if workspace is not None else Workspace(X.device)
    U = torch.empty(a, k, dtype=torch.float64, device=X.device)
    Y = torch.empty(k, b, dtype=torch.float64, device=X.device)
    sv = (C.c_double * k)()
    sweeps = C.c_int(0)
    wp, wb = ws.get(lib().tnx_svd_trunc_workspace_bytes(a, b, k))
    check(lib().tnx_svd_trunc_f64(ptr(X, "X"), a, b, k, ptr(U, "U"), ptr(Y, "Y"), sv, C.byref(sweeps), wp, wb,
                                  stream_ptr(X.device)), "tnx_svd_trunc_f64")
    return U, Y, np.frombuffer(sv, dtype=np.float64).copy()


class Workspace:
    """Grow-only device scratch buffer shared by the calls of one solver."""

    def __init__(self, device: torch.device):
        self.device = device
        self.buf = torch.empty(0, dtype=torch.uint8, device=device)

    def get(self, nbytes: int) -> tuple[int, int]:
        nbytes = int(nbytes)
        if self.buf.numel() < nbytes:
            self.buf = torch.empty(int(nbytes * 1.1) + 4096, dtype=torch.uint8, device=self.device)
        return self.buf.data_ptr(), self.buf.numel()
