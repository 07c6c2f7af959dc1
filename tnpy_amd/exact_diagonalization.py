"""Exact diagonalisation on the GPU (``tnpy/exact_diagonalization.py:10-250``
and ``FullHamiltonian``, ``tnpy/operators.py:119-166``).

``FullHamiltonian(mpo).matrix`` is the dense operator of an MPO, built on
the device site by site (``tnx_mpo_absorb``); ``ExactDiagonalization``
diagonalises it completely (``tnx_syev_f64``: the library's SVD of the
positive-definite shift H + cI) and gives the reference's measurements:
Schmidt spectra from ``reduced_density_matrix`` (``tnx_svd_trunc_f64`` of
the reshaped eigenvector), entanglement entropy, one- and two-point
functions and the variance of every eigenstate. Eigenvalues ascend;
eigenvectors are the columns of ``evecs`` (within a degenerate eigenspace
any orthonormal basis is as good as LAPACK's).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from tnpy_amd import _lib, logger
from tnpy_amd.operators import as_mpo


class FullHamiltonian:
    """Dense matrix of an MPO, rows fusing the ket indices k0..k{n-1},
    columns the bra indices (``tnpy/operators.py:119-154``). Raises
    ``ResourceWarning`` beyond 2^12 basis states, as the reference."""

    def __init__(self, mpo, device: Optional[torch.device] = None):
        mpo = as_mpo(mpo)
        self._n_sites = mpo.n_sites
        self._phys_dim = mpo.phys_dim
        if self._phys_dim**self._n_sites > 2**12:
            raise ResourceWarning(f"Requesting more than {self._n_sites} sites with physical dim {self._phys_dim}.")
        self.device = device if device is not None else _lib.require_gpu()
        lib = _lib.lib()
        stream = _lib.stream_ptr(self.device)
        d = self._phys_dim
        Ws = [torch.as_tensor(w, device=self.device).contiguous() for w in mpo.padded]
        # T[w][K][B] after site 0 is W0[0][w][k][b] (its left bond has size 1)
        T = torch.ones(1, 1, 1, dtype=torch.float64, device=self.device)
        dimK = 1
        for W in Ws:
            Dt, Dr = int(W.shape[0]), int(W.shape[1])
            out = torch.empty(Dr, dimK * d, dimK * d, dtype=torch.float64, device=self.device)
            _lib.check(lib.tnx_mpo_absorb(_lib.ptr(T, "T"), Dt, dimK, _lib.ptr(W, "W"), Dr, d,
                                          _lib.ptr(out, "Tout"), stream), "tnx_mpo_absorb")
            T, dimK = out, dimK * d
        self._matrix_dev = T[0].contiguous()

    @property
    def n_sites(self) -> int:
        return self._n_sites

    @property
    def phys_dim(self) -> int:
        return self._phys_dim

    @property
    def matrix(self) -> np.ndarray:
        return self._matrix_dev.cpu().numpy()

    @property
    def matrix_device(self) -> torch.Tensor:
        return self._matrix_dev


def _gemm(A: torch.Tensor, B: torch.Tensor, transA: bool = False) -> torch.Tensor:
    """C = op(A) B on the device through tnx_gemm_f64 (row-major)."""
    if transA:
        K, M = A.shape
    else:
        M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError("inner dimensions differ")
    Cm = torch.empty(M, N, dtype=torch.float64, device=A.device)
    _lib.check(
        _lib.lib().tnx_gemm_f64(M, N, K, 1, 1.0, _lib.ptr(A, "A"), A.shape[1], 0, int(not transA), _lib.ptr(B, "B"),
                                N, 0, 1, 0.0, _lib.ptr(Cm, "C"), N, 0, _lib.TNX_F64, _lib.stream_ptr(A.device)),
        "tnx_gemm_f64",
    )
    return Cm


class ExactDiagonalization(FullHamiltonian):
    """Complete diagonalisation of the MPO's dense Hamiltonian on the GPU,
    computed at construction (``tnpy/exact_diagonalization.py:11-37``).

    Args:
        mpo: the matrix product operator.
        proj: optional projector P; the matrix becomes P^T H P.
    """

    def __init__(self, mpo, proj: Optional[np.ndarray] = None, device: Optional[torch.device] = None):
        super().__init__(mpo, device=device)
        self._proj = proj
        if proj is not None:
            P = torch.as_tensor(np.ascontiguousarray(proj), dtype=torch.float64, device=self.device)
            self._matrix_dev = _gemm(P, _gemm(self._matrix_dev, P), transA=True).contiguous()
        self._evals, self._evecs_dev = self._eigen_solver()

    @property
    def proj(self) -> Optional[np.ndarray]:
        return self._proj

    @property
    def evals(self) -> np.ndarray:
        """Eigenvalues in ascending order."""
        return self._evals

    @property
    def evecs(self) -> np.ndarray:
        """Eigenvectors as columns, in the order of ``evals``."""
        return self._evecs_dev.cpu().numpy()

    @property
    def evecs_device(self) -> torch.Tensor:
        return self._evecs_dev

    def _eigen_solver(self):
        H = self._matrix_dev
        n = int(H.shape[0])
        lib = _lib.lib()
        evecs = torch.empty(n, n, dtype=torch.float64, device=self.device)
        evals = (C.c_double * n)()
        ws = torch.empty(lib.tnx_syev_workspace_bytes(n), dtype=torch.uint8, device=self.device)
        _lib.check(lib.tnx_syev_f64(_lib.ptr(H, "H"), n, evals, _lib.ptr(evecs, "evecs"), ws.data_ptr(), ws.numel(),
                                    _lib.stream_ptr(self.device)), "tnx_syev_f64")
        return np.frombuffer(evals, dtype=np.float64).copy(), evecs

    def reduced_density_matrix(self, site: int, level_idx: int = 0) -> np.ndarray:
        """Squared Schmidt values of eigenstate ``level_idx`` across bond
        (site, site+1) (``tnpy/exact_diagonalization.py:62-90``)."""
        if not 0 <= site < self.n_sites - 1:
            raise ValueError("Parameter `site` for bi-partition has to be within the system size.")
        if not 0 <= level_idx < len(self.evals):
            raise ValueError("Parameter `level_idx` has to be lower than truncation dimension.")
        d, n = self.phys_dim, self.n_sites
        v = self._evecs_dev[:, level_idx].contiguous()
        rows = d ** (site + 1) if site < n // 2 else v.numel() // d ** (n - site - 1)
        M = v.reshape(rows, -1)
        _, _, s = _lib.svd_trunc(M, min(M.shape))
        return np.square(s)

    def entanglement_entropy(self, site: int, level_idx: int = 0, nan_to_num: bool = False) -> float:
        """von Neumann entropy across bond (site, site+1)
        (``tnpy/exact_diagonalization.py:92-111``)."""
        rho = self.reduced_density_matrix(site, level_idx)
        with np.errstate(divide="ignore", invalid="ignore"):
            entropy = -1 * rho @ np.log(rho)
        return np.nan_to_num(entropy) if nan_to_num else entropy

    @staticmethod
    def kron_operators(operators: Sequence[np.ndarray]) -> np.ndarray:
        """Kronecker product of a sequence of operators (``:113-128``)."""
        opt = operators[0]
        for nxt in operators[1:]:
            opt = np.kron(opt, nxt)
        return opt

    def _expect(self, opt_mat: np.ndarray, level_idx: int) -> float:
        O = torch.as_tensor(np.ascontiguousarray(opt_mat), dtype=torch.float64, device=self.device)
        v = self._evecs_dev[:, level_idx].contiguous().reshape(-1, 1)
        return float(_gemm(v, _gemm(O, v), transA=True).reshape(-1)[0])

    def one_point_function(self, operator: np.ndarray, site: int, level_idx: int = 0) -> float:
        """<O_site> in eigenstate ``level_idx`` (``:130-152``)."""
        if operator.shape != (self.phys_dim, self.phys_dim):
            raise ValueError(f"Operator shape mismatched with physical dims {self.phys_dim}.")
        d = self.phys_dim
        opt = self.kron_operators([np.eye(d**site), operator, np.eye(d ** (self.n_sites - site - 1))])
        return self._expect(opt, level_idx)

    def two_point_function(self, operator1: np.ndarray, operator2: np.ndarray, site1: int, site2: int,
                           level_idx: int) -> float:
        """<O1_site1 O2_site2> in eigenstate ``level_idx`` (``:154-197``)."""
        d = self.phys_dim
        if operator1.shape != (d, d) or operator2.shape != (d, d):
            raise ValueError(f"Operator shape mismatched with physical dims {self.phys_dim}.")
        if site1 == site2:
            raise ValueError("Two sites must not be the same site.")
        site1, site2 = sorted((site1, site2))
        opt = self.kron_operators([np.eye(d**site1), operator1, np.eye(d ** (abs(site1 - site2) - 1)), operator2,
                                   np.eye(d ** (self.n_sites - site2 - 1))])
        return self._expect(opt, level_idx)

    def connected_two_point_function(self, operator1: np.ndarray, operator2: np.ndarray, site1: int, site2: int,
                                     level_idx: int) -> float:
        """<O1 O2> - <O1><O2> (``:199-222``)."""
        return self.two_point_function(operator1, operator2, site1, site2, level_idx) - self.one_point_function(
            operator1, site1, level_idx
        ) * self.one_point_function(operator2, site2, level_idx)

    def variance(self, operator: Optional[np.ndarray] = None, tol: float = 1e-12) -> np.ndarray:
        """<O^2> - <O>^2 in every eigenstate (``:224-250``); warns when the
        off-diagonal part exceeds ``tol``."""
        O = (self._matrix_dev if operator is None
             else torch.as_tensor(np.ascontiguousarray(operator), dtype=torch.float64, device=self.device))
        V = self._evecs_dev
        OV = _gemm(O, V)
        A = _gemm(V, OV, transA=True).cpu().numpy()   # V^T O V
        B = _gemm(OV, OV, transA=True).cpu().numpy()  # V^T O^2 V (O symmetric)
        var = B - A * A
        if not np.allclose(np.zeros(var.shape), var - np.diagflat(np.diag(var)), atol=tol):
            logger.warning("Expectation value may contain large off-diagonal elements.")
        return np.diag(var)


__all__ = ["FullHamiltonian", "ExactDiagonalization"]
