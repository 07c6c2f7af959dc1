"""CPU FP64 restatement of tnpy's finite-DMRG path.

TEST INFRASTRUCTURE ONLY. This module is the oracle: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and only as the checker or the timed CPU baseline. The product
package ``tnpy_amd`` never imports it.

It restates, in plain numpy/scipy, the algorithm of

* ``tnpy/operators.py:10-166``          spin operators, MPO square, dense H
* ``tnpy/model/model_1d.py:25-33``      per-site W -> MPO
* ``tnpy/model/xxz.py:19-30``           XXZ MPO
* ``tnpy/model/random_heisenberg.py:50-58,84-109``  random-field Heisenberg MPO
* ``tnpy/model/utils.py:7-64``          boundary vectors, penalty minors
* ``tnpy/matrix_product_state.py:28-440``  MPS, split_tensor, Environment
* ``tnpy/linalg.py:9-87``               svd / eigh / eigshmv
* ``tnpy/finite_dmrg.py:22-263``        FiniteDMRG sweep / run / convergence
* ``tnpy/exact_diagonalization.py:10-90``  ED and Schmidt spectra

Conventions (identical to the reference):

* MPS site tensors are ``'lpr'`` (left bond, physical, right bond). Edge
  sites carry a bond of size 1 here, where the reference drops the leg
  (``matrix_product_state.py:425-430``); the arithmetic is the same.
* MPO tensors are ``W[wl, wr, k, b]`` (quimb ``'lrud'``, k = ket/upper,
  b = bra/lower), edge tensors padded to a bond of size 1.
* Environments are ``L[l_ket, w, l_bra]`` / ``R[r_ket, w, r_bra]``, the
  (ket, mpo, bra) order the reference relies on at
  ``matrix_product_state.py:386-407,427-437``.

The per-site Krylov solver of the reference is primme 3.2.1 (not vendored,
``tnpy/linalg.py:86``); here ``scipy.sparse.linalg.eigsh`` (ARPACK,
``which='SA'``, ``k=1``, same ``v0`` and ``tol``) stands in for it. Energies
agree to the eigen tolerance; iteration trajectories are not comparable
(parity of the trajectory is unpinned, the converged energies are pinned by
ED and closed forms in ``tests/``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

RIGHTWARD = 1
LEFTWARD = -1


# --------------------------------------------------------------------------
# operators (tnpy/operators.py:10-63)
# --------------------------------------------------------------------------
def spin_operators(spin: float = 0.5):
    """Sp, Sm, Sz, I2, O2 as ``SpinOperators.__post_init__``
    (``tnpy/operators.py:55-60``)."""
    Sp = spin * np.array([[0, 2], [0, 0]], dtype=float)
    Sm = spin * np.array([[0, 0], [2, 0]], dtype=float)
    Sz = spin * np.array([[1, 0], [0, -1]], dtype=float)
    I2 = np.identity(2, dtype=float)
    O2 = np.zeros((2, 2), dtype=float)
    return Sp, Sm, Sz, I2, O2


def _boundary(tensor: np.ndarray, site: int, n: int, row: int = 0, col: int = -1):
    """``boundary_vectors(row=0, col=-1)`` (``tnpy/model/utils.py:17-31``),
    returning 4-leg tensors with a size-1 outer bond."""
    if site == 0:
        r = row % tensor.shape[0]
        return tensor[r : r + 1].copy()
    if site == n - 1:
        c = col % tensor.shape[1]
        return tensor[:, c : c + 1].copy()
    return tensor.copy()


def xxz_bulk(delta: float) -> np.ndarray:
    """Bulk W of ``XXZ._elem`` (``tnpy/model/xxz.py:19-30``):
    H = -sum (0.5 S+S- + 0.5 S-S+ + delta SzSz)."""
    Sp, Sm, Sz, I2, O2 = spin_operators()
    return np.array(
        [
            [I2, -0.5 * Sp, -0.5 * Sm, -delta * Sz, O2],
            [O2, O2, O2, O2, Sm],
            [O2, O2, O2, O2, Sp],
            [O2, O2, O2, O2, Sz],
            [O2, O2, O2, O2, I2],
        ]
    )


def xxz_mpo(n: int, delta: float) -> List[np.ndarray]:
    """``XXZ(n, delta).mpo`` (``model_1d.py:25-33`` + ``xxz.py:19-30``)."""
    W = xxz_bulk(delta)
    return [_boundary(W, i, n) for i in range(n)]


def random_fields(n: int, h: float, seed: Optional[int]) -> np.ndarray:
    """``np.random.RandomState(seed).uniform(-h, h, n)``
    (``tnpy/model/random_heisenberg.py:57-58``)."""
    return np.random.RandomState(seed).uniform(-h, h, size=n)


def random_heisenberg_bulk(
    site: int,
    n: int,
    fields: np.ndarray,
    penalty: float = 0.0,
    s_target: int = 0,
    offset: float = 0.0,
) -> np.ndarray:
    """``RandomHeisenberg._elem`` before the boundary/minor decorators
    (``tnpy/model/random_heisenberg.py:84-109``)."""
    Sp, Sm, Sz, I2, O2 = spin_operators()
    alpha = penalty * (0.25 + s_target**2 / n) - offset / n
    beta = fields[site] - 2.0 * penalty * s_target
    W = np.array(
        [
            [I2, 0.5 * Sp, 0.5 * Sm, 2.0 * penalty * Sz, Sz, alpha * I2 + beta * Sz],
            [O2, O2, O2, O2, O2, Sm],
            [O2, O2, O2, O2, O2, Sp],
            [O2, O2, O2, I2, O2, Sz],
            [O2, O2, O2, O2, O2, Sz],
            [O2, O2, O2, O2, O2, I2],
        ],
        dtype=float,
    )
    if penalty == 0:  # minors_if_no_penalty(row=3, col=3), utils.py:55-61
        W = np.delete(np.delete(W, 3, axis=0), 3, axis=1)
    return W


def random_heisenberg_mpo(
    n: int,
    h: float,
    penalty: float = 0.0,
    s_target: int = 0,
    offset: float = 0.0,
    seed: Optional[int] = None,
) -> List[np.ndarray]:
    """``RandomHeisenberg(n, h, penalty, s_target, offset, seed=seed).mpo``."""
    fields = random_fields(n, h, seed)
    return [
        _boundary(random_heisenberg_bulk(i, n, fields, penalty, s_target, offset), i, n)
        for i in range(n)
    ]


def random_coupling_heisenberg_mpo(n: int, couplings: Sequence[float]) -> List[np.ndarray]:
    """Random-coupling Heisenberg chain H = sum_i J_i S_i . S_{i+1}.

    No reference function exists for this model. It is the RandomHeisenberg
    W at h=0 with J_i multiplying the first-row XY and Z couplings of site
    i (the operators that open bond (i, i+1)). Pinned by brute-force ED in
    the tests.
    """
    J = np.asarray(couplings, dtype=float)
    if J.shape != (n - 1,):
        raise ValueError("need n-1 couplings")
    Sp, Sm, Sz, I2, O2 = spin_operators()
    mpo = []
    for i in range(n):
        j = J[i] if i < n - 1 else 0.0
        W = np.array(
            [
                [I2, 0.5 * j * Sp, 0.5 * j * Sm, j * Sz, O2],
                [O2, O2, O2, O2, Sm],
                [O2, O2, O2, O2, Sp],
                [O2, O2, O2, O2, Sz],
                [O2, O2, O2, O2, I2],
            ]
        )
        mpo.append(_boundary(W, i, n))
    return mpo


def mpo_square(mpo: List[np.ndarray]) -> List[np.ndarray]:
    """``MatrixProductOperator.square`` (``tnpy/operators.py:89-116``):
    W2[(w1,w2), (v1,v2), k, b] = sum_c W[w1,v1,k,c] W[w2,v2,c,b]."""
    out = []
    for W in mpo:
        Dl, Dr, d, _ = W.shape
        W2 = np.einsum("avkc,bwcd->abvwkd", W, W).reshape(Dl * Dl, Dr * Dr, d, d)
        out.append(W2)
    return out


def full_hamiltonian(mpo: List[np.ndarray]) -> np.ndarray:
    """``FullHamiltonian(mpo).matrix`` (``tnpy/operators.py:119-154``):
    rows fuse the ket indices k0..k{n-1}, columns the bra indices."""
    n = len(mpo)
    d = mpo[0].shape[2]
    if d**n > 2**12:
        raise ResourceWarning(f"Requesting more than {n} sites with physical dim {d}.")
    # T[w, K, B] accumulated left to right
    T = mpo[0][0]  # (wr, k, b)
    T = T.reshape(T.shape[0], d, d)
    dimK = d
    for W in mpo[1:]:
        # T[w, K, B] W[w, v, k, b] -> T'[v, K k, B b]
        T = np.einsum("wKB,wvkb->vKkBb", T, W)
        dimK *= d
        T = T.reshape(T.shape[0], dimK, dimK)
    return T[0]


def bond_term_hamiltonian_sparse(n: int, terms) -> sp.csr_matrix:
    """Sparse H from nearest-neighbour terms: ``terms`` is a list over bonds
    of lists ``(coef, opA, opB)`` plus optional on-site fields. Used for ED
    beyond the dense 2^12 cap (sparse Lanczos, n=16)."""
    dim = 2**n
    H = sp.csr_matrix((dim, dim))
    eye = lambda m: sp.identity(2**m, format="csr")
    for i, bond in enumerate(terms):
        for coef, A, B in bond:
            if B is None:
                op = sp.kron(sp.kron(eye(i), sp.csr_matrix(A)), eye(n - i - 1))
            else:
                op = sp.kron(
                    sp.kron(sp.kron(eye(i), sp.csr_matrix(A)), sp.csr_matrix(B)),
                    eye(n - i - 2),
                )
            H = H + coef * op
    return H.tocsr()


def xxz_sparse(n: int, delta: float) -> sp.csr_matrix:
    """Sparse XXZ H with tnpy's sign convention (``xxz.py:19-30``)."""
    Sp, Sm, Sz, _, _ = spin_operators()
    bonds = [[(-0.5, Sp, Sm), (-0.5, Sm, Sp), (-delta, Sz, Sz)] for _ in range(n - 1)]
    return bond_term_hamiltonian_sparse(n, bonds)


# --------------------------------------------------------------------------
# exact diagonalization (tnpy/exact_diagonalization.py:10-90)
# --------------------------------------------------------------------------
class ExactDiagonalization:
    """Dense ``eigh`` of ``FullHamiltonian`` (``exact_diagonalization.py:11-60``)."""

    def __init__(self, mpo: List[np.ndarray]):
        self.n_sites = len(mpo)
        self.phys_dim = mpo[0].shape[2]
        self.matrix = full_hamiltonian(mpo)
        self.evals, self.evecs = np.linalg.eigh(self.matrix)

    def reduced_density_matrix(self, site: int, level_idx: int = 0) -> np.ndarray:
        """Squared Schmidt values across bond (site, site+1)
        (``exact_diagonalization.py:62-90``)."""
        n, d = self.n_sites, self.phys_dim
        if not 0 <= site < n - 1:
            raise ValueError("site out of range")
        shape = (d ** (site + 1), -1) if site < n // 2 else (-1, d ** (n - site - 1))
        return np.square(np.linalg.svd(self.evecs[:, level_idx].reshape(shape), compute_uv=False))


# --------------------------------------------------------------------------
# MPS (tnpy/matrix_product_state.py:28-225)
# --------------------------------------------------------------------------
def bond_dims(n: int, chi: int, d: int) -> List[int]:
    """Bond dims of a compressed random MPS, ``min(d^(i+1), d^(n-i-1), chi)``
    for bond i (between sites i and i+1); the formula of
    ``tests/test_matrix_product_state.py:11-20``."""
    return [int(min(d ** (i + 1), d ** (n - i - 1), chi)) for i in range(n - 1)]


def site_shapes(n: int, chi: int, d: int) -> List[Tuple[int, int, int]]:
    b = [1] + bond_dims(n, chi, d) + [1]
    return [(b[i], d, b[i + 1]) for i in range(n)]


def random_mps(n: int, chi: int, d: int, seed: int) -> List[np.ndarray]:
    """Seeded random MPS: entries N(0,1) from ``numpy.random.default_rng(seed)``
    drawn site by site in C order of ``(l, p, r)``, then right-canonicalised by
    an LQ sweep from the right end and normalised (the state
    ``MPS_rand_state(...).compress()`` produces at
    ``matrix_product_state.py:183-185``, right-canonical with centre at
    site 0, up to its unseeded draw)."""
    rng = np.random.default_rng(seed)
    arrays = [rng.standard_normal(s) for s in site_shapes(n, chi, d)]
    return right_canonicalize(arrays)


def right_canonicalize(arrays: List[np.ndarray]) -> List[np.ndarray]:
    arrays = [a.copy() for a in arrays]
    n = len(arrays)
    for i in range(n - 1, 0, -1):
        l, p, r = arrays[i].shape
        q, rr = np.linalg.qr(arrays[i].reshape(l, p * r).T)  # M^T = Q R -> M = R^T Q^T
        arrays[i] = q.T.reshape(l, p, r)
        pl, pp, pr = arrays[i - 1].shape
        arrays[i - 1] = (arrays[i - 1].reshape(pl * pp, pr) @ rr.T).reshape(pl, pp, pr)
        arrays[i - 1] /= np.linalg.norm(arrays[i - 1])  # rescaling keeps long chains finite
    arrays[0] = arrays[0] / np.linalg.norm(arrays[0])
    return arrays


def mps_to_dense(arrays: List[np.ndarray]) -> np.ndarray:
    v = arrays[0].reshape(-1, arrays[0].shape[-1])
    for a in arrays[1:]:
        l, p, r = a.shape
        v = (v @ a.reshape(l, p * r)).reshape(-1, r)
    return v.reshape(-1)


def split_tensor_svd(arrays: List[np.ndarray], site: int, direction: int) -> None:
    """``MatrixProductState.split_tensor`` (``matrix_product_state.py:187-225``)
    with ``linalg.svd`` (``linalg.py:9-23``): untruncated SVD gauge move."""
    A = arrays[site]
    l, p, r = A.shape
    if direction == RIGHTWARD:
        u, s, vt = np.linalg.svd(A.reshape(l * p, r), full_matrices=False)
        u, s, vt = u[:, :r], s[:r], vt[:r, :]
        arrays[site] = u.reshape(l, p, r)
        nl, npp, nr = arrays[site + 1].shape
        arrays[site + 1] = ((np.diag(s) @ vt) @ arrays[site + 1].reshape(nl, npp * nr)).reshape(
            nl, npp, nr
        )
    elif direction == LEFTWARD:
        u, s, vt = np.linalg.svd(A.reshape(l, p * r), full_matrices=False)
        u, s, vt = u[:, :l], s[:l], vt[:l, :]
        arrays[site] = vt.reshape(l, p, r)
        pl, pp, pr = arrays[site - 1].shape
        arrays[site - 1] = (arrays[site - 1].reshape(pl * pp, pr) @ (u @ np.diag(s))).reshape(
            pl, pp, pr
        )
    else:
        raise KeyError("MatrixProductState only supplies left or right direction.")


def schmidt_values(arrays: List[np.ndarray], normalize: bool = True) -> List[np.ndarray]:
    """Singular values across every bond (gauge invariant). The state is
    left-orthogonalised by QR from the left; at bond i the singular values
    of the centre matrix are the Schmidt values. With ``normalize`` the
    state is scaled to unit norm first (the DMRG state is not normalised,
    ``finite_dmrg.py:139-141``)."""
    arrays = [a.copy() for a in arrays]
    n = len(arrays)
    # right-orthogonalise, then move a centre to the right reading s at each bond
    arrays = right_canonicalize_keep_norm(arrays)
    if normalize:
        arrays[0] = arrays[0] / np.linalg.norm(arrays[0])
    out = []
    for i in range(n - 1):
        l, p, r = arrays[i].shape
        q, rr = np.linalg.qr(arrays[i].reshape(l * p, r))
        out.append(np.linalg.svd(rr, compute_uv=False))
        arrays[i] = q.reshape(l, p, r)
        nl, npp, nr = arrays[i + 1].shape
        arrays[i + 1] = (rr @ arrays[i + 1].reshape(nl, npp * nr)).reshape(nl, npp, nr)
    return out


def right_canonicalize_keep_norm(arrays: List[np.ndarray]) -> List[np.ndarray]:
    arrays = [a.copy() for a in arrays]
    for i in range(len(arrays) - 1, 0, -1):
        l, p, r = arrays[i].shape
        q, rr = np.linalg.qr(arrays[i].reshape(l, p * r).T)
        arrays[i] = q.T.reshape(l, p, r)
        pl, pp, pr = arrays[i - 1].shape
        arrays[i - 1] = (arrays[i - 1].reshape(pl * pp, pr) @ rr.T).reshape(pl, pp, pr)
    return arrays


# --------------------------------------------------------------------------
# Environment (tnpy/matrix_product_state.py:231-440)
# --------------------------------------------------------------------------
def env_left(L: np.ndarray, A: np.ndarray, W: np.ndarray) -> np.ndarray:
    """``Environment.update_left`` (``matrix_product_state.py:296-315``):
    L'[r, v, s] = sum L[a, w, b] A[a, k, r] W[w, v, k, c] A[b, c, s]."""
    a, w, b = L.shape
    _, k, r = A.shape
    Dl, Dr, d, _ = W.shape
    # T1[w, b, k, r] = sum_a L[a, w, b] A[a, k, r]
    T1 = (L.reshape(a, w * b).T @ A.reshape(a, k * r)).reshape(w, b, k, r)
    # T2[v, b, c, r] = sum_{w,k} W[w, v, k, c] T1[w, b, k, r]
    T2 = np.einsum("wvkc,wbkr->vbcr", W, T1)
    # L'[r, v, s] = sum_{b,c} T2[v, b, c, r] A[b, c, s]
    out = np.einsum("vbcr,bcs->rvs", T2, A, optimize=True)
    return out


def env_right(R: np.ndarray, A: np.ndarray, W: np.ndarray) -> np.ndarray:
    """``Environment.update_right`` (``matrix_product_state.py:317-336``):
    R'[a, w, b] = sum R[r, v, s] A[a, k, r] W[w, v, k, c] A[b, c, s]."""
    r, v, s = R.shape
    a, k, _ = A.shape
    # T1[a, k, v, s] = sum_r A[a, k, r] R[r, v, s]
    T1 = (A.reshape(a * k, r) @ R.reshape(r, v * s)).reshape(a, k, v, s)
    # T2[a, w, c, s] = sum_{k,v} W[w, v, k, c] T1[a, k, v, s]
    T2 = np.einsum("wvkc,akvs->awcs", W, T1)
    # R'[a, w, b] = sum_{c,s} T2[a, w, c, s] A[b, c, s]
    bdim = A.shape[0]
    out = T2.reshape(a * T2.shape[1], -1) @ A.reshape(bdim, -1).T
    return out.reshape(a, T2.shape[1], bdim)


def heff_matvec(L: np.ndarray, W: np.ndarray, R: np.ndarray, x: np.ndarray) -> np.ndarray:
    """``Environment.one_site_matvec`` (``matrix_product_state.py:411-440``):
    y[b, c, s] = sum L[a, w, b] W[w, v, k, c] R[r, v, s] x[a, k, r]."""
    a, w, b = L.shape
    r, v, s = R.shape
    Dl, Dr, d, _ = W.shape
    X = x.reshape(a, d, r)
    # T1[a, k, v, s] = sum_r X[a, k, r] R[r, v, s]
    T1 = (X.reshape(a * d, r) @ R.reshape(r, v * s)).reshape(a, d, v, s)
    # T2[a, w, c, s] = sum_{k, v} W[w, v, k, c] T1[a, k, v, s]
    T2 = np.einsum("wvkc,akvs->awcs", W, T1)
    # y[b, c, s] = sum_{a, w} L[a, w, b] T2[a, w, c, s]
    y = L.reshape(a * w, b).T @ T2.reshape(a * w, d * s)
    return y.reshape(-1)


def heff_dense(L: np.ndarray, W: np.ndarray, R: np.ndarray) -> np.ndarray:
    """``Environment.one_site_full_matrix`` (``matrix_product_state.py:372-409``):
    rows fuse (l, k, r) ket, columns (l', b, r') bra."""
    M = np.einsum("awb,wvkc,rvs->akrbcs", L, W, R)
    n = L.shape[0] * W.shape[2] * R.shape[0]
    return M.reshape(n, n)


def two_site_matvec(L, W1, W2, R, x):
    """Two-site H_eff: y[b,c1,c2,s] = sum L[a,w,b] W1[w,u,k1,c1] W2[u,v,k2,c2]
    R[r,v,s] x[a,k1,k2,r]. No reference function (``finite_dmrg.py:113-114``
    is a stub); this is the textbook two-site contraction, as four pairwise
    (BLAS) contractions: R, then W2, then W1, then L."""
    a, w, b = L.shape
    r, v, s = R.shape
    d1, d2 = W1.shape[2], W2.shape[2]
    X = np.asarray(x).reshape(a, d1, d2, r)
    T1 = np.tensordot(X, R, axes=([3], [0]))            # a k1 k2 v s
    T2 = np.tensordot(T1, W2, axes=([2, 3], [2, 1]))    # a k1 s u c2
    T3 = np.tensordot(T2, W1, axes=([1, 3], [2, 1]))    # a s c2 w c1
    y = np.tensordot(L, T3, axes=([0, 1], [0, 3]))      # b s c2 c1
    return y.transpose(0, 3, 2, 1).reshape(-1)


def expectation(arrays: List[np.ndarray], mpo: List[np.ndarray]) -> float:
    """<psi|H|psi> unnormalised, as ``MatrixProductStateMeasurements``
    (``matrix_product_state.py:447-453``) and the second term of
    ``Environment.variance`` (``:367-370``)."""
    L = np.ones((1, 1, 1))
    for A, W in zip(arrays, mpo):
        L = env_left(L, A, W)
    return float(L.reshape(-1)[0])


def variance(arrays: List[np.ndarray], mpo: List[np.ndarray]) -> float:
    """``Environment.variance`` (``matrix_product_state.py:367-370``)."""
    return expectation(arrays, mpo_square(mpo)) - expectation(arrays, mpo) ** 2


# --------------------------------------------------------------------------
# DMRG driver (tnpy/finite_dmrg.py:33-257)
# --------------------------------------------------------------------------
@dataclass
class OracleDMRG:
    """One-site finite DMRG, step for step as ``FiniteDMRG``
    (``finite_dmrg.py:33-257``), on explicit arrays."""

    mpo: List[np.ndarray]
    mps: List[np.ndarray]
    exact_solver_dim: int = 200
    alpha: float = 1e-5
    left: Dict[int, np.ndarray] = field(default_factory=dict)
    right: Dict[int, np.ndarray] = field(default_factory=dict)
    energies: List[float] = field(default_factory=lambda: [np.nan])
    variances: List[float] = field(default_factory=lambda: [np.nan])
    site_energies: List[float] = field(default_factory=list)
    matvecs: int = 0

    def __post_init__(self):
        self.mps = [a.copy() for a in self.mps]
        self.n = len(self.mpo)
        self.left[0] = np.ones((1, 1, 1))
        self.right[self.n - 1] = np.ones((1, 1, 1))
        # Environment.__init__ (matrix_product_state.py:247-250)
        for site in range(1, self.n):
            self.update_left(site)
        for site in range(self.n - 2, -1, -1):
            self.update_right(site)

    def update_left(self, site: int):
        self.left[site] = env_left(self.left[site - 1], self.mps[site - 1], self.mpo[site - 1])

    def update_right(self, site: int):
        self.right[site] = env_right(self.right[site + 1], self.mps[site + 1], self.mpo[site + 1])

    def matvec(self, site: int, x: np.ndarray) -> np.ndarray:
        self.matvecs += 1
        return heff_matvec(self.left[site], self.mpo[site], self.right[site], x)

    def one_site_solver(self, site: int, tol: float) -> Tuple[float, np.ndarray]:
        """``FiniteDMRG.one_site_solver`` (``finite_dmrg.py:97-111``)."""
        v0 = self.mps[site].reshape(-1)
        if v0.size < self.exact_solver_dim:
            M = heff_dense(self.left[site], self.mpo[site], self.right[site])
            evals, evecs = np.linalg.eigh(M)  # linalg.eigh (linalg.py:42-61)
            return float(evals[0]), evecs[:, 0]
        N = v0.size
        op = spla.LinearOperator((N, N), matvec=lambda x: self.matvec(site, np.ravel(x)), dtype=float)
        evals, evecs = spla.eigsh(op, k=1, which="SA", v0=v0, tol=tol)
        return float(evals[0]), evecs[:, 0]

    def perturb(self, site: int):
        """``FiniteDMRG.perturb_wave_function`` (``finite_dmrg.py:116-141``)."""
        psi = self.mps[site].reshape(-1).copy()
        psi += self.alpha * self.matvec(site, psi)
        self.mps[site] = psi.reshape(self.mps[site].shape)

    def sweep(self, direction: int, tol: float) -> float:
        """``FiniteDMRG.sweep`` (``finite_dmrg.py:143-171``)."""
        sites = range(self.n - 1) if direction == RIGHTWARD else range(self.n - 1, 0, -1)
        energy = None
        for site in sites:
            energy, psi = self.one_site_solver(site, tol)
            self.site_energies.append(energy)
            self.mps[site] = psi.reshape(self.mps[site].shape)
            self.perturb(site)
            split_tensor_svd(self.mps, site, direction)
            if direction == RIGHTWARD:
                self.update_left(site + 1)
            else:
                self.update_right(site - 1)
        return energy

    def run(self, tol: float = 1e-8, max_sweep: int = 100, with_variance: bool = True) -> List[float]:
        """``FiniteDMRG.run`` (``finite_dmrg.py:214-257``) with the
        ``Metric.ENERGY`` stopping rule of ``_converged`` (``:173-212``)."""
        direction = RIGHTWARD
        for n_sweep in range(1, max_sweep + 1):
            energy = self.sweep(direction, tol)
            self.energies.append(energy)
            self.variances.append(variance(self.mps, self.mpo) if with_variance else np.nan)
            gradient = np.diff(self.energies[-2:])[0]
            if abs(gradient) < tol:
                break
            direction = -direction
        return self.energies[1:]


def split_two_site_svd(theta: np.ndarray, chi: int, direction: int):
    """Truncated SVD of the two-site tensor theta[l, k1, k2, r] (shape of
    (l*d) x (d*r)) keeping min(chi, l*d, d*r) singular values. Rightward:
    (U, diag(s) V^T); leftward: (U diag(s), V^T). Returns the two site
    tensors and the singular values."""
    l, d1, d2, r = theta.shape
    u, s, vt = np.linalg.svd(theta.reshape(l * d1, d2 * r), full_matrices=False)
    k = min(chi, l * d1, d2 * r)
    u, s, vt = u[:, :k], s[:k], vt[:k, :]
    if direction == RIGHTWARD:
        return u.reshape(l, d1, k), (s[:, None] * vt).reshape(k, d2, r), s
    return (u * s[None, :]).reshape(l, d1, k), vt.reshape(k, d2, r), s


@dataclass
class OracleDMRG2(OracleDMRG):
    """Two-site finite DMRG with SVD truncation to ``chi``. The reference's
    two-site solver is a stub (``finite_dmrg.py:113-114``); this is the
    textbook algorithm on the same environments, eigensolver switch and
    convergence rule as the one-site driver."""

    chi: int = 0

    def solve2(self, site: int, tol: float):
        A, B = self.mps[site], self.mps[site + 1]
        theta = np.tensordot(A, B, axes=(2, 0))
        shape = theta.shape
        v0 = theta.reshape(-1)
        L, R = self.left[site], self.right[site + 1]
        W1, W2 = self.mpo[site], self.mpo[site + 1]
        if v0.size < self.exact_solver_dim:
            M = np.einsum("awb,wukc,uvld,rvs->aklrbcds", L, W1, W2, R).reshape(v0.size, v0.size)
            w, V = np.linalg.eigh(M)
            return float(w[0]), V[:, 0].reshape(shape)
        N = v0.size

        def mv(x):
            self.matvecs += 1
            return two_site_matvec(L, W1, W2, R, np.ravel(x).reshape(shape))

        op = spla.LinearOperator((N, N), matvec=mv, dtype=float)
        w, V = spla.eigsh(op, k=1, which="SA", v0=v0, tol=tol)
        return float(w[0]), V[:, 0].reshape(shape)

    def sweep(self, direction: int, tol: float) -> float:
        sites = range(self.n - 1) if direction == RIGHTWARD else range(self.n - 2, -1, -1)
        energy = None
        for site in sites:
            energy, theta = self.solve2(site, tol)
            self.site_energies.append(energy)
            A, B, _ = split_two_site_svd(theta, self.chi, direction)
            self.mps[site], self.mps[site + 1] = A, B
            if direction == RIGHTWARD:
                self.update_left(site + 1)
            else:
                self.update_right(site)
        return energy


def xx_chain_ground_energy(n: int) -> float:
    """Closed form for XXZ at delta=0 (free fermions, open chain):
    E0 = sum of negative single-particle energies -cos(k pi / (n+1))."""
    eps = -np.cos(np.arange(1, n + 1) * math.pi / (n + 1))
    return float(eps[eps < 0].sum())
