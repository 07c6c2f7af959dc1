"""Spin operators and the matrix product operator (host data).

Mirrors ``tnpy/operators.py``: ``SpinOperators`` (``:10-63``) and
``MatrixProductOperator`` (``:66-116``). The MPO is plain numpy here (quimb
is not required); tensors follow quimb's ``'lrud'`` order,
``W[wl, wr, k, b]`` with k the ket (upper) and b the bra (lower) index, and
edge tensors have the reference's 3-leg shapes ``(wr, k, b)`` at site 0 and
``(wl, k, b)`` at the last site. The device copy used by the kernels pads the
edges to 4 legs with a bond of size 1.
"""
from __future__ import annotations

from dataclasses import InitVar, astuple, dataclass, field
from typing import List, Sequence

import numpy as np


@dataclass
class SpinOperators:
    """Sp, Sm, Sz, I2, O2 for the given spin (``tnpy/operators.py:10-63``).

    Unpack in this order: ``Sp, Sm, Sz, I2, O2 = SpinOperators()``.
    """

    spin: InitVar[float] = field(default=0.5)
    Sp: np.ndarray = field(init=False)
    Sm: np.ndarray = field(init=False)
    Sz: np.ndarray = field(init=False)
    I2: np.ndarray = field(init=False)
    O2: np.ndarray = field(init=False)

    def __post_init__(self, spin: float):
        self.Sp = spin * np.array([[0, 2], [0, 0]], dtype=float)
        self.Sm = spin * np.array([[0, 0], [2, 0]], dtype=float)
        self.Sz = spin * np.array([[1, 0], [0, -1]], dtype=float)
        self.I2 = np.identity(2, dtype=float)
        self.O2 = np.zeros((2, 2), dtype=float)

    def __iter__(self):
        return iter(astuple(self))


def _to_arrays(obj) -> List[np.ndarray]:
    if isinstance(obj, MatrixProductOperator):
        return [a.copy() for a in obj._arrays]
    if hasattr(obj, "arrays") and not isinstance(obj, (list, tuple)):
        # a quimb MatrixProductOperator (default 'lrud' shape)
        return [np.asarray(a, dtype=float) for a in obj.arrays]
    return [np.asarray(a, dtype=float) for a in obj]


class MatrixProductOperator:
    """Finite MPO as a list of site tensors (``tnpy/operators.py:66-87``).

    Args:
        arrays: per-site tensors in ``'lrud'`` order; the first and last may
            be 3-leg (reference shape) or 4-leg with a size-1 outer bond. A
            quimb MPO is accepted as well.
    """

    def __init__(self, arrays):
        arrs = _to_arrays(arrays)
        n = len(arrs)
        if n < 2:
            raise ValueError("an MPO needs at least 2 sites")
        padded = []
        for i, a in enumerate(arrs):
            if a.ndim == 4:
                padded.append(np.ascontiguousarray(a, dtype=float))
            elif a.ndim == 3 and i == 0:
                padded.append(np.ascontiguousarray(a[None], dtype=float))
            elif a.ndim == 3 and i == n - 1:
                padded.append(np.ascontiguousarray(a[:, None], dtype=float))
            else:
                raise ValueError(f"MPO tensor {i} has shape {a.shape}")
        if padded[0].shape[0] != 1 or padded[-1].shape[1] != 1:
            raise ValueError("edge MPO tensors must have an outer bond of size 1")
        for i in range(n - 1):
            if padded[i].shape[1] != padded[i + 1].shape[0]:
                raise ValueError(f"MPO bond mismatch between sites {i} and {i + 1}")
        dims = {a.shape[2] for a in padded} | {a.shape[3] for a in padded}
        if len(dims) != 1:
            raise ValueError("All MPO tensors are assumed to have same physical dims.")
        self._padded = padded
        self._arrays = [padded[0][0]] + padded[1:-1] + [padded[-1][:, 0]]
        self._n_sites = n
        self._phys_dim = padded[0].shape[2]

    @property
    def n_sites(self) -> int:
        return self._n_sites

    @property
    def nsites(self) -> int:
        return self._n_sites

    @property
    def phys_dim(self) -> int:
        return self._phys_dim

    @property
    def arrays(self) -> List[np.ndarray]:
        """Site tensors in the reference's shapes (3-leg edges)."""
        return list(self._arrays)

    @property
    def padded(self) -> List[np.ndarray]:
        """Site tensors as ``(Dl, Dr, d, d)`` with size-1 outer bonds."""
        return list(self._padded)

    def bond_dims(self) -> List[int]:
        return [a.shape[1] for a in self._padded[:-1]]

    def __getitem__(self, site: int) -> np.ndarray:
        return self._arrays[site]

    def __len__(self) -> int:
        return self._n_sites

    def __rmul__(self, scalar: float) -> "MatrixProductOperator":
        arrs = [a.copy() for a in self._padded]
        arrs[0] = arrs[0] * scalar
        return MatrixProductOperator(arrs)

    __mul__ = __rmul__

    def square(self) -> "MatrixProductOperator":
        """Two MPO layers merged into one (``tnpy/operators.py:89-116``):
        ``W2[(w1,w2),(v1,v2),k,b] = sum_c W[w1,v1,k,c] W[w2,v2,c,b]``."""
        out = []
        for W in self._padded:
            Dl, Dr, d, _ = W.shape
            W2 = np.einsum("avkc,bwcd->abvwkd", W, W).reshape(Dl * Dl, Dr * Dr, d, d)
            out.append(W2)
        return MatrixProductOperator(out)


def as_mpo(obj) -> MatrixProductOperator:
    if isinstance(obj, MatrixProductOperator):
        return obj
    return MatrixProductOperator(obj)
