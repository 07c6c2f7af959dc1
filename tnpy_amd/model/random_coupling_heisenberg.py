"""Heisenberg chain with random bond couplings.

H = sum_i J_i S_i . S_{i+1}, J_i ~ Uniform(0, 1) per bond from
``numpy.random.default_rng(seed)``. The reference has no such model
(``tnpy/model/random_heisenberg.py`` randomises fields, ``dimer_xxz.py``
only the XY part with Pauli normalisation); this one reuses the
RandomHeisenberg W at h=0 with J_i multiplying the couplings that open bond
(i, i+1).
"""
from typing import Optional, Sequence

import numpy as np

from tnpy_amd.model.model_1d import Model1D
from tnpy_amd.model.utils import boundary_vectors
from tnpy_amd.operators import SpinOperators


class RandomCouplingHeisenberg(Model1D):
    def __init__(self, n: int, seed: Optional[int] = None, couplings: Optional[Sequence[float]] = None):
        super().__init__(n)
        self._seed = seed
        if couplings is None:
            couplings = np.random.default_rng(seed).uniform(0.0, 1.0, size=n - 1)
        self._J = np.asarray(couplings, dtype=float)
        if self._J.shape != (n - 1,):
            raise ValueError("need n-1 couplings")

    @property
    def couplings(self) -> np.ndarray:
        return self._J.copy()

    @boundary_vectors(row=0, col=-1)
    def _elem(self, site: int) -> np.ndarray:
        Sp, Sm, Sz, I2, O2 = SpinOperators()
        j = self._J[site] if site < self.n - 1 else 0.0
        return np.array(
            [
                [I2, 0.5 * j * Sp, 0.5 * j * Sm, j * Sz, O2],
                [O2, O2, O2, O2, Sm],
                [O2, O2, O2, O2, Sp],
                [O2, O2, O2, O2, Sz],
                [O2, O2, O2, O2, I2],
            ]
        )
