"""XXZ chain (``tnpy/model/xxz.py:8-30``).

H = -sum_i (Sx_i Sx_{i+1} + Sy_i Sy_{i+1} + delta Sz_i Sz_{i+1}), written with
S+/S- as -0.5 (S+S- + S-S+) - delta SzSz; open boundary.
"""
import numpy as np

from tnpy_amd.model.model_1d import Model1D
from tnpy_amd.model.utils import boundary_vectors
from tnpy_amd.operators import SpinOperators


class XXZ(Model1D):
    def __init__(self, n: int, delta: float):
        super().__init__(n)
        self.delta = delta

    @boundary_vectors(row=0, col=-1)
    def _elem(self, site: int) -> np.ndarray:
        Sp, Sm, Sz, I2, O2 = SpinOperators()
        return np.array(
            [
                [I2, -0.5 * Sp, -0.5 * Sm, -self.delta * Sz, O2],
                [O2, O2, O2, O2, Sm],
                [O2, O2, O2, O2, Sp],
                [O2, O2, O2, O2, Sz],
                [O2, O2, O2, O2, I2],
            ]
        )
