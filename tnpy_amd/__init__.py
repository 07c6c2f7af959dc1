"""MI355X-native finite DMRG with tnpy's API.

The contractions, the Krylov eigensolver, the small dense eigensolver and
the QR gauge moves run as hand-written HIP kernels in ``libtnx.so``
(``tnpy_amd/csrc``), called through ctypes; torch owns the device memory.
"""
import logging

__version__ = "0.1.0"

formatter = logging.Formatter(
    "%(asctime)s [%(filename)s] %(levelname)s: %(message)s", datefmt="%Y-%m-%d %H:%M:%S"
)
handler = logging.StreamHandler()
handler.setLevel(logging.INFO)
handler.setFormatter(formatter)

# same logger name and level as tnpy/__init__.py:10-20
logger = logging.getLogger("tnpy")
logger.setLevel(logging.INFO)
if not logger.handlers:
    logger.addHandler(handler)

from tnpy_amd.operators import MatrixProductOperator, SpinOperators  # noqa: E402
from tnpy_amd.matrix_product_state import (  # noqa: E402
    Direction,
    Environment,
    MatrixProductState,
    MatrixProductStateMeasurements,
)
from tnpy_amd.finite_dmrg import FiniteDMRG, Metric, ShiftInvertDMRG  # noqa: E402
from tnpy_amd.batched_dmrg import BatchedFiniteDMRG  # noqa: E402
from tnpy_amd.exact_diagonalization import ExactDiagonalization, FullHamiltonian  # noqa: E402

__all__ = [
    "logger",
    "SpinOperators",
    "MatrixProductOperator",
    "MatrixProductState",
    "Environment",
    "MatrixProductStateMeasurements",
    "Direction",
    "FiniteDMRG",
    "ShiftInvertDMRG",
    "Metric",
    "BatchedFiniteDMRG",
    "ExactDiagonalization",
    "FullHamiltonian",
]
