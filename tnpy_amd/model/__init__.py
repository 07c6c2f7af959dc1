from tnpy_amd.model.model_1d import Model1D
from tnpy_amd.model.random_coupling_heisenberg import RandomCouplingHeisenberg
from tnpy_amd.model.random_heisenberg import RandomHeisenberg
from tnpy_amd.model.xxz import XXZ

__all__ = ["Model1D", "XXZ", "RandomHeisenberg", "RandomCouplingHeisenberg"]
