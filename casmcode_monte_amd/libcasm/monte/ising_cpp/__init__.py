"""libcasm.monte.ising_cpp (python/src/monte_ising_cpp.cpp): the 2D Ising system."""
from casmcode_monte_amd._libcasm_monte import ising_cpp as _i

IsingConfiguration = _i.IsingConfiguration
IsingState = _i.IsingState
IsingFormationEnergy = _i.IsingFormationEnergy
IsingParamComposition = _i.IsingParamComposition
IsingSystem = _i.IsingSystem
