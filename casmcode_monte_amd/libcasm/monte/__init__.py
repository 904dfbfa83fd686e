"""libcasm.monte: ValueMap, MethodLog, RandomNumberEngine (python/src/monte.cpp)."""
from casmcode_monte_amd._libcasm_monte import monte as _m

ValueMap = _m.ValueMap
MethodLog = _m.MethodLog
RandomNumberEngine = _m.RandomNumberEngine

__all__ = ["ValueMap", "MethodLog", "RandomNumberEngine"]
