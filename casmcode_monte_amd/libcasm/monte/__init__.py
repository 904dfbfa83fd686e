"""libcasm.monte: ValueMap, MethodLog, RandomNumberEngine, RandomNumberGenerator
(python/src/monte.cpp)."""
from casmcode_monte_amd._libcasm_monte import monte as _m

ValueMap = _m.ValueMap
MethodLog = _m.MethodLog
RandomNumberEngine = _m.RandomNumberEngine
RandomNumberGenerator = _m.RandomNumberGenerator

__all__ = ["ValueMap", "MethodLog", "RandomNumberEngine", "RandomNumberGenerator"]
