"""libcasm.monte.ising_cpp.semigrand_canonical
(python/src/monte_ising_cpp_semigrand_canonical.cpp:503-576): the SGC
calculator, whose run() executes on the MI355X."""
from casmcode_monte_amd._libcasm_monte import ising_cpp as _i

_g = _i.semigrand_canonical
SemiGrandCanonicalCalculator = _g.SemiGrandCanonicalCalculator
SemiGrandCanonicalEventGenerator = _g.SemiGrandCanonicalEventGenerator
SemiGrandCanonicalData = _g.SemiGrandCanonicalData
