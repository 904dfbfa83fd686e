"""The libcasm.monte Python names over the MI355X drop-in: import
casmcode_monte_amd.libcasm.monte (and its .sampling, .ising_cpp and
.ising_cpp.semigrand_canonical submodules) where scripts import libcasm.monte."""
