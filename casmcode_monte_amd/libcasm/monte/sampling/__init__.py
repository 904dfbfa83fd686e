"""libcasm.monte.sampling (python/src/monte_sampling.cpp): samplers, sampling
functions, completion-check parameters and results, converge()."""
from casmcode_monte_amd._libcasm_monte import sampling as _s

StateSamplingFunction = _s.StateSamplingFunction
StateSamplingFunctionMap = _s.StateSamplingFunctionMap
jsonStateSamplingFunction = _s.jsonStateSamplingFunction
jsonStateSamplingFunctionMap = _s.jsonStateSamplingFunctionMap
Sampler = _s.Sampler
SamplerMap = _s.SamplerMap
SamplerComponent = _s.SamplerComponent
RequestedPrecision = _s.RequestedPrecision
RequestedPrecisionMap = _s.RequestedPrecisionMap
CutoffCheckParams = _s.CutoffCheckParams
CompletionCheckParams = _s.CompletionCheckParams
CompletionCheckResults = _s.CompletionCheckResults
CompletionCheck = _s.CompletionCheck
BasicStatistics = _s.BasicStatistics
EquilibrationCheckResults = _s.EquilibrationCheckResults
ConvergenceCheckResults = _s.ConvergenceCheckResults
converge = _s.converge
get_n_samples = _s.get_n_samples
CorrelationsDataParams = _s.CorrelationsDataParams
SelectedEventFunctionParams = _s.SelectedEventFunctionParams
