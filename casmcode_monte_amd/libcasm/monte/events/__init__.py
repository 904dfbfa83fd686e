"""libcasm.monte.events: occupant candidates, swaps, OccLocation and the event
proposal functions (python/src/monte_events.cpp), over the one-asymmetric-unit
Conversions of this library (Conversions(n_sites, occupants))."""
from casmcode_monte_amd._libcasm_monte import events as _e

Conversions = _e.Conversions
Mol = _e.Mol
OccTransform = _e.OccTransform
OccEvent = _e.OccEvent
OccCandidate = _e.OccCandidate
OccSwap = _e.OccSwap
OccCandidateList = _e.OccCandidateList
OccLocation = _e.OccLocation
is_allowed_canonical_swap = _e.is_allowed_canonical_swap
make_canonical_swaps = _e.make_canonical_swaps
is_allowed_semigrand_canonical_swap = _e.is_allowed_semigrand_canonical_swap
make_semigrand_canonical_swaps = _e.make_semigrand_canonical_swaps
choose_canonical_swap = _e.choose_canonical_swap
propose_canonical_event_from_swap = _e.propose_canonical_event_from_swap
propose_canonical_event = _e.propose_canonical_event
choose_semigrand_canonical_swap = _e.choose_semigrand_canonical_swap
propose_semigrand_canonical_event_from_swap = _e.propose_semigrand_canonical_event_from_swap
propose_semigrand_canonical_event = _e.propose_semigrand_canonical_event

__all__ = [k for k in dir() if not k.startswith("_")]
