"""The libcasm.monte Python names over the drop-in (casmcode_monte_amd.libcasm).

CPU: objects build and the parameter plumbing (cutoffs, converge) matches the
reference's Python API.  GPU: the reference's own script
python/tests/ising_cpp/test_ising_semigrand_canonical_cpp.py, with only the
import root changed, runs on the MI355X and meets its assertions; a seeded run
reproduces the oracle's sampled series bit for bit.
"""
import numpy as np
import pytest

import casmcode_monte_amd.libcasm.monte as monte
import casmcode_monte_amd.libcasm.monte.ising_cpp as ising
import casmcode_monte_amd.libcasm.monte.ising_cpp.semigrand_canonical as sgc
import casmcode_monte_amd.libcasm.monte.sampling as sampling
import oracle as O


def make_calculator(J=0.1):
    return sgc.SemiGrandCanonicalCalculator(
        system=ising.IsingSystem(
            formation_energy_calculator=ising.IsingFormationEnergy(J=J, lattice_type=1),
            param_composition_calculator=ising.IsingParamComposition(),
        )
    )


def test_objects_and_parameters():
    mc = make_calculator()
    f = mc.default_sampling_functions()
    assert sorted(f.keys()) == ["formation_energy", "param_composition", "potential_energy"]
    assert list(mc.default_json_sampling_functions().keys()) == ["configuration"]
    ccp = sampling.CompletionCheckParams()
    assert (ccp.check_begin, ccp.check_period, ccp.log_spacing) == (100, 100, False)
    ccp.cutoff_params.min_sample = 100
    ccp.cutoff_params.max_count = 500
    assert ccp.cutoff_params.min_sample == 100 and ccp.cutoff_params.max_count == 500
    sampling.converge(f, ccp).set_precision("potential_energy", abs=0.001).set_precision(
        "param_composition", abs=0.002)
    req = {k.sampler_name: r.abs_precision for k, r in ccp.requested_precision.items()}
    assert req == {"potential_energy": 0.001, "param_composition": 0.002}
    state = ising.IsingState(
        configuration=ising.IsingConfiguration(shape=(25, 25)),
        conditions=monte.ValueMap.from_dict({"temperature": 2000.0,
                                             "exchange_potential": [0.0]}))
    assert state.configuration.n_sites == 625
    assert state.conditions.to_dict() == {"temperature": 2000.0, "exchange_potential": [0.0]}


@pytest.mark.gpu
def test_reference_script(tmp_path):
    print("case: python surface, reference script", flush=True)
    mc_calculator = make_calculator()
    sampling_functions = mc_calculator.default_sampling_functions()
    json_sampling_functions = mc_calculator.default_json_sampling_functions()
    initial_state = ising.IsingState(
        configuration=ising.IsingConfiguration(shape=(25, 25)),
        conditions=monte.ValueMap.from_dict({"temperature": 2000.0,
                                             "exchange_potential": [0.0]}))
    for linear_site_index in range(initial_state.configuration.n_sites):
        initial_state.configuration.set_occ(linear_site_index, 1)
    event_generator = sgc.SemiGrandCanonicalEventGenerator()
    completion_check_params = sampling.CompletionCheckParams()
    completion_check_params.cutoff_params.min_sample = 100
    completion_check_params.log_spacing = False
    completion_check_params.check_begin = 100
    completion_check_params.check_period = 10
    sampling.converge(sampling_functions, completion_check_params).set_precision(
        "potential_energy", abs=0.001).set_precision("param_composition", abs=0.001)
    method_log = monte.MethodLog(logfile_path=str(tmp_path / "status.json"), log_frequency=0.2)
    mc_calculator.run(
        state=initial_state, sampling_functions=sampling_functions,
        json_sampling_functions=json_sampling_functions,
        completion_check_params=completion_check_params, event_generator=event_generator,
        sample_period=1, method_log=method_log, random_engine=None)
    samplers = mc_calculator.data.samplers
    results = mc_calculator.data.completion_check.results()
    d = results.to_dict()
    assert d["is_complete"] and "convergence_check_results" in d
    assert sampling.get_n_samples(samplers) >= 100
    assert results.is_complete
    assert results.equilibration_check_results.all_equilibrated
    assert len(results.equilibration_check_results.individual_results) == 2
    assert results.convergence_check_results.all_converged
    assert len(results.convergence_check_results.individual_results) == 2
    converge_results = results.convergence_check_results.individual_results
    for key, req in completion_check_params.requested_precision.items():
        assert converge_results[key].stats.calculated_precision < req.abs_precision


@pytest.mark.gpu
def test_seeded_run_equals_oracle():
    print("case: python surface vs oracle", flush=True)
    mc = make_calculator()
    f = mc.default_sampling_functions()
    ccp = sampling.CompletionCheckParams()
    ccp.cutoff_params.max_count = 700
    state = ising.IsingState(
        configuration=ising.IsingConfiguration(shape=(25, 25)),
        conditions=monte.ValueMap.from_dict({"temperature": 1800.0,
                                             "exchange_potential": [0.05]}))
    engine = monte.RandomNumberEngine()
    engine.seed(42)
    mc.run(state=state, sampling_functions=f,
           json_sampling_functions=mc.default_json_sampling_functions(),
           completion_check_params=ccp, event_generator=sgc.SemiGrandCanonicalEventGenerator(),
           sample_period=1, random_engine=engine)
    ref = O.sgc_run(25, 25, J=0.1, use_nlist=True, temperature=1800.0, mu=0.05, seed=42,
                    max_count=700)
    for name in ("formation_energy", "param_composition", "potential_energy"):
        got = mc.data.samplers[name].component(0)
        np.testing.assert_array_equal(got, ref[name])
    assert (mc.data.n_accept, mc.data.n_reject) == (ref["n_accept"], ref["n_reject"])
    assert np.array_equal(np.array(state.configuration.occupation()), ref["final_occ"])
