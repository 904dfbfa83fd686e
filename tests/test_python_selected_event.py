"""libcasm.monte.sampling's selected-event parameters in Python
(python/src/monte_sampling.cpp:3267-3371, 4059-4330): the assertions of the
reference's own test, python/tests/sampling/test_SelectedEventFunctionParams.py,
replayed on the bindings of the drop-in (host only, no device)."""
from casmcode_monte_amd.libcasm.monte import sampling


def test_collect_and_do_not_collect():
    params = sampling.SelectedEventFunctionParams()
    params.collect("event_type")
    assert params.function_names == ["event_type"]
    params.do_not_collect("event_type")
    assert len(params.function_names) == 0


def test_get_parameters():
    params = sampling.SelectedEventFunctionParams()
    params.collect(name="dE_activated", initial_begin=0.01, bin_width=1.0, max_size=10,
                   spacing="linear")
    data = params.get_parameters("dE_activated")
    assert data == {"initial_begin": 0.01, "bin_width": 1.0, "max_size": 10,
                    "spacing": "linear"}
    assert params.get_parameters("other") == {}


def test_collect_hop_correlations():
    params = sampling.SelectedEventFunctionParams()
    assert params.correlations_data_params is None
    params.collect_hop_correlations(jumps_per_position_sample=10)
    assert isinstance(params.correlations_data_params, sampling.CorrelationsDataParams)
    assert params.correlations_data_params.jumps_per_position_sample == 10
    data = params.correlations_data_params.to_dict()
    assert data == {"jumps_per_position_sample": 10, "max_n_position_samples": 100,
                    "output_incomplete_samples": False, "stop_run_when_complete": False}
    back = sampling.CorrelationsDataParams.from_dict(data)
    assert back.jumps_per_position_sample == 10 and back.max_n_position_samples == 100
    params.do_not_collect_hop_correlations()
    assert params.correlations_data_params is None
    params.collect_hop_correlations(**data)
    assert params.correlations_data_params.jumps_per_position_sample == 10


def test_to_and_from_dict():
    params = sampling.SelectedEventFunctionParams()
    params.collect("event_type")
    params.collect(name="dE_activated", initial_begin=0.01, bin_width=1.0, max_size=10,
                   spacing="linear")
    params.collect_hop_correlations(jumps_per_position_sample=10)
    data = params.to_dict()
    assert data["quantities"] == ["event_type", "dE_activated"]
    assert data["spacing"] == {"dE_activated": "linear"}
    assert data["correlations_data_params"]["jumps_per_position_sample"] == 10
    params_in = sampling.SelectedEventFunctionParams.from_dict(data)
    assert params_in.to_dict() == data
    assert params_in.get_parameters("dE_activated") == params.get_parameters("dE_activated")
