// libcasm_monte.cpp -- Python surface of the C++ drop-in, mirroring the names
// of libcasm-monte's own bindings so that scripts written against
// libcasm.monte / libcasm.monte.sampling / libcasm.monte.ising_cpp /
// libcasm.monte.ising_cpp.semigrand_canonical run unchanged apart from the
// import root (casmcode_monte_amd.libcasm.monte ...).
//
// Reference bindings: python/src/monte.cpp (ValueMap, MethodLog,
// RandomNumberEngine), python/src/monte_sampling.cpp (Sampler,
// StateSamplingFunction, CompletionCheck*, RequestedPrecision, converge),
// python/src/monte_ising_cpp.cpp and
// python/src/monte_ising_cpp_semigrand_canonical.cpp:503-576 (the Ising
// system and the SGC calculator), python/src/monte_events.cpp (occupant
// candidates, swaps, OccLocation and the event proposal functions, over the
// one-asymmetric-unit Conversions of this library).  The calculator's run() drives the MI355X
// library through the C++ drop-in (include/casm/monte/ising_cpp/
// basic_semigrand_canonical.hh); nothing here runs a Metropolis step on the
// host.
#include <sstream>

#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>

#include "casm/monte/events/OccEventProposal.hh"
#include "casm/monte/ising_cpp/basic_semigrand_canonical.hh"
#include "casm/monte/sampling/io/json/SelectedEventFunctions_json_io.hh"

namespace py = pybind11;
using namespace CASM;
using namespace CASM::monte;
using namespace CASM::monte::ising_cpp;
namespace sgc = CASM::monte::ising_cpp::basic_semigrand_canonical;

typedef std::mt19937_64 engine_type;
typedef RandomNumberGenerator<engine_type> generator_type;
typedef sgc::SemiGrandCanonicalEventGenerator<engine_type> event_generator_type;
typedef sgc::SemiGrandCanonicalCalculator<IsingSystem, event_generator_type> calculator_type;
typedef std::map<SamplerComponent, RequestedPrecision> RequestedPrecisionMap;
typedef std::map<SamplerComponent, IndividualEquilibrationCheckResult> EquilibrationResultMap;
typedef std::map<SamplerComponent, IndividualConvergenceCheckResult<BasicStatistics>>
    ConvergenceResultMap;

PYBIND11_MAKE_OPAQUE(StateSamplingFunctionMap);
PYBIND11_MAKE_OPAQUE(jsonStateSamplingFunctionMap);
PYBIND11_MAKE_OPAQUE(SamplerMap);
PYBIND11_MAKE_OPAQUE(RequestedPrecisionMap);
PYBIND11_MAKE_OPAQUE(EquilibrationResultMap);
PYBIND11_MAKE_OPAQUE(ConvergenceResultMap);

namespace {

py::array_t<double> to_numpy(VectorXd const &v) {
  py::array_t<double> a(v.size());
  auto m = a.mutable_unchecked<1>();
  for (Index i = 0; i < v.size(); ++i) m(i) = v(i);
  return a;
}

VectorXd from_sequence(std::vector<double> const &v) { return VectorXd(v); }

// JSON values to and from Python objects (through their text)
py::object json_to_py(jsonParser const &j) {
  std::ostringstream ss;
  j.print(ss, -1);
  return py::module_::import("json").attr("loads")(ss.str());
}
jsonParser py_to_json(py::object const &o) {
  return jsonParser::parse(py::module_::import("json").attr("dumps")(o).cast<std::string>());
}

py::dict to_dict(RequestedPrecision const &r) {
  py::dict d;
  if (r.abs_convergence_is_required) d["abs_precision"] = r.abs_precision;
  if (r.rel_convergence_is_required) d["rel_precision"] = r.rel_precision;
  return d;
}

py::dict to_dict(BasicStatistics const &s) {
  py::dict d;
  d["mean"] = s.mean;
  d["calculated_precision"] = s.calculated_precision;
  return d;
}

// shapes of CompletionCheck_json_io.hh:419-438, ConvergenceCheck_json_io.hh:33-62
py::dict to_dict(CompletionCheckResults<BasicStatistics> const &r) {
  py::dict d;
  d["has_all_minimums_met"] = r.has_all_minimums_met;
  d["has_any_maximum_met"] = r.has_any_maximum_met;
  d["count"] = r.count ? py::cast(*r.count) : py::none();
  d["time"] = r.time ? py::cast(*r.time) : py::none();
  d["clocktime"] = r.clocktime;
  d["n_samples"] = r.n_samples;
  d["is_complete"] = r.is_complete;
  if (r.n_samples_at_convergence_check) {
    d["n_samples_at_convergence_check"] = *r.n_samples_at_convergence_check;
    py::dict eq;
    eq["all_equilibrated"] = r.equilibration_check_results.all_equilibrated;
    eq["N_samples_for_all_to_equilibrate"] =
        r.equilibration_check_results.N_samples_for_all_to_equilibrate;
    py::list eqi;
    for (auto const &p : r.equilibration_check_results.individual_results) {
      py::dict t;
      t["is_equilibrated"] = p.second.is_equilibrated;
      t["N_samples_for_equilibration"] = p.second.N_samples_for_equilibration;
      t["sampler_name"] = p.first.sampler_name;
      t["component_name"] = p.first.component_name;
      t["component_index"] = p.first.component_index;
      eqi.append(t);
    }
    eq["individual_results"] = eqi;
    d["equilibration_check_results"] = eq;
    py::dict cv;
    cv["all_converged"] = r.convergence_check_results.all_converged;
    cv["N_samples_for_statistics"] = r.convergence_check_results.N_samples_for_statistics;
    py::list cvi;
    for (auto const &p : r.convergence_check_results.individual_results) {
      py::dict t;
      t["is_converged"] = p.second.is_converged;
      t["requested_precision"] = to_dict(p.second.requested_precision);
      t["stats"] = to_dict(p.second.stats);
      t["sampler_name"] = p.first.sampler_name;
      t["component_name"] = p.first.component_name;
      t["component_index"] = p.first.component_index;
      cvi.append(t);
    }
    cv["individual_results"] = cvi;
    d["convergence_check_results"] = cv;
  }
  return d;
}

}  // namespace

PYBIND11_MODULE(_libcasm_monte, m) {
  m.doc() = "libcasm.monte surface over the MI355X occupation-Metropolis drop-in";

  // ---- libcasm.monte ------------------------------------------------------------
  py::module_ mm = m.def_submodule("monte");
  py::class_<ValueMap>(mm, "ValueMap")
      .def(py::init<>())
      .def_readwrite("boolean_values", &ValueMap::boolean_values)
      .def_readwrite("scalar_values", &ValueMap::scalar_values)
      .def_property(
          "vector_values",
          [](ValueMap const &v) {
            std::map<std::string, std::vector<double>> out;
            for (auto const &p : v.vector_values)
              out[p.first] = std::vector<double>(p.second.begin(), p.second.end());
            return out;
          },
          [](ValueMap &v, std::map<std::string, std::vector<double>> const &in) {
            v.vector_values.clear();
            for (auto const &p : in) v.vector_values[p.first] = from_sequence(p.second);
          })
      .def_static(
          "from_dict",
          [](py::dict d) {
            // booleans, numbers and number sequences (monte.cpp ValueMap.from_dict)
            ValueMap v;
            for (auto const &item : d) {
              std::string key = py::cast<std::string>(item.first);
              py::handle val = item.second;
              if (py::isinstance<py::bool_>(val))
                v.boolean_values[key] = py::cast<bool>(val);
              else if (py::isinstance<py::float_>(val) || py::isinstance<py::int_>(val))
                v.scalar_values[key] = py::cast<double>(val);
              else
                v.vector_values[key] = from_sequence(py::cast<std::vector<double>>(val));
            }
            return v;
          })
      .def("to_dict", [](ValueMap const &v) {
        py::dict d;
        for (auto const &p : v.boolean_values) d[py::str(p.first)] = p.second;
        for (auto const &p : v.scalar_values) d[py::str(p.first)] = p.second;
        for (auto const &p : v.vector_values)
          d[py::str(p.first)] = std::vector<double>(p.second.begin(), p.second.end());
        return d;
      });

  py::class_<MethodLog>(mm, "MethodLog")
      .def(py::init([](std::optional<std::string> logfile_path,
                       std::optional<double> log_frequency) {
             MethodLog l;
             if (logfile_path) l.logfile_path = *logfile_path;
             l.log_frequency = log_frequency;
             l.reset();
             return l;
           }),
           py::arg("logfile_path") = std::nullopt, py::arg("log_frequency") = std::nullopt)
      .def("logfile_path", [](MethodLog const &l) { return l.logfile_path.string(); })
      .def("log_frequency", [](MethodLog const &l) { return l.log_frequency; });

  py::class_<engine_type, std::shared_ptr<engine_type>>(mm, "RandomNumberEngine")
      .def(py::init<>())
      .def("seed", [](engine_type &e, uint64_t s) { e.seed(s); })
      .def("__call__", [](engine_type &e) { return e(); });

  // python/src/monte.cpp RandomNumberGenerator: random_int / random_real over a
  // shared engine
  py::class_<generator_type>(mm, "RandomNumberGenerator")
      .def(py::init([](std::shared_ptr<engine_type> e) { return generator_type(e); }),
           py::arg("engine") = nullptr)
      .def("random_int", [](generator_type &g, uint64_t m) { return g.random_int(m); },
           py::arg("maximum_value"))
      .def("random_real", [](generator_type &g, double m) { return g.random_real(m); },
           py::arg("maximum_value"))
      .def_property_readonly("engine", [](generator_type const &g) { return g.engine; });

  // ---- libcasm.monte.events ------------------------------------------------------
  py::module_ me = m.def_submodule("events");
  py::class_<Conversions, std::shared_ptr<Conversions>>(me, "Conversions")
      .def(py::init<Index, std::vector<std::string>>(), py::arg("n_sites"),
           py::arg("occupants"))
      .def("l_size", &Conversions::l_size)
      .def("l_to_asym", &Conversions::l_to_asym, py::arg("l"))
      .def("asym_size", &Conversions::asym_size)
      .def("occ_size", &Conversions::occ_size, py::arg("asym"))
      .def("species_index",
           py::overload_cast<Index, Index>(&Conversions::species_index, py::const_),
           py::arg("asym"), py::arg("occ_index"))
      .def("occ_index", &Conversions::occ_index, py::arg("asym"), py::arg("species_index"))
      .def("species_allowed", &Conversions::species_allowed, py::arg("asym"),
           py::arg("species_index"))
      .def("species_size", &Conversions::species_size);
  py::class_<Mol>(me, "Mol")
      .def(py::init<>())
      .def_readwrite("id", &Mol::id)
      .def_readwrite("component_id", &Mol::component)
      .def_readwrite("linear_site_index", &Mol::l)
      .def_readwrite("asymmetric_unit_index", &Mol::asym)
      .def_readwrite("species_index", &Mol::species_index)
      .def_readwrite("mol_location_index", &Mol::loc);
  py::class_<OccTransform>(me, "OccTransform")
      .def(py::init<>())
      .def_readwrite("linear_site_index", &OccTransform::l)
      .def_readwrite("mol_id", &OccTransform::mol_id)
      .def_readwrite("asym", &OccTransform::asym)
      .def_readwrite("from_species", &OccTransform::from_species)
      .def_readwrite("to_species", &OccTransform::to_species);
  py::class_<OccEvent>(me, "OccEvent")
      .def(py::init<>())
      .def_readwrite("linear_site_index", &OccEvent::linear_site_index)
      .def_readwrite("new_occ", &OccEvent::new_occ)
      .def_readwrite("occ_transform", &OccEvent::occ_transform)
      .def("__copy__", [](OccEvent const &e) { return OccEvent(e); });
  py::class_<OccCandidate>(me, "OccCandidate")
      .def(py::init<Index, Index>(), py::arg("asymmetric_unit_index"), py::arg("species_index"))
      .def_readwrite("asymmetric_unit_index", &OccCandidate::asym)
      .def_readwrite("species_index", &OccCandidate::species_index)
      .def("__lt__", [](OccCandidate const &a, OccCandidate const &b) { return a < b; })
      .def("__eq__", [](OccCandidate const &a, OccCandidate const &b) { return a == b; })
      .def("__ne__", [](OccCandidate const &a, OccCandidate const &b) { return a != b; })
      .def("__repr__", [](OccCandidate const &c) {
        return "OccCandidate(" + std::to_string(c.asym) + ", " +
               std::to_string(c.species_index) + ")";
      });
  py::class_<OccSwap>(me, "OccSwap")
      .def(py::init<OccCandidate const &, OccCandidate const &>(), py::arg("first"),
           py::arg("second"))
      .def_readwrite("first", &OccSwap::cand_a)
      .def_readwrite("second", &OccSwap::cand_b)
      .def("reverse", &OccSwap::reverse)
      .def("sort", [](OccSwap &s) { s.sort(); })
      .def("sorted", &OccSwap::sorted)
      .def("__lt__", [](OccSwap const &a, OccSwap const &b) { return a < b; })
      .def("__eq__", [](OccSwap const &a, OccSwap const &b) { return a == b; })
      .def("__ne__", [](OccSwap const &a, OccSwap const &b) { return a != b; });
  py::class_<OccCandidateList>(me, "OccCandidateList")
      .def(py::init<Conversions const &>(), py::arg("convert"))
      .def(py::init<std::vector<OccCandidate>, Conversions const &>(), py::arg("candidates"),
           py::arg("convert"))
      .def("index", py::overload_cast<OccCandidate const &>(&OccCandidateList::index,
                                                            py::const_),
           py::arg("candidate"))
      .def("__getitem__", [](OccCandidateList const &l, Index i) { return l[i]; })
      .def("__len__", [](OccCandidateList const &l) { return l.size(); })
      .def(
          "__iter__",
          [](OccCandidateList const &l) { return py::make_iterator(l.begin(), l.end()); },
          py::keep_alive<0, 1>());
  me.def("is_allowed_canonical_swap", &allowed_canonical_swap, py::arg("convert"),
         py::arg("cand_a"), py::arg("cand_b"));
  me.def("make_canonical_swaps", &make_canonical_swaps, py::arg("convert"),
         py::arg("occ_candidate_list"));
  me.def("is_allowed_semigrand_canonical_swap", &allowed_semigrand_canonical_swap,
         py::arg("convert"), py::arg("cand_a"), py::arg("cand_b"));
  me.def("make_semigrand_canonical_swaps", &make_semigrand_canonical_swaps, py::arg("convert"),
         py::arg("occ_candidate_list"));
  // OccLocation keeps references to its Conversions and OccCandidateList
  py::class_<OccLocation, std::shared_ptr<OccLocation>>(me, "OccLocation")
      .def(py::init<Conversions const &, OccCandidateList const &, bool, bool, bool>(),
           py::arg("convert"), py::arg("occ_candidate_list"), py::arg("update_atoms") = false,
           py::arg("track_unique_atoms") = false, py::arg("save_atom_info") = false,
           py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
      .def("initialize",
           [](OccLocation &ol, std::vector<int> const &occ, std::optional<double> t) {
             ol.initialize(occ, t);
           },
           py::arg("occupation"), py::arg("time") = std::nullopt)
      .def("apply",
           [](OccLocation &ol, OccEvent const &e,
              py::array_t<int, py::array::c_style> occupation, std::optional<double> t) {
             // the occupation array is updated in place, as the reference's
             // Eigen::Ref argument is
             auto buf = occupation.mutable_unchecked<1>();
             std::vector<int> v(buf.shape(0));
             for (py::ssize_t i = 0; i < buf.shape(0); ++i) v[i] = buf(i);
             ol.apply(e, v, t);
             for (py::ssize_t i = 0; i < buf.shape(0); ++i) buf(i) = v[i];
           },
           py::arg("event"), py::arg("occupation"), py::arg("time") = std::nullopt)
      .def("choose_mol_by_candidate_index",
           [](OccLocation const &ol, Index c, generator_type &g) { return ol.choose_mol(c, g); },
           py::arg("candidate_index"), py::arg("random_number_generator"))
      .def("choose_mol",
           [](OccLocation const &ol, OccCandidate const &c, generator_type &g) {
             return ol.choose_mol(c, g);
           },
           py::arg("candidate"), py::arg("random_number_generator"))
      .def("mol_size", &OccLocation::mol_size)
      .def("mol", [](OccLocation const &ol, Index id) { return ol.mol(id); }, py::arg("mol_id"))
      .def("candidate_list", &OccLocation::candidate_list, py::return_value_policy::reference_internal)
      .def("cand_size_by_candidate_index",
           [](OccLocation const &ol, Index c) { return ol.cand_size(c); },
           py::arg("candidate_index"))
      .def("cand_size",
           [](OccLocation const &ol, OccCandidate const &c) { return ol.cand_size(c); },
           py::arg("candidate"))
      .def("mol_id_by_candidate_index",
           [](OccLocation const &ol, Index c, Index loc) { return ol.mol_id(c, loc); },
           py::arg("candidate_index"), py::arg("location_index"))
      .def("mol_id",
           [](OccLocation const &ol, OccCandidate const &c, Index loc) {
             return ol.mol_id(c, loc);
           },
           py::arg("candidate"), py::arg("location_index"))
      .def("linear_site_index_to_mol_id", &OccLocation::l_to_mol_id,
           py::arg("linear_site_index"))
      .def("convert", &OccLocation::convert, py::return_value_policy::reference_internal);
  me.def("choose_canonical_swap",
         [](OccLocation const &ol, std::vector<OccSwap> const &swaps, generator_type &g) {
           return choose_canonical_swap(ol, swaps, g);
         },
         py::arg("occ_location"), py::arg("canonical_swaps"), py::arg("random_number_generator"));
  me.def("propose_canonical_event_from_swap",
         [](OccEvent &e, OccLocation const &ol, OccSwap const &swap, generator_type &g) {
           propose_canonical_event_from_swap(e, ol, swap, g);
         },
         py::arg("event"), py::arg("occ_location"), py::arg("swap"),
         py::arg("random_number_generator"));
  me.def("propose_canonical_event",
         [](OccEvent &e, OccLocation const &ol, std::vector<OccSwap> const &swaps,
            generator_type &g) { propose_canonical_event(e, ol, swaps, g); },
         py::arg("event"), py::arg("occ_location"), py::arg("canonical_swaps"),
         py::arg("random_number_generator"));
  me.def("choose_semigrand_canonical_swap",
         [](OccLocation const &ol, std::vector<OccSwap> const &swaps, generator_type &g) {
           return choose_semigrand_canonical_swap(ol, swaps, g);
         },
         py::arg("occ_location"), py::arg("semigrand_canonical_swaps"),
         py::arg("random_number_generator"));
  me.def("propose_semigrand_canonical_event_from_swap",
         [](OccEvent &e, OccLocation const &ol, OccSwap const &swap, generator_type &g) {
           propose_semigrand_canonical_event_from_swap(e, ol, swap, g);
         },
         py::arg("event"), py::arg("occ_location"), py::arg("swap"),
         py::arg("random_number_generator"));
  me.def("propose_semigrand_canonical_event",
         [](OccEvent &e, OccLocation const &ol, std::vector<OccSwap> const &swaps,
            generator_type &g) { propose_semigrand_canonical_event(e, ol, swaps, g); },
         py::arg("event"), py::arg("occ_location"), py::arg("semigrand_canonical_swaps"),
         py::arg("random_number_generator"));

  // ---- libcasm.monte.sampling ----------------------------------------------------
  py::module_ ms = m.def_submodule("sampling");
  py::class_<StateSamplingFunction>(ms, "StateSamplingFunction")
      .def(py::init([](std::string name, std::string description, std::vector<Index> shape,
                       std::function<std::vector<double>()> f) {
             return StateSamplingFunction(name, description, shape,
                                          [f]() { return from_sequence(f()); });
           }),
           py::arg("name"), py::arg("description"), py::arg("shape"), py::arg("function"))
      .def_readonly("name", &StateSamplingFunction::name)
      .def_readonly("description", &StateSamplingFunction::description)
      .def_readonly("shape", &StateSamplingFunction::shape)
      .def_readonly("component_names", &StateSamplingFunction::component_names)
      .def("__call__", [](StateSamplingFunction const &f) { return to_numpy(f()); });
  py::bind_map<StateSamplingFunctionMap>(ms, "StateSamplingFunctionMap");

  // selected-event statistics parameters (python/src/monte_sampling.cpp:
  // 3267-3371, 4059-4330)
  py::class_<CorrelationsDataParams>(ms, "CorrelationsDataParams")
      .def(py::init([](Index j, Index m, bool o, bool st) {
             return CorrelationsDataParams{j, m, o, st};
           }),
           py::arg("jumps_per_position_sample") = 1, py::arg("max_n_position_samples") = 100,
           py::arg("output_incomplete_samples") = false,
           py::arg("stop_run_when_complete") = false)
      .def_readwrite("jumps_per_position_sample", &CorrelationsDataParams::jumps_per_position_sample)
      .def_readwrite("max_n_position_samples", &CorrelationsDataParams::max_n_position_samples)
      .def_readwrite("output_incomplete_samples", &CorrelationsDataParams::output_incomplete_samples)
      .def_readwrite("stop_run_when_complete", &CorrelationsDataParams::stop_run_when_complete)
      .def("copy", [](CorrelationsDataParams const &p) { return CorrelationsDataParams(p); })
      .def("to_dict",
           [](CorrelationsDataParams const &p) {
             jsonParser j;
             to_json(p, j);
             return json_to_py(j);
           })
      .def_static("from_dict", [](py::dict d) {
        return correlations_data_params_from_json(py_to_json(d));
      });
  py::class_<SelectedEventFunctionParams, std::shared_ptr<SelectedEventFunctionParams>>(
      ms, "SelectedEventFunctionParams")
      .def(py::init<>())
      .def_readwrite("correlations_data_params",
                     &SelectedEventFunctionParams::correlations_data_params)
      .def_readonly("function_names", &SelectedEventFunctionParams::function_names)
      .def(
          "collect_hop_correlations",
          [](SelectedEventFunctionParams &self, Index j, Index m, bool o, bool st) {
            self.correlations_data_params = CorrelationsDataParams{j, m, o, st};
            return self;
          },
          py::arg("jumps_per_position_sample") = 1, py::arg("max_n_position_samples") = 100,
          py::arg("output_incomplete_samples") = false, py::arg("stop_run_when_complete") = false)
      .def("do_not_collect_hop_correlations",
           [](SelectedEventFunctionParams &self) {
             self.correlations_data_params = std::nullopt;
             return self;
           })
      .def("evaluate", &SelectedEventFunctionParams::evaluate, py::arg("name"),
           py::arg("order") = std::nullopt)
      .def("collect", &SelectedEventFunctionParams::collect, py::arg("name"),
           py::arg("tol") = std::nullopt, py::arg("bin_width") = std::nullopt,
           py::arg("initial_begin") = std::nullopt, py::arg("spacing") = std::nullopt,
           py::arg("max_size") = std::nullopt)
      .def("do_not_collect", &SelectedEventFunctionParams::do_not_collect, py::arg("name"))
      .def(
          "get_parameters",
          [](SelectedEventFunctionParams const &self, std::string name) {
            jsonParser j = jsonParser::object();
            if (self.tol.count(name)) j["tol"] = self.tol.at(name);
            if (self.bin_width.count(name)) j["bin_width"] = self.bin_width.at(name);
            if (self.initial_begin.count(name)) j["initial_begin"] = self.initial_begin.at(name);
            if (self.is_log.count(name)) j["spacing"] = self.is_log.at(name) ? "log" : "linear";
            if (self.max_size.count(name)) j["max_size"] = self.max_size.at(name);
            return json_to_py(j);
          },
          py::arg("name"))
      .def("reset", &SelectedEventFunctionParams::reset)
      .def("copy",
           [](SelectedEventFunctionParams const &self) {
             return std::make_shared<SelectedEventFunctionParams>(self);
           })
      .def("to_dict",
           [](SelectedEventFunctionParams const &self) {
             jsonParser j;
             to_json(self, j);
             return json_to_py(j);
           })
      .def_static("from_dict", [](py::dict d) {
        return std::make_shared<SelectedEventFunctionParams>(
            selected_event_function_params_from_json(py_to_json(d)));
      });

  py::class_<jsonStateSamplingFunction>(ms, "jsonStateSamplingFunction")
      .def_readonly("name", &jsonStateSamplingFunction::name)
      .def_readonly("description", &jsonStateSamplingFunction::description)
      .def("__call__", [](jsonStateSamplingFunction const &f) {
        return py::module_::import("json").attr("loads")(f());
      });
  py::bind_map<jsonStateSamplingFunctionMap>(ms, "jsonStateSamplingFunctionMap");

  py::class_<Sampler, std::shared_ptr<Sampler>>(ms, "Sampler")
      .def("n_samples", &Sampler::n_samples)
      .def("n_components", &Sampler::n_components)
      .def("shape", &Sampler::shape)
      .def("component_names", &Sampler::component_names)
      .def("component", [](Sampler const &s, Index i) { return to_numpy(s.component(i)); })
      .def("values", [](Sampler const &s) {
        py::array_t<double> a({(py::ssize_t)s.n_samples(), (py::ssize_t)s.n_components()});
        auto w = a.mutable_unchecked<2>();
        for (Index c = 0; c < s.n_components(); ++c) {
          VectorXd col = s.component(c);
          for (Index r = 0; r < col.size(); ++r) w(r, c) = col(r);
        }
        return a;
      });
  py::bind_map<SamplerMap>(ms, "SamplerMap");
  ms.def("get_n_samples", [](SamplerMap const &s) { return get_n_samples(s); });

  py::class_<SamplerComponent>(ms, "SamplerComponent")
      .def(py::init<std::string, Index, std::string>(), py::arg("sampler_name"),
           py::arg("component_index"), py::arg("component_name"))
      .def_readonly("sampler_name", &SamplerComponent::sampler_name)
      .def_readonly("component_index", &SamplerComponent::component_index)
      .def_readonly("component_name", &SamplerComponent::component_name)
      .def("__eq__", [](SamplerComponent const &a, SamplerComponent const &b) {
        return !(a < b) && !(b < a);
      })
      .def("__hash__", [](SamplerComponent const &a) {
        return py::hash(py::make_tuple(a.sampler_name, a.component_index));
      })
      .def("__repr__", [](SamplerComponent const &a) {
        return "SamplerComponent(" + a.sampler_name + ", " + std::to_string(a.component_index) +
               ", " + a.component_name + ")";
      });

  py::class_<RequestedPrecision>(ms, "RequestedPrecision")
      .def(py::init<>())
      .def_readwrite("abs_convergence_is_required",
                     &RequestedPrecision::abs_convergence_is_required)
      .def_readwrite("abs_precision", &RequestedPrecision::abs_precision)
      .def_readwrite("rel_convergence_is_required",
                     &RequestedPrecision::rel_convergence_is_required)
      .def_readwrite("rel_precision", &RequestedPrecision::rel_precision)
      .def("to_dict", [](RequestedPrecision const &r) { return to_dict(r); });
  py::bind_map<RequestedPrecisionMap>(ms, "RequestedPrecisionMap");

  py::class_<CutoffCheckParams>(ms, "CutoffCheckParams")
      .def(py::init<>())
      .def_readwrite("min_count", &CutoffCheckParams::min_count)
      .def_readwrite("max_count", &CutoffCheckParams::max_count)
      .def_readwrite("min_time", &CutoffCheckParams::min_time)
      .def_readwrite("max_time", &CutoffCheckParams::max_time)
      .def_readwrite("min_sample", &CutoffCheckParams::min_sample)
      .def_readwrite("max_sample", &CutoffCheckParams::max_sample)
      .def_readwrite("min_clocktime", &CutoffCheckParams::min_clocktime)
      .def_readwrite("max_clocktime", &CutoffCheckParams::max_clocktime);

  typedef CompletionCheckParams<BasicStatistics> ccp_type;
  py::class_<ccp_type>(ms, "CompletionCheckParams")
      .def(py::init<>())
      .def_readwrite("cutoff_params", &ccp_type::cutoff_params)
      .def_readwrite("requested_precision", &ccp_type::requested_precision)
      .def_readwrite("log_spacing", &ccp_type::log_spacing)
      .def_readwrite("check_begin", &ccp_type::check_begin)
      .def_readwrite("check_period", &ccp_type::check_period)
      .def_readwrite("check_base", &ccp_type::check_base)
      .def_readwrite("check_shift", &ccp_type::check_shift)
      .def_readwrite("check_period_max", &ccp_type::check_period_max);

  // converge(...).set_precision(name, abs=, rel=)
  // (python/libcasm/monte/sampling/_requested_precision_constructor.py:96)
  struct PyConverge {
    StateSamplingFunctionMap const *functions;
    ccp_type *params;
  };
  py::class_<PyConverge>(ms, "RequestedPrecisionConstructor")
      .def(
          "set_precision",
          [](PyConverge &self, std::string const &name, std::optional<double> abs,
             std::optional<double> rel) -> PyConverge & {
            auto it = self.functions->find(name);
            if (it == self.functions->end())
              throw py::key_error("no sampling function named '" + name + "'");
            RequestedPrecision rp;
            if (abs) {
              rp.abs_convergence_is_required = true;
              rp.abs_precision = *abs;
            }
            if (rel) {
              rp.rel_convergence_is_required = true;
              rp.rel_precision = *rel;
            }
            auto const &f = it->second;
            for (Index i = 0; i < (Index)f.component_names.size(); ++i)
              self.params->requested_precision[SamplerComponent(name, i, f.component_names[i])] =
                  rp;
            return self;
          },
          py::arg("name"), py::arg("abs") = std::nullopt, py::arg("rel") = std::nullopt,
          py::return_value_policy::reference_internal);
  ms.def(
      "converge",
      [](StateSamplingFunctionMap const &functions, ccp_type &params) {
        return PyConverge{&functions, &params};
      },
      py::keep_alive<0, 1>(), py::keep_alive<0, 2>());

  py::class_<BasicStatistics>(ms, "BasicStatistics")
      .def_readonly("mean", &BasicStatistics::mean)
      .def_readonly("calculated_precision", &BasicStatistics::calculated_precision)
      .def("to_dict", [](BasicStatistics const &s) { return to_dict(s); });

  py::class_<IndividualEquilibrationCheckResult>(ms, "IndividualEquilibrationCheckResult")
      .def_readonly("is_equilibrated", &IndividualEquilibrationCheckResult::is_equilibrated)
      .def_readonly("N_samples_for_equilibration",
                    &IndividualEquilibrationCheckResult::N_samples_for_equilibration);
  py::bind_map<EquilibrationResultMap>(ms, "EquilibrationResultMap");
  py::class_<EquilibrationCheckResults>(ms, "EquilibrationCheckResults")
      .def_readonly("all_equilibrated", &EquilibrationCheckResults::all_equilibrated)
      .def_readonly("N_samples_for_all_to_equilibrate",
                    &EquilibrationCheckResults::N_samples_for_all_to_equilibrate)
      .def_readonly("individual_results", &EquilibrationCheckResults::individual_results);

  typedef IndividualConvergenceCheckResult<BasicStatistics> icr_type;
  py::class_<icr_type>(ms, "IndividualConvergenceCheckResult")
      .def_readonly("is_converged", &icr_type::is_converged)
      .def_readonly("requested_precision", &icr_type::requested_precision)
      .def_readonly("stats", &icr_type::stats);
  py::bind_map<ConvergenceResultMap>(ms, "ConvergenceResultMap");
  typedef ConvergenceCheckResults<BasicStatistics> ccr_type;
  py::class_<ccr_type>(ms, "ConvergenceCheckResults")
      .def_readonly("all_converged", &ccr_type::all_converged)
      .def_readonly("N_samples_for_statistics", &ccr_type::N_samples_for_statistics)
      .def_readonly("individual_results", &ccr_type::individual_results);

  typedef CompletionCheckResults<BasicStatistics> cres_type;
  py::class_<cres_type>(ms, "CompletionCheckResults")
      .def_readonly("is_complete", &cres_type::is_complete)
      .def_readonly("count", &cres_type::count)
      .def_readonly("n_samples", &cres_type::n_samples)
      .def_readonly("clocktime", &cres_type::clocktime)
      .def_readonly("has_all_minimums_met", &cres_type::has_all_minimums_met)
      .def_readonly("has_any_maximum_met", &cres_type::has_any_maximum_met)
      .def_readonly("n_samples_at_convergence_check",
                    &cres_type::n_samples_at_convergence_check)
      .def_readonly("equilibration_check_results", &cres_type::equilibration_check_results)
      .def_readonly("convergence_check_results", &cres_type::convergence_check_results)
      .def("to_dict", [](cres_type const &r) { return to_dict(r); });

  py::class_<CompletionCheck<BasicStatistics>>(ms, "CompletionCheck")
      .def("results", &CompletionCheck<BasicStatistics>::results,
           py::return_value_policy::reference_internal)
      .def("params", &CompletionCheck<BasicStatistics>::params,
           py::return_value_policy::reference_internal);

  // ---- libcasm.monte.ising_cpp ------------------------------------------------------
  py::module_ mi = m.def_submodule("ising_cpp");
  py::class_<IsingConfiguration>(mi, "IsingConfiguration")
      .def(py::init([](std::vector<int> shape, int fill_value) {
             return IsingConfiguration(shape, fill_value);
           }),
           py::arg("shape") = std::vector<int>{10, 10}, py::arg("fill_value") = 1)
      .def_readonly("shape", &IsingConfiguration::shape)
      .def_readonly("n_sites", &IsingConfiguration::n_sites)
      .def_readonly("n_variable_sites", &IsingConfiguration::n_variable_sites)
      .def_readonly("n_unitcells", &IsingConfiguration::n_unitcells)
      .def("occ", &IsingConfiguration::occ)
      .def("set_occ", &IsingConfiguration::set_occ)
      .def("occupation",
           [](IsingConfiguration const &c) { return std::vector<int>(c.occupation()); })
      .def("set_occupation", &IsingConfiguration::set_occupation);

  py::class_<IsingState>(mi, "IsingState")
      .def(py::init<IsingConfiguration, ValueMap, ValueMap>(), py::arg("configuration"),
           py::arg("conditions"), py::arg("properties") = ValueMap())
      .def_readwrite("configuration", &IsingState::configuration)
      .def_readwrite("conditions", &IsingState::conditions)
      .def_readwrite("properties", &IsingState::properties);

  py::class_<IsingFormationEnergy>(mi, "IsingFormationEnergy")
      .def(py::init<double, int, bool>(), py::arg("J") = 1.0, py::arg("lattice_type") = 1,
           py::arg("use_nlist") = true)
      .def_readonly("J", &IsingFormationEnergy::J)
      .def_readonly("lattice_type", &IsingFormationEnergy::lattice_type);

  py::class_<IsingParamComposition>(mi, "IsingParamComposition").def(py::init([]() {
    return IsingParamComposition();
  }));

  py::class_<IsingSystem, std::shared_ptr<IsingSystem>>(mi, "IsingSystem")
      .def(py::init<IsingFormationEnergy, IsingParamComposition>(),
           py::arg("formation_energy_calculator"), py::arg("param_composition_calculator"))
      .def_readonly("formation_energy_calculator", &IsingSystem::formation_energy_calculator)
      .def_readonly("param_composition_calculator",
                    &IsingSystem::param_composition_calculator);

  // ---- libcasm.monte.ising_cpp.semigrand_canonical -----------------------------------
  py::module_ mg = mi.def_submodule("semigrand_canonical");
  py::class_<event_generator_type>(mg, "SemiGrandCanonicalEventGenerator").def(py::init<>());

  py::class_<sgc::SemiGrandCanonicalData, std::shared_ptr<sgc::SemiGrandCanonicalData>>(
      mg, "SemiGrandCanonicalData")
      .def_readonly("samplers", &sgc::SemiGrandCanonicalData::samplers)
      .def_readonly("n_pass", &sgc::SemiGrandCanonicalData::n_pass)
      .def_readonly("n_steps_per_pass", &sgc::SemiGrandCanonicalData::n_steps_per_pass)
      .def_readonly("n_accept", &sgc::SemiGrandCanonicalData::n_accept)
      .def_readonly("n_reject", &sgc::SemiGrandCanonicalData::n_reject)
      .def_readonly("completion_check", &sgc::SemiGrandCanonicalData::completion_check)
      .def("acceptance_rate", &sgc::SemiGrandCanonicalData::acceptance_rate)
      .def("rejection_rate", &sgc::SemiGrandCanonicalData::rejection_rate)
      .def_property_readonly("json_samplers", [](sgc::SemiGrandCanonicalData const &d) {
        py::dict out;
        auto loads = py::module_::import("json").attr("loads");
        for (auto const &p : d.json_samplers) {
          py::list l;
          for (auto const &v : p.second.values) l.append(loads(v));
          out[py::str(p.first)] = l;
        }
        return out;
      });

  py::class_<calculator_type, std::shared_ptr<calculator_type>>(mg,
                                                                "SemiGrandCanonicalCalculator")
      .def(py::init<std::shared_ptr<IsingSystem>>(), py::arg("system"))
      .def_readonly("system", &calculator_type::system)
      .def_readonly("data", &calculator_type::data)
      .def("default_sampling_functions",
           [](std::shared_ptr<calculator_type> mc) {
             StateSamplingFunctionMap f;
             for (auto const &s : {sgc::make_parametric_composition_f(mc),
                                   sgc::make_formation_energy_f(mc),
                                   sgc::make_potential_energy_f(mc)})
               f.emplace(s.name, s);
             return f;
           })
      .def("default_json_sampling_functions",
           [](std::shared_ptr<calculator_type> mc) {
             jsonStateSamplingFunctionMap f;
             auto s = sgc::make_configuration_json_f(mc);
             f.emplace(s.name, s);
             return f;
           })
      .def(
          "run",
          [](calculator_type &mc, IsingState &state, StateSamplingFunctionMap const &f,
             jsonStateSamplingFunctionMap const &jf, ccp_type const &ccp,
             event_generator_type const &eg, int sample_period,
             std::optional<MethodLog> method_log, std::shared_ptr<engine_type> engine) {
            py::gil_scoped_release release;
            mc.run(state, f, jf, ccp, eg, sample_period, method_log, engine);
          },
          py::arg("state"), py::arg("sampling_functions"), py::arg("json_sampling_functions"),
          py::arg("completion_check_params"), py::arg("event_generator"),
          py::arg("sample_period") = 1, py::arg("method_log") = std::nullopt,
          py::arg("random_engine") = nullptr);
}
