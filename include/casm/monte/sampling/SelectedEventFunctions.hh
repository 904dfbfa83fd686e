// Drop-in counterpart of libcasm-monte's selected-event statistics
// (sampling/SelectedEventFunctions.hh, src/casm/monte/sampling/
// SelectedEventFunctions.cc): histograms of quantities evaluated on each
// selected event of a kinetic run, hop-correlation data (atom positions every
// few jumps), the functions and parameters that choose what is collected, and
// the collector a kinetic loop calls once per selected event.  The
// reference's own loops never call the collector (its callers' calculators
// do); runs store the data in a fixture's results with
// SamplingFixture::set_selected_event_data.
//
// The histogram algorithms follow the reference step for step, because the
// bins, their origin and the out-of-range counts after a sequence of inserts
// are part of the results: a 1d histogram starts its bins at the multiple of
// bin_width from initial_begin just below the first value, prepends bins for
// smaller values and appends bins for larger ones, up to max_size bins; values
// that would need more are counted as out of range.
#ifndef CASM_MI355X_monte_sampling_SelectedEventFunctions
#define CASM_MI355X_monte_sampling_SelectedEventFunctions

#include <cmath>
#include <map>
#include <memory>
#include <optional>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "casm/monte/definitions.hh"
#include "casm/monte/sampling/HistogramFunction.hh"

namespace CASM {
namespace monte {

/// Parameters of hop-correlation collection (SelectedEventFunctions.hh:16-34)
struct CorrelationsDataParams {
  /// every this many jumps of an atom, store its position
  Index jumps_per_position_sample = 1;
  /// at most this many positions per atom
  Index max_n_position_samples = 100;
  /// output matrices with zeros for atoms not yet sampled enough times
  bool output_incomplete_samples = false;
  /// stop the run once every atom has max_n_position_samples positions
  bool stop_run_when_complete = false;
};

/// Hop-correlation data: atom positions at every jumps_per_position_sample-th
/// jump of each atom (SelectedEventFunctions.hh:37-96, .cc:7-66)
struct CorrelationsData {
  Index jumps_per_position_sample = 1;
  CountType max_n_position_samples = 0;
  bool output_incomplete_samples = false;
  bool stop_run_when_complete = false;
  /// positions stored for each atom
  std::vector<CountType> n_position_samples;
  /// position samples every atom has
  CountType n_complete_samples = 0;
  /// when each position was stored: X(position_sample_index, atom_index)
  MatrixXl step;
  MatrixXl pass;
  MatrixXd time;
  /// positions: atom_positions_cart[position_sample_index](x/y/z, atom_index)
  std::vector<MatrixXd> atom_positions_cart;

  void initialize(Index n_atoms, Index _jumps_per_position_sample, Index _max_n_position_samples,
                  bool _output_incomplete_samples) {
    jumps_per_position_sample = _jumps_per_position_sample;
    max_n_position_samples = _max_n_position_samples;
    output_incomplete_samples = _output_incomplete_samples;
    n_position_samples.assign(n_atoms, 0);
    n_complete_samples = 0;
    step = MatrixXl::Zero(_max_n_position_samples, n_atoms);
    pass = MatrixXl::Zero(_max_n_position_samples, n_atoms);
    time = MatrixXd::Zero(_max_n_position_samples, n_atoms);
    atom_positions_cart.assign(_max_n_position_samples, MatrixXd::Zero(3, n_atoms));
  }

  /// Store atom atom_id's position if n_jumps is a multiple of
  /// jumps_per_position_sample and the atom has room left
  void insert(Index atom_id, Index n_jumps, VectorXd const &position_cart, CountType _step,
              CountType _pass, double _time) {
    if (n_jumps % jumps_per_position_sample != 0) return;
    const CountType n = n_position_samples[atom_id];
    if (n >= max_n_position_samples) return;
    step(n, atom_id) = _step;
    pass(n, atom_id) = _pass;
    time(n, atom_id) = _time;
    for (Index k = 0; k < 3; ++k) atom_positions_cart[n](k, atom_id) = position_cart(k);
    ++n_position_samples[atom_id];
    update_n_complete_samples(n);
  }

 private:
  // the atom just sampled had n samples before: if that was the complete
  // count and no atom still has exactly that many, one more sample is complete
  void update_n_complete_samples(CountType n) {
    if (n != n_complete_samples) return;
    for (CountType x : n_position_samples)
      if (x == n_complete_samples) return;
    ++n_complete_samples;
  }
};

/// Histogram of a discrete vector value: counts per distinct value, at most
/// max_size distinct values (SelectedEventFunctions.hh:98-205, .cc:68-202)
template <typename ValueType, typename CompareType>
class DiscreteVectorHistogramT {
 public:
  typedef std::map<ValueType, double, CompareType> map_type;
  typedef std::optional<std::map<ValueType, std::string, CompareType>> labels_type;

  DiscreteVectorHistogramT(std::vector<std::string> const &_component_names,
                           std::vector<Index> _shape, Index _max_size, CompareType compare,
                           labels_type _value_labels)
      : m_shape(_shape),
        m_component_names(_component_names),
        m_max_size(_max_size),
        m_count(compare),
        m_value_labels(_value_labels) {}

  std::vector<Index> const &shape() const { return m_shape; }
  std::vector<std::string> const &component_names() const { return m_component_names; }
  Index size() const { return (Index)m_count.size(); }
  Index max_size() const { return m_max_size; }
  bool max_size_exceeded() const { return m_max_size_exceeded; }

  void insert(ValueType const &value, double weight = 1.0) {
    auto it = m_count.find(value);
    if (it == m_count.end()) {
      if ((Index)m_count.size() == m_max_size) {
        m_max_size_exceeded = true;
        m_out_of_range_count += weight;
        return;
      }
      it = m_count.emplace(value, 0.0).first;
    }
    it->second += weight;
  }

  /// bin counts plus the out-of-range count
  double sum() const {
    double s = m_out_of_range_count;
    for (auto const &x : m_count) s += x.second;
    return s;
  }
  std::vector<ValueType> values() const {
    std::vector<ValueType> v;
    for (auto const &x : m_count) v.push_back(x.first);
    return v;
  }
  std::vector<double> count() const {
    std::vector<double> v;
    for (auto const &x : m_count) v.push_back(x.second);
    return v;
  }
  double out_of_range_count() const { return m_out_of_range_count; }
  std::vector<double> fraction() const {
    std::vector<double> v;
    const double s = sum();
    for (auto const &x : m_count) v.push_back(x.second / s);
    return v;
  }
  labels_type const &value_labels() const { return m_value_labels; }
  map_type const &value_counts() const { return m_count; }

 private:
  std::vector<Index> m_shape;
  std::vector<std::string> m_component_names;
  Index m_max_size;
  bool m_max_size_exceeded = false;
  map_type m_count;
  double m_out_of_range_count = 0.0;
  labels_type m_value_labels;
};

class DiscreteVectorIntHistogram
    : public DiscreteVectorHistogramT<VectorXl, LexicographicalCompare> {
 public:
  DiscreteVectorIntHistogram(std::vector<std::string> const &_component_names,
                             std::vector<Index> _shape, Index _max_size = 10000,
                             labels_type _value_labels = std::nullopt)
      : DiscreteVectorHistogramT(_component_names, _shape, _max_size, LexicographicalCompare(),
                                 _value_labels) {}
};

class DiscreteVectorFloatHistogram
    : public DiscreteVectorHistogramT<VectorXd, FloatLexicographicalCompare> {
 public:
  DiscreteVectorFloatHistogram(std::vector<std::string> const &_component_names,
                               std::vector<Index> _shape, double _tol, Index _max_size = 10000,
                               labels_type _value_labels = std::nullopt)
      : DiscreteVectorHistogramT(_component_names, _shape, _max_size,
                                 FloatLexicographicalCompare(_tol), _value_labels) {}
  double tol() const { return value_counts().key_comp().tol; }
};

/// Histogram of a real value in bins of bin_width, linear or in log10 space
/// (SelectedEventFunctions.hh:290-386, .cc:204-397)
class Histogram1D {
 public:
  Histogram1D(double _initial_begin, double _bin_width, bool _is_log, Index _max_size = 10000)
      : m_initial_begin(_initial_begin),
        m_bin_width(_bin_width),
        m_is_log(_is_log),
        m_max_size(_max_size),
        m_begin(_initial_begin) {}

  double bin_width() const { return m_bin_width; }
  bool is_log() const { return m_is_log; }
  Index size() const { return (Index)m_count.size(); }
  Index max_size() const { return m_max_size; }
  bool max_size_exceeded() const { return m_max_size_exceeded; }
  double begin() const { return m_begin; }
  std::vector<double> const &count() const { return m_count; }
  double out_of_range_count() const { return m_out_of_range_count; }

  /// insert a value (its log10 when the histogram is in log space)
  void insert(double value, double weight = 1.0) {
    bin_value(m_is_log ? std::log10(value) : value, weight);
  }
  /// insert a value already in log10 space
  void insert_log_value(double log_value, double weight = 1.0) {
    if (!m_is_log)
      throw std::runtime_error(
          "Error in Histogram1D::insert_log_value: histogram is not in log space");
    bin_value(log_value, weight);
  }
  /// lower edge of each bin (in log10 space for a log histogram)
  std::vector<double> bin_coords() const {
    std::vector<double> v;
    for (Index i = 0; i < (Index)m_count.size(); ++i) v.push_back(m_begin + i * m_bin_width);
    return v;
  }
  double sum() const {
    double s = m_out_of_range_count;
    for (double x : m_count) s += x;
    return s;
  }
  std::vector<double> density() const {
    std::vector<double> v;
    const double s = sum();
    for (double x : m_count) v.push_back(x / (s * m_bin_width));
    return v;
  }
  /// add another histogram of the same spacing, bin width and initial begin
  void merge(Histogram1D const &other) {
    if (m_is_log != other.m_is_log)
      throw std::runtime_error(
          "Error in Histogram1D::merge: cannot merge histograms with different log settings");
    if (m_bin_width != other.m_bin_width)
      throw std::runtime_error(
          "Error in Histogram1D::merge: cannot merge histograms with different bin_width values");
    if (m_initial_begin != other.m_initial_begin)
      throw std::runtime_error(
          "Error in Histogram1D::merge: cannot merge histograms with different initial_begin "
          "values");
    const std::vector<double> coords = other.bin_coords();
    for (Index i = 0; i < (Index)other.m_count.size(); ++i) {
      if (m_is_log)
        insert_log_value(coords[i], other.m_count[i]);
      else
        insert(coords[i], other.m_count[i]);
    }
    m_out_of_range_count += other.m_out_of_range_count;
  }

 private:
  void bin_value(double value, double weight) {
    if (value < m_begin || m_count.empty()) extend_below(value);
    if (value < m_begin && m_max_size_exceeded) {
      m_out_of_range_count += weight;
      return;
    }
    const double tol = 1e-10;
    const int bin = (int)std::floor((value - m_begin) / m_bin_width + tol);
    while (bin >= (int)m_count.size()) {
      if ((Index)m_count.size() == m_max_size) {
        m_max_size_exceeded = true;
        m_out_of_range_count += weight;
        return;
      }
      m_count.push_back(0.0);
    }
    m_count[bin] += weight;
  }

  // the first value places begin on the bin_width grid just below it; a value
  // below begin prepends empty bins while max_size allows
  void extend_below(double value) {
    if (!std::isfinite(value)) {
      std::stringstream msg;
      msg << "Error in Histogram1D::_reset_bins: value (=" << value << ") is not finite.";
      throw std::runtime_error(msg.str());
    }
    if (m_count.empty()) {
      while (value < m_begin) m_begin -= m_bin_width;
      while (value > m_begin + m_bin_width) m_begin += m_bin_width;
      return;
    }
    Index n_new = 0;
    while (value < m_begin) {
      if (n_new + (Index)m_count.size() == m_max_size) {
        m_max_size_exceeded = true;
        break;
      }
      m_begin -= m_bin_width;
      ++n_new;
    }
    if (n_new == 0) return;
    m_count.insert(m_count.begin(), (size_t)n_new, 0.0);
  }

  double m_initial_begin;
  double m_bin_width;
  bool m_is_log;
  Index m_max_size;
  bool m_max_size_exceeded = false;
  double m_begin;
  std::vector<double> m_count;
  double m_out_of_range_count = 0.0;
};

/// The histograms merged in order (SelectedEventFunctions.cc:399-409)
inline Histogram1D combine(std::vector<Histogram1D> const &histograms) {
  if (histograms.empty())
    throw std::runtime_error("Error in combine: cannot combine empty vector of histograms");
  Histogram1D combined = histograms[0];
  for (size_t i = 1; i < histograms.size(); ++i) combined.merge(histograms[i]);
  return combined;
}

/// One 1d histogram per partition and their combination; asking for the
/// histograms aligns every partition's bins with the combined range
/// (SelectedEventFunctions.hh:393-467, .cc:411-441)
class PartitionedHistogram1D {
 public:
  PartitionedHistogram1D(std::vector<std::string> const &_partition_names, double _initial_begin,
                         double _bin_width, bool _is_log, Index _max_size = 10000)
      : m_partition_names(_partition_names),
        m_histograms(_partition_names.size(),
                     Histogram1D(_initial_begin, _bin_width, _is_log, _max_size)) {}

  std::vector<std::string> const &partition_names() const { return m_partition_names; }
  std::vector<Histogram1D> const &histograms() const {
    if (m_histograms.size() != 1 && !m_up_to_date) make_combined();
    return m_histograms;
  }
  void insert(int partition, double value, double weight = 1.0) {
    m_up_to_date = false;
    if (partition < 0 || partition >= (int)m_histograms.size())
      throw std::runtime_error("Partition index out of range");
    if (!std::isfinite(value)) {
      std::stringstream msg;
      msg << "Error in PartitionedHistogram1D::insert: for partition=\""
          << m_partition_names[partition] << "\" value (=" << value << ") is not finite.";
      throw std::runtime_error(msg.str());
    }
    m_histograms[partition].insert(value, weight);
  }
  Histogram1D const &combined_histogram() const {
    if (m_histograms.size() == 1) return m_histograms.front();
    if (!m_up_to_date) make_combined();
    return *m_combined;
  }

 private:
  void make_combined() const {
    m_combined = combine(m_histograms);
    const std::vector<double> coords = m_combined->bin_coords();
    if (!coords.empty()) {
      for (auto &h : m_histograms) {
        if (m_combined->is_log()) {
          h.insert_log_value(coords.front(), 0.0);
          h.insert_log_value(coords.back(), 0.0);
        } else {
          h.insert(coords.front(), 0.0);
          h.insert(coords.back(), 0.0);
        }
      }
    }
    m_up_to_date = true;
  }

  std::vector<std::string> m_partition_names;
  mutable std::vector<Histogram1D> m_histograms;
  mutable bool m_up_to_date = false;
  mutable std::optional<Histogram1D> m_combined;
};

/// A function evaluated on each selected event for its side effects
/// (SelectedEventFunctions.hh:470-507)
class GenericSelectedEventFunction {
 public:
  GenericSelectedEventFunction(std::string _name, std::string _description,
                               bool _requires_event_state, std::function<void()> _function,
                               std::function<bool()> _has_value_function, Index _order)
      : name(_name),
        description(_description),
        requires_event_state(_requires_event_state),
        function(_function),
        has_value_function(_has_value_function),
        order(_order) {}
  std::string name;
  std::string description;
  bool requires_event_state;
  std::function<void()> function;
  std::function<bool()> has_value_function;
  Index order;
  void operator()() const { function(); }
  bool has_value() const { return has_value_function(); }
};

/// The functions a calculator offers, by name (SelectedEventFunctions.hh:509-540)
struct SelectedEventFunctions {
  std::map<std::string, GenericSelectedEventFunction> generic_functions;
  std::map<std::string, DiscreteVectorIntHistogramFunction> discrete_vector_int_functions;
  std::map<std::string, DiscreteVectorFloatHistogramFunction> discrete_vector_float_functions;
  std::map<std::string, PartitionedHistogramFunction<double>> continuous_1d_functions;
  void insert(GenericSelectedEventFunction f) { generic_functions.emplace(f.name, f); }
  void insert(DiscreteVectorIntHistogramFunction f) {
    discrete_vector_int_functions.emplace(f.name, f);
  }
  void insert(DiscreteVectorFloatHistogramFunction f) {
    discrete_vector_float_functions.emplace(f.name, f);
  }
  void insert(PartitionedHistogramFunction<double> f) {
    continuous_1d_functions.emplace(f.name, f);
  }
  void reset() {
    generic_functions.clear();
    discrete_vector_int_functions.clear();
    discrete_vector_float_functions.clear();
    continuous_1d_functions.clear();
  }
};

/// What to collect, with overrides of the functions' defaults
/// (SelectedEventFunctions.hh:543-631)
struct SelectedEventFunctionParams {
  std::optional<CorrelationsDataParams> correlations_data_params;
  /// histograms to collect and generic functions to evaluate
  std::vector<std::string> function_names;
  std::map<std::string, Index> order;
  std::map<std::string, double> tol;
  std::map<std::string, double> bin_width;
  std::map<std::string, double> initial_begin;
  std::map<std::string, bool> is_log;
  std::map<std::string, Index> max_size;

  SelectedEventFunctionParams &evaluate(std::string name, std::optional<Index> _order) {
    function_names.push_back(name);
    if (_order.has_value()) order[name] = *_order;
    return *this;
  }
  SelectedEventFunctionParams &collect(std::string name, std::optional<double> _tol,
                                       std::optional<double> _bin_width,
                                       std::optional<double> _initial_begin,
                                       std::optional<std::string> spacing,
                                       std::optional<Index> _max_size) {
    function_names.push_back(name);
    if (_tol.has_value()) tol[name] = *_tol;
    if (_bin_width.has_value()) bin_width[name] = *_bin_width;
    if (_initial_begin.has_value()) initial_begin[name] = *_initial_begin;
    if (spacing.has_value()) {
      if (*spacing == "log")
        is_log[name] = true;
      else if (*spacing == "linear")
        is_log[name] = false;
      else
        throw std::runtime_error(
            "Error in SelectedEventFunctionParams::set: spacing must be 'log' or 'linear'");
    }
    if (_max_size.has_value()) max_size[name] = *_max_size;
    return *this;
  }
  SelectedEventFunctionParams &do_not_collect(std::string name) {
    function_names.erase(std::remove(function_names.begin(), function_names.end(), name),
                         function_names.end());
    order.erase(name);
    tol.erase(name);
    bin_width.erase(name);
    initial_begin.erase(name);
    is_log.erase(name);
    max_size.erase(name);
    return *this;
  }
  void reset() {
    correlations_data_params.reset();
    function_names.clear();
    order.clear();
    tol.clear();
    bin_width.clear();
    initial_begin.clear();
    is_log.clear();
    max_size.clear();
  }
};

/// The collected data (SelectedEventFunctions.hh:633-650)
struct SelectedEventData {
  std::optional<CorrelationsData> correlations_data;
  std::map<std::string, DiscreteVectorIntHistogram> discrete_vector_int_histograms;
  std::map<std::string, DiscreteVectorFloatHistogram> discrete_vector_float_histograms;
  std::map<std::string, PartitionedHistogram1D> continuous_1d_histograms;
  void reset() {
    correlations_data.reset();
    discrete_vector_int_histograms.clear();
    discrete_vector_float_histograms.clear();
    continuous_1d_histograms.clear();
  }
};

/// Builds the requested histograms and, once per selected event, evaluates
/// the generic functions (in order of their `order`, then name) and bins the
/// histogram functions' values with weight 1 (SelectedEventFunctions.hh:
/// 654-711, .cc:454-732)
struct SelectedEventDataCollector {
  SelectedEventDataCollector(SelectedEventFunctions const &functions,
                             SelectedEventFunctionParams const &params,
                             std::shared_ptr<SelectedEventData> selected_event_data)
      : data(selected_event_data) {
    if (!data)
      throw std::runtime_error("Error in SelectedEventDataCollector: null selected_event_data");
    data->reset();
    std::map<std::pair<Index, std::string>, GenericSelectedEventFunction> ordered;
    for (std::string const &name : params.function_names) {
      if (functions.generic_functions.count(name)) {
        auto f = functions.generic_functions.at(name);
        if (params.order.count(name)) f.order = params.order.at(name);
        ordered.emplace(std::make_pair(f.order, f.name), f);
        requires_event_state = requires_event_state || f.requires_event_state;
      } else if (functions.discrete_vector_int_functions.count(name)) {
        discrete_vector_int_f.push_back(functions.discrete_vector_int_functions.at(name));
        auto &f = discrete_vector_int_f.back();
        requires_event_state = requires_event_state || f.requires_event_state;
        if (params.max_size.count(name)) f.max_size = params.max_size.at(name);
        data->discrete_vector_int_histograms.emplace(
            name, DiscreteVectorIntHistogram(f.component_names, f.shape, f.max_size,
                                             f.value_labels));
        discrete_vector_int_hist.push_back(&data->discrete_vector_int_histograms.at(name));
      } else if (functions.discrete_vector_float_functions.count(name)) {
        discrete_vector_float_f.push_back(functions.discrete_vector_float_functions.at(name));
        auto &f = discrete_vector_float_f.back();
        requires_event_state = requires_event_state || f.requires_event_state;
        if (params.tol.count(name)) f.tol = params.tol.at(name);
        if (params.max_size.count(name)) f.max_size = params.max_size.at(name);
        data->discrete_vector_float_histograms.emplace(
            name, DiscreteVectorFloatHistogram(f.component_names, f.shape, f.tol, f.max_size,
                                               f.value_labels));
        discrete_vector_float_hist.push_back(&data->discrete_vector_float_histograms.at(name));
      } else if (functions.continuous_1d_functions.count(name)) {
        continuous_1d_f.push_back(functions.continuous_1d_functions.at(name));
        auto &f = continuous_1d_f.back();
        requires_event_state = requires_event_state || f.requires_event_state;
        if (params.initial_begin.count(name)) f.initial_begin = params.initial_begin.at(name);
        if (params.bin_width.count(name)) f.bin_width = params.bin_width.at(name);
        if (params.is_log.count(name)) f.is_log = params.is_log.at(name);
        if (params.max_size.count(name)) f.max_size = params.max_size.at(name);
        data->continuous_1d_histograms.emplace(
            name, PartitionedHistogram1D(f.partition_names, f.initial_begin, f.bin_width,
                                         f.is_log, f.max_size));
        continuous_1d_hist.push_back(&data->continuous_1d_histograms.at(name));
      } else {
        throw std::runtime_error(
            "Error in SelectedEventDataCollector: unknown quantity in "
            "selected_event_data_params: " +
            name);
      }
    }
    for (auto const &pr : ordered) generic_f.push_back(pr.second);
  }

  std::shared_ptr<SelectedEventData> data;
  bool requires_event_state = false;
  std::vector<DiscreteVectorIntHistogramFunction> discrete_vector_int_f;
  std::vector<DiscreteVectorIntHistogram *> discrete_vector_int_hist;
  std::vector<DiscreteVectorFloatHistogramFunction> discrete_vector_float_f;
  std::vector<DiscreteVectorFloatHistogram *> discrete_vector_float_hist;
  std::vector<PartitionedHistogramFunction<double>> continuous_1d_f;
  std::vector<PartitionedHistogram1D *> continuous_1d_hist;
  std::vector<GenericSelectedEventFunction> generic_f;

  void collect_vector_int_data() {
    for (size_t i = 0; i < discrete_vector_int_f.size(); ++i)
      if (discrete_vector_int_f[i].has_value())
        discrete_vector_int_hist[i]->insert(discrete_vector_int_f[i].function(), 1.0);
  }
  void collect_vector_float_data() {
    for (size_t i = 0; i < discrete_vector_float_f.size(); ++i)
      if (discrete_vector_float_f[i].has_value())
        discrete_vector_float_hist[i]->insert(discrete_vector_float_f[i].function(), 1.0);
  }
  void collect_continuous_1d_data() {
    for (size_t i = 0; i < continuous_1d_f.size(); ++i) {
      auto const &f = continuous_1d_f[i];
      const int partition = f.partition();
      const double value = f.function();
      if (partition < 0 || partition >= (int)f.partition_names.size()) {
        std::stringstream msg;
        msg << "Error in PartitionedHistogram1D::insert: function (name=\"" << f.name
            << "\") returned a partition index (=" << partition << ") that is out of range.";
        throw std::runtime_error(msg.str());
      }
      if (f.is_log) {
        const double log_value = std::log10(value);
        if (!std::isfinite(log_value)) {
          std::stringstream msg;
          msg << "Error in PartitionedHistogram1D::insert: function (name=\"" << f.name
              << "\", partition=\"" << f.partition_names[partition] << "\") returned a value (="
              << value << ") with log10(value) (=" << log_value << ") that is not finite.";
          throw std::runtime_error(msg.str());
        }
      } else if (!std::isfinite(value)) {
        std::stringstream msg;
        msg << "Error in PartitionedHistogram1D::insert: function (name=\"" << f.name
            << "\", partition=\"" << f.partition_names[partition] << "\")  returned a value (="
            << value << ") that is not finite.";
        throw std::runtime_error(msg.str());
      }
      continuous_1d_hist[i]->insert(partition, value, 1.0);
    }
  }
  void evaluate_generic_functions() {
    for (auto const &f : generic_f) f.function();
  }
  /// once per selected event
  void collect() {
    evaluate_generic_functions();
    collect_vector_int_data();
    collect_vector_float_data();
    collect_continuous_1d_data();
  }
};

}  // namespace monte
}  // namespace CASM

#endif
