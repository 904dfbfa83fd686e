// Drop-in counterpart of libcasm-monte's histogram functions
// (sampling/HistogramFunction.hh): functions evaluated on a selected event,
// whose values are binned by SelectedEventDataCollector.
#ifndef CASM_MI355X_monte_sampling_HistogramFunction
#define CASM_MI355X_monte_sampling_HistogramFunction

#include <algorithm>
#include <functional>
#include <map>
#include <optional>
#include <string>
#include <vector>

#include "casm/monte/definitions.hh"
#include "casm/monte/sampling/Sampler.hh"

namespace CASM {
namespace monte {

/// Order of integer vectors (misc/LexicographicalCompare.hh:10-22)
struct LexicographicalCompare {
  bool operator()(VectorXl const &A, VectorXl const &B) const {
    return std::lexicographical_compare(A.data(), A.data() + A.size(), B.data(),
                                        B.data() + B.size());
  }
};

/// Order of real vectors with a tolerance: elements within tol of each other
/// compare equal (misc/LexicographicalCompare.hh:24-33, with CASM's
/// float_lexicographical_compare)
struct FloatLexicographicalCompare {
  double tol;
  FloatLexicographicalCompare(double _tol) : tol(_tol) {}
  bool operator()(VectorXd const &A, VectorXd const &B) const {
    const Index n = std::min(A.size(), B.size());
    for (Index i = 0; i < n; ++i) {
      if (A(i) < B(i) - tol) return true;
      if (B(i) < A(i) - tol) return false;
    }
    return A.size() < B.size();
  }
};

/// Histogram function of a discrete vector value (HistogramFunction.hh:14-86)
template <typename ValueType, typename CompareType>
class HistogramFunctionT {
 public:
  HistogramFunctionT(std::string _name, std::string _description, std::vector<Index> _shape,
                     std::optional<std::vector<std::string>> _component_names,
                     bool _requires_event_state, std::function<ValueType()> _function,
                     std::function<bool()> _has_value_function, Index _max_size,
                     double _tol = 1e-5)
      : name(_name),
        description(_description),
        shape(_shape),
        component_names(_component_names.has_value() ? *_component_names
                                                     : default_component_names(shape)),
        requires_event_state(_requires_event_state),
        function(_function),
        has_value_function(_has_value_function),
        max_size(_max_size),
        tol(_tol) {
    if (!function) throw std::runtime_error("HistogramFunction: function is empty");
    if (!has_value_function)
      throw std::runtime_error("HistogramFunction: has_value_function is empty");
  }
  std::string name;
  std::string description;
  std::vector<Index> shape;
  std::vector<std::string> component_names;
  bool requires_event_state;
  std::function<ValueType()> function;
  std::function<bool()> has_value_function;
  Index max_size;
  double tol;
  std::optional<std::map<ValueType, std::string, CompareType>> value_labels;
  ValueType operator()() const { return function(); }
  bool has_value() const { return has_value_function(); }
};

typedef HistogramFunctionT<VectorXl, LexicographicalCompare> DiscreteVectorIntHistogramFunction;
typedef HistogramFunctionT<VectorXd, FloatLexicographicalCompare>
    DiscreteVectorFloatHistogramFunction;

/// Function of a selected event binned into one of several 1d histograms
/// (HistogramFunction.hh:96-146)
template <typename ValueType>
class PartitionedHistogramFunction {
 public:
  PartitionedHistogramFunction(std::string _name, std::string _description,
                               bool _requires_event_state, std::function<ValueType()> _function,
                               std::vector<std::string> const &_partition_names,
                               std::function<int()> _get_partition, bool _is_log,
                               double _initial_begin, double _bin_width, Index _max_size)
      : name(_name),
        description(_description),
        requires_event_state(_requires_event_state),
        function(_function),
        partition_names(_partition_names),
        get_partition(_get_partition),
        is_log(_is_log),
        initial_begin(_initial_begin),
        bin_width(_bin_width),
        max_size(_max_size) {
    if (!function) throw std::runtime_error("PartitionedHistogramFunction: function is empty");
    if (!get_partition)
      throw std::runtime_error("PartitionedHistogramFunction: get_partition is empty");
  }
  std::string name;
  std::string description;
  bool requires_event_state;
  std::function<ValueType()> function;
  ValueType operator()() const { return function(); }
  std::vector<std::string> partition_names;
  std::function<int()> get_partition;
  int partition() const { return get_partition(); }
  bool is_log;
  double initial_begin;
  double bin_width;
  Index max_size;
};

}  // namespace monte
}  // namespace CASM

#endif
