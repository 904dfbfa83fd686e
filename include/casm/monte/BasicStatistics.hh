// Drop-in counterparts of libcasm-monte's BasicStatistics
// (include/casm/monte/BasicStatistics.hh:24-79, src/casm/monte/BasicStatistics.cc:24-188)
// and misc/math.hh:10-72: mean and precision of the mean with the
// Van de Walle & Asta autocorrelation factor, unweighted and weighted.
#ifndef CASM_MI355X_monte_BasicStatistics
#define CASM_MI355X_monte_BasicStatistics

#include <limits>

#include "casm/monte/definitions.hh"

namespace CASM {
namespace monte {

inline double covariance(VectorXd const &x, VectorXd const &y, double mean) {
  Index n = x.size();
  double cov = 0.0;
  for (Index i = 0; i < n; ++i) cov += (x(i) - mean) * (y(i) - mean);
  return cov / n;
}

inline double variance(VectorXd const &x, double x_mean) {
  Index n = x.size();
  double cov = 0.0;
  for (Index i = 0; i < n; ++i) {
    double d = x(i) - x_mean;
    cov += d * d;
  }
  return cov / n;
}

inline double variance(VectorXd const &x) { return variance(x, x.mean()); }

inline double weighted_variance(VectorXd const &x, double x_mean, VectorXd const &w,
                                double w_sum) {
  double cov = 0.0;
  for (Index i = 0; i < x.size(); ++i) {
    double d = x(i) - x_mean;
    cov += w(i) * d * d;
  }
  return cov / w_sum;
}

/// Winitzki's approximation of erf^-1 (a = 0.147), as the reference uses.
inline double approx_erf_inv(double x) {
  const double one = 1.0, PI = 3.141592653589793238463, a = 0.147;
  double sgn = (x < 0.0) ? -one : one;
  double b = std::log((one - x) * (one + x));
  double c = 2.0 / (PI * a) + b * 0.5;
  double d = b / a;
  return sgn * std::sqrt(std::sqrt(c * c - d) - c);
}

struct BasicStatistics {
  BasicStatistics() : mean(0.0), calculated_precision(std::numeric_limits<double>::max()) {}
  double mean;
  double calculated_precision;
};

inline double get_calculated_precision(BasicStatistics const &s) {
  return s.calculated_precision;
}
inline double get_calculated_relative_precision(BasicStatistics const &s) {
  return std::abs(s.calculated_precision / s.mean);
}

/// (1+rho)/(1-rho), rho = 2^(-1/(k*increment)) for the first lag k with
/// |cov(k)/cov(0)| <= 0.5; 1 without variation; DBL_MAX if no lag qualifies.
inline double autocorrelation_factor(VectorXd const &obs, double increment = 1.0) {
  Index N = obs.size();
  double mean = obs.mean();
  double v0 = variance(obs, mean);
  if (std::abs(v0 / mean) < 1e-8 || v0 == 0.0) return 1.0;
  for (CountType i = 1; i < N; ++i) {
    CountType range = N - i;
    double cov = covariance(obs.segment(0, range), obs.segment(i, range), mean);
    if (std::abs(cov / v0) <= 0.5) {
      double rho = std::pow(2.0, (-1.0 / (i * increment)));
      return (1.0 + rho) / (1.0 - rho);
    }
  }
  return std::numeric_limits<double>::max();
}

/// Equally spaced resampling of weighted observations (BasicStatistics.cc:50-73).
inline VectorXd resample(VectorXd const &obs, VectorXd const &w, double w_sum,
                         Index n_equally_spaced) {
  double increment = w_sum / n_equally_spaced;
  VectorXd out(n_equally_spaced);
  Index j = 0;
  double W_j = 0.0;
  for (Index i = 0; i < n_equally_spaced; ++i) {
    double target = i * increment;
    while (W_j + w(j) < target) {
      W_j += w(j);
      ++j;
    }
    out(i) = obs(j);
  }
  return out;
}

struct BasicStatisticsCalculator {
  BasicStatisticsCalculator(double _confidence = 0.95, Index _method = 1,
                            Index _n_resamples = 10000)
      : confidence(_confidence), method(_method), n_resamples(_n_resamples) {}
  double confidence;
  Index method;
  Index n_resamples;

  BasicStatistics operator()(VectorXd const &obs) const {
    if (obs.size() == 0)
      throw std::runtime_error("Error in BasicStatisticsCalculator: observations.size()==0");
    BasicStatistics s;
    s.mean = obs.mean();
    double v0 = variance(obs, s.mean);
    double f = autocorrelation_factor(obs);
    double fc = std::sqrt(2.0) * approx_erf_inv(confidence);
    s.calculated_precision = fc * std::sqrt(f * v0 / obs.size());
    return s;
  }

  BasicStatistics operator()(VectorXd const &obs, VectorXd const &w) const {
    if (obs.size() == 0)
      throw std::runtime_error("Error in BasicStatisticsCalculator: observations.size()==0");
    if (w.size() == 0) return (*this)(obs);
    if (obs.size() != w.size())
      throw std::runtime_error(
          "Error in BasicStatisticsCalculator: observations.size() != sample_weight.size()");
    double W = w.sum();
    double increment = W / n_resamples;
    VectorXd eq = resample(obs, w, W, n_resamples);
    if (method == 1) {
      BasicStatistics s;
      s.mean = obs.dot(w) / W;
      double wv = weighted_variance(obs, s.mean, w, W);
      double f = autocorrelation_factor(eq, increment);
      double fc = std::sqrt(2.0) * approx_erf_inv(confidence);
      s.calculated_precision = fc * std::sqrt(f * wv / W);
      return s;
    } else if (method == 2) {
      return (*this)(eq);
    }
    throw std::runtime_error("Error in BasicStatisticsCalculator: invalid method");
  }
};

template <typename StatisticsType>
using CalcStatisticsFunction =
    std::function<StatisticsType(VectorXd const &observations, VectorXd const &sample_weight)>;

template <typename StatisticsType>
CalcStatisticsFunction<StatisticsType> default_statistics_calculator();

template <>
inline CalcStatisticsFunction<BasicStatistics> default_statistics_calculator<BasicStatistics>() {
  return BasicStatisticsCalculator();
}

}  // namespace monte
}  // namespace CASM

#endif
