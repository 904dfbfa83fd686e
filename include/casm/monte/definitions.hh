// Drop-in counterpart of libcasm-monte's include/casm/monte/definitions.hh
// (typedefs :25-27, default engine :17, SAMPLE_MODE/METHOD :20-23,
// throw_if_null :107-113), plus a minimal dense vector standing in for
// Eigen::VectorXd (the reference's containers; Eigen is not a dependency here).
#ifndef CASM_MI355X_monte_definitions
#define CASM_MI355X_monte_definitions

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <numeric>
#include <optional>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace CASM {

typedef long Index;

// Boltzmann constant, eV/K (the value of CASMcode_global's KB used by the
// reference's beta = 1/(KB*T)).
constexpr double KB = 8.6173303e-05;

namespace monte {

typedef std::mt19937_64 default_engine_type;

enum class SAMPLE_MODE { BY_STEP, BY_PASS, BY_TIME };
enum class SAMPLE_METHOD { LINEAR, LOG, CUSTOM };

typedef long CountType;
typedef long long BigCountType;
typedef double TimeType;

template <typename PtrType>
PtrType throw_if_null(PtrType ptr, std::string const &what) {
  if (ptr == nullptr) throw std::runtime_error(what);
  return ptr;
}

/// Minimal dense vector of doubles with the Eigen::VectorXd operations the
/// Monte Carlo API uses (size, element access, sum/mean/dot, head/tail/segment,
/// scalar arithmetic).
class VectorXd {
 public:
  VectorXd() = default;
  explicit VectorXd(Index n, double value = 0.0) : m_v(static_cast<size_t>(n), value) {}
  VectorXd(std::initializer_list<double> l) : m_v(l) {}
  explicit VectorXd(std::vector<double> v) : m_v(std::move(v)) {}
  static VectorXd Constant(Index n, double value) { return VectorXd(n, value); }
  static VectorXd Zero(Index n) { return VectorXd(n, 0.0); }

  Index size() const { return static_cast<Index>(m_v.size()); }
  double &operator()(Index i) { return m_v[i]; }
  double operator()(Index i) const { return m_v[i]; }
  double &operator[](Index i) { return m_v[i]; }
  double operator[](Index i) const { return m_v[i]; }
  double *data() { return m_v.data(); }
  double const *data() const { return m_v.data(); }
  std::vector<double> const &vector() const { return m_v; }
  auto begin() const { return m_v.begin(); }
  auto end() const { return m_v.end(); }

  double sum() const {
    double s = 0.0;
    for (double x : m_v) s += x;
    return s;
  }
  double mean() const { return sum() / static_cast<double>(m_v.size()); }
  double dot(VectorXd const &o) const {
    if (o.size() != size()) throw std::runtime_error("VectorXd::dot: size mismatch");
    if (m_v.empty()) return 0.0;
    double s = m_v[0] * o.m_v[0];
    for (size_t i = 1; i < m_v.size(); ++i) s += m_v[i] * o.m_v[i];
    return s;
  }
  VectorXd segment(Index start, Index n) const {
    return VectorXd(std::vector<double>(m_v.begin() + start, m_v.begin() + start + n));
  }
  VectorXd head(Index n) const { return segment(0, n); }
  VectorXd tail(Index n) const { return segment(size() - n, n); }
  VectorXd operator/(double d) const {
    VectorXd r(*this);
    for (double &x : r.m_v) x /= d;
    return r;
  }
  VectorXd operator*(double d) const {
    VectorXd r(*this);
    for (double &x : r.m_v) x *= d;
    return r;
  }
  VectorXd operator+(VectorXd const &o) const {
    VectorXd r(*this);
    for (size_t i = 0; i < r.m_v.size(); ++i) r.m_v[i] += o.m_v[i];
    return r;
  }
  VectorXd operator-(VectorXd const &o) const {
    VectorXd r(*this);
    for (size_t i = 0; i < r.m_v.size(); ++i) r.m_v[i] -= o.m_v[i];
    return r;
  }
  bool operator==(VectorXd const &o) const { return m_v == o.m_v; }

 private:
  std::vector<double> m_v;
};

/// Minimal dense column-major matrix of doubles with the Eigen::MatrixXd
/// operations the kinetic Monte Carlo data uses (atom positions, 3 x n).
class MatrixXd {
 public:
  MatrixXd() = default;
  MatrixXd(Index rows, Index cols, double value = 0.0)
      : m_rows(rows), m_cols(cols), m_v(static_cast<size_t>(rows * cols), value) {}
  static MatrixXd Zero(Index rows, Index cols) { return MatrixXd(rows, cols, 0.0); }
  Index rows() const { return m_rows; }
  Index cols() const { return m_cols; }
  double &operator()(Index i, Index j) { return m_v[static_cast<size_t>(i + m_rows * j)]; }
  double operator()(Index i, Index j) const { return m_v[static_cast<size_t>(i + m_rows * j)]; }
  VectorXd col(Index j) const {
    return VectorXd(std::vector<double>(m_v.begin() + m_rows * j, m_v.begin() + m_rows * (j + 1)));
  }
  bool operator==(MatrixXd const &o) const {
    return m_rows == o.m_rows && m_cols == o.m_cols && m_v == o.m_v;
  }

 private:
  Index m_rows = 0, m_cols = 0;
  std::vector<double> m_v;
};

/// Minimal dense vector of longs (Eigen::VectorXl's role in histogram keys)
class VectorXl {
 public:
  VectorXl() = default;
  explicit VectorXl(Index n, long value = 0) : m_v(static_cast<size_t>(n), value) {}
  VectorXl(std::initializer_list<long> l) : m_v(l) {}
  explicit VectorXl(std::vector<long> v) : m_v(std::move(v)) {}
  Index size() const { return static_cast<Index>(m_v.size()); }
  long &operator()(Index i) { return m_v[i]; }
  long operator()(Index i) const { return m_v[i]; }
  long const *data() const { return m_v.data(); }
  std::vector<long> const &vector() const { return m_v; }
  bool operator==(VectorXl const &o) const { return m_v == o.m_v; }

 private:
  std::vector<long> m_v;
};

/// Minimal dense column-major matrix of longs (Eigen::MatrixXl)
class MatrixXl {
 public:
  MatrixXl() = default;
  MatrixXl(Index rows, Index cols, long value = 0)
      : m_rows(rows), m_cols(cols), m_v(static_cast<size_t>(rows * cols), value) {}
  static MatrixXl Zero(Index rows, Index cols) { return MatrixXl(rows, cols, 0); }
  Index rows() const { return m_rows; }
  Index cols() const { return m_cols; }
  long &operator()(Index i, Index j) { return m_v[static_cast<size_t>(i + m_rows * j)]; }
  long operator()(Index i, Index j) const { return m_v[static_cast<size_t>(i + m_rows * j)]; }

 private:
  Index m_rows = 0, m_cols = 0;
  std::vector<long> m_v;
};

typedef std::vector<Index> Shape;

}  // namespace monte
}  // namespace CASM

#endif
