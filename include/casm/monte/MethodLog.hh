// Drop-in counterparts of libcasm-monte's ValueMap (include/casm/monte/ValueMap.hh:14-19)
// and MethodLog (include/casm/monte/MethodLog.hh:13-38), with the part of
// CASMcode_global's Log the Monte Carlo loop uses (restart_clock, begin_lap,
// lap_time, time_s and a stream to write status records to).
#ifndef CASM_MI355X_monte_MethodLog
#define CASM_MI355X_monte_MethodLog

#include <chrono>
#include <filesystem>
#include <fstream>
#include <iostream>

#include "casm/monte/definitions.hh"

namespace CASM {

/// Wall-clock timers and an output stream for status messages.
class Log {
 public:
  Log() : m_out(&std::cout) { restart_clock(); }
  void restart_clock() { m_t0 = clock::now(); m_lap = m_t0; }
  void begin_lap() { m_lap = clock::now(); }
  double time_s() const { return seconds(clock::now() - m_t0); }
  double lap_time() const { return seconds(clock::now() - m_lap); }
  void reset(std::ostream &out = std::cout) { m_out = &out; }
  template <typename T>
  Log &operator<<(T const &v) {
    (*m_out) << v;
    return *this;
  }
  Log &operator<<(std::ostream &(*f)(std::ostream &)) {
    f(*m_out);
    return *this;
  }

 private:
  typedef std::chrono::steady_clock clock;
  static double seconds(clock::duration d) { return std::chrono::duration<double>(d).count(); }
  std::ostream *m_out;
  clock::time_point m_t0, m_lap;
};

namespace monte {

struct ValueMap {
  std::map<std::string, bool> boolean_values;
  std::map<std::string, double> scalar_values;
  std::map<std::string, VectorXd> vector_values;
};

struct MethodLog {
  std::filesystem::path logfile_path;
  std::shared_ptr<std::ofstream> fout;
  Log log;
  std::optional<double> log_frequency;

  void reset() {
    if (!logfile_path.empty()) {
      if (logfile_path.has_parent_path())
        std::filesystem::create_directories(logfile_path.parent_path());
      fout = std::make_shared<std::ofstream>(logfile_path);
      log.reset(*fout);
    }
  }
  void reset_to_stdout() {
    fout.reset();
    log.reset();
  }
};

}  // namespace monte
}  // namespace CASM

#endif
