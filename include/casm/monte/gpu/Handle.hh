// RAII C++ wrapper of the MI355X library's C ABI (include/casm_mi355x/mcgpu.h),
// used by the drop-in calculators.  Errors become std::runtime_error, the
// reference's error model (definitions.hh:107-113).
#ifndef CASM_MI355X_monte_gpu_Handle
#define CASM_MI355X_monte_gpu_Handle

#include <sstream>

#include "casm/monte/definitions.hh"
#include "casm_mi355x/mcgpu.h"

namespace CASM {
namespace monte {
namespace gpu {

inline void check(int rc) {
  if (rc != MCG_OK) throw std::runtime_error(std::string("MI355X backend: ") + mcg_last_error());
}

/// The 313 words std::mt19937_64's operator<< writes (state words, then the
/// index of the next output); the same layout the device chain keeps.
inline std::vector<uint64_t> engine_words(std::mt19937_64 const &e) {
  std::ostringstream os;
  os << e;
  std::istringstream is(os.str());
  std::vector<uint64_t> w(313);
  for (auto &x : w) is >> x;
  return w;
}

inline void set_engine_words(std::mt19937_64 &e, std::vector<uint64_t> const &w) {
  std::ostringstream os;
  for (size_t i = 0; i < w.size(); ++i) os << (i ? " " : "") << w[i];
  std::istringstream is(os.str());
  is >> e;
  if (!is) throw std::runtime_error("MI355X backend: could not restore the engine state");
}

class Handle {
 public:
  explicit Handle(mcg_desc const &d) { check(mcg_create(&d, &m_h)); }
  ~Handle() {
    if (m_h) mcg_destroy(m_h);
  }
  Handle(Handle const &) = delete;
  Handle &operator=(Handle const &) = delete;

  mcg_handle *get() const { return m_h; }

  void set_occupation(Index chain, std::vector<int8_t> const &occ) {
    check(mcg_set_occupation(m_h, chain, occ.data(), (int64_t)occ.size()));
  }
  std::vector<int8_t> occupation(Index chain, Index n) const {
    std::vector<int8_t> occ(n);
    check(mcg_get_occupation(m_h, chain, occ.data(), n));
    return occ;
  }
  /// CANONICAL: the candidate lists, species-0 sites then species-1 sites;
  /// OCCUPANT: the candidate lists in candidate order (n = mutating sites)
  std::vector<int64_t> occ_location(Index chain, Index n) const {
    std::vector<int64_t> loc(n);
    check(mcg_get_occ_location(m_h, chain, loc.data(), n));
    return loc;
  }
  void set_occ_location(Index chain, std::vector<int64_t> const &loc) {
    check(mcg_set_occ_location(m_h, chain, loc.data(), (int64_t)loc.size()));
  }
  void load_engine(Index chain, std::mt19937_64 const &e) {
    auto w = engine_words(e);
    check(mcg_rng_load(m_h, chain, w.data(), (int64_t)w.size()));
  }
  void store_engine(Index chain, std::mt19937_64 &e) const {
    std::vector<uint64_t> w(313);
    check(mcg_rng_dump(m_h, chain, w.data(), (int64_t)w.size()));
    set_engine_words(e, w);
  }
  void run(CountType n_passes) { check(mcg_run(m_h, n_passes)); }
  /// EXACT, OCCUPANT: exactly n_steps more attempts (may end inside a pass).
  void run_steps(CountType n_steps) { check(mcg_run_steps(m_h, n_steps)); }
  /// As run / run_steps, returning the C ABI's code instead of throwing
  int try_run(CountType n_passes) { return mcg_run(m_h, n_passes); }
  int try_run_steps(CountType n_steps) { return mcg_run_steps(m_h, n_steps); }
  /// KMC: apply the pending event and events up to max_events; an event
  /// reaching stop_time stays pending; select_at_end leaves one more selected
  void kmc_advance(CountType max_events, double stop_time, bool select_at_end) {
    check(mcg_kmc_advance(m_h, max_events, stop_time, select_at_end ? 1 : 0));
  }
  /// KMC: record hop correlations on the device (0 turns recording off)
  void kmc_collect_hops(Index jumps_per_position_sample, Index max_n_position_samples) {
    check(mcg_kmc_collect_hops(m_h, jumps_per_position_sample, max_n_position_samples));
  }
  /// KMC: hop-correlation records of a chain: counts, (dx, dy, step, pass) x
  /// max x n_atoms and times
  void kmc_hops(Index chain, Index n_atoms, Index max_n, std::vector<int64_t> &counts,
                std::vector<int64_t> &records, std::vector<double> &times) const {
    counts.assign(n_atoms, 0);
    records.assign(4 * max_n * n_atoms, 0);
    times.assign(max_n * n_atoms, 0.0);
    check(mcg_kmc_get_hops(m_h, chain, n_atoms, counts.data(), records.data(), times.data()));
  }
  /// KMC: per atom (site, displacement along rows, along columns, jumps)
  std::vector<int64_t> atoms(Index chain, Index n_atoms) const {
    std::vector<int64_t> a(4 * n_atoms);
    check(mcg_get_atoms(m_h, chain, a.data(), n_atoms));
    return a;
  }
  /// OCCUPANT: the asymmetric unit of every site (resets occupation and lists)
  void set_site_asym(Index chain, std::vector<int8_t> const &asym) {
    check(mcg_set_site_asym(m_h, chain, asym.data(), (int64_t)asym.size()));
  }
  /// Pass count, steps into the current pass and FP64 running energy per unit cell.
  void position(Index chain, int64_t &n_pass, int64_t &step, double &e_run) const {
    check(mcg_get_position(m_h, chain, &n_pass, &step, &e_run));
  }
  Index n_samples(Index chain) const {
    int64_t n = 0;
    check(mcg_n_samples(m_h, chain, &n));
    return n;
  }
  std::vector<mcg_sample> samples(Index chain, Index first, Index count) const {
    std::vector<mcg_sample> out(count);
    check(mcg_get_samples(m_h, chain, first, count, out.data()));
    return out;
  }
  void counters(Index chain, int64_t &n_pass, int64_t &n_accept, int64_t &n_reject) const {
    check(mcg_get_counters(m_h, chain, &n_pass, &n_accept, &n_reject));
  }
  void clear_samples() { check(mcg_clear_samples(m_h)); }
  /// NFOLD: simulated time, pending event site (-1: none) and its increment
  void nfold_state(Index chain, double &time, int64_t &pending_site, double &pending_dt) const {
    check(mcg_get_nfold_state(m_h, chain, &time, &pending_site, &pending_dt));
  }

 private:
  mcg_handle *m_h = nullptr;
};

}  // namespace gpu
}  // namespace monte
}  // namespace CASM

#endif
