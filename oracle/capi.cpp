// capi.cpp -- extern "C" entry points of the oracle (TEST INFRASTRUCTURE ONLY).
//
// Built into oracle/lib/liboracle.so by oracle/Makefile and loaded with ctypes
// by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Every
// entry point wraps a restatement in refcpu.hh; see that file for the
// reference file:line each one follows.
#include <cstdio>
#include <cstring>
#include <sstream>

#include "refcpu.hh"

using namespace refcpu;

extern "C" {

// All fields are 8 bytes wide so the layout has no padding (mirrored in
// tests/oracle.py with ctypes).
struct orc_sgc_params {
  int64_t rows, cols;
  double J;
  int64_t use_nlist;
  double temperature, mu;
  uint64_t seed;
  int64_t sample_period;
  const int32_t *init_occ;  // NULL -> all +1
  int64_t min_count, min_sample, max_count, max_sample;  // -1: unset
  int64_t check_begin, check_period, log_spacing;
  double check_base, check_shift;
  int64_t check_period_max;
  double prec_formation_energy, prec_param_composition, prec_potential_energy;  // <=0: unset
  int64_t trace_len;
};

struct orc_sgc_output {
  int64_t capacity;  // capacity of the per-sample arrays
  double *formation_energy, *param_composition, *potential_energy;
  int64_t *S, *M;
  int32_t *final_occ;   // rows*cols
  int64_t *trace;       // trace_len*5
  uint64_t *engine;     // 313 words: state[312], position
  int64_t n_samples, n_pass, n_accept, n_reject, n_checks;
  int64_t is_complete, all_equilibrated, N_eq, all_converged, N_stats;
  double mean[3], prec[3];
  int64_t converged[3];
};

static void engine_words(std::mt19937_64 const &e, uint64_t *out) {
  std::ostringstream os;
  os << e;
  std::istringstream is(os.str());
  for (int i = 0; i < 313; ++i) is >> out[i];
}

int orc_sgc_run(const orc_sgc_params *p, orc_sgc_output *o) {
  try {
    IsingConfiguration cfg(p->rows, p->cols, 1);
    if (p->init_occ)
      for (long l = 0; l < cfg.n_sites; ++l) cfg.occupation[l] = p->init_occ[l];
    CompletionCheckParams ccp;
    if (p->min_count >= 0) ccp.cutoff_params.min_count = p->min_count;
    if (p->min_sample >= 0) ccp.cutoff_params.min_sample = p->min_sample;
    if (p->max_count >= 0) ccp.cutoff_params.max_count = p->max_count;
    if (p->max_sample >= 0) ccp.cutoff_params.max_sample = p->max_sample;
    ccp.check_begin = p->check_begin;
    ccp.check_period = p->check_period;
    ccp.log_spacing = p->log_spacing != 0;
    ccp.check_base = p->check_base;
    ccp.check_shift = p->check_shift;
    ccp.check_period_max = p->check_period_max;
    const char *names[3] = {"formation_energy", "param_composition", "potential_energy"};
    double precs[3] = {p->prec_formation_energy, p->prec_param_composition,
                       p->prec_potential_energy};
    for (int k = 0; k < 3; ++k)
      if (precs[k] > 0) {
        RequestedPrecision rp;
        rp.abs_convergence_is_required = true;
        rp.abs_precision = precs[k];
        ccp.requested_precision[SamplerComponent{names[k], 0}] = rp;
      }
    auto eng = std::make_shared<std::mt19937_64>();
    eng->seed(p->seed);
    SgcRunOutput r = sgc_basic_run(cfg, p->J, p->use_nlist != 0, p->temperature, p->mu, eng,
                                   ccp, (int)p->sample_period, (size_t)p->trace_len);
    int64_t ns = get_n_samples(r.samplers);
    o->n_samples = ns;
    if (ns > o->capacity) return 2;
    auto fe = r.samplers.at("formation_energy")->component(0);
    auto pc = r.samplers.at("param_composition")->component(0);
    auto pe = r.samplers.at("potential_energy")->component(0);
    for (int64_t i = 0; i < ns; ++i) {
      if (o->formation_energy) o->formation_energy[i] = fe[i];
      if (o->param_composition) o->param_composition[i] = pc[i];
      if (o->potential_energy) o->potential_energy[i] = pe[i];
      if (o->S) o->S[i] = r.S[i];
      if (o->M) o->M[i] = r.M[i];
    }
    if (o->final_occ)
      for (long l = 0; l < cfg.n_sites; ++l) o->final_occ[l] = r.final_config.occupation[l];
    if (o->trace)
      for (size_t t = 0; t < r.trace.size(); ++t) {
        o->trace[5 * t + 0] = r.trace[t].site;
        o->trace[5 * t + 1] = r.trace[t].old_occ;
        o->trace[5 * t + 2] = r.trace[t].nbr_sum;
        o->trace[5 * t + 3] = r.trace[t].drew;
        o->trace[5 * t + 4] = r.trace[t].accepted;
      }
    if (o->engine) engine_words(*eng, o->engine);
    o->n_pass = r.n_pass;
    o->n_accept = r.n_accept;
    o->n_reject = r.n_reject;
    o->n_checks = r.n_checks;
    o->is_complete = r.results.is_complete;
    o->all_equilibrated = r.results.equilibration_check_results.all_equilibrated;
    o->N_eq = r.results.equilibration_check_results.N_samples_for_all_to_equilibrate;
    o->all_converged = r.results.convergence_check_results.all_converged;
    o->N_stats = r.results.convergence_check_results.N_samples_for_statistics;
    for (int k = 0; k < 3; ++k) {
      o->mean[k] = 0.0;
      o->prec[k] = -1.0;
      o->converged[k] = -1;
      auto it = r.results.convergence_check_results.individual_results.find(
          SamplerComponent{names[k], 0});
      if (it != r.results.convergence_check_results.individual_results.end()) {
        o->mean[k] = it->second.stats.mean;
        o->prec[k] = it->second.stats.calculated_precision;
        o->converged[k] = it->second.is_converged;
      }
    }
    return 0;
  } catch (std::exception const &) {
    return 1;
  }
}

// First n outputs of std::mt19937_64 seeded with `seed`.
void orc_mt64_outputs(uint64_t seed, int64_t n, uint64_t *out) {
  std::mt19937_64 e(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = e();
}

// Engine state words (313: state[312], position) after `n_discard` outputs.
void orc_mt64_state(uint64_t seed, int64_t n_discard, uint64_t *out) {
  std::mt19937_64 e(seed);
  e.discard(n_discard);
  engine_words(e, out);
}

// An engine that replays caller-given 64-bit words, to drive the libstdc++
// distributions with chosen inputs.
struct ReplayEngine {
  typedef uint64_t result_type;
  const uint64_t *xs;
  int64_t n, pos;
  static constexpr uint64_t min() { return 0; }
  static constexpr uint64_t max() { return ~uint64_t(0); }
  uint64_t operator()() { return pos < n ? xs[pos++] : 0; }
};

// std::uniform_int_distribution<long>(0, maximum)(engine) on replayed words;
// returns the value, *consumed = words used (1 unless Lemire rejected).
long orc_uniform_int_replay(const uint64_t *xs, int64_t n, long maximum, int64_t *consumed) {
  ReplayEngine e{xs, n, 0};
  long v = std::uniform_int_distribution<long>(0, maximum)(e);
  *consumed = e.pos;
  return v;
}

// n successive stochastic_count_step(rate) calls on an engine seeded with
// `seed`; dn[i] = the counts drawn, *next_output = the engine's next output
// after them.
void orc_stochastic_count_steps(uint64_t seed, double rate, int64_t n, int64_t *dn,
                                uint64_t *next_output) {
  RandomNumberGenerator rng(std::make_shared<std::mt19937_64>(seed));
  for (int64_t i = 0; i < n; ++i) dn[i] = stochastic_count_step(rate, rng);
  *next_output = (*rng.engine)();
}

// std::uniform_real_distribution<double>(0, 1)(engine) on one replayed word.
double orc_canonical(uint64_t x) {
  ReplayEngine e{&x, 1, 0};
  return std::uniform_real_distribution<double>(0.0, 1.0)(e);
}

// Metropolis acceptance probability std::exp(-dE * beta), beta = 1/(KB*T)
// (basic_occupation_metropolis.hh:361, metropolis.hh:33).
double orc_accept_prob(double dE, double temperature) {
  double beta = 1.0 / (KB * temperature);
  return std::exp(-dE * beta);
}

double orc_kb() { return KB; }

// Statistics restatements (BasicStatistics.cc, EquilibrationCheck.cc).
void orc_basic_statistics(const double *x, int64_t n, double confidence, double *mean,
                          double *prec) {
  std::vector<double> v(x, x + n);
  BasicStatistics s = basic_statistics(v, confidence);
  *mean = s.mean;
  *prec = s.calculated_precision;
}

// 0 on success, 1 on an error (empty or mismatched input, bad method)
int orc_weighted_statistics(const double *x, const double *w, int64_t n, double confidence,
                            int64_t method, int64_t n_resamples, double *out4) {
  try {
    WeightedStatistics s = weighted_statistics(std::vector<double>(x, x + n),
                                               std::vector<double>(w, w + n), confidence,
                                               (int)method, n_resamples);
    out4[0] = s.mean;
    out4[1] = s.calculated_precision;
    out4[2] = s.variance;
    out4[3] = s.autocorr;
    return 0;
  } catch (std::exception const &) {
    return 1;
  }
}

void orc_resample(const double *x, const double *w, int64_t n, double w_sum, int64_t m,
                  double *out) {
  std::vector<double> r = resample(std::vector<double>(x, x + n), std::vector<double>(w, w + n),
                                   w_sum, m);
  std::copy(r.begin(), r.end(), out);
}

double orc_autocorrelation_factor(const double *x, int64_t n) {
  return autocorrelation_factor(std::vector<double>(x, x + n));
}

// autocorrelation_factor with the resampling increment of weighted
// observations (BasicStatistics.cc:24-48, rho = 2^(-1/(lag * increment)))
double orc_autocorrelation_factor_inc(const double *x, int64_t n, double increment) {
  return autocorrelation_factor(std::vector<double>(x, x + n), increment);
}

// The loop of python/tests/sampling/test_CompletionCheck.py:40-124 on the
// restated CompletionCheck: min_sample cutoff, absolute precisions for the
// scalar samplers "e" and "v", a count check every step, one sample of each
// every 10 steps with e = e_mean + random_real(e_amp) - e_amp/2 and likewise
// v (both from one RandomNumberGenerator over std::mt19937_64(seed)).
// out_i: n_steps, n_samples, is_complete, all_equilibrated, n_equilibration
// results, all_converged, n_convergence results, N_samples_for_statistics;
// out_d: calculated_precision of e and v, mean of e and v.
void orc_completion_check_replay(uint64_t seed, int64_t min_sample, double e_mean, double e_amp,
                                 double e_abs, double v_mean, double v_amp, double v_abs,
                                 int64_t *out_i, double *out_d) {
  CompletionCheckParams params;
  params.cutoff_params.min_sample = min_sample;
  SamplerComponent e_key{"e", 0}, v_key{"v", 0};
  RequestedPrecision rp_e, rp_v;
  rp_e.abs_convergence_is_required = true;
  rp_e.abs_precision = e_abs;
  rp_v.abs_convergence_is_required = true;
  rp_v.abs_precision = v_abs;
  params.requested_precision[e_key] = rp_e;
  params.requested_precision[v_key] = rp_v;
  CompletionCheck check(params);
  SamplerMap samplers;
  samplers["e"] = std::make_shared<Sampler>(1);
  samplers["v"] = std::make_shared<Sampler>(1);
  RandomNumberGenerator rng(std::make_shared<std::mt19937_64>(seed));
  CountType n_steps = 0;
  while (!check.is_complete(samplers, n_steps)) {
    n_steps += 1;
    const double e = e_mean + rng.random_real(e_amp) - e_amp / 2.0;
    const double v = v_mean + rng.random_real(v_amp) - v_amp / 2.0;
    if (n_steps % 10 == 0) {
      samplers["e"]->push_back({e});
      samplers["v"]->push_back({v});
    }
  }
  CompletionCheckResults const &r = check.results();
  auto const &conv = r.convergence_check_results;
  out_i[0] = n_steps;
  out_i[1] = get_n_samples(samplers);
  out_i[2] = r.is_complete;
  out_i[3] = r.equilibration_check_results.all_equilibrated;
  out_i[4] = (int64_t)r.equilibration_check_results.individual_results.size();
  out_i[5] = conv.all_converged;
  out_i[6] = (int64_t)conv.individual_results.size();
  out_i[7] = conv.N_samples_for_statistics;
  for (int k = 0; k < 2; ++k) {
    auto it = conv.individual_results.find(k == 0 ? e_key : v_key);
    out_d[k] = it == conv.individual_results.end() ? -1.0 : it->second.stats.calculated_precision;
    out_d[2 + k] = it == conv.individual_results.end() ? 0.0 : it->second.stats.mean;
  }
}

void orc_equilibration(const double *x, int64_t n, double prec, int64_t *is_eq,
                       int64_t *n_eq) {
  EquilibrationResult r = default_equilibration_check_impl(std::vector<double>(x, x + n), prec);
  *is_eq = r.is_equilibrated;
  *n_eq = r.N_samples_for_equilibration;
}

double orc_approx_erf_inv(double x) { return approx_erf_inv(x); }

// Ising energies of a given occupation (model.hh per_supercell, both variants)
double orc_ising_energy(int64_t rows, int64_t cols, const int32_t *occ, double J,
                        int64_t use_nlist) {
  IsingConfiguration c(rows, cols, 1);
  for (long l = 0; l < c.n_sites; ++l) c.occupation[l] = occ[l];
  IsingFormationEnergy fe(J, use_nlist != 0);
  fe.set_state(&c);
  return fe.per_supercell();
}

double orc_ising_delta(int64_t rows, int64_t cols, const int32_t *occ, double J,
                       int64_t use_nlist, double mu, int64_t l, int32_t new_occ) {
  IsingConfiguration c(rows, cols, 1);
  for (long i = 0; i < c.n_sites; ++i) c.occupation[i] = occ[i];
  SemiGrandCanonicalPotential pot{IsingFormationEnergy(J, use_nlist != 0)};
  pot.set_state(&c, mu);
  return pot.delta(l, new_occ);
}

double orc_sgc_potential(int64_t rows, int64_t cols, const int32_t *occ, double J,
                         int64_t use_nlist, double mu) {
  IsingConfiguration c(rows, cols, 1);
  for (long i = 0; i < c.n_sites; ++i) c.occupation[i] = occ[i];
  SemiGrandCanonicalPotential pot{IsingFormationEnergy(J, use_nlist != 0)};
  pot.set_state(&c, mu);
  return pot.per_supercell();
}

// Sequential random-site Metropolis on the generalised Ising lattices; writes
// per-pass integer observables (S1, S2, M); the occupation is updated in place.
int orc_lattice_run(int64_t ndim, const int64_t *L, double J1, double J2, double temperature,
                    double mu, uint64_t seed, int64_t n_passes, int8_t *occ, int64_t *S1,
                    int64_t *S2, int64_t *M) {
  LatticeModel m;
  m.ndim = (int)ndim;
  for (int d = 0; d < 3; ++d) m.L[d] = d < ndim ? L[d] : 1;
  m.J1 = J1;
  m.J2 = J2;
  std::vector<int8_t> s(occ, occ + m.n_sites());
  std::mt19937_64 eng(seed);
  std::vector<long long> a, b, c;
  lattice_sequential_run(m, s, temperature, mu, eng, n_passes, a, b, c);
  for (int64_t p = 0; p < n_passes; ++p) {
    S1[p] = a[p];
    S2[p] = b[p];
    M[p] = c[p];
  }
  std::copy(s.begin(), s.end(), occ);
  return 0;
}

void orc_lattice_observables(int64_t ndim, const int64_t *L, const int8_t *occ, int64_t *out3) {
  LatticeModel m;
  m.ndim = (int)ndim;
  for (int d = 0; d < 3; ++d) m.L[d] = d < ndim ? L[d] : 1;
  std::vector<int8_t> s(occ, occ + m.n_sites());
  long long S1, S2, M;
  lattice_observables(m, s, S1, S2, M);
  out3[0] = S1;
  out3[1] = S2;
  out3[2] = M;
}

// Exact SGC Boltzmann averages by enumeration of all 2^N states (N <= 26).
// out[0..5] = <E>, <E^2>, <Nx>, <(Nx)^2>, <E*Nx>, <E_f>, with E the total SGC
// potential E_f - mu*Nx and Nx = (N+M)/2.
int orc_lattice_enumerate(int64_t ndim, const int64_t *L, double J1, double J2,
                          double temperature, double mu, double *out) {
  LatticeModel m;
  m.ndim = (int)ndim;
  for (int d = 0; d < 3; ++d) m.L[d] = d < ndim ? L[d] : 1;
  m.J1 = J1;
  m.J2 = J2;
  long N = m.n_sites();
  if (N > 26) return 1;
  // Enumerate bonds once.
  std::vector<std::pair<long, long>> b1, b2;
  for (long l = 0; l < N; ++l) {
    long i, j, k;
    m.coords(l, i, j, k);
    for (auto const &o : m.nn_forward()) b1.push_back({l, m.index(i + o[0], j + o[1], k + o[2])});
    for (auto const &o : m.nnn_forward()) b2.push_back({l, m.index(i + o[0], j + o[1], k + o[2])});
  }
  double beta = 1.0 / (KB * temperature);
  long double Z = 0, sE = 0, sE2 = 0, sX = 0, sX2 = 0, sEX = 0, sEf = 0;
  std::vector<double> Es(1L << N), Xs(1L << N), Efs(1L << N);
  double Emin = 1e300;
  for (long st = 0; st < (1L << N); ++st) {
    auto spin = [&](long l) { return ((st >> l) & 1) ? 1 : -1; };
    long long S1 = 0, S2 = 0, M = 0;
    for (auto const &b : b1) S1 += spin(b.first) * spin(b.second);
    for (auto const &b : b2) S2 += spin(b.first) * spin(b.second);
    for (long l = 0; l < N; ++l) M += spin(l);
    double Ef = -J1 * S1 - J2 * S2;
    double Nx = (N + M) / 2.0;
    double E = Ef - mu * Nx;
    Es[st] = E;
    Xs[st] = Nx;
    Efs[st] = Ef;
    Emin = std::min(Emin, E);
  }
  for (long st = 0; st < (1L << N); ++st) {
    long double w = std::exp(-(long double)beta * (Es[st] - Emin));
    Z += w;
    sE += w * Es[st];
    sE2 += w * Es[st] * Es[st];
    sX += w * Xs[st];
    sX2 += w * Xs[st] * Xs[st];
    sEX += w * Es[st] * Xs[st];
    sEf += w * Efs[st];
  }
  out[0] = (double)(sE / Z);
  out[1] = (double)(sE2 / Z);
  out[2] = (double)(sX / Z);
  out[3] = (double)(sX2 / Z);
  out[4] = (double)(sEX / Z);
  out[5] = (double)(sEf / Z);
  return 0;
}

// ---- binary FCC cluster expansion --------------------------------------------
static FccModel fcc_model(const int64_t *L, const double *eci) {
  FccModel m;
  for (int d = 0; d < 3; ++d) m.L[d] = L[d];
  if (eci)
    for (int i = 0; i < 6; ++i) m.eci[i] = eci[i];
  return m;
}

// out5 = N1, P1, P2, P3, T
void orc_fcc_counts(const int64_t *L, const int8_t *occ, int64_t *out5) {
  FccModel m = fcc_model(L, nullptr);
  std::vector<int8_t> s(occ, occ + m.n_sites());
  FccCounts c = fcc_counts(m, s);
  out5[0] = c.N1;
  out5[1] = c.P1;
  out5[2] = c.P2;
  out5[3] = c.P3;
  out5[4] = c.T;
}

// out4 = n1, n2, n3, t of site l
void orc_fcc_local(const int64_t *L, const int8_t *occ, int64_t l, int64_t *out4) {
  FccModel m = fcc_model(L, nullptr);
  std::vector<int8_t> s(occ, occ + m.n_sites());
  int n1, n2, n3, t;
  fcc_local(m, s, (long)l, n1, n2, n3, t);
  out4[0] = n1;
  out4[1] = n2;
  out4[2] = n3;
  out4[3] = t;
}

// shell sizes (12, 6, 24) and triangle pairs per site (24)
void orc_fcc_shell_sizes(int64_t *out4) {
  out4[0] = (int64_t)FccModel::shell(2).size();
  out4[1] = (int64_t)FccModel::shell(4).size();
  out4[2] = (int64_t)FccModel::shell(6).size();
  out4[3] = (int64_t)FccModel::triangle_pairs().size();
}

double orc_fcc_energy(const int64_t *L, const double *eci, const int8_t *occ) {
  FccModel m = fcc_model(L, eci);
  std::vector<int8_t> s(occ, occ + m.n_sites());
  return fcc_energy(m, fcc_counts(m, s));
}

// counts: n_passes x 5 (N1, P1, P2, P3, T after each pass); occ in/out
int orc_fcc_run(const int64_t *L, const double *eci, double temperature, double mu,
                uint64_t seed, int64_t n_passes, int8_t *occ, int64_t *counts) {
  FccModel m = fcc_model(L, eci);
  std::vector<int8_t> s(occ, occ + m.n_sites());
  std::mt19937_64 eng(seed);
  std::vector<FccCounts> out;
  fcc_sequential_run(m, s, temperature, mu, eng, n_passes, out);
  for (int64_t p = 0; p < n_passes; ++p) {
    counts[5 * p + 0] = out[p].N1;
    counts[5 * p + 1] = out[p].P1;
    counts[5 * p + 2] = out[p].P2;
    counts[5 * p + 3] = out[p].P3;
    counts[5 * p + 4] = out[p].T;
  }
  std::copy(s.begin(), s.end(), occ);
  return 0;
}

// Canonical pair-swap run (refcpu::canonical_run).  occ: n = rows*cols spins
// (+1/-1) in, final spins out.  engine: 313 words (state[312], position) in and
// out.  loc_in: NULL (OccLocation::initialize from occ) or n mol ids, the
// species-0 list then the species-1 list; loc_out: the final lists in the same
// form; mol_loc_out: n entries, Mol.loc of each site.  e_run and S: n_passes
// entries.  counters: n_accept, n_reject.  use_nlist: the per_supercell
// variant of the initial energy.  Returns 0, or 1 with the message in
// err (256 bytes) when the reference would stop with an exception.
int orc_canonical_run(int64_t rows, int64_t cols, double J, double temperature,
                      int64_t n_passes, int8_t *occ, uint64_t *engine, const int64_t *loc_in,
                      int64_t *loc_out, int64_t *mol_loc_out, double *e_run, int64_t *S,
                      int64_t *counters, double *initial, int64_t use_nlist, char *err) {
  try {
    const int64_t n = rows * cols;
    std::vector<int> o(n);
    for (int64_t l = 0; l < n; ++l) o[l] = occ[l] > 0 ? 1 : 0;
    std::mt19937_64 eng;
    {
      std::ostringstream os;
      for (int i = 0; i < 313; ++i) os << engine[i] << (i < 312 ? " " : "");
      std::istringstream is(os.str());
      is >> eng;
    }
    BinaryOccLocation ol0;
    const BinaryOccLocation *init = nullptr;
    if (loc_in) {
      ol0.initialize(o);
      int64_t n0 = (int64_t)ol0.loc[0].size();
      for (int c = 0; c < 2; ++c) {
        int64_t base = c ? n0 : 0, cnt = c ? n - n0 : n0;
        for (int64_t i = 0; i < cnt; ++i) {
          Index id = loc_in[base + i];
          ol0.loc[c][i] = id;
          ol0.mol[id].loc = i;
        }
      }
      init = &ol0;
    }
    CanonicalRunOutput r =
        canonical_run(rows, cols, o, J, temperature, eng, n_passes, init, use_nlist != 0);
    for (int64_t l = 0; l < n; ++l) occ[l] = r.final_occ[l] ? 1 : -1;
    engine_words(eng, engine);
    int64_t k = 0;
    for (int c = 0; c < 2; ++c)
      for (Index id : r.ol.loc[c]) loc_out[k++] = id;
    for (int64_t l = 0; l < n; ++l) mol_loc_out[l] = r.ol.mol[l].loc;
    for (int64_t p = 0; p < n_passes; ++p) {
      e_run[p] = r.e_run[p];
      S[p] = r.S[p];
    }
    counters[0] = r.n_accept;
    counters[1] = r.n_reject;
    *initial = r.initial;
    return 0;
  } catch (std::exception &ex) {
    std::snprintf(err, 256, "%s", ex.what());
    return 1;
  }
}

// Occupant-event run over an OccLocation with several asymmetric units
// (refcpu::occ_run).  site_asym: n = rows*cols unit indices; asym_species:
// n_asym x 2 species indices (occupant 0, occupant 1; the second ignored
// where occ_size is 1); occ_size: n_asym entries (1 or 2); mu: n_species
// exchange potentials; swaps: n_swaps pairs of candidate indices.  occ: n
// occupant indices in, final ones out.  engine: 313 words in and out.
// loc_in: NULL (OccLocation::initialize) or the candidate lists' mol ids in
// candidate order, each list's length given by the configuration; loc_out:
// the final lists in the same form (n_mut entries).  Per completed pass
// (at most cap entries): e_run, S, M.  counters: n_accept, n_reject, n_steps,
// passes recorded.  Returns 0, or 1 with the message in err (256 bytes) when
// the reference's loop would stop on an exception (the counters still hold
// the steps made) or the input is invalid.
int orc_occ_run(int64_t rows, int64_t cols, const int8_t *site_asym, int64_t n_asym,
                const int64_t *asym_species, const int64_t *occ_size, int64_t n_species,
                const double *mu, const int64_t *swaps, int64_t n_swaps, int64_t canonical,
                double J, double temperature, double n_unitcells, int64_t n_steps, int8_t *occ,
                uint64_t *engine, const int64_t *loc_in, int64_t *loc_out, double *e_run,
                int64_t *S, int64_t *M, int64_t cap, int64_t *counters, double *initial,
                int64_t use_nlist, char *err) {
  try {
    const int64_t n = rows * cols;
    std::vector<int> sa(site_asym, site_asym + n);
    std::vector<std::vector<int>> spec(n_asym);
    for (int64_t b = 0; b < n_asym; ++b)
      for (int64_t o = 0; o < occ_size[b]; ++o) spec[b].push_back((int)asym_species[2 * b + o]);
    OccSystem sys(sa, spec, (int)n_species);
    std::vector<int> o(occ, occ + n);
    std::vector<double> m(mu, mu + n_species);
    std::vector<std::pair<int, int>> sw;
    for (int64_t i = 0; i < n_swaps; ++i) sw.push_back({(int)swaps[2 * i], (int)swaps[2 * i + 1]});
    std::mt19937_64 eng;
    {
      std::ostringstream os;
      for (int i = 0; i < 313; ++i) os << engine[i] << (i < 312 ? " " : "");
      std::istringstream is(os.str());
      is >> eng;
    }
    OccLocationG ol0;
    const OccLocationG *init = nullptr;
    if (loc_in) {
      ol0.initialize(sys, o);
      int64_t k = 0;
      for (auto &v : ol0.loc)
        for (size_t i = 0; i < v.size(); ++i) {
          Index id = loc_in[k++];
          v[i] = id;
          ol0.mol[id].loc = (Index)i;
        }
      init = &ol0;
    }
    OccRunOutput r = occ_run(rows, cols, sys, o, J, m, sw, canonical != 0, temperature,
                             n_unitcells, eng, n_steps, init, use_nlist != 0);
    for (int64_t l = 0; l < n; ++l) occ[l] = (int8_t)r.final_occ[l];
    engine_words(eng, engine);
    int64_t k = 0;
    for (auto const &v : r.ol.loc)
      for (Index id : v) loc_out[k++] = id;
    const int64_t np = std::min<int64_t>((int64_t)r.e_run.size(), cap);
    for (int64_t p = 0; p < np; ++p) {
      e_run[p] = r.e_run[p];
      S[p] = r.S[p];
      M[p] = r.M[p];
    }
    counters[0] = r.n_accept;
    counters[1] = r.n_reject;
    counters[2] = r.n_steps;
    counters[3] = (int64_t)r.e_run.size();
    *initial = r.initial;
    if (!r.error.empty()) {
      std::snprintf(err, 256, "%s", r.error.c_str());
      return 1;
    }
    return 0;
  } catch (std::exception &ex) {
    std::snprintf(err, 256, "%s", ex.what());
    return 1;
  }
}

// Lattice-gas kinetic Monte Carlo (refcpu::kmc_run).  occ: n = rows*cols
// occupant indices (0 vacancy, 1 atom) in, final ones out.  engine: 313 words
// in and out.  loc_out: n entries, the vacancy list then the atom list.
// atoms_out: 4 per atom (site, dx, dy, n_jumps) in atom order.  Per record
// (at most cap; the return value is the number made): events, time, S,
// sum_sq, sum_dx, sum_dy.  scalars: final time; counts: attempts.
int64_t orc_kmc_run(int64_t rows, int64_t cols, double J, double temperature, int64_t n_events,
                    double time_period, int8_t *occ, uint64_t *engine, int64_t *loc_out,
                    int64_t *atoms_out, int64_t cap, int64_t *rec_events, double *rec_time,
                    int64_t *rec_S, int64_t *rec_sq, int64_t *rec_dx, int64_t *rec_dy,
                    double *final_time, int64_t *attempts) {
  const int64_t n = rows * cols;
  std::vector<int> o(occ, occ + n);
  std::mt19937_64 eng;
  {
    std::ostringstream os;
    for (int i = 0; i < 313; ++i) os << engine[i] << (i < 312 ? " " : "");
    std::istringstream is(os.str());
    is >> eng;
  }
  KmcRunOutput r = kmc_run(rows, cols, o, J, temperature, eng, n_events, time_period);
  for (int64_t l = 0; l < n; ++l) occ[l] = (int8_t)r.final_occ[l];
  engine_words(eng, engine);
  int64_t k = 0;
  for (int c = 0; c < 2; ++c)
    for (Index id : r.ol.loc[c]) loc_out[k++] = id;
  for (size_t a = 0; a < r.atoms.size(); ++a) {
    atoms_out[4 * a + 0] = r.atoms[a].site;
    atoms_out[4 * a + 1] = r.atoms[a].dx;
    atoms_out[4 * a + 2] = r.atoms[a].dy;
    atoms_out[4 * a + 3] = r.atoms[a].n_jumps;
  }
  const int64_t nr = std::min<int64_t>(cap, (int64_t)r.records.size());
  for (int64_t i = 0; i < nr; ++i) {
    rec_events[i] = r.records[i].events;
    rec_time[i] = r.records[i].time;
    rec_S[i] = r.records[i].S;
    rec_sq[i] = r.records[i].sum_sq;
    rec_dx[i] = r.records[i].sum_dx;
    rec_dy[i] = r.records[i].sum_dy;
  }
  *final_time = r.time;
  *attempts = r.attempts;
  return (int64_t)r.records.size();
}

// N-fold way run (refcpu::nfold_run).  occ: rows*cols spins in, final spins
// out; engine: 313 words in and out.  Per sample (at most capacity entries;
// returns the number of samples): S, M, time, weight.  arr_out: the class-partitioned site array
// (rows*cols).  scalars: final time, pending dt; pending site in *pending.
int64_t orc_nfold_run(int64_t rows, int64_t cols, double J, double temperature, double mu,
                      int64_t n_passes, int64_t sample_period, double time_period,
                      int64_t capacity, int8_t *occ, uint64_t *engine, int64_t *S, int64_t *M,
                      double *time, double *weight, int64_t *arr_out, double *scalars,
                      int64_t *pending, int64_t sample_start, double *rate) {
  IsingConfiguration cfg(rows, cols, 1);
  for (long l = 0; l < cfg.n_sites; ++l) cfg.occupation[l] = occ[l] > 0 ? 1 : -1;
  std::mt19937_64 eng;
  {
    std::ostringstream os;
    for (int i = 0; i < 313; ++i) os << engine[i] << (i < 312 ? " " : "");
    std::istringstream is(os.str());
    is >> eng;
  }
  NfoldRunOutput r = nfold_run(cfg, J, temperature, mu, eng, n_passes, sample_period,
                               time_period, sample_start != 0);
  for (long l = 0; l < cfg.n_sites; ++l) occ[l] = (int8_t)r.final_occ[l];
  engine_words(eng, engine);
  for (size_t k = 0; k < r.S.size() && (int64_t)k < capacity; ++k) {
    S[k] = r.S[k];
    M[k] = r.M[k];
    time[k] = r.time[k];
    weight[k] = r.weight[k];
    if (rate && k < r.rate.size()) rate[k] = r.rate[k];
  }
  for (long l = 0; l < cfg.n_sites; ++l) arr_out[l] = r.arr[l];
  scalars[0] = r.final_time;
  scalars[1] = r.pending_dt;
  *pending = r.pending_site;
  return (int64_t)r.S.size();
}

}  // extern "C"
