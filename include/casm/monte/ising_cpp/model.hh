// Drop-in counterpart of libcasm-monte's Ising model (ising_cpp/model.hh:19-452):
// the host-side configuration, state, formation-energy and parametric-
// composition calculators.  The GPU calculators read their parameters (J,
// lattice type, energy variant) and the configuration from these objects; the
// per-attempt work runs on the device.
#ifndef CASM_MI355X_monte_ising_cpp_model
#define CASM_MI355X_monte_ising_cpp_model

#include "casm/monte/MethodLog.hh"

namespace CASM {
namespace monte {
namespace ising_cpp {

/// 2D periodic supercell; column-major linear index l = shape[0]*col + row.
class IsingConfiguration {
 public:
  IsingConfiguration() : IsingConfiguration({0, 0}, 1) {}
  IsingConfiguration(std::vector<int> _shape, int fill_value = 1) : shape(_shape) {
    if (shape.size() != 2) throw std::runtime_error("IsingConfiguration only supports 2d");
    m_occupation.assign(static_cast<size_t>(shape[0]) * shape[1], fill_value);
    n_sites = n_variable_sites = n_unitcells = static_cast<Index>(m_occupation.size());
  }
  std::vector<int> shape;
  Index n_sites, n_variable_sites, n_unitcells;

  std::vector<int> const &occupation() const { return m_occupation; }
  /// mutable occupation, the form OccLocation::apply updates (through
  /// get_occupation in canonical runs)
  std::vector<int> &occupation() { return m_occupation; }
  void set_occupation(std::vector<int> const &occ) {
    if (occ.size() != m_occupation.size())
      throw std::runtime_error("Error in set_occupation: size mismatch");
    m_occupation = occ;
  }
  int occ(Index l) const { return m_occupation[l]; }
  void set_occ(Index l, int v) { m_occupation[l] = v; }
  Index within(Index index, int dim) const {
    Index r = index % shape[dim];
    if (r < 0) r += shape[dim];
    return r;
  }
  std::vector<int> from_linear_site_index(Index l) const {
    return {static_cast<int>(l % shape[0]), static_cast<int>(l / shape[0])};
  }
  Index to_linear_site_index(Index row, Index col) const { return shape[0] * col + row; }

 private:
  std::vector<int> m_occupation;
};

class IsingState {
 public:
  IsingState(IsingConfiguration _configuration, ValueMap _conditions,
             ValueMap _properties = ValueMap())
      : configuration(_configuration), conditions(_conditions), properties(_properties) {}
  IsingConfiguration configuration;
  ValueMap conditions;
  ValueMap properties;
};

class IsingFormationEnergy {
 public:
  typedef IsingState state_type;
  IsingFormationEnergy(double _J = 1.0, int _lattice_type = 1, bool _use_nlist = true,
                       state_type const *_state = nullptr)
      : J(_J), lattice_type(_lattice_type), state(nullptr), m_use_nlist(_use_nlist) {
    if (lattice_type != 1) throw std::runtime_error("Unsupported lattice_type");
    if (_state != nullptr) set_state(_state);
  }
  double J;
  int lattice_type;
  state_type const *state;
  bool use_nlist() const { return m_use_nlist; }

  void set_state(state_type const *_state) {
    state = throw_if_null(_state, "Error in IsingFormationEnergy::set_state: _state==nullptr");
  }

  /// model.hh:259-290 (both variants)
  double per_supercell() const {
    auto const &c = state->configuration;
    auto const &occ = c.occupation();
    Index rows = c.shape[0], cols = c.shape[1];
    auto site = [&](Index i, Index j) { return occ[i + rows * j]; };
    if (m_use_nlist) {
      double e = 0.0;
      for (Index l = 0; l < c.n_sites; ++l) {
        Index i = l % rows, j = l / rows;
        e += site(i, j) * (site(c.within(i + 1, 0), j) + site(i, c.within(j + 1, 1)));
      }
      e *= -J;
      return e;
    }
    double e = 0.0;
    for (Index i = 0; i < rows; ++i) {
      Index in = c.within(i + 1, 0);
      long dot = 0;
      for (Index j = 0; j < cols; ++j) dot += site(i, j) * site(in, j);
      e += -J * static_cast<double>(dot);
    }
    for (Index j = 0; j < cols; ++j) {
      Index jn = c.within(j + 1, 1);
      long dot = 0;
      for (Index i = 0; i < rows; ++i) dot += site(i, j) * site(i, jn);
      e += -J * static_cast<double>(dot);
    }
    return e;
  }
  double per_unitcell() const { return per_supercell() / state->configuration.n_unitcells; }

  /// model.hh:305-345
  double _single_occ_delta_per_supercell(Index l, int new_occ) const {
    auto const &c = state->configuration;
    Index rows = c.shape[0];
    Index i = l % rows, j = l / rows;
    auto const &occ = c.occupation();
    auto site = [&](Index a, Index b) { return occ[a + rows * b]; };
    return -J * (new_occ - site(i, j)) *
           (site(c.within(i + 1, 0), j) + site(i, c.within(j + 1, 1)) +
            site(c.within(i - 1, 0), j) + site(i, c.within(j - 1, 1)));
  }
  double occ_delta_per_supercell(std::vector<Index> const &lsi,
                                 std::vector<int> const &new_occ) const {
    if (lsi.size() == 1) return _single_occ_delta_per_supercell(lsi[0], new_occ[0]);
    auto &c = const_cast<IsingConfiguration &>(state->configuration);
    double dE = 0.0;
    std::vector<int> orig;
    for (size_t i = 0; i < lsi.size(); ++i) {
      dE += _single_occ_delta_per_supercell(lsi[i], new_occ[i]);
      orig.push_back(c.occ(lsi[i]));
      c.set_occ(lsi[i], new_occ[i]);
    }
    for (size_t i = 0; i < orig.size(); ++i) c.set_occ(lsi[i], orig[i]);
    return dE;
  }

 private:
  bool m_use_nlist = true;
};

class IsingParamComposition {
 public:
  typedef IsingState state_type;
  IsingParamComposition(state_type const *_state = nullptr) : state(nullptr) {
    if (_state != nullptr) set_state(_state);
  }
  state_type const *state;
  void set_state(state_type const *_state) {
    state = throw_if_null(_state, "Error in IsingParamComposition::set_state: _state==nullptr");
  }
  Index n_independent_compositions() const { return 1; }
  VectorXd per_supercell() const {  // model.hh:412-417
    auto const &occ = state->configuration.occupation();
    long sum = 0;
    for (int v : occ) sum += v;
    VectorXd r(1);
    r[0] = static_cast<double>(static_cast<long>(occ.size()) + sum) / 2.0;
    return r;
  }
  VectorXd per_unitcell() const {
    return per_supercell() / static_cast<double>(state->configuration.n_unitcells);
  }
  VectorXd occ_delta_per_supercell(std::vector<Index> const &lsi,
                                   std::vector<int> const &new_occ) const {
    VectorXd Ndx(1);
    Ndx[0] = 0.0;
    for (size_t i = 0; i < lsi.size(); ++i)
      Ndx[0] += (new_occ[i] - state->configuration.occ(lsi[i])) / 2.0;
    return Ndx;
  }
};

class IsingSystem {
 public:
  typedef IsingState state_type;
  typedef IsingFormationEnergy formation_energy_f_type;
  typedef IsingParamComposition param_composition_f_type;
  IsingSystem(formation_energy_f_type fe, param_composition_f_type pc)
      : formation_energy_calculator(fe), param_composition_calculator(pc) {}
  formation_energy_f_type formation_energy_calculator;
  param_composition_f_type param_composition_calculator;
};

}  // namespace ising_cpp
}  // namespace monte
}  // namespace CASM

#endif
