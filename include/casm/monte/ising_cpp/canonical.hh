// Canonical (fixed-composition) runs of the square-lattice Ising model through
// the reference's generic occupant-event machinery: occupation_metropolis
// (occupation_metropolis.hh:87-153) with propose_canonical_event over an
// OccLocation.
//
// The reference's generic loop drives a configuration whose occupation holds
// occupant indices (OccLocation::apply writes them, OccLocation.cc:260), so a
// canonical Ising run uses a configuration of occupant indices o in {0, 1}
// (species 0 and 1 of a binary site) and a potential that evaluates the
// reference's Ising formation energy on the spins s = 2o - 1:
// per_supercell as model.hh:259-290 and the multi-site delta of
// model.hh:354-379 (each site's single delta, model.hh:305-345, in event
// order with the earlier sites applied).
#ifndef CASM_MI355X_monte_ising_cpp_canonical
#define CASM_MI355X_monte_ising_cpp_canonical

#include "casm/monte/events/OccEventProposal.hh"
#include "casm/monte/ising_cpp/model.hh"
#include "casm/monte/run_management/RunManager.hh"

namespace CASM {
namespace monte {
namespace ising_cpp {

/// Spins s = 2o - 1 of an occupant-index configuration
inline IsingConfiguration spins_of_occupants(IsingConfiguration const &c) {
  IsingConfiguration s(c.shape, 1);
  for (Index l = 0; l < c.n_sites; ++l) s.set_occ(l, 2 * c.occ(l) - 1);
  return s;
}

/// Formation energy of the square-lattice Ising model over occupant indices,
/// optionally minus an exchange potential per species (semi-grand canonical
/// runs over an OccLocation with several asymmetric units): per supercell,
/// E_f - mu[species(l)] for every site l in site order; the delta of an event,
/// per transformed site in event order with the earlier ones applied,
///   dE = (dE + (-J ((2 o' - 1) - s_l) sum_l)) - (mu[to] - mu[from]),
/// which without exchange potentials is the canonical delta of
/// model.hh:354-379 on the spins 2o - 1.
class IsingOccupantPotential {
 public:
  explicit IsingOccupantPotential(double _J, bool _use_nlist = true)
      : J(_J), use_nlist(_use_nlist) {}
  /// with exchange potentials, one per species of `_convert`
  IsingOccupantPotential(double _J, Conversions const &_convert,
                         std::vector<double> _exchange_potential, bool _use_nlist = true)
      : J(_J), use_nlist(_use_nlist), exchange_potential(std::move(_exchange_potential)),
        m_convert(&_convert) {
    if ((Index)exchange_potential.size() != _convert.species_size())
      throw std::runtime_error("IsingOccupantPotential: one exchange potential per species");
  }
  double J;
  bool use_nlist;
  std::vector<double> exchange_potential;  // per species (empty: none)
  Conversions const *convert() const { return m_convert; }

  void set(State<IsingConfiguration> const *_state) {
    m_state = throw_if_null(_state, "Error in IsingOccupantPotential::set: state==nullptr");
  }
  State<IsingConfiguration> const *state() const { return m_state; }

  double per_supercell() const {
    IsingState st(spins_of_occupants(m_state->configuration), ValueMap());
    double E = IsingFormationEnergy(J, 1, use_nlist, &st).per_supercell();
    if (m_convert) {
      auto const &c = m_state->configuration;
      for (Index l = 0; l < c.n_sites; ++l)
        E -= exchange_potential[m_convert->species_index(m_convert->l_to_asym(l), c.occ(l))];
    }
    return E;
  }
  double per_unitcell() const { return per_supercell() / m_state->configuration.n_unitcells; }

  double occ_delta_per_supercell(std::vector<Index> const &linear_site_index,
                                 std::vector<int> const &new_occ) const {
    auto const &c = m_state->configuration;
    std::vector<std::pair<Index, int>> changed;
    auto spin = [&](Index l) {
      for (auto const &p : changed)
        if (p.first == l) return 2 * p.second - 1;
      return 2 * c.occ(l) - 1;
    };
    double dE = 0.0;
    for (size_t k = 0; k < linear_site_index.size(); ++k) {
      const Index l = linear_site_index[k];
      const Index rows = c.shape[0], i = l % rows, j = l / rows;
      const int sum = spin(c.within(i + 1, 0) + rows * j) + spin(i + rows * c.within(j + 1, 1)) +
                      spin(c.within(i - 1, 0) + rows * j) + spin(i + rows * c.within(j - 1, 1));
      if (m_convert) {
        const Index b = m_convert->l_to_asym(l);
        const int o = (spin(l) + 1) / 2;
        dE = (dE + -J * ((2 * new_occ[k] - 1) - spin(l)) * sum) -
             (exchange_potential[m_convert->species_index(b, new_occ[k])] -
              exchange_potential[m_convert->species_index(b, o)]);
      } else {
        dE += -J * ((2 * new_occ[k] - 1) - spin(l)) * sum;
      }
      changed.emplace_back(l, new_occ[k]);
    }
    return dE;
  }

 private:
  State<IsingConfiguration> const *m_state = nullptr;
  Conversions const *m_convert = nullptr;
};

/// The occupation the occupant events read and write
inline std::vector<int> &get_occupation(State<IsingConfiguration> &state) {
  return state.configuration.occupation();
}

/// A binary square-lattice system with one asymmetric unit: Conversions with
/// occupants {species_0, species_1}, its candidate list [(0,0), (0,1)] and the
/// single canonical swap
struct BinaryLatticeEvents {
  BinaryLatticeEvents(Index n_sites, std::string species_0 = "A", std::string species_1 = "B")
      : convert(n_sites, {species_0, species_1}),
        candidate_list(convert),
        canonical_swaps(make_canonical_swaps(convert, candidate_list)),
        semigrand_canonical_swaps(make_semigrand_canonical_swaps(convert, candidate_list)) {}
  Conversions convert;
  OccCandidateList candidate_list;
  std::vector<OccSwap> canonical_swaps;
  std::vector<OccSwap> semigrand_canonical_swaps;
};

}  // namespace ising_cpp
}  // namespace monte
}  // namespace CASM

#endif
