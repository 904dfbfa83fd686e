// The square lattice gas of the kinetic Monte Carlo drop-in: its hop events
// and its event selector.  The reference's kinetic_monte_carlo
// (methods/kinetic_monte_carlo.hh:136-229) takes the event selector and the
// events from the caller (clexmonte provides them); this is the model the
// MI355X runs (MCG_MODE_KMC):
//   occupant index 0 = vacancy ("Va"), 1 = atom; E = -J sum over nearest-
//   neighbour bonds of n_i n_j;
//   a hop moves the atom at `from` one site in direction `dir` (0: +row,
//   1: -row, 2: +column, 3: -column) onto a vacancy; its rate is
//   min(1, exp(-dE / kT)), dE = J (a - b) (a: atoms next to the origin other
//   than at the target, b: atoms next to the target other than at the
//   origin);
//   events are selected by rejection kinetic Monte Carlo: attempts at 1/R each
//   (R = 4 x atoms), an atom uniform over the OccLocation's atom list, a
//   direction random_int(3), rejected when the target holds an atom, accepted
//   without a draw when dE <= 0, else when random_real(1.0) < exp(-beta dE);
//   the selected event's time increment is fl(k / R) for the k attempts it
//   took.
#ifndef CASM_MI355X_monte_ising_cpp_kmc
#define CASM_MI355X_monte_ising_cpp_kmc

#include <cmath>
#include <memory>

#include "casm/monte/events/OccLocation.hh"
#include "casm/monte/ising_cpp/model.hh"
#include "casm/monte/run_management/RunManager.hh"
#include "casm/monte/sampling/SelectedEventFunctions.hh"

namespace CASM {
namespace monte {
namespace ising_cpp {

/// The event id of a hop: origin site and direction
struct LatticeGasHop {
  Index from = 0;
  int dir = 0;
};

/// Site next to l in direction dir on a rows x cols torus
inline Index lattice_gas_neighbour(Index l, int dir, Index rows, Index cols) {
  const Index i = l % rows, j = l / rows;
  switch (dir) {
    case 0: return (i + 1) % rows + rows * j;
    case 1: return (i + rows - 1) % rows + rows * j;
    case 2: return i + rows * ((j + 1) % cols);
    default: return i + rows * ((j + cols - 1) % cols);
  }
}

/// The lattice gas's event selector: the coupling, the temperature and the
/// engine the selection draws from.  kinetic_monte_carlo() runs its selection
/// on the MI355X; total_rate() and get_event() serve sampling functions and
/// callers that apply events themselves.
class LatticeGasKmcEventSelector {
 public:
  LatticeGasKmcEventSelector(double _J, double _temperature,
                             std::shared_ptr<std::mt19937_64> _engine,
                             State<IsingConfiguration> const *_state,
                             OccLocation const *_occ_location)
      : J(_J),
        temperature(_temperature),
        engine(throw_if_null(_engine, "LatticeGasKmcEventSelector: engine==nullptr")),
        m_state(throw_if_null(_state, "LatticeGasKmcEventSelector: state==nullptr")),
        m_occ_location(throw_if_null(_occ_location, "LatticeGasKmcEventSelector: occ_location==nullptr")) {}

  double J;
  double temperature;
  std::shared_ptr<std::mt19937_64> engine;

  /// Hop correlations to collect during kinetic_monte_carlo(), on the device
  /// (the role of SelectedEventFunctionParams::correlations_data_params); at
  /// the end of the run the data is in selected_event_data->correlations_data
  /// (stop_run_when_complete is not supported)
  std::optional<CorrelationsDataParams> correlations_data_params;
  std::shared_ptr<SelectedEventData> selected_event_data;

  /// Rate of one hop in the current configuration (0 if the target holds an atom)
  double rate(LatticeGasHop h) const {
    auto const &c = m_state->configuration;
    const Index rows = c.shape[0], cols = c.shape[1];
    const Index to = lattice_gas_neighbour(h.from, h.dir, rows, cols);
    if (c.occ(h.from) != 1 || c.occ(to) != 0) return 0.0;
    int a = 0, b = 0;
    for (int e = 0; e < 4; ++e) {
      const Index x = lattice_gas_neighbour(h.from, e, rows, cols);
      const Index y = lattice_gas_neighbour(to, e, rows, cols);
      if (x != to) a += c.occ(x);
      if (y != h.from) b += c.occ(y);
    }
    const double dE = J * (double)(a - b);
    return dE <= 0.0 ? 1.0 : std::exp(-dE * (1.0 / (KB * temperature)));
  }

  /// Sum of the rates of every hop of the current configuration (the
  /// selector's total_rate, kinetic_monte_carlo.hh:195)
  double total_rate() const {
    auto const &c = m_state->configuration;
    double R = 0.0;
    for (Index l = 0; l < c.n_sites; ++l)
      if (c.occ(l) == 1)
        for (int d = 0; d < 4; ++d) R += rate(LatticeGasHop{l, d});
    return R;
  }

  /// The OccEvent of a hop (the get_event_f of kinetic_monte_carlo): the
  /// origin's atom leaves (1 -> 0), the target gains it (0 -> 1), and the atom
  /// moves by one lattice vector
  OccEvent const &get_event(LatticeGasHop h) {
    auto const &c = m_state->configuration;
    auto const &ol = *m_occ_location;
    const Index rows = c.shape[0], cols = c.shape[1];
    const Index to = lattice_gas_neighbour(h.from, h.dir, rows, cols);
    const Index va = ol.convert().species_index(0, 0), at = ol.convert().species_index(0, 1);
    m_event.linear_site_index = {h.from, to};
    m_event.new_occ = {0, 1};
    m_event.occ_transform.resize(2);
    m_event.occ_transform[0] = OccTransform{h.from, ol.l_to_mol_id(h.from), 0, at, va};
    m_event.occ_transform[1] = OccTransform{to, ol.l_to_mol_id(to), 0, va, at};
    AtomTraj traj;
    traj.from = AtomLocation{h.from, ol.l_to_mol_id(h.from), 0};
    traj.to = AtomLocation{to, ol.l_to_mol_id(to), 0};
    traj.delta_ijk = {{h.dir == 0 ? 1L : h.dir == 1 ? -1L : 0L, h.dir == 2 ? 1L : h.dir == 3 ? -1L : 0L, 0L}};
    m_event.atom_traj = {traj};
    return m_event;
  }

  State<IsingConfiguration> const *state() const { return m_state; }

 private:
  State<IsingConfiguration> const *m_state;
  OccLocation const *m_occ_location;
  OccEvent m_event;
};

}  // namespace ising_cpp
}  // namespace monte
}  // namespace CASM

#endif
