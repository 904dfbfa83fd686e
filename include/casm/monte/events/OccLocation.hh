// Drop-in counterpart of libcasm-monte's events/OccLocation.hh and
// src/casm/monte/events/OccLocation.cc: the occupant-location tracker that
// makes choosing an occupant of a given type O(1).
//
// Kept from the reference: the Mol list in site order over the mutating sites
// (OccLocation.cc:39-118), one candidate list per (asymmetric unit, species),
// choose_mol as a uniform draw over a candidate list's positions
// (OccLocation.hh:254-275), and apply's update (OccLocation.cc:254-283): a
// transformed Mol leaves its list by moving the list's last entry into its
// slot, then is appended to the list of its new species.  The order of the
// lists decides which site later draws pick, so it is kept exactly.
//
// Atom tracking (update_atoms, OccLocation.cc:68-106, 216-330): every
// non-vacancy occupant holds one atom, created at initialize in Mol order with
// its site's unit cell as translation; apply moves the atoms of the event's
// trajectories, adding delta_ijk to their translations and counting their
// jumps; atom_positions_cart() gives the positions on the unit square lattice
// of a Conversions that knows its shape.  Not supported here: unique atom ids
// and atom info (track_unique_atoms, save_atom_info), atoms moving to or from
// a reservoir.  Added for moving the lists to and from the device:
// locations(), set_locations() and set_atoms().
#ifndef CASM_MI355X_monte_OccLocation
#define CASM_MI355X_monte_OccLocation

#include <set>

#include "casm/monte/events/OccCandidate.hh"
#include "casm/monte/events/OccEvent.hh"

namespace CASM {
namespace monte {

class OccLocation {
 public:
  typedef Index size_type;

  OccLocation(Conversions const &_convert, OccCandidateList const &_candidate_list,
              bool _update_atoms = false, bool _track_unique_atoms = false,
              bool _save_atom_info = false)
      : m_convert(_convert),
        m_candidate_list(_candidate_list),
        m_loc(_candidate_list.size()),
        m_update_atoms(_update_atoms) {
    if (_track_unique_atoms || _save_atom_info)
      throw std::runtime_error(
          "OccLocation (MI355X build): unique atom ids and atom info are not supported");
  }

  /// Fill the Mol list and the candidate lists from an occupation (occupant indices)
  void initialize(std::vector<int> const &occupation, std::optional<double> time = std::nullopt) {
    m_mol.clear();
    m_l_to_mol.clear();
    m_atoms.clear();
    for (auto &v : m_loc) v.clear();
    Index n_mut = 0;
    for (Index l = 0; l < (Index)occupation.size(); ++l)
      if (m_convert.occ_size(m_convert.l_to_asym(l)) > 1) ++n_mut;
    m_mol.resize(n_mut);
    Index mol_id = 0;
    for (Index l = 0; l < (Index)occupation.size(); ++l) {
      const Index asym = m_convert.l_to_asym(l);
      if (m_convert.occ_size(asym) > 1) {
        const Index species = m_convert.species_index(asym, occupation[l]);
        const Index cand = m_candidate_list.index(asym, species);
        Mol &mol = m_mol[mol_id];
        mol.id = mol_id;
        mol.l = l;
        mol.asym = asym;
        mol.species_index = species;
        mol.loc = (Index)m_loc[cand].size();
        mol.component.clear();
        if (m_update_atoms)
          for (Index k = 0; k < m_convert.components_size(species); ++k) {
            mol.component.push_back((Index)m_atoms.size());
            Atom atom;
            atom.translation = m_convert.l_to_ijk(l);
            atom.n_jumps = 0;
            m_atoms.push_back(atom);
          }
        m_loc[cand].push_back(mol_id);
        m_l_to_mol.push_back(mol_id);
        ++mol_id;
      } else {
        m_l_to_mol.push_back(n_mut);
      }
    }
  }

  /// Update the occupation and the lists for an event that occurred
  void apply(OccEvent const &e, std::vector<int> &occupation,
             std::optional<double> time = std::nullopt) {
    // the moving atoms, read before the transforms resize the Mols' components
    std::vector<Index> moving;
    if (m_update_atoms)
      for (auto const &traj : e.atom_traj) {
        if (traj.from.l == -1 || traj.to.l == -1)
          throw std::runtime_error(
              "OccLocation (MI355X build): atoms moving to or from a reservoir are not supported");
        moving.push_back(m_mol[traj.from.mol_id].component[traj.from.mol_comp]);
      }
    for (auto const &t : e.occ_transform) {
      Mol &mol = m_mol[t.mol_id];
      if (mol.species_index != t.from_species)
        throw std::runtime_error("Error in OccLocation::apply: species mismatch");
      occupation[mol.l] = (int)m_convert.occ_index(mol.asym, t.to_species);
      Index cand = m_candidate_list.index(mol.asym, mol.species_index);
      const Index back = m_loc[cand].back();
      m_loc[cand][mol.loc] = back;
      m_mol[back].loc = mol.loc;
      m_loc[cand].pop_back();
      mol.species_index = t.to_species;
      if (m_update_atoms) mol.component.resize(m_convert.components_size(mol.species_index));
      cand = m_candidate_list.index(mol.asym, mol.species_index);
      mol.loc = (Index)m_loc[cand].size();
      m_loc[cand].push_back(mol.id);
    }
    if (m_update_atoms)
      for (size_t i = 0; i < e.atom_traj.size(); ++i) {
        auto const &traj = e.atom_traj[i];
        const Index id = moving[i];
        m_mol[traj.to.mol_id].component[traj.to.mol_comp] = id;
        for (int d = 0; d < 3; ++d) m_atoms[id].translation[d] += traj.delta_ijk[d];
        m_atoms[id].n_jumps += 1;
      }
  }

  /// Number of tracked atoms and access by atom id
  Index atom_size() const { return (Index)m_atoms.size(); }
  Atom &atom(Index atom_id) { return m_atoms[atom_id]; }
  Atom const &atom(Index atom_id) const { return m_atoms[atom_id]; }

  /// Atom positions, 3 x atom_size(), as if there were no periodic boundary
  /// (OccLocation.cc:385-403 on the unit square lattice: one site per unit
  /// cell at its origin, point atoms)
  MatrixXd atom_positions_cart() const {
    MatrixXd R = MatrixXd::Zero(3, atom_size());
    for (Index a = 0; a < atom_size(); ++a)
      for (int d = 0; d < 3; ++d) R(d, a) = (double)m_atoms[a].translation[d];
    return R;
  }

  /// Replace the atoms and the Mols' components (the device's atom records at
  /// the end of a kinetic run): component[mol] holds the atom ids of each Mol
  void set_atoms(std::vector<Atom> const &atoms,
                 std::vector<std::vector<Index>> const &component) {
    if (component.size() != m_mol.size())
      throw std::runtime_error("OccLocation::set_atoms: one component list per Mol");
    m_atoms = atoms;
    for (size_t i = 0; i < m_mol.size(); ++i) m_mol[i].component = component[i];
  }

  template <typename GeneratorType>
  Mol const &choose_mol(Index cand_index, GeneratorType &random_number_generator) const {
    return mol(m_loc[cand_index][random_number_generator.random_int(m_loc[cand_index].size() - 1)]);
  }
  template <typename GeneratorType>
  Mol const &choose_mol(Index cand_index, std::set<Index> exclude,
                        GeneratorType &random_number_generator) const {
    Index loc;
    do {
      loc = random_number_generator.random_int(m_loc[cand_index].size() - 1);
    } while (exclude.count(loc));
    return mol(m_loc[cand_index][loc]);
  }
  template <typename GeneratorType>
  Mol const &choose_mol(OccCandidate const &cand, GeneratorType &random_number_generator) const {
    return choose_mol(m_candidate_list.index(cand), random_number_generator);
  }
  template <typename GeneratorType>
  Mol const &choose_mol(OccCandidate const &cand, std::set<Index> exclude,
                        GeneratorType &random_number_generator) const {
    return choose_mol(m_candidate_list.index(cand), exclude, random_number_generator);
  }

  size_type mol_size() const { return (size_type)m_mol.size(); }
  Mol &mol(Index mol_id) { return m_mol[mol_id]; }
  Mol const &mol(Index mol_id) const { return m_mol[mol_id]; }
  OccCandidateList const &candidate_list() const { return m_candidate_list; }
  size_type cand_size(Index cand_index) const { return (size_type)m_loc[cand_index].size(); }
  size_type cand_size(OccCandidate const &cand) const {
    return cand_size(m_candidate_list.index(cand));
  }
  Index mol_id(Index cand_index, Index loc) const { return m_loc[cand_index][loc]; }
  Index mol_id(OccCandidate const &cand, Index loc) const {
    return mol_id(m_candidate_list.index(cand), loc);
  }
  Index l_to_mol_id(Index l) const { return m_l_to_mol[l]; }
  Conversions const &convert() const { return m_convert; }

  /// The candidate lists (Mol ids per candidate index, in list order)
  std::vector<std::vector<Index>> const &locations() const { return m_loc; }

  /// Replace the candidate lists (every Mol once, in the list of its current
  /// species) and the Mol positions they imply
  void set_locations(std::vector<std::vector<Index>> const &loc) {
    if (loc.size() != m_loc.size())
      throw std::runtime_error("OccLocation::set_locations: wrong number of lists");
    std::vector<char> seen(m_mol.size(), 0);
    size_t total = 0;
    for (auto const &v : loc) total += v.size();
    if (total != m_mol.size())
      throw std::runtime_error("OccLocation::set_locations: every Mol must be listed once");
    for (size_t c = 0; c < loc.size(); ++c) {
      for (Index id : loc[c]) {
        if (id < 0 || id >= (Index)m_mol.size() || seen[id] ||
            m_candidate_list.index(m_mol[id].asym, m_mol[id].species_index) != (Index)c)
          throw std::runtime_error("OccLocation::set_locations: inconsistent lists");
        seen[id] = 1;
      }
    }
    m_loc = loc;
    for (auto const &v : m_loc)
      for (Index i = 0; i < (Index)v.size(); ++i) m_mol[v[i]].loc = i;
  }

 private:
  Conversions const &m_convert;
  OccCandidateList const &m_candidate_list;
  std::vector<std::vector<Index>> m_loc;
  std::vector<Mol> m_mol;
  std::vector<Index> m_l_to_mol;
  bool m_update_atoms = false;
  std::vector<Atom> m_atoms;
};

}  // namespace monte
}  // namespace CASM

#endif
