// Stand-in for the part of libcasm-monte's Conversions (Conversions.hh:43-131)
// that the occupant-location classes read, for the lattice models this library
// runs: every site belongs to one asymmetric unit, and each unit allows an
// ordered list of occupants (species names).  Species are numbered in the order
// they first appear over the units; a site's occupant index is the position of
// its species in its unit's list.  The reference builds Conversions from an
// xtal::BasicStructure prim and a supercell matrix (CASMcode_crystallography,
// not a dependency of this library), where the units come from the prim's
// basis under symmetry; here the caller gives them.  The member names and
// meanings used by OccCandidateList, OccLocation and the event proposal
// functions are kept.
#ifndef CASM_MI355X_monte_Conversions
#define CASM_MI355X_monte_Conversions

#include <array>

#include "casm/monte/definitions.hh"

namespace CASM {
namespace monte {

class Conversions {
 public:
  /// n_sites sites of one asymmetric unit, each allowing `occupants` (species
  /// names) in this order
  Conversions(Index n_sites, std::vector<std::string> occupants)
      : Conversions(std::vector<Index>(n_sites < 0 ? 0 : n_sites, 0), {std::move(occupants)}) {}

  /// a rows x cols square lattice (site l = i + rows j, lattice constant 1)
  /// of one asymmetric unit allowing `occupants`; it knows its sites'
  /// coordinates (l_to_ijk), as the kinetic Monte Carlo atom tracking needs
  Conversions(Index rows, Index cols, std::vector<std::string> occupants)
      : Conversions(rows * cols, std::move(occupants)) {
    m_rows = rows;
    m_cols = cols;
  }

  /// site l in unit l_to_asym[l]; unit b allows asym_occupants[b] in this order
  Conversions(std::vector<Index> l_to_asym, std::vector<std::vector<std::string>> asym_occupants)
      : m_l_to_asym(std::move(l_to_asym)), m_asym_occ(asym_occupants.size()) {
    if (m_l_to_asym.empty() || asym_occupants.empty())
      throw std::runtime_error("Error in Conversions: no sites or no asymmetric units");
    for (size_t b = 0; b < asym_occupants.size(); ++b) {
      if (asym_occupants[b].empty())
        throw std::runtime_error("Error in Conversions: a unit allows no occupant");
      for (auto const &name : asym_occupants[b]) {
        Index s = species_index(name);
        if (s == species_size()) m_species.push_back(name);
        for (Index x : m_asym_occ[b])
          if (x == s) throw std::runtime_error("Error in Conversions: repeated occupant");
        m_asym_occ[b].push_back(s);
      }
    }
    for (Index b : m_l_to_asym)
      if (b < 0 || b >= asym_size())
        throw std::runtime_error("Error in Conversions: asymmetric unit out of range");
  }

  Index l_size() const { return static_cast<Index>(m_l_to_asym.size()); }
  Index l_to_asym(Index l) const { return m_l_to_asym[l]; }
  Index asym_size() const { return static_cast<Index>(m_asym_occ.size()); }

  Index occ_size(Index asym) const { return static_cast<Index>(m_asym_occ[asym].size()); }
  Index species_index(Index asym, Index occ_index) const { return m_asym_occ[asym][occ_index]; }
  /// occ_size(asym) if the species is not allowed on the unit
  Index occ_index(Index asym, Index species_index) const {
    for (Index o = 0; o < occ_size(asym); ++o)
      if (m_asym_occ[asym][o] == species_index) return o;
    return occ_size(asym);
  }
  bool species_allowed(Index asym, Index species_index) const {
    return asym >= 0 && asym < asym_size() && occ_index(asym, species_index) < occ_size(asym);
  }
  Index species_size() const { return static_cast<Index>(m_species.size()); }
  Index species_index(std::string species_name) const {
    for (Index i = 0; i < species_size(); ++i)
      if (m_species[i] == species_name) return i;
    return species_size();
  }
  std::string const &species_name(Index species_index) const { return m_species.at(species_index); }
  /// atoms of a species: none for a vacancy ("Va"), one otherwise (the
  /// lattice models here have single-atom occupants)
  Index components_size(Index species_index) const {
    return species_name(species_index) == "Va" ? 0 : 1;
  }
  /// unit cell of site l of a square-lattice Conversions (its third index 0)
  std::array<long, 3> l_to_ijk(Index l) const {
    if (m_rows <= 0)
      throw std::runtime_error("Conversions::l_to_ijk: the lattice shape is not known");
    return {{(long)(l % m_rows), (long)(l / m_rows), 0}};
  }

 private:
  Index m_rows = 0, m_cols = 0;
  std::vector<Index> m_l_to_asym;
  std::vector<std::vector<Index>> m_asym_occ;  // species index of each occupant, per unit
  std::vector<std::string> m_species;
};

}  // namespace monte
}  // namespace CASM

#endif
