// Stand-in for the part of libcasm-monte's Conversions (Conversions.hh:43-131)
// that the occupant-location classes read, for the lattice models this library
// runs: one asymmetric unit whose sites all allow the same ordered occupants,
// species index == occupant index.  The reference builds Conversions from an
// xtal::BasicStructure prim and a supercell matrix (CASMcode_crystallography,
// not a dependency of this library); the member names and meanings used by
// OccCandidateList, OccLocation and the event proposal functions are kept.
#ifndef CASM_MI355X_monte_Conversions
#define CASM_MI355X_monte_Conversions

#include "casm/monte/definitions.hh"

namespace CASM {
namespace monte {

class Conversions {
 public:
  /// n_sites sites, each allowing `occupants` (species names) in this order
  Conversions(Index n_sites, std::vector<std::string> occupants)
      : m_n_sites(n_sites), m_species(std::move(occupants)) {
    if (n_sites < 1 || m_species.empty())
      throw std::runtime_error("Error in Conversions: no sites or no occupants");
  }

  Index l_size() const { return m_n_sites; }
  Index l_to_asym(Index l) const { return 0; }
  Index asym_size() const { return 1; }

  Index occ_size(Index asym) const { return static_cast<Index>(m_species.size()); }
  Index species_index(Index asym, Index occ_index) const { return occ_index; }
  Index occ_index(Index asym, Index species_index) const { return species_index; }
  bool species_allowed(Index asym, Index species_index) const {
    return asym == 0 && species_index >= 0 && species_index < species_size();
  }
  Index species_size() const { return static_cast<Index>(m_species.size()); }
  Index species_index(std::string species_name) const {
    for (Index i = 0; i < species_size(); ++i)
      if (m_species[i] == species_name) return i;
    return species_size();
  }
  std::string const &species_name(Index species_index) const { return m_species.at(species_index); }

 private:
  Index m_n_sites;
  std::vector<std::string> m_species;
};

}  // namespace monte
}  // namespace CASM

#endif
