// Drop-in counterpart of libcasm-monte's events/OccEvent.hh: the occupant
// (Mol), transformation and event records the occupant-location classes and
// the event proposal functions exchange.  Same members and meanings; atom
// translations are three integers (Eigen::Vector3l in the reference).
#ifndef CASM_MI355X_monte_OccEvent
#define CASM_MI355X_monte_OccEvent

#include <array>

#include "casm/monte/definitions.hh"

namespace CASM {
namespace monte {

/// One atom of a Mol: translation as if there were no periodic boundary, jumps
struct Atom {
  std::array<long, 3> translation{{0, 0, 0}};
  Index n_jumps = 0;
};

/// An occupant of a mutating site (OccEvent.hh Mol)
struct Mol {
  Index id = 0;             ///< position in OccLocation's Mol list
  Index l = 0;              ///< site (linear index in the configuration)
  Index asym = 0;           ///< asymmetric unit of the site
  Index species_index = 0;  ///< current species
  std::vector<Index> component;  ///< positions of its atoms in OccLocation's atom list
  Index loc = 0;            ///< position in the candidate list of its (asym, species)
};

/// Change of one occupant's species
struct OccTransform {
  Index l = 0;
  Index mol_id = 0;
  Index asym = 0;
  Index from_species = 0;
  Index to_species = 0;
};

struct AtomLocation {
  Index l = 0;
  Index mol_id = 0;
  Index mol_comp = 0;
};

struct AtomTraj {
  AtomLocation from;
  AtomLocation to;
  std::array<long, 3> delta_ijk{{0, 0, 0}};
};

/// A proposed event: the sites that change, their new occupant indices, the
/// occupant transformations and (not used by the canonical and semi-grand
/// canonical proposals) atom trajectories
struct OccEvent {
  std::vector<Index> linear_site_index;
  std::vector<int> new_occ;
  std::vector<OccTransform> occ_transform;
  std::vector<AtomTraj> atom_traj;
};

}  // namespace monte
}  // namespace CASM

#endif
