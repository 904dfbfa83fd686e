// Drop-in counterpart of libcasm-monte's events/OccEventProposal.hh: choose a
// swap type in proportion to the number of events of each type, then the
// sites, through the OccLocation candidate lists.  Same draw order as the
// reference, which fixes the stream a chain consumes:
//   choose_*_swap: one random_real over [0, total) (OccEventProposal.hh:109-152,
//     261-303), then the first type whose cumulative count exceeds it;
//   propose_canonical_event_from_swap: one choose_mol on cand_a, then one on
//     cand_b (:173-208);
//   propose_semigrand_canonical_event_from_swap: one choose_mol on cand_a
//     (:314-336).
// The canonical form is what occupation_metropolis runs on the MI355X.
#ifndef CASM_MI355X_monte_OccEventProposal
#define CASM_MI355X_monte_OccEventProposal

#include "casm/monte/events/OccLocation.hh"

namespace CASM {
namespace monte {

namespace detail {
template <typename CountF, typename GeneratorType>
OccSwap const &choose_swap(std::vector<OccSwap> const &swaps, CountF count,
                           GeneratorType &random_number_generator, const char *who) {
  std::vector<double> tsum(swaps.size() + 1, 0.0);
  for (size_t i = 0; i < swaps.size(); ++i) tsum[i + 1] = tsum[i] + count(swaps[i]);
  if (tsum.back() == 0.0)
    throw std::runtime_error(std::string("Error in ") + who + ": No events possible.");
  const double rand = random_number_generator.random_real(tsum.back());
  for (size_t i = 0; i < swaps.size(); ++i)
    if (rand < tsum[i + 1]) return swaps[i];
  throw std::runtime_error(std::string("Error in ") + who);
}
}  // namespace detail

template <typename GeneratorType>
OccSwap const &choose_canonical_swap(OccLocation const &occ_location,
                                     std::vector<OccSwap> const &canonical_swap,
                                     GeneratorType &random_number_generator) {
  return detail::choose_swap(
      canonical_swap,
      [&](OccSwap const &s) {
        return ((double)occ_location.cand_size(s.cand_a)) *
               ((double)occ_location.cand_size(s.cand_b));
      },
      random_number_generator, "choose_canonical_swap");
}

template <typename GeneratorType>
OccEvent &propose_canonical_event_from_swap(OccEvent &e, OccLocation const &occ_location,
                                            OccSwap const &swap,
                                            GeneratorType &random_number_generator) {
  e.occ_transform.resize(2);
  e.atom_traj.resize(0);
  e.linear_site_index.resize(2);
  e.new_occ.resize(2);
  OccCandidate const *from[2] = {&swap.cand_a, &swap.cand_b};
  OccCandidate const *to[2] = {&swap.cand_b, &swap.cand_a};
  for (int i = 0; i < 2; ++i) {
    Mol const &m = occ_location.choose_mol(*from[i], random_number_generator);
    OccTransform &t = e.occ_transform[i];
    t.mol_id = m.id;
    t.l = m.l;
    t.asym = from[i]->asym;
    t.from_species = from[i]->species_index;
    t.to_species = to[i]->species_index;
  }
  for (int i = 0; i < 2; ++i) {
    OccTransform const &t = e.occ_transform[i];
    e.linear_site_index[i] = t.l;
    e.new_occ[i] = (int)occ_location.convert().occ_index(t.asym, t.to_species);
  }
  return e;
}

template <typename GeneratorType>
OccEvent &propose_canonical_event(OccEvent &e, OccLocation const &occ_location,
                                  std::vector<OccSwap> const &canonical_swap,
                                  GeneratorType &random_number_generator) {
  auto const &swap = choose_canonical_swap(occ_location, canonical_swap, random_number_generator);
  return propose_canonical_event_from_swap(e, occ_location, swap, random_number_generator);
}

template <typename GeneratorType>
OccSwap const &choose_semigrand_canonical_swap(OccLocation const &occ_location,
                                               std::vector<OccSwap> const &semigrand_canonical_swap,
                                               GeneratorType &random_number_generator) {
  return detail::choose_swap(
      semigrand_canonical_swap,
      [&](OccSwap const &s) { return (double)occ_location.cand_size(s.cand_a); },
      random_number_generator, "choose_semigrand_canonical_swap");
}

template <typename GeneratorType>
OccEvent &propose_semigrand_canonical_event_from_swap(OccEvent &e,
                                                      OccLocation const &occ_location,
                                                      OccSwap const &swap,
                                                      GeneratorType &random_number_generator) {
  e.occ_transform.resize(1);
  e.atom_traj.resize(0);
  e.linear_site_index.resize(1);
  e.new_occ.resize(1);
  OccTransform &t = e.occ_transform[0];
  Mol const &m = occ_location.choose_mol(swap.cand_a, random_number_generator);
  t.mol_id = m.id;
  t.l = m.l;
  t.asym = swap.cand_a.asym;
  t.from_species = swap.cand_a.species_index;
  t.to_species = swap.cand_b.species_index;
  e.linear_site_index[0] = t.l;
  e.new_occ[0] = (int)occ_location.convert().occ_index(t.asym, t.to_species);
  return e;
}

template <typename GeneratorType>
OccEvent &propose_semigrand_canonical_event(OccEvent &e, OccLocation const &occ_location,
                                            std::vector<OccSwap> const &semigrand_canonical_swap,
                                            GeneratorType &random_number_generator) {
  auto const &swap = choose_semigrand_canonical_swap(occ_location, semigrand_canonical_swap,
                                                     random_number_generator);
  return propose_semigrand_canonical_event_from_swap(e, occ_location, swap,
                                                     random_number_generator);
}

}  // namespace monte
}  // namespace CASM

#endif
