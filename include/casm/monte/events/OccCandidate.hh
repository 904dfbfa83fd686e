// Drop-in counterpart of libcasm-monte's events/OccCandidate.hh and
// src/casm/monte/events/OccCandidate.cc: occupant types (asymmetric unit,
// species), swap types, the candidate list and the swap-list builders.  Same
// names, orderings and rules: the candidate list enumerates, per asymmetric
// unit with more than one allowed occupant, its occupants in occupant-index
// order (OccCandidate.cc:31-52); make_canonical_swaps keeps each unordered
// pair once, cand_a < cand_b (OccCandidate.cc:99-110);
// make_semigrand_canonical_swaps keeps both orders on one asymmetric unit
// (OccCandidate.cc:126-148).
#ifndef CASM_MI355X_monte_OccCandidate
#define CASM_MI355X_monte_OccCandidate

#include <tuple>

#include "casm/monte/Conversions.hh"

namespace CASM {
namespace monte {

struct OccCandidate {
  OccCandidate(Index _asym, Index _species_index) : asym(_asym), species_index(_species_index) {}
  Index asym;
  Index species_index;
  bool operator<(OccCandidate B) const {
    if (asym != B.asym) return asym < B.asym;
    return species_index < B.species_index;
  }
  bool operator==(OccCandidate B) const { return !(*this < B) && !(B < *this); }
  bool operator!=(OccCandidate B) const { return !(*this == B); }
};

/// A type of change (which occupant types on which asymmetric units), not
/// which sites
class OccSwap {
 public:
  OccSwap(OccCandidate const &_cand_a, OccCandidate const &_cand_b)
      : cand_a(_cand_a), cand_b(_cand_b) {}
  OccCandidate cand_a;
  OccCandidate cand_b;
  void reverse() { std::swap(cand_a, cand_b); }
  OccSwap &sort() {
    OccSwap B(*this);
    B.reverse();
    if (B < *this) *this = B;
    return *this;
  }
  OccSwap sorted() const {
    OccSwap r(*this);
    r.sort();
    return r;
  }
  bool operator<(OccSwap const &B) const {
    return std::make_tuple(cand_a, cand_b) < std::make_tuple(B.cand_a, B.cand_b);
  }
  bool operator==(OccSwap const &B) const { return !(*this < B) && !(B < *this); }
  bool operator!=(OccSwap const &B) const { return !(*this == B); }
};

class OccCandidateList {
 public:
  typedef std::vector<OccCandidate>::const_iterator const_iterator;
  OccCandidateList() {}
  OccCandidateList(std::vector<OccCandidate> candidates, Conversions const &convert)
      : m_candidate(std::move(candidates)) {
    _index(convert);
  }
  explicit OccCandidateList(Conversions const &convert) {
    for (Index asym = 0; asym < convert.asym_size(); ++asym) {
      if (convert.occ_size(asym) < 2) continue;
      for (Index i = 0; i < convert.occ_size(asym); ++i)
        m_candidate.push_back(OccCandidate(asym, convert.species_index(asym, i)));
    }
    _index(convert);
  }
  Index index(OccCandidate const &cand) const {
    return m_species_to_cand_index[cand.asym][cand.species_index];
  }
  Index index(Index asym, Index species_index) const {
    return m_species_to_cand_index[asym][species_index];
  }
  OccCandidate const &operator[](Index candidate_index) const {
    return m_candidate[candidate_index];
  }
  const_iterator begin() const { return m_candidate.begin(); }
  const_iterator end() const { return m_candidate.end(); }
  Index size() const { return m_end; }

 private:
  void _index(Conversions const &convert) {
    m_end = static_cast<Index>(m_candidate.size());
    m_species_to_cand_index.assign(convert.asym_size(),
                                   std::vector<Index>(convert.species_size(), m_end));
    Index i = 0;
    for (auto const &c : m_candidate) m_species_to_cand_index[c.asym][c.species_index] = i++;
  }
  std::vector<std::vector<Index>> m_species_to_cand_index;
  std::vector<OccCandidate> m_candidate;
  Index m_end = 0;
};

inline bool is_valid(Conversions const &convert, OccCandidate const &cand) {
  return cand.asym >= 0 && cand.asym < convert.asym_size() && cand.species_index >= 0 &&
         cand.species_index < convert.species_size() &&
         convert.species_allowed(cand.asym, cand.species_index);
}

inline bool is_valid(Conversions const &convert, OccCandidate const &cand_a,
                     OccCandidate const &cand_b) {
  return is_valid(convert, cand_a) && is_valid(convert, cand_b);
}

inline bool is_valid(Conversions const &convert, OccSwap const &swap) {
  return is_valid(convert, swap.cand_a, swap.cand_b);
}

inline bool allowed_canonical_swap(Conversions const &convert, OccCandidate cand_a,
                                   OccCandidate cand_b) {
  return is_valid(convert, cand_a) && is_valid(convert, cand_b) &&
         cand_a.species_index != cand_b.species_index &&
         convert.species_allowed(cand_a.asym, cand_b.species_index) &&
         convert.species_allowed(cand_b.asym, cand_a.species_index);
}

inline std::vector<OccSwap> make_canonical_swaps(Conversions const &convert,
                                                 OccCandidateList const &occ_candidate_list) {
  std::vector<OccSwap> out;
  for (auto const &a : occ_candidate_list)
    for (auto const &b : occ_candidate_list)
      if (a < b && allowed_canonical_swap(convert, a, b)) out.push_back(OccSwap(a, b));
  return out;
}

inline bool allowed_semigrand_canonical_swap(Conversions const &convert, OccCandidate cand_a,
                                             OccCandidate cand_b) {
  return is_valid(convert, cand_a) && is_valid(convert, cand_b) && cand_a.asym == cand_b.asym &&
         cand_a.species_index != cand_b.species_index &&
         convert.species_allowed(cand_a.asym, cand_b.species_index);
}

inline std::vector<OccSwap> make_semigrand_canonical_swaps(
    Conversions const &convert, OccCandidateList const &occ_candidate_list) {
  std::vector<OccSwap> out;
  for (auto const &a : occ_candidate_list)
    for (auto const &b : occ_candidate_list)
      if (allowed_semigrand_canonical_swap(convert, a, b)) out.push_back(OccSwap(a, b));
  return out;
}

}  // namespace monte
}  // namespace CASM

#endif
