// api.hip -- host implementation of the C ABI declared in include/casm_mi355x/mcgpu.h.
//
// Owns the device memory of a run (lattices, RNG states, counters, sample
// series), builds the per-chain acceptance tables on the host with the same
// libm exp the reference calls per attempt, and launches the kernels of
// exact.hip, checker.hip and stats.hip on the handle's stream.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fcc.hh"
#include "internal.hh"

namespace mcg {
hipError_t exact_set_max_lds(int bytes);
hipError_t checker_set_max_lds(int bytes);
hipError_t fcc_set_max_lds(int bytes);
int checker_units_per_item();
}  // namespace mcg

using namespace mcg;

namespace {

thread_local std::string g_err;

int fail(int code, std::string const &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return fail(MCG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

constexpr int64_t LDS_LIMIT = 160 * 1024;     // gfx950 LDS per workgroup
constexpr int64_t EXACT_LDS_MAX = 144 * 1024;  // lattice + fold scratch (+ ~10 KB static)
constexpr int64_t CHECKER_LDS_MAX = 128 * 1024;
constexpr int64_t FCC_LDS_MAX = 140 * 1024;  // lattice; tables, lists and schedule take 19.2 KB static
// dynamic LDS above half a CU's 160 KB: a workgroup then has its CU to itself
constexpr int FCC_ONE_PER_CU_LDS = 81 * 1024;

// The most lattice lines any step of a chain's class schedule sweeps: the
// class-line kernel runs up to one thread per line of the step.
int fcc_max_step_lines(int L0, int L1, int L2, const FccSched &S) {
  auto per_parity = [](int L, int a, int par) {
    int k = 0;
    for (int v = 0; v < 2 * L; ++v) k += fcc_ring_colour(v, 2 * L) == a && (v & 1) == par;
    return k;
  };
  int most = 0;
  for (int s = 0; s < S.n_steps; ++s) {
    int lines = 0;
    for (int k = S.start[s]; k < S.start[s + 1]; ++k) {
      const int col = S.cls[k], cx = col % FCC_NCOL, cy = (col / FCC_NCOL) % FCC_NCOL,
                cz = col / (FCC_NCOL * FCC_NCOL);
      for (int py = 0; py < 2; ++py)
        for (int pz = 0; pz < 2; ++pz)
          if (per_parity(L0, cx, py ^ pz) > 0)
            lines += per_parity(L1, cy, py) * per_parity(L2, cz, pz);
    }
    most = std::max(most, lines);
  }
  return most;
}
constexpr int64_t OCC_LDS_MAX = 136 * 1024;  // lattice + lists (+ ~23 KB static)
constexpr int64_t OCC_MAX_SITES = 40000;
constexpr int64_t KMC_MAX_SITES = 7500;  // lists, positions, site atoms, atoms: 18 B a site
constexpr int CHECKER_BAND_MAX_L0 = 8192;     // longest line of the 2D band path

// libstdc++ canonical double of a word compared to prob: the first word whose
// canonical value is >= prob (2^64 if none), by bisection on the monotone map.
// Returns false when every word accepts.
bool threshold_for(double prob, uint64_t *thr) {
  if (canonical_from_word(~0ULL) < prob) return false;  // all words accept
  uint64_t lo = 0, hi = ~0ULL;  // invariant: answer in [lo, hi]
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (canonical_from_word(mid) >= prob)
      hi = mid;
    else
      lo = mid + 1;
  }
  *thr = lo;
  return true;
}

// Ising formation energy per supercell of a 2D column-major lattice, both
// variants of model.hh:259-290.
double ising_per_supercell(std::vector<int8_t> const &occ, int64_t rows, int64_t cols,
                           double J, bool use_nlist) {
  auto site = [&](int64_t i, int64_t j) { return (int)occ[i + rows * j]; };
  if (use_nlist) {
    // model.hh:265-270: integer bond sum accumulated in a double, then *= -J
    double e = 0.0;
    for (int64_t l = 0; l < rows * cols; ++l) {
      int64_t i = l % rows, j = l / rows;
      int64_t iu = (i + 1) % rows, ju = (j + 1) % cols;
      e += site(i, j) * (site(iu, j) + site(i, ju));
    }
    e *= -J;
    return e;
  }
  double e = 0.0;
  for (int64_t i = 0; i < rows; ++i) {
    int64_t iu = (i + 1) % rows;
    long dot = 0;
    for (int64_t j = 0; j < cols; ++j) dot += site(i, j) * site(iu, j);
    e += -J * static_cast<double>(dot);
  }
  for (int64_t j = 0; j < cols; ++j) {
    int64_t ju = (j + 1) % cols;
    long dot = 0;
    for (int64_t i = 0; i < rows; ++i) dot += site(i, j) * site(i, ju);
    e += -J * static_cast<double>(dot);
  }
  return e;
}

// CHECKERBOARD device layout: every line of L0 sites stored as its even-i
// sites then its odd-i sites, one byte per site, 1 for spin +1 and 0 for -1.
void to_split_rows(std::vector<int8_t> const &nat, int64_t L0, std::vector<int8_t> &dev) {
  const int64_t half = L0 / 2, n = (int64_t)nat.size();
  for (int64_t l = 0; l < n; ++l) {
    const int64_t line = l / L0, i = l - line * L0;
    dev[line * L0 + (i & 1) * half + (i >> 1)] = nat[l] > 0 ? 1 : 0;
  }
}

void from_split_rows(std::vector<int8_t> const &dev, int64_t L0, std::vector<int8_t> &nat) {
  const int64_t half = L0 / 2, n = (int64_t)dev.size();
  for (int64_t l = 0; l < n; ++l) {
    const int64_t line = l / L0, i = l - line * L0;
    nat[l] = dev[line * L0 + (i & 1) * half + (i >> 1)] ? 1 : -1;
  }
}

struct LatticeGeom {
  int ndim;
  int64_t L[3];
  int64_t n() const { return L[0] * L[1] * L[2]; }
};

// Integer observables of a lattice (forward bonds counted once).
void integer_observables(std::vector<int8_t> const &s, LatticeGeom const &g, bool nnn,
                         int64_t &S1, int64_t &S2, int64_t &M) {
  S1 = S2 = M = 0;
  auto at = [&](int64_t i, int64_t j, int64_t k) {
    i = (i % g.L[0] + g.L[0]) % g.L[0];
    j = (j % g.L[1] + g.L[1]) % g.L[1];
    k = (k % g.L[2] + g.L[2]) % g.L[2];
    return (int)s[i + g.L[0] * (j + g.L[1] * k)];
  };
  for (int64_t k = 0; k < g.L[2]; ++k)
    for (int64_t j = 0; j < g.L[1]; ++j)
      for (int64_t i = 0; i < g.L[0]; ++i) {
        int v = at(i, j, k);
        M += v;
        S1 += v * (at(i + 1, j, k) + at(i, j + 1, k));
        if (g.ndim == 3) S1 += v * at(i, j, k + 1);
        if (nnn) {
          S2 += v * (at(i + 1, j + 1, k) + at(i + 1, j - 1, k));
          if (g.ndim == 3)
            S2 += v * (at(i + 1, j, k + 1) + at(i + 1, j, k - 1) + at(i, j + 1, k + 1) +
                       at(i, j + 1, k - 1));
        }
      }
}

// ---- FCC cluster expansion (host side) ------------------------------------------
struct FccOffset {
  int dx, dy, dz;
};

std::vector<FccOffset> fcc_offsets(int shell) {
  std::vector<FccOffset> v;
#define MCG_FCC_PUSH(k, dx, dy, dz) v.push_back({dx, dy, dz});
  if (shell == 1) {
    MCG_FCC_NN(MCG_FCC_PUSH)
  } else if (shell == 2) {
    MCG_FCC_N2(MCG_FCC_PUSH)
  } else {
    MCG_FCC_N3(MCG_FCC_PUSH)
  }
#undef MCG_FCC_PUSH
  return v;
}

bool fcc_is_nn(int dx, int dy, int dz) { return dx * dx + dy * dy + dz * dz == 2; }

// The class schedule of an L0 x L1 x L2 chain (FccSched, fcc.hh).  Two classes
// interact when some site of one has a site of the other among its 42
// neighbours (shells 1-3; the triangles lie within the first shell).  Per
// ring the pairs (d, v mod 2) with colour(v) = a and colour(v + d) = b are
// listed for |d| <= 2; classes (ax, ay, az) and (bx, by, bz) interact when an
// offset (dx, dy, dz) has such pairs in the three rings whose parities sum to
// an even number (x + y + z is even on every site).  The test agrees with a
// brute-force search over all site pairs on small lattices of even and odd
// edges.  Classes are placed in order of decreasing size, then class id.
void fcc_class_schedule(int L0, int L1, int L2, FccSched &S) {
  constexpr int NC = FCC_NCOL;
  const int L[3] = {L0, L1, L2};
  // pairs[dim][a][b]: bit (d + 2) * 2 + parity; cnt[dim][a][p]: points of colour a, parity p
  uint32_t pairs[3][NC][NC] = {};
  int cnt[3][NC][2] = {};
  for (int dim = 0; dim < 3; ++dim) {
    const int n = 2 * L[dim];
    for (int v = 0; v < n; ++v) {
      const int a = fcc_ring_colour(v, n);
      ++cnt[dim][a][v & 1];
      for (int d = -2; d <= 2; ++d) {
        const int b = fcc_ring_colour(((v + d) % n + n) % n, n);
        pairs[dim][a][b] |= 1u << ((d + 2) * 2 + (v & 1));
      }
    }
  }
  std::vector<FccOffset> off;
  for (int sh = 1; sh <= 3; ++sh)
    for (auto const &o : fcc_offsets(sh)) off.push_back(o);
  auto coords = [](int col, int *c) {
    c[0] = col % NC;
    c[1] = (col / NC) % NC;
    c[2] = col / (NC * NC);
  };
  auto interact = [&](int a, int b) {
    int ac[3], bc[3];
    coords(a, ac);
    coords(b, bc);
    for (auto const &o : off) {
      const int dd[3] = {o.dx, o.dy, o.dz};
      for (int px = 0; px < 2; ++px)
        for (int py = 0; py < 2; ++py) {
          const int pp[3] = {px, py, (px + py) & 1};
          bool ok = true;
          for (int dim = 0; dim < 3 && ok; ++dim)
            ok = (pairs[dim][ac[dim]][bc[dim]] >> ((dd[dim] + 2) * 2 + pp[dim])) & 1u;
          if (ok) return true;
        }
    }
    return false;
  };
  auto sites = [&](int col) {
    int c[3];
    coords(col, c);
    int64_t k = 0;
    for (int px = 0; px < 2; ++px)
      for (int py = 0; py < 2; ++py)
        k += (int64_t)cnt[0][c[0]][px] * cnt[1][c[1]][py] * cnt[2][c[2]][(px + py) & 1];
    return k;
  };
  std::vector<int> order;
  for (int col = 0; col < FCC_NCLASS; ++col)
    if (sites(col) > 0) order.push_back(col);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return sites(a) > sites(b); });
  std::vector<std::vector<int>> steps;
  for (int col : order) {
    bool placed = false;
    for (auto &st : steps) {
      if ((int)st.size() >= FCC_STEP_MAX) continue;
      bool free = true;
      for (int o : st) free = free && !interact(col, o);
      if (free) {
        st.push_back(col);
        placed = true;
        break;
      }
    }
    if (!placed) steps.push_back({col});
  }
  S = FccSched{};
  S.n_steps = (int32_t)steps.size();
  int k = 0;
  for (size_t s = 0; s < steps.size(); ++s) {
    S.start[s] = (uint8_t)k;
    for (int col : steps[s]) S.cls[k++] = (uint8_t)col;
  }
  S.start[steps.size()] = (uint8_t)k;
  S.n_classes = k;
}


// Triangle count of a nearest-neighbour occupancy mask (bit k = neighbour k of
// MCG_FCC_NN): the occupied pairs among the neighbour pairs that are
// themselves nearest neighbours.
std::vector<uint8_t> fcc_tri_table() {
  const auto nn = fcc_offsets(1);
  std::vector<std::pair<int, int>> tp;
  for (int a = 0; a < 12; ++a)
    for (int b = a + 1; b < 12; ++b)
      if (fcc_is_nn(nn[b].dx - nn[a].dx, nn[b].dy - nn[a].dy, nn[b].dz - nn[a].dz))
        tp.push_back({a, b});
  std::vector<uint8_t> tab(4096);
  for (int m = 0; m < 4096; ++m) {
    int t = 0;
    for (auto const &pr : tp) t += ((m >> pr.first) & 1) & ((m >> pr.second) & 1);
    tab[m] = (uint8_t)t;
  }
  return tab;
}

// The triangle table in the nearest-neighbour bit order of the bit-line
// kernel (fcc.hip, fcc_bits_unit): bits 0-7 the lines (dy, dz) = (-1, 0),
// (1, 0), (0, -1), (0, 1) with dx = -1 then +1, bits 8-11 the lines (-1, -1),
// (-1, 1), (1, -1), (1, 1) with dx = 0.  The count is a property of the
// occupied neighbours, so both tables give the same t for the same lattice.
std::vector<uint8_t> fcc_tri_table_lines() {
  const FccOffset nn[12] = {{-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
                            {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
                            {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}};
  std::vector<std::pair<int, int>> tp;
  for (int a = 0; a < 12; ++a)
    for (int b = a + 1; b < 12; ++b)
      if (fcc_is_nn(nn[b].dx - nn[a].dx, nn[b].dy - nn[a].dy, nn[b].dz - nn[a].dz))
        tp.push_back({a, b});
  std::vector<uint8_t> tab(4096);
  for (int m = 0; m < 4096; ++m) {
    int t = 0;
    for (auto const &pr : tp) t += ((m >> pr.first) & 1) & ((m >> pr.second) & 1);
    tab[m] = (uint8_t)t;
  }
  return tab;
}

// Counts of an FCC occupation (site order l = h + L0*(y + 2*L1*z)).
void fcc_counts(std::vector<int8_t> const &occ, const int64_t *L, int64_t &N1, int64_t &P1,
                int64_t &P2, int64_t &P3, int64_t &T) {
  const int64_t L0 = L[0], ry = 2 * L[1], rz = 2 * L[2], rx = 2 * L[0];
  auto at = [&](int64_t x, int64_t y, int64_t z) -> int {
    x = (x % rx + rx) % rx;
    y = (y % ry + ry) % ry;
    z = (z % rz + rz) % rz;
    return occ[(x >> 1) + L0 * (y + ry * z)];
  };
  const auto s1 = fcc_offsets(1), s2 = fcc_offsets(2), s3 = fcc_offsets(3);
  std::vector<std::pair<int, int>> tp;
  for (int a = 0; a < 12; ++a)
    for (int b = a + 1; b < 12; ++b)
      if (fcc_is_nn(s1[b].dx - s1[a].dx, s1[b].dy - s1[a].dy, s1[b].dz - s1[a].dz))
        tp.push_back({a, b});
  int64_t p1 = 0, p2 = 0, p3 = 0, t = 0;
  N1 = 0;
  for (int64_t l = 0; l < (int64_t)occ.size(); ++l) {
    if (!occ[l]) continue;
    const int64_t h = l % L0, line = l / L0, y = line % ry, z = line / ry;
    const int64_t x = ((y + z) & 1) + 2 * h;
    ++N1;
    int nb[12];
    for (int k = 0; k < 12; ++k) {
      nb[k] = at(x + s1[k].dx, y + s1[k].dy, z + s1[k].dz);
      p1 += nb[k];
    }
    for (auto const &o : s2) p2 += at(x + o.dx, y + o.dy, z + o.dz);
    for (auto const &o : s3) p3 += at(x + o.dx, y + o.dy, z + o.dz);
    for (auto const &pr : tp) t += nb[pr.first] & nb[pr.second];
  }
  P1 = p1 / 2;
  P2 = p2 / 2;
  P3 = p3 / 2;
  T = t / 3;
}

}  // namespace

struct mcg_handle {
  int64_t mode = 0, lattice = 0;
  bool use_nlist = true;
  double J1 = 0, J2 = 0, kb = 0;
  int64_t sample_period = 1;
  int64_t options = 0;
  double eci[6] = {0, 0, 0, 0, 0, 0};  // FCC
  int n_chains = 0;
  std::vector<mcg_chain_desc> chains;
  std::vector<int64_t> lat_off;
  int64_t lat_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0, last_attempts = 0.0;

  int8_t *d_lat = nullptr;
  uint8_t *d_active = nullptr;
  std::vector<uint8_t> active;
  mcg_sample *d_series = nullptr;
  int64_t series_cap = 0;
  std::vector<int64_t> n_samples;  // host mirror of the series length
  std::vector<int64_t> n_pass;     // host mirror of the pass counters
  std::vector<int64_t> n_step;     // EXACT: host mirror of the steps into the current pass
  std::vector<int64_t> capacity;   // sites each chain slot was allocated for
  double *d_stats = nullptr;
  int64_t stats_cap = 0;
  // mcg_check_batch: chain list, counts, scratch offsets and observable
  // parameters of one call (bytes)
  unsigned char *d_batch = nullptr;
  int64_t batch_cap = 0;
  int64_t *d_eq = nullptr;

  // EXACT (and CANONICAL, which shares the state records d_xs)
  ExactParams *d_xp = nullptr;
  ExactState *d_xs = nullptr;
  std::vector<ExactParams> xp;
  // per chain: LDS bytes of the block walk and whether they fit (else the
  // chain runs the scalar walk on its HBM lattice); device masks of the active
  // chains on each path, and a second stream and two events that let the HBM
  // chains run beside the LDS ones
  std::vector<int> xlds;
  std::vector<uint8_t> xfit;
  uint8_t *d_mask_lds = nullptr, *d_mask_hbm = nullptr;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  // CANONICAL: per-chain parameters and the candidate lists (int32 entries at
  // the lattice offsets: species-0 sites then species-1 sites)
  CanonParams *d_kp = nullptr;
  std::vector<CanonParams> kp;
  int32_t *d_list = nullptr;
  int canon_lds = 0;
  bool canonical() const { return mode == MCG_MODE_CANONICAL; }
  // NFOLD: per-chain parameters (the class arrays share d_list)
  NfoldParams *d_np = nullptr;
  std::vector<NfoldParams> np_;
  int nfold_lds = 0;
  double sample_time_period = 0.0;
  bool nfold() const { return mode == MCG_MODE_NFOLD; }
  // OCCUPANT: the system (units, candidates, swap types, exchange potentials
  // per chain and species), per-chain parameters and class tables, and each
  // chain's site units (the lists share d_list)
  OccParams *d_op = nullptr;
  OccTables *d_ot = nullptr;
  std::vector<OccParams> op;
  int occ_lds = 0;
  int occ_n_asym = 0, occ_n_species = 0, occ_canonical = 0;
  int64_t occ_spu = 1;
  int occ_size[OCC_MAX_ASYM] = {};
  int occ_species[OCC_MAX_ASYM][2] = {};
  std::vector<int32_t> occ_cand_asym, occ_cand_occ, occ_swap_a, occ_swap_b;
  std::vector<double> occ_mu;  // n_chains x n_species
  std::vector<std::vector<int8_t>> site_asym;
  bool occupant() const { return mode == MCG_MODE_OCCUPANT; }
  // KMC: per-chain parameters, the atom pool (4 int32 per atom: site, dx, dy,
  // jumps; the lists share d_list) and the attempts without an event after
  // which a chain stops
  KmcParams *d_kp_kmc = nullptr;
  std::vector<KmcParams> kp_kmc;
  int32_t *d_atoms = nullptr;
  // hop correlations (mcg_kmc_collect_hops): every kmc_hop_every-th jump, at
  // most kmc_hop_max records per atom; records kmc_hop_max x lat_bytes, counts
  // one per site slot of the lattice pool
  int32_t kmc_hop_every = 0, kmc_hop_max = 0;
  KmcHop *d_hop = nullptr;
  int32_t *d_hop_cnt = nullptr;
  int kmc_lds = 0;
  int64_t kmc_spin_limit = int64_t(1) << 26;
  bool kmc() const { return mode == MCG_MODE_KMC; }
  // one sequential chain per wavefront, MT19937-64 state, native lattice layout
  bool sequential() const {
    return mode == MCG_MODE_EXACT || mode == MCG_MODE_CANONICAL || mode == MCG_MODE_NFOLD ||
           mode == MCG_MODE_OCCUPANT || mode == MCG_MODE_KMC;
  }

  // CHECKERBOARD
  // chains with an odd edge run the ring-coloured kernels (checker_ring.hip) on
  // a natural-layout lattice; masks of the active chains of each kind and the
  // list of active ring chains for the HBM class launches
  std::vector<uint8_t> cring;
  uint8_t *d_mask_even = nullptr, *d_mask_ring = nullptr;
  int32_t *d_ring_list = nullptr;
  int n_ring_active = 0, ring_per_chain = 1, ring_lds_bytes = 0;
  bool any_ring() const { return n_ring_active > 0; }
  CheckerParams *d_cp = nullptr;
  CheckerState *d_cs = nullptr;
  uint32_t *d_tab = nullptr;
  std::vector<CheckerParams> cp;
  int32_t *d_work = nullptr;
  int64_t work_cap = 0, band_cap = 0, pair_cap = 0;  // int32 entries allocated
  int n_work = 0;
  bool work_dirty = true;
  bool checker_lds = true;
  int checker_lds_bytes = 0;
  int checker_block = 256;
  int ncolor = 2;
  int units_per_thread = 2;
  int graph_passes = 8;
  // 2D band path: lines per band (-1 = chosen from the sizes, 0 = off) and the
  // (chain, first line, lines) list
  int band_lines = -1;
  int line_tiles = 0;  // line-class path: LDS tiles (1) or lines from HBM (0)
  int32_t *d_band = nullptr;
  int n_bands = 0;
  int band_ring_bytes = 0;
  bool band_slabs = false;  // the band list holds 3D slabs (checker_slab.hip)
  // line-class path (4/8 colourings): (chain, first line, lines) items
  int32_t *d_pair = nullptr;
  int n_pairs = 0;
  int pair_lds = 0;  // dynamic LDS of the line-class launches (largest tile)
  int tile_grid = 0;  // persistent workgroups of the tiled line-class launches
  int n_cu = 256;     // compute units of the device
  int device = 0;     // the device the handle's memory and stream belong to
  double tuned_ms = 0.0;
  // hipGraph of graph_passes passes of the HBM path (rebuilt when the work
  // list, the series buffer or the schedule change)
  hipGraphExec_t graph_exec = nullptr;
  bool graph_dirty = true;

  // FCC cluster expansion (CHECKERBOARD with lattice FCC; shares cp/d_cp/d_cs)
  double *d_ftab = nullptr;  // FCC_NTAB doubles per chain
  uint8_t *d_tri = nullptr;  // 4096-entry triangle-count table (and the bit-line order one)
  uint8_t *d_tri_lines = nullptr;
  FccSched *d_fcc_sched = nullptr;  // per chain: the class steps of a pass
  // bit-line kernel: every chain L0 <= 32, lattice in LDS as one word per line
  bool fcc_bits = false;
  bool fcc_wide = false;  // bit-line kernel with lines of up to 64 sites
  int fcc_bits_lds = 0;
  // class-line kernel (narrow bit lines, schedule line_tiles = 0): threads per
  // workgroup (the most lines of any step, rounded up to a wavefront, at most
  // 512) and the dynamic LDS it requests (raised so that one chain takes a CU
  // when the batch has no more chains than the device has CUs)
  int fcc_line_block = 64;
  int fcc_line_lds = 0;
  bool fcc() const { return lattice == MCG_LATTICE_FCC; }

  ~mcg_handle() {
    // teardown: errors here have nowhere to go
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (stream2) (void)hipStreamSynchronize(stream2);
    void *ptrs[] = {d_lat, d_active, d_series, d_stats, d_eq, d_xp, d_xs, d_cp, d_cs, d_tab,
                    d_work, d_ftab, d_tri, d_tri_lines, d_band, d_pair, d_kp, d_list, d_np,
                    d_mask_lds, d_mask_even, d_ring_list, d_batch, d_op, d_ot, d_kp_kmc,
                    d_atoms, d_fcc_sched, d_hop, d_hop_cnt};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int64_t sites(int64_t c) const {
    return (fcc() ? 4 : 1) * chains[c].shape[0] * chains[c].shape[1] * chains[c].shape[2];
  }
  // attempts per pass: every site, or in OCCUPANT the mutating sites
  int64_t per_pass(int64_t c) const { return occupant() ? (int64_t)op[c].n_mut : sites(c); }
};

namespace {

// Host<->device copies and fills go through the handle's own non-blocking
// stream (the legacy stream would not be ordered with the handle's kernels).
hipError_t copy_sync(mcg_handle *h, void *dst, const void *src, size_t n, hipMemcpyKind k) {
  hipError_t e = hipMemcpyAsync(dst, src, n, k, h->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(h->stream);
}

hipError_t fill_sync(mcg_handle *h, void *dst, int v, size_t n) {
  hipError_t e = hipMemsetAsync(dst, v, n, h->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(h->stream);
}

int ensure_series(mcg_handle *h, int64_t need) {
  if (need <= h->series_cap) return MCG_OK;
  int64_t cap = std::max<int64_t>(need, h->series_cap * 2);
  cap = std::max<int64_t>(cap, 64);
  mcg_sample *d = nullptr;
  size_t bytes = sizeof(mcg_sample) * (size_t)cap * h->n_chains;
  if (hipMalloc(&d, bytes) != hipSuccess)
    return fail(MCG_ERR_CAPACITY, "out of device memory for sample series");
  if (h->d_series) {
    HIP_TRY(hipMemcpy2DAsync(d, sizeof(mcg_sample) * cap, h->d_series,
                             sizeof(mcg_sample) * h->series_cap,
                             sizeof(mcg_sample) * h->series_cap, h->n_chains,
                             hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    hipFree(h->d_series);
  }
  h->d_series = d;
  h->graph_dirty = true;  // graph nodes hold the old buffer and capacity
  h->series_cap = cap;
  return MCG_OK;
}

// Parameters of one EXACT chain: geometry, speculative window and the
// per-class acceptance tables (returns the LDS bytes the chain needs).
int fill_exact_params(mcg_handle *h, int c, ExactParams &P) {
  {
    auto const &ch = h->chains[c];
    std::memset(&P, 0, sizeof(P));
    P.rows = (int32_t)ch.shape[0];
    P.cols = (int32_t)ch.shape[1];
    P.n = (int32_t)(ch.shape[0] * ch.shape[1]);
    P.fold = h->use_nlist ? 0 : 1;
    P.self_nb = 2 * (P.rows == 1) + 2 * (P.cols == 1);
    // speculative window: ~5*window/n of the lanes see another lane's site in
    // their neighbourhood; keep that near 2%
    P.window = (int32_t)std::min<int64_t>(64, std::max<int64_t>(4, P.n / 256));
    P.div_magic = ((1ULL << 40) + (uint64_t)P.rows - 1) / (uint64_t)P.rows;
    P.negJ = -h->J1;
    P.lat_off = h->lat_off[c];
    const double beta = 1.0 / (h->kb * ch.temperature);  // basic_occupation_metropolis.hh:361
    const double mu = ch.exchange_potential;
    for (int sidx = 0; sidx < 2; ++sidx) {
      const int s_old = sidx ? 1 : -1;
      const int new_occ = -s_old;
      for (int m = 0; m < 5; ++m) {
        const int nsum = 2 * m - 4;
        const int k = sidx * 5 + m;
        // model.hh:312-314: -J * (new - old) * (sum of 4 neighbours)
        const double dEf = -h->J1 * (new_occ - s_old) * nsum;
        // model.hh:432 and basic_semigrand_canonical.hh:188-191
        const double Ndx = (new_occ - s_old) / 2.0;
        const double dE = dEf - mu * Ndx;
        P.dq[k] = dE / (double)P.n;  // occupation_metropolis.hh:141
        if (dE < 0.0) {
          P.neg_mask |= 1u << k;
          continue;
        }
        const double prob = std::exp(-dE * beta);  // metropolis.hh:33
        uint64_t thr = 0;
        if (threshold_for(prob, &thr))
          P.thr[k] = thr;
        else
          P.full_mask |= 1u << k;
      }
    }
    return (int)(((P.n + 15) & ~15) + 4 * (P.rows + P.cols));
  }
}

// Device masks of the active EXACT chains whose lattice fits the block walk's
// LDS and of the others (scalar walk on the HBM lattice).
int exact_update_masks(mcg_handle *h) {
  std::vector<uint8_t> m(2 * (size_t)h->n_chains);
  for (int c = 0; c < h->n_chains; ++c) {
    m[c] = h->active[c] && h->xfit[c];
    m[h->n_chains + c] = h->active[c] && !h->xfit[c];
  }
  HIP_TRY(copy_sync(h, h->d_mask_lds, m.data(), 2 * (size_t)h->n_chains, hipMemcpyHostToDevice));
  return MCG_OK;
}

// Set chain c's EXACT path from its parameters: the block walk in LDS when its
// lattice fits, else the scalar walk on HBM (fold scratch in LDS).
int exact_choose_path(mcg_handle *h, int c, int lds) {
  auto const &P = h->xp[c];
  h->xlds[c] = lds;
  h->xfit[c] = lds <= EXACT_LDS_MAX;
  if (!h->xfit[c] && 4 * (P.rows + P.cols) > EXACT_LDS_MAX)
    return fail(MCG_ERR_ARG, "lattice too large for EXACT mode");
  return MCG_OK;
}

int build_exact(mcg_handle *h) {
  h->xp.resize(h->n_chains);
  h->xlds.assign(h->n_chains, 0);
  h->xfit.assign(h->n_chains, 1);
  for (int c = 0; c < h->n_chains; ++c) {
    const int rc = exact_choose_path(h, c, fill_exact_params(h, c, h->xp[c]));
    if (rc != MCG_OK) return rc;
  }
  HIP_TRY(exact_set_max_lds((int)LDS_LIMIT - 4096));
  HIP_TRY(hipMalloc(&h->d_mask_lds, 2 * (size_t)h->n_chains));
  h->d_mask_hbm = h->d_mask_lds + h->n_chains;
  {
    const int rc = exact_update_masks(h);
    if (rc != MCG_OK) return rc;
  }
  if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(MCG_ERR_HIP, "second stream for the HBM-lattice chains");
  HIP_TRY(hipMalloc(&h->d_xp, sizeof(ExactParams) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_xs, sizeof(ExactState) * h->n_chains));
  HIP_TRY(copy_sync(h, h->d_xp, h->xp.data(), sizeof(ExactParams) * h->n_chains,
                    hipMemcpyHostToDevice));
  std::vector<ExactState> xs(h->n_chains);
  for (int c = 0; c < h->n_chains; ++c) {
    ExactState &S = xs[c];
    std::memset(&S, 0, sizeof(S));
    int pos = 0;
    mt_seed(h->chains[c].seed, S.mt, &pos);
    S.pos = pos;
  }
  HIP_TRY(copy_sync(h, h->d_xs, xs.data(), sizeof(ExactState) * h->n_chains,
                    hipMemcpyHostToDevice));
  return MCG_OK;
}

// Parameters of one CANONICAL chain with nA sites of species 0 (spin -1): the
// 25 swap classes (Sa, Sb) of the pair delta model.hh:354-379 (site a
// -1 -> +1 first, then site b +1 -> -1 with a already flipped).
void fill_canonical_params(mcg_handle *h, int c, int64_t nA, CanonParams &P) {
  auto const &ch = h->chains[c];
  std::memset(&P, 0, sizeof(P));
  P.rows = (int32_t)ch.shape[0];
  P.cols = (int32_t)ch.shape[1];
  P.n = (int32_t)(ch.shape[0] * ch.shape[1]);
  P.nA = (int32_t)nA;
  P.nB = P.n - P.nA;
  P.fold = h->use_nlist ? 0 : 1;
  P.self_nb = 2 * (P.rows == 1) + 2 * (P.cols == 1);
  P.window = (h->options & MCG_OPT_EXACT_SCALAR_WALK) ? 1 : 64;
  uint32_t marks = 64;
  while (marks < (uint32_t)P.n && marks < 8192u) marks <<= 1;
  P.mark_mask = marks - 1;
  P.div_magic = ((1ULL << 40) + (uint64_t)P.rows - 1) / (uint64_t)P.rows;
  P.negJ = -h->J1;
  P.lat_off = h->lat_off[c];
  const double beta = 1.0 / (h->kb * ch.temperature);  // occupation_metropolis.hh:108-109
  for (int ia = 0; ia < 5; ++ia)
    for (int ib = 0; ib < 5; ++ib) {
      const int Sa = 2 * ia - 4, Sb = 2 * ib - 4, k = ia * 5 + ib;
      // model.hh:305-314 on spins: -J * (new - old) * (neighbour sum), summed
      // from 0.0 in event order (model.hh:362-369)
      double dE = 0.0;
      dE += -h->J1 * (1 - (-1)) * Sa;
      dE += -h->J1 * ((-1) - 1) * Sb;
      P.dq[k] = dE / (double)P.n;  // occupation_metropolis.hh:141
      if (dE < 0.0) {
        P.neg_mask |= 1u << k;
        continue;
      }
      uint64_t thr = 0;
      if (threshold_for(std::exp(-dE * beta), &thr))  // metropolis.hh:33
        P.thr[k] = thr;
      else
        P.full_mask |= 1u << k;
    }
}

int build_canonical(mcg_handle *h) {
  h->kp.resize(h->n_chains);
  int max_lds = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    fill_canonical_params(h, c, 0, h->kp[c]);
    max_lds = std::max(max_lds, canon_lds_bytes(h->kp[c]));
  }
  if (max_lds > EXACT_LDS_MAX)
    return fail(MCG_ERR_ARG,
                "CANONICAL mode keeps a chain's lattice, lists and conflict marks in LDS: "
                "lattice too large");
  h->canon_lds = max_lds;
  HIP_TRY(canonical_set_max_lds((int)EXACT_LDS_MAX));
  HIP_TRY(hipMalloc(&h->d_kp, sizeof(CanonParams) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_xs, sizeof(ExactState) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_list, sizeof(int32_t) * (size_t)h->lat_bytes));
  HIP_TRY(copy_sync(h, h->d_kp, h->kp.data(), sizeof(CanonParams) * h->n_chains,
                    hipMemcpyHostToDevice));
  std::vector<ExactState> xs(h->n_chains);
  for (int c = 0; c < h->n_chains; ++c) {
    ExactState &S = xs[c];
    std::memset(&S, 0, sizeof(S));
    int pos = 0;
    mt_seed(h->chains[c].seed, S.mt, &pos);
    S.pos = pos;
  }
  HIP_TRY(copy_sync(h, h->d_xs, xs.data(), sizeof(ExactState) * h->n_chains,
                    hipMemcpyHostToDevice));
  return MCG_OK;
}

// OccLocation::initialize (OccLocation.cc:39-118) for one chain: both lists in
// site order; uploads them and the list sizes.
int canonical_init_lists(mcg_handle *h, int64_t c, std::vector<int8_t> const &occ) {
  std::vector<int32_t> lists;
  lists.reserve(occ.size());
  for (int64_t l = 0; l < (int64_t)occ.size(); ++l)
    if (occ[l] < 0) lists.push_back((int32_t)l);
  const int64_t nA = (int64_t)lists.size();
  for (int64_t l = 0; l < (int64_t)occ.size(); ++l)
    if (occ[l] > 0) lists.push_back((int32_t)l);
  HIP_TRY(copy_sync(h, h->d_list + h->lat_off[c], lists.data(), sizeof(int32_t) * lists.size(),
                    hipMemcpyHostToDevice));
  fill_canonical_params(h, (int)c, nA, h->kp[c]);
  HIP_TRY(copy_sync(h, h->d_kp + c, &h->kp[c], sizeof(CanonParams), hipMemcpyHostToDevice));
  return MCG_OK;
}

// ---------------------------------------------------------------------------
// OCCUPANT: the system's units and candidates, per-chain list segments and
// class tables, and the conversion between the candidate-order lists of the
// C ABI (OccLocation's m_loc) and the device's segments.
// ---------------------------------------------------------------------------

// Geometry, list segments (units in order, sized by their mutating sites),
// candidates and swap types of one chain.
void fill_occupant_params(mcg_handle *h, int c, OccParams &P) {
  auto const &ch = h->chains[c];
  std::memset(&P, 0, sizeof(P));
  P.rows = (int32_t)ch.shape[0];
  P.cols = (int32_t)ch.shape[1];
  P.n = (int32_t)(ch.shape[0] * ch.shape[1]);
  P.canonical = h->occ_canonical;
  P.n_asym = h->occ_n_asym;
  P.n_cand = (int32_t)h->occ_cand_asym.size();
  P.n_swaps = (int32_t)h->occ_swap_a.size();
  P.window = (h->options & MCG_OPT_EXACT_SCALAR_WALK) ? 1 : 64;
  P.div_magic = ((1ULL << 40) + (uint64_t)P.rows - 1) / (uint64_t)P.rows;
  auto const &ua = h->site_asym[c];
  int32_t off = 0;
  for (int b = 0; b < h->occ_n_asym; ++b) {
    int32_t size = 0;
    if (h->occ_size[b] == 2)
      for (int8_t u : ua) size += u == b ? 1 : 0;
    P.asym_off[b] = off;
    P.asym_size[b] = size;
    off += size;
  }
  P.n_mut = off;
  for (int i = 0; i < P.n_cand; ++i) {
    P.cand_asym[i] = h->occ_cand_asym[i];
    P.cand_occ[i] = h->occ_cand_occ[i];
  }
  for (int i = 0; i < P.n_swaps; ++i) {
    P.swap_a[i] = h->occ_swap_a[i];
    P.swap_b[i] = h->occ_swap_b[i];
  }
  P.lat_off = h->lat_off[c];
}

// Class tables of one chain: per swap type and neighbour sums, the potential
// delta in the order of its definition (per transformed site,
// dE = (dE + (-J ((2 o' - 1) - (2 o - 1)) * sum)) - (mu[to] - mu[from]);
// the single-site term as model.hh:305-314 writes it), fl(dE / n_unitcells)
// (occupation_metropolis.hh:141) and the acceptance threshold of
// metropolis.hh:25-35.
void fill_occupant_tables(mcg_handle *h, int c, OccTables &T) {
  std::memset(&T, 0, sizeof(T));
  auto const &ch = h->chains[c];
  const double beta = 1.0 / (h->kb * ch.temperature);  // occupation_metropolis.hh:108-109
  const double n_uc = static_cast<double>(h->sites(c) / h->occ_spu);
  const double *mu = h->occ_mu.data() + (size_t)c * h->occ_n_species;
  const int nb_max = h->occ_canonical ? 5 : 1;
  for (size_t i = 0; i < h->occ_swap_a.size(); ++i) {
    const int ca = h->occ_swap_a[i], cb = h->occ_swap_b[i];
    const int ua = h->occ_cand_asym[ca], oa = h->occ_cand_occ[ca];
    const int ub = h->occ_cand_asym[cb], ob = h->occ_cand_occ[cb];
    const int from_a = h->occ_species[ua][oa], to_a = h->occ_species[ua][1 - oa];
    const int from_b = h->occ_species[ub][ob], to_b = h->occ_species[ub][1 - ob];
    for (int ia = 0; ia < 5; ++ia)
      for (int ib = 0; ib < nb_max; ++ib) {
        const int Sa = 2 * ia - 4, Sb = 2 * ib - 4;
        double dE = 0.0;
        dE = (dE + -h->J1 * ((2 * (1 - oa) - 1) - (2 * oa - 1)) * Sa) - (mu[to_a] - mu[from_a]);
        if (h->occ_canonical)
          dE = (dE + -h->J1 * ((2 * (1 - ob) - 1) - (2 * ob - 1)) * Sb) - (mu[to_b] - mu[from_b]);
        const int k = (int)i * 25 + ia * 5 + ib;
        T.dq[k] = dE / n_uc;
        if (dE < 0.0) {
          T.flags[k] = 1;
          continue;
        }
        uint64_t thr = 0;
        if (threshold_for(std::exp(-dE * beta), &thr))
          T.thr[k] = thr;
        else
          T.flags[k] = 2;
      }
  }
}

// The device segments of candidate-order lists (each list's entries in
// order), and back.
void occ_lists_to_segments(OccParams const &P, std::vector<std::vector<int32_t>> const &lists,
                           std::vector<int32_t> &seg) {
  seg.assign(P.n_mut, 0);
  for (int ci = 0; ci < P.n_cand; ++ci) {
    const int b = P.cand_asym[ci], o = P.cand_occ[ci];
    for (size_t j = 0; j < lists[ci].size(); ++j)
      seg[o ? P.asym_off[b] + P.asym_size[b] - 1 - (int32_t)j : P.asym_off[b] + (int32_t)j] =
          lists[ci][j];
  }
}

void occ_segments_to_lists(OccParams const &P, const int32_t *cnt0,
                           std::vector<int32_t> const &seg,
                           std::vector<std::vector<int32_t>> &lists) {
  lists.assign(P.n_cand, {});
  for (int ci = 0; ci < P.n_cand; ++ci) {
    const int b = P.cand_asym[ci], o = P.cand_occ[ci];
    const int32_t cnt = o ? P.asym_size[b] - cnt0[b] : cnt0[b];
    for (int32_t j = 0; j < cnt; ++j)
      lists[ci].push_back(seg[o ? P.asym_off[b] + P.asym_size[b] - 1 - j : P.asym_off[b] + j]);
  }
}

// candidate index of (unit, occupant), -1 for units with one occupant
int occ_cand_of(mcg_handle const *h, int b, int o) {
  for (size_t i = 0; i < h->occ_cand_asym.size(); ++i)
    if (h->occ_cand_asym[i] == b && h->occ_cand_occ[i] == o) return (int)i;
  return -1;
}

// Upload candidate-order lists and the matching occupant-0 counts of one chain.
int occupant_upload_lists(mcg_handle *h, int64_t c, std::vector<std::vector<int32_t>> const &lists) {
  OccParams const &P = h->op[c];
  std::vector<int32_t> seg;
  occ_lists_to_segments(P, lists, seg);
  if (!seg.empty())
    HIP_TRY(copy_sync(h, h->d_list + h->lat_off[c], seg.data(), sizeof(int32_t) * seg.size(),
                      hipMemcpyHostToDevice));
  ExactState st;
  HIP_TRY(copy_sync(h, &st, h->d_xs + c, sizeof(st), hipMemcpyDeviceToHost));
  for (int b = 0; b < OCC_MAX_ASYM; ++b) st.occ_cnt0[b] = 0;
  for (int ci = 0; ci < P.n_cand; ++ci)
    if (P.cand_occ[ci] == 0) st.occ_cnt0[P.cand_asym[ci]] = (int32_t)lists[ci].size();
  HIP_TRY(copy_sync(h, h->d_xs + c, &st, sizeof(st), hipMemcpyHostToDevice));
  return MCG_OK;
}

// OccLocation::initialize (OccLocation.cc:39-118) for one chain: every
// mutating site, in site order, at the back of its candidate's list.
int occupant_init_lists(mcg_handle *h, int64_t c, std::vector<int8_t> const &occ) {
  std::vector<std::vector<int32_t>> lists(h->op[c].n_cand);
  auto const &ua = h->site_asym[c];
  for (int64_t l = 0; l < (int64_t)occ.size(); ++l) {
    const int ci = occ_cand_of(h, ua[l], occ[l]);
    if (ci >= 0) lists[ci].push_back((int32_t)l);
  }
  return occupant_upload_lists(h, c, lists);
}

// Parameters, tables, state and lists of one chain (occupation all occupant 0).
int occupant_setup_chain(mcg_handle *h, int c, bool fresh_state) {
  OccParams P;
  fill_occupant_params(h, c, P);
  if (occ_lds_bytes(P) > h->occ_lds)
    return fail(MCG_ERR_CAPACITY, "shape needs more LDS than the handle was created with");
  h->op[c] = P;
  HIP_TRY(copy_sync(h, h->d_op + c, &P, sizeof(P), hipMemcpyHostToDevice));
  auto T = std::make_unique<OccTables>();
  fill_occupant_tables(h, c, *T);
  HIP_TRY(copy_sync(h, h->d_ot + c, T.get(), sizeof(OccTables), hipMemcpyHostToDevice));
  if (fresh_state) {
    ExactState S;
    std::memset(&S, 0, sizeof(S));
    int pos = 0;
    mt_seed(h->chains[c].seed, S.mt, &pos);
    S.pos = pos;
    HIP_TRY(copy_sync(h, h->d_xs + c, &S, sizeof(S), hipMemcpyHostToDevice));
  }
  return MCG_OK;
}

int build_occupant(mcg_handle *h) {
  h->op.resize(h->n_chains);
  h->site_asym.resize(h->n_chains);
  int max_lds = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    h->site_asym[c].assign(h->sites(c), 0);
    OccParams P;
    fill_occupant_params(h, c, P);
    // room for every site in a list, whatever the units
    P.n_mut = P.n;
    max_lds = std::max(max_lds, occ_lds_bytes(P));
  }
  if (max_lds > OCC_LDS_MAX)
    return fail(MCG_ERR_ARG, "OCCUPANT mode keeps a chain's lattice and lists in LDS: "
                             "lattice too large");
  h->occ_lds = max_lds;
  HIP_TRY(occupant_set_max_lds((int)OCC_LDS_MAX));
  HIP_TRY(hipMalloc(&h->d_op, sizeof(OccParams) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_ot, sizeof(OccTables) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_xs, sizeof(ExactState) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_list, sizeof(int32_t) * (size_t)h->lat_bytes));
  for (int c = 0; c < h->n_chains; ++c) {
    const int rc = occupant_setup_chain(h, c, true);
    if (rc != MCG_OK) return rc;
  }
  return MCG_OK;
}

// Check and copy the OCCUPANT system of a descriptor.
int occupant_configure(mcg_handle *h, const mcg_desc *desc) {
  const mcg_occupant_desc *o = desc->occupant;
  if (!o) return fail(MCG_ERR_ARG, "OCCUPANT mode needs mcg_desc.occupant");
  if (o->n_asym < 1 || o->n_asym > OCC_MAX_ASYM || !o->occ_size || !o->species)
    return fail(MCG_ERR_ARG, "OCCUPANT: 1..8 asymmetric units with occ_size and species");
  if (o->n_species < 1 || o->n_species > 8)
    return fail(MCG_ERR_ARG, "OCCUPANT: 1..8 species");
  if (o->n_swaps < 1 || o->n_swaps > OCC_MAX_SWAPS || !o->swaps)
    return fail(MCG_ERR_ARG, "OCCUPANT: 1..16 swap types");
  if (o->sites_per_unitcell < 1) return fail(MCG_ERR_ARG, "OCCUPANT: sites_per_unitcell < 1");
  h->occ_n_asym = (int)o->n_asym;
  h->occ_n_species = (int)o->n_species;
  h->occ_canonical = o->canonical ? 1 : 0;
  h->occ_spu = o->sites_per_unitcell;
  for (int b = 0; b < h->occ_n_asym; ++b) {
    const int64_t k = o->occ_size[b];
    if (k != 1 && k != 2) return fail(MCG_ERR_ARG, "OCCUPANT: a unit allows 1 or 2 occupants");
    h->occ_size[b] = (int)k;
    for (int j = 0; j < k; ++j) {
      const int64_t sp = o->species[2 * b + j];
      if (sp < 0 || sp >= o->n_species) return fail(MCG_ERR_ARG, "OCCUPANT: species out of range");
      h->occ_species[b][j] = (int)sp;
    }
    if (k == 2 && h->occ_species[b][0] == h->occ_species[b][1])
      return fail(MCG_ERR_ARG, "OCCUPANT: a unit's two occupants must be different species");
    if (k == 2)
      for (int j = 0; j < 2; ++j) {  // OccCandidateList order (OccCandidate.cc:31-52)
        h->occ_cand_asym.push_back(b);
        h->occ_cand_occ.push_back(j);
      }
  }
  const int nc = (int)h->occ_cand_asym.size();
  auto species_of = [&](int ci) { return h->occ_species[h->occ_cand_asym[ci]][h->occ_cand_occ[ci]]; };
  auto other_of = [&](int ci) { return h->occ_species[h->occ_cand_asym[ci]][1 - h->occ_cand_occ[ci]]; };
  for (int64_t i = 0; i < o->n_swaps; ++i) {
    const int64_t a = o->swaps[2 * i], b = o->swaps[2 * i + 1];
    if (a < 0 || a >= nc || b < 0 || b >= nc)
      return fail(MCG_ERR_ARG, "OCCUPANT: swap candidate out of range");
    const bool ok = h->occ_canonical
                        // allowed_canonical_swap (OccCandidate.cc:90-97)
                        ? (species_of((int)a) != species_of((int)b) &&
                           other_of((int)a) == species_of((int)b) &&
                           other_of((int)b) == species_of((int)a))
                        // allowed_semigrand_canonical_swap (OccCandidate.cc:126-133)
                        : (h->occ_cand_asym[a] == h->occ_cand_asym[b] && a != b);
    if (!ok)
      return fail(MCG_ERR_ARG, h->occ_canonical
                                   ? "OCCUPANT: not an allowed canonical swap"
                                   : "OCCUPANT: not an allowed semi-grand canonical swap");
    h->occ_swap_a.push_back((int32_t)a);
    h->occ_swap_b.push_back((int32_t)b);
  }
  h->occ_mu.assign((size_t)desc->n_chains * o->n_species, 0.0);
  if (o->exchange_potential)
    std::copy(o->exchange_potential, o->exchange_potential + h->occ_mu.size(), h->occ_mu.begin());
  return MCG_OK;
}

// ---------------------------------------------------------------------------
// KMC: lattice-gas kinetic Monte Carlo with atom tracking.
// ---------------------------------------------------------------------------

// Parameters of one KMC chain with n_atoms atoms: geometry, the attempt rate
// R = 4 n_atoms and the seven hop classes' acceptance (dE = J (a - b),
// class a - b + 3: accepted without a draw when dE <= 0, else when
// random_real(1.0) < exp(-beta dE)).
void fill_kmc_params(mcg_handle *h, int c, int64_t n_atoms, KmcParams &P) {
  auto const &ch = h->chains[c];
  std::memset(&P, 0, sizeof(P));
  P.rows = (int32_t)ch.shape[0];
  P.cols = (int32_t)ch.shape[1];
  P.n = (int32_t)(ch.shape[0] * ch.shape[1]);
  P.n_atoms = (int32_t)n_atoms;
  P.window = (h->options & MCG_OPT_EXACT_SCALAR_WALK) ? 1 : 64;
  P.div_magic = ((1ULL << 40) + (uint64_t)P.rows - 1) / (uint64_t)P.rows;
  P.attempt_rate = 4.0 * (double)n_atoms;
  P.time_period = h->sample_time_period;
  P.J = h->J1;
  P.spin_limit = h->kmc_spin_limit;
  P.lat_off = h->lat_off[c];
  P.hop_every = h->kmc_hop_every;
  P.hop_max = h->kmc_hop_max;
  P.hop_off = (int64_t)h->kmc_hop_max * h->lat_off[c];
  const double beta = 1.0 / (h->kb * ch.temperature);
  for (int k = 0; k < KMC_NCLASS; ++k) {
    const double dE = h->J1 * (double)(k - 3);
    if (dE <= 0.0) continue;
    P.draw_mask |= 1u << k;
    uint64_t thr = 0;
    if (threshold_for(std::exp(-dE * beta), &thr))
      P.thr[k] = thr;
    else
      P.full_mask |= 1u << k;
  }
}

// OccLocation::initialize with atoms (OccLocation.cc:39-118) for one chain:
// vacancies and atoms in site order (the vacancy list from the front of the
// list array, the atom list from the back), atoms numbered in site order with
// no displacement; the loop's time and sums start again.
int kmc_init(mcg_handle *h, int64_t c, std::vector<int8_t> const &occ) {
  const int64_t n = (int64_t)occ.size();
  std::vector<int32_t> seg(n), atoms;
  int64_t k0 = 0, k1 = 0;
  for (int64_t l = 0; l < n; ++l) {
    if (occ[l]) {
      seg[n - 1 - k1++] = (int32_t)l;
      atoms.insert(atoms.end(), {(int32_t)l, 0, 0, 0});
    } else {
      seg[k0++] = (int32_t)l;
    }
  }
  HIP_TRY(copy_sync(h, h->d_list + h->lat_off[c], seg.data(), sizeof(int32_t) * n,
                    hipMemcpyHostToDevice));
  if (!atoms.empty())
    HIP_TRY(copy_sync(h, h->d_atoms + 4 * h->lat_off[c], atoms.data(),
                      sizeof(int32_t) * atoms.size(), hipMemcpyHostToDevice));
  fill_kmc_params(h, (int)c, k1, h->kp_kmc[c]);
  HIP_TRY(copy_sync(h, h->d_kp_kmc + c, &h->kp_kmc[c], sizeof(KmcParams), hipMemcpyHostToDevice));
  if (h->d_hop_cnt)  // the chain's hop correlations start again with its atoms
    HIP_TRY(fill_sync(h, h->d_hop_cnt + h->lat_off[c], 0, sizeof(int32_t) * n));
  ExactState st;
  HIP_TRY(copy_sync(h, &st, h->d_xs + c, sizeof(st), hipMemcpyDeviceToHost));
  st.time = 0.0;
  st.next_k = 0;
  st.kmc_k = 0;
  st.kmc_sum_sq = st.kmc_sum_dx = st.kmc_sum_dy = 0;
  st.frozen = 0;
  st.pending_site = -1;
  st.pending_dt = 0.0;
  st.kmc_pdir = st.kmc_pk = 0;
  int64_t S = 0;
  const int64_t rows = h->chains[c].shape[0], cols = h->chains[c].shape[1];
  for (int64_t l = 0; l < n; ++l) {
    const int64_t i = l % rows, j = l / rows;
    if (occ[l]) S += occ[(i + 1) % rows + rows * j] + occ[i + rows * ((j + 1) % cols)];
  }
  st.S = S;
  st.M = k1;
  st.e_run = 0.0;
  HIP_TRY(copy_sync(h, h->d_xs + c, &st, sizeof(st), hipMemcpyHostToDevice));
  return MCG_OK;
}

int build_kmc(mcg_handle *h) {
  h->kp_kmc.resize(h->n_chains);
  int max_lds = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    fill_kmc_params(h, c, 0, h->kp_kmc[c]);
    max_lds = std::max(max_lds, kmc_lds_bytes((int)h->sites(c)));
  }
  if (max_lds > OCC_LDS_MAX)
    return fail(MCG_ERR_ARG, "KMC mode keeps a chain's lists and atoms in LDS: lattice too large");
  h->kmc_lds = max_lds;
  HIP_TRY(kmc_set_max_lds((int)OCC_LDS_MAX));
  HIP_TRY(hipMalloc(&h->d_kp_kmc, sizeof(KmcParams) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_xs, sizeof(ExactState) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_list, sizeof(int32_t) * (size_t)h->lat_bytes));
  HIP_TRY(hipMalloc(&h->d_atoms, 4 * sizeof(int32_t) * (size_t)h->lat_bytes));
  std::vector<ExactState> xs(h->n_chains);
  for (int c = 0; c < h->n_chains; ++c) {
    std::memset(&xs[c], 0, sizeof(ExactState));
    int pos = 0;
    mt_seed(h->chains[c].seed, xs[c].mt, &pos);
    xs[c].pos = pos;
  }
  HIP_TRY(copy_sync(h, h->d_xs, xs.data(), sizeof(ExactState) * h->n_chains,
                    hipMemcpyHostToDevice));
  return MCG_OK;
}

// KMC runs of n_passes passes (n events each) plus n_steps events per active
// chain, a pending event first; an event reaching stop_time stays pending,
// and select_at_end leaves one more selected.  Fails before running when a
// chain has no atom or no vacancy (no hop is possible), and after running
// when a chain froze (no event in the spin limit's attempts) or when the time
// grid's records overflowed the series.
int kmc_run(mcg_handle *h, int64_t n_passes, int64_t n_steps,
            double stop_time = std::numeric_limits<double>::infinity(), bool select_at_end = false) {
  h->last_ms = 0.0;
  h->last_attempts = 0.0;
  if (n_passes == 0 && n_steps == 0 && !select_at_end) return MCG_OK;
  const bool grid = h->sample_time_period > 0.0;
  int64_t need = 0;
  double events = 0.0;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    const int64_t n = h->sites(c), na = h->kp_kmc[c].n_atoms;
    if (na == 0 || na == n)
      return fail(MCG_ERR_STATE, "KMC: no hop is possible (no atom or no vacancy)");
    const int64_t p0 = h->n_pass[c];
    const int64_t passes = n_passes + (h->n_step[c] + n_steps + 1) / n;
    if (!grid)
      need = std::max(need, h->n_samples[c] + (p0 + passes) / h->sample_period -
                                p0 / h->sample_period);
    events += (double)n * (double)n_passes + (double)n_steps;
  }
  int rc = ensure_series(h, need);
  if (rc != MCG_OK) return rc;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  launch_kmc(h->d_kp_kmc, h->d_xs, h->d_lat, h->d_list, h->d_atoms, h->d_hop, h->d_hop_cnt,
             h->d_series, h->series_cap,
             h->d_active, h->n_chains, n_passes, n_steps, h->sample_period, stop_time,
             select_at_end ? 1 : 0, h->kmc_lds, h->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_attempts = events;
  std::vector<ExactState> xs(h->n_chains);
  HIP_TRY(copy_sync(h, xs.data(), h->d_xs, sizeof(ExactState) * h->n_chains,
                    hipMemcpyDeviceToHost));
  bool frozen = false;
  int64_t over = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    h->n_pass[c] = xs[c].n_pass;
    h->n_step[c] = xs[c].step;
    over = std::max(over, xs[c].n_samples - h->series_cap);
    h->n_samples[c] = std::min(xs[c].n_samples, h->series_cap);
    frozen = frozen || xs[c].frozen;
  }
  if (frozen)
    return fail(MCG_ERR_STATE,
                "KMC: a chain made no event in the spin limit's attempts (every hop of its "
                "state has a vanishing probability at this temperature); it stopped there");
  if (over > 0)
    return fail(MCG_ERR_CAPACITY,
                "KMC time-grid samples exceeded the sample series by " + std::to_string(over) +
                    " records (they were dropped); reserve more with mcg_reserve_samples");
  return MCG_OK;
}

// Parameters of one NFOLD chain: geometry and the ten class rates
// min(1, exp(-beta dE_k)) for the SGC potential delta of class k (as the
// exact mode's classes, basic_semigrand_canonical.hh:178-192).
void fill_nfold_params(mcg_handle *h, int c, NfoldParams &P) {
  auto const &ch = h->chains[c];
  std::memset(&P, 0, sizeof(P));
  P.rows = (int32_t)ch.shape[0];
  P.cols = (int32_t)ch.shape[1];
  P.n = (int32_t)(ch.shape[0] * ch.shape[1]);
  P.div_magic = ((1ULL << 40) + (uint64_t)P.rows - 1) / (uint64_t)P.rows;
  P.time_period = h->sample_time_period;
  P.lat_off = h->lat_off[c];
  P.sample_start = (h->options & MCG_OPT_NFOLD_SAMPLE_START) ? 1 : 0;
  const double beta = 1.0 / (h->kb * ch.temperature);
  const double mu = ch.exchange_potential;
  for (int sidx = 0; sidx < 2; ++sidx)
    for (int m = 0; m < 5; ++m) {
      const int s_old = sidx ? 1 : -1, new_occ = -s_old, nsum = 2 * m - 4;
      const double dEf = -h->J1 * (new_occ - s_old) * nsum;
      const double Ndx = (new_occ - s_old) / 2.0;
      const double dE = dEf - mu * Ndx;
      P.rate[sidx * 5 + m] = dE < 0.0 ? 1.0 : std::exp(-dE * beta);
    }
}

void nfold_reset_state(ExactState &S, uint64_t seed) {
  std::memset(&S, 0, sizeof(S));
  int pos = 0;
  mt_seed(seed, S.mt, &pos);
  S.pos = pos;
  S.pending_site = -1;
  S.next_k = 1;
}

int build_nfold(mcg_handle *h) {
  h->np_.resize(h->n_chains);
  int max_lds = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    fill_nfold_params(h, c, h->np_[c]);
    max_lds = std::max(max_lds, nfold_lds_bytes(h->np_[c]));
  }
  if (max_lds > EXACT_LDS_MAX)
    return fail(MCG_ERR_ARG, "NFOLD mode keeps a chain's lattice and class array in LDS: "
                             "lattice too large");
  h->nfold_lds = max_lds;
  HIP_TRY(nfold_set_max_lds((int)EXACT_LDS_MAX));
  HIP_TRY(hipMalloc(&h->d_np, sizeof(NfoldParams) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_xs, sizeof(ExactState) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_list, sizeof(int32_t) * (size_t)h->lat_bytes));
  HIP_TRY(copy_sync(h, h->d_np, h->np_.data(), sizeof(NfoldParams) * h->n_chains,
                    hipMemcpyHostToDevice));
  std::vector<ExactState> xs(h->n_chains);
  for (int c = 0; c < h->n_chains; ++c) nfold_reset_state(xs[c], h->chains[c].seed);
  HIP_TRY(copy_sync(h, h->d_xs, xs.data(), sizeof(ExactState) * h->n_chains,
                    hipMemcpyHostToDevice));
  return MCG_OK;
}

// The class-partitioned site array of an NFOLD chain, classes in order and
// sites in site order within a class; drops the chain's pending event (it was
// selected on the previous occupation).
int nfold_init_array(mcg_handle *h, int64_t c, std::vector<int8_t> const &occ) {
  const int64_t rows = h->chains[c].shape[0], cols = h->chains[c].shape[1], n = rows * cols;
  auto cls = [&](int64_t l) {
    const int64_t i = l % rows, j = l / rows;
    const int nsum = occ[(i + 1) % rows + rows * j] + occ[(i + rows - 1) % rows + rows * j] +
                     occ[i + rows * ((j + 1) % cols)] + occ[i + rows * ((j + cols - 1) % cols)];
    return (occ[l] > 0 ? 5 : 0) + (nsum + 4) / 2;
  };
  std::vector<int32_t> arr;
  arr.reserve(n);
  for (int k = 0; k < 10; ++k)
    for (int64_t l = 0; l < n; ++l)
      if (cls(l) == k) arr.push_back((int32_t)l);
  HIP_TRY(copy_sync(h, h->d_list + h->lat_off[c], arr.data(), sizeof(int32_t) * n,
                    hipMemcpyHostToDevice));
  ExactState st;
  HIP_TRY(copy_sync(h, &st, h->d_xs + c, sizeof(ExactState), hipMemcpyDeviceToHost));
  st.pending_site = -1;
  HIP_TRY(copy_sync(h, h->d_xs + c, &st, sizeof(ExactState), hipMemcpyHostToDevice));
  return MCG_OK;
}

// Entries of a chain's acceptance table (indexed by checker_class).
int checker_ntab(mcg_handle *) { return CHECKER_NTAB; }

// Parameters and acceptance table (ntab entries at tab) of one CHECKERBOARD chain.
int fill_checker_params(mcg_handle *h, int c, CheckerParams &P, uint32_t *tab) {
  const bool nnn = h->J2 != 0.0;
  const int ndim = h->lattice == MCG_LATTICE_SC ? 3 : 2;
  const int z1 = ndim == 3 ? 6 : 4;
  const int z2 = nnn ? (ndim == 3 ? 12 : 4) : 0;
  const int ntab = CHECKER_NTAB;
  std::fill(tab, tab + ntab, 0u);
  {
    auto const &ch = h->chains[c];
    // even edges: parity colourings (checker.hip); an odd edge (>= 3): ring
    // colourings (checker_ring.hip)
    for (int d = 0; d < ndim; ++d)
      if (ch.shape[d] < 2) return fail(MCG_ERR_ARG, "CHECKERBOARD needs edge lengths >= 2");
    std::memset(&P, 0, sizeof(P));
    P.L0 = (int32_t)ch.shape[0];
    P.L1 = (int32_t)ch.shape[1];
    P.L2 = (int32_t)ch.shape[2];
    P.ncolor = h->ncolor;
    P.z1 = z1;
    P.z2 = z2;
    P.ndim = ndim;
    P.key0 = (uint32_t)ch.seed;
    P.key1 = (uint32_t)(ch.seed >> 32);
    P.n = h->sites(c);
    P.lat_off = h->lat_off[c];
    P.tab_off = (int64_t)ntab * c;
    P.quad_only = (h->options & MCG_OPT_CHECKER_QUAD_PATH) ? 1 : 0;
    const double beta = 1.0 / (h->kb * ch.temperature);
    const double mu = ch.exchange_potential;
    for (int sidx = 0; sidx < 2; ++sidx) {
      const int s = sidx ? 1 : -1;
      const double ds = -2.0 * s;
      for (int a = 0; a <= z1; ++a)
        for (int b = 0; b <= z2; ++b) {
          const int nn = 2 * a - z1, nnn_sum = 2 * b - z2;
          const double dE = -h->J1 * ds * nn - h->J2 * ds * nnn_sum - mu * (ds / 2.0);
          const double p = std::exp(-beta * dE);
          uint32_t t;
          if (!(p < 1.0)) {
            t = 0xFFFFFFFFu;
          } else {
            double v = std::ceil(p * 4294967296.0) - 1.0;
            if (v < 0.0) v = 0.0;
            if (v > 4294967295.0) v = 4294967295.0;
            t = (uint32_t)v;
          }
          tab[checker_class(sidx, a, b)] = t;
        }
    }
  }
  return MCG_OK;
}

// A CHECKERBOARD chain with an odd edge runs the ring-coloured kernels.
bool checker_ring_chain(mcg_handle const *h, int c) {
  const int ndim = h->lattice == MCG_LATTICE_SC ? 3 : 2;
  for (int d = 0; d < ndim; ++d)
    if (h->chains[c].shape[d] & 1) return true;
  return false;
}

// Masks of the active parity-coloured and ring-coloured chains, the list of the
// active ring chains and the blocks per chain of their HBM class launches.
int checker_update_masks(mcg_handle *h) {
  std::vector<uint8_t> m(2 * (size_t)h->n_chains);
  std::vector<int32_t> ring;
  int64_t max_units = 1;
  for (int c = 0; c < h->n_chains; ++c) {
    m[c] = h->active[c] && !h->cring[c];
    m[h->n_chains + c] = h->active[c] && h->cring[c];
    if (h->active[c] && h->cring[c]) {
      ring.push_back(c);
      auto const &P = h->cp[c];
      // the largest class: ceil(L0/2)/4 units per line, ceil(L1/2) (x ceil(L2/2)) lines
      const int64_t u = ((P.L0 + 1) / 2 + 3) / 4 * ((P.L1 + 1) / 2) *
                        (P.ndim == 3 ? (P.L2 + 1) / 2 : 1);
      max_units = std::max(max_units, u);
    }
  }
  HIP_TRY(copy_sync(h, h->d_mask_even, m.data(), m.size(), hipMemcpyHostToDevice));
  h->n_ring_active = (int)ring.size();
  // about four units per thread of a 256-thread block
  h->ring_per_chain = (int)std::min<int64_t>(4096, std::max<int64_t>(1, (max_units + 1023) / 1024));
  if (!ring.empty())
    HIP_TRY(copy_sync(h, h->d_ring_list, ring.data(), sizeof(int32_t) * ring.size(),
                      hipMemcpyHostToDevice));
  h->graph_dirty = true;
  return MCG_OK;
}

int build_checker(mcg_handle *h) {
  const bool nnn = h->J2 != 0.0;
  const int ndim = h->lattice == MCG_LATTICE_SC ? 3 : 2;
  h->ncolor = nnn ? (ndim == 3 ? 8 : 4) : 2;
  const int ntab = checker_ntab(h);
  h->cp.resize(h->n_chains);
  std::vector<uint32_t> tabs((size_t)ntab * h->n_chains);
  int64_t max_n = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    int rc = fill_checker_params(h, c, h->cp[c], tabs.data() + (size_t)ntab * c);
    if (rc != MCG_OK) return rc;
    max_n = std::max(max_n, h->cp[c].n);
  }
  h->checker_lds = max_n <= CHECKER_LDS_MAX && !(h->options & MCG_OPT_FORCE_STREAMING);
  h->checker_lds_bytes = (int)((max_n + 15) & ~15);
  HIP_TRY(checker_set_max_lds((int)LDS_LIMIT - 4096));
  HIP_TRY(checker_ring_set_max_lds((int)CHECKER_LDS_MAX));
  h->cring.assign(h->n_chains, 0);
  for (int c = 0; c < h->n_chains; ++c) h->cring[c] = checker_ring_chain(h, c) ? 1 : 0;
  HIP_TRY(hipMalloc(&h->d_mask_even, 2 * (size_t)h->n_chains));
  h->d_mask_ring = h->d_mask_even + h->n_chains;
  HIP_TRY(hipMalloc(&h->d_ring_list, sizeof(int32_t) * (size_t)h->n_chains));
  if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(MCG_ERR_HIP, "second stream for the ring-coloured chains");
  // the HBM path's kernels get their largest dynamic LDS once, here: an
  // 8-line ring of the longest band line (L0 <= 8192) and the largest line
  // tile; rebuild_work only chooses what a launch uses, so handles driven from
  // different host threads never change a kernel attribute under each other
  HIP_TRY(checker_band_set_max_lds(checker_band_ring_lines() * CHECKER_BAND_MAX_L0));
  HIP_TRY(checker_pair_set_max_lds(checker_tile_max_bytes()));
  HIP_TRY(checker_slab_set_max_lds(CHECKER_SLAB_MAX_LDS));
  HIP_TRY(hipMalloc(&h->d_cp, sizeof(CheckerParams) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_cs, sizeof(CheckerState) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_tab, sizeof(uint32_t) * tabs.size()));
  HIP_TRY(copy_sync(h, h->d_cp, h->cp.data(), sizeof(CheckerParams) * h->n_chains,
                    hipMemcpyHostToDevice));
  HIP_TRY(copy_sync(h, h->d_tab, tabs.data(), sizeof(uint32_t) * tabs.size(),
                    hipMemcpyHostToDevice));
  HIP_TRY(fill_sync(h, h->d_cs, 0, sizeof(CheckerState) * h->n_chains));
  return checker_update_masks(h);
}

// Parameters and acceptance tables (FCC_NTAB doubles at tab) of one FCC chain:
// A[s][n1][n2] = exp(-beta ds (V1 - mu + V2/6 n1 + V3/3 n2)) and
// B[s][n3][t] = exp(-beta ds (V4/12 n3 + Vt/8 t)), ds = 1 - 2s.
int fill_fcc_params(mcg_handle *h, int c, CheckerParams &P, double *tab) {
  auto const &ch = h->chains[c];
  for (int d = 0; d < 3; ++d)
    if (ch.shape[d] < 3 || ch.shape[d] > FCC_LMAX)
      return fail(MCG_ERR_ARG, "FCC edges must be 3..127 conventional cells");
  std::memset(&P, 0, sizeof(P));
  P.L0 = (int32_t)ch.shape[0];
  P.L1 = (int32_t)ch.shape[1];
  P.L2 = (int32_t)ch.shape[2];
  P.ncolor = 64;
  P.z1 = FCC_Z1;
  P.z2 = FCC_Z2;
  P.ndim = 3;
  P.key0 = (uint32_t)ch.seed;
  P.key1 = (uint32_t)(ch.seed >> 32);
  P.n = h->sites(c);
  P.lat_off = h->lat_off[c];
  P.tab_off = (int64_t)FCC_NTAB * c;
  const double beta = 1.0 / (h->kb * ch.temperature);
  const double mu = ch.exchange_potential;
  const double *v = h->eci;
  double worst = 0.0;
  for (int sg = 0; sg < 2; ++sg) {
    const double ds = 1.0 - 2.0 * sg;
    for (int n1 = 0; n1 <= FCC_Z1; ++n1)
      for (int n2 = 0; n2 <= FCC_Z2; ++n2) {
        const double x = -beta * ds * (v[1] - mu + v[2] / 6.0 * n1 + v[3] / 3.0 * n2);
        worst = std::max(worst, std::fabs(x));
        tab[sg * FCC_TAB_A + n1 * (FCC_Z2 + 1) + n2] = std::exp(x);
      }
    for (int n3 = 0; n3 <= FCC_Z3; ++n3)
      for (int t = 0; t <= FCC_ZT; ++t) {
        const double x = -beta * ds * (v[4] / 12.0 * n3 + v[5] / 8.0 * t);
        worst = std::max(worst, std::fabs(x));
        tab[2 * FCC_TAB_A + sg * FCC_TAB_B + n3 * (FCC_ZT + 1) + t] = std::exp(x);
      }
  }
  // each factor must be a finite, normal double for the product to decide
  if (!(worst < 700.0))
    return fail(MCG_ERR_ARG, "FCC: beta*|dE| term exceeds 700 (temperature too low for the ECIs)");
  return MCG_OK;
}

// LDS bytes of an FCC lattice: 4*L1*L2 lines of fcc_lds_stride(L0) bytes
int64_t fcc_lds_bytes(CheckerParams const &P) {
  return ((int64_t)4 * P.L1 * P.L2 * fcc_lds_stride(P.L0) + 15) & ~15LL;
}

int build_fcc(mcg_handle *h) {
  h->ncolor = 64;
  h->checker_block = 1024;
  h->cp.resize(h->n_chains);
  std::vector<double> tabs((size_t)FCC_NTAB * h->n_chains);
  int64_t max_lds = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    int rc = fill_fcc_params(h, c, h->cp[c], tabs.data() + (size_t)FCC_NTAB * c);
    if (rc != MCG_OK) return rc;
    max_lds = std::max(max_lds, fcc_lds_bytes(h->cp[c]));
  }
  h->checker_lds = max_lds <= FCC_LDS_MAX && !(h->options & MCG_OPT_FORCE_STREAMING);
  h->checker_lds_bytes = h->checker_lds ? (int)max_lds : 0;
  // the bit-line kernel when every line fits one word
  h->fcc_bits = !(h->options & MCG_OPT_FORCE_STREAMING);
  h->fcc_wide = false;
  h->fcc_bits_lds = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    auto const &P = h->cp[c];
    if (P.L0 > FCC_BITS_MAX_L0) h->fcc_bits = false;
    if (P.L0 > FCC_BITS_NARROW_L0) h->fcc_wide = true;
    h->fcc_bits_lds = std::max(h->fcc_bits_lds, fcc_bits_lds_bytes(P.L1, P.L2));
  }
  if (h->fcc_bits_lds > FCC_LDS_MAX) h->fcc_bits = false;
  if (!h->fcc_bits) h->fcc_wide = false;
  if (h->fcc_bits) h->checker_lds = true;
  if (h->fcc_wide) h->checker_block = 512;  // the wide kernel's launch bound
  HIP_TRY(fcc_set_max_lds((int)FCC_LDS_MAX));
  const std::vector<uint8_t> tri = fcc_tri_table();
  const std::vector<uint8_t> tri_lines = fcc_tri_table_lines();
  if (!fcc_tri_pairs_check(tri_lines))
    return fail(MCG_ERR_STATE, "FCC triangle pairs disagree with the triangle table");
  std::vector<FccSched> sched(h->n_chains);
  for (int c = 0; c < h->n_chains; ++c)
    fcc_class_schedule(h->cp[c].L0, h->cp[c].L1, h->cp[c].L2, sched[c]);
  HIP_TRY(hipMalloc(&h->d_fcc_sched, sizeof(FccSched) * h->n_chains));
  HIP_TRY(copy_sync(h, h->d_fcc_sched, sched.data(), sizeof(FccSched) * h->n_chains,
                    hipMemcpyHostToDevice));
  if (h->fcc_bits && !h->fcc_wide) {
    int most = 1;
    for (int c = 0; c < h->n_chains; ++c)
      most = std::max(most, fcc_max_step_lines(h->cp[c].L0, h->cp[c].L1, h->cp[c].L2, sched[c]));
    h->fcc_line_block = std::min(512, 64 * ((most + 63) / 64));
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    h->fcc_line_lds = h->fcc_bits_lds;
    if (h->n_chains <= cus) h->fcc_line_lds = std::max(h->fcc_line_lds, FCC_ONE_PER_CU_LDS);
  }
  HIP_TRY(hipMalloc(&h->d_tri_lines, tri_lines.size()));
  HIP_TRY(copy_sync(h, h->d_tri_lines, tri_lines.data(), tri_lines.size(),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&h->d_cp, sizeof(CheckerParams) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_cs, sizeof(CheckerState) * h->n_chains));
  HIP_TRY(hipMalloc(&h->d_ftab, sizeof(double) * tabs.size()));
  HIP_TRY(hipMalloc(&h->d_tri, tri.size()));
  HIP_TRY(copy_sync(h, h->d_cp, h->cp.data(), sizeof(CheckerParams) * h->n_chains,
                    hipMemcpyHostToDevice));
  HIP_TRY(copy_sync(h, h->d_ftab, tabs.data(), sizeof(double) * tabs.size(),
                    hipMemcpyHostToDevice));
  HIP_TRY(copy_sync(h, h->d_tri, tri.data(), tri.size(), hipMemcpyHostToDevice));
  HIP_TRY(fill_sync(h, h->d_cs, 0, sizeof(CheckerState) * h->n_chains));
  return MCG_OK;
}

// Recompute a chain's integer observables (and EXACT running energy) from its
// host occupation and store them in the device state.
int reset_observables(mcg_handle *h, int64_t c, std::vector<int8_t> const &occ) {
  auto const &ch = h->chains[c];
  if (h->fcc()) {
    CheckerState st;
    HIP_TRY(copy_sync(h, &st, h->d_cs + c, sizeof(CheckerState), hipMemcpyDeviceToHost));
    fcc_counts(occ, ch.shape, st.M, st.S1, st.S2, st.S3, st.T3);
    HIP_TRY(copy_sync(h, h->d_cs + c, &st, sizeof(CheckerState), hipMemcpyHostToDevice));
    return MCG_OK;
  }
  LatticeGeom g{h->lattice == MCG_LATTICE_SC ? 3 : 2, {ch.shape[0], ch.shape[1], ch.shape[2]}};
  int64_t S1, S2, M;
  if (h->occupant()) {
    // spins 2o - 1 of the occupant indices; the potential's per_supercell:
    // the formation energy, then minus each site's exchange potential in site
    // order; per unit cell (occupation_metropolis.hh:103-106)
    std::vector<int8_t> spins(occ.size());
    for (size_t l = 0; l < occ.size(); ++l) spins[l] = (int8_t)(2 * occ[l] - 1);
    integer_observables(spins, g, false, S1, S2, M);
    double E = ising_per_supercell(spins, ch.shape[0], ch.shape[1], h->J1, h->use_nlist);
    const double *mu = h->occ_mu.data() + (size_t)c * h->occ_n_species;
    for (size_t l = 0; l < occ.size(); ++l) E -= mu[h->occ_species[h->site_asym[c][l]][occ[l]]];
    ExactState st;
    HIP_TRY(copy_sync(h, &st, h->d_xs + c, sizeof(ExactState), hipMemcpyDeviceToHost));
    st.S = S1;
    st.M = M;
    st.e_run = E / static_cast<double>(h->sites(c) / h->occ_spu);
    HIP_TRY(copy_sync(h, h->d_xs + c, &st, sizeof(ExactState), hipMemcpyHostToDevice));
    return MCG_OK;
  }
  integer_observables(occ, g, h->J2 != 0.0, S1, S2, M);
  if (h->sequential()) {
    ExactState st;
    HIP_TRY(copy_sync(h, &st, h->d_xs + c, sizeof(ExactState), hipMemcpyDeviceToHost));
    st.S = S1;
    st.M = M;
    // occupation_metropolis.hh:103-106: potential per_supercell / n_unitcells
    // (the canonical potential is the formation energy alone)
    const int64_t n = g.n();
    const double Ef = ising_per_supercell(occ, ch.shape[0], ch.shape[1], h->J1, h->use_nlist);
    const double Nx = static_cast<double>(n + M) / 2.0;
    const double E = h->canonical() ? Ef : Ef - ch.exchange_potential * Nx;
    // NFOLD keeps S and M only (nfold.hh sets no state properties)
    st.e_run = h->nfold() ? 0.0 : E / (double)n;
    HIP_TRY(copy_sync(h, h->d_xs + c, &st, sizeof(ExactState), hipMemcpyHostToDevice));
  } else {
    CheckerState st;
    HIP_TRY(copy_sync(h, &st, h->d_cs + c, sizeof(CheckerState), hipMemcpyDeviceToHost));
    st.S1 = S1;
    st.S2 = S2;
    st.M = M;
    HIP_TRY(copy_sync(h, h->d_cs + c, &st, sizeof(CheckerState), hipMemcpyHostToDevice));
  }
  return MCG_OK;
}

// Copy a work list to its device buffer, growing the buffer only when the list
// outgrows it (a retired or re-admitted chain changes the lists on every
// active-mask change; the buffers keep their largest size so those changes
// cost no hipFree, which would synchronise the whole device).
int upload_list(mcg_handle *h, int32_t **buf, int64_t *cap, std::vector<int32_t> const &v) {
  if ((int64_t)v.size() > *cap) {
    int32_t *d = nullptr;
    const int64_t want = std::max<int64_t>((int64_t)v.size(), 2 * *cap);
    if (hipMalloc(&d, sizeof(int32_t) * want) != hipSuccess)
      return fail(MCG_ERR_CAPACITY, "out of device memory for work lists");
    if (*buf) (void)hipFree(*buf);
    *buf = d;
    *cap = want;
  }
  HIP_TRY(copy_sync(h, *buf, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice));
  return MCG_OK;
}

int rebuild_work(mcg_handle *h) {
  const int upi = checker_units_per_item();
  std::vector<int32_t> work;
  // ring-coloured chains have their own class launches (launch_checker_ring_pass)
  auto parity_chain = [&](int c) { return h->active[c] && !h->cring[c]; };
  for (int c = 0; c < h->n_chains; ++c) {
    if (!parity_chain(c)) continue;
    auto const &P = h->cp[c];
    // work units of one colour: 16-site chunks when the half line is a
    // multiple of 16, else 4-site quads (checker.hip)
    int64_t half0 = P.L0 / 2, qpl = (half0 + 3) / 4;
    int64_t per_line = (half0 % 16 == 0 && !P.quad_only) ? half0 / 16 : qpl;
    int64_t nl = P.ncolor == 2 ? (int64_t)P.L1 * P.L2
                               : (P.ncolor == 4 ? P.L1 / 2 : (int64_t)(P.L1 / 2) * (P.L2 / 2));
    int64_t nq = nl * per_line;
    int64_t items = (nq + upi - 1) / upi;
    for (int64_t it = 0; it < items; ++it) {
      work.push_back(c);
      work.push_back((int32_t)it);
    }
  }
  h->n_work = (int)(work.size() / 2);
  if (h->n_work) {
    int rc = upload_list(h, &h->d_work, &h->work_cap, work);
    if (rc != MCG_OK) return rc;
  }
  // 2D band path (checker.hip): every active chain square, 16-site chunks,
  // lines short enough for the 8-line ring
  h->n_bands = 0;
  h->band_ring_bytes = 0;
  bool band_ok = h->lattice == MCG_LATTICE_SQUARE && h->band_lines != 0 && !h->fcc() &&
                 !(h->options & MCG_OPT_CHECKER_QUAD_PATH);
  int64_t maxL0 = 0, total_lines = 0;
  for (int c = 0; c < h->n_chains && band_ok; ++c) {
    if (!parity_chain(c)) continue;
    auto const &P = h->cp[c];
    if (P.L0 % 32 != 0 || P.L0 > CHECKER_BAND_MAX_L0) band_ok = false;
    maxL0 = std::max<int64_t>(maxL0, P.L0);
    total_lines += P.L1;
  }
  // 3D J1+J2 slab path (checker_slab.hip): every active chain bit-packable;
  // band_lines counts planes per slab here
  h->band_slabs = false;
  bool slab_ok = h->lattice == MCG_LATTICE_SC && h->ncolor == 8 && h->band_lines != 0 &&
                 !(h->options & MCG_OPT_CHECKER_QUAD_PATH);
  int64_t total_planes = 0;
  int slab_bytes = 0;
  for (int c = 0; c < h->n_chains && slab_ok; ++c) {
    if (!parity_chain(c)) continue;
    auto const &P = h->cp[c];
    if (!checker_slab_fits(P)) slab_ok = false;
    total_planes += P.L2;
    slab_bytes = std::max(slab_bytes, checker_slab_ring_bytes(P.L0, P.L1));
  }
  if (slab_ok && total_planes > 0) {
    int target = h->band_lines;
    if (target < 0) {  // about two slabs per CU or more, slabs of at most 16 planes
      target = 16;
      while (target > 2 && total_planes / target < 2 * h->n_cu) target /= 2;
    }
    std::vector<int32_t> slabs;
    for (int c = 0; c < h->n_chains; ++c) {
      if (!parity_chain(c)) continue;
      const int L2 = h->cp[c].L2;
      int H = std::max(2, target & ~1);
      while (H > 2 && L2 % H != 0) H -= 2;  // an even divisor of L2 (L2 is even)
      for (int k0 = 0; k0 < L2; k0 += H) {
        slabs.push_back(c);
        slabs.push_back(k0);
        slabs.push_back(H);
      }
    }
    h->n_bands = (int)(slabs.size() / 3);
    h->band_ring_bytes = slab_bytes;
    h->band_slabs = true;
    int rc = upload_list(h, &h->d_band, &h->band_cap, slabs);
    if (rc != MCG_OK) return rc;
  }
  if (band_ok && total_lines > 0) {
    int target = h->band_lines;
    if (target < 0) {  // about 2048 bands or more, bands of at most 64 lines
      target = 64;
      while (target > 8 && total_lines / target < 2048) target /= 2;
    }
    std::vector<int32_t> band;
    for (int c = 0; c < h->n_chains; ++c) {
      if (!parity_chain(c)) continue;
      const int L1 = h->cp[c].L1;
      int H = std::max(2, target & ~1);
      while (H > 2 && L1 % H != 0) H -= 2;  // an even divisor of L1 (L1 is even)
      for (int j0 = 0; j0 < L1; j0 += H) {
        band.push_back(c);
        band.push_back(j0);
        band.push_back(H);
      }
    }
    h->n_bands = (int)(band.size() / 3);
    h->band_ring_bytes = (int)(checker_band_ring_lines() * maxL0);
    int rc = upload_list(h, &h->d_band, &h->band_cap, band);
    if (rc != MCG_OK) return rc;
  }
  // line-class path: whole colour-class lines per workgroup (used when the
  // band path is not); LDS tiles of class lines with their neighbour lines on
  // the 16-site chunk path, else lines straight from HBM
  h->n_pairs = 0;
  h->pair_lds = 0;
  if (h->n_bands == 0 && (h->ncolor == 4 || h->ncolor == 8) && !h->fcc()) {
    const int blk = checker_pair_block();
    // LDS per tile: units_per_thread 1, 2, 4 -> 24, 36, 48 KB
    const int64_t budget = std::min<int64_t>(
        checker_tile_max_bytes(),
        1024 * (h->units_per_thread == 1 ? 24 : (h->units_per_thread == 2 ? 36 : 48)));
    // tiles when requested and every active chain runs the 16-site chunk path
    bool tiles = h->line_tiles != 0;
    for (int c = 0; c < h->n_chains; ++c)
      if (parity_chain(c) && (h->cp[c].L0 / 2 % 16 != 0 || h->cp[c].quad_only)) tiles = false;
    std::vector<int32_t> items;
    for (int c = 0; c < h->n_chains; ++c) {
      if (!parity_chain(c)) continue;
      auto const &P = h->cp[c];
      const int64_t half = P.L0 / 2;
      const bool chunked = half % 16 == 0 && !P.quad_only;
      const int64_t upl = chunked ? half / 16 : (half + 3) / 4;
      const bool d3 = P.ndim == 3;
      const int64_t h1 = P.L1 / 2, h2 = d3 ? P.L2 / 2 : 1;
      const int64_t nl = P.ncolor == 4 ? P.L1 / 2 : (int64_t)(P.L1 / 2) * (P.L2 / 2);
      // tile TJ x TK with (2TJ+1)(2TK+1) lines of L0 bytes within the budget,
      // the most class lines per staged line
      int64_t TJ = 0, TK = 0;
      if (tiles) {
        const int64_t lines = budget / P.L0;
        double best = 0.0;
        for (int64_t tk = 1; tk <= (d3 ? h2 : 1); ++tk)
          for (int64_t tj = 1; tj <= h1; ++tj) {
            const int64_t staged = (2 * tj + 1) * (d3 ? 2 * tk + 1 : 1);
            if (staged > lines || tj > 0xFFFF || tk > 0x7FFF) break;
            const double score = (double)(tj * tk) / staged + 1e-9 * (double)(tj * tk);
            if (score > best) {
              best = score;
              TJ = tj;
              TK = tk;
            }
          }
      }
      if (TJ > 0) {
        h->pair_lds = std::max<int>(
            h->pair_lds, (int)((2 * TJ + 1) * (d3 ? 2 * TK + 1 : 1) * P.L0));
        for (int64_t kh = 0; kh < h2; kh += TK)
          for (int64_t jh = 0; jh < h1; jh += TJ) {
            const int64_t tj = std::min(TJ, h1 - jh), tk = std::min(TK, h2 - kh);
            items.push_back(c);
            items.push_back((int32_t)jh);
            items.push_back((int32_t)kh);
            items.push_back((int32_t)(tj | (tk << 16)));
          }
        continue;
      }
      // a workgroup takes units_per_thread units of each colour per thread
      const int64_t lpb = h->units_per_thread * std::max<int64_t>(1, blk / upl);
      for (int64_t l0 = 0; l0 < nl; l0 += lpb) {
        items.push_back(c);
        items.push_back((int32_t)l0);
        items.push_back((int32_t)std::min(lpb, nl - l0));
        items.push_back(0);
      }
    }
    h->n_pairs = (int)(items.size() / 4);
    if (h->pair_lds > 0) {
      const int per_cu = std::max(1, std::min(8, (int)(LDS_LIMIT / (h->pair_lds + 1024))));
      h->tile_grid = std::min(h->n_pairs, h->n_cu * per_cu);
    }
    if (h->n_pairs) {
      int rc = upload_list(h, &h->d_pair, &h->pair_cap, items);
      if (rc != MCG_OK) return rc;
    }
  }
  h->work_dirty = false;
  return MCG_OK;
}

// One pass of the HBM path: one launch per colour over the work list, then
// the pass-counter/sample-record launch.
void enqueue_pass(mcg_handle *h) {
  if (h->n_bands > 0 && h->band_slabs)
    launch_checker_slab(h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_band, h->n_bands,
                        h->band_ring_bytes, h->stream);
  else if (h->n_bands > 0)
    launch_checker_band(h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_band, h->n_bands,
                        h->band_ring_bytes, h->stream);
  else if (h->n_pairs > 0)
    for (int lc = 0; lc < h->ncolor / 2; ++lc)
      launch_checker_pair(h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_pair, h->n_pairs, lc,
                          h->pair_lds, h->tile_grid, h->stream);
  else if (h->n_work > 0)
    for (int color = 0; color < h->ncolor; ++color)
      launch_checker_color(h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_work, h->n_work, color,
                           h->units_per_thread, h->stream);
  if (h->any_ring())
    launch_checker_ring_pass(h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_ring_list,
                             h->n_ring_active, h->ring_per_chain,
                             h->lattice == MCG_LATTICE_SC ? 3 : 2, h->stream);
  launch_checker_record(h->d_cp, h->d_cs, h->d_series, h->series_cap, h->d_active, h->n_chains,
                        h->sample_period, h->stream);
}

// Build a hipGraph of graph_passes passes of the HBM path (kernel nodes added
// explicitly; every launch reads its pass counter from device state, so
// replays are independent).
int build_graph(mcg_handle *h) {
  if (h->graph_exec) {
    HIP_TRY(hipGraphExecDestroy(h->graph_exec));
    h->graph_exec = nullptr;
  }
  hipGraph_t g = nullptr;
  HIP_TRY(hipGraphCreate(&g, 0));
  hipError_t e = checker_graph_add_passes(
      g, h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_work, h->n_work, h->ncolor,
      h->units_per_thread, h->d_series, h->series_cap, h->d_active, h->n_chains,
      h->sample_period, h->graph_passes, h->d_band, h->n_bands, h->band_ring_bytes,
      h->band_slabs, h->d_pair,
      h->n_pairs, h->pair_lds, h->tile_grid);
  if (e == hipSuccess) e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    h->graph_exec = nullptr;
    return fail(MCG_ERR_HIP, std::string("graph build: ") + hipGetErrorString(e));
  }
  h->graph_dirty = false;
  return MCG_OK;
}

// Autotuner cache key: path (LDS-resident or HBM), lattice, colouring and
// the size class (log2 of the sites per chain and of the chain count).
std::string tune_key(const mcg_handle *h) {
  int64_t sites = 0;
  for (int c = 0; c < h->n_chains; ++c) sites = std::max(sites, h->sites(c));
  auto lg = [](int64_t v) {
    int k = 0;
    while ((int64_t(1) << k) < v) ++k;
    return k;
  };
  return std::string(h->checker_lds ? "lds" : "hbm") + ":" + std::to_string(h->lattice) + ":" +
         std::to_string(h->ncolor) + ":" + std::to_string(lg(sites)) + ":" +
         std::to_string(lg(h->n_chains));
}

std::mutex &tune_mutex() {
  static std::mutex m;
  return m;
}

std::map<std::string, mcg_schedule> &tune_cache() {
  static std::map<std::string, mcg_schedule> m;
  return m;
}

// File format, one schedule per line:
//   key lds_block units_per_thread graph_passes trial_ms band_lines line_tiles
bool tune_cache_lookup(const std::string &key, mcg_schedule *out) {
  std::lock_guard<std::mutex> lock(tune_mutex());
  auto &m = tune_cache();
  auto it = m.find(key);
  if (it != m.end()) {
    *out = it->second;
    return true;
  }
  const char *path = std::getenv("MCG_TUNE_CACHE");
  if (!path || !*path) return false;
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  char k[128];
  long long b, u, g, bl, lt;
  double ms;
  bool found = false;
  while (std::fscanf(f, "%127s %lld %lld %lld %lf %lld %lld", k, &b, &u, &g, &ms, &bl, &lt) ==
         7) {
    if (key == k) {  // the last line for a key wins
      *out = mcg_schedule{b, u, g, ms, bl, lt};
      found = true;
    }
  }
  std::fclose(f);
  if (found) m[key] = *out;
  return found;
}

void tune_cache_store(const std::string &key, const mcg_schedule &s) {
  std::lock_guard<std::mutex> lock(tune_mutex());
  tune_cache()[key] = s;
  const char *path = std::getenv("MCG_TUNE_CACHE");
  if (!path || !*path) return;
  if (FILE *f = std::fopen(path, "a")) {
    std::fprintf(f, "%s %lld %lld %lld %.6f %lld %lld\n", key.c_str(), (long long)s.lds_block,
                 (long long)s.units_per_thread, (long long)s.graph_passes, s.trial_ms,
                 (long long)s.band_lines, (long long)s.line_tiles);
    std::fclose(f);
  }
}

bool valid_chain(mcg_handle *h, int64_t c) { return h && c >= 0 && c < h->n_chains; }

// EXACT launches: the chains whose lattice fits in LDS run the block walk on the
// handle's stream; the others run the scalar walk on their HBM lattices on the
// second stream at the same time (the two sets touch disjoint chains), and
// the handle's stream waits for them.
int exact_launch(mcg_handle *h, int64_t n_passes, int64_t n_steps) {
  int lds = 0, scratch = 0;
  bool any_fit = false, any_hbm = false;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    if (h->xfit[c]) {
      any_fit = true;
      lds = std::max(lds, h->xlds[c]);
    } else {
      any_hbm = true;
      scratch = std::max(scratch, 4 * (h->xp[c].rows + h->xp[c].cols));
    }
  }
  if (any_hbm) {
    HIP_TRY(hipEventRecord(h->ev_fork, h->stream));
    HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
    launch_exact(h->d_xp, h->d_xs, h->d_lat, h->d_series, h->series_cap, h->d_mask_hbm,
                 h->n_chains, n_passes, n_steps, h->sample_period, scratch, false, false,
                 h->stream2);
    HIP_TRY(hipGetLastError());
  }
  if (any_fit) {
    launch_exact(h->d_xp, h->d_xs, h->d_lat, h->d_series, h->series_cap, h->d_mask_lds,
                 h->n_chains, n_passes, n_steps, h->sample_period, lds, true,
                 !(h->options & MCG_OPT_EXACT_SCALAR_WALK), h->stream);
    HIP_TRY(hipGetLastError());
  }
  if (any_hbm) {
    HIP_TRY(hipEventRecord(h->ev_join, h->stream2));
    HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
  }
  return MCG_OK;
}

}  // namespace

extern "C" {

const char *mcg_last_error(void) { return g_err.c_str(); }

int mcg_create(const mcg_desc *desc, mcg_handle **out) {
  if (!desc || !out) return fail(MCG_ERR_ARG, "null argument");
  *out = nullptr;
  if (desc->mode != MCG_MODE_EXACT && desc->mode != MCG_MODE_CHECKERBOARD &&
      desc->mode != MCG_MODE_CANONICAL && desc->mode != MCG_MODE_NFOLD &&
      desc->mode != MCG_MODE_OCCUPANT && desc->mode != MCG_MODE_KMC)
    return fail(MCG_ERR_ARG, "unknown mode");
  if (desc->mode == MCG_MODE_KMC && (desc->lattice != MCG_LATTICE_SQUARE || desc->J2 != 0.0))
    return fail(MCG_ERR_ARG, "KMC mode supports the 2D square nearest-neighbour lattice gas only");
  if (desc->mode == MCG_MODE_OCCUPANT &&
      (desc->lattice != MCG_LATTICE_SQUARE || desc->J2 != 0.0))
    return fail(MCG_ERR_ARG, "OCCUPANT mode supports the 2D square nearest-neighbour model only");
  if (desc->mode == MCG_MODE_NFOLD && (desc->lattice != MCG_LATTICE_SQUARE || desc->J2 != 0.0))
    return fail(MCG_ERR_ARG, "NFOLD mode supports the 2D square nearest-neighbour model only");
  if (desc->sample_time_period < 0.0 ||
      (desc->sample_time_period > 0.0 && desc->mode != MCG_MODE_NFOLD &&
       desc->mode != MCG_MODE_KMC))
    return fail(MCG_ERR_ARG, "sample_time_period must be 0, or > 0 in NFOLD and KMC modes");
  if (desc->lattice != MCG_LATTICE_SQUARE && desc->lattice != MCG_LATTICE_SC &&
      desc->lattice != MCG_LATTICE_FCC)
    return fail(MCG_ERR_ARG, "unknown lattice");
  if (desc->lattice == MCG_LATTICE_FCC &&
      (desc->mode != MCG_MODE_CHECKERBOARD || desc->eci == nullptr))
    return fail(MCG_ERR_ARG, "the FCC cluster expansion runs in CHECKERBOARD mode with 6 ECIs");
  if (desc->mode == MCG_MODE_EXACT && (desc->lattice != MCG_LATTICE_SQUARE || desc->J2 != 0.0))
    return fail(MCG_ERR_ARG, "EXACT mode supports the 2D square nearest-neighbour model only");
  if (desc->mode == MCG_MODE_CANONICAL &&
      (desc->lattice != MCG_LATTICE_SQUARE || desc->J2 != 0.0))
    return fail(MCG_ERR_ARG, "CANONICAL mode supports the 2D square nearest-neighbour model only");
  if (desc->n_chains <= 0 || desc->n_chains > (1 << 24) || !desc->chains)
    return fail(MCG_ERR_ARG, "n_chains out of range");
  if (desc->sample_period < 1) return fail(MCG_ERR_ARG, "sample_period < 1");
  if (!(desc->kb > 0.0)) return fail(MCG_ERR_ARG, "kb must be > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(MCG_ERR_HIP, "no HIP device available (the library has no CPU fallback)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
    return fail(MCG_ERR_HIP, std::string("device is not gfx950: ") + prop.gcnArchName);

  // the no-allocation statistics scratch: its device address, once per device
  HIP_TRY(stats_noalloc_scratch_init(dev));

  auto h = new mcg_handle();
  h->device = dev;
  h->n_cu = std::max(1, prop.multiProcessorCount);
  h->mode = desc->mode;
  h->lattice = desc->lattice;
  h->use_nlist = desc->use_nlist != 0;
  h->J1 = desc->J1;
  h->J2 = desc->J2;
  h->kb = desc->kb;
  h->sample_period = desc->sample_period;
  h->options = desc->options;
  h->sample_time_period = desc->sample_time_period;
  if (h->fcc())
    for (int i = 0; i < 6; ++i) h->eci[i] = desc->eci[i];
  if (h->occupant()) {
    const int orc = occupant_configure(h, desc);
    if (orc != MCG_OK) {
      delete h;
      return orc;
    }
  }
  h->n_chains = (int)desc->n_chains;
  h->chains.assign(desc->chains, desc->chains + desc->n_chains);
  for (auto &ch : h->chains) {
    if (h->lattice == MCG_LATTICE_SQUARE) ch.shape[2] = 1;
    for (int d = 0; d < 3; ++d)
      if (ch.shape[d] < 1 || ch.shape[d] > (1 << 20)) {
        delete h;
        return fail(MCG_ERR_ARG, "shape out of range");
      }
    if (!(ch.temperature > 0.0)) {
      delete h;
      return fail(MCG_ERR_ARG, "temperature must be > 0");
    }
  }
  int64_t off = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    int64_t n = h->sites(c);
    // EXACT: the walk divides l < 2^24 by rows with a 40-bit reciprocal, which
    // is exact for rows < 2^16
    if (h->mode == MCG_MODE_EXACT &&
        (n > (1LL << 24) || h->chains[c].shape[0] >= (1LL << 16))) {
      delete h;
      return fail(MCG_ERR_ARG, "EXACT mode supports at most 2^24 sites and rows < 65536");
    }
    if ((h->canonical() && n > (1LL << 16)) || (h->nfold() && n > (1LL << 15))) {
      delete h;
      return fail(MCG_ERR_ARG, "CANONICAL mode supports at most 65536 sites per chain, "
                               "NFOLD at most 32768");
    }
    if (h->kmc() && (n > KMC_MAX_SITES || h->chains[c].shape[0] < 3 || h->chains[c].shape[1] < 3)) {
      delete h;
      return fail(MCG_ERR_ARG, "KMC mode supports 3x3 to 7500 sites per chain");
    }
    if (h->occupant() && (n > OCC_MAX_SITES || n % h->occ_spu != 0)) {
      delete h;
      return fail(MCG_ERR_ARG, "OCCUPANT mode supports at most 40000 sites per chain, a "
                               "multiple of sites_per_unitcell");
    }
    if (n > (1LL << 31)) {
      delete h;
      return fail(MCG_ERR_ARG, "too many sites in one chain");
    }
    h->lat_off.push_back(off);
    off += (n + 255) & ~255LL;  // 256-byte aligned lattices
  }
  h->lat_bytes = off;
  int rc = MCG_OK;
  auto bail = [&](int code) {
    delete h;
    return code;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(MCG_ERR_HIP, "hipStreamCreate failed"));
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return bail(fail(MCG_ERR_HIP, "hipEventCreate failed"));
  if (hipMalloc(&h->d_lat, (size_t)h->lat_bytes) != hipSuccess)
    return bail(fail(MCG_ERR_CAPACITY, "out of device memory for lattices"));
  // Ising lattices start all +1 (stored as 1), FCC lattices empty, OCCUPANT
  // lattices at occupant 0
  const int fill0 = (h->fcc() || h->occupant() || h->kmc()) ? 0 : 1;
  if (fill_sync(h, h->d_lat, fill0, (size_t)h->lat_bytes) != hipSuccess)
    return bail(fail(MCG_ERR_HIP, "hipMemset failed"));
  h->active.assign(h->n_chains, 1);
  if (hipMalloc(&h->d_active, h->n_chains) != hipSuccess ||
      copy_sync(h, h->d_active, h->active.data(), h->n_chains, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(MCG_ERR_HIP, "active mask allocation failed"));
  h->n_samples.assign(h->n_chains, 0);
  h->n_pass.assign(h->n_chains, 0);
  h->n_step.assign(h->n_chains, 0);
  for (int c = 0; c < h->n_chains; ++c) h->capacity.push_back(h->sites(c));
  if (hipMalloc(&h->d_eq, 2 * sizeof(int64_t)) != hipSuccess)
    return bail(fail(MCG_ERR_HIP, "allocation failed"));
  rc = (h->mode == MCG_MODE_EXACT)  ? build_exact(h)
       : h->canonical()              ? build_canonical(h)
       : h->nfold()                  ? build_nfold(h)
       : h->occupant()               ? build_occupant(h)
       : h->kmc()                    ? build_kmc(h)
       : h->fcc()                    ? build_fcc(h)
                                     : build_checker(h);
  if (rc != MCG_OK) return bail(rc);
  // initial lattices: set observables
  for (int c = 0; c < h->n_chains; ++c) {
    std::vector<int8_t> occ(h->sites(c), (int8_t)fill0);
    rc = reset_observables(h, c, occ);
    if (rc == MCG_OK && h->nfold()) rc = nfold_init_array(h, c, occ);
    if (rc == MCG_OK && h->occupant()) rc = occupant_init_lists(h, c, occ);
    if (rc == MCG_OK && h->kmc()) rc = kmc_init(h, c, occ);
    if (rc != MCG_OK) return bail(rc);
  }
  rc = ensure_series(h, 64);
  if (rc != MCG_OK) return bail(rc);
  *out = h;
  return MCG_OK;
}

void mcg_destroy(mcg_handle *h) { delete h; }

int mcg_set_occupation(mcg_handle *h, int64_t chain, const int8_t *occ, int64_t n) {
  if (!valid_chain(h, chain) || !occ) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (n != h->sites(chain)) return fail(MCG_ERR_ARG, "occupation size mismatch");
  std::vector<int8_t> v(occ, occ + n);
  if (h->occupant()) {
    for (int64_t l = 0; l < n; ++l)
      if (v[l] < 0 || v[l] >= h->occ_size[h->site_asym[chain][l]])
        return fail(MCG_ERR_ARG, "OCCUPANT occupation values must be occupant indices of the "
                                 "site's asymmetric unit");
  } else if (h->fcc() || h->kmc()) {
    for (int8_t s : v)
      if (s != 0 && s != 1)
        return fail(MCG_ERR_ARG, "FCC and KMC occupation values must be 0 or 1");
  } else {
    for (int8_t s : v)
      if (s != 1 && s != -1) return fail(MCG_ERR_ARG, "occupation values must be +1 or -1");
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sequential() || h->fcc()) {
    HIP_TRY(copy_sync(h, h->d_lat + h->lat_off[chain], v.data(), n, hipMemcpyHostToDevice));
    if (h->canonical()) {
      const int rc = canonical_init_lists(h, chain, v);
      if (rc != MCG_OK) return rc;
    }
    if (h->nfold()) {
      const int rc = nfold_init_array(h, chain, v);
      if (rc != MCG_OK) return rc;
    }
    if (h->occupant()) {
      const int rc = occupant_init_lists(h, chain, v);
      if (rc != MCG_OK) return rc;
    }
    if (h->kmc()) return kmc_init(h, chain, v);
  } else {
    std::vector<int8_t> dev(n);
    if (h->cring[chain])  // natural layout, 1 for +1
      for (int64_t l = 0; l < n; ++l) dev[l] = v[l] > 0 ? 1 : 0;
    else
      to_split_rows(v, h->chains[chain].shape[0], dev);
    HIP_TRY(copy_sync(h, h->d_lat + h->lat_off[chain], dev.data(), n, hipMemcpyHostToDevice));
  }
  return reset_observables(h, chain, v);
}

int mcg_get_occupation(mcg_handle *h, int64_t chain, int8_t *occ, int64_t n) {
  if (!valid_chain(h, chain) || !occ) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (n != h->sites(chain)) return fail(MCG_ERR_ARG, "occupation size mismatch");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sequential() || h->fcc()) {
    HIP_TRY(copy_sync(h, occ, h->d_lat + h->lat_off[chain], n, hipMemcpyDeviceToHost));
  } else {
    std::vector<int8_t> dev(n), nat(n);
    HIP_TRY(copy_sync(h, dev.data(), h->d_lat + h->lat_off[chain], n, hipMemcpyDeviceToHost));
    if (h->cring[chain])
      for (int64_t l = 0; l < n; ++l) nat[l] = dev[l] ? 1 : -1;
    else
      from_split_rows(dev, h->chains[chain].shape[0], nat);
    std::copy(nat.begin(), nat.end(), occ);
  }
  return MCG_OK;
}

int mcg_rng_dump(mcg_handle *h, int64_t chain, uint64_t *words, int64_t n_words) {
  if (!valid_chain(h, chain) || !words) return fail(MCG_ERR_ARG, "bad chain or buffer");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sequential()) {
    if (n_words < MT_N + 1) return fail(MCG_ERR_CAPACITY, "need 313 words");
    ExactState st;
    HIP_TRY(copy_sync(h, &st, h->d_xs + chain, sizeof(st), hipMemcpyDeviceToHost));
    std::memcpy(words, st.mt, sizeof(uint64_t) * MT_N);
    words[MT_N] = (uint64_t)st.pos;
  } else {
    if (n_words < 4) return fail(MCG_ERR_CAPACITY, "need 4 words");
    CheckerState st;
    HIP_TRY(copy_sync(h, &st, h->d_cs + chain, sizeof(st), hipMemcpyDeviceToHost));
    words[0] = h->cp[chain].key0;
    words[1] = h->cp[chain].key1;
    words[2] = (uint64_t)st.n_pass;
    words[3] = 0;
  }
  return MCG_OK;
}

int mcg_rng_load(mcg_handle *h, int64_t chain, const uint64_t *words, int64_t n_words) {
  if (!valid_chain(h, chain) || !words) return fail(MCG_ERR_ARG, "bad chain or buffer");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sequential()) {
    if (n_words < MT_N + 1) return fail(MCG_ERR_CAPACITY, "need 313 words");
    if (words[MT_N] > (uint64_t)MT_N) return fail(MCG_ERR_ARG, "engine position out of range");
    ExactState st;
    HIP_TRY(copy_sync(h, &st, h->d_xs + chain, sizeof(st), hipMemcpyDeviceToHost));
    std::memcpy(st.mt, words, sizeof(uint64_t) * MT_N);
    st.pos = (int64_t)words[MT_N];
    HIP_TRY(copy_sync(h, h->d_xs + chain, &st, sizeof(st), hipMemcpyHostToDevice));
  } else {
    if (n_words < 4) return fail(MCG_ERR_CAPACITY, "need 4 words");
    h->cp[chain].key0 = (uint32_t)words[0];
    h->cp[chain].key1 = (uint32_t)words[1];
    HIP_TRY(copy_sync(h, h->d_cp + chain, &h->cp[chain], sizeof(CheckerParams),
                      hipMemcpyHostToDevice));
    CheckerState st;
    HIP_TRY(copy_sync(h, &st, h->d_cs + chain, sizeof(st), hipMemcpyDeviceToHost));
    st.n_pass = (int64_t)words[2];
    HIP_TRY(copy_sync(h, h->d_cs + chain, &st, sizeof(st), hipMemcpyHostToDevice));
    h->n_pass[chain] = st.n_pass;
  }
  return MCG_OK;
}

int mcg_assign_chain(mcg_handle *h, int64_t chain, const mcg_chain_desc *desc) {
  if (!valid_chain(h, chain) || !desc) return fail(MCG_ERR_ARG, "bad chain or descriptor");
  mcg_chain_desc nd = *desc;
  if (h->lattice == MCG_LATTICE_SQUARE) nd.shape[2] = 1;
  for (int d = 0; d < 3; ++d)
    if (nd.shape[d] < 1 || nd.shape[d] > (1 << 20)) return fail(MCG_ERR_ARG, "shape out of range");
  if (!(nd.temperature > 0.0)) return fail(MCG_ERR_ARG, "temperature must be > 0");
  const int64_t n = (h->fcc() ? 4 : 1) * nd.shape[0] * nd.shape[1] * nd.shape[2];
  if (n > h->capacity[chain])
    return fail(MCG_ERR_CAPACITY, "shape has more sites than the chain slot holds");
  if (h->mode == MCG_MODE_EXACT && nd.shape[0] >= (1LL << 16))
    return fail(MCG_ERR_ARG, "EXACT mode supports rows < 65536");
  if ((h->canonical() && n > (1LL << 16)) || (h->nfold() && n > (1LL << 15)))
    return fail(MCG_ERR_ARG, "CANONICAL mode supports at most 65536 sites per chain, "
                             "NFOLD at most 32768");
  if (h->occupant() && n % h->occ_spu != 0)
    return fail(MCG_ERR_ARG, "OCCUPANT: sites must be a multiple of sites_per_unitcell");
  if (h->kmc() && (nd.shape[0] < 3 || nd.shape[1] < 3))
    return fail(MCG_ERR_ARG, "KMC mode supports lattices of at least 3x3 sites");

  HIP_TRY(hipStreamSynchronize(h->stream));
  const mcg_chain_desc old = h->chains[chain];
  h->chains[chain] = nd;
  if (h->mode == MCG_MODE_EXACT) {
    ExactParams P;
    const int lds = fill_exact_params(h, (int)chain, P);
    const ExactParams P_old = h->xp[chain];
    h->xp[chain] = P;
    const int prc = exact_choose_path(h, (int)chain, lds);
    if (prc != MCG_OK) {
      h->chains[chain] = old;
      h->xp[chain] = P_old;
      exact_choose_path(h, (int)chain, fill_exact_params(h, (int)chain, h->xp[chain]));
      return prc;
    }
    HIP_TRY(copy_sync(h, h->d_xp + chain, &P, sizeof(P), hipMemcpyHostToDevice));
    {
      const int mrc = exact_update_masks(h);
      if (mrc != MCG_OK) return mrc;
    }
    ExactState S;
    std::memset(&S, 0, sizeof(S));
    int pos = 0;
    mt_seed(nd.seed, S.mt, &pos);
    S.pos = pos;
    HIP_TRY(copy_sync(h, h->d_xs + chain, &S, sizeof(S), hipMemcpyHostToDevice));
  } else if (h->nfold()) {
    NfoldParams P;
    fill_nfold_params(h, (int)chain, P);
    if (nfold_lds_bytes(P) > h->nfold_lds) {
      h->chains[chain] = old;
      return fail(MCG_ERR_CAPACITY, "shape needs more LDS than the handle was created with");
    }
    h->np_[chain] = P;
    HIP_TRY(copy_sync(h, h->d_np + chain, &P, sizeof(P), hipMemcpyHostToDevice));
    ExactState S;
    nfold_reset_state(S, nd.seed);
    HIP_TRY(copy_sync(h, h->d_xs + chain, &S, sizeof(S), hipMemcpyHostToDevice));
  } else if (h->kmc()) {
    ExactState S;
    std::memset(&S, 0, sizeof(S));
    int pos = 0;
    mt_seed(nd.seed, S.mt, &pos);
    S.pos = pos;
    HIP_TRY(copy_sync(h, h->d_xs + chain, &S, sizeof(S), hipMemcpyHostToDevice));
  } else if (h->occupant()) {
    const std::vector<int8_t> old_units = h->site_asym[chain];
    h->site_asym[chain].assign(n, 0);
    const int orc = occupant_setup_chain(h, (int)chain, true);
    if (orc != MCG_OK) {
      h->chains[chain] = old;
      h->site_asym[chain] = old_units;
      return orc;
    }
  } else if (h->canonical()) {
    CanonParams P;
    fill_canonical_params(h, (int)chain, 0, P);
    if (canon_lds_bytes(P) > h->canon_lds) {
      h->chains[chain] = old;
      return fail(MCG_ERR_CAPACITY, "shape needs more LDS than the handle was created with");
    }
    h->kp[chain] = P;
    HIP_TRY(copy_sync(h, h->d_kp + chain, &h->kp[chain], sizeof(CanonParams),
                      hipMemcpyHostToDevice));
    ExactState S;
    std::memset(&S, 0, sizeof(S));
    int pos = 0;
    mt_seed(nd.seed, S.mt, &pos);
    S.pos = pos;
    HIP_TRY(copy_sync(h, h->d_xs + chain, &S, sizeof(S), hipMemcpyHostToDevice));
  } else if (h->fcc()) {
    CheckerParams P;
    std::vector<double> tab(FCC_NTAB);
    const int rc = fill_fcc_params(h, (int)chain, P, tab.data());
    if (rc != MCG_OK) {
      h->chains[chain] = old;
      return rc;
    }
    const bool no_room =
        h->fcc_bits ? (P.L0 > (h->fcc_wide ? FCC_BITS_MAX_L0 : FCC_BITS_NARROW_L0) ||
                       fcc_bits_lds_bytes(P.L1, P.L2) > h->fcc_bits_lds)
                    : (h->checker_lds && fcc_lds_bytes(P) > h->checker_lds_bytes);
    if (no_room) {
      h->chains[chain] = old;
      return fail(MCG_ERR_CAPACITY, "shape needs more LDS than the handle was created with");
    }
    h->cp[chain] = P;
    HIP_TRY(copy_sync(h, h->d_cp + chain, &P, sizeof(P), hipMemcpyHostToDevice));
    HIP_TRY(copy_sync(h, h->d_ftab + P.tab_off, tab.data(), sizeof(double) * tab.size(),
                      hipMemcpyHostToDevice));
    FccSched sched;
    fcc_class_schedule(P.L0, P.L1, P.L2, sched);
    HIP_TRY(copy_sync(h, h->d_fcc_sched + chain, &sched, sizeof(sched), hipMemcpyHostToDevice));
    CheckerState S;
    std::memset(&S, 0, sizeof(S));
    HIP_TRY(copy_sync(h, h->d_cs + chain, &S, sizeof(S), hipMemcpyHostToDevice));
  } else {
    CheckerParams P;
    std::vector<uint32_t> tab(checker_ntab(h));
    const int rc = fill_checker_params(h, (int)chain, P, tab.data());
    if (rc != MCG_OK) {
      h->chains[chain] = old;
      return rc;
    }
    h->cp[chain] = P;
    HIP_TRY(copy_sync(h, h->d_cp + chain, &P, sizeof(P), hipMemcpyHostToDevice));
    HIP_TRY(copy_sync(h, h->d_tab + P.tab_off, tab.data(), sizeof(uint32_t) * tab.size(),
                      hipMemcpyHostToDevice));
    CheckerState S;
    std::memset(&S, 0, sizeof(S));
    HIP_TRY(copy_sync(h, h->d_cs + chain, &S, sizeof(S), hipMemcpyHostToDevice));
    h->work_dirty = true;
    h->cring[chain] = checker_ring_chain(h, (int)chain) ? 1 : 0;
    const int mrc = checker_update_masks(h);
    if (mrc != MCG_OK) return mrc;
  }
  const int fill = (h->fcc() || h->occupant() || h->kmc()) ? 0 : 1;
  HIP_TRY(fill_sync(h, h->d_lat + h->lat_off[chain], fill,
                    (size_t)((h->capacity[chain] + 255) & ~255LL)));
  h->n_samples[chain] = 0;
  h->n_pass[chain] = 0;
  h->n_step[chain] = 0;
  const int rc = reset_observables(h, chain, std::vector<int8_t>(n, (int8_t)fill));
  if (rc == MCG_OK && h->occupant())
    return occupant_init_lists(h, chain, std::vector<int8_t>(n, (int8_t)fill));
  if (rc == MCG_OK && h->kmc()) return kmc_init(h, chain, std::vector<int8_t>(n, (int8_t)fill));
  if (rc != MCG_OK || !h->nfold()) return rc;
  return nfold_init_array(h, chain, std::vector<int8_t>(n, (int8_t)fill));
}

int mcg_get_occ_location(mcg_handle *h, int64_t chain, int64_t *loc, int64_t n) {
  if (!valid_chain(h, chain) || (!loc && n)) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (h->occupant()) {
    OccParams const &P = h->op[chain];
    if (n != P.n_mut) return fail(MCG_ERR_ARG, "list size mismatch (mutating sites)");
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<int32_t> seg(n);
    if (n)
      HIP_TRY(copy_sync(h, seg.data(), h->d_list + h->lat_off[chain], sizeof(int32_t) * n,
                        hipMemcpyDeviceToHost));
    ExactState st;
    HIP_TRY(copy_sync(h, &st, h->d_xs + chain, sizeof(st), hipMemcpyDeviceToHost));
    std::vector<std::vector<int32_t>> lists;
    occ_segments_to_lists(P, st.occ_cnt0, seg, lists);
    int64_t k = 0;
    for (auto const &v : lists)
      for (int32_t l : v) loc[k++] = l;
    return MCG_OK;
  }
  if (!loc) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (h->kmc()) {
    if (n != h->sites(chain)) return fail(MCG_ERR_ARG, "list size mismatch");
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<int32_t> seg(n);
    HIP_TRY(copy_sync(h, seg.data(), h->d_list + h->lat_off[chain], sizeof(int32_t) * n,
                      hipMemcpyDeviceToHost));
    const int64_t na = h->kp_kmc[chain].n_atoms;
    for (int64_t i = 0; i < n - na; ++i) loc[i] = seg[i];
    for (int64_t j = 0; j < na; ++j) loc[n - na + j] = seg[n - 1 - j];
    return MCG_OK;
  }
  if (!h->canonical() && !h->nfold())
    return fail(MCG_ERR_STATE, "site lists exist in CANONICAL and NFOLD modes only");
  if (n != h->sites(chain)) return fail(MCG_ERR_ARG, "list size mismatch");
  HIP_TRY(hipStreamSynchronize(h->stream));
  std::vector<int32_t> v(n);
  HIP_TRY(copy_sync(h, v.data(), h->d_list + h->lat_off[chain], sizeof(int32_t) * n,
                    hipMemcpyDeviceToHost));
  std::copy(v.begin(), v.end(), loc);
  return MCG_OK;
}

int mcg_set_occ_location(mcg_handle *h, int64_t chain, const int64_t *loc, int64_t n) {
  if (!valid_chain(h, chain) || (!loc && n)) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (h->occupant()) {
    OccParams const &P = h->op[chain];
    if (n != P.n_mut) return fail(MCG_ERR_ARG, "list size mismatch (mutating sites)");
    const int64_t ns = h->sites(chain);
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<int8_t> occ(ns);
    HIP_TRY(copy_sync(h, occ.data(), h->d_lat + h->lat_off[chain], ns, hipMemcpyDeviceToHost));
    // list sizes from the occupation, then every entry a mutating site of the
    // list's candidate, each once
    std::vector<int64_t> size(P.n_cand, 0);
    for (int64_t l = 0; l < ns; ++l) {
      const int ci = occ_cand_of(h, h->site_asym[chain][l], occ[l]);
      if (ci >= 0) ++size[ci];
    }
    std::vector<std::vector<int32_t>> lists(P.n_cand);
    std::vector<uint8_t> seen(ns, 0);
    int64_t k = 0;
    for (int ci = 0; ci < P.n_cand; ++ci)
      for (int64_t j = 0; j < size[ci]; ++j, ++k) {
        const int64_t l = loc[k];
        if (l < 0 || l >= ns || seen[l])
          return fail(MCG_ERR_ARG, "occupant lists must hold every mutating site exactly once");
        if (occ_cand_of(h, h->site_asym[chain][l], occ[l]) != ci)
          return fail(MCG_ERR_ARG, "occupant lists do not match the occupation");
        seen[l] = 1;
        lists[ci].push_back((int32_t)l);
      }
    return occupant_upload_lists(h, chain, lists);
  }
  if (!loc) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (!h->canonical()) return fail(MCG_ERR_STATE, "occupant lists exist in CANONICAL mode only");
  if (n != h->sites(chain)) return fail(MCG_ERR_ARG, "list size mismatch");
  HIP_TRY(hipStreamSynchronize(h->stream));
  std::vector<int8_t> occ(n);
  HIP_TRY(copy_sync(h, occ.data(), h->d_lat + h->lat_off[chain], n, hipMemcpyDeviceToHost));
  const int64_t nA = h->kp[chain].nA;
  std::vector<uint8_t> seen(n, 0);
  std::vector<int32_t> v(n);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t l = loc[i];
    if (l < 0 || l >= n || seen[l])
      return fail(MCG_ERR_ARG, "occupant lists must hold every site exactly once");
    if ((occ[l] < 0) != (i < nA))
      return fail(MCG_ERR_ARG, "occupant lists do not match the occupation");
    seen[l] = 1;
    v[i] = (int32_t)l;
  }
  HIP_TRY(copy_sync(h, h->d_list + h->lat_off[chain], v.data(), sizeof(int32_t) * n,
                    hipMemcpyHostToDevice));
  return MCG_OK;
}

int mcg_get_atoms(mcg_handle *h, int64_t chain, int64_t *atoms, int64_t n_atoms) {
  if (!valid_chain(h, chain) || (!atoms && n_atoms)) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (!h->kmc()) return fail(MCG_ERR_STATE, "atoms are tracked in KMC mode only");
  if (n_atoms != h->kp_kmc[chain].n_atoms) return fail(MCG_ERR_ARG, "atom count mismatch");
  HIP_TRY(hipStreamSynchronize(h->stream));
  std::vector<int32_t> v(4 * n_atoms);
  if (n_atoms)
    HIP_TRY(copy_sync(h, v.data(), h->d_atoms + 4 * h->lat_off[chain],
                      sizeof(int32_t) * v.size(), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < v.size(); ++i) atoms[i] = v[i];
  return MCG_OK;
}

// NFOLD event by event (the kernel's max_events >= 0 mode); no records
int nfold_advance(mcg_handle *h, int64_t max_events, double stop_time, bool select_at_end) {
  h->last_ms = 0.0;
  h->last_attempts = 0.0;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  launch_nfold(h->d_np, h->d_xs, h->d_lat, h->d_list, h->d_series, h->series_cap, h->d_active,
               h->n_chains, 0, h->sample_period, 0, max_events, stop_time, select_at_end ? 1 : 0,
               h->nfold_lds, h->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  std::vector<ExactState> xs(h->n_chains);
  HIP_TRY(copy_sync(h, xs.data(), h->d_xs, sizeof(ExactState) * h->n_chains,
                    hipMemcpyDeviceToHost));
  bool frozen = false;
  double events = 0.0;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    events += (double)(xs[c].n_accept - (h->n_pass[c] * h->sites(c) + h->n_step[c]));
    h->n_pass[c] = xs[c].n_pass;
    h->n_step[c] = xs[c].step;
    frozen = frozen || xs[c].frozen;
  }
  h->last_attempts = events;
  if (frozen)
    return fail(MCG_ERR_STATE,
                "NFOLD: the total rate is 0 (every flip has probability 0 at this temperature), "
                "so the next event never comes; the chain stopped there");
  return MCG_OK;
}

int mcg_kmc_advance(mcg_handle *h, int64_t max_events, double stop_time, int64_t select_at_end) {
  if (!h) return fail(MCG_ERR_ARG, "null handle");
  if (!h->kmc() && !h->nfold()) return fail(MCG_ERR_STATE, "mcg_kmc_advance: KMC and NFOLD modes only");
  if (max_events < 0 || std::isnan(stop_time)) return fail(MCG_ERR_ARG, "bad event count or stop time");
  if (h->nfold()) return nfold_advance(h, max_events, stop_time, select_at_end != 0);
  return kmc_run(h, 0, max_events, stop_time, select_at_end != 0);
}

int mcg_kmc_collect_hops(mcg_handle *h, int64_t jumps_per_position_sample,
                         int64_t max_n_position_samples) {
  if (!h) return fail(MCG_ERR_ARG, "null handle");
  if (!h->kmc()) return fail(MCG_ERR_STATE, "hop correlations exist in KMC mode only");
  if (jumps_per_position_sample < 0 || jumps_per_position_sample > (1 << 30))
    return fail(MCG_ERR_ARG, "jumps_per_position_sample must lie in [0, 2^30]");
  if (jumps_per_position_sample > 0 &&
      (max_n_position_samples < 1 || max_n_position_samples > 100000))
    return fail(MCG_ERR_ARG, "max_n_position_samples must lie in [1, 100000]");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (jumps_per_position_sample == 0) {
    h->kmc_hop_every = 0;
  } else {
    if (h->d_hop) (void)hipFree(h->d_hop);
    h->d_hop = nullptr;
    const size_t recs = (size_t)max_n_position_samples * (size_t)h->lat_bytes;
    if (hipMalloc(&h->d_hop, sizeof(KmcHop) * recs) != hipSuccess) {
      h->d_hop = nullptr;
      h->kmc_hop_every = 0;
      return fail(MCG_ERR_CAPACITY, "out of device memory for the hop-correlation records");
    }
    if (!h->d_hop_cnt) HIP_TRY(hipMalloc(&h->d_hop_cnt, sizeof(int32_t) * (size_t)h->lat_bytes));
    HIP_TRY(fill_sync(h, h->d_hop_cnt, 0, sizeof(int32_t) * (size_t)h->lat_bytes));
    h->kmc_hop_every = (int32_t)jumps_per_position_sample;
    h->kmc_hop_max = (int32_t)max_n_position_samples;
  }
  for (int c = 0; c < h->n_chains; ++c) {
    h->kp_kmc[c].hop_every = h->kmc_hop_every;
    h->kp_kmc[c].hop_max = h->kmc_hop_max;
    h->kp_kmc[c].hop_off = (int64_t)h->kmc_hop_max * h->lat_off[c];
  }
  HIP_TRY(copy_sync(h, h->d_kp_kmc, h->kp_kmc.data(), sizeof(KmcParams) * h->n_chains,
                    hipMemcpyHostToDevice));
  return MCG_OK;
}

int mcg_kmc_get_hops(mcg_handle *h, int64_t chain, int64_t n_atoms, int64_t *counts,
                     int64_t *records, double *times) {
  if (!valid_chain(h, chain) || !counts || !records || !times)
    return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (!h->kmc() || !h->d_hop || !h->d_hop_cnt)
    return fail(MCG_ERR_STATE, "no hop correlations collected (mcg_kmc_collect_hops)");
  if (n_atoms != h->kp_kmc[chain].n_atoms) return fail(MCG_ERR_ARG, "atom count mismatch");
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int64_t n = h->sites(chain), m = h->kmc_hop_max;
  std::vector<int32_t> cnt(n_atoms);
  if (n_atoms > 0)
    HIP_TRY(copy_sync(h, cnt.data(), h->d_hop_cnt + h->lat_off[chain], sizeof(int32_t) * n_atoms,
                      hipMemcpyDeviceToHost));
  std::vector<KmcHop> rec((size_t)(m * n));
  HIP_TRY(copy_sync(h, rec.data(), h->d_hop + (size_t)m * (size_t)h->lat_off[chain],
                    sizeof(KmcHop) * rec.size(), hipMemcpyDeviceToHost));
  for (int64_t a = 0; a < n_atoms; ++a) counts[a] = cnt[a];
  for (int64_t k = 0; k < m; ++k)
    for (int64_t a = 0; a < n_atoms; ++a) {
      const bool have = k < cnt[a];
      KmcHop const &r = rec[(size_t)(k * n + a)];
      int64_t *o = records + 4 * (k * n_atoms + a);
      o[0] = have ? r.dx : 0;
      o[1] = have ? r.dy : 0;
      o[2] = have ? r.step : 0;
      o[3] = have ? r.pass : 0;
      times[k * n_atoms + a] = have ? r.time : 0.0;
    }
  return MCG_OK;
}

int mcg_kmc_set_spin_limit(mcg_handle *h, int64_t attempts) {
  if (!h || attempts < 1) return fail(MCG_ERR_ARG, "bad handle or spin limit");
  if (!h->kmc()) return fail(MCG_ERR_STATE, "the spin limit exists in KMC mode only");
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->kmc_spin_limit = attempts;
  for (int c = 0; c < h->n_chains; ++c) h->kp_kmc[c].spin_limit = attempts;
  HIP_TRY(copy_sync(h, h->d_kp_kmc, h->kp_kmc.data(), sizeof(KmcParams) * h->n_chains,
                    hipMemcpyHostToDevice));
  return MCG_OK;
}

int mcg_set_site_asym(mcg_handle *h, int64_t chain, const int8_t *asym, int64_t n) {
  if (!valid_chain(h, chain) || !asym) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (!h->occupant()) return fail(MCG_ERR_STATE, "site units exist in OCCUPANT mode only");
  if (n != h->sites(chain)) return fail(MCG_ERR_ARG, "unit array size mismatch");
  for (int64_t l = 0; l < n; ++l)
    if (asym[l] < 0 || asym[l] >= h->occ_n_asym)
      return fail(MCG_ERR_ARG, "asymmetric unit index out of range");
  HIP_TRY(hipStreamSynchronize(h->stream));
  const std::vector<int8_t> old_units = h->site_asym[chain];
  h->site_asym[chain].assign(asym, asym + n);
  const int rc = occupant_setup_chain(h, (int)chain, false);
  if (rc != MCG_OK) {
    h->site_asym[chain] = old_units;
    return rc;
  }
  HIP_TRY(fill_sync(h, h->d_lat + h->lat_off[chain], 0, (size_t)n));
  std::vector<int8_t> zero(n, 0);
  const int rc2 = reset_observables(h, chain, zero);
  if (rc2 != MCG_OK) return rc2;
  return occupant_init_lists(h, chain, zero);
}

int mcg_set_active(mcg_handle *h, const uint8_t *active, int64_t n_chains) {
  if (!h || !active || n_chains != h->n_chains) return fail(MCG_ERR_ARG, "bad active mask");
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int c = 0; c < h->n_chains; ++c) h->active[c] = active[c] ? 1 : 0;
  HIP_TRY(copy_sync(h, h->d_active, h->active.data(), h->n_chains, hipMemcpyHostToDevice));
  h->work_dirty = true;
  if (h->mode == MCG_MODE_EXACT) return exact_update_masks(h);
  if (h->mode == MCG_MODE_CHECKERBOARD && !h->fcc()) return checker_update_masks(h);
  return MCG_OK;
}

// OCCUPANT runs of n_passes passes plus n_steps steps per active chain.  A
// chain whose swap types have no events (OccEventProposal.hh:136-139,
// 290-293) stops there; the run then fails with the reference's message after
// every chain has gone as far as it can.
int occupant_run(mcg_handle *h, int64_t n_passes, int64_t n_steps) {
  h->last_ms = 0.0;
  h->last_attempts = 0.0;
  if (n_passes == 0 && n_steps == 0) return MCG_OK;
  const char *no_events = h->occ_canonical
                              ? "Error in choose_canonical_swap: No events possible."
                              : "Error in choose_semigrand_canonical_swap: No events possible.";
  HIP_TRY(hipStreamSynchronize(h->stream));
  std::vector<ExactState> xs(h->n_chains);
  HIP_TRY(copy_sync(h, xs.data(), h->d_xs, sizeof(ExactState) * h->n_chains,
                    hipMemcpyDeviceToHost));
  int64_t need = 0;
  double attempts = 0.0;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    OccParams const &P = h->op[c];
    // the swap types' event count at the start (OccEventProposal.hh:127-134)
    int64_t total = 0;
    for (int i = 0; i < P.n_swaps; ++i) {
      auto cnt = [&](int ci) -> int64_t {
        const int b = P.cand_asym[ci];
        return P.cand_occ[ci] ? P.asym_size[b] - xs[c].occ_cnt0[b] : xs[c].occ_cnt0[b];
      };
      total += P.canonical ? cnt(P.swap_a[i]) * cnt(P.swap_b[i]) : cnt(P.swap_a[i]);
    }
    if (total == 0) return fail(MCG_ERR_STATE, no_events);
    const int64_t p0 = h->n_pass[c];
    const int64_t passes = n_passes + (h->n_step[c] + n_steps) / P.n_mut;
    need = std::max(need, h->n_samples[c] + (p0 + passes) / h->sample_period -
                              p0 / h->sample_period);
    attempts += (double)P.n_mut * (double)n_passes + (double)n_steps;
  }
  int rc = ensure_series(h, need);
  if (rc != MCG_OK) return rc;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  launch_occupant(h->d_op, h->d_ot, h->d_xs, h->d_lat, h->d_list, h->d_series, h->series_cap,
                  h->d_active, h->n_chains, n_passes, n_steps, h->sample_period, h->occ_lds,
                  h->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_attempts = attempts;
  HIP_TRY(copy_sync(h, xs.data(), h->d_xs, sizeof(ExactState) * h->n_chains,
                    hipMemcpyDeviceToHost));
  bool frozen = false;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    h->n_pass[c] = xs[c].n_pass;
    h->n_step[c] = xs[c].step;
    h->n_samples[c] = std::min(xs[c].n_samples, h->series_cap);
    frozen = frozen || xs[c].frozen;
  }
  if (frozen) return fail(MCG_ERR_STATE, no_events);
  return MCG_OK;
}

int mcg_run(mcg_handle *h, int64_t n_passes) {
  if (!h) return fail(MCG_ERR_ARG, "null handle");
  if (n_passes < 0) return fail(MCG_ERR_ARG, "n_passes < 0");
  if (h->occupant()) return occupant_run(h, n_passes, 0);
  if (h->kmc()) return kmc_run(h, n_passes, 0);
  h->last_ms = 0.0;
  h->last_attempts = 0.0;
  const bool time_grid = h->nfold() && h->sample_time_period > 0.0;
  // NFOLD by pass with MCG_OPT_NFOLD_SAMPLE_START: the pass-0 sample of a
  // fresh chain (the kernel's condition: no pass and no sample yet); a run of
  // 0 passes takes just that sample
  auto start_sample = [&](int c) {
    return h->nfold() && !time_grid && (h->options & MCG_OPT_NFOLD_SAMPLE_START) &&
           h->n_pass[c] == 0 && h->n_samples[c] == 0;
  };
  // NFOLD by pass: a run of 0 passes selects the next event of each active
  // chain that has none pending and records the current pass
  const bool sample_now = h->nfold() && !time_grid && n_passes == 0;
  if (n_passes == 0 && !sample_now) return MCG_OK;
  // NFOLD records the host cannot count: time-grid records, and a pass-0 or
  // asked-for sample, which the kernel skips when an event is pending
  bool count_on_device = time_grid || sample_now;
  for (int c = 0; c < h->n_chains; ++c) count_on_device |= h->active[c] && start_sample(c);
  // grow the series to hold every sample this run can append
  int64_t need = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    int64_t p0 = h->n_pass[c];
    int64_t add = (p0 + n_passes) / h->sample_period - p0 / h->sample_period +
                  (start_sample(c) || sample_now ? 1 : 0);
    need = std::max(need, h->n_samples[c] + add);
  }
  int rc = ensure_series(h, need);
  if (rc != MCG_OK) return rc;
  double attempts = 0.0;
  for (int c = 0; c < h->n_chains; ++c)
    if (h->active[c]) attempts += (double)h->sites(c) * (double)n_passes;
  if (h->canonical())
    for (int c = 0; c < h->n_chains; ++c)
      if (h->active[c] && (h->kp[c].nA == 0 || h->kp[c].nB == 0))
        // OccEventProposal.hh:136-139 (occupation_metropolis.hh:121-127 stops the run)
        return fail(MCG_ERR_STATE, "Error in choose_canonical_swap: No events possible.");
  const bool hbm_path = !h->sequential() && !h->checker_lds && !h->fcc();
  // the graph holds the parity-coloured launches only: with ring-coloured
  // chains active, passes are enqueued one by one
  const int G = (hbm_path && h->any_ring()) ? 1 : h->graph_passes;
  if (hbm_path) {
    if (h->work_dirty) {
      rc = rebuild_work(h);
      if (rc != MCG_OK) return rc;
      h->graph_dirty = true;
    }
    if (G > 1 && n_passes >= G && h->graph_dirty) {
      rc = build_graph(h);
      if (rc != MCG_OK) return rc;
    }
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  if (h->mode == MCG_MODE_EXACT) {
    rc = exact_launch(h, n_passes, 0);
    if (rc != MCG_OK) return rc;
  } else if (h->canonical()) {
    launch_canonical(h->d_kp, h->d_xs, h->d_lat, h->d_list, h->d_series, h->series_cap,
                     h->d_active, h->n_chains, n_passes, h->sample_period, h->canon_lds,
                     h->stream);
  } else if (h->nfold()) {
    launch_nfold(h->d_np, h->d_xs, h->d_lat, h->d_list, h->d_series, h->series_cap,
                 h->d_active, h->n_chains, n_passes, h->sample_period, sample_now ? 1 : 0, -1,
                 std::numeric_limits<double>::infinity(), 0, h->nfold_lds, h->stream);
  } else if (h->fcc()) {
    HIP_TRY(hipGetLastError());
    const bool lines = h->fcc_bits && !h->fcc_wide && h->line_tiles == 0;
    launch_fcc(h->d_cp, h->d_cs, h->d_lat, h->d_ftab, h->d_tri, h->d_tri_lines,
               h->d_fcc_sched, h->d_series,
               h->series_cap, h->d_active, h->n_chains, n_passes, h->sample_period,
               lines ? h->fcc_line_lds : h->fcc_bits ? h->fcc_bits_lds : h->checker_lds_bytes,
               h->checker_lds, h->fcc_bits, h->fcc_wide, lines,
               lines ? std::min(512, std::max(h->checker_block, h->fcc_line_block))
                     : h->checker_block,
               h->stream);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess)
      return fail(MCG_ERR_HIP, std::string("fcc launch (block ") +
                                   std::to_string(h->checker_block) + ", lds " +
                                   std::to_string(h->checker_lds_bytes) +
                                   "): " + hipGetErrorString(le));
  } else if (h->checker_lds) {
    // parity-coloured chains on the handle's stream, ring-coloured ones beside
    // them on the second stream (they touch disjoint chains)
    if (h->any_ring()) {
      HIP_TRY(hipEventRecord(h->ev_fork, h->stream));
      HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
      launch_checker_ring_lds(h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_series, h->series_cap,
                              h->d_mask_ring, h->n_chains, n_passes, h->sample_period,
                              h->checker_lds_bytes, h->checker_block, h->stream2);
      HIP_TRY(hipGetLastError());
    }
    launch_checker_lds(h->d_cp, h->d_cs, h->d_lat, h->d_tab, h->d_series, h->series_cap,
                       h->d_mask_even, h->n_chains, n_passes, h->sample_period,
                       h->checker_lds_bytes, h->checker_block, h->stream);
    if (h->any_ring()) {
      HIP_TRY(hipEventRecord(h->ev_join, h->stream2));
      HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
    }
  } else {
    int64_t done = 0;
    if (G > 1 && n_passes >= G)
      for (; done + G <= n_passes; done += G) HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
    for (; done < n_passes; ++done) enqueue_pass(h);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_attempts = attempts;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    int64_t p0 = h->n_pass[c];
    if (!count_on_device)
      h->n_samples[c] += (p0 + n_passes) / h->sample_period - p0 / h->sample_period;
    h->n_pass[c] = p0 + n_passes;
  }
  if (count_on_device) {
    std::vector<ExactState> xs(h->n_chains);
    HIP_TRY(copy_sync(h, xs.data(), h->d_xs, sizeof(ExactState) * h->n_chains,
                      hipMemcpyDeviceToHost));
    int64_t over = 0;
    for (int c = 0; c < h->n_chains; ++c) {
      if (!h->active[c]) continue;
      if (time_grid && xs[c].frozen)
        return fail(MCG_ERR_STATE,
                    "NFOLD chain " + std::to_string(c) +
                        ": the total rate is 0 (every flip has probability 0 at this "
                        "temperature), so the next event never comes and the time grid "
                        "would never end; the chain's state is undefined");
      over = std::max(over, xs[c].n_samples - h->series_cap);
      h->n_samples[c] = std::min(xs[c].n_samples, h->series_cap);
    }
    if (over > 0)
      return fail(MCG_ERR_CAPACITY,
                  "NFOLD time-grid samples exceeded the sample series by " + std::to_string(over) +
                      " records (they were dropped); reserve more with mcg_reserve_samples");
  }
  return MCG_OK;
}

int mcg_run_steps(mcg_handle *h, int64_t n_steps) {
  if (!h) return fail(MCG_ERR_ARG, "null handle");
  if (h->mode != MCG_MODE_EXACT && !h->occupant() && !h->kmc())
    return fail(MCG_ERR_STATE,
                "mcg_run_steps: single steps are run in EXACT, OCCUPANT and KMC modes only");
  if (n_steps < 0) return fail(MCG_ERR_ARG, "n_steps < 0");
  if (h->occupant()) return occupant_run(h, 0, n_steps);
  if (h->kmc()) return kmc_run(h, 0, n_steps);
  h->last_ms = 0.0;
  h->last_attempts = 0.0;
  if (n_steps == 0) return MCG_OK;
  // passes each chain completes and the records they append
  std::vector<int64_t> passes(h->n_chains, 0), added(h->n_chains, 0);
  int64_t need = 0;
  double attempts = 0.0;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    const int64_t p0 = h->n_pass[c];
    passes[c] = (h->n_step[c] + n_steps) / h->sites(c);
    added[c] = (p0 + passes[c]) / h->sample_period - p0 / h->sample_period;
    need = std::max(need, h->n_samples[c] + added[c]);
    attempts += (double)n_steps;
  }
  int rc = ensure_series(h, need);
  if (rc != MCG_OK) return rc;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  rc = exact_launch(h, 0, n_steps);
  if (rc != MCG_OK) return rc;
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_attempts = attempts;
  for (int c = 0; c < h->n_chains; ++c) {
    if (!h->active[c]) continue;
    const int64_t s = h->n_step[c] + n_steps;
    h->n_step[c] = s - passes[c] * h->sites(c);
    h->n_pass[c] += passes[c];
    h->n_samples[c] += added[c];
  }
  return MCG_OK;
}

int mcg_get_position(mcg_handle *h, int64_t chain, int64_t *n_pass, int64_t *step,
                     double *e_run) {
  if (!valid_chain(h, chain)) return fail(MCG_ERR_ARG, "bad chain");
  if (!h->sequential()) return fail(MCG_ERR_STATE, "mcg_get_position: sequential modes only");
  HIP_TRY(hipStreamSynchronize(h->stream));
  ExactState st;
  HIP_TRY(copy_sync(h, &st, h->d_xs + chain, sizeof(st), hipMemcpyDeviceToHost));
  if (n_pass) *n_pass = st.n_pass;
  if (step)
    *step = (h->mode == MCG_MODE_EXACT || h->occupant() || h->kmc() || h->nfold()) ? st.step : 0;
  if (e_run) *e_run = st.e_run;
  return MCG_OK;
}

int mcg_get_nfold_state(mcg_handle *h, int64_t chain, double *time, int64_t *pending_site,
                        double *pending_dt) {
  if (!valid_chain(h, chain)) return fail(MCG_ERR_ARG, "bad chain");
  if (!h->nfold() && !h->kmc())
    return fail(MCG_ERR_ARG, "mcg_get_nfold_state: not an NFOLD or KMC handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  ExactState st;
  HIP_TRY(copy_sync(h, &st, h->d_xs + chain, sizeof(st), hipMemcpyDeviceToHost));
  if (time) *time = st.time;
  if (pending_site) *pending_site = st.pending_site;
  if (pending_dt) *pending_dt = st.pending_dt;
  return MCG_OK;
}

int mcg_get_counters(mcg_handle *h, int64_t chain, int64_t *n_pass, int64_t *n_accept,
                     int64_t *n_reject) {
  if (!valid_chain(h, chain)) return fail(MCG_ERR_ARG, "bad chain");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sequential()) {
    ExactState st;
    HIP_TRY(copy_sync(h, &st, h->d_xs + chain, sizeof(st), hipMemcpyDeviceToHost));
    if (n_pass) *n_pass = st.n_pass;
    if (n_accept) *n_accept = st.n_accept;
    if (n_reject) *n_reject = st.n_reject;
  } else {
    CheckerState st;
    HIP_TRY(copy_sync(h, &st, h->d_cs + chain, sizeof(st), hipMemcpyDeviceToHost));
    if (n_pass) *n_pass = st.n_pass;
    if (n_accept) *n_accept = st.n_accept;
    if (n_reject) *n_reject = st.n_pass * h->sites(chain) - st.n_accept;
  }
  return MCG_OK;
}

int mcg_n_samples(mcg_handle *h, int64_t chain, int64_t *n_samples) {
  if (!valid_chain(h, chain) || !n_samples) return fail(MCG_ERR_ARG, "bad chain");
  *n_samples = h->n_samples[chain];
  return MCG_OK;
}

int mcg_get_samples(mcg_handle *h, int64_t chain, int64_t first, int64_t count,
                    mcg_sample *out) {
  if (!valid_chain(h, chain) || (!out && count)) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (first < 0 || count < 0 || first + count > h->n_samples[chain])
    return fail(MCG_ERR_ARG, "sample range out of bounds");
  if (!count) return MCG_OK;
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(copy_sync(h, out, h->d_series + chain * h->series_cap + first,
                    sizeof(mcg_sample) * count, hipMemcpyDeviceToHost));
  return MCG_OK;
}

int mcg_reserve_samples(mcg_handle *h, int64_t n_records) {
  if (!h || n_records < 0 || n_records > (int64_t(1) << 40))
    return fail(MCG_ERR_ARG, "bad handle or record count");
  return ensure_series(h, n_records);
}

int mcg_clear_samples(mcg_handle *h) {
  if (!h) return fail(MCG_ERR_ARG, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sequential()) {
    std::vector<ExactState> xs(h->n_chains);
    HIP_TRY(copy_sync(h, xs.data(), h->d_xs, sizeof(ExactState) * h->n_chains,
                      hipMemcpyDeviceToHost));
    for (auto &s : xs) s.n_samples = 0;
    HIP_TRY(copy_sync(h, h->d_xs, xs.data(), sizeof(ExactState) * h->n_chains,
                      hipMemcpyHostToDevice));
  } else {
    std::vector<CheckerState> cs(h->n_chains);
    HIP_TRY(copy_sync(h, cs.data(), h->d_cs, sizeof(CheckerState) * h->n_chains,
                      hipMemcpyDeviceToHost));
    for (auto &s : cs) s.n_samples = 0;
    HIP_TRY(copy_sync(h, h->d_cs, cs.data(), sizeof(CheckerState) * h->n_chains,
                      hipMemcpyHostToDevice));
  }
  std::fill(h->n_samples.begin(), h->n_samples.end(), 0);
  return MCG_OK;
}

// Grow the statistics scratch to `need` doubles.  The new buffer is allocated
// before the old one is released, and the handle's pointer and capacity change
// together, so a failed allocation leaves the previous (valid) buffer in place.
static int ensure_stats(mcg_handle *h, int64_t need) {
  if (h->d_stats && need <= h->stats_cap) return MCG_OK;
  const int64_t cap = std::max<int64_t>(need, 2 * h->stats_cap);
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * cap) != hipSuccess)
    return fail(MCG_ERR_CAPACITY, "out of device memory for the statistics scratch");
  if (h->d_stats) (void)hipFree(h->d_stats);
  h->d_stats = d;
  h->stats_cap = cap;
  return MCG_OK;
}

static ObsParams obs_params(mcg_handle *h, int64_t chain, int64_t observable) {
  ObsParams op;
  op.n_sites = h->sites(chain);
  op.negJ1 = -h->J1;
  op.negJ2 = -h->J2;
  op.mu = h->canonical() ? 0.0 : h->chains[chain].exchange_potential;
  // only EXACT and CANONICAL records carry the row/column fold
  op.fold = ((h->mode == MCG_MODE_EXACT || h->canonical()) && !h->use_nlist) ? 1 : 0;
  op.observable = (int32_t)observable;
  op.fcc = h->fcc() ? 1 : 0;
  op.running = h->occupant() ? 1 : 0;
  op.coef[0] = h->eci[0];
  op.coef[1] = h->eci[1];
  op.coef[2] = h->eci[2] / 6.0;
  op.coef[3] = h->eci[3] / 3.0;
  op.coef[4] = h->eci[4] / 12.0;
  op.coef[5] = h->eci[5] / 8.0;
  return op;
}

// f_confidence = sqrt(2) * approx_erf_inv(confidence) (BasicStatistics.cc:127,
// misc/math.hh:62-72; host libm)
static double confidence_factor(double confidence) {
  const double one = 1.0, PI = 3.141592653589793238463, a = 0.147;
  const double x = confidence;
  const double sgn = (x < 0.0) ? -one : one;
  const double b = std::log((one - x) * (one + x));
  const double cc = 2.0 / (PI * a) + b * 0.5;
  const double d = b / a;
  return std::sqrt(2.0) * (sgn * std::sqrt(std::sqrt(cc * cc - d) - cc));
}

int mcg_stats_weighted(mcg_handle *h, int64_t chain, int64_t observable, int64_t first,
                       double confidence, int64_t method, int64_t n_resamples, mcg_stats *out) {
  if (!valid_chain(h, chain) || !out) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (observable < 0 || observable > 2) return fail(MCG_ERR_ARG, "bad observable");
  if (!h->nfold())
    return fail(MCG_ERR_STATE, "weighted statistics need weighted records (NFOLD mode)");
  if (method != 1 && method != 2)
    return fail(MCG_ERR_ARG, "Error in BasicStatisticsCalculator: invalid method");
  // the lag search is quadratic in n_resamples: bound it
  if (n_resamples < 1 || n_resamples > (1LL << 22))
    return fail(MCG_ERR_ARG, "n_resamples must be in [1, 2^22]");
  const int64_t count = h->n_samples[chain] - first;
  if (first < 0 || count <= 0)
    return fail(MCG_ERR_ARG, "Error in BasicStatisticsCalculator: observations.size()==0");
  {
    int rc = ensure_stats(h, n_resamples + 8);
    if (rc != MCG_OK) return rc;
  }
  launch_weighted_stats(h->d_series + chain * h->series_cap, first, count,
                        obs_params(h, chain, observable), (int)method, n_resamples, h->d_stats,
                        h->stream);
  HIP_TRY(hipGetLastError());
  double r[8];
  HIP_TRY(hipMemcpyAsync(r, h->d_stats, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  out->mean = r[0];
  out->variance = r[1];
  out->autocorr_factor = r[2];
  out->n = count;
  // BasicStatistics.cc:176-178 (method 1: / W) and :127-128 via :183 (method 2: / n)
  out->calculated_precision = confidence_factor(confidence) * std::sqrt(r[2] * r[1] / r[4]);
  return MCG_OK;
}

int mcg_stats_compute(mcg_handle *h, int64_t chain, int64_t observable, int64_t first,
                      double confidence, mcg_stats *out) {
  if (!valid_chain(h, chain) || !out) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (observable < 0 || observable > 2) return fail(MCG_ERR_ARG, "bad observable");
  // weighted records: the reference calculator's defaults (BasicStatistics.hh:
  // method 1, 10000 resamples)
  if (h->nfold()) return mcg_stats_weighted(h, chain, observable, first, confidence, 1, 10000, out);
  const int64_t count = h->n_samples[chain] - first;
  if (first < 0 || count <= 0)
    return fail(MCG_ERR_ARG, "Error in BasicStatisticsCalculator: observations.size()==0");
  {
    int rc = ensure_stats(h, count + 4);
    if (rc != MCG_OK) return rc;
  }
  launch_stats(h->d_series + chain * h->series_cap, first, count,
               obs_params(h, chain, observable), h->d_stats, h->stream);
  HIP_TRY(hipGetLastError());
  double r[4];
  HIP_TRY(hipMemcpyAsync(r, h->d_stats, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  // BasicStatistics.cc:127-128 (host libm for the confidence factor)
  const double f_conf = confidence_factor(confidence);
  out->mean = r[0];
  out->variance = r[1];
  out->autocorr_factor = r[2];
  out->n = count;
  out->calculated_precision = f_conf * std::sqrt(r[2] * r[1] / (double)count);
  return MCG_OK;
}

int mcg_stats_compute_noalloc(mcg_handle *h, int64_t chain, int64_t observable, int64_t first,
                              double confidence, mcg_stats *out) {
  if (!valid_chain(h, chain) || !out) return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (observable < 0 || observable > 2) return fail(MCG_ERR_ARG, "bad observable");
  const int64_t count = h->n_samples[chain] - first;
  if (first < 0 || count <= 0)
    return fail(MCG_ERR_ARG, "Error in BasicStatisticsCalculator: observations.size()==0");
  // the weighted branch reads the series at the calculator's default 10000
  // resamples (as mcg_stats_compute does for NFOLD records)
  const int64_t n_resamples = 10000;
  // 8 result words at most, then the series (or the resamples)
  const int64_t need = (h->nfold() ? n_resamples : count) + 8;
  if (need > MCG_STATS_NOALLOC_MAX)
    return fail(MCG_ERR_CAPACITY,
                "mcg_stats_compute_noalloc: " + std::to_string(count) +
                    " observations exceed the fixed scratch (at most " +
                    std::to_string(MCG_STATS_NOALLOC_MAX - 8) +
                    "); use mcg_stats_compute");
  double *ptr = stats_noalloc_scratch(h->device);
  if (!ptr) return fail(MCG_ERR_STATE, "statistics scratch not initialised on this device");
  if (h->nfold())
    launch_weighted_stats(h->d_series + chain * h->series_cap, first, count,
                          obs_params(h, chain, observable), 1, n_resamples, ptr, h->stream);
  else
    launch_stats(h->d_series + chain * h->series_cap, first, count,
                 obs_params(h, chain, observable), ptr, h->stream);
  HIP_TRY(hipGetLastError());
  double r[8];
  const size_t nr = h->nfold() ? 8 : 4;
  HIP_TRY(hipMemcpyAsync(r, ptr, sizeof(double) * nr, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  out->mean = r[0];
  out->variance = r[1];
  out->autocorr_factor = r[2];
  out->n = count;
  // as mcg_stats_weighted (method 1: / W) and mcg_stats_compute (/ n)
  const double denom = h->nfold() ? r[4] : (double)count;
  out->calculated_precision = confidence_factor(confidence) * std::sqrt(r[2] * r[1] / denom);
  return MCG_OK;
}

int mcg_check_batch(mcg_handle *h, int64_t n, const int64_t *chains, int64_t observable,
                    double prec, double confidence, mcg_check *out) {
  if (!h || n < 0 || (n > 0 && (!chains || !out))) return fail(MCG_ERR_ARG, "bad arguments");
  if (observable < 0 || observable > 2) return fail(MCG_ERR_ARG, "bad observable");
  if (h->nfold())
    return fail(MCG_ERR_STATE, "mcg_check_batch: weighted (NFOLD) records are checked one by "
                               "one (mcg_equilibration, mcg_stats_weighted)");
  if (n == 0) return MCG_OK;
  if (n > h->n_chains) return fail(MCG_ERR_ARG, "more chains than the handle holds");
  std::vector<int64_t> ints(3 * (size_t)n);  // chains, counts, scratch offsets
  std::vector<ObsParams> ops((size_t)n);
  int64_t scratch = 0;
  for (int64_t b = 0; b < n; ++b) {
    if (!valid_chain(h, chains[b])) return fail(MCG_ERR_ARG, "bad chain in the list");
    const int64_t count = h->n_samples[chains[b]];
    if (count <= 0)
      return fail(MCG_ERR_ARG, "Error in equilibration_check: observations.size()==0");
    ints[b] = chains[b];
    ints[n + b] = count;
    ints[2 * n + b] = scratch;
    scratch += count;
    ops[b] = obs_params(h, chains[b], observable);
  }
  const size_t ibytes = sizeof(int64_t) * ints.size(), obytes = sizeof(ObsParams) * ops.size();
  if ((int64_t)(ibytes + obytes) > h->batch_cap) {
    unsigned char *d = nullptr;
    const int64_t cap = std::max<int64_t>((int64_t)(ibytes + obytes), 2 * h->batch_cap);
    if (hipMalloc(&d, cap) != hipSuccess)
      return fail(MCG_ERR_CAPACITY, "out of device memory for the check list");
    if (h->d_batch) (void)hipFree(h->d_batch);
    h->d_batch = d;
    h->batch_cap = cap;
  }
  {
    const int rc = ensure_stats(h, 8 * n + scratch);
    if (rc != MCG_OK) return rc;
  }
  std::vector<unsigned char> host(ibytes + obytes);
  std::memcpy(host.data(), ints.data(), ibytes);
  std::memcpy(host.data() + ibytes, ops.data(), obytes);
  HIP_TRY(hipMemcpyAsync(h->d_batch, host.data(), host.size(), hipMemcpyHostToDevice,
                         h->stream));
  const int64_t *d_ints = reinterpret_cast<const int64_t *>(h->d_batch);
  launch_check_batch(h->d_series, h->series_cap, d_ints, d_ints + n,
                     reinterpret_cast<const ObsParams *>(h->d_batch + ibytes), d_ints + 2 * n,
                     (int)n, prec, h->d_stats + 8 * n, h->d_stats, h->stream);
  HIP_TRY(hipGetLastError());
  std::vector<double> r(8 * (size_t)n);
  HIP_TRY(hipMemcpyAsync(r.data(), h->d_stats, sizeof(double) * r.size(), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const double f_conf = confidence_factor(confidence);
  for (int64_t b = 0; b < n; ++b) {
    const double *o = r.data() + 8 * b;
    mcg_check &c = out[b];
    c.is_equilibrated = (int64_t)o[0];
    c.n_equilibration = (int64_t)o[1];
    c.stats.n = (int64_t)o[6];
    c.stats.mean = o[2];
    c.stats.variance = o[3];
    c.stats.autocorr_factor = o[4];
    c.stats.calculated_precision =
        c.stats.n > 0 ? f_conf * std::sqrt(o[4] * o[3] / (double)c.stats.n) : 0.0;
  }
  return MCG_OK;
}

int mcg_equilibration(mcg_handle *h, int64_t chain, int64_t observable, double prec,
                      int64_t *is_equilibrated, int64_t *n_samples_for_equilibration) {
  if (!valid_chain(h, chain) || !is_equilibrated || !n_samples_for_equilibration)
    return fail(MCG_ERR_ARG, "bad chain or buffer");
  if (observable < 0 || observable > 2) return fail(MCG_ERR_ARG, "bad observable");
  const int64_t count = h->n_samples[chain];
  if (count <= 0) return fail(MCG_ERR_ARG, "Error in equilibration_check: observations.size()==0");
  // NFOLD records are weighted: EquilibrationCheck.cc:134-160
  launch_equilibration(h->d_series + chain * h->series_cap, count,
                       obs_params(h, chain, observable), prec, h->nfold() ? 1 : 0, h->d_eq,
                       h->stream);
  HIP_TRY(hipGetLastError());
  int64_t r[2];
  HIP_TRY(hipMemcpyAsync(r, h->d_eq, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *is_equilibrated = r[0];
  *n_samples_for_equilibration = r[1];
  return MCG_OK;
}

int mcg_get_schedule(mcg_handle *h, mcg_schedule *out) {
  if (!h || !out) return fail(MCG_ERR_ARG, "null argument");
  out->lds_block = h->checker_block;
  out->units_per_thread = h->units_per_thread;
  out->graph_passes = h->graph_passes;
  out->band_lines = h->band_lines;
  out->line_tiles = h->line_tiles;
  out->trial_ms = h->tuned_ms;
  return MCG_OK;
}

int mcg_set_schedule(mcg_handle *h, const mcg_schedule *s) {
  if (!h || !s) return fail(MCG_ERR_ARG, "null argument");
  if (s->lds_block != 256 && s->lds_block != 512 && s->lds_block != 1024)
    return fail(MCG_ERR_ARG, "lds_block must be 256, 512 or 1024");
  if (s->units_per_thread != 1 && s->units_per_thread != 2 && s->units_per_thread != 4)
    return fail(MCG_ERR_ARG, "units_per_thread must be 1, 2 or 4");
  if (s->graph_passes < 1 || s->graph_passes > 256)
    return fail(MCG_ERR_ARG, "graph_passes must be in [1, 256]");
  if (s->band_lines != -1 && (s->band_lines < 0 || s->band_lines > 4096 || (s->band_lines & 1)))
    return fail(MCG_ERR_ARG, "band_lines must be -1 (automatic), 0 (off) or even <= 4096");
  if (s->line_tiles != 0 && s->line_tiles != 1) return fail(MCG_ERR_ARG, "line_tiles must be 0 or 1");
  if (h->fcc_wide && s->lds_block > 512)
    return fail(MCG_ERR_ARG, "lds_block: FCC lines of more than 32 sites run at most 512 threads");
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->checker_block = (int)s->lds_block;
  if (h->band_lines != (int)s->band_lines || h->units_per_thread != (int)s->units_per_thread ||
      h->line_tiles != (int)s->line_tiles)
    h->work_dirty = true;
  h->line_tiles = (int)s->line_tiles;
  h->units_per_thread = (int)s->units_per_thread;
  h->graph_passes = (int)s->graph_passes;
  h->band_lines = (int)s->band_lines;
  h->graph_dirty = true;
  return MCG_OK;
}

int mcg_autotune(mcg_handle *h, int64_t trial_passes, mcg_schedule *chosen) {
  if (!h) return fail(MCG_ERR_ARG, "null handle");
  if (h->sequential()) return chosen ? mcg_get_schedule(h, chosen) : MCG_OK;
  const std::string key = tune_key(h);
  mcg_schedule best{};
  if (!tune_cache_lookup(key, &best)) {
    trial_passes = std::max<int64_t>(trial_passes, 16);
    HIP_TRY(hipStreamSynchronize(h->stream));
    // snapshot of everything a trial run changes
    int8_t *lat_snap = nullptr;
    CheckerState *cs_snap = nullptr;
    HIP_TRY(hipMalloc(&lat_snap, (size_t)h->lat_bytes));
    HIP_TRY(hipMalloc(&cs_snap, sizeof(CheckerState) * h->n_chains));
    HIP_TRY(copy_sync(h, lat_snap, h->d_lat, (size_t)h->lat_bytes, hipMemcpyDeviceToDevice));
    HIP_TRY(copy_sync(h, cs_snap, h->d_cs, sizeof(CheckerState) * h->n_chains,
                      hipMemcpyDeviceToDevice));
    const auto ns_snap = h->n_samples, np_snap = h->n_pass;
    auto restore = [&]() -> int {
      HIP_TRY(copy_sync(h, h->d_lat, lat_snap, (size_t)h->lat_bytes, hipMemcpyDeviceToDevice));
      HIP_TRY(copy_sync(h, h->d_cs, cs_snap, sizeof(CheckerState) * h->n_chains,
                        hipMemcpyDeviceToDevice));
      h->n_samples = ns_snap;
      h->n_pass = np_snap;
      return MCG_OK;
    };
    std::vector<mcg_schedule> cands;
    if (h->fcc() && h->fcc_bits && !h->fcc_wide) {
      // the class-line kernel (its block follows from the lines per class),
      // then the 4-site unit kernel at three block sizes
      for (int64_t b : {256, 512})
        cands.push_back({b, h->units_per_thread, 1, 0.0, h->band_lines, 0});
      for (int64_t b : {256, 512, 1024})
        cands.push_back({b, h->units_per_thread, 1, 0.0, h->band_lines, 1});
    } else if (h->checker_lds || h->fcc()) {
      for (int64_t b : {256, 512, 1024})
        if (!(h->fcc_wide && b > 512))
          cands.push_back({b, h->units_per_thread, 1, 0.0, h->band_lines, h->line_tiles});
    } else {
      const bool line_class = h->ncolor == 4 || h->ncolor == 8;
      for (int64_t lt : {0, 1}) {
        if (lt && !line_class) continue;
        for (int64_t u : {1, 2, 4})
          for (int64_t g : {1, 8}) cands.push_back({h->checker_block, u, g, 0.0, 0, lt});
      }
      if (h->lattice == MCG_LATTICE_SQUARE)
        for (int64_t bl : {8, 16, 32, 64})
          for (int64_t g : {1, 8})
            cands.push_back({h->checker_block, h->units_per_thread, g, 0.0, bl, 0});
      if (h->lattice == MCG_LATTICE_SC && h->ncolor == 8)  // slabs of 4 to 16 planes
        for (int64_t bl : {4, 8, 16})
          for (int64_t g : {1, 8})
            cands.push_back({h->checker_block, h->units_per_thread, g, 0.0, bl, 0});
    }
    double best_ms = 1e300;
    int rc = MCG_OK;
    for (auto &c : cands) {
      if ((rc = mcg_set_schedule(h, &c)) != MCG_OK) break;
      if ((rc = mcg_run(h, std::min<int64_t>(trial_passes, 8))) != MCG_OK) break;  // warm-up
      if ((rc = restore()) != MCG_OK) break;
      if ((rc = mcg_run(h, trial_passes)) != MCG_OK) break;
      c.trial_ms = h->last_ms;
      if ((rc = restore()) != MCG_OK) break;
      if (c.trial_ms < best_ms) {
        best_ms = c.trial_ms;
        best = c;
      }
    }
    (void)hipFree(lat_snap);
    (void)hipFree(cs_snap);
    if (rc != MCG_OK) return rc;
    tune_cache_store(key, best);
  }
  int rc = mcg_set_schedule(h, &best);
  if (rc != MCG_OK) return rc;
  h->tuned_ms = best.trial_ms;
  if (chosen) *chosen = best;
  return MCG_OK;
}

int mcg_last_run_time(mcg_handle *h, double *ms, double *attempts) {
  if (!h) return fail(MCG_ERR_ARG, "null handle");
  if (ms) *ms = h->last_ms;
  if (attempts) *attempts = h->last_attempts;
  return MCG_OK;
}

int mcg_fcc_class_schedule(int64_t L0, int64_t L1, int64_t L2, int64_t *n_steps,
                           int64_t *start, int64_t *cls) {
  if (!n_steps || !start || !cls) return fail(MCG_ERR_ARG, "null argument");
  if (L0 < 1 || L1 < 1 || L2 < 1 || L0 > FCC_LMAX || L1 > FCC_LMAX || L2 > FCC_LMAX)
    return fail(MCG_ERR_ARG, "FCC edges must lie in [1, 127]");
  FccSched S;
  fcc_class_schedule((int)L0, (int)L1, (int)L2, S);
  *n_steps = S.n_steps;
  for (int i = 0; i <= S.n_steps; ++i) start[i] = S.start[i];
  for (int i = 0; i < S.n_classes; ++i) cls[i] = S.cls[i];
  return MCG_OK;
}

}  // extern "C"
