// internal.hh -- device data layout shared by the kernels and the C ABI host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/casm_mi355x/mcgpu.h"
#include "rng.hh"

namespace mcg {

// ---------------------------------------------------------------------------
// EXACT mode: one chain per 64-lane workgroup (one wavefront).
// ---------------------------------------------------------------------------

// Number of (old spin, neighbour sum) classes of the 2D square NN Ising flip:
// class k = (s_old > 0) * 5 + (nbr_sum + 4) / 2.
constexpr int EXACT_NCLASS = 10;

// Read-only per-chain parameters (built on the host).
struct ExactParams {
  int32_t rows, cols;  // shape[0], shape[1]
  int32_t n;           // sites
  int32_t fold;        // 1: compute the row/column-fold formation energy at samples
  int32_t self_nb;     // neighbour slots that are the site itself (rows==1, cols==1)
  int32_t window;      // stream positions evaluated per block by the speculative walk
  uint64_t div_magic;  // floor(l / rows) = (l * div_magic) >> 40 for l < 2^24
  uint64_t thr[EXACT_NCLASS];  // accept a drawn word x iff x < thr[k]
  uint32_t neg_mask;   // bit k: dE < 0, accept without drawing (metropolis.hh:28-30)
  uint32_t full_mask;  // bit k: every drawn word accepts (thr == 2^64)
  double dq[EXACT_NCLASS];  // fl(dE_k / n_unitcells), occupation_metropolis.hh:141
  double negJ;               // -J, for the row/column fold
  int64_t lat_off;           // byte offset of this chain's lattice in the pool
};

// Read-write per-chain state (EXACT, CANONICAL and NFOLD).
struct ExactState {
  uint64_t mt[MT_N];  // std::mt19937_64 state words
  int64_t pos;        // index of the next output (312: twist first)
  int64_t n_pass, n_accept, n_reject;
  int64_t S, M;       // exact bond sum and magnetisation
  double e_run;       // running potential energy per unit cell (NFOLD: the total rate
                      // at the pending event's selection)
  int64_t n_samples;  // records the run produced (written while below the capacity)
  // NFOLD
  double time;           // simulated time
  int64_t pending_site;  // event selected at the end of a sampled pass, or -1
  double pending_dt;     // its time increment
  int64_t next_k;        // next time-grid sample index (sample at next_k * period)
  int64_t frozen;        // NFOLD: the total rate reached 0 on the time grid; OCCUPANT: the
                         // swap types ran out of events
  // EXACT: attempts done in the current pass (0 <= step < n; runs of whole
  // passes keep it, runs of single steps move it); OCCUPANT: steps done in the
  // current pass (0 <= step < n_mut); NFOLD and KMC: events done in the
  // current pass
  int64_t step;
  // OCCUPANT: sites of each asymmetric unit holding its occupant 0
  int32_t occ_cnt0[8];
  // KMC: attempts since the last event, and the sums over atoms of dx^2 +
  // dy^2, dx and dy (time, next_k and frozen are shared with NFOLD)
  int64_t kmc_k;
  int64_t kmc_sum_sq, kmc_sum_dx, kmc_sum_dy;
  // KMC pending event (pending_site >= 0): its direction and attempts; its
  // time is in pending_dt
  int64_t kmc_pdir, kmc_pk;
};

// Every active chain makes n_passes * n + n_steps attempts from where it stands.
void launch_exact(const ExactParams *P, ExactState *S, int8_t *lat_pool, mcg_sample *series,
                  int64_t series_cap, const uint8_t *active, int n_chains, int64_t n_passes,
                  int64_t n_steps, int64_t sample_period, int lds_bytes, bool lattice_in_lds,
                  bool block_walk, hipStream_t stream);

// ---------------------------------------------------------------------------
// OCCUPANT mode: occupant events over an OccLocation with several asymmetric
// units (semi-grand canonical single-site events or canonical pair swaps),
// one chain per wavefront.
// ---------------------------------------------------------------------------
constexpr int OCC_MAX_ASYM = 8;
constexpr int OCC_MAX_CAND = 2 * OCC_MAX_ASYM;
constexpr int OCC_MAX_SWAPS = 16;
// classes: swap type i, neighbour sum of a, and (pairs) of b after a changed,
// k = i * 25 + ((Sa + 4) / 2) * 5 + (Sb + 4) / 2
constexpr int OCC_NCLASS = OCC_MAX_SWAPS * 25;

struct OccParams {
  int32_t rows, cols, n, n_mut;
  int32_t canonical;  // 1: pair swaps (propose_canonical_event), 0: single sites
  int32_t n_asym, n_cand, n_swaps;
  int32_t window;     // stream positions evaluated per block
  int32_t pad;
  uint64_t div_magic; // floor(l / rows) = (l * div_magic) >> 40
  // per unit: segment of the list array (units with one occupant: size 0)
  int32_t asym_off[OCC_MAX_ASYM], asym_size[OCC_MAX_ASYM];
  // per candidate (OccCandidateList order): unit and occupant index
  int32_t cand_asym[OCC_MAX_CAND], cand_occ[OCC_MAX_CAND];
  // swap types: candidate indices (cand_a, cand_b)
  int32_t swap_a[OCC_MAX_SWAPS], swap_b[OCC_MAX_SWAPS];
  int64_t lat_off;    // lattice byte offset; the lists start at entry lat_off of the list pool
};

// per chain: the class thresholds (first word whose canonical double reaches
// exp(-beta dE)), fl(dE / n_unitcells), and flags (bit 0: dE < 0, accepted
// without a draw; bit 1: every word accepts)
struct OccTables {
  uint64_t thr[OCC_NCLASS];
  double dq[OCC_NCLASS];
  uint8_t flags[OCC_NCLASS];
};

// LDS bytes of one chain: lattice and the list array.
inline int occ_lds_bytes(const OccParams &P) {
  return ((P.n + 15) & ~15) + ((2 * P.n_mut + 15) & ~15);
}

// Every active chain makes n_passes * n_mut + n_steps steps from where it
// stands (fewer if its swap types run out of events: ExactState::frozen).
void launch_occupant(const OccParams *P, const OccTables *T, ExactState *S, int8_t *lat_pool,
                     int32_t *list_pool, mcg_sample *series, int64_t series_cap,
                     const uint8_t *active, int n_chains, int64_t n_passes, int64_t n_steps,
                     int64_t sample_period, int lds_bytes, hipStream_t stream);
// Dynamic LDS limit of the occupant kernel (its static LDS is ~23 KB).
hipError_t occupant_set_max_lds(int bytes);

// ---------------------------------------------------------------------------
// KMC mode: kinetic Monte Carlo with atom tracking for a square lattice gas,
// one chain per wavefront.
// ---------------------------------------------------------------------------
// classes of a hop: a - b + 3 (a, b: atoms next to the origin other than at
// the target, next to the target other than at the origin), dE = J (a - b)
constexpr int KMC_NCLASS = 7;

struct KmcParams {
  int32_t rows, cols, n, n_atoms;
  int32_t window;      // stream positions evaluated per block
  uint32_t draw_mask;  // bit k: dE > 0, the class draws an acceptance word
  uint32_t full_mask;  // bit k: every drawn word accepts
  int32_t pad;
  uint64_t div_magic;  // floor(l / rows) = (l * div_magic) >> 40
  double attempt_rate; // R = 4 * n_atoms: attempts per unit time
  double time_period;  // > 0: samples at k * time_period; 0: by pass
  double J;
  int64_t spin_limit;  // attempts without an event after which a chain stops (frozen)
  int64_t lat_off;     // lattice byte offset; the lists start at entry lat_off of the
                       // list pool, the atoms (4 int32 each) at entry 4 lat_off of the atom pool
  uint64_t thr[KMC_NCLASS];
  // hop correlations (0: off): every hop_every-th jump of an atom is recorded,
  // at most hop_max per atom; record k of atom a at hop pool entry
  // hop_off + k * n + a, the atoms' record counts at count pool entry lat_off + a
  int32_t hop_every, hop_max;
  int64_t hop_off;
};

// One hop-correlation record: the atom's displacement since its occupation was
// set, the step within the pass and the pass of the event, and its time.
struct KmcHop {
  int32_t dx, dy;
  int64_t step, pass;
  double time;
};

// LDS bytes of one chain: list, positions and site atoms (16-bit), atom
// displacements and jumps (32-bit).
inline int kmc_lds_bytes(int n) {
  const int a = (n + 7) & ~7;
  return 3 * 2 * a + 3 * 4 * a;
}

// Every active chain applies its pending event, if any, then makes events up
// to n_passes * n + n_steps in all (fewer if it freezes: ExactState::frozen).
// An event whose time reaches stop_time is left pending instead of applied
// (the chain stops there); with select_at_end a chain that reached the count
// selects one more event and leaves it pending.
void launch_kmc(const KmcParams *P, ExactState *S, int8_t *lat_pool, int32_t *list_pool,
                int32_t *atom_pool, KmcHop *hop_pool, int32_t *hop_count_pool,
                mcg_sample *series, int64_t series_cap, const uint8_t *active, int n_chains,
                int64_t n_passes, int64_t n_steps, int64_t sample_period, double stop_time,
                int select_at_end, int lds_bytes, hipStream_t stream);
hipError_t kmc_set_max_lds(int bytes);

// ---------------------------------------------------------------------------
// CANONICAL mode: pair swaps at fixed composition, one chain per wavefront.
// ---------------------------------------------------------------------------

// Classes of a swap (site a: -1 -> +1, then site b: +1 -> -1): the neighbour
// sum of a before the swap and the neighbour sum of b after a has flipped,
// k = ((Sa + 4) / 2) * 5 + (Sb + 4) / 2.
constexpr int CANON_NCLASS = 25;

struct CanonParams {
  int32_t rows, cols, n;
  int32_t nA, nB;        // sizes of the species-0 (spin -1) and species-1 (+1) lists
  int32_t fold;          // 1: compute the row/column-fold formation energy at samples
  int32_t self_nb;       // neighbour slots that are the site itself
  int32_t window;        // stream positions evaluated per block
  uint32_t mark_mask;    // conflict marks: entry (site & mark_mask)
  uint32_t neg_mask;     // bit k: dE < 0, accept without drawing
  uint32_t full_mask;    // bit k: every drawn word accepts
  int32_t pad;
  uint64_t div_magic;    // floor(l / rows) = (l * div_magic) >> 40
  uint64_t thr[CANON_NCLASS];
  double dq[CANON_NCLASS];  // fl(dE_k / n_unitcells)
  double negJ;
  int64_t lat_off;       // lattice byte offset; the lists start at entry lat_off of the list pool
};

// LDS bytes of one chain: lattice, conflict marks, the two lists, fold scratch.
inline int canon_lds_bytes(const CanonParams &P) {
  return ((P.n + 15) & ~15) + 4 * (int)(P.mark_mask + 1) + ((2 * P.n + 15) & ~15) +
         4 * (P.rows + P.cols);
}

void launch_canonical(const CanonParams *P, ExactState *S, int8_t *lat_pool, int32_t *list_pool,
                      mcg_sample *series, int64_t series_cap, const uint8_t *active,
                      int n_chains, int64_t n_passes, int64_t sample_period, int lds_bytes,
                      hipStream_t stream);
// Dynamic LDS limit of the canonical kernel (its static LDS is ~10 KB).
hipError_t canonical_set_max_lds(int bytes);

// ---------------------------------------------------------------------------
// NFOLD mode: rejection-free single-site flips, one chain per wavefront.
// ---------------------------------------------------------------------------
struct NfoldParams {
  int32_t rows, cols, n;
  int32_t pad;
  uint64_t div_magic;   // floor(l / rows) = (l * div_magic) >> 40
  double rate[10];      // Metropolis rate of class k = (s > 0) * 5 + (nbr_sum + 4) / 2
  double time_period;   // > 0: samples on the time grid k * time_period; else by pass
  int64_t lat_off;      // lattice byte offset; the class array starts at entry lat_off
  int32_t sample_start; // by pass: also sample pass 0 (the first event selected then)
  int32_t pad2;
};

// class array and positions (16-bit each; a position's bit 15 is the spin)
inline int nfold_lds_bytes(const NfoldParams &P) { return 4 * P.n + 64; }

// max_events >= 0: event by event (no records, n_passes ignored): apply the
// pending event unless its time reaches stop_time, then up to max_events
// events, leaving pending the one that reaches stop_time or, with
// select_at_end, the one after the last; max_events < 0: whole passes.
void launch_nfold(const NfoldParams *P, ExactState *S, int8_t *lat_pool, int32_t *list_pool,
                  mcg_sample *series, int64_t series_cap, const uint8_t *active, int n_chains,
                  int64_t n_passes, int64_t sample_period, int sample_now, int64_t max_events,
                  double stop_time, int select_at_end, int lds_bytes, hipStream_t stream);
hipError_t nfold_set_max_lds(int bytes);

// ---------------------------------------------------------------------------
// CHECKERBOARD mode
// ---------------------------------------------------------------------------

// Acceptance classes: (old spin, NN sum, NNN sum).  For coordination numbers
// z1 (NN) and z2 (NNN) the sums take z+1 values each (steps of 2).
struct CheckerParams {
  int32_t L0, L1, L2;      // shape
  int32_t ncolor;          // 2, 4 or 8
  int32_t z1, z2;          // coordination numbers (z2 = 0 without J2)
  int32_t ndim;            // 2 or 3
  int32_t quad_only;       // 1: use the 4-site path even when 16-site chunks fit
  uint32_t key0, key1;     // Philox key (chain seed)
  int64_t n;               // sites
  int64_t lat_off;         // byte offset in the lattice pool
  int64_t tab_off;         // offset of the chain's threshold table (uint32 entries)
};

// Acceptance table of a chain: entry (b << 7) | (c1 << 4) | c2 for old spin
// bit b and up-neighbour counts c1 <= 6, c2 <= 12 (unused entries 0), so the
// class bytes of four sites come from one packed-word shift-and-add.
constexpr int CHECKER_NTAB = 256;
__host__ __device__ inline int checker_class(int b, int c1, int c2) {
  return (b << 7) | (c1 << 4) | c2;
}

struct CheckerState {
  int64_t n_pass;
  int64_t n_accept;
  int64_t S1, S2, M;       // exact integer observables (FCC: P1, P2, N1)
  int64_t n_samples;
  int64_t S3, T3;          // FCC: P3 and the triangle count
};

void launch_checker_lds(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                        const uint32_t *tables, mcg_sample *series, int64_t series_cap,
                        const uint8_t *active, int n_chains, int64_t n_passes,
                        int64_t sample_period, int lds_bytes, int block, hipStream_t stream);

void launch_checker_color(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                          const uint32_t *tables, const int32_t *work, int n_work_items,
                          int color, int upt, hipStream_t stream);

// Chains with an odd edge (checker_ring.hip): ring colourings, natural layout.
void launch_checker_ring_lds(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                             const uint32_t *tables, mcg_sample *series, int64_t series_cap,
                             const uint8_t *active, int n_chains, int64_t n_passes,
                             int64_t sample_period, int lds_bytes, int block,
                             hipStream_t stream);
// the class launches of one HBM pass of the listed chains (the pass counter and
// the record follow in launch_checker_record)
void launch_checker_ring_pass(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                              const uint32_t *tables, const int32_t *chains, int n_ring,
                              int per_chain, int ndim, hipStream_t stream);
hipError_t checker_ring_set_max_lds(int bytes);

void launch_checker_record(const CheckerParams *P, CheckerState *St, mcg_sample *series,
                           int64_t series_cap, const uint8_t *active, int n_chains,
                           int64_t sample_period, hipStream_t stream);

hipError_t checker_graph_add_passes(hipGraph_t g, const CheckerParams *P, CheckerState *St,
                                    int8_t *lat_pool, const uint32_t *tables,
                                    const int32_t *work, int n_work_items, int ncolor, int upt,
                                    mcg_sample *series, int64_t series_cap,
                                    const uint8_t *active, int n_chains, int64_t sample_period,
                                    int passes, const int32_t *band_work, int n_bands,
                                    int ring_bytes, bool slabs, const int32_t *pair_work,
                                    int n_pairs, int pair_lds, int tile_grid);

// Line-class path (4 and 8 colourings): one launch per line class, both
// halves of each line.  lds_bytes == 0: items (chain, first colour-class line,
// lines, 0) read straight from HBM, one workgroup per item.  lds_bytes > 0:
// tiles (chain, jh0, kh0, TJ | TK << 16) of class lines staged with their
// neighbour lines in LDS ((2TJ+1) x (2TK+1) x L0 bytes, 2D (2TJ+1) x L0, at
// most checker_tile_max_bytes()), tile_grid persistent workgroups.
void launch_checker_pair(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                         const uint32_t *tables, const int32_t *pair_work, int n_items,
                         int line_class, int lds_bytes, int tile_grid, hipStream_t stream);
hipError_t checker_pair_set_max_lds(int bytes);
int checker_pair_block();
int checker_tile_max_bytes();

// 2D band path: one launch over (chain, first line, lines) bands, then the
// band-edge launch; ring_bytes = checker_band_ring_lines() * L0 (largest chain).
void launch_checker_band(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                         const uint32_t *tables, const int32_t *band_work, int n_bands,
                         int ring_bytes, hipStream_t stream);
int checker_band_ring_lines();
hipError_t checker_band_set_max_lds(int bytes);

// 3D J1+J2 slab path (checker_slab.hip): one launch over (chain, first plane,
// planes) slabs with the planes bit-packed in an LDS ring, then the slab-edge
// launch; ring_bytes = checker_slab_ring_bytes(L0, L1) of the largest plane.
// checker_slab_fits: 8 colours, L0 % 64 == 0, even L1 and L2, the ring within
// CHECKER_SLAB_MAX_LDS.
constexpr int CHECKER_SLAB_MAX_LDS = 144 * 1024;
int checker_slab_ring_bytes(int L0, int L1);
bool checker_slab_fits(const CheckerParams &P);
void launch_checker_slab(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                         const uint32_t *tables, const int32_t *slab_work, int n_slabs,
                         int ring_bytes, hipStream_t stream);
void *checker_slab_kernel_fn();
void *checker_slab_seam_kernel_fn();
int checker_slab_block();
int checker_slab_ring_planes();
hipError_t checker_slab_set_max_lds(int bytes);

// FCC cluster expansion (fcc.hip): CheckerParams with ncolor = 64, tab_off in
// doubles into the per-chain acceptance tables, tri = the 4096-entry
// triangle-count table of the nearest-neighbour occupancy mask.
// bit_lines: the bit-line kernels (every chain L0 <= 64, lattice in LDS as one
// 64-bit word per line, tri_lines the triangle table in its mask bit order);
// class_lines (with bit_lines, every chain L0 <= 32): one thread per class
// line with bit-sliced counts, else 4-site units
struct FccSched;
void launch_fcc(const CheckerParams *P, CheckerState *St, int8_t *lat_pool, const double *tables,
                const uint8_t *tri, const uint8_t *tri_lines, const FccSched *sched,
                mcg_sample *series, int64_t series_cap, const uint8_t *active, int n_chains,
                int64_t n_passes, int64_t sample_period, int lds_bytes, bool lattice_in_lds,
                bool bit_lines, bool wide_lines, bool class_lines, int block,
                hipStream_t stream);
// the triangle pairs of the class-line kernel against the host's table in the
// bit-line mask order (true: they agree for every mask)
bool fcc_tri_pairs_check(const std::vector<uint8_t> &tri_lines);

// ---------------------------------------------------------------------------
// Statistics on the device sample series
// ---------------------------------------------------------------------------
struct ObsParams {
  int64_t n_sites;
  double negJ1, negJ2;
  double mu;
  int32_t fold;  // use e_fold as E_f
  int32_t observable;
  int32_t fcc;   // FCC cluster expansion: E_f from the counts and coef
  int32_t running;  // OCCUPANT: the potential energy is the record's running e_run
  double coef[6];  // FCC: V0, V1, V2/6, V3/3, V4/12, Vt/8
};

void launch_stats(const mcg_sample *series, int64_t first, int64_t count, ObsParams op,
                  double *d_out, hipStream_t stream);

void launch_weighted_stats(const mcg_sample *series, int64_t first, int64_t count, ObsParams op,
                           int method, int64_t n_resamples, double *d_out, hipStream_t stream);
void launch_equilibration(const mcg_sample *series, int64_t count, ObsParams op, double prec,
                          int weighted, int64_t *d_out, hipStream_t stream);
// Equilibration and statistics of n chains in one launch (check_batch_kernel).
void launch_check_batch(const mcg_sample *series, int64_t series_cap, const int64_t *chains,
                        const int64_t *counts, const ObsParams *ops, const int64_t *scratch_off,
                        int n, double prec, double *scratch, double *out, hipStream_t stream);
// Device address of the fixed statistics scratch (MCG_STATS_NOALLOC_MAX doubles)
// on device `dev`: looked up once per device by the init call, with that
// device current; nullptr before it.
hipError_t stats_noalloc_scratch_init(int dev);
double *stats_noalloc_scratch(int dev);

}  // namespace mcg
