/* mcgpu.h -- C ABI of the MI355X occupation-Metropolis library
 * (casmcode_monte_amd/lib/libcasm_monte_mi355x.so).
 *
 * The library replaces the inner loop of libcasm-monte's occupation
 * Metropolis method for the semi-grand-canonical Ising model (and the
 * build-defined lattice models of the statistical modes) with hand-written HIP
 * kernels for gfx950.  The C++ drop-in classes under include/casm/monte/ call
 * these entry points; nothing else in the reference's API changes.  Each entry
 * point names the reference code it replaces (paths relative to the reference
 * repository root).
 *
 * Conventions:
 *  - every function returns MCG_OK (0) or an error code; mcg_last_error()
 *    returns a thread-local message for the last failure on this thread;
 *  - a handle owns its device memory and one HIP stream; the caller owns every
 *    host buffer; there are no callbacks across the ABI;
 *  - a handle is not thread-safe; separate handles may be driven from separate
 *    host threads;
 *  - there is no CPU fallback: without a usable gfx950 device mcg_create fails.
 */
#ifndef CASM_MI355X_MCGPU_H
#define CASM_MI355X_MCGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define MCG_OK 0
#define MCG_ERR_ARG 1      /* invalid argument */
#define MCG_ERR_HIP 2      /* HIP runtime error (no device, launch failure, ...) */
#define MCG_ERR_STATE 3    /* call not valid in the handle's current state */
#define MCG_ERR_CAPACITY 4 /* caller buffer too small / device memory exhausted */

/* ---- modes and lattices ----------------------------------------------------- */
/* EXACT: one sequential chain per wavefront that reproduces the reference's
 *   random-site order and std::mt19937_64 stream bit for bit
 *   (basic_occupation_metropolis.hh:381-422 with the event generator of
 *   basic_semigrand_canonical.hh:308-320).  Square lattice, J1 only.
 * CHECKERBOARD: sublattice-coloured sweeps with a Philox4x32-10 counter RNG;
 *   one pass updates every site once.  Square (2D) or simple-cubic (3D)
 *   lattices with nearest (J1) and next-nearest (J2) couplings.
 * CANONICAL: one sequential chain per wavefront running the reference's
 *   canonical pair-swap Metropolis bit for bit: the generic loop of
 *   occupation_metropolis.hh:87-153 with propose_canonical_event
 *   (OccEventProposal.hh:109-236) over an OccLocation (OccLocation.cc:39-118,
 *   216-283), std::mt19937_64.  The lattice is binary with one asymmetric
 *   unit: species 0 = occupation -1, species 1 = +1, one canonical swap type;
 *   the potential is the square-lattice formation energy (J1 only, the
 *   multi-site delta of model.hh:354-379); the composition never changes and
 *   exchange_potential is ignored.  At most 65536 sites per chain, held in LDS
 *   with the two candidate lists. */
#define MCG_MODE_EXACT 0
#define MCG_MODE_CHECKERBOARD 1
#define MCG_MODE_CANONICAL 2
/* NFOLD: the N-fold way (rejection-free kinetic Monte Carlo) of nfold.hh:78-149
 *   for the SGC square-lattice Ising model (J1), one sequential chain per
 *   wavefront.  Every event flips one site chosen with probability
 *   proportional to its Metropolis rate min(1, exp(-beta dE)); the time
 *   advances by -log1p(-u)/R (R the total rate).  The reference takes its event
 *   selector from outside the repository; the one here groups sites by rate
 *   class (10 classes, one class-partitioned site array), draws a class with
 *   one canonical word, a site in it with one Lemire word and the time with one
 *   more word.  One pass = N events.  Samples: by pass (the state after the
 *   pass, weighted by the time increment of the next event, which is selected
 *   then) or, with mcg_desc.sample_time_period > 0, on the time grid (one record
 *   per residence interval holding grid points, weighted by their number).
 *   Occupations, counts and the event sequence are bit-identical to the
 *   oracle's; times agree to the rounding of log1p.  At most 32768 sites per
 *   chain, held in LDS as the class array and the positions (4 bytes a site).
 *   On the time grid, a chain whose total rate reaches 0 (every flip's
 *   probability underflows) fails the run with MCG_ERR_STATE. */
#define MCG_MODE_NFOLD 3
/* OCCUPANT: one sequential chain per wavefront running the reference's generic
 *   occupant-event Metropolis bit for bit: the loop of
 *   occupation_metropolis.hh:87-153 over an OccLocation (OccLocation.cc:39-118,
 *   216-283) with several asymmetric units, proposing either semi-grand
 *   canonical single-site events (propose_semigrand_canonical_event,
 *   OccEventProposal.hh:261-352) or canonical pair swaps
 *   (propose_canonical_event, :109-236) from the caller's swap types,
 *   std::mt19937_64.  Square lattice; every site belongs to one asymmetric unit
 *   (mcg_set_site_asym), and a unit allows one or two occupants
 *   (mcg_desc.occupant).  Occupations are occupant indices (0 or 1).  The
 *   potential: the square-lattice formation energy (J1, model.hh:259-290) on
 *   the spins 2o - 1 of the occupant indices o, minus the exchange potential of
 *   each site's species; the delta of an event, per transformed site in event
 *   order with the earlier ones applied, is
 *     dE = (dE + (-J1 ((2 o' - 1) - (2 o - 1)) * nbr_sum)) - (mu[to] - mu[from]).
 *   One pass = the number of mutating sites (OccLocation::mol_size) steps.  At
 *   most 40000 sites per chain (lattice and lists held in LDS). */
#define MCG_MODE_OCCUPANT 4
/* KMC: kinetic Monte Carlo with atom tracking, one sequential chain per
 *   wavefront: the loop of kinetic_monte_carlo.hh:136-229 (select an event and
 *   its time increment, sample by time at every grid time the event's time
 *   reaches, apply the event through the OccLocation with atom tracking,
 *   advance the time) for a square lattice gas of at least 3x3 and at most
 *   7500 sites: occupant 0 = vacancy, 1 = atom, E = -J1 sum over nearest-
 *   neighbour bonds of n_i n_j.  The reference takes its event selector from
 *   outside the repository; the one here is rejection kinetic Monte Carlo:
 *   attempts at 1/R each (R = 4 x atoms), an atom chosen uniformly from the
 *   OccLocation's atom list (one Lemire word), a direction random_int(3)
 *   (0: +row, 1: -row, 2: +column, 3: -column), rejected when the target holds
 *   an atom, otherwise dE = J1 (a - b) (atoms next to the origin other than at
 *   the target, next to the target other than at the origin) accepted without
 *   a draw when dE <= 0, else when random_real(1.0) < exp(-beta dE); the
 *   event's time increment is fl(k / R) for the k attempts it took.  One pass
 *   = N events.  Samples: by pass (the state after the pass with the time of
 *   its last event) or, with mcg_desc.sample_time_period > 0, at the times
 *   k * period (k = 0, 1, ...), each taken before the event whose time reaches
 *   it.  Records: pass = passes (by pass) or events applied (time grid),
 *   S1 = sum over bonds of n_i n_j, M = atoms, S2 = sum over atoms of
 *   dx^2 + dy^2, S3 = sum of dx, T3 = sum of dy (displacements in lattice
 *   units since the occupation was set), time.  Occupations, lists, atoms,
 *   records, times and engine words are bit-identical to the oracle's.  A
 *   chain that makes no event in the spin limit's attempts (2^26 unless set
 *   with mcg_kmc_set_spin_limit; every hop of its state has a vanishing
 *   probability) stops, and the run fails with MCG_ERR_STATE. */
#define MCG_MODE_KMC 5

#define MCG_LATTICE_SQUARE 2 /* 2D, shape[2] == 1 */
#define MCG_LATTICE_SC 3     /* 3D simple cubic */
/* Binary FCC cluster expansion (CHECKERBOARD only; a build-defined model with
 * no counterpart among the reference's lattice models).  shape = conventional
 * cells per edge (each 3..127), 4*L0*L1*L2 sites at the doubled coordinates
 * (x, y, z), x+y+z even, of a periodic 2L0 x 2L1 x 2L2 grid; linear site
 * order l = h + L0*(y + 2*L1*z) with x = ((y+z)&1) + 2h.  Occupations are
 * 0/1.  Energy per supercell, with N sites, N1 = sum of occupations and Pk /
 * T the occupied pair (shells 1-3) and nearest-neighbour triangle counts:
 *   E = N*V0 + V1*N1 + V2/6*P1 + V3/3*P2 + V4/12*P3 + Vt/8*T
 * (orbit averages per primitive cell; per-site multiplicities 6, 3, 12, 8),
 * ECIs given in mcg_desc.eci = {V0, V1, V2, V3, V4, Vt}.  Colouring: each
 * doubled coordinate is coloured mod 4 (a seam of two sites gets colours 1,
 * 2 when 2L = 2 mod 4), at most 64 colour classes per pass. */
#define MCG_LATTICE_FCC 4

/* One chain: one supercell at one (T, mu) condition.
 * Replaces the (configuration, conditions) pair of IsingState
 * (ising_cpp/model.hh:141-157) and SemiGrandCanonicalConditions
 * (ising_cpp/basic_semigrand_canonical.hh:36-73). */
typedef struct mcg_chain_desc {
  int64_t shape[3];          /* supercell edge lengths; shape[2] == 1 in 2D  */
  double temperature;        /* K                                            */
  double exchange_potential; /* eV, conjugate to the parametric composition  */
  uint64_t seed;             /* std::mt19937_64::seed value (EXACT) or Philox key */
} mcg_chain_desc;

/* OCCUPANT system.  Replaces the Conversions / OccCandidateList / swap-list
 * set-up of a generic occupation_metropolis run (OccCandidate.cc:31-155): the
 * candidates are, in OccCandidateList order, (unit, occupant) of every unit
 * with two occupants, occupant 0 first; swap types name two candidates. */
typedef struct mcg_occupant_desc {
  int64_t n_asym;            /* asymmetric units, 1..8                            */
  const int64_t *occ_size;   /* n_asym entries: occupants each unit allows, 1 or 2 */
  const int64_t *species;    /* n_asym x 2: species index of occupant 0 and 1 of
                                each unit (the second ignored if occ_size is 1)  */
  int64_t n_species;         /* species, 1..8                                     */
  int64_t canonical;         /* 1: pair swaps (propose_canonical_event); 0: single
                                sites (propose_semigrand_canonical_event)         */
  int64_t n_swaps;           /* swap types, 1..16                                 */
  const int64_t *swaps;      /* n_swaps x 2 candidate indices (cand_a, cand_b), in
                                the order of the caller's swap list (e.g. what
                                make_semigrand_canonical_swaps builds)            */
  const double *exchange_potential; /* n_chains x n_species, eV (NULL: all 0)    */
  int64_t sites_per_unitcell; /* n_unitcells = sites / sites_per_unitcell (>= 1)  */
} mcg_occupant_desc;

/* Run description.  Replaces IsingFormationEnergy / IsingParamComposition /
 * SemiGrandCanonicalPotential construction (model.hh:164-436,
 * basic_semigrand_canonical.hh:104-199) and the beta = 1/(KB*T) set-up of
 * basic_occupation_metropolis.hh:361. */
typedef struct mcg_desc {
  int64_t mode;          /* MCG_MODE_*                                        */
  int64_t lattice;       /* MCG_LATTICE_*                                     */
  int64_t use_nlist;     /* EXACT: formation-energy variant sampled: 1 = the
                            neighbour-list sum fl(S*(-J)) (model.hh:261-270),
                            0 = the row/column fold (model.hh:271-286)        */
  double J1;             /* nearest-neighbour coupling, eV                    */
  double J2;             /* next-nearest coupling, eV (CHECKERBOARD only)     */
  double kb;             /* Boltzmann constant, eV/K                          */
  int64_t sample_period; /* passes per sample (basic_occupation_metropolis
                            sample_period)                                   */
  int64_t n_chains;
  const mcg_chain_desc *chains; /* n_chains entries */
  int64_t options;       /* bit mask of MCG_OPT_*                             */
  const double *eci;     /* MCG_LATTICE_FCC: 6 ECIs (V0, V1, V2, V3, V4, Vt), eV;
                            otherwise ignored                                 */
  double sample_time_period; /* NFOLD: > 0 samples on the time grid k*period
                                (nfold.hh:105-108 BY_TIME); 0 samples by pass   */
  const mcg_occupant_desc *occupant; /* OCCUPANT: the system; otherwise ignored */
} mcg_desc;

/* CHECKERBOARD: always use the HBM work-list kernels, even for lattices that
 * fit in LDS (the two paths produce identical trajectories). */
#define MCG_OPT_FORCE_STREAMING 1
/* EXACT: run the one-attempt-at-a-time scalar walk instead of the
 * block-speculative walk (both reproduce the reference stream bit for bit;
 * the scalar walk is the cross-check of the speculative one). */
#define MCG_OPT_EXACT_SCALAR_WALK 2
/* CHECKERBOARD: update 4 sites per thread with byte loads even where the
 * 16-site vector path applies (both give identical trajectories; this is the
 * cross-check of the vector path). */
#define MCG_OPT_CHECKER_QUAD_PATH 4
/* NFOLD, by-pass sampling: also sample the initial state (pass 0), weighted by
 * the time increment of the first event, which is selected then -- the
 * reference's sample at count 0 (nfold.hh:120-122 with a sampling fixture
 * whose first sample count is 0).  Takes effect in the first mcg_run after
 * mcg_create or mcg_assign_chain (see mcg_run for runs of 0 passes). */
#define MCG_OPT_NFOLD_SAMPLE_START 8

/* One sampled record of one chain.  Integer observables are exact; the
 * reference's FP64 sampling functions (basic_semigrand_canonical.hh:485-590)
 * are recomputed from them on the host with the reference's rounding. */
typedef struct mcg_sample {
  int64_t pass;   /* pass count at which the sample was taken                */
  int64_t S1;     /* sum over nearest-neighbour bonds of s_i s_j             */
  int64_t S2;     /* sum over next-nearest bonds of s_i s_j                  */
  int64_t M;      /* sum of s_i                                              */
  double e_fold;  /* EXACT, use_nlist == 0: the reference's row/column fold of
                     the formation energy per supercell, computed on device in
                     the reference's order; otherwise 0                      */
  double e_run;   /* EXACT: running potential energy per unit cell, updated by
                     fl(dE / n_unitcells) per accepted step
                     (occupation_metropolis.hh:141); otherwise 0             */
  int64_t S3;     /* FCC: third-shell pair count; otherwise 0                 */
  int64_t T3;     /* FCC: nearest-neighbour triangle count; otherwise 0       */
  double time;    /* NFOLD: simulated time of the sampled state (time grid: the
                     first grid point of the record); otherwise 0           */
  double weight;  /* NFOLD: sample weight (by pass: residence time of the state;
                     time grid: number of grid points); otherwise 0         */
} mcg_sample;
/* FCC records: M = N1 (occupied sites), S1 = P1, S2 = P2, S3 = P3 (pairs of
 * shells 1-3 with both sites occupied), T3 = occupied nearest-neighbour
 * triangles. */

typedef struct mcg_handle mcg_handle;

/* Create a handle: validates the description, allocates device memory for
 * every chain (lattice, RNG state, counters, sample series), fills every
 * lattice with +1 (IsingConfiguration default, model.hh:23-33) and seeds every
 * RNG.  Replaces IsingFormationEnergy::set_state (model.hh:210-256; the
 * neighbour list becomes an implicit stencil) and RandomNumberGenerator
 * construction (RandomNumberGenerator.hh:20-28). */
int mcg_create(const mcg_desc *desc, mcg_handle **out);

/* Release all device memory and the stream. */
void mcg_destroy(mcg_handle *h);

/* Message for the last error on the calling thread ("" if none). */
const char *mcg_last_error(void);

/* Set / get one chain's occupation (values +1/-1, reference linear order
 * l = i + shape0*(j + shape1*k), column-major as model.hh:82-99).
 * Replaces IsingConfiguration::set_occupation / occupation (model.hh:52-60).
 * Setting recomputes the chain's integer observables and, for EXACT, the
 * running energy from the initial per-supercell potential
 * (occupation_metropolis.hh:106).  FCC: values 0/1 in the FCC site order;
 * lattices start empty (all 0).  CANONICAL: setting also rebuilds the
 * candidate lists in site order, as OccLocation::initialize does
 * (OccLocation.cc:39-118). */
int mcg_set_occupation(mcg_handle *h, int64_t chain, const int8_t *occ, int64_t n);
int mcg_get_occupation(mcg_handle *h, int64_t chain, int8_t *occ, int64_t n);

/* CANONICAL: get / set one chain's candidate lists, n = sites entries: the
 * sites of species 0 (occupation -1) in list order, then the sites of species
 * 1.  They are OccLocation's m_loc[0] and m_loc[1] (OccLocation.hh:204-206;
 * Mol ids are site indices because every site mutates) and decide which site
 * each later choose_mol draw picks (OccLocation.hh:254-259).  Setting checks
 * that every site appears once, in the list of its species.  NFOLD (get only):
 * the site array partitioned by rate class, in class order. */
int mcg_get_occ_location(mcg_handle *h, int64_t chain, int64_t *loc, int64_t n);
int mcg_set_occ_location(mcg_handle *h, int64_t chain, const int64_t *loc, int64_t n);
/* OCCUPANT lists (get and set): n = the chain's mutating sites, the candidate
 * lists in candidate order, each list in its order, as site indices (Mol.l of
 * OccLocation's m_loc entries; the caller maps Mol ids).  KMC (get): n =
 * sites, the vacancy list then the atom list. */

/* KMC: the atoms of one chain, 4 entries each in atom order (atoms are
 * numbered in site order when the occupation is set, OccLocation.cc:68-106):
 * current site, displacement along rows and along columns since then, jumps
 * (Atom::translation and n_jumps, OccEvent.hh:13-20; OccLocation.cc:286-330).
 * n_atoms must equal the chain's atom count. */
int mcg_get_atoms(mcg_handle *h, int64_t chain, int64_t *atoms, int64_t n_atoms);
/* KMC: hop correlations collected on the device (the CorrelationsData of
 * sampling/SelectedEventFunctions.hh:37-96): from now on, every
 * jumps_per_position_sample-th jump of each atom records the atom's
 * displacement since its chain's occupation was set, the index of the event
 * within its pass, the pass and the event's time, at most
 * max_n_position_samples records per atom (later jumps are not recorded).
 * Every call starts the records afresh; setting a chain's occupation starts
 * that chain's afresh; 0 turns recording off (the records stay readable).
 * Device memory: 32 B x max_n_position_samples per lattice site. */
int mcg_kmc_collect_hops(mcg_handle *h, int64_t jumps_per_position_sample,
                         int64_t max_n_position_samples);
/* KMC: one chain's hop-correlation records: counts[a] records of atom a;
 * record k of atom a is records[4 (k n_atoms + a) + 0..3] = (displacement
 * along rows, along columns, step within the pass, pass) and
 * times[k n_atoms + a], for k < max_n_position_samples (zeros past counts[a]).
 * n_atoms must equal the chain's atom count. */
int mcg_kmc_get_hops(mcg_handle *h, int64_t chain, int64_t n_atoms, int64_t *counts,
                     int64_t *records, double *times);
/* KMC: attempts without an event after which a chain stops (default 2^26). */
int mcg_kmc_set_spin_limit(mcg_handle *h, int64_t attempts);
/* KMC and NFOLD: the reference loops' select-then-sample order
 * (kinetic_monte_carlo.hh:185-226, nfold.hh:121-141).  Every active chain
 * applies its pending event, if any, unless that event's time reaches
 * stop_time, then selects and applies events until it has applied max_events
 * in this call; an event whose time reaches stop_time is selected but left
 * pending (the chain stops there), and with select_at_end a chain that
 * applied max_events selects one more and leaves it pending.  A pending
 * event's site is read with mcg_get_nfold_state, with its time (KMC) or its
 * time increment (NFOLD; the total rate at its selection is the running
 * energy of mcg_get_position); the next mcg_run, mcg_run_steps or
 * mcg_kmc_advance applies it first (and counts it).  No records are written
 * (NFOLD: a chain whose total rate is 0 stops and the call fails with
 * MCG_ERR_STATE).  The drop-ins kinetic_monte_carlo() and nfold() sample by
 * count and by time between the selection and the application, as the
 * reference does. */
int mcg_kmc_advance(mcg_handle *h, int64_t max_events, double stop_time, int64_t select_at_end);

/* OCCUPANT: the asymmetric unit of every site of one chain (n = sites; the
 * reference's Conversions::l_to_asym).  Resets the chain's occupation to
 * occupant 0 everywhere and its lists (OccLocation::initialize); set the
 * occupation afterwards.  A new handle starts with every site in unit 0. */
int mcg_set_site_asym(mcg_handle *h, int64_t chain, const int8_t *asym, int64_t n);

/* Dump / load the RNG state of one chain.  EXACT: 313 words, the
 * std::mt19937_64 state words followed by the index of the next output (the
 * layout operator<< writes, used by the reference's engine dump/load,
 * python/src/monte.cpp:495-513).  CHECKERBOARD: 4 words (key lo, key hi,
 * pass counter, 0). */
int mcg_rng_dump(mcg_handle *h, int64_t chain, uint64_t *words, int64_t n_words);
int mcg_rng_load(mcg_handle *h, int64_t chain, const uint64_t *words, int64_t n_words);

/* Advance every active chain by n_passes passes (one pass = one attempt per
 * site: N random-site attempts in EXACT, N pair-swap attempts in CANONICAL,
 * one coloured sweep in CHECKERBOARD; mol_size() steps in OCCUPANT).
 * CANONICAL fails with MCG_ERR_STATE when an active chain has no site of one
 * species (the reference's "No events possible"); OCCUPANT fails the same way
 * when a chain's swap types have no events at the start, or stops a chain at
 * the step where they run out (the reference's loop stops there,
 * occupation_metropolis.hh:121-127) and then fails with MCG_ERR_STATE after
 * every chain has run (mcg_get_position gives where each chain stands).
 * A sample record is appended to a chain's device series whenever its pass
 * count reaches a multiple of sample_period.  Replaces the main loop of
 * basic_occupation_metropolis.hh:381-422 between two completion checks (the
 * drop-in classes call it with the number of passes to the next point at
 * which the reference's completion check can change its answer).
 * NFOLD by pass: mcg_run(h, 0) selects the next event of every active chain
 * that has none pending and appends the record of the current pass (the
 * state, the selected event's time increment as weight and the total rate),
 * as a sampled pass does; the next mcg_run applies that event first.  This is
 * the reference's select-then-sample at a count where a fixture is due
 * (nfold.hh:121-126), taken only once the caller knows the run goes on. */
int mcg_run(mcg_handle *h, int64_t n_passes);

/* EXACT, OCCUPANT and KMC: advance every active chain by exactly n_steps
 * attempts (KMC: events) from where it stands, which may end inside a pass; the next mcg_run or mcg_run_steps goes
 * on from that step (mcg_run then makes whole passes' worth of attempts, so a
 * chain keeps its offset into the pass).  Records are appended at pass ends as
 * in mcg_run.  The counterpart of the reference's per-step counting for
 * fixtures that sample by step (MonteCounter::increment_step with
 * SAMPLE_MODE::BY_STEP, run_management/SamplingFixture.hh:154-169, sampled
 * after every step by methods/occupation_metropolis.hh:150).  Other modes:
 * MCG_ERR_STATE. */
int mcg_run_steps(mcg_handle *h, int64_t n_steps);

/* Sequential modes: a chain's pass count, its steps into the current pass
 * (EXACT, OCCUPANT, KMC and NFOLD; 0 otherwise) and its FP64 running
 * potential energy per unit cell (occupation_metropolis.hh:103-106, :141;
 * EXACT, CANONICAL and OCCUPANT; NFOLD: the total rate at the selection of
 * the pending event) -- the state a sample taken between two steps reads. */
int mcg_get_position(mcg_handle *h, int64_t chain, int64_t *n_pass, int64_t *step,
                     double *e_run);

/* Mark chains active (1) or inactive (0); inactive chains are skipped by
 * mcg_run.  Used by the continuous-batching server to retire converged
 * chains while others keep running. */
int mcg_set_active(mcg_handle *h, const uint8_t *active, int64_t n_chains);

/* Re-initialise one chain slot for a new run: a new shape (at most as many
 * sites as the slot was created with, same lattice type), conditions and seed;
 * the lattice becomes all +1 and the counters, sample series and RNG restart.
 * Replaces building a new state/calculator for the next run of a condition
 * scan (basic_semigrand_canonical.hh:427-462); used by the continuous-batching
 * server to admit a request into the slot a finished request freed. */
int mcg_assign_chain(mcg_handle *h, int64_t chain, const mcg_chain_desc *desc);

/* Per-chain counters: pass count, accepted and rejected attempts
 * (BasicOccupationMetropolisData n_pass/n_accept/n_reject,
 * basic_occupation_metropolis.hh:71-82). */
int mcg_get_counters(mcg_handle *h, int64_t chain, int64_t *n_pass, int64_t *n_accept,
                     int64_t *n_reject);

/* Number of sample records held in a chain's device series. */
int mcg_n_samples(mcg_handle *h, int64_t chain, int64_t *n_samples);

/* NFOLD: a chain's simulated time and its pending event -- the site of the
 * event selected at the last by-pass sample and its time increment, which the
 * next mcg_run applies first (site -1: none pending).  The reference's loop
 * keeps these in locals of nfold.hh:83-88 (selected_event_id, time_increment)
 * between selecting an event and applying it; the drop-in nfold() reads them
 * when a completion check stops the run right after a sample, where the
 * reference applies that event before it stops.  KMC: the loop's time
 * (kinetic_monte_carlo.hh:218, the time of the last applied event), the
 * origin site of the pending event left by mcg_kmc_advance (-1: none) and, in
 * pending_dt, that event's time. */
int mcg_get_nfold_state(mcg_handle *h, int64_t chain, double *time, int64_t *pending_site,
                        double *pending_dt);

/* Copy sample records [first, first+count) of a chain to the host.
 * Replaces the Sampler storage (sampling/Sampler.hh:252-267). */
int mcg_get_samples(mcg_handle *h, int64_t chain, int64_t first, int64_t count,
                    mcg_sample *out);

/* Drop all stored sample records of every chain (device memory is kept). */
int mcg_clear_samples(mcg_handle *h);

/* Grow the sample series so every chain can hold n_records records without a
 * reallocation inside later mcg_run calls (the counterpart of the Sampler's
 * capacity increment, sampling/Sampler.hh:252-267). */
int mcg_reserve_samples(mcg_handle *h, int64_t n_records);

/* Observables computed on the device from a chain's sample series:
 *   MCG_OBS_PARAM_COMPOSITION  x  = fl(fl((N+M)/2)/N)        (model.hh:412-422)
 *   MCG_OBS_FORMATION_ENERGY   ef = fl(E_f/N)                 (model.hh:293-295)
 *   MCG_OBS_POTENTIAL_ENERGY   e  = fl(fl(E_f - fl(mu*Nx))/N) (basic_semigrand_canonical.hh:165-174)
 * where E_f = fl(-J1*S1) - ... per the model (EXACT use_nlist==0: e_fold).
 * OCCUPANT: MCG_OBS_POTENTIAL_ENERGY is the record's running potential energy
 * per unit cell (e_run; the potential has an exchange potential per species),
 * and x is the fraction of sites at occupant 1. */
#define MCG_OBS_FORMATION_ENERGY 0
#define MCG_OBS_PARAM_COMPOSITION 1
#define MCG_OBS_POTENTIAL_ENERGY 2

/* Statistics of one observable over a chain's samples [first, n_samples):
 * mean and calculated_precision as BasicStatisticsCalculator::operator()
 * (BasicStatistics.cc:114-131, with autocorrelation_factor :24-48 and
 * approx_erf_inv math.hh:62-72), computed by a device reduction. */
typedef struct mcg_stats {
  double mean;
  double calculated_precision;
  double variance;        /* CoVar0 = sum (x - mean)^2 / N                     */
  double autocorr_factor; /* (1+rho)/(1-rho), 1 if no variation, DBL_MAX if none */
  int64_t n;              /* number of observations used                        */
} mcg_stats;

/* NFOLD records are weighted: there mcg_stats_compute is mcg_stats_weighted
 * with the calculator's defaults (method 1, 10000 resamples). */
int mcg_stats_compute(mcg_handle *h, int64_t chain, int64_t observable, int64_t first,
                      double confidence, mcg_stats *out);

/* Statistics of weighted records (NFOLD mode; weight = residence time by pass,
 * number of grid points on the time grid), as BasicStatisticsCalculator's
 * weighted branch (BasicStatistics.cc:144-188): the series is read at
 * n_resamples equally spaced points of its cumulative weight (resample,
 * :50-73); method 1 gives the weighted mean and variance
 * (weighted_variance, misc/math.hh:45-54) with the autocorrelation factor of
 * the resampled series (increment W/n_resamples) and precision
 * f * sqrt(factor * variance / W); method 2 gives the plain statistics of the
 * resampled series.  Sums are reduced in another order than the reference's
 * (results agree to rounding).  Errors: MCG_ERR_STATE outside NFOLD mode,
 * MCG_ERR_ARG for another method or n_resamples outside [1, 2^22]. */
int mcg_stats_weighted(mcg_handle *h, int64_t chain, int64_t observable, int64_t first,
                       double confidence, int64_t method, int64_t n_resamples, mcg_stats *out);

/* No counterpart in the reference: mcg_stats_compute for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and results as
 * mcg_stats_compute (NFOLD records: the weighted statistics with the
 * calculator's defaults, method 1 and 10000 resamples), computed in a fixed
 * scratch array compiled into the library -- MCG_STATS_NOALLOC_MAX doubles
 * (2 MiB) on every GPU the library runs on, held until the process ends --
 * instead of the handle's growable scratch; no hipMalloc or hipFree is called.
 * Limit: at most MCG_STATS_NOALLOC_MAX - 8 observations (NFOLD: resamples);
 * a longer series fails with MCG_ERR_CAPACITY.  All calls on one GPU share
 * the array and each call leaves its contents there: calls that overlap on
 * one GPU (from other host threads, on any handle) race. */
#define MCG_STATS_NOALLOC_MAX 262144
int mcg_stats_compute_noalloc(mcg_handle *h, int64_t chain, int64_t observable, int64_t first,
                              double confidence, mcg_stats *out);

/* Equilibration check of one observable over a chain's whole series,
 * as _default_equilibration_check (src/casm/monte/checks/EquilibrationCheck.cc:50-117)
 * with absolute precision prec; the partition scan runs on the device.  NFOLD
 * records are weighted: each observation is scaled by (N/W) w_i first, W the
 * in-order sum of the weights (EquilibrationCheck.cc:134-160). */
int mcg_equilibration(mcg_handle *h, int64_t chain, int64_t observable, double prec,
                      int64_t *is_equilibrated, int64_t *n_samples_for_equilibration);

/* The completion check of many chains at once: for each listed chain, the
 * equilibration check of one observable over its whole series at absolute
 * precision prec (as mcg_equilibration), then, when it is equilibrated with
 * samples left, the statistics of the samples after the equilibration index
 * (as mcg_stats_compute with first = n_equilibration) -- the sequence of
 * CompletionCheck::_is_complete (checks/CompletionCheck.hh:272-376: the
 * equilibration check of EquilibrationCheck.cc:180-225, then the convergence
 * statistics of ConvergenceCheck.hh:139-184 on the samples after it), in one
 * launch with one workgroup per chain.  stats.n = 0 when no statistics were
 * computed.  Results equal the single-chain entry points' bit for bit.
 * Unweighted records only (NFOLD: MCG_ERR_STATE). */
typedef struct mcg_check {
  int64_t is_equilibrated;
  int64_t n_equilibration;
  mcg_stats stats;
} mcg_check;
int mcg_check_batch(mcg_handle *h, int64_t n, const int64_t *chains, int64_t observable,
                    double prec, double confidence, mcg_check *out);

/* Launch schedule of a CHECKERBOARD handle (EXACT handles have nothing to
 * tune: one wavefront per chain).  Every schedule produces the same
 * trajectories; only the speed changes. */
typedef struct mcg_schedule {
  int64_t lds_block;        /* LDS-resident and FCC kernels: threads per workgroup
                               (256, 512, 1024)                                         */
  int64_t units_per_thread; /* HBM kernels: work units per thread and colour (1, 2, 4)    */
  int64_t graph_passes;     /* HBM path: passes captured per hipGraph replay (1 = none)   */
  double trial_ms;          /* mcg_autotune: time of the chosen schedule on its trial     */
  int64_t band_lines;       /* 2D HBM path: lines per band of the two-launch band kernels
                               (-1 = chosen from the sizes, 0 = one launch per colour)  */
  int64_t line_tiles;       /* 4/8-colour HBM path (not banded): 1 = stage tiles of class
                               lines with their neighbour lines in LDS, 0 = read lines
                               from HBM per class line.  FCC chains of lines of at most
                               32 sites: 0 = class lines with bit-sliced counts, one to
                               four threads per line (at least as many threads per
                               workgroup as the most lines of a class, lds_block if
                               larger, at most 512), 1 = 4-site units at lds_block  */
} mcg_schedule;

/* Autotuner: times the candidate schedules on `trial_passes` passes of this
 * handle's own chains (from a device snapshot that is restored afterwards, so
 * the run's state and trajectory are unchanged), keeps the fastest and caches
 * it per (lattice, colouring, size class) for the process and, when the
 * environment variable MCG_TUNE_CACHE names a file, across processes.
 * No counterpart in the reference (its loop has no launch parameters). */
int mcg_autotune(mcg_handle *h, int64_t trial_passes, mcg_schedule *chosen);
int mcg_get_schedule(mcg_handle *h, mcg_schedule *out);
int mcg_set_schedule(mcg_handle *h, const mcg_schedule *s);

/* Time of the last mcg_run's kernels on the handle's stream, in ms
 * (HIP events around the launches), and the number of site attempts done. */
int mcg_last_run_time(mcg_handle *h, double *ms, double *attempts);

/* FCC: the class schedule a chain of L0 x L1 x L2 conventional cells sweeps
 * each pass (no counterpart in the reference, whose models have no FCC
 * checkerboard; a host computation, no device needed): the non-empty colour
 * classes (col = cx + 6 cy + 36 cz; colours 4 and 5 are the two seam points
 * of a ring of odd L) grouped into steps of classes no two of which interact,
 * updated together.  n_steps; start[0..n_steps] (129 entries suffice): step s
 * holds cls[start[s] .. start[s+1]); cls: 128 entries. */
int mcg_fcc_class_schedule(int64_t L0, int64_t L1, int64_t L2, int64_t *n_steps,
                           int64_t *start, int64_t *cls);

#ifdef __cplusplus
}
#endif

#endif /* CASM_MI355X_MCGPU_H */
