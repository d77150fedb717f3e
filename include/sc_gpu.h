/*
 * sc_gpu.h -- C ABI of the MI355X (gfx950) Monte Carlo hot path for patchy
 * spherocylinders, a drop-in for the energy / move interface of
 * robertvacha/SC (scOOP 4.0).
 *
 * The reference's plug point is the compile-time typedef
 *   typedef TotalEMatrix<PairE> TotalEnergyCalculator;
 * (scOOP/mc/totalenergycalculator.h:1227) whose methods are called by
 * Updater (updater.h:33), MoveCreator (movecreator.h:36) and ClusterSampler
 * (clust.h:23).  Each entry point below names the reference method it
 * replaces.  A C++ class with the reference's method set (GpuTotalE in
 * include/sc_gpu_totale.hpp) wraps this ABI for in-process C++ callers.
 *
 * Conventions:
 *   - every function returns 0 on success and a negative SC_E* code on
 *     failure; sc_last_error() gives the message.  Nothing aborts;
 *   - the caller owns every host array, the context owns all device memory
 *     and one HIP stream;
 *   - energies cross the ABI in FP64 whatever the internal precision;
 *   - positions cross the ABI as box fractions (as Particle::pos,
 *     scOOP/structures/particle.h:26); inside they are 32-bit fixed point;
 *   - a context is not thread safe: one context per host thread / replica;
 *   - particle indices are the reference's indices (order of config.init).
 */
#ifndef SC_GPU_H
#define SC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SC_OK 0
#define SC_EINVAL (-1)      /* bad argument or malformed input text */
#define SC_EHIP (-2)        /* HIP runtime error */
#define SC_ESTATE (-3)      /* call out of order (e.g. no topology yet) */
#define SC_ECAPACITY (-4)   /* a fixed device capacity was exceeded */
#define SC_ENODEV (-5)      /* no gfx950 device / HIP runtime unavailable */

typedef struct sc_ctx sc_ctx;

typedef struct {
    uint64_t trans_acc, trans_try, rot_acc, rot_try;   /* Statistics trans/rot acc/rej */
    uint64_t left_cell;       /* translations rejected for leaving their cell */
    uint64_t stage1_acc, stage1_try;                    /* delayed acceptance: screen on the exact cheap terms */
    uint64_t stage2_acc, stage2_try;                    /* delayed acceptance: network screen (trials it ran for) */
    double energy_change;     /* sum of accepted dE (for the drift check, updater.cpp:122-150) */
    uint64_t bond_rejected;   /* moves rejected because a bonded chain partner lay outside the stencil */
    /* whole-molecule moves (sc_set_chain_moves): Statistics chainm / chainr acc and tries */
    uint64_t chain_disp_acc, chain_disp_try, chain_rot_acc, chain_rot_try;
    uint64_t chain_rejected;  /* of those tries: molecule wider than the extent, or anchor leaving its coarse cell */
    uint64_t chain_cells_skipped;   /* coarse cells whose neighbourhood outgrew the LDS stage: no move that pass */
} sc_sweep_stats;

/* --- lifetime --------------------------------------------------------- */
int sc_create(sc_ctx** out, int device);
void sc_destroy(sc_ctx* ctx);
const char* sc_last_error(const sc_ctx* ctx);

/* --- inputs in SC's own text formats ------------------------------------ */
/* top.init text: [Types] [Molecules] [System] [EXCLUDE] sections
 * (Inicializer::readTopoFile, scOOP/mc/inicializer.cpp:453-585, fillTypes
 * :660-852, fillMol :929-1108, fillSystem :631-657) followed by
 * Topo::genParamPairs / genTopoParams (scOOP/structures/topo.cpp:5-153). */
int sc_load_topology(sc_ctx* ctx, const char* top_init_text);
/* options text, key = value (Sim::readOptions, scOOP/structures/sim.h:123-406).
 * Stored values keep the reference's post-processing: transmx doubled
 * (sim.h:365), rotmx turned into the quaternion half-angle (sim.h:360). */
int sc_load_options(sc_ctx* ctx, const char* options_text);
int sc_get_option(const sc_ctx* ctx, const char* key, double* value);
/* config.init text (Inicializer::initConfig, inicializer.cpp:89-298) for the
 * particles the topology's [System] declares. */
int sc_load_config(sc_ctx* ctx, const char* config_init_text);
/* Conf::draw-format text of the current state (structures/Conf.h:388-425);
 * returns the length written (or needed, when buf is NULL / cap too small). */
long sc_write_config(sc_ctx* ctx, char* buf, long cap);

/* The run's report files, text per frame (the caller appends it to the file);
 * each returns the length written (or needed, when buf is NULL / cap too small):
 *   movie frame        Updater::dumpMovie (mc/updater.h:82-93): "N\nsweep s; box x y z\n" + Conf::draw
 *   stat.dat line      " s; volume\n"                      (mc/updater.cpp:322)
 *   energy.dat line    " s; allToAll  alignmentOrder \n"   (mc/updater.cpp:323, :466-482) */
long sc_write_movie_frame(sc_ctx* ctx, long sweep, char* buf, long cap);
long sc_write_stat_line(sc_ctx* ctx, long sweep, char* buf, long cap);
long sc_write_energy_line(sc_ctx* ctx, long sweep, char* buf, long cap);
/* Updater::alignmentOrder (mc/updater.cpp:466-482): sum of dir_i . dir_j over
 * pairs with centre distance < 1.5 */
int sc_alignment_order(sc_ctx* ctx, double* out);
/* ClusterSampler::writeCluster without decoration (mc/clust.cpp:7-49): the
 * frames of cluster_stat.dat (size histogram) and cluster.dat (clusters in
 * order of their smallest member, with their energy = sum of p2p over their
 * pairs, clust.h:100-111, and 1-based members); lengths as above. */
int sc_write_clusters(sc_ctx* ctx, double threshold, long sweep, char* stat_buf, long stat_cap, char* cl_buf,
                      long cl_cap, long* stat_len, long* cl_len);

/* --- topology queries ---------------------------------------------------- */
int sc_num_particles(const sc_ctx* ctx);
int sc_particle_types(const sc_ctx* ctx, int32_t* types, int cap);
/* maxcut = 1.1 * (max cutoff + max length), Topo::genTopoParams (topo.cpp:140-153) */
int sc_maxcut(const sc_ctx* ctx, double* maxcut);
/* FP64 pair-parameter table row for reference types (ti,tj): 44 doubles in
 * the order documented in sc_host.cpp (ia_params[ti][tj], structures.h:187-222) */
int sc_pair_params(const sc_ctx* ctx, int ti, int tj, double* out44);

/* --- raw state (the counterpart of Conf::pvec) --------------------------- */
/* pos: box fractions n*3; dir, pd0: n*3 (normalised / orthogonalised here as
 * initConfig does); types: reference type numbers. */
int sc_set_config(sc_ctx* ctx, int n, const double box[3], const double* pos, const double* dir,
                  const double* pd0, const int32_t* types);
int sc_get_config(sc_ctx* ctx, double box[3], double* pos, double* dir, double* pd0);
/* fixed-point positions (uint32 fractions of the box), n*3 */
int sc_get_fixed_positions(sc_ctx* ctx, uint32_t* u);

/* --- cell list ------------------------------------------------------------ */
/* Builds the cell list for a grid of cells at least maxcut/r wide (r = 1 or 2:
 * stencil radius), each axis count a multiple of p = r+1, shifted by
 * shift[3] (fixed-point units).  Optional outputs: cell id per particle
 * (x fastest), CSR starts (ncell+1), particle indices sorted by cell and, in a
 * cell, by index; ncell_out[3] the grid. */
int sc_build_cells(sc_ctx* ctx, int r, const uint32_t shift[3], int32_t* cell_of, int32_t* cell_start,
                   int32_t* sorted, int32_t ncell_out[3]);
/* Neighbour pairs (i<j) with minimum-image centre distance <= rc (FP64 from
 * the fixed-point positions, no FMA contraction); the GPU counterpart of
 * genSimplePairList's distance test (scOOP/mc/updater.cpp:484-552),
 * enumerated on the device over a cell list.  Pairs in row-major order. */
int sc_neighbor_pairs(sc_ctx* ctx, double rc, int32_t* pairs, long cap, long* count);

/* --- energies ------------------------------------------------------------- */
/* pair energies E(i,j) with i as part1 (TotalE::p2p, totalenergycalculator.h:200);
 * fp64 != 0 evaluates the same device code in FP64 (cross-check mode). */
int sc_pair_energies(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, double* out, int fp64);
/* The same pair energies through an entry point that allocates nothing (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops).  Its staging
 * (indices and energies of up to SC_NOALLOC_MAX_PAIRS pairs, 1 MiB) is one
 * fixed device array per GPU that lives until the process ends; more pairs
 * return SC_ECAPACITY.  Results equal sc_pair_energies' bit for bit.  Calls
 * that overlap on one GPU share that array and race. */
#define SC_NOALLOC_MAX_PAIRS 65536
int sc_pair_energies_noalloc(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, double* out, int fp64);
/* whole-system energy: FP32 pair terms, FP64 reduction
 * (TotalEMatrix::allToAll / initEM, totalenergycalculator.h:314-353,502-521) */
int sc_energy_all(sc_ctx* ctx, double* e);
/* energy of one particle with all others (TotalEFull::oneToAll, :563-583;
 * TotalEMatrix::oneToAll :355-381 returns the same value from its cache) */
int sc_energy_one(sc_ctx* ctx, int target, double* e);
/* energy of particle `target` placed at a trial state (TotalEMatrix::oneToAllTrial,
 * :383-415); pos is a box fraction */
int sc_energy_one_trial(sc_ctx* ctx, int target, const double pos[3], const double dir[3], const double pd0[3],
                        double* e);
/* commit a particle trial (TotalEMatrix::update(int), :326-328) */
int sc_commit_particle(sc_ctx* ctx, int target, const double pos[3], const double dir[3], const double pd0[3]);
/* non-bonded energy of the molecule [first, first + size) with every particle
 * outside it, part1 = the molecule's particle (TotalEMatrix::mol2others,
 * totalenergycalculator.h:417-453) */
int sc_energy_mol(sc_ctx* ctx, int first, int size, double* e);
/* the same with the molecule's particles at trial states, size x 3 values
 * each, pos as box fractions (TotalEMatrix::mol2othersTrial, :455-500) */
int sc_energy_mol_trial(sc_ctx* ctx, int first, int size, const double* pos, const double* dir, const double* pd0,
                        double* e);
/* commit a molecule trial (TotalEMatrix::update(Molecule), :330-332) */
int sc_commit_mol(sc_ctx* ctx, int first, int size, const double* pos, const double* dir, const double* pd0);
/* every particle's energy with all others, original order */
int sc_energy_per_particle(sc_ctx* ctx, double* out);
/* whole-system energy at a trial box (TotalEMatrix::allToAllTrial, :334-336) */
int sc_energy_all_trial(sc_ctx* ctx, const double box_new[3], double* e);
/* commit a box change (TotalEMatrix::update(), :318-320) */
int sc_commit_box(sc_ctx* ctx, const double box_new[3]);

/* --- hard-core overlap ---------------------------------------------------- */
/* overlap(i,j) <=> closest segment (or sphere) distance < sigma_ij.  FP32,
 * with the distance recomputed in FP64 when the FP32 one lies within 1e-5
 * sigma of contact (ten times the FP32 error at stencil distances), so flags
 * match an FP64 evaluation except within FP64 rounding of contact.
 * out[k] = 1 on overlap; dmin[k] (optional) = the distance.  Build-defined:
 * the reference's Conf::overlap (structures/Conf.cpp:104-239) is not called
 * anywhere in scOOP. */
int sc_overlap_flags(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, uint8_t* out, double* dmin);
/* count of overlapping pairs in the whole system */
int sc_overlap_count(sc_ctx* ctx, long* count);

/* --- Monte Carlo ---------------------------------------------------------- */
/* nsweeps checkerboard sweeps (N trial moves each, the step loop of
 * Updater::simulate, updater.cpp:206-230, with MoveCreator::particleMove,
 * movecreator.cpp:11-33) at temperature temper.  transmx is the translation
 * length (already doubled, sim.h:365), rot_half the maximum quaternion
 * half-angle in radians (sim.h:360).  r selects the checkerboard (1: 2x2x2
 * colours, 2: 3x3x3).  seed/counter key the counter-based RNG; the same
 * (seed, counter) reproduces a run bit for bit.  stats may be NULL. */
int sc_sweep(sc_ctx* ctx, int nsweeps, double temper, double transmx, double rot_half, int r, uint64_t seed,
             uint64_t counter, sc_sweep_stats* stats);   /* r = 0: the autotuned radius (default 2) */
/* How many trial moves an active cell makes in one checkerboard pass.
 * mode 0 (default): as many as the cell holds particles, so every particle is
 * chosen once per sweep on average as in Updater::simulate's step loop
 * (updater.cpp:206-230, target = ran2()*N).  mode 1: the mean occupancy of the
 * non-empty cells, randomly rounded, so a sweep still averages N trials but no
 * pass waits for its fullest cell.  Both are Metropolis chains with the same
 * stationary distribution (the count never depends on the moves of the pass). */
int sc_set_trial_mode(sc_ctx* ctx, int mode);
/* Waves that work together on one trial of an active cell: 1, 2, 4, or 0
 * (default: four waves, each evaluating a trial of its own; see
 * sc_set_speculation).  With 1, 2 or 4 the stencil scan, the pair lists and
 * the patch terms of every trial are spread over the block's threads and the
 * chain of trials runs one trial at a time.  Results are deterministic for
 * each setting; they differ between settings only by the order of FP64 sums. */
int sc_set_block_waves(sc_ctx* ctx, int nw);
/* Speculative trials in sc_sweep (on = 1, the default, with block waves 0):
 * each of the four waves of an active cell's block evaluates one of the next
 * four trials against the cell's current state; the trials are then kept in
 * order until one depends on an earlier accepted trial of the round (the same
 * particle, the two within maxcut in any of their old/new states, or chain
 * partners), and evaluation resumes from there.  The chain of states and
 * decisions is the one-trial-at-a-time sweep's (Updater::simulate's sequence,
 * updater.cpp:206-230, within each cell).  A trial whose moving particle has
 * more than 320 particles within maxcut, or more than 192 (pair, state)
 * entries within the attraction's range, fails the sweep with SC_ECAPACITY;
 * on = 0 evaluates such systems.  No counterpart in the reference (its loop is
 * sequential). */
int sc_set_speculation(sc_ctx* ctx, int on);
/* The sweep keeps the derived vectors (patch sides, second patch, chiral
 * axes: Particle::init, particle.cpp:3-79) of a pass's home-cell particles in
 * LDS, computed once per pass and refreshed on acceptance, for cells of up to
 * `cap` particles (default and maximum 64); larger cells, or cap = 0, derive
 * them at every trial.  Both give the same trajectories; the setting exists
 * so that tests can compare them. */
int sc_set_home_cache(sc_ctx* ctx, int cap);
/* Colour passes per sweep of sc_sweep: a sweep visits the first 1/v of its
 * randomly ordered colours (v must divide their number: 8 for 2x2x2), and
 * an active cell makes v times the trials of sc_set_trial_mode per visit,
 * so a sweep still averages N trial moves (the step loop of
 * Updater::simulate, updater.cpp:206-230, picks its particles at random; a
 * sweep need not touch every particle).  Longer chains per pass average the
 * cells' uneven trial costs, so the slowest cell of a pass lags the mean
 * less, and a sweep launches fewer passes.  Default 1.  Replica batches
 * (sc_batch_sweep) use v for every replica whose number of colours it
 * divides and 1 for the others.  Every pass is a product of independent
 * Metropolis chains, whatever v. */
int sc_set_visit_factor(sc_ctx* ctx, int v);
/* Largest number of active cells of one colour in sc_sweep's grid: when the
 * grid of cells >= maxcut/r would give more, it takes coarser cells (fewer
 * along its longest axis) until a colour has at most n, so that a pass runs
 * all its cells at once (the sweep kernel runs two blocks per CU).  n = 0: no
 * limit; n < 0: the default, two per CU of the device.  Coarser cells keep
 * every property of the checkerboard (cells are still >= maxcut/r wide). */
int sc_set_max_active_cells(sc_ctx* ctx, int n);
/* Whole-molecule moves in sc_sweep (MoveCreator::chainMove, movecreator.cpp:
 * 305-328): besides its N particle trials, every sweep makes chainprob * N
 * molecule trials (chainprob is the per-step probability of Updater's step
 * loop, updater.cpp:206-217), each on a molecule of more than one particle
 * chosen uniformly, displaced by exactly chain_mx in a random direction
 * (chainDisplace, :1075-1137; chain_mx is the doubled chainmmx, sim.h:366)
 * or rotated about its centre by the quaternion half-angle chain_angle * U
 * (chainRotate / clusterRotate, :1139-1305; sim.h:362), half and half,
 * accepted by moveTry on mol2others.  The centre is weighted by
 * Ia_param::volume (clusterCM, :1307-1324) and taken on minimum-image
 * unwrapped coordinates.  Moves run on a coarse 2x2x2 checkerboard of cells
 * at least 2 extent + maxcut wide; molecules wider than `extent` (0: 1.2 x
 * the widest molecule now) do not move, nor does a move leave the anchor's
 * coarse cell (the extent is fixed at this call; call again after loading
 * another configuration).  Box edges must be at least 2 extent + maxcut and above 4
 * extent (no molecule reaches its own image, which the reference's
 * mol2others has no term for, and the minimum image unwraps every molecule).
 * Molecules of at most 64 particles.  chainprob = 0 turns the moves off. */
int sc_set_chain_moves(sc_ctx* ctx, double chainprob, double chain_mx, double chain_angle, double extent);
/* the extent the chain moves use */
int sc_get_chain_extent(sc_ctx* ctx, double* extent);
/* how the last sweep ran its chain moves: the coarse colour period (2 or 3,
 * 0 before any) and whether the trials read an LDS copy of each cell's
 * neighbourhood (1) or global memory (0) */
int sc_chain_pass_info(const sc_ctx* ctx, int* period, int* staged);
/* Visit factor of the chain-move passes (1..27, default 1): a sweep's chain
 * moves visit the first 1/v' of the coarse colours in the sweep's random
 * order, each with v' times the molecule trials, where v' is the largest
 * divisor of the colour count (8 or 27, by the box) not above v; chainprob * N
 * trials per sweep on average either way.  A move never leaves its anchor's
 * coarse cell, so every pass is a product of independent Metropolis chains;
 * fewer passes pay the neighbourhood staging and the cell-list rebuild fewer
 * times.  No counterpart in the reference (its chain moves are sequential). */
int sc_set_chain_visit_factor(sc_ctx* ctx, int v);
/* the chain visit factor v' the last sweep used (0 before any chain pass) */
int sc_chain_visit_info(const sc_ctx* ctx, int* visit);
/* Volume move (MoveCreator::pressureMove, movecreator.cpp:330-550) for
 * ptype 0 (one random edge) and 1 (isotropic); returns accepted=1/0 and the
 * new box.  Whole-system energies are FP32 pair terms (FP64 bonds) with FP64
 * reduction; dE is FP64.  E_old is the energy carried through the sweeps
 * since the last full evaluation (the reference's allToAll sums its cached
 * pair energies likewise), re-evaluated every 16 volume moves and after any
 * other change of the state; E_new is always evaluated in full. */
int sc_volume_move(sc_ctx* ctx, int ptype, double press, double temper, double edge_mx, uint64_t seed,
                   uint64_t counter, int* accepted, double box_out[3]);
/* The whole-system energy the volume moves take as E_old: the last full
 * evaluation plus the energy changes the sweeps accepted since (the role of
 * the reference's cached pair energies in allToAll, totalenergycalculator.h:
 * 338-353).  age: volume moves since that evaluation (a new one every 16).
 * SC_ESTATE when no volume move has run since the state last changed. */
int sc_carried_energy(sc_ctx* ctx, double* e, int* age);

/* --- learned surrogate and delayed acceptance ------------------------------ */
/* No counterpart in the reference, which evaluates every trial move with the
 * exact pair functors (paire.h:466-827, paire.cpp:244-301).  A small network
 * (12 body-frame features -> 64 -> 64 -> 64 -> 1, SiLU) stands in for the
 * patch attraction of one spherocylinder type with itself; the sweep then
 * screens each move with exact repulsion + network attraction and evaluates
 * the exact energy only for moves the screen accepts (two-stage delayed
 * acceptance, Christen & Fox 2005), which leaves the stationary distribution
 * exactly the reference's.
 *
 * Flat FP32 weight vector (SC_MLP_NPARAM floats), row-major [in][out]:
 *   W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4.
 * Features of a pair (moving particle a as part1, neighbour b; v = pos_b -
 * pos_a minimum image; e1 = dir_a, e2 = patchdir0_a, e3 = e1 x e2;
 * lsc = half_len_a + half_len_b + rcut):
 *   v.e1/lsc, v.e2/lsc, v.e3/lsc, dir_b.(e1,e2,e3), patchdir0_b.(e1,e2,e3),
 *   d_seg/rcut, |v|/lsc, w(d_seg) (the cos^2 switch of the attraction). */
#define SC_MLP_NPARAM 9217
#define SC_MLP_NF 12
/* Allocate the network for reference type `type` (same type both sides) with
 * seeded random weights and reset the optimiser state. */
int sc_mlp_init(sc_ctx* ctx, int type, uint64_t seed);
int sc_mlp_set_weights(sc_ctx* ctx, const float* w, int n);
int sc_mlp_get_weights(sc_ctx* ctx, float* w, int n);
/* Generate ns training pairs on the device: uniform orientations, centre
 * offsets with closest segment distance <= dmax (a fraction near_frac within
 * 1.1 rcut); target = exact FP32 patch attraction (0 beyond rcut). */
int sc_mlp_generate(sc_ctx* ctx, long ns, uint64_t seed, double dmax, double near_frac);
/* copy rows [first, first+count) of the training set: features (count*12),
 * target attraction and full exact pair energy; any pointer may be NULL */
int sc_mlp_training_data(sc_ctx* ctx, long first, long count, float* feats, float* target, float* efull);
/* Adam on mean squared error, FP32 forward/backward kernels, cosine learning
 * rate from lr to lr_final; the last val_frac of the set is held out and its
 * RMSE (BF16 MFMA inference) returned. */
int sc_mlp_train(sc_ctx* ctx, int epochs, int batch, double lr, double lr_final, uint64_t seed, double val_frac,
                 double* train_loss, double* val_rmse);
/* Loss (mean squared error) and its gradient w.r.t. every weight for one batch
 * (x: B*12, y: B) with the current weights, by the training kernels; grads
 * holds ngrads = SC_MLP_NPARAM floats. */
int sc_mlp_gradients(sc_ctx* ctx, int batch, const float* x, const float* y, float* grads, int ngrads, double* loss);
/* SC_MLP_NPARAM as compiled into the library (bindings check it at load). */
int sc_mlp_nparam(void);
/* Network output for m feature rows: bf16 = 1 the MFMA tile the sweep uses,
 * bf16 = 0 FP32 throughout. */
int sc_mlp_forward(sc_ctx* ctx, int m, const float* x, float* out, int bf16);
/* Features and exact FP32 patch attraction of pairs (i[k] as part1, j[k]) of
 * the loaded configuration. */
int sc_mlp_pair_features(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, float* feats, float* eattr);
/* Acceptance scheme of sc_sweep.  0: exact Metropolis.  1: delayed
 * acceptance in three stages: a Metropolis screen on the exact cheap terms of
 * the energy change (segment-distance repulsion, sphere pairs), a screen on
 * the network's patch attraction (a one-type system of the network's type),
 * then the exact energy, which accepts with min(1, exp(-(dE - screened)/T)).
 * 2: the same without the network screen.  All three sample the reference's
 * Boltzmann distribution; sc_sweep_stats.stage1_* / stage2_* count the
 * screens. */
int sc_set_surrogate(sc_ctx* ctx, int mode);
/* The network screen of mode 1 runs for trials with at least m (pair, state)
 * entries within the attraction's range (both states counted, so a move and
 * its reverse are screened alike and the scheme stays exact); the others go
 * from the first screen to the exact energy.  m < 0 (default): the entries
 * one wave's round of patch-pair items holds, plus one (64 / (2 x patch pairs
 * per pair) + 1: 9 for two-patch particles), so the network stands in for
 * phase C where that takes more than one round.  m = 0: every trial.  No
 * counterpart in the reference. */
int sc_set_network_min_entries(sc_ctx* ctx, int m);

/* --- replica batches (parallel tempering) ------------------------------------ */
/* The replicas of MoveCreator::replicaExchangeMove (movecreator.cpp:555-760),
 * which the reference runs one per MPI process, held in one context: load
 * their particles back to back with sc_set_config (positions as fractions of
 * each replica's own box), then declare the cut.  Replica r gets its own cell
 * grid; one checkerboard pass of every replica is one kernel launch.  A new
 * sc_set_config ends the batch; single-system calls refuse a batch. */
int sc_batch_set(sc_ctx* ctx, int nrep, const int32_t* n_per_rep, const double* boxes /* nrep x 3 */);
/* nsweeps sweeps of every replica r at temperature temper[r] (r = 0: the
 * autotuned stencil radius); stats has nrep entries (may be NULL).  Replica r
 * draws from seed sc_replica_seed(seed, r), so a replica swept alone with that
 * seed follows the same trajectory. */
int sc_batch_sweep(sc_ctx* ctx, int nsweeps, const double* temper, double transmx, double rot_half, int r,
                   uint64_t seed, uint64_t counter, sc_sweep_stats* stats);
uint64_t sc_replica_seed(uint64_t seed, int r);
/* whole-system energy of every replica (FP32 pair terms, FP64 sums) */
int sc_batch_energies(sc_ctx* ctx, double* e);
/* One replica-exchange step: replicas ordered by temperature, neighbours
 * (k, k+1) with k % 2 == parity swap temperatures when
 * change = (1/T_k - 1/T_k+1)(E_k - E_k+1) > 0 or U < exp(change)
 * (movecreator.cpp:695-713, canonical ensemble).  temper is updated in place;
 * accepted[k] (nrep - 1 entries, may be NULL) marks the swapped ladder pairs. */
int sc_batch_exchange(sc_ctx* ctx, double* temper, const double* energy, int parity, uint64_t seed, uint64_t counter,
                      int32_t* accepted);
int sc_batch_get_config(sc_ctx* ctx, int r, double* pos, double* dir, double* pd0);
/* Autotuner: times `sweeps` sweeps of the loaded system or batch in each of
 * stencil radius 1 / 2 x trial mode 0 / 1 (state restored after each), keeps
 * the fastest for later calls with r = 0; rates[4] (may be NULL) receives
 * sweeps/s in the order (1,0), (1,1), (2,0), (2,1).  temper: one temperature
 * for a single system, one per replica for a batch.  No counterpart in the
 * reference. */
int sc_autotune(sc_ctx* ctx, int sweeps, const double* temper, double transmx, double rot_half, uint64_t seed,
                double* rates, int* r_best, int* mode_best);

/* --- observables ------------------------------------------------------------ */
/* Cluster analysis of one-particle molecules (ClusterSampler::sameCluster /
 * genClusterList, clust.cpp:51-163): particles i, j share a cluster when
 * p2p(i, j) <= threshold (the reference uses -0.10); labels[i] = the smallest
 * index of i's cluster (NULL allowed), nclusters = their number.  FP64 pair
 * energies over the device pair list within maxcut, connected components on
 * the device. */
int sc_cluster_labels(sc_ctx* ctx, double threshold, int32_t* labels, int* nclusters);
/* Nematic order of the particle axes: S2 = -2 x the middle eigenvalue of
 * Q = (3<u u> - I)/2 (Tests/sc35.c:8563-8603); director = the eigenvector of
 * the largest eigenvalue (may be NULL).  Q is reduced on the device in FP64. */
int sc_nematic_order(sc_ctx* ctx, double* s2, double director[3]);

/* --- synthetic inputs ----------------------------------------------------- */
/* Seeded overlap-free configuration by random sequential addition (host,
 * FP64).  No counterpart in the scOOP library: the reference's generator
 * scripts/config-scripts/psc-genconfig.py:16-29 places particles without an
 * overlap check.  Outputs box-fraction positions and unit vectors (n*3). */
int sc_generate_rsa(sc_ctx* ctx, int n, const int32_t* types, const double box[3], uint64_t seed, long max_attempts,
                    double* pos, double* dir, double* pd0);

/* --- device timing -------------------------------------------------------- */
/* With profiling on, sc_sweep brackets every checkerboard-pass launch with
 * HIP events on the context's stream (no counterpart in the reference, whose
 * only timer is clock() around the sweep loop, updater.cpp:137-149). */
int sc_set_profiling(sc_ctx* ctx, int on);
/* A plain device copy of `bytes` (read + write), `reps` times: the known-size
 * kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE counts.  No
 * counterpart in the reference. */
int sc_copy_probe(sc_ctx* ctx, long bytes, int reps);
/* milliseconds of the last sc_sweep / sc_energy_all call measured with HIP
 * events on the context's stream around its kernels, the number of launches
 * of the dominant kernel in it and (with profiling on) their summed time */
int sc_last_timing(const sc_ctx* ctx, double* ms, long* launches, double* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif
