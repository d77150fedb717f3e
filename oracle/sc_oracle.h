/*
 * sc_oracle.h -- FP64 CPU oracle for the patchy-spherocylinder Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of robertvacha/SC
 * (scOOP 4.0) energy routines, used as the checker for the HIP product path
 * and as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never links it.
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed in the header of sc_oracle.c.
 */
#ifndef SC_ORACLE_H
#define SC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXT 40          /* scOOP/structures/macros.h:87 MAXT */
/* geometric types, scOOP/structures/macros.h:71-85 */
#define ORC_SCN 10
#define ORC_SCA 11
#define ORC_PSC 12
#define ORC_CPSC 13
#define ORC_CHPSC 14
#define ORC_CHCPSC 15
#define ORC_TPSC 16
#define ORC_TCPSC 17
#define ORC_TCHPSC 18
#define ORC_TCHCPSC 19
#define ORC_SP 30
#define ORC_SPN 30
#define ORC_SPA 31

typedef struct { double x, y, z; } orc_vec;

/* Pair parameters, field-for-field the quantities of Ia_param
 * (scOOP/structures/structures.h:187-222) that the energy path reads. */
typedef struct {
    int geotype[2];
    double sigma, epsilon, A, B;
    double pdis, pswitch, pswitchINV;
    double pangl[4], panglsw[4], pcangl[4], pcanglsw[4];
    double rcut, rcutSq, rcutwca, rcutwcaSq;
    double pcoshalfi[4], psinhalfi[4];
    double csecpatchrot[2], ssecpatchrot[2];
    double len[2], half_len[2];
    double chiral_cos[2], chiral_sin[2];
    double parallel;
    double volume;
    int exclude;
} orc_ia;

#define ORC_MAXMT 100        /* scOOP/structures/macros.h:88 MAXMT */

/* Bond and angle parameters of one molecule type, the fields of
 * MoleculeParams (scOOP/structures/moleculeparams.h:16-32) that the energy
 * reads; -1 marks an absent term (the constructor's default, :42-43).
 * Angles in radians. */
typedef struct {
    int size;              /* particles per molecule (molSize) */
    double bond1c, bond1eq, bond2c, bond2eq, bonddc, bonddeq, bondhc, bondheq;
    double angle1c, angle1eq, angle2c, angle2eq;
} orc_mol;

typedef struct {
    orc_ia ia[ORC_MAXT][ORC_MAXT];
    double sqmaxcut;   /* squared after finalize, as Topo::sqmaxcut */
    double maxcut;
    orc_mol mol[ORC_MAXMT];
    int nmol;
    /* per particle (index order): molecule type, position in its molecule.
     * Caller-owned; NULL or nchain == 0 means every particle is atomic. */
    int nchain;
    const int* chain;
} orc_topo;

typedef struct {
    orc_vec pos;          /* box fractions, as Particle::pos */
    orc_vec dir;
    orc_vec patchdir[2];
    orc_vec patchsides[4];
    orc_vec chdir[2];
    int type;
    int idx;              /* index in the configuration (identity for the connectivity list) */
    int moltype, molpos;  /* molecule type and position in the molecule; moltype -1: atomic */
} orc_part;

/* ---- topology ---- */
void orc_topo_clear(orc_topo* t);
/* One [Types] line: geotype code and the numeric parameters after GEOTYPE
 * (EPSILON SIGMA ...); returns 1 on success, 0 on an arity error. */
int orc_topo_add_type(orc_topo* t, int type, int geotype, const double* params, int nparams);
/* exclusions: ORC_MAXT*ORC_MAXT bytes (row-major) or NULL */
void orc_topo_finalize(orc_topo* t, const unsigned char* exclusions);
const orc_ia* orc_topo_ia(const orc_topo* t, int i, int j);

/* ---- molecules ----
 * m: molecule type; p: size then the 12 parameters in orc_mol order.  Returns
 * 1, or 0 if m is out of range. */
int orc_topo_set_mol(orc_topo* t, int m, int size, const double* p);
/* chain: 2 ints per particle (molecule type, position); kept by pointer. */
void orc_topo_set_chain(orc_topo* t, int n, const int* chain);

/* ---- particles ---- */
void orc_part_init(const orc_topo* t, orc_part* p);
orc_vec orc_image(const double box[3], orc_vec r1, orc_vec r2);

/* ---- energies ---- */
/* PairE::operator() with p1's connectivity list: bonds, angles and the
 * non-bonded functor */
double orc_pair_energy(const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2);
/* the bond and angle part alone (HarmonicSc/HarmonicSp + AngleSc) */
double orc_bond_energy(const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2);
orc_vec orc_min_dist_segments(orc_vec segA, orc_vec segB, double halfl1, double halfl2, orc_vec r_cm);
double orc_energy_all(const orc_topo* t, const double box[3], const orc_part* p, int n);
double orc_energy_one(const orc_topo* t, const double box[3], const orc_part* p, int n, int target);
/* whole-system energy through a cell grid: same pairs (each once, part1 =
 * larger index) in O(N) time; used at sizes where O(N^2) is too slow. */
double orc_energy_all_cells(const orc_topo* t, const double box[3], const orc_part* p, int n);

/* ---- SoA convenience for the ctypes binding ----
 * pos/dir/pd0/pd1 are n*3 doubles; pd1 may be NULL.  Positions are box
 * fractions.  Each particle is initialised with orc_part_init. */
int orc_soa_pair_energies(const orc_topo* t, const double box[3], int n,
                          const double* pos, const double* dir, const double* pd0, const double* pd1,
                          const int* type, int m, const int* pi, const int* pj, double* out);
double orc_soa_energy_all(const orc_topo* t, const double box[3], int n,
                          const double* pos, const double* dir, const double* pd0, const double* pd1,
                          const int* type, int use_cells);
/* bond + angle energy of each pair (i[k], j[k]) with i[k] as part1 */
int orc_soa_bond_energies(const orc_topo* t, const double box[3], int n,
                          const double* pos, const double* dir, const double* pd0, const double* pd1,
                          const int* type, int m, const int* pi, const int* pj, double* out);
int orc_soa_energy_one_each(const orc_topo* t, const double box[3], int n,
                            const double* pos, const double* dir, const double* pd0, const double* pd1,
                            const int* type, int m, const int* targets, double* out);
int orc_soa_initconfig(int n, const double box[3], double* pos, double* dir, double* pd0);
/* derived vectors in the reference's Particle layout: out is n*30 doubles:
 * dir, patchdir[0..1], patchsides[0..3], chdir[0..1] */
int orc_soa_derived(const orc_topo* t, int n, const double* dir, const double* pd0, const double* pd1,
                    const int* type, double* out);

/* ---- cell lists (bit-exact reference for the GPU cell list) ----
 * Fixed-point fractional positions u = frac(s) * 2^32.  The grid is shifted
 * by (sx,sy,sz) in the same units; cell index along an axis is
 * floor(((u - shift) mod 2^32) * nc / 2^32).  Output: cell id per particle
 * (x fastest), CSR starts (ncells+1) and particle indices sorted by cell,
 * ascending index inside a cell. */
int orc_cell_list(int n, const uint32_t* ux, const uint32_t* uy, const uint32_t* uz,
                  int ncx, int ncy, int ncz, uint32_t sx, uint32_t sy, uint32_t sz,
                  int32_t* cell_of, int32_t* cell_start, int32_t* sorted);
/* Brute-force neighbour list: pairs (i<j) whose minimum-image centre distance
 * (FP64, from the fixed-point positions) is <= rc.  Returns the pair count;
 * pairs are written as (i,j) in row-major order if out != NULL and count<=cap. */
long orc_neighbor_pairs(int n, const uint32_t* ux, const uint32_t* uy, const uint32_t* uz,
                        const double box[3], double rc, int32_t* out, long cap);

/* ---- sequential Metropolis MC (CPU baseline and ensemble reference) ----
 * Follows MoveCreator::particleMove/partDisplace/partRotate/moveTry
 * (scOOP/mc/movecreator.cpp:11-33,947-1028; movecreator.h:175-187) with a
 * counter RNG of this file (trajectories are not the reference's).
 * transmx is the full translation length (already doubled, sim.h:365), rot_half
 * is the maximum quaternion half-angle in radians (sim.h:360).
 * State arrays are updated in place.  stats[0..3] = trans acc, trans tries,
 * rot acc, rot tries. Returns the energy drift check value E_end - E_start - sum dE. */
double orc_mc_run(const orc_topo* t, const double box[3], int n,
                  double* pos, double* dir, double* pd0, double* pd1, const int* type,
                  long nsweeps, double temper, double transmx, double rot_half,
                  uint64_t seed, long* stats, double* e_final);

/* NPT: as orc_mc_run plus one volume move (MoveCreator::pressureMove,
 * movecreator.cpp:330-421; ptype 0 one random edge, 1 isotropic) at the end
 * of every sweep; box is updated in place; stats[4..5] = volume acc, tries.
 * The drift check covers the energy changes of both kinds of move. */
double orc_mc_run_npt(const orc_topo* t, double box[3], int n,
                      double* pos, double* dir, double* pd0, double* pd1, const int* type,
                      long nsweeps, double temper, double transmx, double rot_half,
                      uint64_t seed, long* stats, double* e_final, int ptype, double press, double edge_mx);

/* As orc_mc_run_npt (ptype < 0: no volume moves) with whole-molecule moves
 * (MoveCreator::chainMove / chainDisplace / chainRotate, movecreator.cpp:
 * 305-328, 1075-1305): each step is a chain move with probability chainprob;
 * chain_mx is the displacement length (already doubled, sim.h:366),
 * chain_angle the maximum quaternion half-angle (sim.h:362).  The rotation
 * centre is taken on minimum-image unwrapped coordinates.  stats needs 10
 * entries: [6..9] = chain displace acc, tries, rotate acc, tries. */
double orc_mc_run_chains(const orc_topo* t, double box[3], int n,
                         double* pos, double* dir, double* pd0, double* pd1, const int* type,
                         long nsweeps, double temper, double transmx, double rot_half,
                         double chainprob, double chain_mx, double chain_angle,
                         uint64_t seed, long* stats, double* e_final, int ptype, double press, double edge_mx);
/* mol2others / mol2othersTrial (totalenergycalculator.h:438-452, 484-497):
 * non-bonded energy of the molecule [first, first + size) with every
 * particle outside it (part1 = the molecule's particle). */
double orc_mol2others(const orc_topo* t, const double box[3], const orc_part* p, int n, int first, int size);
int orc_soa_mol2others(const orc_topo* t, const double box[3], int n, const double* pos, const double* dir,
                       const double* pd0, const double* pd1, const int* type, int m, const int* first,
                       const int* size, double* out);

/* ---- observables ---- */
/* nematic order S2 = -2 * middle eigenvalue of Q (Tests/sc35.c:8563-8603) */
double orc_nematic_s2(int n, const double* dir);

#ifdef __cplusplus
}
#endif
#endif
