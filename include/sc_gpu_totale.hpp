// sc_gpu_totale.hpp -- drop-in TotalEnergyCalculator for robertvacha/SC
// (scOOP 4.0) backed by the gfx950 library (include/sc_gpu.h).
//
// The reference selects its energy backend with one typedef
// (scOOP/mc/totalenergycalculator.h:1225-1238):
//     typedef TotalEMatrix<PairE> TotalEnergyCalculator;
// GpuTotalE<Sim, Conf> has the method set its callers use
// (Updater, updater.h:33; MoveCreator, movecreator.h:36; ClusterSampler,
// clust.h:23) with the same call protocol:
//   particle move: E_old = oneToAll(t); caller mutates conf->pvec[t];
//                  E_new = oneToAllTrial(t); accept -> update(t);
//                  reject -> caller restores pvec[t] (device state untouched)
//                  (MoveCreator::partDisplace / partRotate,
//                   movecreator.cpp:947-1028)
//   volume move:   E_old = allToAll(); caller mutates conf->geo.box;
//                  E_new = allToAllTrial(); accept -> update()
//                  (MoveCreator::pressureMove, movecreator.cpp:330-421)
//   chain move:    E_old = mol2others(chain); caller mutates the chain's
//                  pvec entries; E_new = mol2othersTrial(chain);
//                  accept -> update(chain)
//                  (MoveCreator::chainDisplace / chainRotate,
//                   movecreator.cpp:1075-1256)
// It only reads the members the reference's Particle / Conf carry (pos as box
// fractions, dir, patchdir[0], type; geo.box; pvec.size()).  Errors are
// reported on stderr and return 0.0 / leave the state, as the reference's
// functors do for unprogrammed pairs (paire.h:207-212).
//
// The rest of TotalE's interface (totalenergycalculator.h:135-296) is there
// so that the whole program compiles with this class at the typedef: the
// (Sim*, Conf*) constructor Updater uses (updater.h:17; it reads top.init
// from the working directory, as the Inicializer does), update(EMResize)
// after grand-canonical insertions and deletions (the configuration is
// uploaded again), and pairE.pbc / pairE.verbose, which chainRotate reads
// (movecreator.cpp:1155-1222).  The energies of particles outside pvec
// (p2p(Particle*, int), mol2others(vector<Particle>&), chainInner) belong to
// the cluster and grand-canonical moves, which this backend does not cover:
// they report that on stderr and return 0.
#pragma once

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "sc_gpu.h"

template <typename SimT, typename ConfT>
class GpuTotalE {
  public:
    // top_init_text: the top.init the Inicializer read (the library derives
    // the same pair table, Topo::genParamPairs / genTopoParams)
    GpuTotalE(SimT* sim, ConfT* conf, const std::string& top_init_text, int device = 0)
        : sim_(sim), conf_(conf) {
        pairE.pbc = &conf->geo;
        open(top_init_text, device);
    }
    // the constructor of the typedef's users (Updater, updater.h:17): the
    // topology from top.init in the working directory (FileNames, files.h)
    GpuTotalE(SimT* sim, ConfT* conf) : sim_(sim), conf_(conf) {
        pairE.pbc = &conf->geo;
        std::ifstream in("top.init");
        std::stringstream text;
        text << in.rdbuf();
        if (!in) fprintf(stderr, "GpuTotalE: cannot read top.init\n");
        open(text.str(), 0);
    }
    ~GpuTotalE() { sc_destroy(ctx_); }
    GpuTotalE(const GpuTotalE&) = delete;
    GpuTotalE& operator=(const GpuTotalE&) = delete;

    // TotalEMatrix::initEM (totalenergycalculator.h:314): upload Conf::pvec
    void initEM() {
        size_t n = conf_->pvec.size();
        std::vector<double> pos(3 * n), dir(3 * n), pd0(3 * n);
        std::vector<int32_t> types(n);
        for (size_t i = 0; i < n; i++) {
            const auto& p = conf_->pvec[i];
            pos[3 * i] = p.pos.x; pos[3 * i + 1] = p.pos.y; pos[3 * i + 2] = p.pos.z;
            dir[3 * i] = p.dir.x; dir[3 * i + 1] = p.dir.y; dir[3 * i + 2] = p.dir.z;
            pd0[3 * i] = p.patchdir[0].x; pd0[3 * i + 1] = p.patchdir[0].y; pd0[3 * i + 2] = p.patchdir[0].z;
            types[i] = p.type;
        }
        double box[3] = {conf_->geo.box.x, conf_->geo.box.y, conf_->geo.box.z};
        check(sc_set_config(ctx_, (int)n, box, pos.data(), dir.data(), pd0.data(), types.data()), "sc_set_config");
    }
    // TotalEMatrix::allToAll (:338-353)
    double allToAll() { double e = 0; check(sc_energy_all(ctx_, &e), "sc_energy_all"); return e; }
    // TotalEMatrix::allToAllTrial (:334-336): the box the caller just changed
    double allToAllTrial() {
        double box[3] = {conf_->geo.box.x, conf_->geo.box.y, conf_->geo.box.z}, e = 0;
        check(sc_energy_all_trial(ctx_, box, &e), "sc_energy_all_trial");
        return e;
    }
    // TotalEMatrix::oneToAll (:355-381)
    double oneToAll(int target) { double e = 0; check(sc_energy_one(ctx_, target, &e), "sc_energy_one"); return e; }
    // TotalEMatrix::oneToAllTrial (:383-415): the state the caller just wrote into pvec[target]
    double oneToAllTrial(int target) {
        double pos[3], dir[3], pd0[3], e = 0;
        read(target, pos, dir, pd0);
        check(sc_energy_one_trial(ctx_, target, pos, dir, pd0, &e), "sc_energy_one_trial");
        return e;
    }
    // TotalEMatrix::update(int) (:326-328)
    void update(int target) {
        double pos[3], dir[3], pd0[3];
        read(target, pos, dir, pd0);
        check(sc_commit_particle(ctx_, target, pos, dir, pd0), "sc_commit_particle");
    }
    // TotalEMatrix::update() (:318-320): commit the box of a volume move
    void update() {
        double box[3] = {conf_->geo.box.x, conf_->geo.box.y, conf_->geo.box.z};
        check(sc_commit_box(ctx_, box), "sc_commit_box");
    }
    // TotalEMatrix::update(EMResize) (:322-324): the particle count changed
    // (grand-canonical moves): upload the configuration again
    template <typename R>
    typename std::enable_if<std::is_empty<R>::value>::type update(const R&) {
        initEM();
    }
    // TotalEMatrix::mol2others (:417-453); a Molecule (structures.h:15, a
    // vector<int>) lists consecutive particle indices
    template <typename MolT>
    typename std::enable_if<std::is_convertible<decltype(std::declval<MolT&>()[0]), int>::value, double>::type
    mol2others(MolT& mol) {
        double e = 0;
        check(sc_energy_mol(ctx_, mol[0], (int)mol.size(), &e), "sc_energy_mol");
        return e;
    }
    // TotalEMatrix::mol2othersTrial (:455-500): the states the caller just
    // wrote into the molecule's pvec entries (chainDisplace / chainRotate,
    // movecreator.cpp:1075-1256)
    template <typename MolT>
    double mol2othersTrial(MolT& mol) {
        int m = (int)mol.size();
        std::vector<double> pos(3 * m), dir(3 * m), pd0(3 * m);
        for (int a = 0; a < m; a++) read(mol[a], &pos[3 * a], &dir[3 * a], &pd0[3 * a]);
        double e = 0;
        check(sc_energy_mol_trial(ctx_, mol[0], m, pos.data(), dir.data(), pd0.data(), &e), "sc_energy_mol_trial");
        return e;
    }
    // TotalEMatrix::update(Molecule) (:330-332)
    template <typename MolT>
    auto update(const MolT& mol) -> decltype((void)mol.size(), void()) {
        int m = (int)mol.size();
        std::vector<double> pos(3 * m), dir(3 * m), pd0(3 * m);
        for (int a = 0; a < m; a++) read(mol[a], &pos[3 * a], &dir[3 * a], &pd0[3 * a]);
        check(sc_commit_mol(ctx_, mol[0], m, pos.data(), dir.data(), pd0.data()), "sc_commit_mol");
    }
    // TotalE::p2p (:200-203)
    double p2p(int part1, int part2) {
        double e = 0;
        int32_t i = part1, j = part2;
        check(sc_pair_energies(ctx_, 1, &i, &j, &e, 0), "sc_pair_energies");
        return e;
    }
    // TotalE::p2p(Particle*, int), mol2others(vector<Particle>&) and
    // chainInner (:204-296): particles outside pvec (cluster and
    // grand-canonical moves), not covered
    template <typename P>
    double p2p(P* part1, int part2) {
        (void)part1; (void)part2;
        return unsupported("p2p(Particle*, int)");
    }
    template <typename PV>
    typename std::enable_if<!std::is_convertible<decltype(std::declval<PV&>()[0]), int>::value, double>::type
    mol2others(PV& particles) {
        (void)particles;
        return unsupported("mol2others(vector<Particle>&)");
    }
    template <typename C>
    double chainInner(C& chain) {
        (void)chain;
        return unsupported("chainInner");
    }
    sc_ctx* context() { return ctx_; }

    // what chainRotate reads of TotalE::pairE (movecreator.cpp:1155-1222)
    struct PairView {
        bool verbose = false;
        decltype(&std::declval<ConfT&>().geo) pbc = nullptr;
    } pairE;

  private:
    void open(const std::string& top_init_text, int device) {
        if (sc_create(&ctx_, device) != SC_OK) {
            fprintf(stderr, "GpuTotalE: no usable gfx950 device %d\n", device);
            ctx_ = nullptr;
            return;
        }
        check(sc_load_topology(ctx_, top_init_text.c_str()), "sc_load_topology");
    }
    double unsupported(const char* what) const {
        fprintf(stderr, "GpuTotalE: %s is not covered by the GPU backend (cluster / grand-canonical moves)\n", what);
        return 0.0;
    }
    void read(int t, double* pos, double* dir, double* pd0) const {
        const auto& p = conf_->pvec[t];
        pos[0] = p.pos.x; pos[1] = p.pos.y; pos[2] = p.pos.z;
        dir[0] = p.dir.x; dir[1] = p.dir.y; dir[2] = p.dir.z;
        pd0[0] = p.patchdir[0].x; pd0[1] = p.patchdir[0].y; pd0[2] = p.patchdir[0].z;
    }
    void check(int rc, const char* what) const {
        if (rc != SC_OK) fprintf(stderr, "GpuTotalE: %s failed: %s\n", what, sc_last_error(ctx_));
    }
    SimT* sim_;
    ConfT* conf_;
    sc_ctx* ctx_ = nullptr;
};
