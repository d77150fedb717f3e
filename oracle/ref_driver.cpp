// ref_driver.cpp -- test harness around the compiled reference (robertvacha/SC,
// scOOP 4.0), built by oracle/Makefile into oracle/_ref/scref together with the
// reference's own, unmodified sources.  TEST INFRASTRUCTURE: it is used only
// to write the golden vectors under tests/golden/ and by the live reference
// tests; nothing in the product links it.
//
// It reads options / top.init / config.init from the working directory exactly
// as scOOP's main() does (scOOP/main.cpp:60-90), then answers one query:
//   scref pairs  <out> <pairs.txt>  -> p2p(i,j) per line "i j" (TotalE::p2p,
//                                      totalenergycalculator.h:200-203)
//   scref total  <out>              -> allToAll after initEM (TotalEMatrix,
//                                      totalenergycalculator.h:314-353,502-521)
//                                      then oneToAll(i) for every particle
//   scref topo   <out>              -> ia_params of every defined type pair
//   scref derived <out>             -> dir, patchdir[2], patchsides[4], chdir[2]
//                                      after Conf::partVecInit
//   scref mindist <out> <cases.txt> -> EPatch::minDistSegments on raw inputs
//   scref molecules <out>           -> mol2others / mol2othersTrial of every
//                                      molecule of more than one particle
//                                      (totalenergycalculator.h:417-500)
//   scref run <out>                 -> the reference's whole program (its own
//                                      main(), scOOP/main.cpp:33-330: options,
//                                      sequential MC sweeps, energy.dat every
//                                      `report` sweeps, config.last); <out>
//                                      gets the wall-clock seconds of the run
// Values are printed with %.17g so the fixtures carry full FP64 precision.
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <vector>

#include "mc/inicializer.h"
#include "mc/updater.h"
#include "mc/totalenergycalculator.h"

extern Topo topo;
int scoop_reference_main(int argc, char** argv);   // scOOP/main.cpp, renamed at compile time

static void dumpVec(FILE* f, const Vector& v) { fprintf(f, " %.17g %.17g %.17g", v.x, v.y, v.z); }

// EPatch::minDistSegments is a public member template; any instantiation
// exposes it.
struct MinDistProbe : public EPatch<WcaTruncSq> {};

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: scref pairs|total|topo|derived|mindist <out> [in]\n");
        return 2;
    }
    std::string mode(argv[1]);
    FILE* out = fopen(argv[2], "w");
    if (!out) { perror("out"); return 2; }

    if (mode == "run") {
        char* args[] = {argv[0], nullptr};
        timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        int rc = scoop_reference_main(1, args);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        fprintf(out, "%.6f\n", (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
        fclose(out);
        return rc;
    }

    if (mode == "mindist") {
        // each line: ax ay az bx by bz h1 h2 rx ry rz
        FILE* in = fopen(argv[3], "r");
        double a[11];
        MinDistProbe probe;
        while (fscanf(in, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5],
                      &a[6], &a[7], &a[8], &a[9], &a[10]) == 11) {
            Vector r = probe.minDistSegments(Vector(a[0], a[1], a[2]), Vector(a[3], a[4], a[5]), a[6], a[7],
                                             Vector(a[8], a[9], a[10]));
            fprintf(out, "%.17g %.17g %.17g\n", r.x, r.y, r.z);
        }
        fclose(in);
        fclose(out);
        return 0;
    }

    FileNames files(0);
    Conf conf;
    Sim sim(&conf, &files, 0, 1);
    Inicializer init(&sim, &conf, &files);
    init.initTop();
    FILE* infile = fopen(files.configurationInFile, "r");
    if (!infile) { fprintf(stderr, "no config.init\n"); return 2; }
    if (!init.initConfig(&infile, conf.pvec)) return 3;
    fclose(infile);
    conf.partVecInit();

    if (mode == "topo") {
        for (int i = 0; i < MAXT; i++)
            for (int j = 0; j < MAXT; j++) {
                const Ia_param& p = topo.ia_params[i][j];
                if (p.geotype[0] == 0) continue;
                fprintf(out, "%d %d %d %d", i, j, p.geotype[0], p.geotype[1]);
                const double vals[] = {p.sigma, p.epsilon, p.A, p.B, p.pdis, p.pswitch, p.pswitchINV, p.rcut,
                                       p.rcutSq, p.rcutwca, p.rcutwcaSq, p.len[0], p.len[1], p.half_len[0],
                                       p.half_len[1], p.parallel, p.volume,
                                       p.pcangl[0], p.pcangl[1], p.pcangl[2], p.pcangl[3],
                                       p.pcanglsw[0], p.pcanglsw[1], p.pcanglsw[2], p.pcanglsw[3],
                                       p.pcoshalfi[0], p.pcoshalfi[1], p.pcoshalfi[2], p.pcoshalfi[3],
                                       p.psinhalfi[0], p.psinhalfi[1], p.psinhalfi[2], p.psinhalfi[3],
                                       p.csecpatchrot[0], p.csecpatchrot[1], p.ssecpatchrot[0], p.ssecpatchrot[1],
                                       p.chiral_cos[0], p.chiral_cos[1], p.chiral_sin[0], p.chiral_sin[1]};
                for (double v : vals) fprintf(out, " %.17g", v);
                fprintf(out, " %d\n", p.exclude ? 1 : 0);
            }
        fprintf(out, "sqmaxcut %.17g maxcut %.17g\n", topo.sqmaxcut, topo.maxcut);
    } else if (mode == "derived") {
        for (size_t i = 0; i < conf.pvec.size(); i++) {
            const Particle& p = conf.pvec[i];
            fprintf(out, "%d", p.type);
            dumpVec(out, p.pos);
            dumpVec(out, p.dir);
            dumpVec(out, p.patchdir[0]);
            dumpVec(out, p.patchdir[1]);
            for (int k = 0; k < 4; k++) dumpVec(out, p.patchsides[k]);
            dumpVec(out, p.chdir[0]);
            dumpVec(out, p.chdir[1]);
            fprintf(out, "\n");
        }
    } else if (mode == "pairs") {
        PairE pairE(&conf.geo);
        FILE* in = fopen(argv[3], "r");
        int i, j;
        while (fscanf(in, "%d %d", &i, &j) == 2) {
            ConList conlist = conf.pvec.getConlist(i);
            fprintf(out, "%.17g\n", pairE(&conf.pvec[i], &conf.pvec[j], &conlist));
        }
        fclose(in);
    } else if (mode == "total") {
        TotalEMatrix<PairE> calc(&sim, &conf);
        calc.initEM();
        fprintf(out, "%.17g\n", calc.allToAll());
        for (size_t i = 0; i < conf.pvec.size(); i++) fprintf(out, "%.17g\n", calc.oneToAll((int)i));
    } else if (mode == "molecules") {
        // every molecule of more than one particle (ParticleVector::getChain):
        // first index, size, TotalEMatrix::mol2others and mol2othersTrial on
        // the unchanged state
        TotalEMatrix<PairE> calc(&sim, &conf);
        calc.initEM();
        for (int k = 0; k < conf.pvec.getChainCount(); k++) {
            Molecule m = conf.pvec.getChain(k);
            double e0 = calc.mol2others(m);
            double e1 = calc.mol2othersTrial(m);
            fprintf(out, "%d %d %.17g %.17g\n", m[0], (int)m.size(), e0, e1);
        }
    } else {
        fprintf(stderr, "unknown mode\n");
        return 2;
    }
    fclose(out);
    return 0;
}
