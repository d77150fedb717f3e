// Compile-and-run check of include/sc_gpu_totale.hpp against minimal
// stand-ins with the member layout the adapter reads from SC's Particle /
// Conf (pos, dir, patchdir[0], type; geo.box; pvec).  It exercises the host
// path only: without a GPU sc_create fails and the adapter reports it.
#include <cstdio>
#include <vector>
#include "../../include/sc_gpu_totale.hpp"

struct Vec { double x, y, z; };
struct Part { Vec pos, dir, patchdir[2]; int type; };
struct Geo { Vec box; };
struct Conf { std::vector<Part> pvec; Geo geo; };
struct Sim {};

int main() {
    Sim sim;
    Conf conf;
    conf.geo.box = {10, 10, 10};
    conf.pvec.push_back({{0.1, 0.1, 0.1}, {1, 0, 0}, {{0, 1, 0}, {0, 0, 0}}, 1});
    GpuTotalE<Sim, Conf> e(&sim, &conf, "[Types]\nA 1 PSC 1 1 1 0.2 80 5 3 0\n[Molecules]\nA: {\nparticles: 1\n}\n[System]\nA 1\n", 0);
    printf("context %s\n", e.context() ? "created" : "absent");
    if (!e.context()) {   // the molecule methods report the missing context and return 0
        std::vector<int> mol(1, 0);
        double a = e.mol2others(mol) + e.mol2othersTrial(mol);
        e.update(mol);
        e.update(0);
        printf("molecule calls %g\n", a);
    }
    return 0;
}
