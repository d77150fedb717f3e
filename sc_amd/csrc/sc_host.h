// sc_host.h -- host side of the drop-in: SC's input formats, the FP64 pair
// parameter table, and the device parameter tables built from it.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace schost {

constexpr int MAXT = 40;   // scOOP/structures/macros.h:87

// geometric type codes, scOOP/structures/macros.h:71-85
enum Geo { SCN = 10, SCA = 11, PSC = 12, CPSC = 13, CHPSC = 14, CHCPSC = 15, TPSC = 16, TCPSC = 17,
           TCHPSC = 18, TCHCPSC = 19, SPN = 30, SPA = 31 };

int geotypeFromName(const std::string& s);   // Inicializer::convertGeotype, inicializer.cpp:855-885

// FP64 pair parameters: the Ia_param fields the energy path uses
// (scOOP/structures/structures.h:187-222).
struct Ia {
    int geotype[2] = {0, 0};
    double sigma = 0, epsilon = 0, A = 0, B = 0, pdis = 0, pswitch = 0, pswitchINV = 0;
    double pangl[4] = {0, 0, 0, 0}, panglsw[4] = {0, 0, 0, 0}, pcangl[4] = {0, 0, 0, 0}, pcanglsw[4] = {0, 0, 0, 0};
    double rcut = 0, rcutSq = 0, rcutwca = 0, rcutwcaSq = 0;
    double pcoshalfi[4] = {0, 0, 0, 0}, psinhalfi[4] = {0, 0, 0, 0};
    double csecpatchrot[2] = {0, 0}, ssecpatchrot[2] = {0, 0};
    double len[2] = {0, 0}, half_len[2] = {0, 0};
    double chiral_cos[2] = {0, 0}, chiral_sin[2] = {0, 0};
    double parallel = 0, volume = 0;
    bool exclude = false;
    std::string name;
};

struct Molecule {
    std::string name;
    std::vector<int> types;
    double bond1c = -1, bond1eq = -1, bond2c = -1, bond2eq = -1, bonddc = -1, bonddeq = -1, bondhc = -1,
           bondheq = -1, angle1c = -1, angle1eq = 0, angle2c = -1, angle2eq = 0;
};

struct Topology {
    Ia ia[MAXT][MAXT];
    std::vector<Molecule> molecules;
    std::vector<std::pair<std::string, long>> system;
    double sqmaxcut = 0, maxcut = 0;
    bool loaded = false;

    // parses top.init text; returns "" on success or the error message
    std::string parse(const std::string& text);
    // particle types in config order (Inicializer::setParticlesParamss, :409-450)
    std::vector<int> particleTypes() const;
    std::vector<int> particleMolecules() const;

  private:
    std::string fillType(const std::string& line);
    void genParamPairs(const bool excl[MAXT][MAXT]);
    void genTopoParams();
};

struct Options {
    std::map<std::string, double> values;   // after the reference's post-processing
    std::string parse(const std::string& text);
    bool get(const std::string& k, double& v) const;
};

// config.init: box line then one line per particle (Inicializer::initConfig,
// inicializer.cpp:89-298).  Returns "" or the error.  pos comes back in box
// fractions folded into [0,1] (Cuboid::usePBC), dir normalised, patchdir
// orthogonalised to dir and normalised.
std::string parseConfig(const std::string& text, int n, double box[3], std::vector<double>& pos,
                        std::vector<double>& dir, std::vector<double>& pd0);
// Conf::draw (structures/Conf.h:388-425) for one state: the particle lines;
// drawConfig prefixes the box line of config.last (main.cpp:304-310),
// movieFrame the header of Updater::dumpMovie (mc/updater.h:82-93).
std::string drawParticles(const double box[3], const std::vector<double>& pos, const std::vector<double>& dir,
                          const std::vector<double>& pd0, const std::vector<int>& mol);
std::string drawConfig(const double box[3], const std::vector<double>& pos, const std::vector<double>& dir,
                       const std::vector<double>& pd0, const std::vector<int>& mol);
std::string movieFrame(long sweep, const double box[3], const std::vector<double>& pos, const std::vector<double>& dir,
                       const std::vector<double>& pd0, const std::vector<int>& mol);
// stat.dat / energy.dat lines (mc/updater.cpp:322-323)
std::string statLine(long sweep, double volume);
std::string energyLine(long sweep, double energy, double alignment);
// cluster_stat.dat and cluster.dat frames of ClusterSampler::writeCluster
// (mc/clust.cpp:7-49); energy_of_root[r] = energy of the cluster whose
// smallest member is r
void clusterFiles(long sweep, const std::vector<int>& labels, const std::vector<double>& energy_of_root,
                  std::string& stat, std::string& clusters);

// row of 44 doubles describing ia[ti][tj] (for parity dumps)
void iaRow(const Ia& p, double* out44);

}  // namespace schost
