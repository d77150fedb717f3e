// sc_host.cpp -- SC input formats and FP64 pair-parameter derivation (host).
#include "sc_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

namespace schost {

static const double kPI = 3.141592653589793238462643383279;   // macros.h:65
static const double kDEG2RAD = 0.017453292519943295769236907683;

int geotypeFromName(const std::string& s) {
    static const std::map<std::string, int> m = {{"SCN", SCN},   {"SCA", SCA},       {"PSC", PSC},     {"CPSC", CPSC},
                                                 {"CHPSC", CHPSC}, {"CHCPSC", CHCPSC}, {"TPSC", TPSC},   {"TCPSC", TCPSC},
                                                 {"TCHPSC", TCHPSC}, {"TCHCPSC", TCHCPSC}, {"SP", 30}, {"SPN", SPN},
                                                 {"SPA", SPA}};
    auto it = m.find(s);
    return it == m.end() ? 0 : it->second;
}

static bool isChiral(int g) { return g == CHPSC || g == CHCPSC || g == TCHPSC || g == TCHCPSC; }
static bool isTwo(int g) { return g == TPSC || g == TCPSC || g == TCHPSC || g == TCHCPSC; }

static std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
static std::string upper(std::string s) {
    for (auto& c : s) c = (char)toupper((unsigned char)c);
    return s;
}
static std::vector<std::string> split_ws(const std::string& s) {
    std::istringstream is(s);
    std::vector<std::string> v;
    std::string t;
    while (is >> t) v.push_back(t);
    return v;
}

// One [Types] line "NAME NUMBER GEOTYPE EPSILON SIGMA ..."
// (Inicializer::fillTypes, inicializer.cpp:660-852).  The arity of the
// numeric fields decides which parameters the type carries.
std::string Topology::fillType(const std::string& line) {
    std::string body = line;
    size_t colon = body.find(':');
    if (colon != std::string::npos) body = body.substr(colon + 1);
    auto f = split_ws(body);
    if (f.size() < 3) return "TOPOLOGY ERROR: short [Types] line: " + line;
    int type = atoi(f[1].c_str());
    int g = geotypeFromName(f[2]);
    if (!g) return "TOPOLOGY ERROR: Unknown GEOTYPE: " + f[2];
    if (type < 0 || type >= MAXT) return "TOPOLOGY ERROR: type number out of range: " + f[1];
    std::vector<double> p;
    for (size_t k = 3; k < f.size() && p.size() < 12; k++) {
        char* end = nullptr;
        double v = strtod(f[k].c_str(), &end);
        if (end == f[k].c_str()) break;
        p.push_back(v);
    }
    int fields = (int)p.size() - 2;
    bool bad = (g == SPN && fields != 0) || (g == SCN && fields != 1) || (g == SPA && fields != 2) ||
               (g == SCA && fields != 3) || ((g == PSC || g == CPSC) && fields != 6) ||
               ((g == CHPSC || g == CHCPSC) && fields != 7) || ((g == TPSC || g == TCPSC) && fields != 9) ||
               ((g == TCHPSC || g == TCHCPSC) && fields != 10);
    if (bad) return "TOPOLOGY ERROR: wrong number of parameters for " + f[2];
    p.resize(12, 0.0);
    Ia& t = ia[type][type];
    t = Ia();
    t.name = f[0];
    t.geotype[0] = t.geotype[1] = g;
    t.epsilon = p[0];
    t.sigma = p[1];
    t.A = 4 * t.epsilon * pow(t.sigma, 12);
    t.B = 4 * t.epsilon * pow(t.sigma, 6);
    t.rcutwca = t.sigma * pow(2.0, 1.0 / 6.0);
    t.rcutwcaSq = t.rcutwca * t.rcutwca;
    if (fields > 0 && fields != 1 && fields != 3) {     // everything with an attraction range
        t.pdis = p[2];
        t.pswitch = p[3];
        t.pswitchINV = 1.0 / p[3];
        t.rcut = (g != SPN) ? t.pswitch + t.pdis : 0.0;
        t.rcutSq = t.rcut * t.rcut;
    }
    if (fields == 1) for (int i = 0; i < 2; i++) { t.len[i] = p[2]; t.half_len[i] = p[2] / 2; }
    if (fields == 3) for (int i = 0; i < 2; i++) { t.len[i] = p[4]; t.half_len[i] = p[4] / 2; }
    if (fields > 2 && fields != 3) {
        for (int i = 0; i < 2; i++) {
            t.len[i] = p[6];
            t.half_len[i] = p[6] / 2;
            t.pangl[i] = p[4];
            t.panglsw[i] = p[5];
            t.pcangl[i] = cos(p[4] / 2.0 / 180 * kPI);
            t.pcanglsw[i] = cos((p[4] / 2.0 + p[5]) / 180 * kPI);
            t.pcoshalfi[i] = cos((p[4] / 2.0 + p[5]) / 2.0 / 180 * kPI);
            t.psinhalfi[i] = sqrt(1.0 - t.pcoshalfi[i] * t.pcoshalfi[i]);
        }
        t.parallel = p[7];
    }
    if (fields == 7)
        for (int i = 0; i < 2; i++) {
            t.chiral_cos[i] = cos(p[8] / 360 * kPI);
            t.chiral_sin[i] = sqrt(1 - t.chiral_cos[i] * t.chiral_cos[i]);
        }
    if (fields == 9 || fields == 10)
        for (int i = 0; i < 2; i++) {
            t.csecpatchrot[i] = cos(p[8] / 360 * kPI);
            t.ssecpatchrot[i] = sqrt(1 - t.csecpatchrot[i] * t.csecpatchrot[i]);
            t.pangl[i + 2] = p[9];
            t.panglsw[i + 2] = p[10];
            t.pcangl[i + 2] = cos(p[9] / 2.0 / 180 * kPI);
            t.pcanglsw[i + 2] = cos((p[9] / 2.0 + p[10]) / 180 * kPI);
            t.pcoshalfi[i + 2] = cos((p[9] / 2.0 + p[10]) / 2.0 / 180 * kPI);
            t.psinhalfi[i + 2] = sqrt(1.0 - t.pcoshalfi[i + 2] * t.pcoshalfi[i + 2]);
        }
    if (fields == 10)
        for (int i = 0; i < 2; i++) {
            t.chiral_cos[i] = cos(p[11] / 360 * kPI);
            t.chiral_sin[i] = sqrt(1 - t.chiral_cos[i] * t.chiral_cos[i]);
        }
    t.volume = (g < 30) ? 4.0 / 3.0 * kPI * pow(t.sigma / 2.0, 3.0) + kPI / 2.0 * t.len[0] * pow(t.sigma / 2.0, 2.0)
                        : 4.0 / 3.0 * kPI * pow(t.sigma / 2.0, 3.0);
    if (t.rcutwca > sqmaxcut) sqmaxcut = t.rcutwca;
    if (t.rcut > sqmaxcut) sqmaxcut = t.rcut;
    return "";
}

// Mixing rules for unlike pairs, Topo::genParamPairs (topo.cpp:5-138):
// sigma arithmetic, epsilon geometric, the flat part of the attraction
// averaged, patch geometry copied from each side.
void Topology::genParamPairs(const bool excl[MAXT][MAXT]) {
    for (int i = 0; i < MAXT; i++)
        for (int j = 0; j < MAXT; j++) {
            if (i == j || ia[i][i].geotype[0] == 0 || ia[j][j].geotype[0] == 0) continue;
            Ia& p = ia[i][j];
            const Ia* side[2] = {&ia[i][i], &ia[j][j]};
            for (int k = 0; k < 2; k++) {
                const Ia& s = *side[k];
                p.geotype[k] = s.geotype[0];
                p.len[k] = s.len[0];
                p.half_len[k] = s.half_len[0];
                if (p.geotype[k] >= PSC && p.geotype[k] < 30) {
                    p.pangl[k] = s.pangl[0];
                    p.panglsw[k] = s.panglsw[0];
                    p.pcangl[k] = cos(p.pangl[k] / 2.0 / 180 * kPI);
                    p.pcanglsw[k] = cos((p.pangl[k] / 2.0 + p.panglsw[k]) / 180 * kPI);
                    p.pcoshalfi[k] = cos((p.pangl[k] / 2.0 + p.panglsw[k]) / 2.0 / 180 * kPI);
                    p.psinhalfi[k] = sqrt(1.0 - p.pcoshalfi[k] * p.pcoshalfi[k]);
                }
                if (isChiral(p.geotype[k])) {
                    p.chiral_cos[k] = s.chiral_cos[0];
                    p.chiral_sin[k] = s.chiral_sin[0];
                }
                if (isTwo(p.geotype[k])) {
                    p.csecpatchrot[k] = s.csecpatchrot[0];
                    p.ssecpatchrot[k] = s.ssecpatchrot[0];
                    p.pangl[k + 2] = s.pangl[2];
                    p.panglsw[k + 2] = s.panglsw[2];
                    p.pcangl[k + 2] = cos(p.pangl[k + 2] / 2.0 / 180 * kPI);
                    p.pcanglsw[k + 2] = cos((p.pangl[k + 2] / 2.0 + p.panglsw[k + 2]) / 180 * kPI);
                    p.pcoshalfi[k + 2] = cos((p.pangl[k + 2] / 2.0 + p.panglsw[k + 2]) / 2.0 / 180 * kPI);
                    p.psinhalfi[k + 2] = sqrt(1.0 - p.pcoshalfi[k + 2] * p.pcoshalfi[k + 2]);
                }
            }
            p.name = ia[i][i].name;
            p.sigma = (ia[i][i].sigma + ia[j][j].sigma) * 0.5;
            p.epsilon = sqrt(ia[i][i].epsilon * ia[j][j].epsilon);
            p.A = 4 * p.epsilon * pow(p.sigma, 12);
            p.B = 4 * p.epsilon * pow(p.sigma, 6);
            p.pswitch = (ia[i][i].pswitch + ia[j][j].pswitch) * 0.5;
            p.pswitchINV = 1.0 / p.pswitch;
            p.rcutwca = p.sigma * pow(2.0, 1.0 / 6.0);
            p.rcutwcaSq = p.rcutwca * p.rcutwca;
            if (ia[i][i].parallel > 0 && ia[j][j].parallel > 0) p.parallel = sqrt(ia[i][i].parallel * ia[j][j].parallel);
            if (ia[i][i].parallel < 0 && ia[j][j].parallel < 0) p.parallel = -sqrt(ia[i][i].parallel * ia[j][j].parallel);
            p.pdis = ((ia[i][i].pdis - ia[i][i].rcutwca) + (ia[j][j].pdis - ia[j][j].rcutwca)) * 0.5 + p.rcutwca;
            p.rcut = (p.geotype[0] == SPN || p.geotype[1] == SPN) ? 0.0 : p.pswitch + p.pdis;
            p.rcutSq = p.rcut * p.rcut;
            if (p.rcutwca > sqmaxcut) sqmaxcut = p.rcutwca;
            if (p.rcut > sqmaxcut) sqmaxcut = p.rcut;
        }
    for (int i = 0; i < MAXT; i++)
        for (int j = 0; j < MAXT; j++) ia[i][j].exclude = excl[i][j];
}

// Topo::genTopoParams (topo.cpp:140-153): maxcut = 1.1*(largest cutoff + largest length)
void Topology::genTopoParams() {
    double maxlength = 0;
    for (int i = 0; i < MAXT; i++)
        if (maxlength < ia[i][i].len[0]) maxlength = ia[i][i].len[0];
    sqmaxcut += maxlength;
    sqmaxcut *= 1.1;
    maxcut = sqmaxcut;
    sqmaxcut = sqmaxcut * sqmaxcut;
}

std::string Topology::parse(const std::string& text) {
    *this = Topology();
    static bool excl[MAXT][MAXT];
    memset(excl, 0, sizeof(excl));
    // join continuation lines ('\' at the end, simlib.cpp:53-62), strip comments
    std::vector<std::string> lines;
    {
        std::istringstream is(text);
        std::string raw, acc;
        while (std::getline(is, raw)) {
            std::string r = raw;
            while (!r.empty() && (r.back() == ' ' || r.back() == '\t' || r.back() == '\r')) r.pop_back();
            if (!r.empty() && r.back() == '\\') { acc += r.substr(0, r.size() - 1); continue; }
            acc += r;
            size_t c = acc.find('#');
            if (c != std::string::npos) acc = acc.substr(0, c);
            lines.push_back(trim(acc));
            acc.clear();
        }
        if (!acc.empty()) lines.push_back(trim(acc));
    }
    std::string key, molname;
    bool inMol = false;
    for (const auto& line : lines) {
        if (line.empty()) continue;
        if (line[0] == '[') {
            size_t e = line.find(']');
            key = upper(trim(line.substr(1, e == std::string::npos ? std::string::npos : e - 1)));
            continue;
        }
        if (key == "TYPES") {
            std::string err = fillType(line);
            if (!err.empty()) return err;
        } else if (key == "MOLECULES") {
            // "NAME: {" opens a molecule, "}" closes it; inside "command: params"
            // (Inicializer::fillMol, inicializer.cpp:929-1108)
            std::string rest = line;
            if (!inMol) {
                size_t c = rest.find(':');
                molname = trim(rest.substr(0, c));
                Molecule m;
                m.name = molname;
                molecules.push_back(m);
                inMol = true;
                size_t ob = rest.find('{');
                rest = (ob == std::string::npos) ? std::string() : rest.substr(ob + 1);
            }
            bool closes = rest.find('}') != std::string::npos;
            rest = trim(rest.substr(0, rest.find('}')));
            if (!rest.empty()) {
                size_t c = rest.find(':');
                if (c == std::string::npos) return "TOPOLOGY ERROR: unknown parameter: " + rest;
                std::string cmd = upper(trim(rest.substr(0, c)));
                auto v = split_ws(rest.substr(c + 1));
                Molecule& m = molecules.back();
                auto two = [&](double& a, double& b) -> bool {
                    if (v.size() < 2) return false;
                    a = atof(v[0].c_str());
                    b = atof(v[1].c_str());
                    return b >= 0;
                };
                if (cmd == "PARTICLES") {
                    if (v.empty()) return "TOPOLOGY ERROR: could not read a pacticle.";
                    int t = atoi(v[0].c_str());
                    if (t < 0 || t > MAXT) return "TOPOLOGY ERROR: pacticles include type out of range";
                    m.types.push_back(t);
                } else if (cmd == "BOND1") {
                    if (!two(m.bond1c, m.bond1eq)) return "TOPOLOGY ERROR: wrong number of parameters for bond1";
                } else if (cmd == "BOND2") {
                    if (!two(m.bond2c, m.bond2eq)) return "TOPOLOGY ERROR: wrong number of parameters for bond2";
                } else if (cmd == "BONDD") {
                    if (!two(m.bonddc, m.bonddeq)) return "TOPOLOGY ERROR: wrong number of parameters for bondd";
                } else if (cmd == "BONDH") {
                    if (!two(m.bondhc, m.bondheq)) return "TOPOLOGY ERROR: wrong number of parameters for bondh";
                } else if (cmd == "ANGLE1") {
                    if (!two(m.angle1c, m.angle1eq)) return "TOPOLOGY ERROR: wrong number of parameters for angle1";
                    m.angle1eq *= kDEG2RAD;
                } else if (cmd == "ANGLE2") {
                    if (!two(m.angle2c, m.angle2eq)) return "TOPOLOGY ERROR: wrong number of parameters for angle2";
                    m.angle2eq *= kDEG2RAD;
                } else if (cmd == "ACTIVITY") {
                    // grand-canonical activity: not part of this path, accepted and ignored
                } else {
                    return "TOPOLOGY ERROR: unknown parameter: " + cmd;
                }
            }
            if (closes) inMol = false;
        } else if (key == "SYSTEM") {
            auto v = split_ws(line);
            if (v.size() < 2) return "TOPOLOGY ERROR: failed reading system from (" + line + ")";
            system.push_back({v[0], atol(v[1].c_str())});
        } else if (key == "EXCLUDE") {
            auto v = split_ws(line);
            if (v.size() % 2) return "Error in readin Topology exclusions, probably there is not even number of types";
            for (size_t k = 0; k + 1 < v.size(); k += 2) {
                int a = atoi(v[k].c_str()), b = atoi(v[k + 1].c_str());
                if (a < 0 || a >= MAXT || b < 0 || b >= MAXT) return "TOPOLOGY ERROR: exclusion type out of range";
                excl[a][b] = excl[b][a] = true;
            }
        } else if (key == "POOL" || key == "EXTER") {
            // grand-canonical pool and external wall: outside this path
        } else {
            return "TOPOLOGY ERROR: invalid keyword:" + key;
        }
    }
    for (const auto& s : system) {
        bool found = false;
        for (const auto& m : molecules) found |= (m.name == s.first);
        if (!found) return "TOPOLOGY ERROR: molecules " + s.first + " is not defined.";
    }
    genParamPairs(excl);
    genTopoParams();
    loaded = true;
    return "";
}

std::vector<int> Topology::particleTypes() const {
    std::vector<int> t;
    for (const auto& s : system)
        for (const auto& m : molecules)
            if (m.name == s.first)
                for (long c = 0; c < s.second; c++)
                    for (int ty : m.types) t.push_back(ty);
    return t;
}

std::vector<int> Topology::particleMolecules() const {
    std::vector<int> t;
    for (const auto& s : system)
        for (size_t mi = 0; mi < molecules.size(); mi++)
            if (molecules[mi].name == s.first)
                for (long c = 0; c < s.second; c++)
                    for (size_t k = 0; k < molecules[mi].types.size(); k++) t.push_back((int)mi);
    return t;
}

// Sim::readOptions (sim.h:123-406): every key of the table is required except
// transmx_type and coneAngle; unknown keys are reported and ignored.
std::string Options::parse(const std::string& text) {
    values.clear();
    static const char* required[] = {"write_cluster", "adjust", "movie", "nequil", "nsweeps", "nrepchange",
                                     "nGrandCanon", "nClustMove", "paramfrq", "report", "seed", "pairlist_update",
                                     "ptype", "wlm", "wlmtype", "press", "paralpress", "edge_mx", "shave",
                                     "chainprob", "switchprob", "temper", "paraltemper", "transmx", "rotmx",
                                     "chainmmx", "chainrmx"};
    std::istringstream is(text);
    std::string raw;
    while (std::getline(is, raw)) {
        size_t c = raw.find('#');
        std::string line = trim(c == std::string::npos ? raw : raw.substr(0, c));
        if (line.empty()) continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) return "error parsing Configuration line (" + line + ")";
        std::string id = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (val.empty()) return "error parsing Configuration line (" + line + ")";
        char* end = nullptr;
        double v = strtod(val.c_str(), &end);
        if (end == val.c_str()) return "could not read value of " + id;
        values[id] = v;
    }
    for (const char* k : required)
        if (!values.count(k)) return std::string("option '") + k + "' is not set!";
    double rotmx = values["rotmx"];
    if (rotmx < 0.0 || rotmx > 180) return "ERROR: Maximum orientation change must be in range 0 to 180.";
    double chainrmx = values["chainrmx"];
    if (chainrmx < 0.0 || chainrmx > 180)
        return "ERROR: Maximum orientation change for chains must be in range 0 to 180.";
    int ptype = (int)values["ptype"];
    if (ptype < 0 || ptype > 5) return "ERROR: Unknown pressure coupling";
    const double PIH = 1.57079632679489661923132169163975;
    values["rot_half"] = rotmx / 180.0 * PIH * 0.5;    // sim.h:360 (quaternion half-angle)
    values["transmx"] *= 2.0;                          // sim.h:365
    values["edge_mx"] *= 2.0;                          // sim.h:364
    values["chainmmx"] *= 2.0;                         // sim.h:366
    values["chain_angle"] = chainrmx / 180.0 * PIH;    // sim.h:362 (quaternion half-angle)
    return "";
}

bool Options::get(const std::string& k, double& v) const {
    auto it = values.find(k);
    if (it == values.end()) return false;
    v = it->second;
    return true;
}

std::string parseConfig(const std::string& text, int n, double box[3], std::vector<double>& pos,
                        std::vector<double>& dir, std::vector<double>& pd0) {
    std::istringstream is(text);
    std::string raw;
    auto next = [&](std::string& out) -> bool {
        while (std::getline(is, raw)) {
            size_t c = raw.find('#');
            out = trim(c == std::string::npos ? raw : raw.substr(0, c));
            return true;
        }
        return false;
    };
    std::string line;
    if (!next(line)) return "ERROR: Could not read box size";
    if (sscanf(line.c_str(), "%le %le %le", &box[0], &box[1], &box[2]) != 3) {
        // second form: a line "... x Lx Ly Lz" after the first (inicializer.cpp:131-143)
        if (!next(line)) return "ERROR: Could not read box size2.";
        size_t x = line.find('x');
        std::string l2 = x == std::string::npos ? line : line.substr(x + 1);
        if (sscanf(l2.c_str(), "%le %le %le", &box[0], &box[1], &box[2]) != 3) return "ERROR: Could not read box size3.";
    }
    pos.assign(3 * (size_t)n, 0.0);
    dir.assign(3 * (size_t)n, 0.0);
    pd0.assign(3 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++) {
        if (!next(line)) return "ERROR: Could not read coordinates for particle " + std::to_string(i + 1);
        double v[9];
        int sw = 0;
        int f = sscanf(line.c_str(), "%le %le %le %le %le %le %le %le %le %d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                       &v[6], &v[7], &v[8], &sw);
        if (f < 9) return "ERROR: Could not read coordinates for particle " + std::to_string(i + 1);
        for (int k = 0; k < 3; k++) {
            double s = v[k] / box[k];
            while (s < 0.0) s += 1.0;       // Cuboid::usePBC, geometry.h:78-97
            while (s > 1.0) s -= 1.0;
            pos[3 * i + k] = s;
        }
        double d[3] = {v[3], v[4], v[5]}, q[3] = {v[6], v[7], v[8]};
        double t = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (t != 0.0) { t = 1.0 / t; d[0] *= t; d[1] *= t; d[2] *= t; }
        double dp = q[0] * d[0] + q[1] * d[1] + q[2] * d[2];
        q[0] -= dp * d[0]; q[1] -= dp * d[1]; q[2] -= dp * d[2];
        t = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (t != 0.0) { t = 1.0 / t; q[0] *= t; q[1] *= t; q[2] *= t; }
        for (int k = 0; k < 3; k++) { dir[3 * i + k] = d[k]; pd0[3 * i + k] = q[k]; }
    }
    return "";
}

// Conf::draw (structures/Conf.h:413-421): one line per particle
std::string drawParticles(const double box[3], const std::vector<double>& pos, const std::vector<double>& dir,
                          const std::vector<double>& pd0, const std::vector<int>& mol) {
    std::string out;
    char buf[512];
    size_t n = pos.size() / 3;
    for (size_t i = 0; i < n; i++) {
        double r[3];
        for (int k = 0; k < 3; k++) r[k] = box[k] * (pos[3 * i + k] - nearbyint(pos[3 * i + k]));
        snprintf(buf, sizeof buf, "%15.8e %15.8e %15.8e   %15.8e %15.8e %15.8e   %15.8e %15.8e %15.8e %d %d\n", r[0],
                 r[1], r[2], dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], pd0[3 * i], pd0[3 * i + 1], pd0[3 * i + 2], 0,
                 i < mol.size() ? mol[i] : 0);
        out += buf;
    }
    return out;
}

// config.last (main.cpp:304-310): the box line, then Conf::draw
std::string drawConfig(const double box[3], const std::vector<double>& pos, const std::vector<double>& dir,
                       const std::vector<double>& pd0, const std::vector<int>& mol) {
    char buf[128];
    snprintf(buf, sizeof buf, "%15.8e %15.8e %15.8e\n", box[0], box[1], box[2]);
    return std::string(buf) + drawParticles(box, pos, dir, pd0, mol);
}

// Updater::dumpMovie (mc/updater.h:82-93): particle count, sweep and box, then Conf::draw
std::string movieFrame(long sweep, const double box[3], const std::vector<double>& pos, const std::vector<double>& dir,
                       const std::vector<double>& pd0, const std::vector<int>& mol) {
    char buf[256];
    snprintf(buf, sizeof buf, "%ld\nsweep %ld; box %.10f %.10f %.10f\n", (long)(pos.size() / 3), sweep, box[0], box[1],
             box[2]);
    return std::string(buf) + drawParticles(box, pos, dir, pd0, mol);
}

// stat.dat and energy.dat lines of Updater::simulate (mc/updater.cpp:322-323)
std::string statLine(long sweep, double volume) {
    char buf[128];
    snprintf(buf, sizeof buf, " %ld; %.10f\n", sweep, volume);
    return buf;
}
std::string energyLine(long sweep, double energy, double alignment) {
    char buf[160];
    snprintf(buf, sizeof buf, " %ld; %.10f  %f \n", sweep, energy, alignment);
    return buf;
}

// ClusterSampler::writeCluster without decoration (mc/clust.cpp:7-49):
// clusters in order of their smallest member (sortClusterList, :165-244),
// cluster_stat.dat = "Sweep: s | Maximal size: m" + the size histogram
// (printClusterStat, clust.h:161-176), cluster.dat = "Sweep: s | Number of
// clusters: c" + per cluster its energy and 1-based members (printClusters,
// clust.h:138-158).  labels[i] = the smallest index of i's cluster.
void clusterFiles(long sweep, const std::vector<int>& labels, const std::vector<double>& energy_of_root,
                  std::string& stat, std::string& clusters) {
    size_t n = labels.size();
    std::vector<std::vector<int>> members;
    std::vector<int> root_of_cluster;
    std::vector<int> cluster_of_root(n, -1);
    for (size_t i = 0; i < n; i++) {
        int r = labels[i];
        if (cluster_of_root[r] < 0) {
            cluster_of_root[r] = (int)members.size();
            members.emplace_back();
            root_of_cluster.push_back(r);
        }
        members[cluster_of_root[r]].push_back((int)i);
    }
    long maxc = 0;
    for (auto& m : members) maxc = std::max(maxc, (long)m.size());
    std::vector<long> hist(maxc, 0);
    for (auto& m : members) hist[m.size() - 1]++;
    char buf[128];
    snprintf(buf, sizeof buf, "Sweep: %ld | Maximal size: %ld\n", sweep, maxc);
    stat = buf;
    for (long i = 0; i < maxc; i++) {
        snprintf(buf, sizeof buf, "%5d\t%5ld\n", (int)(i + 1), hist[i]);
        stat += buf;
    }
    snprintf(buf, sizeof buf, "Sweep: %ld | Number of clusters: %ld\n", sweep, (long)members.size());
    clusters = buf;
    for (size_t c = 0; c < members.size(); c++) {
        snprintf(buf, sizeof buf, "%3d(%f):", (int)(c + 1), energy_of_root[root_of_cluster[c]]);
        clusters += buf;
        for (int p : members[c]) {
            snprintf(buf, sizeof buf, "%5ld", (long)p + 1);
            clusters += buf;
        }
        clusters += "\n";
    }
}

void iaRow(const Ia& p, double* o) {
    const double v[] = {(double)p.geotype[0], (double)p.geotype[1], p.sigma, p.epsilon, p.A, p.B, p.pdis, p.pswitch,
                        p.pswitchINV, p.rcut, p.rcutSq, p.rcutwca, p.rcutwcaSq, p.len[0], p.len[1], p.half_len[0],
                        p.half_len[1], p.parallel, p.volume, p.pcangl[0], p.pcangl[1], p.pcangl[2], p.pcangl[3],
                        p.pcanglsw[0], p.pcanglsw[1], p.pcanglsw[2], p.pcanglsw[3], p.pcoshalfi[0], p.pcoshalfi[1],
                        p.pcoshalfi[2], p.pcoshalfi[3], p.psinhalfi[0], p.psinhalfi[1], p.psinhalfi[2],
                        p.psinhalfi[3], p.csecpatchrot[0], p.csecpatchrot[1], p.ssecpatchrot[0], p.ssecpatchrot[1],
                        p.chiral_cos[0], p.chiral_cos[1], p.chiral_sin[0], p.chiral_sin[1], p.exclude ? 1.0 : 0.0};
    static_assert(sizeof(v) / sizeof(v[0]) == 44, "row width");
    for (int k = 0; k < 44; k++) o[k] = v[k];
}

}  // namespace schost
