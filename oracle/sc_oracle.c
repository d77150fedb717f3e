/*
 * sc_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE, never shipped).
 *
 * A plain-C restatement of the robertvacha/SC (scOOP 4.0) pair-energy path.
 * The arithmetic of each routine follows the cited reference lines step for
 * step, because the branch structure (0.0 "no intersection" sentinel, exact
 * equality de-duplication of intersection parameters, clamping order) is part
 * of the energy definition.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * What pins each part (tests/test_oracle.py, tests/test_oracle_ref.py):
 *   image, minDistSegments, WCA, cos^2 well, fanglScale, scparallel:
 *       closed forms, brute-force (s,t) minimisation, golden vectors from the
 *       compiled reference (oracle/_ref), symmetry E(i,j)=E(j,i).
 *   pscIntersect / cpscIntersect / atrE / MixSpSc / Sphere functors and the
 *   TCHPSC four-term composition: golden vectors from the compiled reference
 *   (tests/golden/pair_golden.npz) and the live reference when it is built.
 *   parameter derivation (fillTypes/genParamPairs/genTopoParams): the
 *   reference's own ia_params dumped into tests/golden/topo_golden.json.
 *   cell list / neighbour pairs: brute force over integers (bit-exact).
 *   bonds and angles (HarmonicSc, HarmonicSp, AngleSc, getConlist): golden
 *       vectors of five chain systems from the compiled reference
 *       (tests/golden/chain_golden.npz: bond1/bond2/bondd/bondh/angle1/angle2,
 *       spheres at either chain end) and closed forms of each term
 *       (tests/test_oracle_chains.py).
 *   orc_mc_run: energy-drift invariant (E_end - E_start - sum dE = 0) and the
 *   acceptance rule restated from movecreator.h:175-187; its trajectories are
 *   not the reference's (own RNG) -- parity of sampled ensembles only.
 */
#include "sc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793238462643383279   /* macros.h:65 */
#define PIH_D 1.57079632679489661923132169163975 /* macros.h:66 */

/* ------------------------------------------------------------------ vectors */
static inline orc_vec V(double x, double y, double z) { orc_vec r = {x, y, z}; return r; }
static inline double DOT(orc_vec a, orc_vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline orc_vec SCL(double s, orc_vec a) { return V(a.x * s, a.y * s, a.z * s); }
static inline orc_vec NEG(orc_vec a) { return V(-1.0 * a.x, -1.0 * a.y, -1.0 * a.z); }
/* Vector::cross / vecCrossProduct, Vector.h:129, math_calc.h:39 */
static inline orc_vec CROSS(orc_vec a, orc_vec b) {
    return V(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
/* Vector::size, Vector.h:48 */
static inline double SIZE(orc_vec a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* Vector::perpProject, Vector.h:233-244 */
static inline orc_vec PERP(orc_vec a, orc_vec b) {
    double dp = DOT(a, b);
    return V(a.x - b.x * dp, a.y - b.y * dp, a.z - b.z * dp);
}
/* Vector::normalise, Vector.h:56-64 */
static inline orc_vec NORM(orc_vec a) {
    double tot = SIZE(a);
    if (tot != 0.0) { tot = 1.0 / tot; a.x *= tot; a.y *= tot; a.z *= tot; }
    return a;
}
/* Vector::ortogonalise, Vector.h:133-135: a -= (a.b) b */
static inline orc_vec ORTHO(orc_vec a, orc_vec b) {
    double dp = DOT(a, b);
    a.x -= dp * b.x; a.y -= dp * b.y; a.z -= dp * b.z;
    return a;
}
/* Vector::rotate(axis, cos, sin): rotation by the quaternion (c, s*axis),
 * Vector.h:138-160 */
static orc_vec ROT(orc_vec v, orc_vec ax, double c, double s) {
    double qw = c, qx = ax.x * s, qy = ax.y * s, qz = ax.z * s;
    double t2 = qw * qx, t3 = qw * qy, t4 = qw * qz, t5 = -qx * qx, t6 = qx * qy;
    double t7 = qx * qz, t8 = -qy * qy, t9 = qy * qz, t10 = -qz * qz;
    orc_vec r;
    r.x = 2.0 * ((t8 + t10) * v.x + (t6 - t4) * v.y + (t3 + t7) * v.z) + v.x;
    r.y = 2.0 * ((t4 + t6) * v.x + (t5 + t10) * v.y + (t9 - t2) * v.z) + v.y;
    r.z = 2.0 * ((t7 - t3) * v.x + (t2 + t9) * v.y + (t5 + t8) * v.z) + v.z;
    return r;
}

/* ----------------------------------------------------------------- geotypes */
static inline int is_psc(int g) { return g == ORC_PSC || g == ORC_CHPSC || g == ORC_TPSC || g == ORC_TCHPSC; }
static inline int is_cpsc(int g) { return g == ORC_CPSC || g == ORC_CHCPSC || g == ORC_TCPSC || g == ORC_TCHCPSC; }
static inline int is_chiral(int g) { return g == ORC_CHPSC || g == ORC_CHCPSC || g == ORC_TCHPSC || g == ORC_TCHCPSC; }
static inline int is_two(int g) { return g == ORC_TPSC || g == ORC_TCPSC || g == ORC_TCHPSC || g == ORC_TCHCPSC; }

/* ----------------------------------------------------------------- topology */
void orc_topo_clear(orc_topo* t) { memset(t, 0, sizeof(*t)); }

const orc_ia* orc_topo_ia(const orc_topo* t, int i, int j) { return &t->ia[i][j]; }

/* Inicializer::fillTypes, scOOP/mc/inicializer.cpp:660-852.  params are the
 * fields after GEOTYPE; nparams is the reference's "fields" after -= 5 plus 2
 * (EPSILON SIGMA), i.e. fields == nparams - 2. */
int orc_topo_add_type(orc_topo* t, int type, int g, const double* param, int nparams) {
    int fields = nparams - 2;
    orc_ia* ia;
    int i;
    if (type < 0 || type >= ORC_MAXT) return 0;
    if (g == ORC_SPN && fields != 0) return 0;                             /* :700 */
    if (g == ORC_SCN && fields != 1) return 0;                             /* :705 */
    if (g == ORC_SPA && fields != 2) return 0;                             /* :710 */
    if (g == ORC_SCA && fields != 3) return 0;                             /* :715 */
    if ((g == ORC_PSC || g == ORC_CPSC) && fields != 6) return 0;          /* :720 */
    if ((g == ORC_CHPSC || g == ORC_CHCPSC) && fields != 7) return 0;      /* :726 */
    if ((g == ORC_TPSC || g == ORC_TCPSC) && fields != 9) return 0;        /* :731 */
    if ((g == ORC_TCHPSC || g == ORC_TCHCPSC) && fields != 10) return 0;   /* :736 */
    ia = &t->ia[type][type];
    ia->geotype[0] = ia->geotype[1] = g;
    ia->epsilon = param[0];
    ia->sigma = param[1];
    ia->A = 4 * ia->epsilon * pow(ia->sigma, 12);
    ia->B = 4 * ia->epsilon * pow(ia->sigma, 6);
    ia->rcutwca = ia->sigma * pow(2.0, 1.0 / 6.0);
    ia->rcutwcaSq = ia->rcutwca * ia->rcutwca;
    if (fields > 0 && fields != 1 && fields != 3) {                        /* :762 */
        ia->pdis = param[2];
        ia->pswitch = param[3];
        ia->pswitchINV = 1.0 / param[3];
        ia->rcut = (g != ORC_SPN) ? ia->pswitch + ia->pdis : 0.0;
        ia->rcutSq = ia->rcut * ia->rcut;
    }
    if (fields == 1) { for (i = 0; i < 2; i++) { ia->len[i] = param[2]; ia->half_len[i] = param[2] / 2; } }
    if (fields == 3) { for (i = 0; i < 2; i++) { ia->len[i] = param[4]; ia->half_len[i] = param[4] / 2; } }
    if (fields > 2 && fields != 3) {                                       /* :787-804 */
        for (i = 0; i < 2; i++) {
            ia->len[i] = param[6];
            ia->half_len[i] = param[6] / 2;
            ia->pangl[i] = param[4];
            ia->panglsw[i] = param[5];
            ia->pcangl[i] = cos(param[4] / 2.0 / 180 * PI_D);
            ia->pcanglsw[i] = cos((param[4] / 2.0 + param[5]) / 180 * PI_D);
            ia->pcoshalfi[i] = cos((param[4] / 2.0 + param[5]) / 2.0 / 180 * PI_D);
            ia->psinhalfi[i] = sqrt(1.0 - ia->pcoshalfi[i] * ia->pcoshalfi[i]);
            ia->parallel = param[7];
        }
    }
    if (fields == 7) {                                                     /* :805-812 */
        for (i = 0; i < 2; i++) {
            ia->chiral_cos[i] = cos(param[8] / 360 * PI_D);
            ia->chiral_sin[i] = sqrt(1 - ia->chiral_cos[i] * ia->chiral_cos[i]);
        }
    }
    if (fields == 9 || fields == 10) {                                     /* :813-830 */
        for (i = 0; i < 2; i++) {
            ia->csecpatchrot[i] = cos(param[8] / 360 * PI_D);
            ia->ssecpatchrot[i] = sqrt(1 - ia->csecpatchrot[i] * ia->csecpatchrot[i]);
            ia->pangl[i + 2] = param[9];
            ia->panglsw[i + 2] = param[10];
            ia->pcangl[i + 2] = cos(param[9] / 2.0 / 180 * PI_D);
            ia->pcanglsw[i + 2] = cos((param[9] / 2.0 + param[10]) / 180 * PI_D);
            ia->pcoshalfi[i + 2] = cos((param[9] / 2.0 + param[10]) / 2.0 / 180 * PI_D);
            ia->psinhalfi[i + 2] = sqrt(1.0 - ia->pcoshalfi[i + 2] * ia->pcoshalfi[i + 2]);
        }
    }
    if (fields == 10) {                                                    /* :831-838 */
        for (i = 0; i < 2; i++) {
            ia->chiral_cos[i] = cos(param[11] / 360 * PI_D);
            ia->chiral_sin[i] = sqrt(1 - ia->chiral_cos[i] * ia->chiral_cos[i]);
        }
    }
    if (g < ORC_SP)                                                        /* :841-844 */
        ia->volume = 4.0 / 3.0 * PI_D * pow(ia->sigma / 2.0, 3.0) + PI_D / 2.0 * ia->len[0] * pow(ia->sigma / 2.0, 2.0);
    else
        ia->volume = 4.0 / 3.0 * PI_D * pow(ia->sigma / 2.0, 3.0);
    if (ia->rcutwca > t->sqmaxcut) t->sqmaxcut = ia->rcutwca;              /* :845-848 */
    if (ia->rcut > t->sqmaxcut) t->sqmaxcut = ia->rcut;
    return 1;
}

/* Topo::genParamPairs (scOOP/structures/topo.cpp:5-138) followed by
 * Topo::genTopoParams (topo.cpp:140-153). */
void orc_topo_finalize(orc_topo* t, const unsigned char* excl) {
    int i, j, k;
    double maxlength = 0;
    for (i = 0; i < ORC_MAXT; i++) {
        for (j = 0; j < ORC_MAXT; j++) {
            orc_ia* p;
            int a[2];
            if (i == j) continue;
            if (t->ia[j][j].geotype[0] == 0 || t->ia[i][i].geotype[0] == 0) continue;
            p = &t->ia[i][j];
            a[0] = i; a[1] = j;
            for (k = 0; k < 2; k++) {
                const orc_ia* s = &t->ia[a[k]][a[k]];
                p->geotype[k] = s->geotype[0];
                p->len[k] = s->len[0];
                p->half_len[k] = s->half_len[0];
                if (p->geotype[k] >= ORC_PSC && p->geotype[k] < ORC_SP) {   /* :37-46 */
                    p->pangl[k] = s->pangl[0];
                    p->panglsw[k] = s->panglsw[0];
                    p->pcangl[k] = cos(p->pangl[k] / 2.0 / 180 * PI_D);
                    p->pcanglsw[k] = cos((p->pangl[k] / 2.0 + p->panglsw[k]) / 180 * PI_D);
                    p->pcoshalfi[k] = cos((p->pangl[k] / 2.0 + p->panglsw[k]) / 2.0 / 180 * PI_D);
                    p->psinhalfi[k] = sqrt(1.0 - p->pcoshalfi[k] * p->pcoshalfi[k]);
                }
                if (is_chiral(p->geotype[k])) {                            /* :49-53 */
                    p->chiral_cos[k] = s->chiral_cos[0];
                    p->chiral_sin[k] = s->chiral_sin[0];
                }
                if (is_two(p->geotype[k])) {                               /* :55-69 */
                    p->csecpatchrot[k] = s->csecpatchrot[0];
                    p->ssecpatchrot[k] = s->ssecpatchrot[0];
                    p->pangl[k + 2] = s->pangl[2];
                    p->panglsw[k + 2] = s->panglsw[2];
                    p->pcangl[k + 2] = cos(p->pangl[k + 2] / 2.0 / 180 * PI_D);
                    p->pcanglsw[k + 2] = cos((p->pangl[k + 2] / 2.0 + p->panglsw[k + 2]) / 180 * PI_D);
                    p->pcoshalfi[k + 2] = cos((p->pangl[k + 2] / 2.0 + p->panglsw[k + 2]) / 2.0 / 180 * PI_D);
                    p->psinhalfi[k + 2] = sqrt(1.0 - p->pcoshalfi[k + 2] * p->pcoshalfi[k + 2]);
                }
            }
            p->sigma = (t->ia[i][i].sigma + t->ia[j][j].sigma) * 0.5;     /* :76-84 */
            p->epsilon = sqrt(t->ia[i][i].epsilon * t->ia[j][j].epsilon);
            p->A = 4 * p->epsilon * pow(p->sigma, 12);
            p->B = 4 * p->epsilon * pow(p->sigma, 6);
            p->pswitch = (t->ia[i][i].pswitch + t->ia[j][j].pswitch) * 0.5;
            p->pswitchINV = 1.0 / p->pswitch;
            p->rcutwca = p->sigma * pow(2.0, 1.0 / 6.0);
            p->rcutwcaSq = p->rcutwca * p->rcutwca;
            if (t->ia[i][i].parallel > 0 && t->ia[j][j].parallel > 0)      /* :85-86 */
                p->parallel = sqrt(t->ia[i][i].parallel * t->ia[j][j].parallel);
            if (t->ia[i][i].parallel < 0 && t->ia[j][j].parallel < 0)
                p->parallel = -sqrt(t->ia[i][i].parallel * t->ia[j][j].parallel);
            p->pdis = ((t->ia[i][i].pdis - t->ia[i][i].rcutwca) + (t->ia[j][j].pdis - t->ia[j][j].rcutwca)) * 0.5
                      + p->rcutwca;                                        /* :91-92 */
            if (p->geotype[0] == ORC_SPN || p->geotype[1] == ORC_SPN)      /* :95-98 */
                p->rcut = 0.0;
            else
                p->rcut = p->pswitch + p->pdis;
            p->rcutSq = p->rcut * p->rcut;
            if (p->rcutwca > t->sqmaxcut) t->sqmaxcut = p->rcutwca;        /* :111-114 */
            if (p->rcut > t->sqmaxcut) t->sqmaxcut = p->rcut;
        }
    }
    for (i = 0; i < ORC_MAXT; i++)
        for (j = 0; j < ORC_MAXT; j++)
            t->ia[i][j].exclude = excl ? (excl[i * ORC_MAXT + j] != 0) : 0;
    for (i = 0; i < ORC_MAXT; i++)                                        /* genTopoParams */
        if (maxlength < t->ia[i][i].len[0]) maxlength = t->ia[i][i].len[0];
    t->sqmaxcut += maxlength;
    t->sqmaxcut *= 1.1;
    t->maxcut = t->sqmaxcut;
    t->sqmaxcut = t->sqmaxcut * t->sqmaxcut;
}

/* ---------------------------------------------------------------- particles */
/* Particle::init, scOOP/structures/particle.cpp:3-79, with the diagonal
 * parameters ia[type][type] (Conf::partVecInit, Conf.cpp:97-102). */
void orc_part_init(const orc_topo* t, orc_part* p) {
    const orc_ia* ia = &t->ia[p->type][p->type];
    int g = ia->geotype[0];
    if (g >= ORC_SP) return;                   /* spheres are not initialised */
    if (g == ORC_SCA || g == ORC_SCN) return;
    p->dir = NORM(p->dir);
    p->patchdir[0] = NORM(ORTHO(p->patchdir[0], p->dir));
    if (g == ORC_PSC || g == ORC_CPSC || g == ORC_TPSC || g == ORC_TCPSC) {
        p->patchsides[0] = ROT(p->patchdir[0], p->dir, ia->pcoshalfi[0], ia->psinhalfi[0]);
        p->patchsides[1] = ROT(p->patchdir[0], p->dir, ia->pcoshalfi[0], -1.0 * ia->psinhalfi[0]);
    }
    if (is_two(g)) {
        orc_vec q = ROT(p->patchdir[0], p->dir, ia->csecpatchrot[0], ia->ssecpatchrot[0]);
        p->patchdir[1] = NORM(ORTHO(q, p->dir));
    }
    if (g == ORC_TPSC || g == ORC_TCPSC) {
        p->patchsides[2] = ROT(p->patchdir[1], p->dir, ia->pcoshalfi[2], ia->psinhalfi[2]);
        p->patchsides[3] = ROT(p->patchdir[1], p->dir, ia->pcoshalfi[2], -1.0 * ia->psinhalfi[2]);
    }
    if (is_chiral(g)) {
        p->chdir[0] = ROT(p->dir, p->patchdir[0], ia->chiral_cos[0], ia->chiral_sin[0]);
        p->patchsides[0] = ROT(p->patchdir[0], p->chdir[0], ia->pcoshalfi[0], ia->psinhalfi[0]);
        p->patchsides[1] = ROT(p->patchdir[0], p->chdir[0], ia->pcoshalfi[0], -1.0 * ia->psinhalfi[0]);
    }
    if (g == ORC_TCHPSC || g == ORC_TCHCPSC) {
        p->chdir[1] = ROT(p->dir, p->patchdir[1], ia->chiral_cos[0], ia->chiral_sin[0]);
        p->patchsides[2] = ROT(p->patchdir[1], p->chdir[1], ia->pcoshalfi[2], ia->psinhalfi[2]);
        p->patchsides[3] = ROT(p->patchdir[1], p->chdir[1], ia->pcoshalfi[2], -1.0 * ia->psinhalfi[2]);
    }
}

/* Cuboid::image, scOOP/structures/geometry.h:110-128.  The reference adds
 * 1.5*2^52 and reads the low 32 bits, i.e. round-half-to-even; nearbyint() in
 * the default rounding mode is the same function. */
orc_vec orc_image(const double box[3], orc_vec r1, orc_vec r2) {
    orc_vec d = V(r1.x - r2.x, r1.y - r2.y, r1.z - r2.z);
    d.x = box[0] * (d.x - nearbyint(d.x));
    d.y = box[1] * (d.y - nearbyint(d.y));
    d.z = box[2] * (d.z - nearbyint(d.z));
    return d;
}

/* ------------------------------------------------------ segment distances */
/* One clamped (s,t) solve shared by both passes of minDistSegments
 * (scOOP/mc/paire.cpp:113-168 and :192-234): returns the closest vector
 * w + sc*u - tc*v, where aa = |u|^2 is the length^2 of the s-segment. */
static orc_vec clamp_solve(orc_vec u, orc_vec v, orc_vec w, double aa, double b, double cc,
                           double d, double e, int first_pass, int* paralel) {
    double D = aa * cc - b * b;
    double sc, sN = D, sD = D, tc, tN = D, tD = D;
    if (D < 0.00000001) {
        if (first_pass) *paralel = 1;
        sN = 0.0; sD = 1.0; tN = e; tD = cc;
    } else if (first_pass) {
        sN = (b * e - cc * d);
        tN = (aa * e - b * d);
        if (sN < 0.0) { sN = 0.0; tN = e; tD = cc; }
        else if (sN > sD) { sN = sD; tN = e + b; tD = cc; }
    }
    /* second pass (paire.cpp:192-198) leaves sN=tN=D when D>=1e-8 */
    if (tN < 0.0) {
        tN = 0.0;
        if (-d < 0.0) sN = 0.0;
        else if (-d > aa) sN = sD;
        else { sN = -d; sD = aa; }
    } else if (tN > tD) {
        tN = tD;
        if ((-d + b) < 0.0) sN = 0;
        else if ((-d + b) > aa) sN = sD;
        else { sN = (-d + b); sD = aa; }
    }
    sc = (fabs(sN) < 0.00000001) ? 0.0 : sN / sD;
    tc = (fabs(tN) < 0.00000001) ? 0.0 : tN / tD;
    return V(u.x * sc + w.x - v.x * tc, u.y * sc + w.y - v.y * tc, u.z * sc + w.z - v.z * tc);
}

/* EPatch::minDistSegments, scOOP/mc/paire.cpp:86-241. */
orc_vec orc_min_dist_segments(orc_vec segA, orc_vec segB, double halfl1, double halfl2, orc_vec r_cm) {
    orc_vec u = SCL(2.0 * halfl1, segA), v = SCL(2.0 * halfl2, segB), w, vec;
    double a, b, c, d, e;
    int paralel = 0;
    w.x = segB.x * halfl2 - segA.x * halfl1 - r_cm.x;
    w.y = segB.y * halfl2 - segA.y * halfl1 - r_cm.y;
    w.z = segB.z * halfl2 - segA.z * halfl1 - r_cm.z;
    a = DOT(u, u); b = DOT(u, v); c = DOT(v, v); d = DOT(u, w); e = DOT(v, w);
    vec = clamp_solve(u, v, w, a, b, c, d, e, 1, &paralel);
    if (paralel) {
        /* roles of the two segments swapped (paire.cpp:170-238) */
        orc_vec vec2;
        int dummy = 0;
        w.x = segA.x * halfl1 - segB.x * halfl2 + r_cm.x;
        w.y = segA.y * halfl1 - segB.y * halfl2 + r_cm.y;
        w.z = segA.z * halfl1 - segB.z * halfl2 + r_cm.z;
        d = DOT(v, w);
        e = DOT(u, w);
        /* in the reference the swapped pass keeps D = a*c - b*b with tD = a
         * (paire.cpp:184-198), i.e. the s-segment is v with length^2 c */
        vec2 = clamp_solve(v, u, w, c, b, a, d, e, 0, &dummy);
        if (DOT(vec2, vec2) < DOT(vec, vec)) return vec2;
    }
    return vec;
}

/* ------------------------------------------------------- patch machinery */
/* fanglScale, scOOP/mc/paire.h:12-17 */
static inline double fangl_scale(double a, double pcangl, double pcanglsw) {
    if (a <= pcanglsw) return 0.0;
    return (a >= pcangl) ? 1.0 : (0.5 - ((pcanglsw + pcangl) * 0.5 - a) / (pcangl - pcanglsw));
}
/* EPatch::scparallel, paire.h:77-84 */
static inline double sc_parallel(double epar, orc_vec d1, orc_vec d2) {
    double cosa = DOT(d1, d2);
    if ((epar > 0 && cosa > 0) || (epar < 0 && cosa < 0)) return 1.0 + epar * cosa;
    return 1.0;
}

typedef struct { orc_vec dir; orc_vec sides[2]; } opatch;

/* EPatch::testIntrPatch, paire.h:86-101 (ints[] uses 0.0 as "empty") */
static void test_intr_patch(orc_vec dir, orc_vec patchdir, orc_vec vec, double cospatch, double ti, double* ints) {
    vec = PERP(vec, dir);
    if (DOT(patchdir, vec) >= cospatch * SIZE(vec)) {
        if (ints[0] == 0) { ints[0] = ti; return; }
        if (ints[1] == 0 && ints[0] != ti) { ints[1] = ti; return; }
    }
}

/* EPatch::scToInfiIntr, paire.h:103-119 */
static void sc_to_infi_intr(orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, orc_vec r_cm, double pcanglsw,
                            double halfl1, double halfl2, double* ints, double x1) {
    orc_vec v1;
    double e;
    if ((x1 >= halfl2) || (x1 <= -halfl2)) return;
    v1 = V(p2Dir.x * x1 - r_cm.x, p2Dir.y * x1 - r_cm.y, p2Dir.z * x1 - r_cm.z);
    e = DOT(p1Dir, v1);
    if ((e >= halfl1) || (e <= -halfl1)) return;
    test_intr_patch(p1Dir, p1P->dir, v1, pcanglsw, x1, ints);
}

/* EPatch::testIntrAtC, paire.h:121-148 */
static void test_intr_at_c(orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, orc_vec r_cm, double pcanglsw,
                           double rcutSq, double halfl1, double halfl2, double* ints) {
    orc_vec v1 = CROSS(NEG(r_cm), p1Dir), v2 = CROSS(p2Dir, p1Dir);
    double a = DOT(v2, v2), b = 2 * DOT(v1, v2), c = -rcutSq + DOT(v1, v1), d = b * b - 4 * a * c;
    if (d >= 0) {
        d = sqrt(d);
        a = 0.5 / a;
        sc_to_infi_intr(p1Dir, p2Dir, p1P, r_cm, pcanglsw, halfl1, halfl2, ints, (-b + d) * a);
        if (d > 0) sc_to_infi_intr(p1Dir, p2Dir, p1P, r_cm, pcanglsw, halfl1, halfl2, ints, (-b - d) * a);
    }
}

/* EPatch::findIntersectPlaneUni, paire.h:150-181 */
static int find_intersect_plane(orc_vec dirA, orc_vec dirB, double halfl, orc_vec r_cm, orc_vec w_vec,
                                double cospatch, double* ti, double* c, double* d) {
    orc_vec nplane = CROSS(dirA, w_vec), dv;
    double a = DOT(nplane, dirB);
    *c = 1.0; *d = 1.0;
    if (a == 0.0) return 0;
    *ti = DOT(nplane, r_cm) / a;
    if ((*ti > halfl) || (*ti < -halfl)) return 0;
    dv = V(*ti * dirB.x - r_cm.x, *ti * dirB.y - r_cm.y, *ti * dirB.z - r_cm.z);
    *c = DOT(dv, w_vec);
    if (*c * cospatch < 0) return 0;
    *d = fabs(DOT(dv, dirA)) - halfl;
    return 1;
}

/* Psc::scToEndSpIntr, paire.h:588-602 */
static void sc_to_end_sp_intr(orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, orc_vec r_cm, double pcanglsw,
                              double halfl1, double halfl2, double* ints, double x1) {
    orc_vec v1;
    double e;
    if ((x1 >= halfl2) || (x1 <= -halfl2)) return;
    v1 = V(p2Dir.x * x1 - r_cm.x, p2Dir.y * x1 - r_cm.y, p2Dir.z * x1 - r_cm.z);
    e = DOT(p1Dir, v1);
    if ((e >= halfl1) || (e <= -halfl1)) test_intr_patch(p1Dir, p1P->dir, v1, pcanglsw, x1, ints);
}

/* Psc::calcIntersections, paire.h:605-623 */
static void calc_intersections(orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, orc_vec r_cm, double* ints,
                               double pcanglsw, double halfl1, double halfl2, double b, double c) {
    double d = b * b - 4 * c;
    if (d >= 0) {
        d = sqrt(d);
        sc_to_end_sp_intr(p1Dir, p2Dir, p1P, r_cm, pcanglsw, halfl1, halfl2, ints, (-b + d) * 0.5);
        if (d > 0) sc_to_end_sp_intr(p1Dir, p2Dir, p1P, r_cm, pcanglsw, halfl1, halfl2, ints, (-b - d) * 0.5);
    }
}

/* Psc::testIntrA, paire.h:625-652 */
static void test_intr_a(orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, orc_vec r_cm, double pcanglsw,
                        double rcutSq, double halfl1, double halfl2, double* ints) {
    orc_vec v1 = V(p2Dir.x * halfl2 - r_cm.x, p2Dir.y * halfl2 - r_cm.y, p2Dir.z * halfl2 - r_cm.z), v2;
    double a = DOT(v1, p1Dir), b, c, d;
    v2 = V(v1.x - p1Dir.x * a, v1.y - p1Dir.y * a, v1.z - p1Dir.z * a);
    b = DOT(v2, v2);
    d = fabs(a) - halfl1;
    c = (d <= 0) ? b : d * d + b;
    if (c < rcutSq) test_intr_patch(p1Dir, p1P->dir, v1, pcanglsw, halfl2, ints);
}

/* Psc::pscIntersect, paire.h:499-583 */
static int psc_intersect(orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, orc_vec r_cm, double* in1, double* in2,
                         double pcanglsw, double rcutSq, double halfl1, double halfl2) {
    double c, d, ti, disti, ints[2] = {0, 0};
    orc_vec v1, v2;
    if (find_intersect_plane(p1Dir, p2Dir, halfl2, r_cm, p1P->sides[0], pcanglsw, &ti, &c, &d)) {
        disti = (d <= 0) ? c * c : d * d + c * c;
        if (disti <= rcutSq) ints[0] = ti;
    }
    if (find_intersect_plane(p1Dir, p2Dir, halfl2, r_cm, p1P->sides[1], pcanglsw, &ti, &c, &d)) {
        disti = (d <= 0) ? c * c : d * d + c * c;
        if (disti <= rcutSq) {
            if (ints[0] == 0.0) ints[0] = ti;
            else if (ti != ints[0]) ints[1] = ti;
        }
    }
    if (ints[1] != 0.0) { *in1 = ints[0]; *in2 = ints[1]; return 2; }
    test_intr_at_c(p1Dir, p2Dir, p1P, r_cm, pcanglsw, rcutSq, halfl1, halfl2, ints);
    if (ints[1] == 0.0) {
        v1 = V(p1Dir.x * halfl1 - r_cm.x, p1Dir.y * halfl1 - r_cm.y, p1Dir.z * halfl1 - r_cm.z);
        v2 = V(-p1Dir.x * halfl1 - r_cm.x, -p1Dir.y * halfl1 - r_cm.y, -p1Dir.z * halfl1 - r_cm.z);
        calc_intersections(p1Dir, p2Dir, p1P, r_cm, ints, pcanglsw, halfl1, halfl2, 2.0 * DOT(v1, p2Dir), DOT(v1, v1) - rcutSq);
        calc_intersections(p1Dir, p2Dir, p1P, r_cm, ints, pcanglsw, halfl1, halfl2, 2.0 * DOT(v2, p2Dir), DOT(v2, v2) - rcutSq);
    } else {
        *in1 = ints[0]; *in2 = ints[1];
        return 2;
    }
    if (ints[1] == 0.0) {
        test_intr_a(p1Dir, p2Dir, p1P, r_cm, pcanglsw, rcutSq, halfl1, halfl2, ints);
        if (ints[1] == 0.0) test_intr_a(p1Dir, p2Dir, p1P, r_cm, pcanglsw, rcutSq, halfl1, -halfl2, ints);
    }
    *in1 = ints[0]; *in2 = ints[1];
    return (ints[1] == 0.0) ? 0 : 2;
}

/* CPsc::cpscIntersect, paire.h:687-826 */
static int cpsc_intersect(orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, orc_vec r_cm, double* in1, double* in2,
                          double pcanglsw, double rcutSq, double halfl1, double halfl2) {
    double a, b, c, d, x1, ti, ints[2] = {0, 0};
    orc_vec v1, v2;
    if (find_intersect_plane(p1Dir, p2Dir, halfl2, r_cm, p1P->sides[0], pcanglsw, &ti, &c, &d)) {
        if (d <= 0 && c * c <= rcutSq) ints[0] = ti;
    }
    if (find_intersect_plane(p1Dir, p2Dir, halfl2, r_cm, p1P->sides[1], pcanglsw, &ti, &c, &d)) {
        if (d <= 0 && c * c <= rcutSq) {
            if (ints[0] == 0.0) ints[0] = ti;
            else if (ti != ints[0]) ints[1] = ti;
        }
    }
    if (ints[1] != 0.0) { *in1 = ints[0]; *in2 = ints[1]; return 2; }
    test_intr_at_c(p1Dir, p2Dir, p1P, r_cm, pcanglsw, rcutSq, halfl1, halfl2, ints);
    if (ints[1] == 0.0) {
        a = DOT(p1Dir, p2Dir);
        if (a != 0.0) {
            int cap;
            for (cap = 0; cap < 2; cap++) {       /* plane caps at +halfl1 then -halfl1 (:741-776) */
                if (cap == 0) v1 = V(r_cm.x + halfl1 * p1Dir.x, r_cm.y + halfl1 * p1Dir.y, r_cm.z + halfl1 * p1Dir.z);
                else v1 = V(r_cm.x - halfl1 * p1Dir.x, r_cm.y - halfl1 * p1Dir.y, r_cm.z - halfl1 * p1Dir.z);
                x1 = DOT(p1Dir, v1) / a;
                if (!((x1 > halfl2) || (x1 < -halfl2))) {
                    v2 = V(x1 * p2Dir.x - v1.x, x1 * p2Dir.y - v1.y, x1 * p2Dir.z - v1.z);
                    b = DOT(v2, v2);
                    if (!(b > rcutSq)) test_intr_patch(p1Dir, p1P->dir, v2, pcanglsw, x1, ints);
                }
            }
        }
    } else {
        *in1 = ints[0]; *in2 = ints[1];
        return 2;
    }
    if (ints[1] == 0.0) {                          /* :786-821 */
        v1 = V(p2Dir.x * halfl2 - r_cm.x, p2Dir.y * halfl2 - r_cm.y, p2Dir.z * halfl2 - r_cm.z);
        a = DOT(v1, p1Dir);
        v2 = V(v1.x - p1Dir.x * a, v1.y - p1Dir.y * a, v1.z - p1Dir.z * a);
        b = DOT(v2, v2);
        d = fabs(a) - halfl1;
        if (d <= 0 && b < rcutSq) test_intr_patch(p1Dir, p1P->dir, v1, pcanglsw, halfl2, ints);
        if (ints[1] == 0.0) {
            v1 = V(-p2Dir.x * halfl2 - r_cm.x, -p2Dir.y * halfl2 - r_cm.y, -p2Dir.z * halfl2 - r_cm.z);
            a = DOT(v1, p1Dir);
            v2 = V(v1.x - p1Dir.x * a, v1.y - p1Dir.y * a, v1.z - p1Dir.z * a);
            b = DOT(v2, v2);
            d = fabs(a) - halfl1;
            if (d <= 0 && b < rcutSq) test_intr_patch(p1Dir, p1P->dir, v1, pcanglsw, -1.0 * halfl2, ints);
        }
    }
    *in1 = ints[0]; *in2 = ints[1];
    return (ints[1] == 0.0) ? 0 : 2;
}

/* EPatch::atrE, scOOP/mc/paire.cpp:244-301 */
static double atr_e(const orc_ia* ia, orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P, const opatch* p2P,
                    orc_vec r_cm, int pn1, int pn2, double S1, double S2, double T1, double T2) {
    double v1 = fabs(S1 - S2), v2 = fabs(T1 - T2), f0 = 0.5 * (v1 + v2);
    orc_vec a1 = SCL((S1 + S2) * 0.5, p1Dir), a2 = SCL((T1 + T2) * 0.5, p2Dir), vi, vm, w;
    double e, ndist, a, f1, f2, paral;
    vi = V(a2.x - a1.x - r_cm.x, a2.y - a1.y - r_cm.y, a2.z - a1.z - r_cm.z);
    vm = orc_min_dist_segments(p1Dir, p2Dir, v1, v2, vi);
    ndist = sqrt(DOT(vm, vm));
    if (ndist < ia->pdis) e = -ia->epsilon;
    else {
        e = cos(PIH_D * (ndist - ia->pdis) / ia->pswitch);
        e *= -e * ia->epsilon;
    }
    w = PERP(vi, p1Dir);
    a = DOT(w, p1P->dir) / SIZE(w);
    f1 = fangl_scale(a, ia->pcangl[0 + 2 * pn1], ia->pcanglsw[0 + 2 * pn1]);
    w = PERP(SCL(-1.0, vi), p2Dir);
    a = DOT(w, p2P->dir) / SIZE(w);
    f2 = fangl_scale(a, ia->pcangl[1 + 2 * pn2], ia->pcanglsw[1 + 2 * pn2]);
    paral = 1.0;
    if (ia->parallel != 0.0) paral = sc_parallel(ia->parallel, p1Dir, p2Dir);
    e *= f0 * f1 * f2 * paral;
    return e;
}

/* Patch-pair functor: Psc::operator() (paire.h:469-484), CPsc::operator()
 * (:658-672), PscCPsc::operator() (:864-889).  mode: 0 PSC-PSC, 1 CPSC-CPSC,
 * 2 mixed. */
static double patch_pair(int mode, const orc_ia* ia, orc_vec p1Dir, orc_vec p2Dir, const opatch* p1P,
                         const opatch* p2P, orc_vec r_cm, int pn1, int pn2) {
    double T1 = 0, T2 = 0, S1 = 0, S2 = 0;
    orc_vec mr = SCL(-1.0, r_cm);
    int first_psc = (mode == 0) || (mode == 2 && is_psc(ia->geotype[0]));
    int second_psc = (mode == 0) || (mode == 2 && !is_psc(ia->geotype[0]));
    int r;
    r = first_psc ? psc_intersect(p1Dir, p2Dir, p1P, r_cm, &T1, &T2, ia->pcanglsw[2 * pn1], ia->rcutSq, ia->half_len[0], ia->half_len[1])
                  : cpsc_intersect(p1Dir, p2Dir, p1P, r_cm, &T1, &T2, ia->pcanglsw[2 * pn1], ia->rcutSq, ia->half_len[0], ia->half_len[1]);
    if (r < 2) return 0.0;
    r = second_psc ? psc_intersect(p2Dir, p1Dir, p2P, mr, &S1, &S2, ia->pcanglsw[2 * pn2 + 1], ia->rcutSq, ia->half_len[1], ia->half_len[0])
                   : cpsc_intersect(p2Dir, p1Dir, p2P, mr, &S1, &S2, ia->pcanglsw[2 * pn2 + 1], ia->rcutSq, ia->half_len[1], ia->half_len[0]);
    if (r < 2) return 0.0;
    return atr_e(ia, p1Dir, p2Dir, p1P, p2P, r_cm, pn1, pn2, S1, S2, T1, T2);
}

/* WcaTruncSq, paire.h:385-393 */
static inline double wca_sq(double distSq, const orc_ia* ia) {
    if (distSq > ia->rcutwcaSq) return 0.0;
    return ia->epsilon + ia->A * pow(distSq, -6) - ia->B * pow(distSq, -3);
}
/* WcaTrunc, paire.h:376-383 */
static inline double wca(double dist, const orc_ia* ia) {
    if (dist > ia->rcutwca) return 0.0;
    return ia->A * pow(dist, -12) - ia->B * pow(dist, -6) + ia->epsilon;
}
/* WcaCos2Taylor, paire.h:416-439 */
static double wca_cos2_taylor(double dist, const orc_ia* ia) {
    double e;
    if (dist > ia->rcut || ia->epsilon == 0.0 || ia->exclude) return 0.0;
    if (dist > ia->pdis) {
        e = PIH_D * (dist - ia->pdis) * ia->pswitchINV;
        e *= e;
        e = (1 - e + e * e * (1.0 / 3.0) - e * e * e * (2.0 / 45.0) + e * e * e * e * (1.0 / 315.0)
             - e * e * e * e * e * (2.0 / 14175.0) + e * e * e * e * e * e * (2.0 / 467775.0)
             - e * e * e * e * e * e * e * (4.0 / 42567525) + e * e * e * e * e * e * e * e * (1.0 / 638512875))
            * -ia->epsilon;
    } else {
        e = -ia->epsilon;
    }
    if (dist > ia->rcutwca) return e;
    return ia->A * pow(dist, -12) - ia->B * pow(dist, -6);
}

/* SpheroCylinder::operator(), paire.h:1122-1196; abE: its bond + angle terms */
static double sc_sc(int mode, const orc_ia* ia, const orc_part* p1, const orc_part* p2, orc_vec r_cm, double abE) {
    double distSq, rep, atr = 0.0;
    orc_vec d1, d2;
    int fCH, sCH, fT, sT;
    opatch P10, P11, P20, P21;
    {
        orc_vec dv = orc_min_dist_segments(p1->dir, p2->dir, ia->half_len[0], ia->half_len[1], r_cm);
        distSq = DOT(dv, dv);
    }
    rep = wca_sq(distSq, ia);
    if (mode == 3 || mode == 4) {                  /* Scn / Sca patch functors */
        if (!(distSq > ia->rcutSq || ia->epsilon == 0.0 || ia->exclude) && mode == 4) {
            /* Sca::operator(), paire.h:845-852: WCA of the same distance again */
            double dist = sqrt(distSq);
            if (!(dist > ia->rcutwca)) atr = ia->A * pow(dist, -12) - ia->B * pow(dist, -6) + ia->epsilon;
        }
        return abE + rep + atr;
    }
    if (distSq > ia->rcutSq || ia->epsilon == 0.0 || ia->exclude) return abE + rep + atr;
    /* getGeoTypes of Psc (:486-497), CPsc (:674-685), PscCPsc (:891-906) */
    if (mode == 0) {
        fCH = ia->geotype[0] == ORC_CHPSC || ia->geotype[0] == ORC_TCHPSC;
        sCH = ia->geotype[1] == ORC_CHPSC || ia->geotype[1] == ORC_TCHPSC;
        fT = ia->geotype[0] == ORC_TPSC || ia->geotype[0] == ORC_TCHPSC;
        sT = ia->geotype[1] == ORC_TPSC || ia->geotype[1] == ORC_TCHPSC;
    } else if (mode == 1) {
        fCH = ia->geotype[0] == ORC_CHCPSC || ia->geotype[0] == ORC_TCHCPSC;
        sCH = ia->geotype[1] == ORC_CHCPSC || ia->geotype[1] == ORC_TCHCPSC;
        fT = ia->geotype[0] == ORC_TCPSC || ia->geotype[0] == ORC_TCHCPSC;
        sT = ia->geotype[1] == ORC_TCPSC || ia->geotype[1] == ORC_TCHCPSC;
    } else {
        fCH = is_chiral(ia->geotype[0]); sCH = is_chiral(ia->geotype[1]);
        fT = is_two(ia->geotype[0]); sT = is_two(ia->geotype[1]);
    }
    P10.dir = p1->patchdir[0]; P10.sides[0] = p1->patchsides[0]; P10.sides[1] = p1->patchsides[1];
    P11.dir = p1->patchdir[1]; P11.sides[0] = p1->patchsides[2]; P11.sides[1] = p1->patchsides[3];
    P20.dir = p2->patchdir[0]; P20.sides[0] = p2->patchsides[0]; P20.sides[1] = p2->patchsides[1];
    P21.dir = p2->patchdir[1]; P21.sides[0] = p2->patchsides[2]; P21.sides[1] = p2->patchsides[3];
    d1 = fCH ? p1->chdir[0] : p1->dir;
    d2 = sCH ? p2->chdir[0] : p2->dir;
    atr = patch_pair(mode, ia, d1, d2, &P10, &P20, r_cm, 0, 0);
    if (fT) atr += patch_pair(mode, ia, fCH ? p1->chdir[1] : p1->dir, d2, &P11, &P20, r_cm, 1, 0);
    if (sT) atr += patch_pair(mode, ia, d1, sCH ? p2->chdir[1] : p2->dir, &P10, &P21, r_cm, 0, 1);
    if (fT && sT) atr += patch_pair(mode, ia, fCH ? p1->chdir[1] : p1->dir, sCH ? p2->chdir[1] : p2->dir, &P11, &P21, r_cm, 1, 1);
    return abE + rep + atr;
}

/* MixSpSc::closestDist, paire.h:1077-1093 */
static double sp_sc_dist(orc_vec r_cm, orc_vec dir1, double halfl, double* contt, orc_vec* distvec) {
    double d;
    *contt = DOT(dir1, r_cm);
    d = -halfl;
    if (*contt >= halfl) d = halfl;
    else if (*contt > -halfl) d = *contt;
    *distvec = V(-r_cm.x + dir1.x * d, -r_cm.y + dir1.y * d, -r_cm.z + dir1.z * d);
    return DOT(*distvec, *distvec);
}

/* ScaSpa / PscSpa::operator(), paire.h:914-940, :946-979.  sca: no angular factor */
static double patch_to_sphere(int sca, double dist, double contt, orc_vec distvec, const orc_ia* ia, orc_vec p1Dir,
                              orc_vec patchdir) {
    double e, a = 0, b, f0, halfl;
    if (dist < ia->pdis) e = -ia->epsilon;
    else {
        e = cos(PIH_D * (dist - ia->pdis) / ia->pswitch);
        e *= -e * ia->epsilon;
    }
    if (!sca) {
        orc_vec w = PERP(distvec, p1Dir);
        a = DOT(w, patchdir) / SIZE(w);
    }
    halfl = ia->half_len[0];
    b = sqrt(ia->rcutSq - dist * dist);
    f0 = (contt + b > halfl) ? halfl : contt + b;
    if (contt - b < -halfl) f0 -= -halfl;
    else f0 -= contt - b;
    if (sca) return e * f0;
    return e * (fangl_scale(a, ia->pcangl[0], ia->pcanglsw[0]) * f0);
}

/* MixSpSc::operator(), paire.h:1022-1067.  kind: 0 ScaSpa, 1 PscSpa,
 * 2 CPscSpa; abE: its bond + angle terms */
static double sp_sc(int kind, const orc_topo* t, const orc_part* p1, const orc_part* p2, orc_vec r_cm, double abE) {
    const orc_ia* ia12 = &t->ia[p1->type][p2->type];
    int isp1 = ia12->geotype[0] < ORC_SP;
    const orc_part* spc = isp1 ? p1 : p2;
    const orc_ia* ia = isp1 ? ia12 : &t->ia[p2->type][p1->type];
    orc_vec rr = isp1 ? r_cm : SCL(-1.0, r_cm), distvec;
    double contt = 0.0, distSq, dist, rep = 0.0, atr = 0.0;
    int chiral = 0, sec = 0, far;
    distSq = sp_sc_dist(rr, spc->dir, ia->half_len[0], &contt, &distvec);
    if (distSq < ia->rcutwcaSq) rep = wca_sq(distSq, ia12);
    far = (kind == 2) && ((contt > ia->half_len[0]) || (contt < -ia->half_len[0]));
    if (!(distSq > ia->rcutSq || ia->epsilon == 0.0 || ia->exclude || far)) {
        if (kind == 1) {
            chiral = ia->geotype[0] == ORC_CHPSC || ia->geotype[0] == ORC_TCHPSC;
            sec = ia->geotype[0] == ORC_TPSC || ia->geotype[0] == ORC_TCHPSC;
        } else if (kind == 2) {
            chiral = ia->geotype[0] == ORC_CHCPSC || ia->geotype[0] == ORC_TCHCPSC;
            sec = ia->geotype[0] == ORC_TCPSC || ia->geotype[0] == ORC_TCHCPSC;
        }
        if (chiral) distSq = sp_sc_dist(rr, spc->chdir[0], ia->half_len[0], &contt, &distvec);
        dist = sqrt(distSq);
        if (dist < ia->rcut)
            atr = patch_to_sphere(kind == 0, dist, contt, distvec, ia, chiral ? spc->chdir[0] : spc->dir, spc->patchdir[0]);
        if (sec) {
            distSq = sp_sc_dist(rr, chiral ? spc->chdir[1] : spc->dir, ia->half_len[0], &contt, &distvec);
            dist = sqrt(distSq);
            if (dist < ia->rcut)
                atr += patch_to_sphere(kind == 0, dist, contt, distvec, ia, chiral ? spc->chdir[1] : spc->dir, spc->patchdir[1]);
        }
    }
    return abE + rep + atr;
}

/* ------------------------------------------------------- bonds and angles */
/* ConList (scOOP/structures/Conf.h:17-31) with particle indices in place of
 * pointers (-1: none): [0] previous, [1] next, [2] second previous, [3]
 * second next in the chain. */
typedef struct {
    int empty;
    int cl[4];
    double sp, mod[2], c[2], eq[2];
} oconlist;

int orc_topo_set_mol(orc_topo* t, int m, int size, const double* p) {
    orc_mol* q;
    if (m < 0 || m >= ORC_MAXMT || size < 1) return 0;
    q = &t->mol[m];
    q->size = size;
    q->bond1c = p[0]; q->bond1eq = p[1]; q->bond2c = p[2]; q->bond2eq = p[3];
    q->bonddc = p[4]; q->bonddeq = p[5]; q->bondhc = p[6]; q->bondheq = p[7];
    q->angle1c = p[8]; q->angle1eq = p[9]; q->angle2c = p[10]; q->angle2eq = p[11];
    if (m + 1 > t->nmol) t->nmol = m + 1;
    return 1;
}

void orc_topo_set_chain(orc_topo* t, int n, const int* chain) {
    t->nchain = chain ? n : 0;
    t->chain = chain;
}

/* ParticleVector::getConlist, Conf.h:90-147 */
static oconlist conlist_of(const orc_topo* t, const orc_part* p) {
    oconlist c;
    const orc_mol* m;
    int i = p->idx, pos = p->molpos;
    memset(&c, 0, sizeof(c));
    c.empty = 1;
    c.cl[0] = c.cl[1] = c.cl[2] = c.cl[3] = -1;
    if (p->moltype < 0) return c;
    m = &t->mol[p->moltype];
    if (m->size == 1) return c;                                         /* isAtomic, :91-92 */
    if (m->bond1c >= 0.0 || m->bonddc >= 0.0 || m->bondhc >= 0.0) {    /* :99-131 */
        if (pos > 0) c.cl[0] = i - 1;
        if (pos + 1 < m->size) c.cl[1] = i + 1;
        if (m->bond1c >= 0.0) {
            c.eq[0] = m->bond1eq; c.c[0] = m->bond1c;
            c.mod[0] = 0.0; c.mod[1] = 0.0; c.sp = 0.0;
        }
        if (m->bonddc >= 0.0) {
            c.eq[0] = 0.0; c.c[0] = m->bonddc;
            c.mod[0] = m->bonddeq; c.mod[1] = 0.0; c.sp = 0.0;
        }
        if (m->bondhc >= 0.0) {
            c.eq[0] = 0.0; c.c[0] = m->bondhc;
            c.mod[0] = m->bondheq; c.mod[1] = m->bondheq; c.sp = m->bondheq;
        }
        c.empty = 0;
    }
    if (m->bond2c >= 0.0) {                                             /* :132-143 */
        if (pos > 1) c.cl[2] = i - 2;
        if (pos + 2 < m->size) c.cl[3] = i + 2;
        c.eq[1] = m->bond2eq; c.c[1] = m->bond2c;
        c.empty = 0;
    }
    return c;
}

/* EBond::harmonicPotential, paire.h:202-204 */
static inline double harmonic(double a, double eq, double k) { return k * (a - eq) * (a - eq) * 0.5; }

/* HarmonicSp, paire.h:221-236: centre distance for both neighbour shells */
static double bond_sp(const oconlist* c, const orc_part* p2, double dist) {
    if (p2->idx == c->cl[1] || p2->idx == c->cl[0]) return harmonic(dist, c->eq[0], c->c[0]);
    if (p2->idx == c->cl[2] || p2->idx == c->cl[3]) return harmonic(dist, c->eq[1], c->c[1]);
    return 0.0;
}

/* HarmonicSc, paire.h:239-280: nearest neighbours bond end to end (particle
 * i's -dir end to i+1's +dir end; spheres bond at sp along dir), second
 * neighbours by centre distance */
static double bond_sc(const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2,
                      const oconlist* c, double dist) {
    const orc_ia* ia = &t->ia[p1->type][p2->type];
    if (p2->idx == c->cl[1] || p2->idx == c->cl[0]) {
        int prev = p2->idx == c->cl[0];
        double h1, h2;
        orc_vec v1, v2, b;
        if (ia->geotype[0] < ORC_SP) h1 = (ia->half_len[0] + c->mod[prev ? 0 : 1]) * (prev ? 1.0 : -1.0);
        else h1 = c->sp;
        if (ia->geotype[1] < ORC_SP) h2 = (ia->half_len[1] + c->mod[prev ? 1 : 0]) * (prev ? -1.0 : 1.0);
        else h2 = c->sp;
        v1 = V(p1->pos.x + (p1->dir.x * h1 / box[0]), p1->pos.y + (p1->dir.y * h1 / box[1]),
               p1->pos.z + (p1->dir.z * h1 / box[2]));
        v2 = V(p2->pos.x + (p2->dir.x * h2 / box[0]), p2->pos.y + (p2->dir.y * h2 / box[1]),
               p2->pos.z + (p2->dir.z * h2 / box[2]));
        b = orc_image(box, v1, v2);
        return harmonic(sqrt(DOT(b, b)), c->eq[0], c->c[0]);
    }
    if (p2->idx == c->cl[2] || p2->idx == c->cl[3]) return harmonic(dist, c->eq[1], c->c[1]);
    return 0.0;
}

/* AngleSc::angleEnergyAngle2, paire.h:335-350: angle between the two patch
 * directions projected on the planes normal to dir_a x dir_b */
static double angle2_of(const orc_part* a, const orc_part* b) {
    orc_vec ax = CROSS(a->dir, b->dir), x1 = CROSS(a->dir, ax), x2 = CROSS(b->dir, ax);
    double v1x = DOT(x1, a->patchdir[0]), v1y = DOT(ax, a->patchdir[0]);
    double v2x = DOT(x2, b->patchdir[0]), v2y = DOT(ax, b->patchdir[0]);
    return acos((v1x * v2x + v1y * v2y) / sqrt((v1x * v1x + v1y * v1y) * (v2x * v2x + v2y * v2y)));
}

/* AngleSc::operator(), paire.h:283-333 */
static double angle_sc(const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2,
                       const oconlist* c) {
    const orc_ia* ia = &t->ia[p1->type][p2->type];
    const orc_mol* m;
    double e = 0.0;
    int near = p2->idx == c->cl[0] || p2->idx == c->cl[1];
    if (p1->moltype < 0 || !near) return 0.0;
    m = &t->mol[p1->moltype];
    if (m->angle1c >= 0) {
        int prev = p2->idx == c->cl[0];
        orc_vec v1, v2;
        if (ia->geotype[0] < ORC_SP) v1 = p1->dir;
        else {
            double h = ia->half_len[1] * (prev ? -1.0 : 1.0);
            v1 = V(p2->pos.x + p2->dir.x * h / box[0], p2->pos.y + p2->dir.y * h / box[1],
                   p2->pos.z + p2->dir.z * h / box[2]);
            v1 = NORM(orc_image(box, v1, p1->pos));
        }
        if (ia->geotype[1] < ORC_SP) v2 = p2->dir;
        else {
            double h = ia->half_len[0] * (prev ? 1.0 : -1.0);
            v2 = V(p1->pos.x + p1->dir.x * h / box[0], p1->pos.y + p1->dir.y * h / box[1],
                   p1->pos.z + p1->dir.z * h / box[2]);
            v2 = NORM(orc_image(box, v2, p2->pos));
        }
        e += harmonic(acos(DOT(v1, v2)), m->angle1eq, m->angle1c);
    }
    if (m->angle2c >= 0 && ia->geotype[0] < ORC_SP && ia->geotype[1] < ORC_SP) {
        double a = (p2->idx == c->cl[0]) ? angle2_of(p2, p1) : angle2_of(p1, p2);
        e += harmonic(a, m->angle2eq, m->angle2c);
    }
    return e;
}

/* the functor of a type pair, PairE::initIntFCE (paire.cpp:6-80): later
 * rules override earlier ones.  0..4 SpheroCylinder (Psc, CPsc, PscCPsc,
 * Scn, Sca), 10 Sphere<WcaTrunc>, 11 Sphere<WcaCos2Taylor>, 20..22 MixSpSc
 * (ScaSpa, PscSpa, CPscSpa), -1 EBasic */
static int functor_kind(const orc_ia* ia) {
    int g = ia->geotype[0], o = ia->geotype[1];
    int kind = -1;
    if ((is_cpsc(g) && is_psc(o)) || (is_psc(g) && is_cpsc(o))) kind = 2;
    if (is_cpsc(g) && is_cpsc(o)) kind = 1;
    if (is_psc(g) && is_psc(o)) kind = 0;
    if (g == ORC_SCN && o == ORC_SCN) kind = 3;
    if (g == ORC_SCA && o == ORC_SCA) kind = 4;
    if (g == ORC_SPN || o == ORC_SPN) kind = 10;
    if (g == ORC_SPA && o == ORC_SPA) kind = 11;
    if ((g == ORC_SCA && o == ORC_SPA) || (g == ORC_SPA && o == ORC_SCA)) kind = 20;
    if ((is_psc(g) && (o == ORC_SPA || o == ORC_SPN)) || ((g == ORC_SPA || g == ORC_SPN) && is_psc(o))) kind = 21;
    if ((is_cpsc(g) && (o == ORC_SPA || o == ORC_SPN)) || ((g == ORC_SPA || g == ORC_SPN) && is_cpsc(o))) kind = 22;
    return kind;
}

/* bond + angle terms of the functor: HarmonicSc + AngleSc for
 * SpheroCylinder and MixSpSc, HarmonicSp for Sphere, none for EBasic */
static double bonds_of(int kind, const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2,
                       const oconlist* c, double dist) {
    double ab;
    if (c->empty) return 0.0;
    if (kind == 10 || kind == 11) return bond_sp(c, p2, dist);
    if (kind < 0) return 0.0;
    ab = bond_sc(t, box, p1, p2, c, dist);
    ab += angle_sc(t, box, p1, p2, c);
    return ab;
}

/* the functor with its bond + angle part abE */
static double functor(int kind, const orc_topo* t, const orc_ia* ia, const orc_part* p1, const orc_part* p2,
                      orc_vec r_cm, double dist, double abE) {
    switch (kind) {
    case 0: case 1: case 2: case 3: case 4: return sc_sc(kind, ia, p1, p2, r_cm, abE);
    case 10: return abE + wca(dist, ia);
    case 11: return abE + wca_cos2_taylor(dist, ia);
    case 20: return sp_sc(0, t, p1, p2, r_cm, abE);
    case 21: return sp_sc(1, t, p1, p2, r_cm, abE);
    case 22: return sp_sc(2, t, p1, p2, r_cm, abE);
    default: return 0.0;   /* EBasic: "interaction not programmed" (paire.h:207-212) */
    }
}

/* PairE::operator(), paire.h:1209-1220, with p1's connectivity list
 * (TotalEMatrix passes getConlist(part1)): the early exit holds only for an
 * empty list. */
double orc_pair_energy(const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2) {
    orc_vec r_cm = orc_image(box, p1->pos, p2->pos);
    double dotrcm = DOT(r_cm, r_cm), dist;
    const orc_ia* ia = &t->ia[p1->type][p2->type];
    oconlist c = conlist_of(t, p1);
    int kind;
    if (dotrcm > t->sqmaxcut && c.empty) return 0.0;
    dist = sqrt(dotrcm);
    kind = functor_kind(ia);
    return functor(kind, t, ia, p1, p2, r_cm, dist, bonds_of(kind, t, box, p1, p2, &c, dist));
}

/* non-bonded part alone (zero beyond sqmaxcut) */
static double pair_nb(const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2) {
    orc_vec r_cm = orc_image(box, p1->pos, p2->pos);
    double dotrcm = DOT(r_cm, r_cm);
    const orc_ia* ia = &t->ia[p1->type][p2->type];
    if (dotrcm > t->sqmaxcut) return 0.0;
    return functor(functor_kind(ia), t, ia, p1, p2, r_cm, sqrt(dotrcm), 0.0);
}

double orc_bond_energy(const orc_topo* t, const double box[3], const orc_part* p1, const orc_part* p2) {
    orc_vec r_cm = orc_image(box, p1->pos, p2->pos);
    oconlist c = conlist_of(t, p1);
    return bonds_of(functor_kind(&t->ia[p1->type][p2->type]), t, box, p1, p2, &c, sqrt(DOT(r_cm, r_cm)));
}

/* bond energy of p (as part1) with its chain partners: all of them, or only
 * those of smaller index (each pair once in whole-system sums) */
static double bonds_with_partners(const orc_topo* t, const double box[3], const orc_part* p, const orc_part* q,
                                  int lower_only) {
    oconlist c = conlist_of(t, q);
    double e = 0.0;
    int k;
    if (c.empty) return 0.0;
    for (k = 0; k < 4; k++) {
        int j = c.cl[k];
        orc_vec r_cm;
        if (j < 0 || (lower_only && j > q->idx)) continue;
        r_cm = orc_image(box, q->pos, p[j].pos);
        e += bonds_of(functor_kind(&t->ia[q->type][p[j].type]), t, box, q, &p[j], &c, sqrt(DOT(r_cm, r_cm)));
    }
    return e;
}

/* TotalEMatrix::allToAll(EM*) order, totalenergycalculator.h:502-521:
 * pairs (i, j<i) with particle i as part1. */
double orc_energy_all(const orc_topo* t, const double box[3], const orc_part* p, int n) {
    double e = 0.0;
    int i, j;
    for (i = 1; i < n; ++i)
        for (j = 0; j < i; ++j) e += orc_pair_energy(t, box, &p[i], &p[j]);
    return e;
}

/* TotalEFull::oneToAll without a pair list, totalenergycalculator.h:563-583 */
double orc_energy_one(const orc_topo* t, const double box[3], const orc_part* p, int n, int target) {
    double e = 0.0;
    int i;
    for (i = 0; i < n; i++)
        if (i != target) e += orc_pair_energy(t, box, &p[target], &p[i]);
    return e;
}

/* TotalEMatrix::mol2othersTrial without a pair list, totalenergycalculator.h:
 * 484-497: the molecule [first, first + size) with every particle outside it,
 * part1 = the molecule's particle, an empty connectivity list (so only the
 * non-bonded functor, zero beyond sqmaxcut).  mol2others (:438-452) sums the
 * same pairs from the cached matrix. */
double orc_mol2others(const orc_topo* t, const double box[3], const orc_part* p, int n, int first, int size) {
    double e = 0.0;
    int a, j;
    for (a = first; a < first + size; a++)
        for (j = 0; j < n; j++)
            if (j < first || j >= first + size) e += pair_nb(t, box, &p[a], &p[j]);
    return e;
}

/* ------------------------------------------------------------- cell grid */
typedef struct {
    int nc[3];
    int* head;   /* CSR start per cell */
    int* idx;
    int* cell;
} ogrid;

static double frac01(double s) { return s - floor(s); }

static int grid_build(ogrid* g, const orc_part* p, int n, const double box[3], double rc) {
    int k, i, ncell;
    int* cnt;
    for (k = 0; k < 3; k++) {
        g->nc[k] = (int)floor(box[k] / rc);
        if (g->nc[k] < 1) g->nc[k] = 1;
    }
    ncell = g->nc[0] * g->nc[1] * g->nc[2];
    g->head = (int*)calloc((size_t)ncell + 1, sizeof(int));
    g->idx = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    g->cell = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    cnt = (int*)calloc((size_t)ncell, sizeof(int));
    for (i = 0; i < n; i++) {
        int c[3];
        double s[3] = {frac01(p[i].pos.x), frac01(p[i].pos.y), frac01(p[i].pos.z)};
        for (k = 0; k < 3; k++) {
            c[k] = (int)(s[k] * g->nc[k]);
            if (c[k] >= g->nc[k]) c[k] = g->nc[k] - 1;
        }
        g->cell[i] = c[0] + g->nc[0] * (c[1] + g->nc[1] * c[2]);
        g->head[g->cell[i] + 1]++;
    }
    for (i = 0; i < ncell; i++) g->head[i + 1] += g->head[i];
    for (i = 0; i < n; i++) g->idx[g->head[g->cell[i]] + cnt[g->cell[i]]++] = i;
    free(cnt);
    return ncell;
}

static void grid_free(ogrid* g) { free(g->head); free(g->idx); free(g->cell); }

/* list of distinct neighbour cells (handles grids narrower than 3 cells) */
static int grid_stencil(const ogrid* g, int cell, int* out) {
    int cx = cell % g->nc[0], cy = (cell / g->nc[0]) % g->nc[1], cz = cell / (g->nc[0] * g->nc[1]);
    int dx, dy, dz, m = 0, q;
    for (dz = -1; dz <= 1; dz++)
        for (dy = -1; dy <= 1; dy++)
            for (dx = -1; dx <= 1; dx++) {
                int x = (cx + dx + g->nc[0]) % g->nc[0], y = (cy + dy + g->nc[1]) % g->nc[1];
                int z = (cz + dz + g->nc[2]) % g->nc[2], c = x + g->nc[0] * (y + g->nc[1] * z), dup = 0;
                for (q = 0; q < m; q++) if (out[q] == c) dup = 1;
                if (!dup) out[m++] = c;
            }
    return m;
}

double orc_energy_all_cells(const orc_topo* t, const double box[3], const orc_part* p, int n) {
    ogrid g;
    int i, ncell, st[27];
    double e = 0.0;
    if (n < 2) return 0.0;
    ncell = grid_build(&g, p, n, box, t->maxcut);
    (void)ncell;
    for (i = 0; i < n; i++) {
        int m = grid_stencil(&g, g.cell[i], st), q, k;
        for (q = 0; q < m; q++)
            for (k = g.head[st[q]]; k < g.head[st[q] + 1]; k++) {
                int j = g.idx[k];
                if (j < i) e += pair_nb(t, box, &p[i], &p[j]);
            }
        e += bonds_with_partners(t, box, p, &p[i], 1);
    }
    grid_free(&g);
    return e;
}

/* ----------------------------------------------------------- SoA binding */
static void load_part(const orc_topo* t, orc_part* q, int i, const double* pos, const double* dir,
                      const double* pd0, const double* pd1, const int* type) {
    memset(q, 0, sizeof(*q));
    q->pos = V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    q->dir = V(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    q->patchdir[0] = V(pd0[3 * i], pd0[3 * i + 1], pd0[3 * i + 2]);
    if (pd1) q->patchdir[1] = V(pd1[3 * i], pd1[3 * i + 1], pd1[3 * i + 2]);
    q->type = type[i];
    q->idx = i;
    q->moltype = -1;
    if (t->chain && i < t->nchain) {
        q->moltype = t->chain[2 * i];
        q->molpos = t->chain[2 * i + 1];
    }
    orc_part_init(t, q);
}

static orc_part* load_all(const orc_topo* t, int n, const double* pos, const double* dir, const double* pd0,
                          const double* pd1, const int* type) {
    orc_part* p = (orc_part*)malloc(sizeof(orc_part) * (size_t)(n > 0 ? n : 1));
    int i;
    for (i = 0; i < n; i++) load_part(t, &p[i], i, pos, dir, pd0, pd1, type);
    return p;
}

int orc_soa_pair_energies(const orc_topo* t, const double box[3], int n, const double* pos, const double* dir,
                          const double* pd0, const double* pd1, const int* type, int m, const int* pi,
                          const int* pj, double* out) {
    orc_part* p = load_all(t, n, pos, dir, pd0, pd1, type);
    int k;
    for (k = 0; k < m; k++) out[k] = orc_pair_energy(t, box, &p[pi[k]], &p[pj[k]]);
    free(p);
    return 0;
}

int orc_soa_bond_energies(const orc_topo* t, const double box[3], int n, const double* pos, const double* dir,
                          const double* pd0, const double* pd1, const int* type, int m, const int* pi,
                          const int* pj, double* out) {
    orc_part* p = load_all(t, n, pos, dir, pd0, pd1, type);
    int k;
    for (k = 0; k < m; k++) out[k] = orc_bond_energy(t, box, &p[pi[k]], &p[pj[k]]);
    free(p);
    return 0;
}

double orc_soa_energy_all(const orc_topo* t, const double box[3], int n, const double* pos, const double* dir,
                          const double* pd0, const double* pd1, const int* type, int use_cells) {
    orc_part* p = load_all(t, n, pos, dir, pd0, pd1, type);
    double e = use_cells ? orc_energy_all_cells(t, box, p, n) : orc_energy_all(t, box, p, n);
    free(p);
    return e;
}

int orc_soa_energy_one_each(const orc_topo* t, const double box[3], int n, const double* pos, const double* dir,
                            const double* pd0, const double* pd1, const int* type, int m, const int* targets,
                            double* out) {
    orc_part* p = load_all(t, n, pos, dir, pd0, pd1, type);
    int k;
    for (k = 0; k < m; k++) out[k] = orc_energy_one(t, box, p, n, targets[k]);
    free(p);
    return 0;
}

int orc_soa_mol2others(const orc_topo* t, const double box[3], int n, const double* pos, const double* dir,
                       const double* pd0, const double* pd1, const int* type, int m, const int* first,
                       const int* size, double* out) {
    orc_part* p = load_all(t, n, pos, dir, pd0, pd1, type);
    int k;
    for (k = 0; k < m; k++) out[k] = orc_mol2others(t, box, p, n, first[k], size[k]);
    free(p);
    return 0;
}

int orc_soa_derived(const orc_topo* t, int n, const double* dir, const double* pd0, const double* pd1,
                    const int* type, double* out) {
    int i, k;
    for (i = 0; i < n; i++) {
        orc_part q;
        const orc_vec* vs[9];
        double zero[3] = {0, 0, 0};
        load_part(t, &q, 0, zero, dir + 3 * i, pd0 + 3 * i, pd1 ? pd1 + 3 * i : NULL, type + i);
        vs[0] = &q.dir; vs[1] = &q.patchdir[0]; vs[2] = &q.patchdir[1];
        vs[3] = &q.patchsides[0]; vs[4] = &q.patchsides[1]; vs[5] = &q.patchsides[2]; vs[6] = &q.patchsides[3];
        vs[7] = &q.chdir[0]; vs[8] = &q.chdir[1];
        for (k = 0; k < 9; k++) {
            out[30 * i + 3 * k] = vs[k]->x; out[30 * i + 3 * k + 1] = vs[k]->y; out[30 * i + 3 * k + 2] = vs[k]->z;
        }
        out[30 * i + 27] = out[30 * i + 28] = out[30 * i + 29] = 0.0;
    }
    return 0;
}

/* Inicializer::initConfig per-particle transformations (inicializer.cpp:201-225):
 * position scaled to box fractions and folded by Cuboid::usePBC
 * (geometry.h:78-97), dir normalised, patchdir orthogonalised to dir and
 * normalised.  Arrays are n*3 and updated in place; pos becomes fractions. */
int orc_soa_initconfig(int n, const double box[3], double* pos, double* dir, double* pd0) {
    int i, k;
    for (i = 0; i < n; i++) {
        orc_vec d = V(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        orc_vec q = V(pd0[3 * i], pd0[3 * i + 1], pd0[3 * i + 2]);
        for (k = 0; k < 3; k++) {
            double s = pos[3 * i + k] / box[k];
            while (s < 0.0) s += 1.0;
            while (s > 1.0) s -= 1.0;
            pos[3 * i + k] = s;
        }
        d = NORM(d);
        q = NORM(ORTHO(q, d));
        dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
        pd0[3 * i] = q.x; pd0[3 * i + 1] = q.y; pd0[3 * i + 2] = q.z;
    }
    return 0;
}

/* ------------------------------------------------------------ cell lists */
static inline int axis_cell(uint32_t u, uint32_t shift, int nc) {
    uint32_t s = u - shift;   /* wraps modulo 2^32 */
    return (int)(((uint64_t)s * (uint64_t)nc) >> 32);
}

int orc_cell_list(int n, const uint32_t* ux, const uint32_t* uy, const uint32_t* uz, int ncx, int ncy, int ncz,
                  uint32_t sx, uint32_t sy, uint32_t sz, int32_t* cell_of, int32_t* cell_start, int32_t* sorted) {
    int ncell = ncx * ncy * ncz, i;
    int32_t* fill = (int32_t*)calloc((size_t)ncell, sizeof(int32_t));
    for (i = 0; i <= ncell; i++) cell_start[i] = 0;
    for (i = 0; i < n; i++) {
        int c = axis_cell(ux[i], sx, ncx) + ncx * (axis_cell(uy[i], sy, ncy) + ncy * axis_cell(uz[i], sz, ncz));
        cell_of[i] = c;
        cell_start[c + 1]++;
    }
    for (i = 0; i < ncell; i++) cell_start[i + 1] += cell_start[i];
    for (i = 0; i < n; i++) sorted[cell_start[cell_of[i]] + fill[cell_of[i]]++] = i;
    free(fill);
    return 0;
}

long orc_neighbor_pairs(int n, const uint32_t* ux, const uint32_t* uy, const uint32_t* uz, const double box[3],
                        double rc, int32_t* out, long cap) {
    long cnt = 0;
    int i, j;
    double rc2 = rc * rc;
    const double s32 = 1.0 / 4294967296.0;
    for (i = 0; i < n; i++)
        for (j = i + 1; j < n; j++) {
            double dx = box[0] * ((double)(int32_t)(ux[i] - ux[j]) * s32);
            double dy = box[1] * ((double)(int32_t)(uy[i] - uy[j]) * s32);
            double dz = box[2] * ((double)(int32_t)(uz[i] - uz[j]) * s32);
            double r2 = dx * dx;
            r2 = r2 + dy * dy;
            r2 = r2 + dz * dz;
            if (r2 <= rc2) {
                if (out && cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                cnt++;
            }
        }
    return cnt;
}

/* ------------------------------------------------------------- CPU MC */
/* splitmix64 stream: the oracle's own generator (not the reference's dSFMT) */
static inline uint64_t sm64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double urand(uint64_t* s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Vector::randomUnitSphere, Vector.h:190-205 (Marsaglia 1972) */
static orc_vec rand_sphere(uint64_t* s) {
    double a, b, x1, x2;
    do {
        x1 = 1.0 - 2.0 * urand(s);
        x2 = 1.0 - 2.0 * urand(s);
        a = x1 * x1 + x2 * x2;
    } while (a > 1.0);
    b = 2.0 * sqrt(1.0 - a);
    return V(x1 * b, x2 * b, 1.0 - 2.0 * a);
}

/* non-bonded energy of q (as part1) with all particles in the 27 cells around
 * it except those of index lo..hi-1 */
static double grid_excl(const orc_topo* t, const double box[3], const ogrid* g, const orc_part* p, const orc_part* q,
                        int lo, int hi) {
    int st[27], m, a, k, cx, cy, cz, c;
    double e = 0.0, s[3] = {frac01(q->pos.x), frac01(q->pos.y), frac01(q->pos.z)};
    cx = (int)(s[0] * g->nc[0]); if (cx >= g->nc[0]) cx = g->nc[0] - 1;
    cy = (int)(s[1] * g->nc[1]); if (cy >= g->nc[1]) cy = g->nc[1] - 1;
    cz = (int)(s[2] * g->nc[2]); if (cz >= g->nc[2]) cz = g->nc[2] - 1;
    c = cx + g->nc[0] * (cy + g->nc[1] * cz);
    m = grid_stencil(g, c, st);
    for (a = 0; a < m; a++)
        for (k = g->head[st[a]]; k < g->head[st[a] + 1]; k++) {
            int j = g->idx[k];
            if (j < lo || j >= hi) e += pair_nb(t, box, q, &p[j]);
        }
    return e;
}

/* energy of p (as part1) with all particles in the 27 cells around it, skipping self */
static double grid_one(const orc_topo* t, const double box[3], const ogrid* g, const orc_part* p, const orc_part* q,
                       int self) {
    return grid_excl(t, box, g, p, q, self, self + 1) + bonds_with_partners(t, box, p, q, 0);
}

/* move particle i between cells of the CSR grid by re-building only if its cell changed */
static void grid_move(ogrid* g, const orc_part* p, int n, int i) {
    double s[3] = {frac01(p[i].pos.x), frac01(p[i].pos.y), frac01(p[i].pos.z)};
    int c[3], k, nc, oc = g->cell[i], a, pos = -1, ncell = g->nc[0] * g->nc[1] * g->nc[2];
    (void)n;
    for (k = 0; k < 3; k++) { c[k] = (int)(s[k] * g->nc[k]); if (c[k] >= g->nc[k]) c[k] = g->nc[k] - 1; }
    nc = c[0] + g->nc[0] * (c[1] + g->nc[1] * c[2]);
    if (nc == oc) return;
    /* remove i from oc: shift the tail of idx left by one over cells oc..nc, or right */
    for (a = g->head[oc]; a < g->head[oc + 1]; a++) if (g->idx[a] == i) { pos = a; break; }
    if (nc > oc) {
        int cc;
        for (a = pos; a < g->head[nc + 1] - 1; a++) g->idx[a] = g->idx[a + 1];
        for (cc = oc + 1; cc <= nc; cc++) g->head[cc]--;
        g->idx[g->head[nc + 1] - 1] = i;
    } else {
        int cc;
        for (a = pos; a > g->head[nc + 1]; a--) g->idx[a] = g->idx[a - 1];
        for (cc = nc + 1; cc <= oc; cc++) g->head[cc]++;
        g->idx[g->head[nc + 1] - 1] = i;
    }
    (void)ncell;
    g->cell[i] = nc;
}

typedef struct {
    double chainprob, chain_mx, chain_angle;
} ochainmv;

static double mc_run(const orc_topo* t, double box[3], int n, double* pos, double* dir, double* pd0, double* pd1,
                     const int* type, long nsweeps, double temper, double transmx, double rot_half, uint64_t seed,
                     long* stats, double* e_final, int ptype, double press, double edge_mx, const ochainmv* cm);

/* Planted mistakes for the sampler's own tests (0 = none, the default): they
 * show that the two-particle Boltzmann check fails for plausible errors.
 *   1: the Metropolis test at twice the temperature;
 *   2: rotations about the particle's own axis dir x patchdir[0], always in
 *      the same sense (turning dir towards the patch), so no move has a reverse;
 *   3: translations always along the particle's own +dir axis (no reverse move). */
static int orc_fault = 0;
void orc_set_sampler_fault(int f) { orc_fault = f; }

double orc_mc_run(const orc_topo* t, const double box[3], int n, double* pos, double* dir, double* pd0, double* pd1,
                  const int* type, long nsweeps, double temper, double transmx, double rot_half, uint64_t seed,
                  long* stats, double* e_final) {
    double b[3] = {box[0], box[1], box[2]};
    return mc_run(t, b, n, pos, dir, pd0, pd1, type, nsweeps, temper, transmx, rot_half, seed, stats, e_final, -1,
                  0.0, 0.0, NULL);
}

double orc_mc_run_npt(const orc_topo* t, double box[3], int n, double* pos, double* dir, double* pd0, double* pd1,
                      const int* type, long nsweeps, double temper, double transmx, double rot_half, uint64_t seed,
                      long* stats, double* e_final, int ptype, double press, double edge_mx) {
    return mc_run(t, box, n, pos, dir, pd0, pd1, type, nsweeps, temper, transmx, rot_half, seed, stats, e_final,
                  ptype, press, edge_mx, NULL);
}

double orc_mc_run_chains(const orc_topo* t, double box[3], int n, double* pos, double* dir, double* pd0, double* pd1,
                         const int* type, long nsweeps, double temper, double transmx, double rot_half,
                         double chainprob, double chain_mx, double chain_angle, uint64_t seed, long* stats,
                         double* e_final, int ptype, double press, double edge_mx) {
    ochainmv cm;
    cm.chainprob = chainprob;
    cm.chain_mx = chain_mx;
    cm.chain_angle = chain_angle;
    return mc_run(t, box, n, pos, dir, pd0, pd1, type, nsweeps, temper, transmx, rot_half, seed, stats, e_final,
                  ptype, press, edge_mx, &cm);
}

/* non-bonded energy of the molecule [first, first + size) with every other
 * particle, through the grid (mol2others / mol2othersTrial) */
static double grid_mol(const orc_topo* t, const double box[3], const ogrid* g, const orc_part* p, int first,
                       int size) {
    double e = 0.0;
    int a;
    for (a = first; a < first + size; a++) e += grid_excl(t, box, g, p, &p[a], first, first + size);
    return e;
}

/* MoveCreator::clusterCM, movecreator.cpp:1307-1324: the centre of the
 * molecule weighted by Ia_param::volume, in box fractions.  It is taken on
 * coordinates unwrapped by the minimum image from the first particle (off[a]
 * = offset of particle a); the reference averages the stored fractions, which
 * it wraps into the box only on input, so a molecule that straddles the box
 * edge there rotates about a point far from it. */
static orc_vec chain_cm(const orc_topo* t, const orc_part* p, int first, int size, orc_vec* off) {
    orc_vec c = V(0.0, 0.0, 0.0);
    double vsum = 0.0;
    int a;
    for (a = 0; a < size; a++) {
        const orc_part* q = &p[first + a];
        double v = t->ia[q->type][q->type].volume;
        orc_vec d = V(q->pos.x - p[first].pos.x, q->pos.y - p[first].pos.y, q->pos.z - p[first].pos.z);
        d.x -= nearbyint(d.x);
        d.y -= nearbyint(d.y);
        d.z -= nearbyint(d.z);
        off[a] = d;
        c.x += d.x * v; c.y += d.y * v; c.z += d.z * v;
        vsum += v;
    }
    return SCL(1.0 / vsum, c);
}

/* MoveCreator::chainMove, movecreator.cpp:305-328: a molecule of more than one
 * particle chosen uniformly, displaced (chainDisplace, :1075-1137) or rotated
 * about its centre (chainRotate + clusterRotate, :1139-1305) with equal
 * probability; moveTry on mol2others before and mol2othersTrial after.
 * Returns the accepted energy change (0 on rejection); stats[6..9] = displace
 * acc, tries, rotate acc, tries. */
static double chain_move(const orc_topo* t, const double box[3], ogrid* g, orc_part* p, int n, const int* cfirst,
                         const int* csize, int nch, double temper, const ochainmv* cm, uint64_t* rs, long* stats,
                         orc_part* save, orc_vec* off) {
    int c = (int)(urand(rs) * nch), first, size, a, disp;
    double eo, en;
    if (c >= nch) c = nch - 1;
    first = cfirst[c];
    size = csize[c];
    disp = urand(rs) < 0.5;
    eo = grid_mol(t, box, g, p, first, size);
    for (a = 0; a < size; a++) save[a] = p[first + a];
    if (disp) {
        /* |dr| = chain_mx exactly in a uniform direction, :1088-1091 */
        orc_vec dr = rand_sphere(rs);
        for (a = 0; a < size; a++) {
            p[first + a].pos.x += dr.x * cm->chain_mx / box[0];
            p[first + a].pos.y += dr.y * cm->chain_mx / box[1];
            p[first + a].pos.z += dr.z * cm->chain_mx / box[2];
        }
    } else {
        /* clusterRotate, :1258-1305: axis, then the quaternion half-angle
         * chain_angle * U and a random sign; positions rotate about the
         * centre in real units, with the axes and patch directions */
        orc_vec ax = rand_sphere(rs), cen, base = p[first].pos;
        double vc = cos(cm->chain_angle * urand(rs));
        double vs = (urand(rs) < 0.5) ? sqrt(1.0 - vc * vc) : -sqrt(1.0 - vc * vc);
        cen = chain_cm(t, p, first, size, off);
        for (a = 0; a < size; a++) {
            orc_part* q = &p[first + a];
            orc_vec o = V((off[a].x - cen.x) * box[0], (off[a].y - cen.y) * box[1], (off[a].z - cen.z) * box[2]);
            o = ROT(o, ax, vc, vs);
            q->pos = V(base.x + cen.x + o.x / box[0], base.y + cen.y + o.y / box[1], base.z + cen.z + o.z / box[2]);
            q->dir = ROT(q->dir, ax, vc, vs);
            q->patchdir[0] = ROT(q->patchdir[0], ax, vc, vs);
            orc_part_init(t, q);
        }
    }
    en = grid_mol(t, box, g, p, first, size);
    stats[disp ? 7 : 9]++;
    if (!(en <= eo) && !(exp(-1.0 * (en - eo) / temper) > urand(rs))) {   /* moveTry, movecreator.h:175-187 */
        for (a = 0; a < size; a++) p[first + a] = save[a];
        return 0.0;
    }
    stats[disp ? 6 : 8]++;
    for (a = 0; a < size; a++) grid_move(g, p, n, first + a);
    return en - eo;
}

/* MoveCreator::pressureMove, movecreator.cpp:330-421 (ptype 0: one random
 * edge, 1: isotropic); box fractions stay, so every position scales.
 * Returns the change of the energy (without the pressure-volume terms) of an
 * accepted move, 0 on rejection; *e_cur is the current whole-system energy. */
static double pressure_move(const orc_topo* t, double box[3], const orc_part* p, int n, int ptype, double press,
                            double temper, double edge_mx, uint64_t* rs, double* e_cur, long* vstats) {
    double nb[3] = {box[0], box[1], box[2]}, enermove, e_new;
    int k;
    if (ptype == 0) {
        double rsave = urand(rs), area, old;
        int side = rsave < 1.0 / 3.0 ? 0 : (rsave < 2.0 / 3.0 ? 1 : 2);
        area = nb[(side + 1) % 3] * nb[(side + 2) % 3];
        old = nb[side];
        nb[side] += edge_mx * (urand(rs) - 0.5);
        if (nb[side] <= 0.0) { vstats[1]++; return 0.0; }
        enermove = press * area * (nb[side] - old) - (double)n * temper * log(nb[side] / old);
    } else {
        double psch = edge_mx * (urand(rs) - 0.5), pvol = nb[0] * nb[1] * nb[2], pvoln;
        for (k = 0; k < 3; k++) nb[k] += psch;
        pvoln = nb[0] * nb[1] * nb[2];
        enermove = press * (pvoln - pvol) - (double)n * temper * log(pvoln / pvol);
    }
    vstats[1]++;
    /* the cell grid needs edges above the cutoff (the device's limit too) */
    for (k = 0; k < 3; k++) if (!(nb[k] > t->maxcut)) return 0.0;
    e_new = orc_energy_all_cells(t, nb, p, n);
    enermove += e_new;
    /* moveTry, movecreator.h:175-187 */
    if (!(enermove <= *e_cur) && !(exp(-1.0 * (enermove - *e_cur) / temper) > urand(rs))) return 0.0;
    vstats[0]++;
    {
        double de = e_new - *e_cur;
        for (k = 0; k < 3; k++) box[k] = nb[k];
        *e_cur = e_new;
        return de;
    }
}

static double mc_run(const orc_topo* t, double box[3], int n, double* pos, double* dir, double* pd0, double* pd1,
                     const int* type, long nsweeps, double temper, double transmx, double rot_half, uint64_t seed,
                     long* stats, double* e_final, int ptype, double press, double edge_mx, const ochainmv* cm) {
    orc_part* p = load_all(t, n, pos, dir, pd0, pd1, type);
    ogrid g;
    uint64_t rs = seed * 0x2545F4914F6CDD1Dull + 12345u;
    double e_start, e_end, edrift = 0.0, e_cur = 0.0;
    long sweep, step;
    int i, nch = 0, maxsize = 1;
    int* cfirst = NULL;
    int* csize = NULL;
    orc_part* save = NULL;
    orc_vec* off = NULL;
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    if (ptype >= 0) stats[4] = stats[5] = 0;
    if (cm) {
        stats[6] = stats[7] = stats[8] = stats[9] = 0;
        /* the molecules of more than one particle (ParticleVector::getChain,
         * Conf.h:150-163); none: no chain moves (Inicializer::testChains) */
        cfirst = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        csize = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (i = 0; i < n; i++) {
            int m;
            if (p[i].moltype < 0 || p[i].molpos != 0) continue;
            m = t->mol[p[i].moltype].size;
            if (m < 2 || i + m > n) continue;
            cfirst[nch] = i;
            csize[nch] = m;
            if (m > maxsize) maxsize = m;
            nch++;
        }
        save = (orc_part*)malloc(sizeof(orc_part) * (size_t)maxsize);
        off = (orc_vec*)malloc(sizeof(orc_vec) * (size_t)maxsize);
    }
    grid_build(&g, p, n, box, t->maxcut);
    e_start = (e_final || ptype >= 0) ? orc_energy_all_cells(t, box, p, n) : 0.0;
    for (sweep = 0; sweep < nsweeps; sweep++) {
        for (step = 0; step < n; step++) {
            /* the step's move kind, Updater's step loop (updater.cpp:206-226):
             * a chain move with probability chainprob (a per-step probability
             * there, despite its "per sweep" description) */
            if (cm && cm->chainprob > 0.0 && nch > 0 && urand(&rs) < cm->chainprob) {
                edrift += chain_move(t, box, &g, p, n, cfirst, csize, nch, temper, cm, &rs, stats, save, off);
                continue;
            }
            /* MoveCreator::particleMove, movecreator.cpp:16-29 */
            int target = (int)(urand(&rs) * n);
            orc_part old = p[target];
            int g0 = t->ia[p[target].type][p[target].type].geotype[0];
            int translate = (urand(&rs) < 0.5) || (g0 >= ORC_SP);
            double eo, en, de;
            int reject;
            eo = grid_one(t, box, &g, p, &p[target], target);
            if (translate) {
                /* partDisplace, movecreator.cpp:956-964: |dr| = transmx exactly */
                orc_vec dr = rand_sphere(&rs);
                if (orc_fault == 3) {
                    double dn = sqrt(DOT(p[target].dir, p[target].dir));
                    dr.x = p[target].dir.x / dn; dr.y = p[target].dir.y / dn; dr.z = p[target].dir.z / dn;
                }
                p[target].pos.x += dr.x * transmx / box[0];
                p[target].pos.y += dr.y * transmx / box[1];
                p[target].pos.z += dr.z * transmx / box[2];
            } else {
                /* Particle::rotateRandom / pscRotate (particle.h:171-272): the
                 * quaternion (cos a, sin a * axis), a = rot_half*U, random sign;
                 * derived vectors are rebuilt by Particle::init on the rotated
                 * dir/patchdir (rotation preserves them up to rounding). */
                double ang = rot_half * urand(&rs), vc, vs;
                orc_vec ax = rand_sphere(&rs);
                vc = cos(ang);
                vs = (urand(&rs) < 0.5) ? sqrt(1.0 - vc * vc) : -sqrt(1.0 - vc * vc);
                if (orc_fault == 2) {
                    orc_vec d0 = p[target].dir, q0 = p[target].patchdir[0];
                    orc_vec nx = {d0.y * q0.z - d0.z * q0.y, d0.z * q0.x - d0.x * q0.z, d0.x * q0.y - d0.y * q0.x};
                    double nn = sqrt(DOT(nx, nx));
                    ax.x = nx.x / nn; ax.y = nx.y / nn; ax.z = nx.z / nn;
                    vs = fabs(vs);
                }
                p[target].dir = ROT(p[target].dir, ax, vc, vs);
                p[target].patchdir[0] = ROT(p[target].patchdir[0], ax, vc, vs);
                p[target].patchdir[0] = ORTHO(p[target].patchdir[0], p[target].dir);  /* partRotate:1007 */
                orc_part_init(t, &p[target]);
            }
            en = grid_one(t, box, &g, p, &p[target], target);
            /* moveTry, movecreator.h:175-187 */
            reject = 0;
            if (!(en <= eo)) reject = !(exp(-1.0 * (en - eo) / (orc_fault == 1 ? 2.0 * temper : temper)) > urand(&rs));
            stats[translate ? 1 : 3]++;
            if (reject) {
                p[target] = old;
            } else {
                stats[translate ? 0 : 2]++;
                de = en - eo;
                edrift += de;
                if (translate) grid_move(&g, p, n, target);
            }
        }
        if (ptype >= 0) {   /* one volume move at the end of every sweep (shave = 1) */
            long acc0 = stats[4];
            e_cur = orc_energy_all_cells(t, box, p, n);
            edrift += pressure_move(t, box, p, n, ptype, press, temper, edge_mx, &rs, &e_cur, stats + 4);
            if (stats[4] != acc0) {
                grid_free(&g);
                grid_build(&g, p, n, box, t->maxcut);
            }
        }
    }
    e_end = (e_final || ptype >= 0) ? orc_energy_all_cells(t, box, p, n) : 0.0;
    if (e_final) *e_final = e_end;
    for (i = 0; i < n; i++) {
        pos[3 * i] = p[i].pos.x; pos[3 * i + 1] = p[i].pos.y; pos[3 * i + 2] = p[i].pos.z;
        dir[3 * i] = p[i].dir.x; dir[3 * i + 1] = p[i].dir.y; dir[3 * i + 2] = p[i].dir.z;
        pd0[3 * i] = p[i].patchdir[0].x; pd0[3 * i + 1] = p[i].patchdir[0].y; pd0[3 * i + 2] = p[i].patchdir[0].z;
        if (pd1) { pd1[3 * i] = p[i].patchdir[1].x; pd1[3 * i + 1] = p[i].patchdir[1].y; pd1[3 * i + 2] = p[i].patchdir[1].z; }
    }
    grid_free(&g);
    free(p);
    free(cfirst);
    free(csize);
    free(save);
    free(off);
    return e_end - e_start - edrift;
}

/* ------------------------------------------------------------ observables */
/* nematic(), Tests/sc35.c:8563-8603: S2 = -2 * middle eigenvalue of
 * Q = (3<uu> - I)/2.  Eigenvalues by cyclic Jacobi rotations. */
double orc_nematic_s2(int n, const double* dir) {
    double q[3][3] = {{0}}, ev[3];
    int i, a, b, it;
    if (n <= 0) return 0.0;
    for (i = 0; i < n; i++)
        for (a = 0; a < 3; a++)
            for (b = 0; b < 3; b++) q[a][b] += dir[3 * i + a] * dir[3 * i + b];
    for (a = 0; a < 3; a++)
        for (b = 0; b < 3; b++) q[a][b] = (q[a][b] * 3.0 / n - (a == b ? 1.0 : 0.0)) / 2.0;
    for (it = 0; it < 64; it++) {
        double off = q[0][1] * q[0][1] + q[0][2] * q[0][2] + q[1][2] * q[1][2];
        int p, r;
        if (off < 1e-30) break;
        for (p = 0; p < 2; p++)
            for (r = p + 1; r < 3; r++) {
                double th, c, s, t2;
                int k;
                if (fabs(q[p][r]) < 1e-300) continue;
                th = (q[r][r] - q[p][p]) / (2.0 * q[p][r]);
                t2 = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                c = 1.0 / sqrt(t2 * t2 + 1.0);
                s = t2 * c;
                for (k = 0; k < 3; k++) {
                    double kp = q[k][p], kr = q[k][r];
                    q[k][p] = c * kp - s * kr;
                    q[k][r] = s * kp + c * kr;
                }
                for (k = 0; k < 3; k++) {
                    double pk = q[p][k], rk = q[r][k];
                    q[p][k] = c * pk - s * rk;
                    q[r][k] = s * pk + c * rk;
                }
            }
    }
    ev[0] = q[0][0]; ev[1] = q[1][1]; ev[2] = q[2][2];
    for (a = 0; a < 2; a++)
        for (b = 0; b < 2 - a; b++)
            if (ev[b] > ev[b + 1]) { double tmp = ev[b]; ev[b] = ev[b + 1]; ev[b + 1] = tmp; }
    return -2.0 * ev[1];
}
