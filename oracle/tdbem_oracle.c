/*
 * tdbem_oracle.c -- CPU FP64 restatement of the reference multi-domain TDBEM
 * GMRES hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, the smoke entry
 * and bench.py's cpu_baseline leg; never by the product library.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference root).  The arithmetic keeps the reference's operation order
 * so that, compiled with -ffp-contract=off, results are the reference's own
 * FP64 values up to libm differences in table construction.
 *
 * Pinning (what fixes each part independently of this restatement):
 *   - Gauss tables: the reference-held text tables DBEM1/gaussposition,
 *     DBEM1/gaussweight (test vs numpy leggauss and the golden strings).
 *   - DynaMat: closed forms c1, c2 (tests).
 *   - U kernel: telescoping sum over n equals the static Kelvin solution.
 *   - Causality judge: brute force sampling (tests).
 *   - Quadrature: convergence vs dense tensor Gauss of the same integrand.
 *   - DOF map: closed-form counts on the rod fixture (120/88/104).
 *   - GMRES: dense solve on random systems.
 *   - End to end: the reference-held known answer DBEM1/analytic_validation.txt
 *     (rod120_100 loaded-end ux at step 100) and the 1-D rod series.
 */
#include "tdbem_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* DBEM1/precompiler.h:6-27 */
static const double kPi = 3.141592653589793239;
static const double kOneOver4Pi = 0.0795774715459476679;
#define NG 4      /* GAUSSPOINT */
#define NG2 16
#define NSG 4     /* SINGAUSSPOINT */
#define NSG2 16
#define NPWG 3    /* GAUSSPOINTPW */
#define NPWG2 9
#define NPW 6     /* NUMPW */
#define NPW2 36
#define NSPWG 3   /* SINGAUSSPOINTPW */
#define NSPWG2 9
#define NSPW 20   /* SINNUMPW */
#define NSPW2 400

/* ------------------------------------------------------------------------ */
/* DynaMat  (DBEM1/dsquareelement.h:22-56, DBEM1/dsquareelement.cpp:63-97)   */
/* ------------------------------------------------------------------------ */
typedef struct {
    double G, v, Rou, C1, C2, Dt, Dt2, Dt_2, Dt_3, C1_2, C2_2, OneOverC1_2, OneOverC2_2,
        OneOverC1_3, OneOverC2_3, C2_2OverC1_2, C2_2OverC1_3, C1Dt, C2Dt, OneOverFourPaiRou,
        OneOverThree_1OC23_1OC13, Half_1OC22_1OC12, STCoef1, STCoef2;
} dmat;

static dmat build_dmat(double v, double G, double Rou, double dt)
{
    dmat m;
    m.v = v;
    m.G = G;
    m.Rou = Rou;
    m.C1 = sqrt((2.0 - 2.0 * v) / (1.0 - 2.0 * v) * G / Rou);
    m.C2 = sqrt(G / Rou);
    m.Dt = dt;
    m.Dt2 = 2.0 * dt;
    m.Dt_2 = dt * dt;
    m.Dt_3 = dt * dt * dt;
    m.OneOverFourPaiRou = kOneOver4Pi / Rou;
    m.C1_2 = m.C1 * m.C1;
    m.C2_2 = m.C2 * m.C2;
    m.OneOverC1_2 = 1.0 / m.C1_2;
    m.OneOverC2_2 = 1.0 / m.C2_2;
    m.OneOverC1_3 = 1.0 / m.C1_2 / m.C1;
    m.OneOverC2_3 = 1.0 / m.C2_2 / m.C2;
    m.C2_2OverC1_2 = m.C2_2 / m.C1_2;
    m.C2_2OverC1_3 = m.C2_2OverC1_2 / m.C1;
    m.C1Dt = m.C1 * m.Dt;
    m.C2Dt = m.C2 * m.Dt;
    m.STCoef1 = -1.0 / (8.0 * kPi * (1.0 - m.v));
    m.STCoef2 = 1.0 - 2.0 * m.v;
    m.OneOverThree_1OC23_1OC13 = 1.0 / 3.0 * (m.OneOverC2_3 - m.OneOverC1_3);
    m.Half_1OC22_1OC12 = 0.5 * (m.OneOverC2_2 - m.OneOverC1_2);
    return m;
}

/* ------------------------------------------------------------------------ */
/* Quadrature tables (DBEM1/gaussnumematrix.cpp:4-85, DBEM1/quadinfo.cpp)    */
/* ------------------------------------------------------------------------ */
/* The reference reads these from the text files DBEM1/gaussposition and
 * DBEM1/gaussweight (orders 3 and 4 at offsets 40*(order-1)); the 15-decimal
 * strings are kept verbatim because w3 = 0.555555555555555 is not 5/9. */
static const char* kPos3[3] = {"-0.774596669241484", "0.000000000000000", "0.774596669241484"};
static const char* kW3[3] = {"0.555555555555555", "0.888888888888889", "0.555555555555555"};
static const char* kPos4[4] = {"-0.861136311594052", "-0.339981043584856", "0.339981043584856",
                               "0.861136311594052"};
static const char* kW4[4] = {"0.347854845137454", "0.652145154862546", "0.652145154862546",
                             "0.347854845137454"};

typedef struct {
    double gap;
    double lc[8][2];             /* m_LocalCoord */
    double rp[NG2][2], ra[NG2];  /* regular 4x4 */
    double sp[NSG2][2], sa[NSG2], spos[NSG], sb[NSG];
    double pwrp[NPWG2][2], pwra[NPWG2];
    double pwsp[NSPWG2][2], pwsa[NSPWG2];
    double nrgv0[NG2];           /* m_NRGV[0][*] (Q1 weights) */
    double cenpw[NPW2][2];
    double lpospw[NPW2][NPWG2][2];
    double pwgv[NPW2][NPWG2];
    double sinlpospw[NSPW2][NSPWG2][2];
    double sinnpwgv[NSPWG2];
    /* constant-element singular tables */
    double secconstpos[4][2][NSG2];
    double secconstval[4][NSG2];
    double constlinepos[4][NSG];
    double secconstpospw[4][2][NSPW2][NSPWG2];
    double secconstvalpw[4][NSPW2][NSPWG2];
    double constfm1secval[4][NSG2];
    double constfm1lineval[4][NSG];
} qtables;

/* quadinfo::N, first component only (DBEM1/quadinfo.h:149-172) */
static double quad_n0(const qtables* q, double s1, double s2)
{
    double lx = q->lc[1][0] - q->lc[0][0];
    double ly = q->lc[3][1] - q->lc[0][1];
    double lxly = lx * ly;
    double hx = 0.25 * lx * lxly;
    double hy = 0.25 * ly * lxly;
    double r0 = (s1 - q->lc[1][0]) * (s2 - q->lc[3][1]) / lxly;
    double r4 = (s1 - q->lc[0][0]) * (s1 - q->lc[1][0]) * (s2 - q->lc[6][1]) / (hx);
    double r7 = (s1 - q->lc[5][0]) * (s2 - q->lc[0][1]) * (s2 - q->lc[3][1]) / (hy);
    r0 -= 0.5 * (r4 + r7);
    return r0;
}

/* Separate libm calls for sin and cos (no sincos fusion by the compiler). */
__attribute__((noinline)) static double ocos(double x) { return cos(x); }
__attribute__((noinline)) static double osin(double x) { return sin(x); }

/* DBEM1/quadinfo.cpp:5-48 (constant-element polar helpers) */
static double const_theta_start(int a)
{
    switch (a) {
    case 0: return -0.75 * kPi;
    case 1: return -0.25 * kPi;
    case 2: return 0.25 * kPi;
    case 3: return 0.75 * kPi;
    default: return 0.0;
    }
}
static double const_theta_diff(void) { return 0.5 * kPi; }
static double const_theta_from_local(double lt, int a)
{
    return const_theta_start(a) + 0.5 * (lt + 1.0) * const_theta_diff();
}
static double const_alpha(double th, int a)
{
    switch (a) {
    case 0: return -1.0 / osin(th);
    case 1: return 1.0 / ocos(th);
    case 2: return 1.0 / osin(th);
    case 3: return -1.0 / ocos(th);
    default: return 0.0;
    }
}
static double const_polar_from_local(double lr, double th, int a)
{
    return 0.5 * (lr + 1.0) * const_alpha(th, a);
}

static void build_tables(qtables* q, double gap)
{
    double pos4[4], w4[4], pos3[3], w3[3];
    int i, j, m, p, a, n;
    for (i = 0; i < 4; ++i) {
        pos4[i] = strtod(kPos4[i], 0);
        w4[i] = strtod(kW4[i], 0);
    }
    for (i = 0; i < 3; ++i) {
        pos3[i] = strtod(kPos3[i], 0);
        w3[i] = strtod(kW3[i], 0);
    }
    memset(q, 0, sizeof(*q));
    q->gap = gap;
    /* GaussNumeMatrix ctor (DBEM1/gaussnumematrix.cpp:43-81) */
    for (i = 0; i < NG; ++i)
        for (j = 0; j < NG; ++j) {
            m = i * NG + j;
            q->rp[m][0] = pos4[j];
            q->rp[m][1] = pos4[i];
            q->ra[m] = w4[j] * w4[i];
        }
    for (i = 0; i < NSG; ++i) {
        q->spos[i] = pos4[i];
        q->sb[i] = w4[i];
    }
    for (i = 0; i < NSG; ++i)
        for (j = 0; j < NSG; ++j) {
            m = i * NSG + j;
            q->sp[m][0] = pos4[j];
            q->sp[m][1] = pos4[i];
            q->sa[m] = w4[j] * w4[i];
        }
    for (i = 0; i < NPWG; ++i)
        for (j = 0; j < NPWG; ++j) {
            m = i * NPWG + j;
            q->pwrp[m][0] = pos3[j];
            q->pwrp[m][1] = pos3[i];
            q->pwra[m] = w3[j] * w3[i];
        }
    for (i = 0; i < NSPWG; ++i)
        for (j = 0; j < NSPWG; ++j) {
            m = i * NSPWG + j;
            q->pwsp[m][0] = pos3[j];
            q->pwsp[m][1] = pos3[i];
            q->pwsa[m] = w3[j] * w3[i];
        }
    /* quadinfo::GetLocalCoord (DBEM1/quadinfo.h:80-107) */
    q->lc[0][0] = -gap; q->lc[0][1] = -gap;
    q->lc[1][0] = gap;  q->lc[1][1] = -gap;
    q->lc[2][0] = gap;  q->lc[2][1] = gap;
    q->lc[3][0] = -gap; q->lc[3][1] = gap;
    q->lc[4][0] = 0.0;  q->lc[4][1] = -gap;
    q->lc[5][0] = gap;  q->lc[5][1] = 0.0;
    q->lc[6][0] = 0.0;  q->lc[6][1] = gap;
    q->lc[7][0] = -gap; q->lc[7][1] = 0.0;
    /* quadinfo::init (DBEM1/quadinfo.cpp:92-98) */
    for (m = 0; m < NG2; ++m)
        q->nrgv0[m] = quad_n0(q, q->rp[m][0], q->rp[m][1]) * q->ra[m];
    /* patches (DBEM1/quadinfo.cpp:102-138) */
    {
        double LocLength = 2.0 / NPW;
        for (i = 0; i < NPW; ++i)
            for (j = 0; j < NPW; ++j) {
                p = i * NPW + j;
                q->cenpw[p][0] = -1.0 + LocLength / 2.0 + LocLength * j;
                q->cenpw[p][1] = -1.0 + LocLength / 2.0 + LocLength * i;
                for (m = 0; m < NPWG2; ++m) {
                    q->lpospw[p][m][0] = q->cenpw[p][0] + q->pwrp[m][0] / NPW;
                    q->lpospw[p][m][1] = q->cenpw[p][1] + q->pwrp[m][1] / NPW;
                }
            }
        for (i = 0; i < NPW2; ++i)
            for (m = 0; m < NPWG2; ++m)
                q->pwgv[i][m] = q->pwra[m] * (LocLength / 2.0) * (LocLength / 2.0);
    }
    /* singular patches (DBEM1/quadinfo.cpp:142-164) */
    {
        double LocLength = 2.0 / NSPW;
        double cen[2];
        for (i = 0; i < NSPW; ++i)
            for (j = 0; j < NSPW; ++j) {
                p = i * NSPW + j;
                cen[0] = -1.0 + LocLength / 2.0 + LocLength * j;
                cen[1] = -1.0 + LocLength / 2.0 + LocLength * i;
                for (m = 0; m < NSPWG2; ++m) {
                    q->sinlpospw[p][m][0] = cen[0] + q->pwsp[m][0] / NSPW;
                    q->sinlpospw[p][m][1] = cen[1] + q->pwsp[m][1] / NSPW;
                }
            }
        for (m = 0; m < NSPWG2; ++m)
            q->sinnpwgv[m] = q->pwsa[m] * (LocLength / 2.0) * (LocLength / 2.0);
    }
    /* quadinfo::AssignSecInfo constant part (DBEM1/quadinfo.cpp:230-263,291-306) */
    for (a = 0; a < 4; ++a)
        for (n = 0; n < NSG2; ++n) {
            double th = const_theta_from_local(q->sp[n][0], a);
            double tempFA = 0.25 * const_alpha(th, a) * const_theta_diff();
            double rou = const_polar_from_local(q->sp[n][1], th, a);
            q->secconstpos[a][0][n] = rou * ocos(th);
            q->secconstpos[a][1][n] = rou * osin(th);
            q->secconstval[a][n] = rou * tempFA * q->sa[n];
            /* AssignFValue (DBEM1/quadinfo.cpp:335-345) */
            q->constfm1secval[a][n] = (rou * tempFA * q->sa[n]) / rou / rou;
        }
    for (a = 0; a < 4; ++a)
        for (n = 0; n < NSG; ++n) {
            double th = const_theta_from_local(q->spos[n], a);
            double rou = const_polar_from_local(1.0, th, a);
            q->constlinepos[a][n] = th;
            /* DBEM1/quadinfo.cpp:367-376 */
            q->constfm1lineval[a][n] = 0.5 * const_theta_diff() * log(rou) * q->sb[n];
        }
    for (a = 0; a < 4; ++a)
        for (n = 0; n < NSPW2; ++n)
            for (m = 0; m < NSPWG2; ++m) {
                double th = const_theta_from_local(q->sinlpospw[n][m][0], a);
                double tempFA = 0.25 * const_alpha(th, a) * const_theta_diff();
                double rou = const_polar_from_local(q->sinlpospw[n][m][1], th, a);
                q->secconstpospw[a][0][n][m] = rou * ocos(th);
                q->secconstpospw[a][1][n][m] = rou * osin(th);
                q->secconstvalpw[a][n][m] = rou * tempFA * q->sinnpwgv[m];
            }
}

/* ------------------------------------------------------------------------ */
/* Element geometry (DBEM1/dsquareelement.h:359-507,582-620;                 */
/*                   DBEM1/dsquareelement.cpp:332-418,990-1003)              */
/* ------------------------------------------------------------------------ */
typedef struct {
    double cc[3][8];       /* m_CoordCoef */
    double center[3];
    double radius, radiusPW;
    double rcenpw[NPW2][3];
    double trans[9];       /* m_transmat of the element's node */
    double u1[9], t1[9], t2[9]; /* m_OnElement{U1,T1,T2}ij[0][0] */
    int bcid;
    int domain;
    int surface;           /* 0 outer, 1 interface */
} elem;

static void get_coef(double* c, const double* p)
{
    c[0] = 0.25 * (-p[0] - p[3] + 2 * p[5] - p[2] - p[1] + 2 * p[4] + 2 * p[7] + 2 * p[6]);
    c[1] = 0.5 * (p[5] - p[7]);
    c[2] = 0.5 * (-p[4] + p[6]);
    c[3] = 0.25 * (p[0] - p[1] + p[2] - p[3]);
    c[4] = 0.25 * (-2 * p[4] + p[3] + p[2] - 2 * p[6] + p[1] + p[0]);
    c[5] = 0.25 * (-2 * p[5] + p[1] - 2 * p[7] + p[2] + p[0] + p[3]);
    c[6] = 0.25 * (2 * p[4] + p[3] + p[2] - 2 * p[6] - p[1] - p[0]);
    c[7] = 0.25 * (2 * p[7] - 2 * p[5] + p[1] - p[3] + p[2] - p[0]);
}

static void from_local(const elem* e, double s1, double s2, double* pt)
{
    double s12 = s1 * s1, s22 = s2 * s2;
    double s1s22 = s1 * s22, s12s2 = s12 * s2;
    int i;
    for (i = 0; i < 3; ++i) {
        const double* c = e->cc[i];
        pt[i] = c[0] + c[1] * s1 + c[2] * s2 + c[3] * s1 * s2 + c[4] * s12 + c[5] * s22 + c[6] * s12s2 +
                c[7] * s1s22;
    }
}

static void rdiffs1(const elem* e, double s1, double s2, double* r)
{
    double s1_2 = 2.0 * s1, s1s2_2 = 2.0 * s1 * s2, s22 = s2 * s2;
    int i;
    for (i = 0; i < 3; ++i) {
        const double* c = e->cc[i];
        r[i] = c[1] + c[3] * s2 + c[4] * s1_2 + c[6] * s1s2_2 + c[7] * s22;
    }
}

static void rdiffs2(const elem* e, double s1, double s2, double* r)
{
    double s2_2 = 2.0 * s2, s1s2_2 = 2.0 * s1 * s2, s12 = s1 * s1;
    int i;
    for (i = 0; i < 3; ++i) {
        const double* c = e->cc[i];
        r[i] = c[2] + c[3] * s1 + c[5] * s2_2 + c[6] * s12 + c[7] * s1s2_2;
    }
}

static void normal_at(const elem* e, double s1, double s2, double* n)
{
    double r1[3], r2[3], t;
    rdiffs1(e, s1, s2, r1);
    rdiffs2(e, s1, s2, r2);
    n[0] = r1[1] * r2[2] - r2[1] * r1[2];
    n[1] = r1[2] * r2[0] - r2[2] * r1[0];
    n[2] = r1[0] * r2[1] - r2[0] * r1[1];
    t = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] = n[0] / t;
    n[1] = n[1] / t;
    n[2] = n[2] / t;
}

static double jacobi(const elem* e, double s1, double s2)
{
    double ru[3], rv[3], E, F, G;
    rdiffs1(e, s1, s2, ru);
    rdiffs2(e, s1, s2, rv);
    E = ru[0] * ru[0] + ru[1] * ru[1] + ru[2] * ru[2];
    F = ru[0] * rv[0] + ru[1] * rv[1] + ru[2] * rv[2];
    G = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
    return sqrt(fabs(E * G - F * F));
}

static void translate_matrix(const elem* e, double s1, double s2, double* T)
{
    double res[3], len, temp;
    rdiffs1(e, s1, s2, res);
    len = sqrt(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
    T[0] = res[0] / len;
    T[1] = res[1] / len;
    T[2] = res[2] / len;
    rdiffs2(e, s1, s2, res);
    len = sqrt(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
    T[3] = res[0] / len;
    T[4] = res[1] / len;
    T[5] = res[2] / len;
    T[6] = T[1] * T[5] - T[4] * T[2];
    T[7] = T[3] * T[2] - T[0] * T[5];
    T[8] = T[0] * T[4] - T[3] * T[1];
    temp = sqrt(T[6] * T[6] + T[7] * T[7] + T[8] * T[8]);
    T[6] /= temp;
    T[7] /= temp;
    T[8] /= temp;
    T[3] = T[7] * T[2] - T[8] * T[1];
    T[4] = T[8] * T[0] - T[6] * T[2];
    T[5] = T[6] * T[1] - T[7] * T[0];
}

/* DBEM1/geo.h Dist */
static double dist3(const double* a, const double* b)
{
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

/* ------------------------------------------------------------------------ */
/* Kernels (DBEM1/dsquareelement.h:631-903, DBEM1/dsquareelement.cpp:1070-1590) */
/* ------------------------------------------------------------------------ */
static int fai(double n, double dt, double r, double c)
{
    double ndt = n * dt;
    double roverc = r / c;
    return ((roverc >= ndt) && (roverc < (ndt + dt))) ? 1 : 0;
}

static int iseries(const dmat* m, double n, double dt, const double* R, double* res)
{
    double ndt = n * dt;
    double np1dt = ndt + dt;
    double rc1 = R[1] / m->C1;
    double rc2 = R[1] / m->C2;
    double t1;
    if (np1dt <= rc1 || ndt >= rc2) {
        *res = 0.0;
        return 0;
    } else if (ndt >= rc1 && np1dt <= rc2) {
        *res = (n + 0.5) * m->Dt_2 / R[2];
        return 1;
    } else if (ndt <= rc1 && (np1dt >= rc1 && np1dt <= rc2)) {
        t1 = np1dt / R[1];
        *res = 0.5 * (t1 * t1 - m->OneOverC1_2);
        return 1;
    } else if (ndt <= rc1 && np1dt >= rc2) {
        *res = m->Half_1OC22_1OC12;
        return 1;
    } else if (np1dt >= rc2 && (ndt >= rc1 && ndt <= rc2)) {
        t1 = ndt / R[1];
        *res = 0.5 * (m->OneOverC2_2 - t1 * t1);
        return 1;
    }
    *res = 0.0;
    return 0;
}

static int jseries(const dmat* m, double n, double dt, const double* R, double* res)
{
    double ndt = n * dt;
    double np1dt = ndt + dt;
    double rc1 = R[1] / m->C1;
    double rc2 = R[1] / m->C2;
    double t1;
    if (np1dt <= rc1 || ndt >= rc2) {
        *res = 0.0;
        return 0;
    } else if (ndt >= rc1 && np1dt <= rc2) {
        t1 = m->Dt_3 / R[3];
        *res = (n * n + n + 0.33333333333333333) * t1;
        return 1;
    } else if (ndt <= rc1 && (np1dt >= rc1 && np1dt <= rc2)) {
        t1 = np1dt / R[1];
        *res = 0.33333333333333 * (t1 * t1 * t1 - m->OneOverC1_3);
        return 1;
    } else if (ndt <= rc1 && np1dt >= rc2) {
        *res = m->OneOverThree_1OC23_1OC13;
        return 1;
    } else if (np1dt >= rc2 && (ndt >= rc1 && ndt <= rc2)) {
        t1 = ndt / R[1];
        *res = 0.333333333333333 * (m->OneOverC2_3 - t1 * t1 * t1);
        return 1;
    }
    *res = 0.0;
    return 0;
}

static void rirj(const double* RI, double* x)
{
    x[0] = RI[1] * RI[1];
    x[1] = RI[1] * RI[2];
    x[2] = RI[1] * RI[3];
    x[3] = RI[2] * RI[2];
    x[4] = RI[2] * RI[3];
    x[5] = RI[3] * RI[3];
}

static void get_aij(const dmat* m, double* a, const double* R, const double* X)
{
    double C1 = m->OneOverC1_2 / R[3];
    a[0] = C1 * X[0];
    a[1] = C1 * X[1];
    a[2] = C1 * X[2];
    a[3] = a[1];
    a[4] = C1 * X[3];
    a[5] = C1 * X[4];
    a[6] = a[2];
    a[7] = a[5];
    a[8] = C1 * X[5];
}

static void get_bij(const dmat* m, double* b, const double* R, const double* X)
{
    b[0] = m->OneOverC2_2 * (1.0 / R[1] - X[0] / R[3]);
    b[1] = m->OneOverC2_2 * (-X[1] / R[3]);
    b[2] = m->OneOverC2_2 * (-X[2] / R[3]);
    b[3] = b[1];
    b[4] = m->OneOverC2_2 * (1.0 / R[1] - X[3] / R[3]);
    b[5] = m->OneOverC2_2 * (-X[4] / R[3]);
    b[6] = b[2];
    b[7] = b[5];
    b[8] = m->OneOverC2_2 * (1.0 / R[1] - X[5] / R[3]);
}

static void get_cij(double* c, const double* R, const double* X)
{
    double C1 = 3.0 / R[3];
    double C2 = 1.0 / R[1];
    c[0] = C1 * X[0] - C2;
    c[1] = C1 * X[1];
    c[2] = C1 * X[2];
    c[3] = c[1];
    c[4] = C1 * X[3] - C2;
    c[5] = C1 * X[4];
    c[6] = c[2];
    c[7] = c[5];
    c[8] = C1 * X[5] - C2;
}

static void get_dij(const dmat* m, double* d, const double* R, const double* RI, const double* N, const double* X)
{
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = (4.0 * m->C2_2OverC1_2 - 1.0) / R[3];
    double C2 = m->C2_2OverC1_2 * 12.0 * RmNm / R[5];
    double C3 = m->C2_2OverC1_2 * 2.0 / R[3];
    d[0] = C1 * RI[1] * N[0] - (C2 * X[0] - C3 * (RI[1] * N[0] + RmNm));
    d[1] = C1 * RI[2] * N[0] - (C2 * X[1] - C3 * (RI[1] * N[1]));
    d[2] = C1 * RI[3] * N[0] - (C2 * X[2] - C3 * (RI[1] * N[2]));
    d[3] = C1 * RI[1] * N[1] - (C2 * X[1] - C3 * (RI[2] * N[0]));
    d[4] = C1 * RI[2] * N[1] - (C2 * X[3] - C3 * (RI[2] * N[1] + RmNm));
    d[5] = C1 * RI[3] * N[1] - (C2 * X[4] - C3 * (RI[2] * N[2]));
    d[6] = C1 * RI[1] * N[2] - (C2 * X[2] - C3 * (RI[3] * N[0]));
    d[7] = C1 * RI[2] * N[2] - (C2 * X[4] - C3 * (RI[3] * N[1]));
    d[8] = C1 * RI[3] * N[2] - (C2 * X[5] - C3 * (RI[3] * N[2] + RmNm));
}

static void get_eij(double* e, const double* R, const double* RI, const double* N, const double* X)
{
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = 12.0 * RmNm / R[5];
    double C2 = 2.0 / R[3];
    double C3 = 3.0 / R[3];
    e[0] = C1 * X[0] - C2 * RI[1] * N[0] - C3 * (RI[1] * N[0] + RmNm);
    e[1] = C1 * X[1] - C2 * RI[2] * N[0] - C3 * (RI[1] * N[1]);
    e[2] = C1 * X[2] - C2 * RI[3] * N[0] - C3 * (RI[1] * N[2]);
    e[3] = C1 * X[1] - C2 * RI[1] * N[1] - C3 * (RI[2] * N[0]);
    e[4] = C1 * X[3] - C2 * RI[2] * N[1] - C3 * (RI[2] * N[1] + RmNm);
    e[5] = C1 * X[4] - C2 * RI[3] * N[1] - C3 * (RI[2] * N[2]);
    e[6] = C1 * X[2] - C2 * RI[1] * N[2] - C3 * (RI[3] * N[0]);
    e[7] = C1 * X[4] - C2 * RI[2] * N[2] - C3 * (RI[3] * N[1]);
    e[8] = C1 * X[5] - C2 * RI[3] * N[2] - C3 * (RI[3] * N[2] + RmNm);
}

static void get_fij(const dmat* m, double* f, const double* R, const double* RI, const double* N, const double* X)
{
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = -6.0 * m->C2_2;
    double C2 = 5.0 * RmNm / R[5];
    f[0] = C1 * (C2 * X[0] - (RI[1] * N[0] + RI[1] * N[0] + RmNm) / R[3]);
    f[1] = C1 * (C2 * X[1] - (RI[2] * N[0] + RI[1] * N[1]) / R[3]);
    f[2] = C1 * (C2 * X[2] - (RI[3] * N[0] + RI[1] * N[2]) / R[3]);
    f[3] = f[1];
    f[4] = C1 * (C2 * X[3] - (RI[2] * N[1] + RI[2] * N[1] + RmNm) / R[3]);
    f[5] = C1 * (C2 * X[4] - (RI[3] * N[1] + RI[2] * N[2]) / R[3]);
    f[6] = f[2];
    f[7] = f[5];
    f[8] = C1 * (C2 * X[5] - (RI[3] * N[2] + RI[3] * N[2] + RmNm) / R[3]);
}

static void get_gij(const dmat* m, double* g, const double* R, const double* RI, const double* N, const double* X)
{
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = -2.0 * m->C2_2OverC1_3 * RmNm / R[4];
    double C2 = 1.0 / m->C1 * (2.0 * m->C2_2OverC1_2 - 1.0) / R[2];
    g[0] = C1 * X[0] + C2 * RI[1] * N[0];
    g[1] = C1 * X[1] + C2 * RI[2] * N[0];
    g[2] = C1 * X[2] + C2 * RI[3] * N[0];
    g[3] = C1 * X[1] + C2 * RI[1] * N[1];
    g[4] = C1 * X[3] + C2 * RI[2] * N[1];
    g[5] = C1 * X[4] + C2 * RI[3] * N[1];
    g[6] = C1 * X[2] + C2 * RI[1] * N[2];
    g[7] = C1 * X[4] + C2 * RI[2] * N[2];
    g[8] = C1 * X[5] + C2 * RI[3] * N[2];
}

static void get_hij(const dmat* m, double* h, const double* R, const double* RI, const double* N, const double* X)
{
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = 2.0 / m->C2 * RmNm / R[4];
    double C2 = -1.0 / m->C2 / R[2];
    h[0] = C1 * X[0] + C2 * (RI[1] * N[0] + RmNm);
    h[1] = C1 * X[1] + C2 * RI[1] * N[1];
    h[2] = C1 * X[2] + C2 * RI[1] * N[2];
    h[3] = C1 * X[1] + C2 * RI[2] * N[0];
    h[4] = C1 * X[3] + C2 * (RI[2] * N[1] + RmNm);
    h[5] = C1 * X[4] + C2 * RI[2] * N[2];
    h[6] = C1 * X[2] + C2 * RI[3] * N[0];
    h[7] = C1 * X[4] + C2 * RI[3] * N[1];
    h[8] = C1 * X[5] + C2 * (RI[3] * N[2] + RmNm);
}

/* GetDynaUij (DBEM1/dsquareelement.cpp:1238-1342) */
static int dyna_u(const dmat* m, double* U, double n, const double* R, const double* RI)
{
    double K[9], X[6], res, t1;
    int flag = 0, k, have = 0;
    for (k = 0; k < 9; ++k)
        U[k] = 0.0;
    if (fai(n, m->Dt, R[1], m->C1)) {
        flag++;
        rirj(RI, X);
        have = 1;
        get_aij(m, K, R, X);
        t1 = m->OneOverFourPaiRou;
        for (k = 0; k < 9; ++k)
            U[k] += K[k] * t1;
    }
    if (fai(n, m->Dt, R[1], m->C2)) {
        flag++;
        if (!have) {
            rirj(RI, X);
            have = 1;
        }
        get_bij(m, K, R, X);
        t1 = m->OneOverFourPaiRou;
        for (k = 0; k < 9; ++k)
            U[k] += K[k] * t1;
    }
    if (iseries(m, n, m->Dt, R, &res)) {
        flag++;
        if (!have) {
            rirj(RI, X);
            have = 1;
        }
        get_cij(K, R, X);
        t1 = m->OneOverFourPaiRou * res;
        for (k = 0; k < 9; ++k)
            U[k] += K[k] * t1;
    }
    return flag;
}

/* GetDynaT1ij (DBEM1/dsquareelement.cpp:1344-1462) */
static int dyna_t1(const dmat* m, double* T, double n, const double* R, const double* RI, const double* N)
{
    double K1[9], K2[9], X[6], ri, rj, t1, t2;
    int flag = 0, k;
    for (k = 0; k < 9; ++k)
        T[k] = 0.0;
    rirj(RI, X);
    if (fai(n, m->Dt, R[1], m->C1)) {
        flag++;
        get_dij(m, K1, R, RI, N, X);
        get_gij(m, K2, R, RI, N, X);
        t1 = kOneOver4Pi * (n + 1.0 - R[1] / m->C1Dt);
        t2 = kOneOver4Pi / m->Dt;
        for (k = 0; k < 9; ++k)
            T[k] += K1[k] * t1 + K2[k] * t2;
    }
    if (fai(n, m->Dt, R[1], m->C2)) {
        flag++;
        get_eij(K1, R, RI, N, X);
        get_hij(m, K2, R, RI, N, X);
        t1 = kOneOver4Pi * (n + 1.0 - R[1] / m->C2Dt);
        t2 = kOneOver4Pi / m->Dt;
        for (k = 0; k < 9; ++k)
            T[k] += K1[k] * t1 + K2[k] * t2;
    }
    if (iseries(m, n, m->Dt, R, &ri)) {
        flag++;
        jseries(m, n, m->Dt, R, &rj);
        get_fij(m, K1, R, RI, N, X);
        t1 = kOneOver4Pi * ((n + 1.0) * ri - R[1] * rj / m->Dt);
        for (k = 0; k < 9; ++k)
            T[k] += K1[k] * t1;
    }
    return flag;
}

/* GetDynaT2ij (DBEM1/dsquareelement.cpp:1464-1582) */
static int dyna_t2(const dmat* m, double* T, double n, const double* R, const double* RI, const double* N)
{
    double K1[9], K2[9], X[6], ri, rj, t1, t2;
    int flag = 0, k;
    for (k = 0; k < 9; ++k)
        T[k] = 0.0;
    rirj(RI, X);
    if (fai(n - 1, m->Dt, R[1], m->C1)) {
        flag++;
        get_dij(m, K1, R, RI, N, X);
        get_gij(m, K2, R, RI, N, X);
        t1 = kOneOver4Pi * (1.0 - n + R[1] / m->C1Dt);
        t2 = -kOneOver4Pi / m->Dt;
        for (k = 0; k < 9; ++k)
            T[k] += K1[k] * t1 + K2[k] * t2;
    }
    if (fai(n - 1, m->Dt, R[1], m->C2)) {
        flag++;
        get_eij(K1, R, RI, N, X);
        get_hij(m, K2, R, RI, N, X);
        t1 = kOneOver4Pi * (1.0 - n + R[1] / m->C2Dt);
        t2 = -kOneOver4Pi / m->Dt;
        for (k = 0; k < 9; ++k)
            T[k] += K1[k] * t1 + K2[k] * t2;
    }
    if (iseries(m, n - 1, m->Dt, R, &ri)) {
        flag++;
        jseries(m, n - 1, m->Dt, R, &rj);
        get_fij(m, K1, R, RI, N, X);
        t1 = kOneOver4Pi * ((1.0 - n) * ri + R[1] * rj / m->Dt);
        for (k = 0; k < 9; ++k)
            T[k] += K1[k] * t1;
    }
    return flag;
}

static int dyna_t(const dmat* m, int typeT, double* T, double n, const double* R, const double* RI,
                  const double* N)
{
    return typeT == 1 ? dyna_t1(m, T, n, R, RI, N) : dyna_t2(m, T, n, R, RI, N);
}

/* GetR with powers (DBEM1/dsquareelement.h:631-645) */
static void get_r5(const double* src, const double* pt, double* R, double* RI)
{
    RI[1] = pt[0] - src[0];
    RI[2] = pt[1] - src[1];
    RI[3] = pt[2] - src[2];
    R[1] = sqrt(RI[1] * RI[1] + RI[2] * RI[2] + RI[3] * RI[3]);
    R[2] = R[1] * R[1];
    R[3] = R[2] * R[1];
    R[4] = R[3] * R[1];
    R[5] = R[4] * R[1];
    RI[0] = R[1];
}

/* GetR (4-value form, DBEM1/dsquareelement.h:511-523) */
static void get_r4(const elem* e, const double* src, double s1, double s2, double* R)
{
    double p[3];
    from_local(e, s1, s2, p);
    R[1] = p[0] - src[0];
    R[2] = p[1] - src[1];
    R[3] = p[2] - src[2];
    R[0] = sqrt(R[1] * R[1] + R[2] * R[2] + R[3] * R[3]);
}

/* GetTij, static traction kernel (DBEM1/dsquareelement.cpp:496-543) */
static void static_tij(const dmat* m, double* T, const double* R, const double* n)
{
    double t1 = R[0] * R[0];
    double t2 = m->STCoef1 / t1;
    double t3 = m->STCoef2 / R[0];
    double drdn = (R[1] * n[0] + R[2] * n[1] + R[3] * n[2]) / R[0];
    double t4, t5, t6, t7, t8, t9, t10;
    t5 = t3 * (R[2] * n[0] - R[1] * n[1]);
    t6 = t3 * (R[3] * n[0] - R[1] * n[2]);
    t7 = t3 * (R[3] * n[1] - R[2] * n[2]);
    t8 = t2 * drdn;
    t9 = t8 * m->STCoef2;
    t10 = t8 * 3.0 / t1;
    t4 = t10 * R[1];
    T[0] = t9 + t4 * R[1];
    T[1] = t4 * R[2] - t2 * t5;
    T[2] = t4 * R[3] - t2 * t6;
    T[3] = t4 * R[2] + t2 * t5;
    T[4] = t9 + t10 * R[2] * R[2];
    t4 = t10 * R[3];
    T[5] = t4 * R[2] - t2 * t7;
    T[6] = t4 * R[1] + t2 * t6;
    T[7] = t4 * R[2] + t2 * t7;
    T[8] = t9 + t10 * R[3] * R[3];
}

/* GetFijMinusOne (DBEM1/dsquareelement.cpp:609-639) */
static void fij_minus_one(double* F, const double* A, const double* J, double coef1, double coef2)
{
    double AkJk0 = A[1] * J[0] + A[2] * J[1] + A[3] * J[2];
    double t1, t2, t3, t4, temp;
    t1 = 3.0 * A[1] * AkJk0 / A[5];
    t2 = coef2 * (A[2] * J[0] - A[1] * J[1]) / A[4];
    t3 = coef2 * (A[3] * J[0] - A[1] * J[2]) / A[4];
    t4 = coef2 * (A[3] * J[1] - A[2] * J[2]) / A[4];
    temp = coef2 * AkJk0 / A[4];
    F[0] = coef1 * (temp + t1 * A[1]);
    F[1] = coef1 * (t1 * A[2] - t2);
    F[2] = coef1 * (t1 * A[3] - t3);
    F[3] = coef1 * (t1 * A[2] + t2);
    F[4] = coef1 * (temp + 3.0 * A[2] * A[2] * AkJk0 / A[5]);
    t1 = 3.0 * A[3] * AkJk0 / A[5];
    F[5] = coef1 * (t1 * A[2] - t4);
    F[6] = coef1 * (t1 * A[1] + t3);
    F[7] = coef1 * (t1 * A[2] + t4);
    F[8] = coef1 * (temp + t1 * A[3]);
}

/* ------------------------------------------------------------------------ */
/* Causality judges (DBEM1/dsquareelement.cpp:1156-1236, .h:926-943)         */
/* ------------------------------------------------------------------------ */
static int need_cal(const dmat* m, const elem* e, int type, double n, const double* o)
{
    double nn = (type == 2) ? n - 1 : n;
    double ndt = nn * m->Dt;
    double np1dt = (nn + 1) * m->Dt;
    double r = dist3(e->center, o);
    double Rmax = r + e->radius;
    double Rmin = r - e->radius;
    double a1, a2, b1, b2;
    if (Rmin < 0.0)
        Rmin = 0.0;
    a1 = Rmax / m->C1;
    a2 = Rmax / m->C2;
    b1 = Rmin / m->C1;
    b2 = Rmin / m->C2;
    if (ndt >= a2 || np1dt <= b1)
        return 0;
    else if (ndt <= b1 && (np1dt >= a1 && np1dt <= b2))
        return 2;
    else if (ndt <= b1 && np1dt >= a2)
        return 2;
    else if (ndt >= a1 && np1dt <= b2)
        return 2;
    else if ((ndt >= a1 && ndt <= b2) && np1dt >= a2)
        return 2;
    return 1;
}

static int pw_need_cal(const dmat* m, const elem* e, int type, double n, const double* pwc, const double* o)
{
    double nn = (type == 2 ? n - 1 : n);
    double c2ndt = nn * m->C2Dt;
    double c1np1dt = (nn + 1) * m->C1Dt;
    double r = dist3(pwc, o);
    double rpr = r + e->radiusPW;
    double rmr = r - e->radiusPW;
    return (rmr >= c1np1dt || rpr < c2ndt) ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* Quadrature (DBEM1/dsquareelementdyna.cpp:794-848,940-994;                 */
/*             DBEM1/dsquareelement.cpp:1955-2009,2233-2290)                 */
/* ------------------------------------------------------------------------ */
static void int_u_pw(const qtables* q, const dmat* m, const elem* e, const double* src, double n, double* out)
{
    double R[6], RI[4], UT[9], pt[3], J, temp;
    int i, j;
    out[0] = out[1] = out[2] = out[4] = out[5] = out[8] = 0.0;
    for (i = 0; i < NPW2; ++i) {
        if (!pw_need_cal(m, e, 1, n, e->rcenpw[i], src))
            continue;
        for (j = 0; j < NPWG2; ++j) {
            from_local(e, q->lpospw[i][j][0], q->lpospw[i][j][1], pt);
            get_r5(src, pt, R, RI);
            dyna_u(m, UT, n, R, RI);
            J = jacobi(e, q->lpospw[i][j][0], q->lpospw[i][j][1]);
            temp = q->pwgv[i][j] * J;
            out[0] += UT[0] * temp;
            out[1] += UT[1] * temp;
            out[2] += UT[2] * temp;
            out[4] += UT[4] * temp;
            out[5] += UT[5] * temp;
            out[8] += UT[8] * temp;
        }
    }
    out[3] = out[1];
    out[6] = out[2];
    out[7] = out[5];
}

static void int_t_pw(const qtables* q, const dmat* m, const elem* e, int typeT, const double* src, double n,
                     double* out)
{
    double R[6], RI[4], UT[9], pt[3], nor[3], J, temp;
    int i, j, k;
    for (k = 0; k < 9; ++k)
        out[k] = 0.0;
    for (i = 0; i < NPW2; ++i) {
        if (!pw_need_cal(m, e, typeT, n, e->rcenpw[i], src))
            continue;
        for (j = 0; j < NPWG2; ++j) {
            from_local(e, q->lpospw[i][j][0], q->lpospw[i][j][1], pt);
            get_r5(src, pt, R, RI);
            normal_at(e, q->lpospw[i][j][0], q->lpospw[i][j][1], nor);
            dyna_t(m, typeT, UT, n, R, RI, nor);
            J = jacobi(e, q->lpospw[i][j][0], q->lpospw[i][j][1]);
            temp = q->pwgv[i][j] * J;
            for (k = 0; k < 9; ++k)
                out[k] += UT[k] * temp;
        }
    }
}

/* regular path: weights m_NRGV[0] = N0(gap)*w (quirk Q1) */
static void int_u_reg(const qtables* q, const dmat* m, const elem* e, const double* src, double n, double* out)
{
    double R[6], RI[4], UT[9], pt[3], J, temp;
    int i;
    out[0] = out[1] = out[2] = out[4] = out[5] = out[8] = 0.0;
    for (i = 0; i < NG2; ++i) {
        from_local(e, q->rp[i][0], q->rp[i][1], pt);
        get_r5(src, pt, R, RI);
        if (dyna_u(m, UT, n, R, RI)) {
            J = jacobi(e, q->rp[i][0], q->rp[i][1]);
            temp = q->nrgv0[i] * J;
            out[0] += UT[0] * temp;
            out[1] += UT[1] * temp;
            out[2] += UT[2] * temp;
            out[4] += UT[4] * temp;
            out[5] += UT[5] * temp;
            out[8] += UT[8] * temp;
        }
    }
    out[3] = out[1];
    out[6] = out[2];
    out[7] = out[5];
}

/* regular T path: T1(n)+T2(n) whatever typeT is (quirk Q2) */
static void int_t_reg(const qtables* q, const dmat* m, const elem* e, const double* src, double n, double* out)
{
    double R[6], RI[4], U1[9], U2[9], pt[3], nor[3], J, temp;
    int i, k, flag;
    for (k = 0; k < 9; ++k)
        out[k] = 0.0;
    for (i = 0; i < NG2; ++i) {
        from_local(e, q->rp[i][0], q->rp[i][1], pt);
        get_r5(src, pt, R, RI);
        normal_at(e, q->rp[i][0], q->rp[i][1], nor);
        flag = dyna_t(m, 1, U1, n, R, RI, nor);
        flag += dyna_t(m, 2, U2, n, R, RI, nor);
        if (flag) {
            J = jacobi(e, q->rp[i][0], q->rp[i][1]);
            temp = q->nrgv0[i] * J;
            for (k = 0; k < 9; ++k)
                out[k] += (U1[k] + U2[k]) * temp;
        }
    }
}

/* IntDynaUijJudge / IntDynaTijJudge (DBEM1/dsquareelementdyna.cpp:1603-1616,1677-1690) */
static int judge_u(const qtables* q, const dmat* m, const elem* e, const double* src, double n, double* out)
{
    int f = need_cal(m, e, 1, n, src);
    if (f == 1)
        int_u_pw(q, m, e, src, n, out);
    else if (f == 2)
        int_u_reg(q, m, e, src, n, out);
    return f;
}

static int judge_t(const qtables* q, const dmat* m, const elem* e, int typeT, const double* src, double n,
                   double* out)
{
    int f = need_cal(m, e, typeT, n, src);
    if (f == 1)
        int_t_pw(q, m, e, typeT, src, n, out);
    else if (f == 2)
        int_t_reg(q, m, e, src, n, out);
    return f;
}

/* ------------------------------------------------------------------------ */
/* Singular self integrals (DBEM1/dsquareelement.cpp:2292-2427,              */
/*                          DBEM1/dsquareelementdyna.cpp:239-490)            */
/* ------------------------------------------------------------------------ */
static void self_integrals(const qtables* q, const dmat* m, elem* e)
{
    double node[3], st[9], R[6], RI[4], SR[4], UT[9], DT1[9], DT2[9], nor[3];
    double oneJ[4][NSG2], oneN[4][NSG2][3], oneR[4][NSG2][4];
    double coef1, coef2, Jh[3], D1[3], D2[3], F[9], A[6], temp, s1, s2, Jac;
    int a, j, l, k, kk;
    from_local(e, 0.0, 0.0, node);
    /* InElementStaticTij */
    for (a = 0; a < 4; ++a)
        for (j = 0; j < NSG2; ++j) {
            s1 = q->secconstpos[a][0][j];
            s2 = q->secconstpos[a][1][j];
            oneJ[a][j] = jacobi(e, s1, s2);
            normal_at(e, s1, s2, oneN[a][j]);
            get_r4(e, node, s1, s2, oneR[a][j]);
        }
    coef1 = -1.0 / (8.0 * kPi * (1.0 - m->v));
    coef2 = 1.0 - 2.0 * m->v;
    for (k = 0; k < 9; ++k)
        st[k] = 0.0;
    {
        double tj = jacobi(e, q->lc[0][0], q->lc[0][1]);
        double n0[3];
        normal_at(e, q->lc[0][0], q->lc[0][1], n0);
        Jh[0] = tj * n0[0];
        Jh[1] = tj * n0[1];
        Jh[2] = tj * n0[2];
    }
    rdiffs1(e, 0.0, 0.0, D1);
    rdiffs2(e, 0.0, 0.0, D2);
    for (a = 0; a < 4; ++a)
        for (j = 0; j < NSG2; ++j) {
            static_tij(m, UT, oneR[a][j], oneN[a][j]);
            temp = q->secconstval[a][j] * oneJ[a][j];
            for (l = 0; l < 9; ++l)
                st[l] += UT[l] * temp;
        }
    for (a = 0; a < 4; ++a)
        for (j = 0; j < NSG; ++j) {
            double th = q->constlinepos[a][j];
            double FVL1;
            A[1] = D1[0] * ocos(th) + D2[0] * osin(th);
            A[2] = D1[1] * ocos(th) + D2[1] * osin(th);
            A[3] = D1[2] * ocos(th) + D2[2] * osin(th);
            A[0] = sqrt(A[1] * A[1] + A[2] * A[2] + A[3] * A[3]);
            A[4] = A[0] * A[0] * A[0];
            A[5] = A[4] * A[0] * A[0];
            fij_minus_one(F, A, Jh, coef1, coef2);
            for (l = 0; l < NSG; ++l) {
                double FVS1 = q->constfm1secval[a][l * NSG + j];
                for (kk = 0; kk < 9; ++kk)
                    st[kk] -= F[kk] * FVS1;
            }
            FVL1 = q->constfm1lineval[a][j];
            for (kk = 0; kk < 9; ++kk)
                st[kk] += F[kk] * FVL1;
        }
    /* PthInElementIntDynamic: U1(0) */
    e->u1[0] = e->u1[1] = e->u1[2] = e->u1[4] = e->u1[5] = e->u1[8] = 0.0;
    for (a = 0; a < 4; ++a)
        for (k = 0; k < NSPW2; ++k)
            for (l = 0; l < NSPWG2; ++l) {
                s1 = q->secconstpospw[a][0][k][l];
                s2 = q->secconstpospw[a][1][k][l];
                Jac = jacobi(e, s1, s2);
                get_r4(e, node, s1, s2, SR);
                R[1] = SR[0];
                R[2] = R[1] * R[1];
                R[3] = R[2] * R[1];
                R[4] = R[3] * R[1];
                R[5] = R[4] * R[1];
                RI[1] = SR[1];
                RI[2] = SR[2];
                RI[3] = SR[3];
                dyna_u(m, UT, 0.0, R, RI);
                temp = q->secconstvalpw[a][k][l] * Jac;
                e->u1[0] += UT[0] * temp;
                e->u1[1] += UT[1] * temp;
                e->u1[2] += UT[2] * temp;
                e->u1[4] += UT[4] * temp;
                e->u1[5] += UT[5] * temp;
                e->u1[8] += UT[8] * temp;
            }
    e->u1[3] = e->u1[1];
    e->u1[6] = e->u1[2];
    e->u1[7] = e->u1[5];
    /* T1(0) and T2(1) with the static part subtracted */
    for (kk = 0; kk < 9; ++kk) {
        e->t1[kk] = 0.0;
        e->t2[kk] = 0.0;
    }
    for (a = 0; a < 4; ++a)
        for (k = 0; k < NSPW2; ++k)
            for (l = 0; l < NSPWG2; ++l) {
                double ST[9];
                s1 = q->secconstpospw[a][0][k][l];
                s2 = q->secconstpospw[a][1][k][l];
                Jac = jacobi(e, s1, s2);
                normal_at(e, s1, s2, nor);
                get_r4(e, node, s1, s2, SR);
                R[1] = SR[0];
                R[2] = R[1] * R[1];
                R[3] = R[2] * R[1];
                R[4] = R[3] * R[1];
                R[5] = R[4] * R[1];
                RI[1] = SR[1];
                RI[2] = SR[2];
                RI[3] = SR[3];
                dyna_t1(m, DT1, 0.0, R, RI, nor);
                dyna_t2(m, DT2, 1.0, R, RI, nor);
                static_tij(m, ST, SR, nor);
                temp = q->secconstvalpw[a][k][l] * Jac;
                for (kk = 0; kk < 9; ++kk) {
                    e->t1[kk] += (DT1[kk] - ST[kk]) * temp;
                    e->t2[kk] += (DT2[kk] - ST[kk]) * temp;
                }
            }
    for (kk = 0; kk < 9; ++kk) {
        e->t1[kk] += st[kk];
        e->t2[kk] += st[kk];
    }
    e->t1[0] += 0.5;
    e->t1[4] += 0.5;
    e->t1[8] += 0.5;
}

/* ------------------------------------------------------------------------ */
/* Builder: MultiDomainCCSRBuilder semantics (DBEM1/multidomain.cpp:1120-1303) */
/* ------------------------------------------------------------------------ */
typedef struct {
    long row, col;
    double v[9];
    int alive;
} bentry;

typedef struct {
    long rows, cols;
    bentry* ent;
    long n, cap;
    long* hash; /* index into ent, -1 empty */
    long hcap;
} builder;

static void bld_init(builder* b, long rows, long cols)
{
    b->rows = rows;
    b->cols = cols;
    b->ent = 0;
    b->n = b->cap = 0;
    b->hcap = 0;
    b->hash = 0;
}

static void bld_free(builder* b)
{
    free(b->ent);
    free(b->hash);
    b->ent = 0;
    b->hash = 0;
    b->n = b->cap = b->hcap = 0;
}

static uint64_t bkey_hash(long r, long c)
{
    uint64_t x = ((uint64_t)r << 32) ^ (uint64_t)c;
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    return x;
}

static void bld_rehash(builder* b, long ncap)
{
    long i;
    free(b->hash);
    b->hash = (long*)malloc(sizeof(long) * (size_t)ncap);
    for (i = 0; i < ncap; ++i)
        b->hash[i] = -1;
    b->hcap = ncap;
    for (i = 0; i < b->n; ++i) {
        uint64_t h = bkey_hash(b->ent[i].row, b->ent[i].col) & (uint64_t)(ncap - 1);
        while (b->hash[h] >= 0)
            h = (h + 1) & (uint64_t)(ncap - 1);
        b->hash[h] = i;
    }
}

static bentry* bld_get(builder* b, long r, long c)
{
    uint64_t h;
    if (b->hcap == 0 || (b->n + 1) * 2 > b->hcap)
        bld_rehash(b, b->hcap ? b->hcap * 2 : 64);
    h = bkey_hash(r, c) & (uint64_t)(b->hcap - 1);
    while (b->hash[h] >= 0) {
        bentry* e = &b->ent[b->hash[h]];
        if (e->row == r && e->col == c)
            return e;
        h = (h + 1) & (uint64_t)(b->hcap - 1);
    }
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 64;
        b->ent = (bentry*)realloc(b->ent, sizeof(bentry) * (size_t)b->cap);
    }
    b->ent[b->n].row = r;
    b->ent[b->n].col = c;
    b->ent[b->n].alive = 0;
    b->hash[h] = b->n;
    b->n++;
    return &b->ent[b->n - 1];
}

static double max_abs9(const double* v)
{
    double m = 0.0, a;
    int i;
    for (i = 0; i < 9; ++i) {
        a = fabs(v[i]);
        if (a > m)
            m = a;
    }
    return m;
}

static int finite9(const double* v)
{
    int i;
    for (i = 0; i < 9; ++i)
        if (!(v[i] == v[i] && fabs(v[i]) <= 1.7976931348623157e308))
            return 0;
    return 1;
}

/* AddBlock: stores the transpose of the column-major input times sign,
 * with the absolute 1e-14 drop rule (DBEM1/multidomain.cpp:1120-1144). */
static void bld_add(builder* b, long row, long col, double sign, const double* blk)
{
    bentry* e;
    int r, c, k, allzero = 1;
    if (row < 0 || row >= b->rows || col < 0 || col >= b->cols)
        return;
    for (k = 0; k < 9; ++k)
        if (blk[k] != 0.0)
            allzero = 0;
    if (sign == 0.0 || allzero || max_abs9(blk) <= 1.0e-14)
        return;
    if (!finite9(blk))
        return;
    e = bld_get(b, row, col);
    if (!e->alive) {
        for (k = 0; k < 9; ++k)
            e->v[k] = 0.0;
        e->alive = 1;
    }
    for (r = 0; r < 3; ++r)
        for (c = 0; c < 3; ++c)
            e->v[r * 3 + c] += sign * blk[c * 3 + r];
    if (max_abs9(e->v) <= 1.0e-14)
        e->alive = 0;
}

typedef struct {
    long n, nnz;
    int* row_ptr;
    int* col;
    double* val; /* 9 per block, row-major */
} ccsr;

static int cmp_bent(const void* a, const void* b)
{
    const bentry* x = *(const bentry* const*)a;
    const bentry* y = *(const bentry* const*)b;
    if (x->row != y->row)
        return x->row < y->row ? -1 : 1;
    if (x->col != y->col)
        return x->col < y->col ? -1 : 1;
    return 0;
}

static void bld_build(const builder* b, ccsr* out)
{
    long i, nnz = 0, k;
    const bentry** list;
    for (i = 0; i < b->n; ++i)
        if (b->ent[i].alive)
            nnz++;
    list = (const bentry**)malloc(sizeof(bentry*) * (size_t)(nnz + 1));
    k = 0;
    for (i = 0; i < b->n; ++i)
        if (b->ent[i].alive)
            list[k++] = &b->ent[i];
    qsort(list, (size_t)nnz, sizeof(bentry*), cmp_bent);
    out->n = b->rows;
    out->nnz = nnz;
    out->row_ptr = (int*)calloc((size_t)b->rows + 1, sizeof(int));
    out->col = (int*)malloc(sizeof(int) * (size_t)(nnz + 1));
    out->val = (double*)malloc(sizeof(double) * 9 * (size_t)(nnz + 1));
    for (i = 0; i < nnz; ++i)
        out->row_ptr[list[i]->row + 1]++;
    for (i = 0; i < b->rows; ++i)
        out->row_ptr[i + 1] += out->row_ptr[i];
    for (i = 0; i < nnz; ++i) {
        out->col[i] = (int)list[i]->col;
        memcpy(&out->val[9 * i], list[i]->v, sizeof(double) * 9);
    }
    free(list);
}

static void ccsr_free(ccsr* m)
{
    free(m->row_ptr);
    free(m->col);
    free(m->val);
    m->row_ptr = 0;
    m->col = 0;
    m->val = 0;
    m->n = m->nnz = 0;
}

/* SparseMul CCSR (DBEM1/mat_vec.cpp:2273-2300), 0-based vectors */
static void spmv(const ccsr* M, const double* x, double* b)
{
    long i;
    int j;
    for (i = 0; i < M->n; ++i) {
        double b0 = 0.0, b1 = 0.0, b2 = 0.0;
        for (j = M->row_ptr[i]; j < M->row_ptr[i + 1]; ++j) {
            const double* v = &M->val[9 * (long)j];
            double x0 = x[3 * (long)M->col[j]], x1 = x[3 * (long)M->col[j] + 1], x2 = x[3 * (long)M->col[j] + 2];
            b0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
            b1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
            b2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
        }
        b[3 * i] = b0;
        b[3 * i + 1] = b1;
        b[3 * i + 2] = b2;
    }
}

/* SparseMul SymCCSR (DBEM1/mat_vec.cpp:2407-2434): (b00 b01 b02 b11 b12 b22) */
static void spmv_sym(const ccsr* M, const double* x, double* b)
{
    long i;
    int j;
    for (i = 0; i < M->n; ++i) {
        double b0 = 0.0, b1 = 0.0, b2 = 0.0;
        for (j = M->row_ptr[i]; j < M->row_ptr[i + 1]; ++j) {
            const double* f = &M->val[9 * (long)j];
            double v[6];
            double x0 = x[3 * (long)M->col[j]], x1 = x[3 * (long)M->col[j] + 1], x2 = x[3 * (long)M->col[j] + 2];
            v[0] = f[0];
            v[1] = f[1];
            v[2] = f[2];
            v[3] = f[4];
            v[4] = f[5];
            v[5] = f[8];
            b0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
            b1 += v[1] * x0 + v[3] * x1 + v[4] * x2;
            b2 += v[2] * x0 + v[4] * x1 + v[5] * x2;
        }
        b[3 * i] = b0;
        b[3 * i + 1] = b1;
        b[3 * i + 2] = b2;
    }
}

/* BuildSym symmetry test (DBEM1/multidomain.cpp:1216-1269) */
static int ccsr_is_sym(const ccsr* M, double tolerance)
{
    long p;
    for (p = 0; p < M->nnz; ++p) {
        const double* b = &M->val[9 * p];
        double scale = max_abs9(b);
        double th = tolerance;
        double mixed = 1.0e-14 + 1.0e-10 * scale;
        if (mixed > th)
            th = mixed;
        if (fabs(b[1] - b[3]) > th || fabs(b[2] - b[6]) > th || fabs(b[5] - b[7]) > th)
            return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Case, DOF map (DBEM1/multidomain.cpp:777-900,1364-1531)                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    long block;
    double sign;
    int known;
} vref;

typedef struct {
    int da, ea, db, eb, ns;
} itf_t;

struct orc_case {
    int ne, nd, ni;
    qtables q;
    dmat* mats;
    elem* el;
    long* dom_begin;
    long* dom_count;
    itf_t* itf;
    double dt;
    /* dof map */
    int dof_ok;
    long counts[9];
    vref *u, *t, *hu, *ht;
    long* pref;
    /* matrices */
    int assembled;
    int maxn;
    ccsr A, K;
    ccsr* mts;
    ccsr* mgs;
    int* mgs_sym;
    /* preconditioner */
    long leaf_count;
    long* leaf_size;
    long* leaf_begin;
    long* repid;
    double** prem; /* per leaf, (3c x 3c) row-major */
};

static void dof_build(orc_case* c)
{
    long N = c->ne, unknown = 0, known = 0, hU = 0, hT = 0, i;
    int* side = (int*)calloc((size_t)N, sizeof(int));
    long* cnt = c->counts;
    memset(cnt, 0, sizeof(long) * 9);
    c->u = (vref*)calloc((size_t)N, sizeof(vref));
    c->t = (vref*)calloc((size_t)N, sizeof(vref));
    c->hu = (vref*)calloc((size_t)N, sizeof(vref));
    c->ht = (vref*)calloc((size_t)N, sizeof(vref));
    c->pref = (long*)malloc(sizeof(long) * (size_t)N);
    for (i = 0; i < N; ++i) {
        c->u[i].block = c->t[i].block = c->hu[i].block = c->ht[i].block = -1;
        c->u[i].sign = c->t[i].sign = c->hu[i].sign = c->ht[i].sign = 1.0;
        c->u[i].known = c->t[i].known = 1;
        c->pref[i] = -1;
    }
    for (i = 0; i < c->ni; ++i) {
        const itf_t* f = &c->itf[i];
        double sB = (double)(f->ns == 0 ? -1 : f->ns);
        long ub = unknown++, tb = unknown++;
        long a = f->ea, b = f->eb;
        c->u[a].block = ub; c->u[a].sign = 1.0; c->u[a].known = 0;
        c->u[b].block = ub; c->u[b].sign = 1.0; c->u[b].known = 0;
        c->t[a].block = tb; c->t[a].sign = 1.0; c->t[a].known = 0;
        c->t[b].block = tb; c->t[b].sign = sB;  c->t[b].known = 0;
        c->pref[a] = ub;
        c->pref[b] = tb;
        c->hu[a].block = hU; c->hu[a].sign = 1.0;
        c->hu[b].block = hU; c->hu[b].sign = 1.0;
        hU++;
        c->ht[a].block = hT; c->ht[a].sign = 1.0;
        c->ht[b].block = hT; c->ht[b].sign = sB;
        hT++;
        side[a] = side[b] = 1;
        cnt[5]++;
        cnt[6]++;
    }
    c->dof_ok = 1;
    for (i = 0; i < N; ++i) {
        if (c->el[i].bcid != 123 && c->el[i].bcid != 456) {
            c->dof_ok = 0;
            break;
        }
        if (side[i])
            continue;
        if (c->el[i].bcid == 123) {
            long tb = unknown++;
            c->u[i].block = known++; c->u[i].sign = 1.0; c->u[i].known = 1;
            c->t[i].block = tb; c->t[i].sign = 1.0; c->t[i].known = 0;
            c->pref[i] = tb;
            cnt[4]++;
        } else {
            long ub = unknown++;
            c->u[i].block = ub; c->u[i].sign = 1.0; c->u[i].known = 0;
            c->t[i].block = known++; c->t[i].sign = 1.0; c->t[i].known = 1;
            c->pref[i] = ub;
            cnt[3]++;
        }
        c->hu[i].block = hU++; c->hu[i].sign = 1.0;
        c->ht[i].block = hT++; c->ht[i].sign = 1.0;
    }
    cnt[0] = unknown;
    cnt[1] = known;
    cnt[2] = N;
    cnt[7] = hU;
    cnt[8] = hT;
    if (unknown != N)
        c->dof_ok = 0;
    free(side);
}

orc_case* orc_create(int np, const double* points, int ne, const int* ele_pid, const int* bcid, int nd,
                     const double* mats, int ni, const int* itf, double dt, double gap)
{
    orc_case* c;
    long b, d, i;
    (void)np;
    if (ne <= 0 || nd <= 0)
        return 0;
    c = (orc_case*)calloc(1, sizeof(orc_case));
    c->ne = ne;
    c->nd = nd;
    c->ni = ni;
    c->dt = dt;
    build_tables(&c->q, gap);
    c->mats = (dmat*)malloc(sizeof(dmat) * (size_t)nd);
    for (d = 0; d < nd; ++d) {
        double E = mats[3 * d], v = mats[3 * d + 1], rho = mats[3 * d + 2];
        double G = E / 2.0 / (1.0 + v);
        c->mats[d] = build_dmat(v, G, rho, dt);
    }
    c->dom_begin = (long*)malloc(sizeof(long) * (size_t)nd);
    c->dom_count = (long*)malloc(sizeof(long) * (size_t)nd);
    b = 0;
    for (d = 0; d < nd; ++d) {
        long base = ne / nd, extra = ne % nd;
        c->dom_begin[d] = b;
        c->dom_count[d] = base + (d < extra ? 1 : 0);
        b += c->dom_count[d];
    }
    c->el = (elem*)calloc((size_t)ne, sizeof(elem));
    c->itf = (itf_t*)malloc(sizeof(itf_t) * (size_t)(ni > 0 ? ni : 1));
    for (i = 0; i < ni; ++i) {
        c->itf[i].da = itf[5 * i];
        c->itf[i].ea = itf[5 * i + 1];
        c->itf[i].db = itf[5 * i + 2];
        c->itf[i].eb = itf[5 * i + 3];
        c->itf[i].ns = itf[5 * i + 4] == 0 ? -1 : itf[5 * i + 4];
    }
    for (d = 0; d < nd; ++d)
        for (i = c->dom_begin[d]; i < c->dom_begin[d] + c->dom_count[d]; ++i)
            c->el[i].domain = (int)d;
    for (i = 0; i < ni; ++i) {
        c->el[c->itf[i].ea].surface = 1;
        c->el[c->itf[i].eb].surface = 1;
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (i = 0; i < ne; ++i) {
        elem* e = &c->el[i];
        double coord[3][8], dd;
        int k, j;
        e->bcid = bcid[i];
        for (k = 0; k < 3; ++k) {
            for (j = 0; j < 8; ++j)
                coord[k][j] = points[3 * (long)ele_pid[8 * i + j] + k];
            get_coef(e->cc[k], coord[k]);
        }
        for (k = 0; k < NPW2; ++k)
            from_local(e, c->q.cenpw[k][0], c->q.cenpw[k][1], e->rcenpw[k]);
        from_local(e, 0.0, 0.0, e->center);
        e->radius = -1.0;
        for (j = 0; j < 8; ++j) {
            const double* p = &points[3 * (long)ele_pid[8 * i + j]];
            dd = dist3(e->center, p);
            if (e->radius < dd)
                e->radius = dd;
        }
        e->radiusPW = e->radius / NPW;
        translate_matrix(e, c->q.lc[0][0], c->q.lc[0][1], e->trans);
        self_integrals(&c->q, &c->mats[e->domain], e);
    }
    dof_build(c);
    return c;
}

static void free_matrices(orc_case* c)
{
    int k;
    long l;
    if (!c->assembled)
        return;
    ccsr_free(&c->A);
    ccsr_free(&c->K);
    for (k = 0; k <= c->maxn; ++k) {
        ccsr_free(&c->mts[k]);
        ccsr_free(&c->mgs[k]);
    }
    free(c->mts);
    free(c->mgs);
    free(c->mgs_sym);
    for (l = 0; l < c->leaf_count; ++l)
        free(c->prem[l]);
    free(c->prem);
    free(c->leaf_size);
    free(c->leaf_begin);
    free(c->repid);
    c->assembled = 0;
}

void orc_destroy(orc_case* c)
{
    if (!c)
        return;
    free_matrices(c);
    free(c->mats);
    free(c->dom_begin);
    free(c->dom_count);
    free(c->el);
    free(c->itf);
    free(c->u);
    free(c->t);
    free(c->hu);
    free(c->ht);
    free(c->pref);
    free(c);
}

int orc_num_elements(const orc_case* c) { return c ? c->ne : -1; }

/* dt = Beta * AveEleSize / c1 of the global material line, AveEleSize the
 * mean |P0 - P1| over elements (DBEM1/DBEM.cpp:910-914,2383-2396). */
double orc_reference_dt(const double* points, const int* ele_pid, int ne, double beta, double E, double v,
                        double rho)
{
    double ave = 0.0, G, c1;
    long i;
    for (i = 0; i < ne; ++i)
        ave += dist3(&points[3 * (long)ele_pid[8 * i]], &points[3 * (long)ele_pid[8 * i + 1]]);
    ave /= ne;
    G = E / 2.0 / (1 + v);
    c1 = sqrt((2.0 - 2.0 * v) / (1.0 - 2.0 * v) * G / rho);
    return beta * ave / c1;
}

int orc_element_info(const orc_case* c, int e, double* centre, double* radius, double* frame)
{
    if (!c || e < 0 || e >= c->ne)
        return -1;
    memcpy(centre, c->el[e].center, sizeof(double) * 3);
    *radius = c->el[e].radius;
    memcpy(frame, c->el[e].trans, sizeof(double) * 9);
    return 0;
}

int orc_self_tables(const orc_case* c, int e, double* u1, double* t1, double* t2)
{
    if (!c || e < 0 || e >= c->ne)
        return -1;
    memcpy(u1, c->el[e].u1, sizeof(double) * 9);
    memcpy(t1, c->el[e].t1, sizeof(double) * 9);
    memcpy(t2, c->el[e].t2, sizeof(double) * 9);
    return 0;
}

int orc_dynamat(const orc_case* c, int d, double* out24)
{
    const dmat* m;
    if (!c || d < 0 || d >= c->nd)
        return -1;
    m = &c->mats[d];
    out24[0] = m->G; out24[1] = m->v; out24[2] = m->Rou; out24[3] = m->C1; out24[4] = m->C2;
    out24[5] = m->Dt; out24[6] = m->Dt2; out24[7] = m->Dt_2; out24[8] = m->Dt_3;
    out24[9] = m->C1_2; out24[10] = m->C2_2; out24[11] = m->OneOverC1_2; out24[12] = m->OneOverC2_2;
    out24[13] = m->OneOverC1_3; out24[14] = m->OneOverC2_3; out24[15] = m->C2_2OverC1_2;
    out24[16] = m->C2_2OverC1_3; out24[17] = m->C1Dt; out24[18] = m->C2Dt; out24[19] = m->OneOverFourPaiRou;
    out24[20] = m->OneOverThree_1OC23_1OC13; out24[21] = m->Half_1OC22_1OC12; out24[22] = m->STCoef1;
    out24[23] = m->STCoef2;
    return 0;
}

int orc_dofmap(orc_case* c, long* counts, long* u_block, double* u_sign, int* u_known, long* t_block,
               double* t_sign, int* t_known, long* hu_block, double* hu_sign, long* ht_block, double* ht_sign,
               long* preferred)
{
    long i;
    if (!c)
        return -1;
    memcpy(counts, c->counts, sizeof(long) * 9);
    for (i = 0; i < c->ne; ++i) {
        u_block[i] = c->u[i].block; u_sign[i] = c->u[i].sign; u_known[i] = c->u[i].known;
        t_block[i] = c->t[i].block; t_sign[i] = c->t[i].sign; t_known[i] = c->t[i].known;
        hu_block[i] = c->hu[i].block; hu_sign[i] = c->hu[i].sign;
        ht_block[i] = c->ht[i].block; ht_sign[i] = c->ht[i].sign;
        preferred[i] = c->pref[i];
    }
    return c->dof_ok ? 0 : -2;
}

int orc_kernel(orc_case* c, int d, int which, double n, const double* ri, const double* nor, double* out9)
{
    double R[6], RI[4], src[3] = {0, 0, 0};
    if (!c || d < 0 || d >= c->nd)
        return -1;
    get_r5(src, ri, R, RI);
    if (which == 0)
        return dyna_u(&c->mats[d], out9, n, R, RI);
    if (which == 1)
        return dyna_t1(&c->mats[d], out9, n, R, RI, nor);
    return dyna_t2(&c->mats[d], out9, n, R, RI, nor);
}

/* Regular-path table probe: the 16 points (s1, s2) of the 4x4 rule, their
 * plain weights m_RGV and the Q1 weights m_NRGV[0] = N0(s1, s2; gap) * w
 * (DBEM1/quadinfo.cpp:83-97). */
int orc_regular_table(const orc_case* c, double* pos32, double* ra16, double* nrgv0_16)
{
    int i;
    if (!c)
        return -1;
    for (i = 0; i < NG2; ++i) {
        pos32[2 * i] = c->q.rp[i][0];
        pos32[2 * i + 1] = c->q.rp[i][1];
        ra16[i] = c->q.ra[i];
        nrgv0_16[i] = c->q.nrgv0[i];
    }
    return 0;
}

int orc_needcal(orc_case* c, int field, int source, int type, double n)
{
    const elem* f;
    if (!c || field < 0 || field >= c->ne || source < 0 || source >= c->ne)
        return -1;
    f = &c->el[field];
    return need_cal(&c->mats[f->domain], f, type, n, c->el[source].center);
}

int orc_quadrature(orc_case* c, int field, int source, int which, double n, double* out9)
{
    const elem* f;
    int k;
    if (!c || field < 0 || field >= c->ne || source < 0 || source >= c->ne)
        return -1;
    f = &c->el[field];
    for (k = 0; k < 9; ++k)
        out9[k] = 0.0;
    if (which == 0)
        return judge_u(&c->q, &c->mats[f->domain], f, c->el[source].center, n, out9);
    return judge_t(&c->q, &c->mats[f->domain], f, which, c->el[source].center, n, out9);
}

/* ------------------------------------------------------------------------ */
/* Step 0 coefficients (DBEM1/dsquareelement.cpp:2756-2868,3281-3394,        */
/* 4215-4550,5157-5491,5922-6078,6476-6632,6719-6774,6891-6911)              */
/* ------------------------------------------------------------------------ */
/* Convert{Unknown,Known}Coef for BCID 123/456: Coef[3j+i] = +-sum_k K[3k+i] R[3j+k] */
static void convert_coef(const double* K, const double* R, double sgn, double* out)
{
    int i, j, k;
    for (i = 0; i < 9; ++i)
        out[i] = 0.0;
    for (i = 0; i < 3; ++i)
        for (j = 0; j < 3; ++j)
            for (k = 0; k < 3; ++k) {
                if (sgn > 0)
                    out[3 * j + i] += K[3 * k + i] * R[3 * j + k];
                else
                    out[3 * j + i] -= K[3 * k + i] * R[3 * j + k];
            }
}

/* Returns the flag (0 = no block) and the column-major coefficient block.
 * unknown_pass selects UnknownSubMatrix vs knownSubMatrix; bcid is the
 * (possibly temporarily overridden) field BCID. */
static int step0_coef(orc_case* c, int field, int source, int unknown_pass, int bcid, double* coef)
{
    const elem* f = &c->el[field];
    const dmat* m = &c->mats[f->domain];
    const qtables* q = &c->q;
    const double* src = c->el[source].center;
    /* which kernel: unknown pass -> U for 123, T for 456; known pass swaps */
    int use_u = unknown_pass ? (bcid == 123) : (bcid == 456);
    double sgn = unknown_pass ? (bcid == 123 ? -1.0 : 1.0) : (bcid == 456 ? 1.0 : -1.0);
    double TU[9], TT[9], tmp[9];
    int k, flag;
    for (k = 0; k < 9; ++k) {
        TU[k] = 0.0;
        TT[k] = 0.0;
        coef[k] = 0.0;
    }
    if (field == source) {
        /* InElementInt{Unknown,Known}Coef, FlagDynaMethod 2 branch */
        if (use_u) {
            for (k = 0; k < 9; ++k)
                TU[k] += 4.0 * f->u1[k];
            int_u_pw(q, m, f, src, 1.0, tmp);
            for (k = 0; k < 9; ++k)
                TU[k] += tmp[k];
        } else {
            for (k = 0; k < 9; ++k) {
                TT[k] += 4.0 * f->t1[k];
                TT[k] += f->t2[k];
            }
            int_t_pw(q, m, f, 1, src, 1.0, tmp);
            for (k = 0; k < 9; ++k)
                TT[k] += tmp[k];
        }
        convert_coef(use_u ? TU : TT, f->trans, sgn, coef);
        return 1;
    }
    /* RegularInt{Unknown,Known}Coef, FlagDynaMethod 2 branch */
    flag = need_cal(m, f, 1, 0.0, src);
    {
        int fsum = flag;
        if (flag == 1 || flag == 2) {
            if (use_u) {
                if (flag == 1)
                    int_u_pw(q, m, f, src, 0.0, tmp);
                else
                    int_u_reg(q, m, f, src, 0.0, tmp);
                for (k = 0; k < 9; ++k)
                    TU[k] += 4.0 * tmp[k];
            } else {
                if (flag == 1)
                    int_t_pw(q, m, f, 1, src, 0.0, tmp);
                else
                    int_t_reg(q, m, f, src, 0.0, tmp);
                for (k = 0; k < 9; ++k)
                    TT[k] += 4.0 * tmp[k];
                if (flag == 1)
                    int_t_pw(q, m, f, 2, src, 1.0, tmp);
                else
                    int_t_reg(q, m, f, src, 1.0, tmp);
                for (k = 0; k < 9; ++k)
                    TT[k] += tmp[k];
            }
        }
        flag = need_cal(m, f, 1, 1.0, src);
        fsum += flag;
        if (flag == 1 || flag == 2) {
            if (use_u) {
                if (flag == 1)
                    int_u_pw(q, m, f, src, 1.0, tmp);
                else
                    int_u_reg(q, m, f, src, 1.0, tmp);
                for (k = 0; k < 9; ++k)
                    TU[k] += tmp[k];
            } else {
                if (flag == 1)
                    int_t_pw(q, m, f, 1, src, 1.0, tmp);
                else
                    int_t_reg(q, m, f, src, 1.0, tmp);
                for (k = 0; k < 9; ++k)
                    TT[k] += tmp[k];
            }
        }
        if (fsum)
            convert_coef(use_u ? TU : TT, f->trans, sgn, coef);
        return fsum;
    }
}

/* TransformColumnBlock9 (DBEM1/multidomain.cpp:1668-1680) */
static void transform_col_block(const double* b, const double* t, double* out)
{
    int r, qq, cc;
    for (r = 0; r < 3; ++r)
        for (qq = 0; qq < 3; ++qq) {
            double v = 0.0;
            for (cc = 0; cc < 3; ++cc)
                v += b[cc * 3 + r] * t[cc * 3 + qq];
            out[qq * 3 + r] = v;
        }
}

/* GetElementLocalTransformFromInterfaceReference (DBEM1/multidomain.cpp:1594-1645) */
static void itf_transform(const orc_case* c, int domain, int ele, double* T)
{
    int i, s, r, ax;
    for (i = 0; i < 9; ++i)
        T[i] = 0.0;
    T[0] = T[4] = T[8] = 1.0;
    for (i = 0; i < c->ni; ++i) {
        const itf_t* f = &c->itf[i];
        if (f->da == domain && f->ea == ele)
            return;
        if (f->db == domain && f->eb == ele) {
            const double* side = c->el[f->eb].trans;
            const double* ref = c->el[f->ea].trans;
            for (s = 0; s < 3; ++s)
                for (r = 0; r < 3; ++r) {
                    double dot = 0.0;
                    for (ax = 0; ax < 3; ++ax)
                        dot += side[s * 3 + ax] * ref[r * 3 + ax];
                    T[s * 3 + r] = dot;
                }
            return;
        }
    }
}

static int itf_reference_side(const orc_case* c, int domain, int ele)
{
    int i;
    for (i = 0; i < c->ni; ++i) {
        if (c->itf[i].da == domain && c->itf[i].ea == ele)
            return 1;
        if (c->itf[i].db == domain && c->itf[i].eb == ele)
            return 0;
    }
    return 1;
}

static void assemble_step0(orc_case* c, builder* A, builder* K)
{
    int d;
    long s, f;
    double blk[9], tb[9], T[9];
    for (d = 0; d < c->nd; ++d)
        for (s = c->dom_begin[d]; s < c->dom_begin[d] + c->dom_count[d]; ++s) {
            long row = s;
            for (f = c->dom_begin[d]; f < c->dom_begin[d] + c->dom_count[d]; ++f) {
                const elem* fe = &c->el[f];
                if (fe->surface == 1) {
                    itf_transform(c, d, (int)f, T);
                    if (step0_coef(c, (int)f, (int)s, 1, 456, blk)) {
                        transform_col_block(blk, T, tb);
                        bld_add(A, row, c->u[f].block, c->u[f].sign, tb);
                    }
                    if (step0_coef(c, (int)f, (int)s, 1, 123, blk)) {
                        transform_col_block(blk, T, tb);
                        bld_add(A, row, c->t[f].block, c->t[f].sign, tb);
                    }
                } else {
                    vref r;
                    if (step0_coef(c, (int)f, (int)s, 1, fe->bcid, blk)) {
                        r = fe->bcid == 123 ? c->t[f] : c->u[f];
                        if (!r.known)
                            bld_add(A, row, r.block, r.sign, blk);
                    }
                    if (step0_coef(c, (int)f, (int)s, 0, fe->bcid, blk)) {
                        r = fe->bcid == 123 ? c->u[f] : c->t[f];
                        if (r.known)
                            bld_add(K, row, r.block, r.sign, blk);
                    }
                }
            }
        }
}

/* ComputeT2/T1/U BlockCurrentMat + AssembleMultiDomainHistoryStep
 * (DBEM1/multidomain.cpp:1849-1905,2017-2078,2329-2378) */
static void assemble_history_rows(orc_case* c, int k, long row0, long row1, builder* mts, builder* mgs)
{
    int d;
    long s, f;
    double blk[9];
    double n = (double)k;
    for (d = 0; d < c->nd; ++d) {
        const dmat* m = &c->mats[d];
        for (s = c->dom_begin[d]; s < c->dom_begin[d] + c->dom_count[d]; ++s) {
            if (s < row0 || s >= row1)
                continue;
            const double* src = c->el[s].center;
            for (f = c->dom_begin[d]; f < c->dom_begin[d] + c->dom_count[d]; ++f) {
                const elem* fe = &c->el[f];
                int flag;
                if (k <= 1 && f == s) {
                    memcpy(blk, fe->t2, sizeof(blk));
                    bld_add(mts, s, c->hu[f].block, c->hu[f].sign, blk);
                } else {
                    flag = judge_t(&c->q, m, fe, 2, src, n, blk);
                    if (flag)
                        bld_add(mts, s, c->hu[f].block, c->hu[f].sign, blk);
                }
                flag = judge_t(&c->q, m, fe, 1, src, n, blk);
                if (flag)
                    bld_add(mts, s, c->hu[f].block, c->hu[f].sign, blk);
                flag = judge_u(&c->q, m, fe, src, n, blk);
                if (flag)
                    bld_add(mgs, s, c->ht[f].block, c->ht[f].sign, blk);
            }
        }
    }
}

static void assemble_history(orc_case* c, int k, builder* mts, builder* mgs)
{
    assemble_history_rows(c, k, 0, c->ne, mts, mgs);
}

/* ------------------------------------------------------------------------ */
/* Octree leaves (DBEM1/tree.cpp:12-165,242-323; DBEM1/geo.cpp:8-40,82-113)  */
/* ------------------------------------------------------------------------ */
typedef struct {
    double center[3], length;
} cube;

static int isin(const double* p, const cube* cb)
{
    double ds[3], l;
    int i;
    for (i = 0; i < 3; ++i)
        ds[i] = cb->center[i] - p[i];
    l = 0.5 * cb->length;
    for (i = 0; i < 3; ++i)
        if (ds[i] * ds[i] > l * l * 1.0001)
            return 0;
    return 1;
}

static void sub_cube(const cube* fa, cube* ch, int pos)
{
    double half;
    int i, bit = 1;
    ch->length = 0.5 * fa->length;
    half = 0.5 * ch->length;
    for (i = 0; i < 3; ++i) {
        ch->center[i] = fa->center[i] + ((bit & pos) == 0 ? -1 : 1) * half;
        bit <<= 1;
    }
}

typedef struct {
    long* leaf_size;
    long* leaf_begin;
    long* repid;
    long nleaf;
} leafset;

/* pts: list of point ids in the node's linked-list order */
static void tree_rec(const orc_case* c, const cube* cb, long* pts, long cnt, long begin, long maxleaf, leafset* out)
{
    cube ch[8];
    long* lists[8];
    long counts[8];
    int i;
    long p, b;
    if (cnt <= maxleaf) {
        out->leaf_size[out->nleaf] = cnt;
        out->leaf_begin[out->nleaf] = begin;
        for (p = 0; p < cnt; ++p)
            out->repid[begin + p] = pts[p];
        out->nleaf++;
        return;
    }
    for (i = 0; i < 8; ++i) {
        sub_cube(cb, &ch[i], i);
        lists[i] = (long*)malloc(sizeof(long) * (size_t)cnt);
        counts[i] = 0;
    }
    for (p = 0; p < cnt; ++p) {
        const double* x = c->el[pts[p]].center;
        int child = -1;
        for (i = 0; i < 8; ++i)
            if (isin(x, &ch[i])) {
                child = i;
                break;
            }
        if (child < 0) {
            double best = dist3(x, ch[0].center);
            child = 0;
            for (i = 1; i < 8; ++i) {
                double dd = dist3(x, ch[i].center);
                if (best > dd) {
                    best = dd;
                    child = i;
                }
            }
        }
        lists[child][counts[child]++] = pts[p];
    }
    b = begin;
    for (i = 0; i < 8; ++i) {
        long j, n = counts[i];
        /* linked lists are built by prepending: reverse insertion order */
        for (j = 0; j < n / 2; ++j) {
            long t = lists[i][j];
            lists[i][j] = lists[i][n - 1 - j];
            lists[i][n - 1 - j] = t;
        }
        if (n > 0) {
            tree_rec(c, &ch[i], lists[i], n, b, maxleaf, out);
            b += n;
        }
        free(lists[i]);
    }
}

static void build_leaves(const orc_case* c, long maxleaf, leafset* out)
{
    long N = c->ne, i;
    int j;
    double rmin[3], rmax[3], ml;
    cube root;
    long* pts = (long*)malloc(sizeof(long) * (size_t)N);
    for (j = 0; j < 3; ++j)
        rmin[j] = rmax[j] = c->el[0].center[j];
    for (i = 1; i < N; ++i)
        for (j = 0; j < 3; ++j) {
            if (rmin[j] > c->el[i].center[j])
                rmin[j] = c->el[i].center[j];
            if (rmax[j] < c->el[i].center[j])
                rmax[j] = c->el[i].center[j];
        }
    ml = rmax[0] - rmin[0];
    for (j = 0; j < 3; ++j) {
        if (ml < (rmax[j] - rmin[j]))
            ml = rmax[j] - rmin[j];
        root.center[j] = rmin[j] + 0.5 * (rmax[j] - rmin[j]);
    }
    root.length = 1.02 * ml;
    for (i = 0; i < N; ++i)
        pts[i] = i;
    out->leaf_size = (long*)malloc(sizeof(long) * (size_t)N);
    out->leaf_begin = (long*)malloc(sizeof(long) * (size_t)N);
    out->repid = (long*)malloc(sizeof(long) * (size_t)N);
    out->nleaf = 0;
    tree_rec(c, &root, pts, N, 0, maxleaf, out);
    free(pts);
}

/* inv_mat: full-pivot Gauss-Jordan (DBEM1/lin_comp.cpp:83-190); row-major n x n */
static int inv_mat(double* A, long n)
{
    long k, i, j;
    long* is = (long*)malloc(sizeof(long) * (size_t)(n + 1));
    long* js = (long*)malloc(sizeof(long) * (size_t)(n + 1));
    double t, d, temp;
#define AA(r, cc) A[((r)-1) * n + ((cc)-1)]
    for (k = 1; k <= n; k++) {
        d = 0.0;
        for (i = k; i <= n; i++)
            for (j = k; j <= n; j++) {
                t = sqrt(AA(i, j) * AA(i, j));
                if (t > d) {
                    d = t;
                    is[k] = i;
                    js[k] = j;
                }
            }
        if (d <= 0.0000000001) {
            free(is);
            free(js);
            return -1;
        }
        for (j = 1; j <= n; j++) {
            i = is[k];
            temp = AA(k, j);
            AA(k, j) = AA(i, j);
            AA(i, j) = temp;
        }
        for (i = 1; i <= n; i++) {
            j = js[k];
            temp = AA(i, k);
            AA(i, k) = AA(i, j);
            AA(i, j) = temp;
        }
        AA(k, k) = 1.0 / AA(k, k);
        for (j = 1; j <= n; j++)
            if (j != k)
                AA(k, j) *= AA(k, k);
        for (i = 1; i <= n; i++)
            if (i != k)
                for (j = 1; j <= n; j++)
                    if (j != k)
                        AA(i, j) -= AA(i, k) * AA(k, j);
        for (i = 1; i <= n; i++)
            if (i != k)
                AA(i, k) *= -1.0 * AA(k, k);
    }
    for (k = n; k >= 1; k--) {
        for (j = 1; j <= n; j++) {
            i = js[k];
            temp = AA(k, j);
            AA(k, j) = AA(i, j);
            AA(i, j) = temp;
        }
        for (i = 1; i <= n; i++) {
            j = is[k];
            temp = AA(i, k);
            AA(i, k) = AA(i, j);
            AA(i, j) = temp;
        }
    }
#undef AA
    free(is);
    free(js);
    return 1;
}

static long find_block(const ccsr* M, long row, long col)
{
    int p;
    if (row < 0 || row >= M->n || col < 0 || col >= M->n)
        return -1;
    for (p = M->row_ptr[row]; p < M->row_ptr[row + 1]; ++p)
        if (M->col[p] == col)
            return p;
    return -1;
}

/* BuildMappedLeafPreConditioner (DBEM1/multidomain.cpp:3493-3565) */
static void build_precond(orc_case* c, long maxleaf)
{
    leafset ls;
    long l, r, cc;
    build_leaves(c, maxleaf, &ls);
    c->leaf_count = ls.nleaf;
    c->leaf_size = ls.leaf_size;
    c->leaf_begin = ls.leaf_begin;
    c->repid = ls.repid;
    c->prem = (double**)malloc(sizeof(double*) * (size_t)ls.nleaf);
    for (l = 0; l < ls.nleaf; ++l) {
        long cnt = ls.leaf_size[l], n3 = 3 * cnt, b = ls.leaf_begin[l];
        double* P = (double*)calloc((size_t)(n3 * n3), sizeof(double));
        for (r = 0; r < cnt; ++r) {
            long row = c->repid[b + r];
            for (cc = 0; cc < cnt; ++cc) {
                long col = c->pref[c->repid[b + cc]];
                long p = find_block(&c->A, row, col);
                int i, j;
                if (p >= 0)
                    for (i = 0; i < 3; ++i)
                        for (j = 0; j < 3; ++j)
                            P[(3 * r + i) * n3 + 3 * cc + j] = c->A.val[9 * p + 3 * i + j];
            }
        }
        inv_mat(P, n3);
        c->prem[l] = P;
    }
}

/* GMRES_M_X (DBEM1/pre_gmres.cpp:837-895) */
static void precond_apply(const orc_case* c, const double* X, double* B)
{
    long l, j, k, N3 = 3 * (long)c->ne, i;
    int m, n;
    for (i = 0; i < N3; ++i)
        B[i] = 0.0;
    for (l = 0; l < c->leaf_count; ++l) {
        long cnt = c->leaf_size[l], b = c->leaf_begin[l], n3 = 3 * cnt;
        const double* P = c->prem[l];
        for (j = 0; j < cnt; ++j) {
            long a1 = 3 * c->pref[c->repid[b + j]];
            for (k = 0; k < cnt; ++k) {
                long a2 = 3 * c->repid[b + k];
                for (m = 0; m < 3; ++m)
                    for (n = 0; n < 3; ++n)
                        B[a1 + m] += P[(3 * j + m) * n3 + 3 * k + n] * X[a2 + n];
            }
        }
    }
}

static double norm2(const double* a, long n)
{
    double r = 0.0;
    long i;
    for (i = 0; i < n; ++i)
        r += a[i] * a[i];
    return sqrt(r);
}

/* gmres(..., CCSRMat&, ...) with maxit = 1 and x0 = 0 (DBEM1/pre_gmres.cpp:1369-1605).
 * Returns flag (0 converged, 1 not converged, 3 stagnated); *iters inner count. */
static int gmres_solve(const orc_case* c, const ccsr* A, const double* b, long restart, double tol, double* x,
                       int* iters, int use_pre)
{
    long n = 3 * A->n, i, j, k, l;
    double n2b, tolb, normr, phibar, rt, cs, sn, temp;
    double *r, *vh, *u, *u2, *q, *V, *h, *QT, *R, *f, *W;
    int flag = 1, stag = 0;
    long rs = restart;
    *iters = 0;
    n2b = norm2(b, n);
    if (n2b == 0.0) {
        for (i = 0; i < n; ++i)
            x[i] = 0.0;
        return 0;
    }
    r = (double*)calloc((size_t)n, sizeof(double));
    vh = (double*)calloc((size_t)n, sizeof(double));
    u = (double*)calloc((size_t)n, sizeof(double));
    u2 = (double*)calloc((size_t)n, sizeof(double));
    q = (double*)calloc((size_t)rs + 2, sizeof(double));
    V = (double*)calloc((size_t)n * (size_t)(rs + 2), sizeof(double));
    h = (double*)calloc((size_t)rs + 2, sizeof(double));
    QT = (double*)calloc((size_t)(rs + 2) * (size_t)(rs + 2), sizeof(double));
    R = (double*)calloc((size_t)(rs + 1) * (size_t)(rs + 1), sizeof(double));
    f = (double*)calloc((size_t)rs + 2, sizeof(double));
    W = (double*)calloc((size_t)n * (size_t)(rs + 1), sizeof(double));
#define VV(row, col) V[(size_t)((col)-1) * (size_t)n + (size_t)((row)-1)]
#define WW(row, col) W[(size_t)((col)-1) * (size_t)n + (size_t)((row)-1)]
#define QQ(a, bb) QT[(size_t)((a)-1) * (size_t)(rs + 1) + (size_t)((bb)-1)]
#define RR(a, bb) R[(size_t)((a)-1) * (size_t)rs + (size_t)((bb)-1)]
    for (i = 0; i < n; ++i)
        x[i] = 0.0;
    tolb = tol * n2b;
    spmv(A, x, r);
    for (i = 0; i < n; ++i)
        r[i] = b[i] - r[i];
    normr = norm2(r, n);
    if (normr <= tolb) {
        flag = 0;
        goto done;
    }
    /* one outer iteration (maxit = 1) */
    if (use_pre)
        precond_apply(c, r, vh);
    else
        memcpy(vh, r, sizeof(double) * (size_t)n);
    h[1] = norm2(vh, n);
    for (j = 1; j <= n; j++)
        VV(j, 1) = vh[j - 1] / h[1];
    QQ(1, 1) = 1.0;
    phibar = h[1];
    for (j = 1; j <= rs; j++) {
        for (k = 1; k <= n; k++)
            u[k - 1] = VV(k, j);
        spmv(A, u, u2);
        if (use_pre)
            precond_apply(c, u2, u);
        else
            memcpy(u, u2, sizeof(double) * (size_t)n);
        for (k = 1; k <= j; k++) {
            h[k] = 0.0;
            for (l = 1; l <= n; l++)
                h[k] += VV(l, k) * u[l - 1];
            for (l = 1; l <= n; l++)
                u[l - 1] -= h[k] * VV(l, k);
        }
        h[j + 1] = norm2(u, n);
        if (h[j + 1] <= 1.0e-12) {
            for (k = 1; k <= n; k++)
                VV(k, j + 1) = 0.0;
        } else {
            for (k = 1; k <= n; k++)
                VV(k, j + 1) = u[k - 1] / h[j + 1];
        }
        for (k = 1; k <= j; k++) {
            RR(k, j) = 0.0;
            for (l = 1; l <= j; l++)
                RR(k, j) += QQ(k, l) * h[l];
        }
        rt = RR(j, j);
        if (h[j + 1] == 0.0) {
            cs = 1.0;
            sn = 0.0;
        } else if (h[j + 1] * h[j + 1] > rt * rt) {
            temp = rt / h[j + 1];
            sn = 1.0 / sqrt(1.0 + fabs(temp) * fabs(temp));
            cs = -temp * sn;
        } else {
            temp = h[j + 1] / rt;
            cs = 1.0 / sqrt(1.0 + fabs(temp) * fabs(temp));
            sn = -temp * cs;
        }
        RR(j, j) = cs * rt - sn * h[j + 1];
        for (k = 1; k <= j; k++)
            q[k] = QQ(j, k);
        for (k = 1; k <= j; k++)
            QQ(j, k) = cs * q[k];
        for (k = 1; k <= j; k++)
            QQ(j + 1, k) = sn * q[k];
        QQ(j, j + 1) = -sn;
        QQ(j + 1, j + 1) = cs;
        f[j] = cs * phibar;
        phibar = sn * phibar;
        if (j < rs) {
            if (f[j] == 0.0)
                stag = 1;
            if (j == 1) {
                for (k = 1; k <= n; k++)
                    WW(k, j) = VV(k, j) / RR(j, j);
            } else {
                for (k = 1; k <= n; k++) {
                    WW(k, j) = VV(k, j) / RR(j, j);
                    for (l = 1; l <= j - 1; l++)
                        WW(k, j) -= WW(k, l) * RR(l, j) / RR(j, j);
                }
            }
            for (k = 1; k <= n; k++)
                x[k - 1] += f[j] * WW(k, j);
        } else {
            /* j == restart: x = V(:,1:j) * (R^-1 f) + x0, x0 = 0.  The reference's
             * mat_vec_product reads one entry past q here; it is taken as 0. */
            double* Rc = (double*)malloc(sizeof(double) * (size_t)(j * j));
            long a, bb;
            for (a = 1; a <= j; ++a)
                for (bb = 1; bb <= j; ++bb)
                    Rc[(a - 1) * j + (bb - 1)] = RR(a, bb);
            inv_mat(Rc, j);
            for (a = 1; a <= j; ++a) {
                q[a] = 0.0;
                for (bb = 1; bb <= j; ++bb)
                    q[a] += Rc[(a - 1) * j + (bb - 1)] * f[bb];
            }
            for (k = 1; k <= n; ++k) {
                x[k - 1] = 0.0;
                for (a = 1; a <= j; ++a)
                    x[k - 1] += VV(k, a) * q[a];
            }
            free(Rc);
        }
        spmv(A, x, vh);
        for (k = 0; k < n; k++)
            vh[k] = b[k] - vh[k];
        normr = norm2(vh, n);
        *iters = (int)j;
        if (normr <= tolb) {
            flag = 0;
            goto done;
        }
        if (stag == 1) {
            flag = 3;
            goto done;
        }
    }
done:
#undef VV
#undef WW
#undef QQ
#undef RR
    free(r);
    free(vh);
    free(u);
    free(u2);
    free(q);
    free(V);
    free(h);
    free(QT);
    free(R);
    free(f);
    free(W);
    return flag;
}

int orc_gmres_ccsr(long nrows, const int* row_ptr, const int* col, const double* val, const double* b,
                   long restart, double tol, double* x, int* iters)
{
    ccsr A;
    A.n = nrows;
    A.nnz = row_ptr[nrows];
    A.row_ptr = (int*)row_ptr;
    A.col = (int*)col;
    A.val = (double*)val;
    return gmres_solve(0, &A, b, restart, tol, x, iters, 0);
}

/* structure check (DBEM1/multidomain.cpp:3315-3443) */
static int structure_ok(const ccsr* A)
{
    long row, zr = 0, zc = 0, bad = 0;
    int p;
    long* cc = (long*)calloc((size_t)A->n, sizeof(long));
    for (row = 0; row < A->n; ++row) {
        if (A->row_ptr[row] == A->row_ptr[row + 1])
            zr++;
        for (p = A->row_ptr[row]; p < A->row_ptr[row + 1]; ++p) {
            if (A->col[p] >= 0 && A->col[p] < A->n)
                cc[A->col[p]]++;
            else {
                bad++;
                continue;
            }
            if (!finite9(&A->val[9 * (long)p]))
                bad++;
        }
    }
    for (row = 0; row < A->n; ++row)
        if (cc[row] == 0)
            zc++;
    free(cc);
    return zr == 0 && zc == 0 && bad == 0;
}

int orc_assemble(orc_case* c, int maxn)
{
    builder bA, bK;
    int k;
    if (!c || !c->dof_ok)
        return -1;
    free_matrices(c);
    bld_init(&bA, c->counts[2], c->counts[0]);
    bld_init(&bK, c->counts[2], c->counts[1]);
    assemble_step0(c, &bA, &bK);
    bld_build(&bA, &c->A);
    bld_build(&bK, &c->K);
    bld_free(&bA);
    bld_free(&bK);
    c->maxn = maxn;
    c->mts = (ccsr*)calloc((size_t)maxn + 1, sizeof(ccsr));
    c->mgs = (ccsr*)calloc((size_t)maxn + 1, sizeof(ccsr));
    c->mgs_sym = (int*)calloc((size_t)maxn + 1, sizeof(int));
    /* history steps are independent builders: parallel over steps, each
     * step's sums stay sequential, so results do not depend on threads */
#pragma omp parallel for schedule(dynamic, 1)
    for (k = 1; k <= maxn; ++k) {
        builder bt, bg;
        bld_init(&bt, c->counts[2], c->counts[7]);
        bld_init(&bg, c->counts[2], c->counts[8]);
        assemble_history(c, k, &bt, &bg);
        bld_build(&bt, &c->mts[k]);
        bld_build(&bg, &c->mgs[k]);
        c->mgs_sym[k] = ccsr_is_sym(&c->mgs[k], 1.0e-10);
        bld_free(&bt);
        bld_free(&bg);
    }
    c->assembled = 1;
    build_precond(c, 10);
    if (!structure_ok(&c->A))
        return -3;
    return 0;
}

int orc_ccsr(orc_case* c, int which, int step, long* nrows, long* nnz, int* row_ptr, int* col, double* val)
{
    const ccsr* M;
    if (!c || !c->assembled)
        return -1;
    if (which == 0)
        M = &c->A;
    else if (which == 1)
        M = &c->K;
    else if ((which == 2 || which == 3) && step >= 1 && step <= c->maxn)
        M = which == 2 ? &c->mts[step] : &c->mgs[step];
    else
        return -2;
    *nrows = M->n;
    *nnz = M->nnz;
    if (row_ptr)
        memcpy(row_ptr, M->row_ptr, sizeof(int) * (size_t)(M->n + 1));
    if (col)
        memcpy(col, M->col, sizeof(int) * (size_t)M->nnz);
    if (val)
        memcpy(val, M->val, sizeof(double) * 9 * (size_t)M->nnz);
    return 0;
}

int orc_history_rows(orc_case* c, int k, long row0, long row1, int which, long* nnz, int* row_ptr, int* col,
                     double* val)
{
    builder b;
    ccsr M;
    long r;
    if (!c || !c->dof_ok || k < 1 || row0 < 0 || row1 > c->ne || row0 >= row1 || (which != 2 && which != 3))
        return -1;
    {
        builder bt, bg;
        bld_init(&bt, c->counts[2], c->counts[7]);
        bld_init(&bg, c->counts[2], c->counts[8]);
        assemble_history_rows(c, k, row0, row1, &bt, &bg);
        if (which == 2) {
            b = bt;
            bld_free(&bg);
        } else {
            b = bg;
            bld_free(&bt);
        }
    }
    bld_build(&b, &M);
    bld_free(&b);
    *nnz = M.row_ptr[row1] - M.row_ptr[row0];
    if (row_ptr)
        for (r = row0; r <= row1; ++r)
            row_ptr[r - row0] = M.row_ptr[r] - M.row_ptr[row0];
    if (col)
        memcpy(col, M.col + M.row_ptr[row0], sizeof(int) * (size_t)*nnz);
    if (val)
        memcpy(val, M.val + 9 * (size_t)M.row_ptr[row0], sizeof(double) * 9 * (size_t)*nnz);
    ccsr_free(&M);
    return 0;
}

int orc_precond_info(orc_case* c, long* leaf_count, long* leaf_size, long* repid)
{
    if (!c || !c->assembled)
        return -1;
    *leaf_count = c->leaf_count;
    if (leaf_size)
        memcpy(leaf_size, c->leaf_size, sizeof(long) * (size_t)c->leaf_count);
    if (repid)
        memcpy(repid, c->repid, sizeof(long) * (size_t)c->ne);
    return 0;
}

static void rebuild_precond(orc_case* c, long maxleaf)
{
    long l;
    for (l = 0; l < c->leaf_count; ++l)
        free(c->prem[l]);
    free(c->prem);
    free(c->leaf_size);
    free(c->leaf_begin);
    free(c->repid);
    build_precond(c, maxleaf);
}

/* ------------------------------------------------------------------------ */
/* Time loop (DBEM1/multidomain.cpp:3047-3225,3815-3893)                     */
/* ------------------------------------------------------------------------ */
static void loc_to_abs(const double* T, double* v)
{
    double a0 = T[0] * v[0] + T[3] * v[1] + T[6] * v[2];
    double a1 = T[1] * v[0] + T[4] * v[1] + T[7] * v[2];
    double a2 = T[2] * v[0] + T[5] * v[1] + T[8] * v[2];
    v[0] = a0;
    v[1] = a1;
    v[2] = a2;
}

static void physical_state(const orc_case* c, const double* lu, const double* lt, double* gU, double* gT)
{
    int d;
    long i, k;
    for (i = 0; i < 3 * c->counts[7]; ++i)
        gU[i] = 0.0;
    for (i = 0; i < 3 * c->counts[8]; ++i)
        gT[i] = 0.0;
    for (d = 0; d < c->nd; ++d)
        for (i = c->dom_begin[d]; i < c->dom_begin[d] + c->dom_count[d]; ++i) {
            if (c->el[i].surface == 1 && !itf_reference_side(c, d, (int)i))
                continue;
            for (k = 0; k < 3; ++k) {
                gU[3 * c->hu[i].block + k] = c->hu[i].sign * lu[3 * i + k];
                gT[3 * c->ht[i].block + k] = c->ht[i].sign * lt[3 * i + k];
            }
        }
}

int orc_solve(orc_case* c, int nstep, int maxn, const double* known_lu, long restart, double tol, long maxleaf,
              double* lu, double* lt, int* iters, int* flags)
{
    long N = c ? c->ne : 0, nU, nK, nHU, nHT, i, k;
    int step, conv, solved = 0, d;
    double *gU, *gT, *known, *rhs, *temp, *cvT, *cvU, *x;
    if (!c || !c->dof_ok)
        return -1;
    if (!c->assembled || c->maxn < maxn) {
        int rc = orc_assemble(c, maxn);
        if (rc < 0)
            return rc;
    }
    if (maxleaf != 10)
        rebuild_precond(c, maxleaf);
    if (!structure_ok(&c->A))
        return -3;
    nU = c->counts[0];
    nK = c->counts[1];
    nHU = c->counts[7];
    nHT = c->counts[8];
    gU = (double*)calloc((size_t)(nstep + 1) * 3 * (size_t)nHU, sizeof(double));
    gT = (double*)calloc((size_t)(nstep + 1) * 3 * (size_t)nHT, sizeof(double));
    known = (double*)calloc(3 * (size_t)nK + 1, sizeof(double));
    rhs = (double*)calloc(3 * (size_t)N, sizeof(double));
    temp = (double*)calloc(3 * (size_t)N, sizeof(double));
    cvT = (double*)calloc(3 * (size_t)nHT, sizeof(double));
    cvU = (double*)calloc(3 * (size_t)nHU, sizeof(double));
    x = (double*)calloc(3 * (size_t)nU, sizeof(double));
    /* step 0: lt = 0, Convert (swap lu/lt for 456), AllLocToAbs, state */
    for (i = 0; i < N; ++i)
        for (k = 0; k < 3; ++k) {
            lu[3 * i + k] = known_lu[3 * i + k];
            lt[3 * i + k] = 0.0;
        }
    for (i = 0; i < N; ++i)
        if (c->el[i].bcid == 456)
            for (k = 0; k < 3; ++k) {
                double a = lu[3 * i + k];
                lu[3 * i + k] = lt[3 * i + k];
                lt[3 * i + k] = a;
            }
    for (i = 0; i < N; ++i) {
        loc_to_abs(c->el[i].trans, &lu[3 * i]);
        loc_to_abs(c->el[i].trans, &lt[3 * i]);
    }
    physical_state(c, lu, lt, gU, gT);
    for (step = 1; step <= nstep; ++step) {
        const double* bl = &known_lu[(size_t)step * 3 * (size_t)N];
        double* slu = &lu[(size_t)step * 3 * (size_t)N];
        double* slt = &lt[(size_t)step * 3 * (size_t)N];
        int it = 0, flag;
        for (i = 0; i < 3 * nK; ++i)
            known[i] = 0.0;
        for (d = 0; d < c->nd; ++d)
            for (i = c->dom_begin[d]; i < c->dom_begin[d] + c->dom_count[d]; ++i) {
                vref r;
                if (c->el[i].surface == 1)
                    continue;
                r = c->el[i].bcid == 123 ? c->u[i] : c->t[i];
                if (!r.known)
                    continue;
                for (k = 0; k < 3; ++k)
                    known[3 * r.block + k] = r.sign * bl[3 * i + k];
            }
        spmv(&c->K, known, rhs);
        for (conv = 1; step > 1 && conv <= step; ++conv) {
            long ids[3];
            double co[3] = {1.0, 2.0, 1.0};
            int kk;
            if (conv > maxn)
                break;
            ids[0] = step - conv + 1;
            ids[1] = step - conv;
            ids[2] = step - conv - 1;
            for (i = 0; i < 3 * nHT; ++i)
                cvT[i] = 0.0;
            for (kk = 0; kk < 3; ++kk)
                if (ids[kk] > 0 && ids[kk] < step)
                    for (i = 0; i < 3 * nHT; ++i)
                        cvT[i] += co[kk] * gT[(size_t)ids[kk] * 3 * (size_t)nHT + (size_t)i];
            if (c->mgs_sym[conv])
                spmv_sym(&c->mgs[conv], cvT, temp);
            else
                spmv(&c->mgs[conv], cvT, temp);
            for (i = 0; i < 3 * N; ++i)
                rhs[i] += 1.0 * temp[i];
            for (i = 0; i < 3 * nHU; ++i)
                cvU[i] = 0.0;
            for (kk = 0; kk < 3; ++kk)
                if (ids[kk] > 0 && ids[kk] < step)
                    for (i = 0; i < 3 * nHU; ++i)
                        cvU[i] += co[kk] * gU[(size_t)ids[kk] * 3 * (size_t)nHU + (size_t)i];
            spmv(&c->mts[conv], cvU, temp);
            for (i = 0; i < 3 * N; ++i)
                rhs[i] += -1.0 * temp[i];
        }
        flag = gmres_solve(c, &c->A, rhs, restart, tol, x, &it, 1);
        iters[step - 1] = it;
        flags[step - 1] = flag;
        solved++;
        if (flag != 0)
            break;
        /* ScatterUnknownToBoundary (DBEM1/multidomain.cpp:3086-3144) */
        for (d = 0; d < c->nd; ++d)
            for (i = c->dom_begin[d]; i < c->dom_begin[d] + c->dom_count[d]; ++i) {
                double T[9], inU[3], inT[3], ref[3], loc[3];
                vref ur = c->u[i], tr = c->t[i];
                int r2, q2;
                itf_transform(c, d, (int)i, T);
                for (k = 0; k < 3; ++k) {
                    inU[k] = bl[3 * i + k];
                    inT[k] = bl[3 * i + k];
                }
                if (!ur.known) {
                    for (k = 0; k < 3; ++k)
                        ref[k] = x[3 * ur.block + k];
                    for (r2 = 0; r2 < 3; ++r2)
                        loc[r2] = T[r2 * 3 + 0] * ref[0] + T[r2 * 3 + 1] * ref[1] + T[r2 * 3 + 2] * ref[2];
                    for (k = 0; k < 3; ++k)
                        slu[3 * i + k] = ur.sign * loc[k];
                }
                if (!tr.known) {
                    for (k = 0; k < 3; ++k)
                        ref[k] = x[3 * tr.block + k];
                    for (q2 = 0; q2 < 3; ++q2)
                        loc[q2] = T[q2 * 3 + 0] * ref[0] + T[q2 * 3 + 1] * ref[1] + T[q2 * 3 + 2] * ref[2];
                    for (k = 0; k < 3; ++k)
                        slt[3 * i + k] = tr.sign * loc[k];
                }
                if (ur.known)
                    for (k = 0; k < 3; ++k)
                        slu[3 * i + k] = ur.sign * inU[k];
                if (tr.known)
                    for (k = 0; k < 3; ++k)
                        slt[3 * i + k] = tr.sign * inT[k];
            }
        for (i = 0; i < N; ++i) {
            loc_to_abs(c->el[i].trans, &slu[3 * i]);
            loc_to_abs(c->el[i].trans, &slt[3 * i]);
        }
        physical_state(c, slu, slt, &gU[(size_t)step * 3 * (size_t)nHU], &gT[(size_t)step * 3 * (size_t)nHT]);
    }
    free(gU);
    free(gT);
    free(known);
    free(rhs);
    free(temp);
    free(cvT);
    free(cvU);
    free(x);
    return solved;
}
