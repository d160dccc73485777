// Shared plain-old-data types and geometry routines used by both the host
// library and the HIP kernels of the multi-domain TDBEM path.
//
// Every routine keeps the reference's floating-point operation order; all
// sources are compiled with -ffp-contract=off so products and sums are
// rounded separately, as the reference's MSVC build does.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace tdbem {

// Quadrature sizes (DBEM1/precompiler.h:13-27).
constexpr int kNG2 = 16;     // regular 4x4 Gauss points
constexpr int kNPW = 6;      // patches per side
constexpr int kNPW2 = 36;    // patches per element
constexpr int kNPWG2 = 9;    // 3x3 points per patch
constexpr int kNSG = 4;      // singular line points
constexpr int kNSG2 = 16;    // singular area points (static part)
constexpr int kNSPW2 = 400;  // singular polar patches
constexpr int kNSPWG2 = 9;   // 3x3 points per singular patch
constexpr int kMaxDomains = 64;

constexpr double kPi = 3.141592653589793239;
constexpr double kOneOver4Pi = 0.0795774715459476679;
constexpr double kAbsTol = 1.0e-14;  // MultiDomainAbsTolerance
constexpr double kRelTol = 1.0e-10;  // MultiDomainRelTolerance

// Per-domain material constants in the reference's DynaMat layout
// (DBEM1/dsquareelement.h:22-56).
struct DynaMat {
    double G, v, Rou, C1, C2, Dt, Dt2, Dt_2, Dt_3, C1_2, C2_2, OneOverC1_2, OneOverC2_2, OneOverC1_3,
        OneOverC2_3, C2_2OverC1_2, C2_2OverC1_3, C1Dt, C2Dt, OneOverFourPaiRou, OneOverThree_1OC23_1OC13,
        Half_1OC22_1OC12, STCoef1, STCoef2;
};

// BuildDynaMat (DBEM1/dsquareelement.cpp:63-97).
inline DynaMat BuildDynaMat(double o_v, double o_G, double o_Rou, double o_dt)
{
    DynaMat m;
    m.v = o_v;
    m.G = o_G;
    m.Rou = o_Rou;
    m.C1 = sqrt((2.0 - 2.0 * o_v) / (1.0 - 2.0 * o_v) * o_G / o_Rou);
    m.C2 = sqrt(o_G / o_Rou);
    m.Dt = o_dt;
    m.Dt2 = 2.0 * o_dt;
    m.Dt_2 = o_dt * o_dt;
    m.Dt_3 = o_dt * o_dt * o_dt;
    m.OneOverFourPaiRou = kOneOver4Pi / o_Rou;
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

// Quadrature tables shared with the device (built once on the host).
struct QuadTables {
    double gap;
    double lc[8][2];               // serendipity node local coordinates at +-gap
    double rp[kNG2][2];            // regular 4x4 points
    double nrgv0[kNG2];            // N0(gap) * w  (regular-path weight)
    double lpospw[kNPW2][kNPWG2][2];
    double pwgv[kNPW2][kNPWG2];    // patch weights * (L/2)^2
    double cenpw[kNPW2][2];
    // constant-element singular tables
    double secconstpos[4][2][kNSG2];
    double secconstval[4][kNSG2];
    double constfm1secval[4][kNSG2];
    double constfm1lineval[4][kNSG];
    double linecos[4][kNSG];       // cos/sin of the line angles, host libm
    double linesin[4][kNSG];
    double secconstpospw[4][2][kNSPW2][kNSPWG2];
    double secconstvalpw[4][kNSPW2][kNSPWG2];
};

void BuildQuadTables(QuadTables& q, double gap);

// Element geometry needed by the kernels.
struct ElemGeo {
    double cc[3][8];        // serendipity coordinate polynomial
    double center[3];       // element centre == collocation node
    double radius;
    double radiusPW;
    double rcenpw[kNPW2][3];
    double trans[9];        // local frame rows t1, t2, n (from local to global via transpose)
};

// ---------------------------------------------------------------------------
// Geometry (DBEM1/dsquareelement.h:359-507,582-620)
// ---------------------------------------------------------------------------
__host__ __device__ inline void GetCoef(double* c, const double* p)
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

__host__ __device__ inline void FromLocal(const double (*cc)[8], double s1, double s2, double* pt)
{
    double s12 = s1 * s1, s22 = s2 * s2;
    double s1s22 = s1 * s22, s12s2 = s12 * s2;
    for (int i = 0; i < 3; ++i) {
        const double* c = cc[i];
        pt[i] = c[0] + c[1] * s1 + c[2] * s2 + c[3] * s1 * s2 + c[4] * s12 + c[5] * s22 + c[6] * s12s2 +
                c[7] * s1s22;
    }
}

__host__ __device__ inline void RDiffs1(const double (*cc)[8], double s1, double s2, double* r)
{
    double s1_2 = 2.0 * s1, s1s2_2 = 2.0 * s1 * s2, s22 = s2 * s2;
    for (int i = 0; i < 3; ++i) {
        const double* c = cc[i];
        r[i] = c[1] + c[3] * s2 + c[4] * s1_2 + c[6] * s1s2_2 + c[7] * s22;
    }
}

__host__ __device__ inline void RDiffs2(const double (*cc)[8], double s1, double s2, double* r)
{
    double s2_2 = 2.0 * s2, s1s2_2 = 2.0 * s1 * s2, s12 = s1 * s1;
    for (int i = 0; i < 3; ++i) {
        const double* c = cc[i];
        r[i] = c[2] + c[3] * s1 + c[5] * s2_2 + c[6] * s12 + c[7] * s1s2_2;
    }
}

__host__ __device__ inline void NormalAt(const double (*cc)[8], double s1, double s2, double* n)
{
    double r1[3], r2[3];
    RDiffs1(cc, s1, s2, r1);
    RDiffs2(cc, s1, s2, r2);
    n[0] = r1[1] * r2[2] - r2[1] * r1[2];
    n[1] = r1[2] * r2[0] - r2[2] * r1[0];
    n[2] = r1[0] * r2[1] - r2[0] * r1[1];
    double t = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] = n[0] / t;
    n[1] = n[1] / t;
    n[2] = n[2] / t;
}

__host__ __device__ inline double Jacobi(const double (*cc)[8], double s1, double s2)
{
    double ru[3], rv[3];
    RDiffs1(cc, s1, s2, ru);
    RDiffs2(cc, s1, s2, rv);
    double E = ru[0] * ru[0] + ru[1] * ru[1] + ru[2] * ru[2];
    double F = ru[0] * rv[0] + ru[1] * rv[1] + ru[2] * rv[2];
    double G = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
    return sqrt(fabs(E * G - F * F));
}

__host__ __device__ inline void TranslateMatrix(const double (*cc)[8], double s1, double s2, double* T)
{
    double res[3];
    RDiffs1(cc, s1, s2, res);
    double len = sqrt(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
    T[0] = res[0] / len;
    T[1] = res[1] / len;
    T[2] = res[2] / len;
    RDiffs2(cc, s1, s2, res);
    len = sqrt(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
    T[3] = res[0] / len;
    T[4] = res[1] / len;
    T[5] = res[2] / len;
    T[6] = T[1] * T[5] - T[4] * T[2];
    T[7] = T[3] * T[2] - T[0] * T[5];
    T[8] = T[0] * T[4] - T[3] * T[1];
    double temp = sqrt(T[6] * T[6] + T[7] * T[7] + T[8] * T[8]);
    T[6] /= temp;
    T[7] /= temp;
    T[8] /= temp;
    T[3] = T[7] * T[2] - T[8] * T[1];
    T[4] = T[8] * T[0] - T[6] * T[2];
    T[5] = T[6] * T[1] - T[7] * T[0];
}

__host__ __device__ inline double Dist3(const double* a, const double* b)
{
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

// ---------------------------------------------------------------------------
// Causality interval of a (field element, source point) pair.
// NeedCal(type 1, n) is non-zero exactly for n in [k0, k1]:
//   zero iff n*Dt >= Rmax/C2  or  (n+1)*Dt <= max(r-R,0)/C1
// (DBEM1/dsquareelement.cpp:1156-1184).  Both tests are monotone in n, so the
// bounds are found from a floating estimate and corrected with the exact
// comparisons the reference evaluates.
// ---------------------------------------------------------------------------
struct CullBounds {
    double a2;  // Rmax / C2
    double b1;  // Rmin / C1
};

__host__ __device__ inline CullBounds PairBounds(const DynaMat& m, const ElemGeo& f, const double* src)
{
    double r = Dist3(f.center, src);
    double Rmax = r + f.radius;
    double Rmin = r - f.radius;
    if (Rmin < 0.0)
        Rmin = 0.0;
    CullBounds b;
    b.a2 = Rmax / m.C2;
    b.b1 = Rmin / m.C1;
    return b;
}

// Largest n >= -1 with n*Dt < a2 (all n below it also satisfy it).
__host__ __device__ inline long LastActive(const DynaMat& m, double a2, long cap)
{
    double est = a2 / m.Dt;
    long n = est < (double)cap ? (long)est : cap;
    if (n < 0)
        n = 0;
    while (n >= 0 && (double)n * m.Dt >= a2)
        --n;
    while (n + 1 <= cap && (double)(n + 1) * m.Dt < a2)
        ++n;
    return n;
}

// Smallest n >= 0 with (n+1)*Dt > b1.
__host__ __device__ inline long FirstActive(const DynaMat& m, double b1, long cap)
{
    double est = b1 / m.Dt - 1.0;
    long n = est > 0.0 ? (est < (double)cap ? (long)est : cap) : 0;
    while (n > 0 && ((double)(n - 1) + 1) * m.Dt > b1)
        --n;
    while (n <= cap && ((double)n + 1) * m.Dt <= b1)
        ++n;
    return n;
}

}  // namespace tdbem
