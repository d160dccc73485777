// HIP kernels of the multi-domain TDBEM path for gfx950 (CDNA4).
//
// FP64 arithmetic keeps the reference's operation order (see common.hpp);
// the quadrature of one (field, source, step) triple is accumulated by one
// lane in the reference's point order, so block values and the 1e-14 drop
// decisions match the CPU restatement bit for bit.
#include "kernels.hpp"

#include <stdexcept>

#pragma clang fp contract(off)

namespace tdbem {


// ---------------------------------------------------------------------------
// Time-integrated elastodynamic kernels (DBEM1/dsquareelement.h:631-903,
// DBEM1/dsquareelement.cpp:1070-1590)
// ---------------------------------------------------------------------------
namespace dev {

__device__ __forceinline__ int Fai(double n, double dt, double r, double c)
{
    double ndt = n * dt;
    double rc = r / c;
    return ((rc >= ndt) && (rc < (ndt + dt))) ? 1 : 0;
}

// ISeries / JSeries: the five-way window of the between-fronts term.
__device__ __forceinline__ int Window(double n, double dt, double r, const DynaMat& m, double& ndt, double& np1dt)
{
    ndt = n * dt;
    np1dt = ndt + dt;
    double rc1 = r / m.C1, rc2 = r / m.C2;
    if (np1dt <= rc1 || ndt >= rc2)
        return 0;
    if (ndt >= rc1 && np1dt <= rc2)
        return 1;
    if (ndt <= rc1 && (np1dt >= rc1 && np1dt <= rc2))
        return 2;
    if (ndt <= rc1 && np1dt >= rc2)
        return 3;
    if (np1dt >= rc2 && (ndt >= rc1 && ndt <= rc2))
        return 4;
    return 0;
}

__device__ __forceinline__ double ISer(int w, double n, double ndt, double np1dt, const double* R, const DynaMat& m)
{
    double t1;
    switch (w) {
    case 1: return (n + 0.5) * m.Dt_2 / R[2];
    case 2: t1 = np1dt / R[1]; return 0.5 * (t1 * t1 - m.OneOverC1_2);
    case 3: return m.Half_1OC22_1OC12;
    case 4: t1 = ndt / R[1]; return 0.5 * (m.OneOverC2_2 - t1 * t1);
    default: return 0.0;
    }
}

__device__ __forceinline__ double JSer(int w, double n, double ndt, double np1dt, const double* R, const DynaMat& m)
{
    double t1;
    switch (w) {
    case 1: t1 = m.Dt_3 / R[3]; return (n * n + n + 0.33333333333333333) * t1;
    case 2: t1 = np1dt / R[1]; return 0.33333333333333 * (t1 * t1 * t1 - m.OneOverC1_3);
    case 3: return m.OneOverThree_1OC23_1OC13;
    case 4: t1 = ndt / R[1]; return 0.333333333333333 * (m.OneOverC2_3 - t1 * t1 * t1);
    default: return 0.0;
    }
}

struct Dist {
    double R[6];
    double RI[4];
    double X[6];  // ri*rj products
};

__device__ __forceinline__ void MakeDist(const double* src, const double* pt, Dist& d)
{
    d.RI[1] = pt[0] - src[0];
    d.RI[2] = pt[1] - src[1];
    d.RI[3] = pt[2] - src[2];
    d.R[1] = sqrt(d.RI[1] * d.RI[1] + d.RI[2] * d.RI[2] + d.RI[3] * d.RI[3]);
    d.R[2] = d.R[1] * d.R[1];
    d.R[3] = d.R[2] * d.R[1];
    d.R[4] = d.R[3] * d.R[1];
    d.R[5] = d.R[4] * d.R[1];
    d.X[0] = d.RI[1] * d.RI[1];
    d.X[1] = d.RI[1] * d.RI[2];
    d.X[2] = d.RI[1] * d.RI[3];
    d.X[3] = d.RI[2] * d.RI[2];
    d.X[4] = d.RI[2] * d.RI[3];
    d.X[5] = d.RI[3] * d.RI[3];
}

// GetDynaUij: P-front (Aij), S-front (Bij) and between-fronts (Cij) terms.
__device__ int DynaU(const DynaMat& m, double n, const Dist& d, double* U)
{
    const double* R = d.R;
    const double* X = d.X;
    int flag = 0;
    for (int k = 0; k < 9; ++k)
        U[k] = 0.0;
    if (Fai(n, m.Dt, R[1], m.C1)) {
        flag++;
        double C1 = m.OneOverC1_2 / R[3];
        double t1 = m.OneOverFourPaiRou;
        double a0 = C1 * X[0], a1 = C1 * X[1], a2 = C1 * X[2], a4 = C1 * X[3], a5 = C1 * X[4], a8 = C1 * X[5];
        U[0] += a0 * t1; U[1] += a1 * t1; U[2] += a2 * t1;
        U[3] += a1 * t1; U[4] += a4 * t1; U[5] += a5 * t1;
        U[6] += a2 * t1; U[7] += a5 * t1; U[8] += a8 * t1;
    }
    if (Fai(n, m.Dt, R[1], m.C2)) {
        flag++;
        double c = m.OneOverC2_2;
        double b0 = c * (1.0 / R[1] - X[0] / R[3]);
        double b1 = c * (-X[1] / R[3]);
        double b2 = c * (-X[2] / R[3]);
        double b4 = c * (1.0 / R[1] - X[3] / R[3]);
        double b5 = c * (-X[4] / R[3]);
        double b8 = c * (1.0 / R[1] - X[5] / R[3]);
        double t1 = m.OneOverFourPaiRou;
        U[0] += b0 * t1; U[1] += b1 * t1; U[2] += b2 * t1;
        U[3] += b1 * t1; U[4] += b4 * t1; U[5] += b5 * t1;
        U[6] += b2 * t1; U[7] += b5 * t1; U[8] += b8 * t1;
    }
    double ndt, np1dt;
    int w = Window(n, m.Dt, R[1], m, ndt, np1dt);
    if (w) {
        flag++;
        double res = ISer(w, n, ndt, np1dt, R, m);
        double C1 = 3.0 / R[3], C2 = 1.0 / R[1];
        double c0 = C1 * X[0] - C2, c1 = C1 * X[1], c2 = C1 * X[2], c4 = C1 * X[3] - C2, c5 = C1 * X[4],
               c8 = C1 * X[5] - C2;
        double t1 = m.OneOverFourPaiRou * res;
        U[0] += c0 * t1; U[1] += c1 * t1; U[2] += c2 * t1;
        U[3] += c1 * t1; U[4] += c4 * t1; U[5] += c5 * t1;
        U[6] += c2 * t1; U[7] += c5 * t1; U[8] += c8 * t1;
    }
    return flag;
}

__device__ __forceinline__ void Dij(const DynaMat& m, const Dist& d, const double* N, double* o)
{
    const double *R = d.R, *RI = d.RI, *X = d.X;
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = (4.0 * m.C2_2OverC1_2 - 1.0) / R[3];
    double C2 = m.C2_2OverC1_2 * 12.0 * RmNm / R[5];
    double C3 = m.C2_2OverC1_2 * 2.0 / R[3];
    o[0] = C1 * RI[1] * N[0] - (C2 * X[0] - C3 * (RI[1] * N[0] + RmNm));
    o[1] = C1 * RI[2] * N[0] - (C2 * X[1] - C3 * (RI[1] * N[1]));
    o[2] = C1 * RI[3] * N[0] - (C2 * X[2] - C3 * (RI[1] * N[2]));
    o[3] = C1 * RI[1] * N[1] - (C2 * X[1] - C3 * (RI[2] * N[0]));
    o[4] = C1 * RI[2] * N[1] - (C2 * X[3] - C3 * (RI[2] * N[1] + RmNm));
    o[5] = C1 * RI[3] * N[1] - (C2 * X[4] - C3 * (RI[2] * N[2]));
    o[6] = C1 * RI[1] * N[2] - (C2 * X[2] - C3 * (RI[3] * N[0]));
    o[7] = C1 * RI[2] * N[2] - (C2 * X[4] - C3 * (RI[3] * N[1]));
    o[8] = C1 * RI[3] * N[2] - (C2 * X[5] - C3 * (RI[3] * N[2] + RmNm));
}

__device__ __forceinline__ void Eij(const Dist& d, const double* N, double* o)
{
    const double *R = d.R, *RI = d.RI, *X = d.X;
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = 12.0 * RmNm / R[5];
    double C2 = 2.0 / R[3];
    double C3 = 3.0 / R[3];
    o[0] = C1 * X[0] - C2 * RI[1] * N[0] - C3 * (RI[1] * N[0] + RmNm);
    o[1] = C1 * X[1] - C2 * RI[2] * N[0] - C3 * (RI[1] * N[1]);
    o[2] = C1 * X[2] - C2 * RI[3] * N[0] - C3 * (RI[1] * N[2]);
    o[3] = C1 * X[1] - C2 * RI[1] * N[1] - C3 * (RI[2] * N[0]);
    o[4] = C1 * X[3] - C2 * RI[2] * N[1] - C3 * (RI[2] * N[1] + RmNm);
    o[5] = C1 * X[4] - C2 * RI[3] * N[1] - C3 * (RI[2] * N[2]);
    o[6] = C1 * X[2] - C2 * RI[1] * N[2] - C3 * (RI[3] * N[0]);
    o[7] = C1 * X[4] - C2 * RI[2] * N[2] - C3 * (RI[3] * N[1]);
    o[8] = C1 * X[5] - C2 * RI[3] * N[2] - C3 * (RI[3] * N[2] + RmNm);
}

__device__ __forceinline__ void Fij(const DynaMat& m, const Dist& d, const double* N, double* o)
{
    const double *R = d.R, *RI = d.RI, *X = d.X;
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = -6.0 * m.C2_2;
    double C2 = 5.0 * RmNm / R[5];
    o[0] = C1 * (C2 * X[0] - (RI[1] * N[0] + RI[1] * N[0] + RmNm) / R[3]);
    o[1] = C1 * (C2 * X[1] - (RI[2] * N[0] + RI[1] * N[1]) / R[3]);
    o[2] = C1 * (C2 * X[2] - (RI[3] * N[0] + RI[1] * N[2]) / R[3]);
    o[3] = o[1];
    o[4] = C1 * (C2 * X[3] - (RI[2] * N[1] + RI[2] * N[1] + RmNm) / R[3]);
    o[5] = C1 * (C2 * X[4] - (RI[3] * N[1] + RI[2] * N[2]) / R[3]);
    o[6] = o[2];
    o[7] = o[5];
    o[8] = C1 * (C2 * X[5] - (RI[3] * N[2] + RI[3] * N[2] + RmNm) / R[3]);
}

__device__ __forceinline__ void Gij(const DynaMat& m, const Dist& d, const double* N, double* o)
{
    const double *R = d.R, *RI = d.RI, *X = d.X;
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = -2.0 * m.C2_2OverC1_3 * RmNm / R[4];
    double C2 = 1.0 / m.C1 * (2.0 * m.C2_2OverC1_2 - 1.0) / R[2];
    o[0] = C1 * X[0] + C2 * RI[1] * N[0];
    o[1] = C1 * X[1] + C2 * RI[2] * N[0];
    o[2] = C1 * X[2] + C2 * RI[3] * N[0];
    o[3] = C1 * X[1] + C2 * RI[1] * N[1];
    o[4] = C1 * X[3] + C2 * RI[2] * N[1];
    o[5] = C1 * X[4] + C2 * RI[3] * N[1];
    o[6] = C1 * X[2] + C2 * RI[1] * N[2];
    o[7] = C1 * X[4] + C2 * RI[2] * N[2];
    o[8] = C1 * X[5] + C2 * RI[3] * N[2];
}

__device__ __forceinline__ void Hij(const DynaMat& m, const Dist& d, const double* N, double* o)
{
    const double *R = d.R, *RI = d.RI, *X = d.X;
    double RmNm = RI[1] * N[0] + RI[2] * N[1] + RI[3] * N[2];
    double C1 = 2.0 / m.C2 * RmNm / R[4];
    double C2 = -1.0 / m.C2 / R[2];
    o[0] = C1 * X[0] + C2 * (RI[1] * N[0] + RmNm);
    o[1] = C1 * X[1] + C2 * RI[1] * N[1];
    o[2] = C1 * X[2] + C2 * RI[1] * N[2];
    o[3] = C1 * X[1] + C2 * RI[2] * N[0];
    o[4] = C1 * X[3] + C2 * (RI[2] * N[1] + RmNm);
    o[5] = C1 * X[4] + C2 * RI[2] * N[2];
    o[6] = C1 * X[2] + C2 * RI[3] * N[0];
    o[7] = C1 * X[4] + C2 * RI[3] * N[1];
    o[8] = C1 * X[5] + C2 * (RI[3] * N[2] + RmNm);
}

// GetDynaT1ij (type 1, time level n) and GetDynaT2ij (type 2, level n-1).
__device__ int DynaT(const DynaMat& m, int type, double n, const Dist& d, const double* N, double* T)
{
    const double* R = d.R;
    double K1[9], K2[9];
    int flag = 0;
    for (int k = 0; k < 9; ++k)
        T[k] = 0.0;
    double nn = type == 1 ? n : n - 1;
    if (Fai(nn, m.Dt, R[1], m.C1)) {
        flag++;
        Dij(m, d, N, K1);
        Gij(m, d, N, K2);
        double t1, t2;
        if (type == 1) {
            t1 = kOneOver4Pi * (n + 1.0 - R[1] / m.C1Dt);
            t2 = kOneOver4Pi / m.Dt;
        } else {
            t1 = kOneOver4Pi * (1.0 - n + R[1] / m.C1Dt);
            t2 = -kOneOver4Pi / m.Dt;
        }
        for (int k = 0; k < 9; ++k)
            T[k] += K1[k] * t1 + K2[k] * t2;
    }
    if (Fai(nn, m.Dt, R[1], m.C2)) {
        flag++;
        Eij(d, N, K1);
        Hij(m, d, N, K2);
        double t1, t2;
        if (type == 1) {
            t1 = kOneOver4Pi * (n + 1.0 - R[1] / m.C2Dt);
            t2 = kOneOver4Pi / m.Dt;
        } else {
            t1 = kOneOver4Pi * (1.0 - n + R[1] / m.C2Dt);
            t2 = -kOneOver4Pi / m.Dt;
        }
        for (int k = 0; k < 9; ++k)
            T[k] += K1[k] * t1 + K2[k] * t2;
    }
    double ndt, np1dt;
    int w = Window(nn, m.Dt, R[1], m, ndt, np1dt);
    if (w) {
        flag++;
        double ri = ISer(w, nn, ndt, np1dt, R, m);
        double rj = JSer(w, nn, ndt, np1dt, R, m);
        Fij(m, d, N, K1);
        double t1;
        if (type == 1)
            t1 = kOneOver4Pi * ((n + 1.0) * ri - R[1] * rj / m.Dt);
        else
            t1 = kOneOver4Pi * ((1.0 - n) * ri + R[1] * rj / m.Dt);
        for (int k = 0; k < 9; ++k)
            T[k] += K1[k] * t1;
    }
    return flag;
}

// Static traction kernel GetTij (DBEM1/dsquareelement.cpp:496-543).
__device__ void StaticT(const DynaMat& m, const double* R, const double* n, double* T)
{
    double t1 = R[0] * R[0];
    double t2 = m.STCoef1 / t1;
    double t3 = m.STCoef2 / R[0];
    double drdn = (R[1] * n[0] + R[2] * n[1] + R[3] * n[2]) / R[0];
    double t5 = t3 * (R[2] * n[0] - R[1] * n[1]);
    double t6 = t3 * (R[3] * n[0] - R[1] * n[2]);
    double t7 = t3 * (R[3] * n[1] - R[2] * n[2]);
    double t8 = t2 * drdn;
    double t9 = t8 * m.STCoef2;
    double t10 = t8 * 3.0 / t1;
    double t4 = t10 * R[1];
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

// ---------------------------------------------------------------------------
// Causality judges (DBEM1/dsquareelement.cpp:1156-1236, .h:926-943)
// ---------------------------------------------------------------------------
__device__ int NeedCal(const DynaMat& m, const ElemGeo& e, int type, double n, const double* o)
{
    double nn = (type == 2) ? n - 1 : n;
    double ndt = nn * m.Dt;
    double np1dt = (nn + 1) * m.Dt;
    double r = Dist3(e.center, o);
    double Rmax = r + e.radius;
    double Rmin = r - e.radius;
    if (Rmin < 0.0)
        Rmin = 0.0;
    double a1 = Rmax / m.C1, a2 = Rmax / m.C2, b1 = Rmin / m.C1, b2 = Rmin / m.C2;
    if (ndt >= a2 || np1dt <= b1)
        return 0;
    if (ndt <= b1 && (np1dt >= a1 && np1dt <= b2))
        return 2;
    if (ndt <= b1 && np1dt >= a2)
        return 2;
    if (ndt >= a1 && np1dt <= b2)
        return 2;
    if ((ndt >= a1 && ndt <= b2) && np1dt >= a2)
        return 2;
    return 1;
}

__device__ __forceinline__ int PWNeedCal(const DynaMat& m, const ElemGeo& e, int type, double n, const double* c,
                                         const double* o)
{
    double nn = (type == 2 ? n - 1 : n);
    double c2ndt = nn * m.C2Dt;
    double c1np1dt = (nn + 1) * m.C1Dt;
    double r = Dist3(c, o);
    double rpr = r + e.radiusPW;
    double rmr = r - e.radiusPW;
    return (rmr >= c1np1dt || rpr < c2ndt) ? 0 : 1;
}

// ---------------------------------------------------------------------------
// Quadrature: patch-wise (DBEM1/dsquareelementdyna.cpp:794-848,940-994) and
// regular with the N0(gap) weight and T1+T2 sum (DBEM1/dsquareelement.cpp:
// 1955-2009,2233-2290).  Column-major 3x3 output.
// ---------------------------------------------------------------------------
__device__ void IntUPW(const DynaMat& m, const ElemGeo& e, const ElemPoints& P, const double* src, double n,
                       double* out)
{
    double o0 = 0.0, o1 = 0.0, o2 = 0.0, o4 = 0.0, o5 = 0.0, o8 = 0.0;
    for (int i = 0; i < kNPW2; ++i) {
        if (!PWNeedCal(m, e, 1, n, e.rcenpw[i], src))
            continue;
        for (int j = 0; j < kNPWG2; ++j) {
            int q = i * kNPWG2 + j;
            Dist d;
            MakeDist(src, P.pw_pos[q], d);
            double U[9];
            DynaU(m, n, d, U);
            double w = P.pw_w[q];
            o0 += U[0] * w;
            o1 += U[1] * w;
            o2 += U[2] * w;
            o4 += U[4] * w;
            o5 += U[5] * w;
            o8 += U[8] * w;
        }
    }
    out[0] = o0; out[1] = o1; out[2] = o2; out[3] = o1; out[4] = o4;
    out[5] = o5; out[6] = o2; out[7] = o5; out[8] = o8;
}

__device__ void IntTPW(const DynaMat& m, const ElemGeo& e, const ElemPoints& P, int type, const double* src,
                       double n, double* out)
{
    double acc[9];
    for (int k = 0; k < 9; ++k)
        acc[k] = 0.0;
    for (int i = 0; i < kNPW2; ++i) {
        if (!PWNeedCal(m, e, type, n, e.rcenpw[i], src))
            continue;
        for (int j = 0; j < kNPWG2; ++j) {
            int q = i * kNPWG2 + j;
            Dist d;
            MakeDist(src, P.pw_pos[q], d);
            double T[9];
            DynaT(m, type, n, d, P.pw_nor[q], T);
            double w = P.pw_w[q];
            for (int k = 0; k < 9; ++k)
                acc[k] += T[k] * w;
        }
    }
    for (int k = 0; k < 9; ++k)
        out[k] = acc[k];
}

__device__ void IntUReg(const DynaMat& m, const ElemPoints& P, const double* src, double n, double* out)
{
    double o0 = 0.0, o1 = 0.0, o2 = 0.0, o4 = 0.0, o5 = 0.0, o8 = 0.0;
    for (int i = 0; i < kNG2; ++i) {
        Dist d;
        MakeDist(src, P.rg_pos[i], d);
        double U[9];
        if (DynaU(m, n, d, U)) {
            double w = P.rg_w[i];
            o0 += U[0] * w;
            o1 += U[1] * w;
            o2 += U[2] * w;
            o4 += U[4] * w;
            o5 += U[5] * w;
            o8 += U[8] * w;
        }
    }
    out[0] = o0; out[1] = o1; out[2] = o2; out[3] = o1; out[4] = o4;
    out[5] = o5; out[6] = o2; out[7] = o5; out[8] = o8;
}

__device__ void IntTReg(const DynaMat& m, const ElemPoints& P, const double* src, double n, double* out)
{
    double acc[9];
    for (int k = 0; k < 9; ++k)
        acc[k] = 0.0;
    for (int i = 0; i < kNG2; ++i) {
        Dist d;
        MakeDist(src, P.rg_pos[i], d);
        double T1[9], T2[9];
        int flag = DynaT(m, 1, n, d, P.rg_nor[i], T1);
        flag += DynaT(m, 2, n, d, P.rg_nor[i], T2);
        if (flag) {
            double w = P.rg_w[i];
            for (int k = 0; k < 9; ++k)
                acc[k] += (T1[k] + T2[k]) * w;
        }
    }
    for (int k = 0; k < 9; ++k)
        out[k] = acc[k];
}

__device__ int JudgeU(const DynaMat& m, const ElemGeo& e, const ElemPoints& P, const double* src, double n,
                      double* out)
{
    int f = NeedCal(m, e, 1, n, src);
    if (f == 1)
        IntUPW(m, e, P, src, n, out);
    else if (f == 2)
        IntUReg(m, P, src, n, out);
    return f;
}

__device__ int JudgeT(const DynaMat& m, const ElemGeo& e, const ElemPoints& P, int type, const double* src,
                      double n, double* out)
{
    int f = NeedCal(m, e, type, n, src);
    if (f == 1)
        IntTPW(m, e, P, type, src, n, out);
    else if (f == 2)
        IntTReg(m, P, src, n, out);
    return f;
}

// MultiDomainCCSRBuilder::AddBlock semantics for one key (transpose * sign,
// 1e-14 drop of the input and of the accumulated value).
struct BlockAcc {
    double v[9];
    bool alive = false;
    __device__ void Add(double sign, const double* b)
    {
        bool zero = true, finite = true;
        double mx = 0.0;
        for (int i = 0; i < 9; ++i) {
            if (b[i] != 0.0)
                zero = false;
            double a = fabs(b[i]);
            if (a > mx)
                mx = a;
            if (!(b[i] == b[i] && a <= 1.7976931348623157e308))
                finite = false;
        }
        if (sign == 0.0 || zero || mx <= kAbsTol || !finite)
            return;
        if (!alive) {
            for (int i = 0; i < 9; ++i)
                v[i] = 0.0;
            alive = true;
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                v[r * 3 + c] += sign * b[c * 3 + r];
        double m2 = 0.0;
        for (int i = 0; i < 9; ++i) {
            double a = fabs(v[i]);
            if (a > m2)
                m2 = a;
        }
        if (m2 <= kAbsTol)
            alive = false;
    }
};

__device__ __forceinline__ void DecodePair(const PairGrid& g, long p, int& s, int& f, int& d)
{
    DecodePairOf(g, p, s, f, d);
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Setup kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_points(const ElemGeo* __restrict__ geo, const QuadTables* __restrict__ q,
                                                ElemPoints* __restrict__ pts, int ne)
{
    constexpr int kPer = kNPW2 * kNPWG2 + kNG2;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ne * kPer)
        return;
    int e = (int)(t / kPer), i = (int)(t % kPer);
    const ElemGeo& g = geo[e];
    ElemPoints& P = pts[e];
    if (i < kNPW2 * kNPWG2) {
        int p = i / kNPWG2, m = i % kNPWG2;
        double s1 = q->lpospw[p][m][0], s2 = q->lpospw[p][m][1];
        FromLocal(g.cc, s1, s2, P.pw_pos[i]);
        NormalAt(g.cc, s1, s2, P.pw_nor[i]);
        P.pw_w[i] = q->pwgv[p][m] * Jacobi(g.cc, s1, s2);
    } else {
        int r = i - kNPW2 * kNPWG2;
        double s1 = q->rp[r][0], s2 = q->rp[r][1];
        FromLocal(g.cc, s1, s2, P.rg_pos[r]);
        NormalAt(g.cc, s1, s2, P.rg_nor[r]);
        P.rg_w[r] = q->nrgv0[r] * Jacobi(g.cc, s1, s2);
    }
}

void LaunchPoints(const ElemGeo* geo, const QuadTables* q, ElemPoints* pts, int ne, hipStream_t s)
{
    long n = (long)ne * (kNPW2 * kNPWG2 + kNG2);
    k_points<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(geo, q, pts, ne);
}

// Singular self integrals: InElementStaticTij + PthInElementIntDynamic
// (DBEM1/dsquareelement.cpp:2292-2427, DBEM1/dsquareelementdyna.cpp:239-490).
// One lane per element keeps the reference's summation order.
__global__ void __launch_bounds__(64) k_singular(const ElemGeo* __restrict__ geo, const int* __restrict__ dom,
                                                 const DynaMat* __restrict__ mats, const QuadTables* __restrict__ q, double* u1o, double* t1o,
                                                 double* t2o, int ne)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    const ElemGeo& g = geo[e];
    const DynaMat& m = mats[dom[e]];
    double node[3];
    FromLocal(g.cc, 0.0, 0.0, node);
    double st[9];
    for (int k = 0; k < 9; ++k)
        st[k] = 0.0;
    double coef1 = -1.0 / (8.0 * kPi * (1.0 - m.v));
    double coef2 = 1.0 - 2.0 * m.v;
    // strong-singular static part: area terms, then F(-1) subtraction
    for (int a = 0; a < 4; ++a)
        for (int j = 0; j < kNSG2; ++j) {
            double s1 = q->secconstpos[a][0][j], s2 = q->secconstpos[a][1][j];
            double J = Jacobi(g.cc, s1, s2);
            double nor[3], p[3], R[4], T[9];
            NormalAt(g.cc, s1, s2, nor);
            FromLocal(g.cc, s1, s2, p);
            R[1] = p[0] - node[0];
            R[2] = p[1] - node[1];
            R[3] = p[2] - node[2];
            R[0] = sqrt(R[1] * R[1] + R[2] * R[2] + R[3] * R[3]);
            dev::StaticT(m, R, nor, T);
            double temp = q->secconstval[a][j] * J;
            for (int l = 0; l < 9; ++l)
                st[l] += T[l] * temp;
        }
    double Jh[3], D1[3], D2[3];
    {
        double tj = Jacobi(g.cc, q->lc[0][0], q->lc[0][1]);
        double n0[3];
        NormalAt(g.cc, q->lc[0][0], q->lc[0][1], n0);
        Jh[0] = tj * n0[0];
        Jh[1] = tj * n0[1];
        Jh[2] = tj * n0[2];
    }
    RDiffs1(g.cc, 0.0, 0.0, D1);
    RDiffs2(g.cc, 0.0, 0.0, D2);
    for (int a = 0; a < 4; ++a)
        for (int j = 0; j < kNSG; ++j) {
            double cs = q->linecos[a][j], sn = q->linesin[a][j];
            double A[6], F[9];
            A[1] = D1[0] * cs + D2[0] * sn;
            A[2] = D1[1] * cs + D2[1] * sn;
            A[3] = D1[2] * cs + D2[2] * sn;
            A[0] = sqrt(A[1] * A[1] + A[2] * A[2] + A[3] * A[3]);
            A[4] = A[0] * A[0] * A[0];
            A[5] = A[4] * A[0] * A[0];
            // GetFijMinusOne (DBEM1/dsquareelement.cpp:609-639)
            double AkJk0 = A[1] * Jh[0] + A[2] * Jh[1] + A[3] * Jh[2];
            double t1 = 3.0 * A[1] * AkJk0 / A[5];
            double t2 = coef2 * (A[2] * Jh[0] - A[1] * Jh[1]) / A[4];
            double t3 = coef2 * (A[3] * Jh[0] - A[1] * Jh[2]) / A[4];
            double t4 = coef2 * (A[3] * Jh[1] - A[2] * Jh[2]) / A[4];
            double temp = coef2 * AkJk0 / A[4];
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
            for (int l = 0; l < kNSG; ++l) {
                double fv = q->constfm1secval[a][l * kNSG + j];
                for (int k = 0; k < 9; ++k)
                    st[k] -= F[k] * fv;
            }
            double fl = q->constfm1lineval[a][j];
            for (int k = 0; k < 9; ++k)
                st[k] += F[k] * fl;
        }
    // dynamic parts over the 4 x 400 x 9 polar points
    double u0 = 0.0, u1 = 0.0, u2 = 0.0, u4 = 0.0, u5 = 0.0, u8 = 0.0;
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < kNSPW2; ++k)
            for (int l = 0; l < kNSPWG2; ++l) {
                double s1 = q->secconstpospw[a][0][k][l], s2 = q->secconstpospw[a][1][k][l];
                double J = Jacobi(g.cc, s1, s2);
                double p[3];
                FromLocal(g.cc, s1, s2, p);
                dev::Dist d;
                dev::MakeDist(node, p, d);
                double U[9];
                dev::DynaU(m, 0.0, d, U);
                double temp = q->secconstvalpw[a][k][l] * J;
                u0 += U[0] * temp;
                u1 += U[1] * temp;
                u2 += U[2] * temp;
                u4 += U[4] * temp;
                u5 += U[5] * temp;
                u8 += U[8] * temp;
            }
    double* uo = u1o + 9 * (long)e;
    uo[0] = u0; uo[1] = u1; uo[2] = u2; uo[3] = u1; uo[4] = u4;
    uo[5] = u5; uo[6] = u2; uo[7] = u5; uo[8] = u8;
    double a1[9], a2[9];
    for (int k = 0; k < 9; ++k) {
        a1[k] = 0.0;
        a2[k] = 0.0;
    }
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < kNSPW2; ++k)
            for (int l = 0; l < kNSPWG2; ++l) {
                double s1 = q->secconstpospw[a][0][k][l], s2 = q->secconstpospw[a][1][k][l];
                double J = Jacobi(g.cc, s1, s2);
                double nor[3], p[3];
                NormalAt(g.cc, s1, s2, nor);
                FromLocal(g.cc, s1, s2, p);
                dev::Dist d;
                dev::MakeDist(node, p, d);
                double SR[4] = {d.R[1], d.RI[1], d.RI[2], d.RI[3]};
                double T1[9], T2[9], ST[9];
                dev::DynaT(m, 1, 0.0, d, nor, T1);
                dev::DynaT(m, 2, 1.0, d, nor, T2);
                dev::StaticT(m, SR, nor, ST);
                double temp = q->secconstvalpw[a][k][l] * J;
                for (int c = 0; c < 9; ++c) {
                    a1[c] += (T1[c] - ST[c]) * temp;
                    a2[c] += (T2[c] - ST[c]) * temp;
                }
            }
    double* to1 = t1o + 9 * (long)e;
    double* to2 = t2o + 9 * (long)e;
    for (int c = 0; c < 9; ++c) {
        to1[c] = a1[c] + st[c];
        to2[c] = a2[c] + st[c];
    }
    to1[0] += 0.5;
    to1[4] += 0.5;
    to1[8] += 0.5;
}

// The same self integrals with one wavefront per element: the lanes share the
// 64 static points, the 16 line points and the 1600 polar rows (strided),
// and the partial sums are added in a fixed tree order (deterministic; equal
// to the lane-per-element order up to rounding).
__global__ void __launch_bounds__(64) k_singular_wave(const ElemGeo* __restrict__ geo, const int* __restrict__ dom,
                                                      const DynaMat* __restrict__ mats,
                                                      const QuadTables* __restrict__ q, double* u1o, double* t1o,
                                                      double* t2o, int ne)
{
    constexpr int kVals = 33;  // st 9, u 6, a1 9, a2 9
    __shared__ double red[kVals][64];
    int e = blockIdx.x;
    int lane = threadIdx.x;
    const ElemGeo& g = geo[e];
    const DynaMat& m = mats[dom[e]];
    double node[3];
    FromLocal(g.cc, 0.0, 0.0, node);
    double v[kVals];
    for (int k = 0; k < kVals; ++k)
        v[k] = 0.0;
    double* st = v;
    double coef1 = -1.0 / (8.0 * kPi * (1.0 - m.v));
    double coef2 = 1.0 - 2.0 * m.v;
    for (int idx = lane; idx < 4 * kNSG2; idx += 64) {
        int a = idx / kNSG2, j = idx - a * kNSG2;
        double s1 = q->secconstpos[a][0][j], s2 = q->secconstpos[a][1][j];
        double J = Jacobi(g.cc, s1, s2);
        double nor[3], p[3], R[4], T[9];
        NormalAt(g.cc, s1, s2, nor);
        FromLocal(g.cc, s1, s2, p);
        R[1] = p[0] - node[0];
        R[2] = p[1] - node[1];
        R[3] = p[2] - node[2];
        R[0] = sqrt(R[1] * R[1] + R[2] * R[2] + R[3] * R[3]);
        dev::StaticT(m, R, nor, T);
        double temp = q->secconstval[a][j] * J;
        for (int l = 0; l < 9; ++l)
            st[l] += T[l] * temp;
    }
    if (lane < 4 * kNSG) {
        double Jh[3], D1[3], D2[3];
        double tj = Jacobi(g.cc, q->lc[0][0], q->lc[0][1]);
        double n0[3];
        NormalAt(g.cc, q->lc[0][0], q->lc[0][1], n0);
        Jh[0] = tj * n0[0];
        Jh[1] = tj * n0[1];
        Jh[2] = tj * n0[2];
        RDiffs1(g.cc, 0.0, 0.0, D1);
        RDiffs2(g.cc, 0.0, 0.0, D2);
        int a = lane / kNSG, j = lane - a * kNSG;
        double cs = q->linecos[a][j], sn = q->linesin[a][j];
        double A[6], F[9];
        A[1] = D1[0] * cs + D2[0] * sn;
        A[2] = D1[1] * cs + D2[1] * sn;
        A[3] = D1[2] * cs + D2[2] * sn;
        A[0] = sqrt(A[1] * A[1] + A[2] * A[2] + A[3] * A[3]);
        A[4] = A[0] * A[0] * A[0];
        A[5] = A[4] * A[0] * A[0];
        double AkJk0 = A[1] * Jh[0] + A[2] * Jh[1] + A[3] * Jh[2];
        double t1 = 3.0 * A[1] * AkJk0 / A[5];
        double t2 = coef2 * (A[2] * Jh[0] - A[1] * Jh[1]) / A[4];
        double t3 = coef2 * (A[3] * Jh[0] - A[1] * Jh[2]) / A[4];
        double t4 = coef2 * (A[3] * Jh[1] - A[2] * Jh[2]) / A[4];
        double temp = coef2 * AkJk0 / A[4];
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
        for (int l = 0; l < kNSG; ++l) {
            double fv = q->constfm1secval[a][l * kNSG + j];
            for (int k = 0; k < 9; ++k)
                st[k] -= F[k] * fv;
        }
        double fl = q->constfm1lineval[a][j];
        for (int k = 0; k < 9; ++k)
            st[k] += F[k] * fl;
    }
    double* u = v + 9;
    double* a1 = v + 15;
    double* a2 = v + 24;
    for (int row = lane; row < 4 * kNSPW2; row += 64) {
        int a = row / kNSPW2, k = row - a * kNSPW2;
        for (int l = 0; l < kNSPWG2; ++l) {
            double s1 = q->secconstpospw[a][0][k][l], s2 = q->secconstpospw[a][1][k][l];
            double J = Jacobi(g.cc, s1, s2);
            double nor[3], p[3];
            NormalAt(g.cc, s1, s2, nor);
            FromLocal(g.cc, s1, s2, p);
            dev::Dist d;
            dev::MakeDist(node, p, d);
            double U[9];
            dev::DynaU(m, 0.0, d, U);
            double temp = q->secconstvalpw[a][k][l] * J;
            u[0] += U[0] * temp;
            u[1] += U[1] * temp;
            u[2] += U[2] * temp;
            u[3] += U[4] * temp;
            u[4] += U[5] * temp;
            u[5] += U[8] * temp;
            double SR[4] = {d.R[1], d.RI[1], d.RI[2], d.RI[3]};
            double T1[9], T2[9], ST[9];
            dev::DynaT(m, 1, 0.0, d, nor, T1);
            dev::DynaT(m, 2, 1.0, d, nor, T2);
            dev::StaticT(m, SR, nor, ST);
            for (int c = 0; c < 9; ++c) {
                a1[c] += (T1[c] - ST[c]) * temp;
                a2[c] += (T2[c] - ST[c]) * temp;
            }
        }
    }
    for (int k = 0; k < kVals; ++k)
        red[k][lane] = v[k];
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if (lane < w)
            for (int k = 0; k < kVals; ++k)
                red[k][lane] += red[k][lane + w];
        __syncthreads();
    }
    if (lane == 0) {
        double* uo = u1o + 9 * (long)e;
        uo[0] = red[9][0];
        uo[1] = red[10][0];
        uo[2] = red[11][0];
        uo[3] = red[10][0];
        uo[4] = red[12][0];
        uo[5] = red[13][0];
        uo[6] = red[11][0];
        uo[7] = red[13][0];
        uo[8] = red[14][0];
        double* to1 = t1o + 9 * (long)e;
        double* to2 = t2o + 9 * (long)e;
        for (int c = 0; c < 9; ++c) {
            to1[c] = red[15 + c][0] + red[c][0];
            to2[c] = red[24 + c][0] + red[c][0];
        }
        to1[0] += 0.5;
        to1[4] += 0.5;
        to1[8] += 0.5;
    }
}

void LaunchSingular(const ElemGeo* geo, const int* dom, const DynaMat* mats, const QuadTables* q, double* u1,
                    double* t1, double* t2, int ne, bool exact, hipStream_t s)
{
    if (ne <= 0)
        return;
    if (exact)
        k_singular<<<(ne + 63) / 64, 64, 0, s>>>(geo, dom, mats, q, u1, t1, t2, ne);
    else
        k_singular_wave<<<ne, 64, 0, s>>>(geo, dom, mats, q, u1, t1, t2, ne);
}

// ---------------------------------------------------------------------------
// Causality cull: one lane per (source, field) pair.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cull(const ElemGeo* __restrict__ geo, const int* __restrict__ dom,
                                              PairGrid grid, long npairs, int maxn, int* kb, int* len,
                                              unsigned char* nearFlag)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs)
        return;
    int s, f, d;
    dev::DecodePair(grid, p, s, f, d);
    const DynaMat& m = grid.mats[d];
    CullBounds b = PairBounds(m, geo[f], geo[s].center);
    long cap = (long)maxn + 2;
    long k1 = LastActive(m, b.a2, cap);
    long k0 = FirstActive(m, b.b1, cap);
    long kbeg = k0 > 1 ? k0 : 1;
    long kend = k1 + 1 < maxn ? k1 + 1 : maxn;
    if (s == f)
        kbeg = 1;
    long L = (k1 >= k0 || s == f) ? kend - kbeg + 1 : 0;
    if (L < 0)
        L = 0;
    kb[p] = (int)kbeg;
    len[p] = (int)L;
    nearFlag[p] = (unsigned char)(((k0 <= 1 && k1 >= k0) || s == f) ? 1 : 0);
}

void LaunchCull(const ElemGeo* geo, const int* dom, PairGrid grid, long npairs, int maxn, int* kb, int* len,
                unsigned char* nearFlag, hipStream_t s)
{
    k_cull<<<(unsigned)((npairs + 255) / 256), 256, 0, s>>>(geo, dom, grid, npairs, maxn, kb, len, nearFlag);
}

// ---------------------------------------------------------------------------
// Band bookkeeping on the device: exclusive prefix sums of the band lengths
// (entry offsets) and of their 64-lag chunk counts over all pairs, and the
// per-row lists of near pairs.  Three-phase scan: per-block totals of 4096
// consecutive values, an in-order scan of those totals, then each block
// writes its values' offsets.
// ---------------------------------------------------------------------------
constexpr int kScanPer = 16;                  // values per thread
constexpr int kScanBlock = 256 * kScanPer;    // values per block

__device__ __forceinline__ long BlockExclusiveScan(long v, long* sh, long* total)
{
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        long u = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    long incl = sh[t];
    *total = sh[255];
    __syncthreads();
    return incl - v;
}

__device__ __forceinline__ long ChunkCount(int l) { return (l + 63) / 64; }

__global__ void __launch_bounds__(256) k_scan_reduce(const int* __restrict__ len, long n, long* __restrict__ bsum,
                                                     long* __restrict__ bcsum)
{
    __shared__ long sh[256];
    long base = (long)blockIdx.x * kScanBlock + (long)threadIdx.x * kScanPer;
    long a = 0, c = 0;
    for (int i = 0; i < kScanPer; ++i) {
        long p = base + i;
        int l = p < n ? len[p] : 0;
        a += l;
        c += ChunkCount(l);
    }
    long ta, tc;
    BlockExclusiveScan(a, sh, &ta);
    BlockExclusiveScan(c, sh, &tc);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = ta;
        bcsum[blockIdx.x] = tc;
    }
}

// in-place exclusive scan of the nb block totals (one block); the grand
// totals go to bsum[nb], bcsum[nb]
__global__ void __launch_bounds__(256) k_scan_top(long* __restrict__ bsum, long* __restrict__ bcsum, long nb)
{
    __shared__ long sh[256];
    long per = (nb + 255) / 256;
    long b0 = (long)threadIdx.x * per;
    long a = 0, c = 0;
    for (long i = 0; i < per; ++i)
        if (b0 + i < nb) {
            a += bsum[b0 + i];
            c += bcsum[b0 + i];
        }
    long ta, tc;
    long ea = BlockExclusiveScan(a, sh, &ta);
    long ec = BlockExclusiveScan(c, sh, &tc);
    for (long i = 0; i < per; ++i)
        if (b0 + i < nb) {
            long va = bsum[b0 + i], vc = bcsum[b0 + i];
            bsum[b0 + i] = ea;
            bcsum[b0 + i] = ec;
            ea += va;
            ec += vc;
        }
    if (threadIdx.x == 0) {
        bsum[nb] = ta;
        bcsum[nb] = tc;
    }
}

__global__ void __launch_bounds__(256) k_scan_apply(const int* __restrict__ len, long n,
                                                    const long* __restrict__ bsum, const long* __restrict__ bcsum,
                                                    long nb, long* __restrict__ off, long* __restrict__ coff)
{
    __shared__ long sh[256];
    long base = (long)blockIdx.x * kScanBlock + (long)threadIdx.x * kScanPer;
    long a = 0, c = 0;
    for (int i = 0; i < kScanPer; ++i) {
        long p = base + i;
        int l = p < n ? len[p] : 0;
        a += l;
        c += ChunkCount(l);
    }
    long ta, tc;
    long ea = BlockExclusiveScan(a, sh, &ta) + bsum[blockIdx.x];
    long ec = BlockExclusiveScan(c, sh, &tc) + bcsum[blockIdx.x];
    for (int i = 0; i < kScanPer; ++i) {
        long p = base + i;
        if (p >= n)
            break;
        int l = len[p];
        off[p] = ea;
        coff[p] = ec;
        ea += l;
        ec += ChunkCount(l);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        off[n] = bsum[nb];
        coff[n] = bcsum[nb];
    }
}

void LaunchBandOffsets(const int* len, long n, long* work, long* off, long* coff, hipStream_t s)
{
    long nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb < 1)
        nb = 1;
    long* bsum = work;
    long* bcsum = work + nb + 1;
    k_scan_reduce<<<(unsigned)nb, 256, 0, s>>>(len, n, bsum, bcsum);
    k_scan_top<<<1, 256, 0, s>>>(bsum, bcsum, nb);
    k_scan_apply<<<(unsigned)nb, 256, 0, s>>>(len, n, bsum, bcsum, nb, off, coff);
}

long BandOffsetsWork(long n)
{
    long nb = (n + kScanBlock - 1) / kScanBlock;
    return 2 * (std::max(nb, 1L) + 1);
}

// Near pairs of each source row (band not empty, first lag <= tile + 1), in
// field order: one wavefront per row counts them (cnt) and, in the second
// pass, writes them from the row's offset ptr.
__global__ void __launch_bounds__(64) k_near_lists(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                   const int* __restrict__ kb, const int* __restrict__ len, int tile,
                                                   const long* __restrict__ ptr, int* __restrict__ cnt,
                                                   long* __restrict__ pairOut, int* __restrict__ fieldOut)
{
    int s = blockIdx.x;
    int lane = threadIdx.x;
    int d = refs[s].domain;
    int nd = grid.eleCount[d];
    int eb = grid.eleBegin[d];
    long rowStart = grid.pairBase[d] + (long)(s - eb) * nd;
    long w = ptr ? ptr[s] : 0;
    int total = 0;
    if (!ptr) {
        for (int q0 = 0; q0 < nd; q0 += 64) {
            int q = q0 + lane;
            long p = rowStart + q;
            bool near = q < nd && len[p] > 0 && kb[p] <= tile + 1;
            total += __popcll(__ballot(near));
        }
        if (lane == 0)
            cnt[s] = total;
        return;
    }
    // the list ordered by band start, field order within one start, so the
    // near kernels stop at the first pair whose band starts past the step's
    // lag limit
    for (int v = 0; v <= tile + 1; ++v)
        for (int q0 = 0; q0 < nd; q0 += 64) {
            int q = q0 + lane;
            long p = rowStart + q;
            bool near = q < nd && len[p] > 0 && kb[p] == v;
            unsigned long long m = __ballot(near);
            if (near) {
                int rank = __popcll(m & ((1ull << lane) - 1ull));
                pairOut[w + rank] = p;
                fieldOut[w + rank] = eb + q;
            }
            w += __popcll(m);
        }
}

void LaunchNearLists(PairGrid grid, const NodeRefs* refs, const int* kb, const int* len, int tile, int ne,
                     const long* ptr, int* cnt, long* pairOut, int* fieldOut, hipStream_t s)
{
    k_near_lists<<<ne, 64, 0, s>>>(grid, refs, kb, len, tile, ptr, cnt, pairOut, fieldOut);
}

// ---------------------------------------------------------------------------
// History quadrature: one lane per (pair, lag) band entry.
// MTS entry = AddBlock(T2) then AddBlock(T1); MGS entry = AddBlock(U)
// (DBEM1/multidomain.cpp:2017-2078).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_history(const ElemGeo* __restrict__ geo, const ElemPoints* __restrict__ pts,
                                                 const int* __restrict__ dom, const NodeRefs* __restrict__ refs,
                                                 const double* __restrict__ selfT2, PairGrid grid,
                                                 const long* __restrict__ off, const int* __restrict__ kb, long npairs,
                                                 long e0, long e1, long ebase, int maxLag,
                                                 double* __restrict__ mts, double* __restrict__ mgs,
                                                 unsigned char* __restrict__ alive)
{
    long e = e0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1)
        return;
    // last pair with off[p] <= e (that pair has a non-empty band)
    long lo = 0, hi = npairs;  // off has npairs+1 entries
    while (hi - lo > 1) {
        long mid = (lo + hi) >> 1;
        if (off[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    long p = lo;
    int s, f, d;
    dev::DecodePair(grid, p, s, f, d);
    int k = kb[p] + (int)(e - off[p]);
    if (k > maxLag)
        return;
    const DynaMat& m = grid.mats[d];
    const ElemGeo& fe = geo[f];
    const ElemPoints& P = pts[f];
    const double* src = geo[s].center;
    double n = (double)k;
    double blk[9];
    dev::BlockAcc acc;
    double su = refs[f].hu_sign;
    if (k <= 1 && s == f) {
        for (int i = 0; i < 9; ++i)
            blk[i] = selfT2[9 * (long)f + i];
        acc.Add(su, blk);
    } else if (dev::JudgeT(m, fe, P, 2, src, n, blk)) {
        acc.Add(su, blk);
    }
    if (dev::JudgeT(m, fe, P, 1, src, n, blk))
        acc.Add(su, blk);
    const long eo = e - ebase;  // the entry's place in the output arrays
    double* out = mts + 9 * eo;
    for (int i = 0; i < 9; ++i)
        out[i] = acc.alive ? acc.v[i] : 0.0;
    unsigned char bits = acc.alive ? 1 : 0;
    dev::BlockAcc accg;
    if (dev::JudgeU(m, fe, P, src, n, blk))
        accg.Add(refs[f].ht_sign, blk);
    double* og = mgs + 6 * eo;
    const int pick[6] = {0, 1, 2, 4, 5, 8};
    for (int i = 0; i < 6; ++i)
        og[i] = accg.alive ? accg.v[pick[i]] : 0.0;
    bits |= accg.alive ? 2 : 0;
    alive[eo] = bits;
}

void LaunchHistory(const ElemGeo* geo, const ElemPoints* pts, const int* dom, const NodeRefs* refs,
                   const double* selfT2, PairGrid grid, const long* off, const int* kb, long npairs, long e0, long e1,
                   double* mts, double* mgs, unsigned char* alive, hipStream_t s, long ebase, int maxLag)
{
    long n = e1 - e0;
    if (n <= 0)
        return;
    k_history<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(geo, pts, dom, refs, selfT2, grid, off, kb, npairs, e0, e1,
                                                          ebase, maxLag, mts, mgs, alive);
}

// Near blocks of a slab: one thread per pair of [p0, p1) with a near part
// (nearOff[p + 1] > nearOff[p]); its first nearOff[p + 1] - nearOff[p] lags.
__global__ void __launch_bounds__(256) k_near_gather(const long* __restrict__ off, const long* __restrict__ nearOff,
                                                     long p0, long p1, long ebase, const double* __restrict__ mts,
                                                     const double* __restrict__ mgs,
                                                     const unsigned char* __restrict__ alive, double* __restrict__ nmts,
                                                     double* __restrict__ nmgs, unsigned char* __restrict__ nalive)
{
    const long p = p0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= p1)
        return;
    const long n0 = nearOff[p], cnt = nearOff[p + 1] - n0;
    const long e0 = off[p] - ebase;
    for (long i = 0; i < cnt; ++i) {
        for (int v = 0; v < 9; ++v)
            nmts[9 * (n0 + i) + v] = mts[9 * (e0 + i) + v];
        for (int v = 0; v < 6; ++v)
            nmgs[6 * (n0 + i) + v] = mgs[6 * (e0 + i) + v];
        nalive[n0 + i] = alive[e0 + i];
    }
}

void LaunchNearGather(const long* off, const int* kb, const long* nearOff, long p0, long p1, long ebase,
                      const double* mts, const double* mgs, const unsigned char* alive, double* nmts, double* nmgs,
                      unsigned char* nalive, hipStream_t s)
{
    (void)kb;
    const long n = p1 - p0;
    if (n <= 0)
        return;
    k_near_gather<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(off, nearOff, p0, p1, ebase, mts, mgs, alive, nmts, nmgs,
                                                            nalive);
}

// ---------------------------------------------------------------------------
// Staged history quadrature: one wavefront per (pair, chunk of up to 64 lags),
// one lag per lane.  The lanes of a wave share the field element and the
// source, so everything about a quadrature point that does not depend on the
// time level -- the distance powers and ratios R/c, R/(c dt), the geometric
// tensors D, G, E, H, F of GetDynaT1ij / GetDynaT2ij and the P-front,
// S-front and between-fronts tensors of GetDynaUij -- is computed once per
// point by one lane and staged in LDS.  Each lane then forms its own time
// coefficients (front flags, ISeries / JSeries window, the n-dependent
// factors) and adds term by term in the reference's order: per point
// T = 0 + P term + S term + window term, and each accumulator adds the points
// patch by patch in ascending order.  Every staged value is the same
// expression of the same inputs as in the per-point kernels, so the blocks
// equal the lane-per-entry kernel's bit for bit.  The regular path's
// T1(n) + T2(n) (quirk Q2) is formed once and added for each judge that
// selected it.
// ---------------------------------------------------------------------------
namespace dev {

// Staged (source, point) record of the wave-per-pair quadrature.  Every
// kernel tensor at one point is a combination of three basis tensors,
//   X_ab = r_a r_b,  Q_ab = r_b n_a,  B_ab = r_a n_b + delta_ab (r . n),
//   Dij = sXd X + sQd Q + sBd B,   Eij = sXe X + sQe Q + sBe B,
//   Fij = sXf X + sQBf (Q + B),    Gij = sXg X + sQg Q,   Hij = sXh X + sBh B,
// (DBEM1/dsquareelement.h:631-903), and every U term is a combination of X
// and delta_ab.  The record keeps the basis and those scalars (times the
// point weight, with 1/(4 pi dt) folded into G and H and 1/(4 pi rho) into
// U), so a lane's contribution at a point is cX X + cQ Q + cB B with three
// scalars built from its time coefficients.
constexpr int kRec = 42;  // doubles per staged point (16-byte multiple)
enum RecField : int {
    qRc1 = 0,   // R / C1
    qRc2 = 1,   // R / C2
    qPc1 = 2,   // R / (C1 dt) / (4 pi)
    qPc2 = 3,   // R / (C2 dt) / (4 pi)
    qIR2h = 4,  // 1 / (2 R^2)
    qIR3t = 5,  // 1 / (3 R^3)
    qRk = 6,    // R / dt / (4 pi)
    qSXd = 7,
    qSQd = 8,
    qSBd = 9,
    qSXe = 10,
    qSQe = 11,
    qSBe = 12,
    qSXf = 13,
    qSQBf = 14,
    qSXg = 15,
    qSQg = 16,
    qSXh = 17,
    qSBh = 18,
    qUP = 19,   // P-front U = uP X
    qUS1 = 20,  // S-front U = uS1 X + uS0 delta
    qUS0 = 21,
    qUC1 = 22,  // between-fronts U = ri (uC1 X + uC0 delta)
    qUC0 = 23,
    qX = 24,    // X: 00 01 02 11 12 22
    qQ = 30,    // Q: 9 entries, row-major
    qBd = 39,   // diagonal of B: Q_aa + r . n
};

__device__ void StageRecord(const DynaMat& m, const double* src, const double* pos, const double* nor, double w,
                            double* g)
{
    double r[3] = {pos[0] - src[0], pos[1] - src[1], pos[2] - src[2]};
    double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double iR = 1.0 / R;
    double iR2 = iR * iR, iR3 = iR2 * iR, iR4 = iR2 * iR2, iR5 = iR4 * iR;
    double rn = r[0] * nor[0] + r[1] * nor[1] + r[2] * nor[2];
    g[qRc1] = R / m.C1;
    g[qRc2] = R / m.C2;
    g[qPc1] = kOneOver4Pi * (R / m.C1Dt);
    g[qPc2] = kOneOver4Pi * (R / m.C2Dt);
    g[qIR2h] = 0.5 * iR2;
    g[qIR3t] = iR3 * (1.0 / 3.0);
    g[qRk] = kOneOver4Pi * (R / m.Dt);
    double c = m.C2_2OverC1_2;
    double tg = kOneOver4Pi / m.Dt * w;
    g[qSXd] = -(c * 12.0 * rn * iR5) * w;
    g[qSQd] = (4.0 * c - 1.0) * iR3 * w;
    g[qSBd] = c * 2.0 * iR3 * w;
    g[qSXe] = 12.0 * rn * iR5 * w;
    g[qSQe] = -2.0 * iR3 * w;
    g[qSBe] = -3.0 * iR3 * w;
    g[qSXf] = -30.0 * m.C2_2 * rn * iR5 * w;
    g[qSQBf] = 6.0 * m.C2_2 * iR3 * w;
    g[qSXg] = -2.0 * m.C2_2OverC1_3 * rn * iR4 * tg;
    g[qSQg] = (2.0 * c - 1.0) / m.C1 * iR2 * tg;
    g[qSXh] = 2.0 / m.C2 * rn * iR4 * tg;
    g[qSBh] = -(iR2 / m.C2) * tg;
    double tu = m.OneOverFourPaiRou * w;
    g[qUP] = m.OneOverC1_2 * iR3 * tu;
    g[qUS1] = -(m.OneOverC2_2 * iR3) * tu;
    g[qUS0] = m.OneOverC2_2 * iR * tu;
    g[qUC1] = 3.0 * iR3 * tu;
    g[qUC0] = -iR * tu;
    g[qX + 0] = r[0] * r[0];
    g[qX + 1] = r[0] * r[1];
    g[qX + 2] = r[0] * r[2];
    g[qX + 3] = r[1] * r[1];
    g[qX + 4] = r[1] * r[2];
    g[qX + 5] = r[2] * r[2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            g[qQ + 3 * a + b] = r[b] * nor[a];
    for (int a = 0; a < 3; ++a)
        g[qBd + a] = r[a] * nor[a] + rn;
}

// One lane's time span: GetDynaT1ij at level n (type 1) or GetDynaT2ij at
// level n - 1 (type 2); nfac = (n + 1) / (4 pi) or (1 - n) / (4 pi).
struct Span {
    double ndt, np1, nfac;
};

__device__ __forceinline__ Span MakeSpan(const DynaMat& m, int type, double n)
{
    Span s;
    double nn = type == 1 ? n : n - 1;
    s.ndt = nn * m.Dt;
    s.np1 = s.ndt + m.Dt;
    s.nfac = type == 1 ? kOneOver4Pi * (n + 1.0) : kOneOver4Pi * (1.0 - n);
    return s;
}

// Basis scalars of one span at one staged point (all zero when !on).  The
// front flags are the reference's (Fai); the ISeries / JSeries window
// integrals are evaluated in their common closed form
//   ri = (hi^2 - lo^2) / (2 R^2),  rj = (hi^3 - lo^3) / (3 R^3),
//   lo = max(ndt, R / C1), hi = min(np1, R / C2), zero when hi <= lo,
// which is each of the four window cases of DBEM1/dsquareelement.h:662-716
// up to rounding.
struct SpanMix {
    double cX, cQ, cB, ri;
    bool P, S;
};

template <int TYPE>
__device__ __forceinline__ SpanMix MixAt(const double* g, const Span& s, bool on)
{
    SpanMix o;
    double rc1 = g[qRc1], rc2 = g[qRc2];
    o.P = on && rc1 >= s.ndt && rc1 < s.np1;
    o.S = on && rc2 >= s.ndt && rc2 < s.np1;
    double lo = fmax(s.ndt, rc1), hi = fmin(s.np1, rc2);
    double d = on ? fmax(hi - lo, 0.0) : 0.0;
    double sm = hi + lo;
    double ri = d * sm * g[qIR2h];
    double rj = d * fma(hi, sm, lo * lo) * g[qIR3t];
    double rq = g[qRk] * rj;
    double gam = TYPE == 1 ? fma(s.nfac, ri, -rq) : fma(s.nfac, ri, rq);
    double aP = o.P ? (TYPE == 1 ? s.nfac - g[qPc1] : s.nfac + g[qPc1]) : 0.0;
    double aS = o.S ? (TYPE == 1 ? s.nfac - g[qPc2] : s.nfac + g[qPc2]) : 0.0;
    double bP = o.P ? (TYPE == 1 ? 1.0 : -1.0) : 0.0;
    double bS = o.S ? (TYPE == 1 ? 1.0 : -1.0) : 0.0;
    o.cX = fma(aP, g[qSXd], fma(aS, g[qSXe], fma(bP, g[qSXg], fma(bS, g[qSXh], gam * g[qSXf]))));
    o.cQ = fma(aP, g[qSQd], fma(aS, g[qSQe], fma(bP, g[qSQg], gam * g[qSQBf])));
    o.cB = fma(aP, g[qSBd], fma(aS, g[qSBe], fma(bS, g[qSBh], gam * g[qSQBf])));
    o.ri = ri;
    return o;
}

// acc += cX X + cQ Q + cB B (3x3, row-major)
__device__ __forceinline__ void AddBasis(const double* g, double cX, double cQ, double cB, double* acc)
{
    const int xm[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            int k = 3 * a + b;
            double bab = a == b ? g[qBd + a] : g[qQ + 3 * b + a];
            acc[k] = fma(cX, g[qX + xm[k]], fma(cQ, g[qQ + k], fma(cB, bab, acc[k])));
        }
}

// accU (00 01 02 11 12 22) += U with the type-1 flags and window integral.
__device__ __forceinline__ void AddU(const double* g, bool P, bool S, double ri, double* accU)
{
    double uX = fma(ri, g[qUC1], (P ? g[qUP] : 0.0) + (S ? g[qUS1] : 0.0));
    double uD = fma(ri, g[qUC0], S ? g[qUS0] : 0.0);
    accU[0] = fma(uX, g[qX + 0], accU[0] + uD);
    accU[1] = fma(uX, g[qX + 1], accU[1]);
    accU[2] = fma(uX, g[qX + 2], accU[2]);
    accU[3] = fma(uX, g[qX + 3], accU[3] + uD);
    accU[4] = fma(uX, g[qX + 4], accU[4]);
    accU[5] = fma(uX, g[qX + 5], accU[5] + uD);
}

__device__ __forceinline__ bool WaveAny(bool v) { return __ballot(v) != 0; }

__device__ __forceinline__ long long WaveSumLL(long long v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Evaluation counts of the staged quadrature (qcount non-null): [0] patch-wise
// T2 (entry, point) evaluations, [1] patch-wise T1 + U ones, [2] regular
// (entry, point) ones, [3] staged (source, point) records; one atomic per
// counter and wavefront.
__device__ __forceinline__ void CountEvaluations(unsigned long long* qcount, long long t2, long long t1, long long rg,
                                                 long long records, int lane)
{
    t2 = WaveSumLL(t2);
    t1 = WaveSumLL(t1);
    rg = WaveSumLL(rg);
    if (lane == 0) {
        atomicAdd(qcount + 0, (unsigned long long)t2);
        atomicAdd(qcount + 1, (unsigned long long)t1);
        atomicAdd(qcount + 2, (unsigned long long)rg);
        atomicAdd(qcount + 3, (unsigned long long)records);
    }
}

}  // namespace dev

namespace dev {

// NeedCal with the centre distance r already formed (same operations).
__device__ __forceinline__ int NeedCalR(const DynaMat& m, double radius, double r, int type, double n)
{
    double nn = (type == 2) ? n - 1 : n;
    double ndt = nn * m.Dt;
    double np1dt = (nn + 1) * m.Dt;
    double Rmax = r + radius;
    double Rmin = r - radius;
    if (Rmin < 0.0)
        Rmin = 0.0;
    double a1 = Rmax / m.C1, a2 = Rmax / m.C2, b1 = Rmin / m.C1, b2 = Rmin / m.C2;
    if (ndt >= a2 || np1dt <= b1)
        return 0;
    if (ndt <= b1 && (np1dt >= a1 && np1dt <= b2))
        return 2;
    if (ndt <= b1 && np1dt >= a2)
        return 2;
    if (ndt >= a1 && np1dt <= b2)
        return 2;
    if ((ndt >= a1 && ndt <= b2) && np1dt >= a2)
        return 2;
    return 1;
}

// Judges of one band entry: f2 (T2 at level k - 1, unless the self table
// replaces it), f1 (T1 and U at level k).
struct EntryFlags {
    bool self1;
    int f2, f1;
    __device__ bool patchwise() const { return f2 == 1 || f1 == 1; }
};

__device__ __forceinline__ EntryFlags Judge(const DynaMat& m, double radius, double r, int k, bool same)
{
    EntryFlags e;
    double n = (double)k;
    e.self1 = k <= 1 && same;
    e.f2 = e.self1 ? 0 : NeedCalR(m, radius, r, 2, n);
    e.f1 = NeedCalR(m, radius, r, 1, n);
    return e;
}

// Judges of a band entry, or the frozen ones (f2 | f1 << 2) when cls is set.
__device__ __forceinline__ EntryFlags JudgeAt(const DynaMat& m, double radius, double r, int k, bool same,
                                              const unsigned char* cls, long e)
{
    if (!cls)
        return Judge(m, radius, r, k, same);
    EntryFlags f;
    f.self1 = k <= 1 && same;
    unsigned char c = cls[e];
    f.f2 = c & 3;
    f.f1 = (c >> 2) & 3;
    return f;
}

// Regular-path contribution of one staged point: the T1 (level n) + T2
// (level n - 1) sum and the U term, into accR / accRU.
__device__ __forceinline__ void ConsumeRegular(const double* g, const Span& s1, const Span& s2, bool rt, double* accR,
                                               double* accRU)
{
    SpanMix a = MixAt<1>(g, s1, rt);
    SpanMix b = MixAt<2>(g, s2, rt);
    AddBasis(g, a.cX + b.cX, a.cQ + b.cQ, a.cB + b.cB, accR);
    AddU(g, a.P, a.S, a.ri, accRU);
}

// Patch-wise contribution of one staged point (patch flags u2 for T2, u1 for
// T1 and U).
__device__ __forceinline__ void ConsumePatchwise(const double* g, const Span& s2, const Span& s1, bool u2, bool u1,
                                                 double* accA, double* accB, double* accU)
{
    SpanMix a = MixAt<2>(g, s2, u2);
    SpanMix b = MixAt<1>(g, s1, u1);
    AddBasis(g, a.cX, a.cQ, a.cB, accA);
    AddBasis(g, b.cX, b.cQ, b.cB, accB);
    AddU(g, b.P, b.S, b.ri, accU);
}

// Split storage of the mixed mode: an FP64 value v is kept as hi = (float)v
// and, optionally, lo = bf16(v - hi) (round to nearest even), so hi + lo
// carries about 32 significant bits.
__device__ __forceinline__ unsigned short Bf16Residual(double v, double hi)
{
    unsigned int bits = __float_as_uint((float)(v - hi));
    bits += 0x7fffu + ((bits >> 16) & 1u);
    return (unsigned short)(bits >> 16);
}

__device__ __forceinline__ double Bf16ToDouble(unsigned short lo)
{
    return (double)__uint_as_float(((unsigned int)lo) << 16);
}

// MTS entry = AddBlock(T2 or the self table) then AddBlock(T1); MGS entry =
// AddBlock(U) (DBEM1/multidomain.cpp:2017-2078).
template <typename TB>
__device__ __forceinline__ void FinalizeEntry(const EntryFlags& fl, const double* selfT2f, double hu_sign,
                                              double ht_sign, const double* accA, const double* accB,
                                              const double* accU, long e, TB* mts, TB* mgs, unsigned short* mtsLo,
                                              unsigned short* mgsLo, unsigned char* alive)
{
    BlockAcc acc;
    if (fl.self1) {
        double blk[9];
        for (int i = 0; i < 9; ++i)
            blk[i] = selfT2f[i];
        acc.Add(hu_sign, blk);
    } else if (fl.f2) {
        acc.Add(hu_sign, accA);
    }
    if (fl.f1)
        acc.Add(hu_sign, accB);
    TB* out = mts + 9 * e;
    for (int i = 0; i < 9; ++i) {
        double v = acc.alive ? acc.v[i] : 0.0;
        out[i] = (TB)v;
        if (mtsLo)
            mtsLo[9 * e + i] = Bf16Residual(v, (double)(TB)v);
    }
    unsigned char bits = acc.alive ? 1 : 0;
    BlockAcc accg;
    if (fl.f1) {
        double u9[9] = {accU[0], accU[1], accU[2], accU[1], accU[3], accU[4], accU[2], accU[4], accU[5]};
        accg.Add(ht_sign, u9);
    }
    TB* og = mgs + 6 * e;
    const int pick[6] = {0, 1, 2, 4, 5, 8};
    for (int i = 0; i < 6; ++i) {
        double v = accg.alive ? accg.v[pick[i]] : 0.0;
        og[i] = (TB)v;
        if (mgsLo)
            mgsLo[6 * e + i] = Bf16Residual(v, (double)(TB)v);
    }
    bits |= accg.alive ? 2 : 0;
    alive[e] = bits;
}

}  // namespace dev

// Number of patch-wise entries (f2 == 1 or f1 == 1) in each pair's band.
__global__ void __launch_bounds__(256) k_pwcount(const ElemGeo* __restrict__ geo, PairGrid grid,
                                                 const int* __restrict__ kb, const int* __restrict__ len,
                                                 const long* __restrict__ off, const unsigned char* __restrict__ cls,
                                                 long npairs, int* __restrict__ cnt)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs)
        return;
    int s, f, d;
    dev::DecodePair(grid, p, s, f, d);
    const DynaMat& m = grid.mats[d];
    double r = Dist3(geo[f].center, geo[s].center);
    double radius = geo[f].radius;
    int c = 0, k0 = kb[p], L = len[p];
    long e0 = off[p];
    for (int t = 0; t < L; ++t)
        c += dev::JudgeAt(m, radius, r, k0 + t, s == f, cls, e0 + t).patchwise() ? 1 : 0;
    cnt[p] = c;
}

void LaunchPwCount(const ElemGeo* geo, PairGrid grid, const int* kb, const int* len, const long* off,
                   const unsigned char* cls, long npairs, int* cnt, hipStream_t s)
{
    if (npairs <= 0)
        return;
    k_pwcount<<<(unsigned)((npairs + 255) / 256), 256, 0, s>>>(geo, grid, kb, len, off, cls, npairs, cnt);
}

// Greedy packing of one field element's patch-wise lags into PwWaves: its
// sources in order, up to kPwSlots pairs and 64 lanes per wave, a pair that
// does not fit continuing in the next wave.  One thread per field element;
// the count pass (waves == nullptr) writes nw[f], the write pass fills the
// waves from woff[f] on.
__global__ void __launch_bounds__(256) k_pwpack(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                const int* __restrict__ cnt, int ne, const long* __restrict__ woff,
                                                long* __restrict__ nw, PwWave* __restrict__ waves, long rowBegin,
                                                long rowEnd)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= ne)
        return;
    int d = refs[f].domain;
    long nd = grid.eleCount[d];
    long fl = f - grid.eleBegin[d];
    PwWave cur{};
    int lanes = 0;
    long w = waves ? woff[f] : 0, count = 0;
    // the sources of the domain inside [rowBegin, rowEnd)
    const long sLo = rowBegin > grid.eleBegin[d] ? rowBegin - grid.eleBegin[d] : 0;
    const long sHi = rowEnd - grid.eleBegin[d] < nd ? rowEnd - grid.eleBegin[d] : nd;
    for (long sl = sLo; sl < sHi; ++sl) {
        long p = grid.pairBase[d] + sl * nd + fl;
        int c = cnt[p], r0 = 0;
        while (c > 0) {
            if (cur.nseg == kPwSlots || lanes == 64) {
                if (waves)
                    waves[w + count] = cur;
                ++count;
                cur = PwWave{};
                lanes = 0;
            }
            int take = c < 64 - lanes ? c : 64 - lanes;
            cur.pair[cur.nseg] = p;
            cur.rank0[cur.nseg] = r0;
            cur.cnt[cur.nseg] = take;
            cur.nseg++;
            lanes += take;
            c -= take;
            r0 += take;
        }
    }
    if (cur.nseg > 0) {
        if (waves)
            waves[w + count] = cur;
        ++count;
    }
    if (!waves)
        nw[f] = count;
}

void LaunchPwPack(PairGrid grid, const NodeRefs* refs, const int* cnt, int ne, const long* woff, long* nw,
                  PwWave* waves, hipStream_t s, long rowBegin, long rowEnd)
{
    if (rowEnd < 0)
        rowEnd = ne;
    if (ne > 0)
        k_pwpack<<<(ne + 255) / 256, 256, 0, s>>>(grid, refs, cnt, ne, woff, nw, waves, rowBegin, rowEnd);
}

// Judges of every band entry (f2 | f1 << 2) with the current materials.
__global__ void __launch_bounds__(256) k_classify(const ElemGeo* __restrict__ geo, PairGrid grid,
                                                  const int* __restrict__ kb, const int* __restrict__ len,
                                                  const long* __restrict__ off, long npairs,
                                                  unsigned char* __restrict__ cls)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs)
        return;
    int s, f, d;
    dev::DecodePair(grid, p, s, f, d);
    const DynaMat& m = grid.mats[d];
    double r = Dist3(geo[f].center, geo[s].center);
    double radius = geo[f].radius;
    int k0 = kb[p], L = len[p];
    long e0 = off[p];
    for (int t = 0; t < L; ++t) {
        dev::EntryFlags e = dev::Judge(m, radius, r, k0 + t, s == f);
        cls[e0 + t] = (unsigned char)(e.f2 | (e.f1 << 2));
    }
}

void LaunchClassify(const ElemGeo* geo, PairGrid grid, const int* kb, const int* len, const long* off, long npairs,
                    unsigned char* cls, hipStream_t s)
{
    if (npairs > 0)
        k_classify<<<(unsigned)((npairs + 255) / 256), 256, 0, s>>>(geo, grid, kb, len, off, npairs, cls);
}

// Regular-path entries: one wavefront per (pair, chunk of up to 64 lags), one
// lag per lane; the lanes of patch-wise entries are left to k_history_pw.
// The 16 regular points are staged once per wave.
template <typename TB>
__global__ void __launch_bounds__(64) k_history_reg(
    const ElemGeo* __restrict__ geo, const ElemPoints* __restrict__ pts, const NodeRefs* __restrict__ refs,
    const double* __restrict__ selfT2, PairGrid grid, const long* __restrict__ coff, const long* __restrict__ off,
    const int* __restrict__ kb, const int* __restrict__ len, const unsigned char* __restrict__ cls, long npairs,
    long c0, long c1, TB* __restrict__ mts, TB* __restrict__ mgs, unsigned short* __restrict__ mtsLo,
    unsigned short* __restrict__ mgsLo, unsigned char* __restrict__ alive, long ebase, int maxLag,
    unsigned long long* __restrict__ qcount)
{
    __shared__ double sgeo[kNG2 * dev::kRec];
    using namespace dev;
    long wv = c0 + (long)blockIdx.x;
    if (wv >= c1)
        return;
    int lane = threadIdx.x;
    long lo = 0, hi = npairs;
    while (hi - lo > 1) {
        long mid = (lo + hi) >> 1;
        if (coff[mid] <= wv)
            lo = mid;
        else
            hi = mid;
    }
    long p = lo;
    int s, f, d;
    DecodePair(grid, p, s, f, d);
    int t = (int)(wv - coff[p]) * 64 + lane;
    if (kb[p] + (int)(wv - coff[p]) * 64 > maxLag)
        return;  // every lag of the chunk is past the lags asked for (wave-uniform)
    const DynaMat& m = grid.mats[d];
    const ElemGeo& fe = geo[f];
    const ElemPoints& P = pts[f];
    const double* src = geo[s].center;
    int k = kb[p] + t;
    double n = (double)k;
    bool inBand = t < len[p];
    EntryFlags fl = JudgeAt(m, fe.radius, Dist3(fe.center, src), k, s == f, inBand ? cls : nullptr, off[p] + t);
    bool mine = inBand && !fl.patchwise();
    // a regular entry's T2 and T1 parts are the same T1 + T2 sum, so one
    // accumulator serves both
    double accR[9], accRU[6];
    for (int i = 0; i < 9; ++i)
        accR[i] = 0.0;
    for (int i = 0; i < 6; ++i)
        accRU[i] = 0.0;
    bool rt = mine && (fl.f2 == 2 || fl.f1 == 2);
    if (qcount) {
        const bool any = WaveAny(rt);
        CountEvaluations(qcount, 0, 0, rt ? kNG2 : 0, any ? kNG2 : 0, lane);
    }
    if (WaveAny(rt)) {
        if (lane < kNG2)
            StageRecord(m, src, P.rg_pos[lane], P.rg_nor[lane], P.rg_w[lane], sgeo + lane * kRec);
        __syncthreads();
        Span s1 = MakeSpan(m, 1, n), s2 = MakeSpan(m, 2, n);
#pragma unroll 1
        for (int i = 0; i < kNG2; ++i)
            ConsumeRegular(sgeo + i * kRec, s1, s2, rt, accR, accRU);
    }
    if (!mine)
        return;
    FinalizeEntry(fl, selfT2 + 9 * (long)f, refs[f].hu_sign, refs[f].ht_sign, accR, accR, accRU, off[p] + t - ebase,
                  mts, mgs, mtsLo, mgsLo, alive);
}

// Patch-wise entries: one wavefront per PwWave, whose lanes are the
// patch-wise lags of up to kPwSlots pairs sharing one field element (so the
// patch points are shared); the point terms of every (source, point) are
// staged once per patch.  Regular parts of these entries (the other judge
// picking the regular path) are added here too, so each entry is finished by
// exactly one kernel.
template <typename TB>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_history_pw(
    const ElemGeo* __restrict__ geo, const ElemPoints* __restrict__ pts, const NodeRefs* __restrict__ refs,
    const double* __restrict__ selfT2, PairGrid grid, const PwWave* __restrict__ waves, long w0, long w1,
    const long* __restrict__ off, const int* __restrict__ kb, const int* __restrict__ len,
    const unsigned char* __restrict__ cls, TB* __restrict__ mts, TB* __restrict__ mgs,
    unsigned short* __restrict__ mtsLo, unsigned short* __restrict__ mgsLo, unsigned char* __restrict__ alive,
    long ebase, int maxLag, unsigned long long* __restrict__ qcount)
{
    // the patch radii are read only by the patch-mask pass, before the first
    // point record is staged, so they share the record buffer (12.1 KB of LDS
    // per wavefront: twelve wavefronts, three per SIMD, fit the 160 KB of a CU)
    static_assert(kPwSlots * 2 * kNPW2 <= kPwSlots * kNPWG2 * dev::kRec, "radii inside the record buffer");
    __shared__ double sgeo[kPwSlots * kNPWG2 * dev::kRec];
    double* const spr = sgeo;
    __shared__ int ssrc[kPwSlots];
    using namespace dev;
    long wi = w0 + (long)blockIdx.x;
    if (wi >= w1)
        return;
    int lane = threadIdx.x;
    const PwWave& W = waves[wi];
    int nseg = W.nseg;
    {
        // the wave's pairs all start past the lags asked for (wave-uniform)
        int first = 1 << 30;
        for (int i = 0; i < nseg; ++i)
            first = kb[W.pair[i]] < first ? kb[W.pair[i]] : first;
        if (first > maxLag)
            return;
    }
    // field element and domain (the same for every slot)
    int s0, f, d;
    DecodePair(grid, W.pair[0], s0, f, d);
    const DynaMat& m = grid.mats[d];
    const ElemGeo& fe = geo[f];
    const ElemPoints& P = pts[f];
    if (lane < nseg) {
        int ss, ff, dd;
        DecodePair(grid, W.pair[lane], ss, ff, dd);
        ssrc[lane] = ss;
    }
    // lane -> (slot, rank among the pair's patch-wise lags)
    int seg = -1, rank = 0, base = 0;
    for (int i = 0; i < kPwSlots; ++i) {
        int c = i < nseg ? W.cnt[i] : 0;
        if (seg < 0 && lane >= base && lane < base + c) {
            seg = i;
            rank = W.rank0[i] + (lane - base);
        }
        base += c;
    }
    __syncthreads();
    // patch radii of every slot's source
    for (int idx = lane; idx < nseg * kNPW2; idx += 64) {
        int slot = idx / kNPW2, i = idx - slot * kNPW2;
        double r = Dist3(fe.rcenpw[i], geo[ssrc[slot]].center);
        spr[slot * 2 * kNPW2 + 2 * i] = r + fe.radiusPW;
        spr[slot * 2 * kNPW2 + 2 * i + 1] = r - fe.radiusPW;
    }
    bool active = seg >= 0;
    int myseg = active ? seg : 0;
    long p = W.pair[myseg];
    int s = ssrc[myseg];
    const double* src = geo[s].center;
    double r0 = Dist3(fe.center, src);
    int t = -1, k = 0;
    EntryFlags fl{false, 0, 0};
    if (active) {
        int k0 = kb[p], L = len[p], c = 0;
        long eb0 = off[p];
        for (int tt = 0; tt < L; ++tt) {
            EntryFlags e = JudgeAt(m, fe.radius, r0, k0 + tt, s == f, cls, eb0 + tt);
            if (e.patchwise()) {
                if (c == rank) {
                    t = tt;
                    fl = e;
                    break;
                }
                ++c;
            }
        }
        k = k0 + t;
        if (t < 0)
            active = false;
    }
    double n = (double)k;
    __syncthreads();
    unsigned long long m2 = 0, m1 = 0;
    {
        double nn2 = n - 1, nn1 = n;
        double lo2 = nn2 * m.C2Dt, hi2 = (nn2 + 1) * m.C1Dt;
        double lo1 = nn1 * m.C2Dt, hi1 = (nn1 + 1) * m.C1Dt;
        bool pw2 = active && fl.f2 == 1, pw1 = active && fl.f1 == 1;
        const double* sp = spr + myseg * 2 * kNPW2;
        for (int i = 0; i < kNPW2; ++i) {
            double rpr = sp[2 * i], rmr = sp[2 * i + 1];
            if (pw2 && !(rmr >= hi2 || rpr < lo2))
                m2 |= 1ull << i;
            if (pw1 && !(rmr >= hi1 || rpr < lo1))
                m1 |= 1ull << i;
        }
    }
    unsigned long long need = 0;
    for (int i = 0; i < kNPW2; ++i)
        if (WaveAny(((m2 | m1) >> i) & 1ull))
            need |= 1ull << i;
    // regular parts first (the other judge picking the regular path; at most
    // one of r2 / r1 holds for a patch-wise entry), 8 points per half
    bool r2 = active && fl.f2 == 2, r1 = active && fl.f1 == 2, rt = r2 || r1;
    if (qcount) {
        const bool any = WaveAny(rt);
        const long long rec = (long long)nseg * (kNPWG2 * __popcll(need) + (any ? kNG2 : 0));
        CountEvaluations(qcount, kNPWG2 * __popcll(m2), kNPWG2 * __popcll(m1), rt ? kNG2 : 0, rec, lane);
    }
    Span s1 = MakeSpan(m, 1, n), s2 = MakeSpan(m, 2, n);
    double accA[9], accB[9], accU[6];
    {
        double accR[9], accRU[6];
        for (int i = 0; i < 9; ++i)
            accR[i] = 0.0;
        for (int i = 0; i < 6; ++i)
            accRU[i] = 0.0;
        if (WaveAny(rt)) {
            int rgSlot = lane >> 3, rgJ = lane & 7;
            for (int half = 0; half < 2; ++half) {
                if (rgSlot < nseg) {
                    int q = half * 8 + rgJ;
                    StageRecord(m, geo[ssrc[rgSlot]].center, P.rg_pos[q], P.rg_nor[q], P.rg_w[q], sgeo + lane * kRec);
                }
                __syncthreads();
                const double* gs = sgeo + myseg * 8 * kRec;
#pragma unroll 1
                for (int j = 0; j < 8; ++j)
                    ConsumeRegular(gs + j * kRec, s1, s2, rt, accR, accRU);
                __syncthreads();
            }
        }
        for (int i = 0; i < 9; ++i) {
            accA[i] = r2 ? accR[i] : 0.0;
            accB[i] = r1 ? accR[i] : 0.0;
        }
        for (int i = 0; i < 6; ++i)
            accU[i] = r1 ? accRU[i] : 0.0;
    }
    // one patch at a time, all slots' sources staged together
    int stSlot = lane / kNPWG2, stJ = lane - stSlot * kNPWG2;
    while (need) {
        int i = __builtin_ctzll(need);
        need &= need - 1;
        if (stSlot < nseg) {
            int q = i * kNPWG2 + stJ;
            StageRecord(m, geo[ssrc[stSlot]].center, P.pw_pos[q], P.pw_nor[q], P.pw_w[q], sgeo + lane * kRec);
        }
        __syncthreads();
        bool u2 = (m2 >> i) & 1ull, u1 = (m1 >> i) & 1ull;
        const double* gs = sgeo + myseg * kNPWG2 * kRec;
#pragma unroll 1
        for (int j = 0; j < kNPWG2; ++j)
            ConsumePatchwise(gs + j * kRec, s2, s1, u2, u1, accA, accB, accU);
        __syncthreads();
    }
    if (!active)
        return;
    FinalizeEntry(fl, selfT2 + 9 * (long)f, refs[f].hu_sign, refs[f].ht_sign, accA, accB, accU, off[p] + t - ebase,
                  mts, mgs, mtsLo, mgsLo, alive);
}

template <typename TB>
void LaunchHistoryReg(const ElemGeo* geo, const ElemPoints* pts, const NodeRefs* refs, const double* selfT2,
                      PairGrid grid, const long* coff, const long* off, const int* kb, const int* len,
                      const unsigned char* cls, long npairs, long c0, long c1, TB* mts, TB* mgs,
                      unsigned short* mtsLo, unsigned short* mgsLo, unsigned char* alive, hipStream_t s, long ebase,
                      int maxLag, unsigned long long* qcount)
{
    long waves = c1 - c0;
    if (waves <= 0)
        return;
    k_history_reg<TB><<<(unsigned)waves, 64, 0, s>>>(geo, pts, refs, selfT2, grid, coff, off, kb, len, cls, npairs,
                                                     c0, c1, mts, mgs, mtsLo, mgsLo, alive, ebase, maxLag, qcount);
}

template <typename TB>
void LaunchHistoryPw(const ElemGeo* geo, const ElemPoints* pts, const NodeRefs* refs, const double* selfT2,
                     PairGrid grid, const PwWave* waves, long w0, long w1, const long* off, const int* kb,
                     const int* len, const unsigned char* cls, TB* mts, TB* mgs, unsigned short* mtsLo,
                     unsigned short* mgsLo, unsigned char* alive, hipStream_t s, long ebase, int maxLag,
                     unsigned long long* qcount)
{
    long n = w1 - w0;
    if (n <= 0)
        return;
    k_history_pw<TB><<<(unsigned)n, 64, 0, s>>>(geo, pts, refs, selfT2, grid, waves, w0, w1, off, kb, len, cls, mts,
                                                mgs, mtsLo, mgsLo, alive, ebase, maxLag, qcount);
}

#define TDBEM_HIST_INSTANTIATE(TB)                                                                                   \
    template void LaunchHistoryReg<TB>(const ElemGeo*, const ElemPoints*, const NodeRefs*, const double*, PairGrid,   \
                                       const long*, const long*, const int*, const int*, const unsigned char*, long,   \
                                       long, long, TB*, TB*, unsigned short*, unsigned short*, unsigned char*,         \
                                       hipStream_t, long, int, unsigned long long*);                                   \
    template void LaunchHistoryPw<TB>(const ElemGeo*, const ElemPoints*, const NodeRefs*, const double*, PairGrid,    \
                                      const PwWave*, long, long, const long*, const int*, const int*,                  \
                                      const unsigned char*, TB*, TB*, unsigned short*, unsigned short*,                \
                                      unsigned char*, hipStream_t, long, int, unsigned long long*);
TDBEM_HIST_INSTANTIATE(double)
TDBEM_HIST_INSTANTIATE(float)
#undef TDBEM_HIST_INSTANTIATE

// ---------------------------------------------------------------------------
// Material derivatives of the history blocks, contracted on the fly.
// For the adjoint gradient only sum_e (dMGS_e . CT_e - dMTS_e : CU_e) is
// needed (CU, CT: the correlation of the adjoint with the entry's filtered
// states, k_corr_store).  A block is sum over points of cX X + cQ Q + cB B
// (and uX X + uD I for U), X, Q, B material-free, so its derivative
// contracted with CU is sum over points of dcX (r.CU^T r) + dcQ (n.CU^T r) +
// dcB (r.CU^T n + (r.n) tr CU): three scalars per point and lane.  The
// derivatives are those of the closed forms of the staged quadrature with the
// band structure, the judges, the patch masks and every point's front / window
// case held fixed (the piece of the piecewise-smooth block the current
// materials sit on), taken with respect to q = (C1, C2, rho) and mapped to
// (E, v, rho) per domain.
// ---------------------------------------------------------------------------
namespace dev {

constexpr int kRecD = 26;  // doubles per staged point of the derivative pass
// geometry of the point (as in the forward record) and the point scalars the
// material-dependent terms are built from: R/C1, R/C2, R/(4 pi C1 dt),
// R/(4 pi C2 dt) and b3 = w/R^3, b5 = w (r.n)/R^5, b4 = tg (r.n)/R^4,
// b2 = tg/R^2, u3 = tu/R^3, u1 = tu/R with tg = w/(4 pi dt), tu = w/(4 pi rho)
enum RecDGeo : int {
    gIR2h = 0, gIR3t = 1, gRk = 2, gR = 3, gN = 6, gRn = 9, gX = 10,
    gRc1 = 16, gRc2, gPc1, gPc2, gB3, gB5, gB4, gB2, gU3, gU1
};

// Material constants of the derivative pass (one domain): the tensor and U
// scalars are (constant) x (point scalar), and so are their derivatives with
// respect to q = (C1, C2, rho).
struct MatD {
    double iC1, iC2, iRho, c, dc1, dc2, C2sq, twoC2, kXg, kQg, kQg1, kQg2, iC1sq, iC2sq;
};

__device__ __forceinline__ MatD MakeMatD(const DynaMat& m)
{
    MatD k;
    k.iC1 = 1.0 / m.C1;
    k.iC2 = 1.0 / m.C2;
    k.iRho = 1.0 / m.Rou;
    k.c = m.C2_2OverC1_2;
    k.dc1 = -2.0 * k.c / m.C1;
    k.dc2 = 2.0 * k.c / m.C2;
    k.C2sq = m.C2_2;
    k.twoC2 = 2.0 * m.C2;
    k.kXg = -2.0 * m.C2_2OverC1_3;
    k.kQg = (2.0 * k.c - 1.0) / m.C1;
    double C1_2 = m.C1 * m.C1;
    k.kQg1 = -6.0 * k.c / C1_2 + 1.0 / C1_2;
    k.kQg2 = 4.0 * m.C2 / (C1_2 * m.C1);
    k.iC1sq = m.OneOverC1_2;
    k.iC2sq = m.OneOverC2_2;
    return k;
}

__device__ void StageRecordD(const DynaMat& m, const double* src, const double* pos, const double* nor, double w,
                             double* g)
{
    double r[3] = {pos[0] - src[0], pos[1] - src[1], pos[2] - src[2]};
    double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double iR = 1.0 / R;
    double iR2 = iR * iR, iR3 = iR2 * iR, iR4 = iR2 * iR2, iR5 = iR4 * iR;
    double rn = r[0] * nor[0] + r[1] * nor[1] + r[2] * nor[2];
    g[gIR2h] = 0.5 * iR2;
    g[gIR3t] = iR3 * (1.0 / 3.0);
    g[gRk] = kOneOver4Pi * (R / m.Dt);
    for (int a = 0; a < 3; ++a) {
        g[gR + a] = r[a];
        g[gN + a] = nor[a];
    }
    g[gRn] = rn;
    g[gX + 0] = r[0] * r[0];
    g[gX + 1] = r[0] * r[1];
    g[gX + 2] = r[0] * r[2];
    g[gX + 3] = r[1] * r[1];
    g[gX + 4] = r[1] * r[2];
    g[gX + 5] = r[2] * r[2];
    g[gRc1] = R / m.C1;
    g[gRc2] = R / m.C2;
    g[gPc1] = kOneOver4Pi * (R / m.C1Dt);
    g[gPc2] = kOneOver4Pi * (R / m.C2Dt);
    double tg = kOneOver4Pi / m.Dt * w;
    double tu = m.OneOverFourPaiRou * w;
    g[gB3] = iR3 * w;
    g[gB5] = rn * iR5 * w;
    g[gB4] = rn * iR4 * tg;
    g[gB2] = iR2 * tg;
    g[gU3] = iR3 * tu;
    g[gU1] = iR * tu;
}

// Derivatives (q = C1, C2, rho) of the basis scalars of one span at one
// staged point, the piece (front flags, window bounds) held at its current
// values; for TYPE 1 also those of the U scalars.  The tensor scalars do not
// depend on rho, so their q = 2 derivatives are zero; every term is formed
// from the point scalars and the material constants.
struct SpanDer {
    double dX[3], dQ[3], dB[3], duX[3], duD[3];
};

template <int TYPE>
__device__ __forceinline__ void MixDerAt(const double* g, const Span& s, bool on, const MatD& mc, SpanDer& o)
{
    double rc1 = g[gRc1], rc2 = g[gRc2];
    bool P = on && rc1 >= s.ndt && rc1 < s.np1;
    bool S = on && rc2 >= s.ndt && rc2 < s.np1;
    double lo = fmax(s.ndt, rc1), hi = fmin(s.np1, rc2);
    bool win = on && hi - lo > 0.0;
    double d = win ? hi - lo : 0.0;
    double sm = hi + lo;
    double ri = d * sm * g[gIR2h];
    double phi = fma(hi, sm, lo * lo);
    double rj = d * phi * g[gIR3t];
    double rk = g[gRk];
    double gam = TYPE == 1 ? fma(s.nfac, ri, -(rk * rj)) : fma(s.nfac, ri, rk * rj);
    double pc1 = g[gPc1], pc2 = g[gPc2];
    double aP = P ? (TYPE == 1 ? s.nfac - pc1 : s.nfac + pc1) : 0.0;
    double bP = P ? (TYPE == 1 ? 1.0 : -1.0) : 0.0;
    double bS = S ? (TYPE == 1 ? 1.0 : -1.0) : 0.0;
    bool loR = rc1 > s.ndt, hiR = rc2 < s.np1;
    double b3 = g[gB3], b5 = g[gB5], b4 = g[gB4], b2 = g[gB2];
    // values needed with the span derivatives
    double sXd = -12.0 * mc.c * b5, sQd = (4.0 * mc.c - 1.0) * b3, sBd = 2.0 * mc.c * b3;
    double sXe = 12.0 * b5, sQe = -2.0 * b3, sBe = -3.0 * b3;
    double sXf = -30.0 * mc.C2sq * b5, sQBf = 6.0 * mc.C2sq * b3;
    double sXg = mc.kXg * b4, sXh = 2.0 * mc.iC2 * b4, sBh = -(mc.iC2 * b2);
    // d(lo), d(hi) and the window terms for q = C1 (0) and C2 (1)
    double dlo[2] = {loR ? -rc1 * mc.iC1 : 0.0, 0.0};
    double dhi[2] = {0.0, hiR ? -rc2 * mc.iC2 : 0.0};
    double dri[2], dgam[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        double dd = win ? dhi[q] - dlo[q] : 0.0;
        double dsm = dhi[q] + dlo[q];
        dri[q] = (dd * sm + d * dsm) * g[gIR2h];
        double dphi = dhi[q] * sm + hi * dsm + 2.0 * lo * dlo[q];
        double drj = (dd * phi + d * dphi) * g[gIR3t];
        dgam[q] = TYPE == 1 ? s.nfac * dri[q] - rk * drj : s.nfac * dri[q] + rk * drj;
    }
    // q = C1: aP and the c-dependent scalars move
    {
        double daP = P ? (TYPE == 1 ? pc1 * mc.iC1 : -pc1 * mc.iC1) : 0.0;
        double dsXd = -12.0 * mc.dc1 * b5, dsQd = 4.0 * mc.dc1 * b3, dsBd = 2.0 * mc.dc1 * b3;
        double dsXg = -3.0 * sXg * mc.iC1, dsQg = mc.kQg1 * b2;
        o.dX[0] = daP * sXd + aP * dsXd + bP * dsXg + dgam[0] * sXf;
        o.dQ[0] = daP * sQd + aP * dsQd + bP * dsQg + dgam[0] * sQBf;
        o.dB[0] = daP * sBd + aP * dsBd + dgam[0] * sQBf;
    }
    // q = C2: aS, c and C2^2 move
    {
        double daS = S ? (TYPE == 1 ? pc2 * mc.iC2 : -pc2 * mc.iC2) : 0.0;
        double dsXd = -12.0 * mc.dc2 * b5, dsQd = 4.0 * mc.dc2 * b3, dsBd = 2.0 * mc.dc2 * b3;
        double dsXf = -30.0 * mc.twoC2 * b5, dsQBf = 6.0 * mc.twoC2 * b3;
        double dsXg = 2.0 * sXg * mc.iC2, dsQg = mc.kQg2 * b2;
        double dsXh = -sXh * mc.iC2, dsBh = -sBh * mc.iC2;
        o.dX[1] = aP * dsXd + daS * sXe + bP * dsXg + bS * dsXh + dgam[1] * sXf + gam * dsXf;
        o.dQ[1] = aP * dsQd + daS * sQe + bP * dsQg + dgam[1] * sQBf + gam * dsQBf;
        o.dB[1] = aP * dsBd + daS * sBe + bS * dsBh + dgam[1] * sQBf + gam * dsQBf;
    }
    o.dX[2] = o.dQ[2] = o.dB[2] = 0.0;
    if (TYPE == 1) {
        double u3 = g[gU3], u1 = g[gU1];
        double uC1 = 3.0 * u3, uC0 = -u1;
        double uP = mc.iC1sq * u3, uS1 = -(mc.iC2sq * u3), uS0 = mc.iC2sq * u1;
        // C1: uP; C2: uS1, uS0; rho: every U scalar scales with 1/rho
        o.duX[0] = dri[0] * uC1 + (P ? -2.0 * uP * mc.iC1 : 0.0);
        o.duD[0] = dri[0] * uC0;
        o.duX[1] = dri[1] * uC1 + (S ? -2.0 * uS1 * mc.iC2 : 0.0);
        o.duD[1] = dri[1] * uC0 + (S ? -2.0 * uS0 * mc.iC2 : 0.0);
        o.duX[2] = -mc.iRho * (ri * uC1 + (P ? uP : 0.0) + (S ? uS1 : 0.0));
        o.duD[2] = -mc.iRho * (ri * uC0 + (S ? uS0 : 0.0));
    } else {
        for (int q = 0; q < 3; ++q)
            o.duX[q] = o.duD[q] = 0.0;
    }
}

__device__ __forceinline__ double WaveTotal(double v)
{
    for (int w = 32; w > 0; w >>= 1)
        v += __shfl_xor(v, w);
    return v;
}

// The lane's correlation of one entry: CU transposed (9) and the symmetric
// sums of CT (6).
struct EntryCorr {
    double cut[9], cts[6], trU, trT;
};

__device__ __forceinline__ void LoadCorr(const double* C, long e, bool valid, EntryCorr& k)
{
    const double* c = C + 15 * (valid ? e : 0);
    for (int i = 0; i < 9; ++i)
        k.cut[i] = valid ? c[i] : 0.0;
    for (int i = 0; i < 6; ++i)
        k.cts[i] = valid ? c[9 + i] : 0.0;
    k.trU = k.cut[0] + k.cut[4] + k.cut[8];
    k.trT = k.cts[0] + k.cts[3] + k.cts[5];
}

// Contractions of the basis tensors at one point with the lane's correlation.
__device__ __forceinline__ void PointContract(const double* g, const EntryCorr& k, double& xc, double& qc,
                                              double& bc, double& xt)
{
    const double* r = g + gR;
    const double* n = g + gN;
    double w1[3], w2[3];
    for (int a = 0; a < 3; ++a) {
        w1[a] = k.cut[3 * a] * r[0] + k.cut[3 * a + 1] * r[1] + k.cut[3 * a + 2] * r[2];
        w2[a] = k.cut[3 * a] * n[0] + k.cut[3 * a + 1] * n[1] + k.cut[3 * a + 2] * n[2];
    }
    xc = r[0] * w1[0] + r[1] * w1[1] + r[2] * w1[2];
    qc = n[0] * w1[0] + n[1] * w1[1] + n[2] * w1[2];
    bc = r[0] * w2[0] + r[1] * w2[1] + r[2] * w2[2] + g[gRn] * k.trU;
    const double* X = g + gX;
    xt = X[0] * k.cts[0] + X[1] * k.cts[1] + X[2] * k.cts[2] + X[3] * k.cts[3] + X[4] * k.cts[4] + X[5] * k.cts[5];
}

// Entry finish: the contracted derivatives with the AddBlock signs and the
// stored presence bits, mapped from q = (C1, C2, rho) to (E, v, rho); self
// entries take the self table's derivative for their T2 block.
__device__ __forceinline__ void FinishEntryGrad(const EntryCorr& k, const double* gM, const double* gU,
                                                unsigned char bits, double hs, double ts, bool self1,
                                                const double* dT2self, const double* J, double* out)
{
    double gq[3];
    for (int q = 0; q < 3; ++q)
        gq[q] = ((bits & 1) ? -hs * gM[q] : 0.0) + ((bits & 2) ? ts * gU[q] : 0.0);
    for (int p = 0; p < 3; ++p) {
        double v = J[p] * gq[0] + J[3 + p] * gq[1] + J[6 + p] * gq[2];
        if (self1 && (bits & 1)) {
            const double* dt = dT2self + 9 * p;
            double sdot = 0.0;
            for (int i = 0; i < 9; ++i)
                sdot += dt[i] * k.cut[i];
            v -= hs * sdot;
        }
        out[p] = v;
    }
}

}  // namespace dev

// Derivative pass of the regular-path entries (layout of k_history_reg);
// part[4 wv + p] = the wave's sum of d/d(E, v, rho)_p, part[4 wv + 3] its
// domain.  J: per domain the
// 3 x 3 matrix d(C1, C2, rho)/d(E, v, rho) (row q); dT2: per element the
// derivative of the T2 self table, 27 values (p-major).
__global__ void __launch_bounds__(64) k_hgrad_reg(
    const ElemGeo* __restrict__ geo, const ElemPoints* __restrict__ pts, const NodeRefs* __restrict__ refs,
    PairGrid grid, const long* __restrict__ coff, const long* __restrict__ off, const int* __restrict__ kb,
    const int* __restrict__ len, const unsigned char* __restrict__ cls, long npairs, long c0, long c1,
    const unsigned char* __restrict__ alive, const double* __restrict__ C, const double* __restrict__ J,
    const double* __restrict__ dT2, double* __restrict__ part)
{
    __shared__ double sgeo[kNG2 * dev::kRecD];
    using namespace dev;
    long wv = c0 + (long)blockIdx.x;
    if (wv >= c1)
        return;
    int lane = threadIdx.x;
    long lo = 0, hi = npairs;
    while (hi - lo > 1) {
        long mid = (lo + hi) >> 1;
        if (coff[mid] <= wv)
            lo = mid;
        else
            hi = mid;
    }
    long p = lo;
    int s, f, d;
    DecodePair(grid, p, s, f, d);
    int t = (int)(wv - coff[p]) * 64 + lane;
    const DynaMat& m = grid.mats[d];
    const dev::MatD mc = dev::MakeMatD(m);
    const ElemGeo& fe = geo[f];
    const ElemPoints& P = pts[f];
    const double* src = geo[s].center;
    int k = kb[p] + t;
    double n = (double)k;
    bool inBand = t < len[p];
    long e = off[p] + t;
    EntryFlags fl = JudgeAt(m, fe.radius, Dist3(fe.center, src), k, s == f, inBand ? cls : nullptr, e);
    bool mine = inBand && !fl.patchwise();
    EntryCorr kc;
    LoadCorr(C, e, mine, kc);
    double gM[3] = {0.0, 0.0, 0.0}, gU[3] = {0.0, 0.0, 0.0};
    bool rt = mine && (fl.f2 == 2 || fl.f1 == 2);
    double fr = (mine && fl.f2 == 2 ? 1.0 : 0.0) + (mine && fl.f1 == 2 ? 1.0 : 0.0);
    double fu = mine && fl.f1 == 2 ? 1.0 : 0.0;
    if (WaveAny(rt)) {
        if (lane < kNG2)
            StageRecordD(m, src, P.rg_pos[lane], P.rg_nor[lane], P.rg_w[lane], sgeo + lane * kRecD);
        __syncthreads();
        Span s1 = MakeSpan(m, 1, n), s2 = MakeSpan(m, 2, n);
#pragma unroll 1
        for (int i = 0; i < kNG2; ++i) {
            const double* g = sgeo + i * kRecD;
            SpanDer a, b;
            MixDerAt<1>(g, s1, rt, mc, a);
            MixDerAt<2>(g, s2, rt, mc, b);
            double xc, qc, bc, xt;
            PointContract(g, kc, xc, qc, bc, xt);
            for (int q = 0; q < 3; ++q) {
                gM[q] += fr * ((a.dX[q] + b.dX[q]) * xc + (a.dQ[q] + b.dQ[q]) * qc + (a.dB[q] + b.dB[q]) * bc);
                gU[q] += fu * (a.duX[q] * xt + a.duD[q] * kc.trT);
            }
        }
    }
    double out[3] = {0.0, 0.0, 0.0};
    if (mine)
        FinishEntryGrad(kc, gM, gU, alive[e], refs[f].hu_sign, refs[f].ht_sign, fl.self1, dT2 + 27 * (long)f,
                        J + 9 * d, out);
    for (int q = 0; q < 3; ++q) {
        double v = WaveTotal(out[q]);
        if (lane == 0)
            part[4 * wv + q] = v;
    }
    if (lane == 0)
        part[4 * wv + 3] = (double)d;
}

// Derivative pass of the patch-wise entries (layout of k_history_pw).
__global__ void __launch_bounds__(64) k_hgrad_pw(
    const ElemGeo* __restrict__ geo, const ElemPoints* __restrict__ pts, const NodeRefs* __restrict__ refs,
    PairGrid grid, const PwWave* __restrict__ waves, long w0, long w1, const long* __restrict__ off,
    const int* __restrict__ kb, const int* __restrict__ len, const unsigned char* __restrict__ cls,
    const unsigned char* __restrict__ alive, const double* __restrict__ C, const double* __restrict__ J,
    const double* __restrict__ dT2, double* __restrict__ part)
{
    __shared__ double sgeo[kPwSlots * kNPWG2 * dev::kRecD];
    __shared__ double spr[kPwSlots * 2 * kNPW2];
    __shared__ int ssrc[kPwSlots];
    using namespace dev;
    long wi = w0 + (long)blockIdx.x;
    if (wi >= w1)
        return;
    int lane = threadIdx.x;
    const PwWave& W = waves[wi];
    int nseg = W.nseg;
    int s0, f, d;
    DecodePair(grid, W.pair[0], s0, f, d);
    const DynaMat& m = grid.mats[d];
    const dev::MatD mc = dev::MakeMatD(m);
    const ElemGeo& fe = geo[f];
    const ElemPoints& P = pts[f];
    if (lane < nseg) {
        int ss, ff, dd;
        DecodePair(grid, W.pair[lane], ss, ff, dd);
        ssrc[lane] = ss;
    }
    int seg = -1, rank = 0, base = 0;
    for (int i = 0; i < kPwSlots; ++i) {
        int c = i < nseg ? W.cnt[i] : 0;
        if (seg < 0 && lane >= base && lane < base + c) {
            seg = i;
            rank = W.rank0[i] + (lane - base);
        }
        base += c;
    }
    __syncthreads();
    for (int idx = lane; idx < nseg * kNPW2; idx += 64) {
        int slot = idx / kNPW2, i = idx - slot * kNPW2;
        double r = Dist3(fe.rcenpw[i], geo[ssrc[slot]].center);
        spr[slot * 2 * kNPW2 + 2 * i] = r + fe.radiusPW;
        spr[slot * 2 * kNPW2 + 2 * i + 1] = r - fe.radiusPW;
    }
    bool active = seg >= 0;
    int myseg = active ? seg : 0;
    long p = W.pair[myseg];
    int s = ssrc[myseg];
    const double* src = geo[s].center;
    double r0 = Dist3(fe.center, src);
    int t = -1, k = 0;
    EntryFlags fl{false, 0, 0};
    if (active) {
        int k0 = kb[p], L = len[p], c = 0;
        long eb0 = off[p];
        for (int tt = 0; tt < L; ++tt) {
            EntryFlags e = JudgeAt(m, fe.radius, r0, k0 + tt, s == f, cls, eb0 + tt);
            if (e.patchwise()) {
                if (c == rank) {
                    t = tt;
                    fl = e;
                    break;
                }
                ++c;
            }
        }
        k = k0 + t;
        if (t < 0)
            active = false;
    }
    double n = (double)k;
    long e = active ? off[p] + t : 0;
    __syncthreads();
    unsigned long long m2 = 0, m1 = 0;
    {
        double nn2 = n - 1, nn1 = n;
        double lo2 = nn2 * m.C2Dt, hi2 = (nn2 + 1) * m.C1Dt;
        double lo1 = nn1 * m.C2Dt, hi1 = (nn1 + 1) * m.C1Dt;
        bool pw2 = active && fl.f2 == 1, pw1 = active && fl.f1 == 1;
        const double* sp = spr + myseg * 2 * kNPW2;
        for (int i = 0; i < kNPW2; ++i) {
            double rpr = sp[2 * i], rmr = sp[2 * i + 1];
            if (pw2 && !(rmr >= hi2 || rpr < lo2))
                m2 |= 1ull << i;
            if (pw1 && !(rmr >= hi1 || rpr < lo1))
                m1 |= 1ull << i;
        }
    }
    unsigned long long need = 0;
    for (int i = 0; i < kNPW2; ++i)
        if (WaveAny(((m2 | m1) >> i) & 1ull))
            need |= 1ull << i;
    EntryCorr kc;
    LoadCorr(C, e, active, kc);
    double gM[3] = {0.0, 0.0, 0.0}, gU[3] = {0.0, 0.0, 0.0};
    Span s1 = MakeSpan(m, 1, n), s2 = MakeSpan(m, 2, n);
    bool r2 = active && fl.f2 == 2, r1 = active && fl.f1 == 2, rt = r2 || r1;
    double fr = (r2 ? 1.0 : 0.0) + (r1 ? 1.0 : 0.0);
    double fu = r1 ? 1.0 : 0.0;
    if (WaveAny(rt)) {
        int rgSlot = lane >> 3, rgJ = lane & 7;
        for (int half = 0; half < 2; ++half) {
            if (rgSlot < nseg) {
                int q = half * 8 + rgJ;
                StageRecordD(m, geo[ssrc[rgSlot]].center, P.rg_pos[q], P.rg_nor[q], P.rg_w[q], sgeo + lane * kRecD);
            }
            __syncthreads();
            const double* gs = sgeo + myseg * 8 * kRecD;
#pragma unroll 1
            for (int j = 0; j < 8; ++j) {
                const double* g = gs + j * kRecD;
                SpanDer a, b;
                MixDerAt<1>(g, s1, rt, mc, a);
                MixDerAt<2>(g, s2, rt, mc, b);
                double xc, qc, bc, xt;
                PointContract(g, kc, xc, qc, bc, xt);
                for (int q = 0; q < 3; ++q) {
                    gM[q] += fr * ((a.dX[q] + b.dX[q]) * xc + (a.dQ[q] + b.dQ[q]) * qc + (a.dB[q] + b.dB[q]) * bc);
                    gU[q] += fu * (a.duX[q] * xt + a.duD[q] * kc.trT);
                }
            }
            __syncthreads();
        }
    }
    int stSlot = lane / kNPWG2, stJ = lane - stSlot * kNPWG2;
    while (need) {
        int i = __builtin_ctzll(need);
        need &= need - 1;
        if (stSlot < nseg) {
            int q = i * kNPWG2 + stJ;
            StageRecordD(m, geo[ssrc[stSlot]].center, P.pw_pos[q], P.pw_nor[q], P.pw_w[q], sgeo + lane * kRecD);
        }
        __syncthreads();
        bool u2 = (m2 >> i) & 1ull, u1 = (m1 >> i) & 1ull;
        const double* gs = sgeo + myseg * kNPWG2 * kRecD;
#pragma unroll 1
        for (int j = 0; j < kNPWG2; ++j) {
            const double* g = gs + j * kRecD;
            SpanDer a, b;
            MixDerAt<2>(g, s2, u2, mc, a);
            MixDerAt<1>(g, s1, u1, mc, b);
            double xc, qc, bc, xt;
            PointContract(g, kc, xc, qc, bc, xt);
            for (int q = 0; q < 3; ++q) {
                gM[q] += (a.dX[q] + b.dX[q]) * xc + (a.dQ[q] + b.dQ[q]) * qc + (a.dB[q] + b.dB[q]) * bc;
                gU[q] += b.duX[q] * xt + b.duD[q] * kc.trT;
            }
        }
        __syncthreads();
    }
    double out[3] = {0.0, 0.0, 0.0};
    if (active)
        FinishEntryGrad(kc, gM, gU, alive[e], refs[f].hu_sign, refs[f].ht_sign, fl.self1, dT2 + 27 * (long)f,
                        J + 9 * d, out);
    for (int q = 0; q < 3; ++q) {
        double v = WaveTotal(out[q]);
        if (lane == 0)
            part[4 * wi + q] = v;
    }
    if (lane == 0)
        part[4 * wi + 3] = (double)d;
}

// sums[3 d + p] += sum of the wave partials part[4 w + p] of domain d
// (part[4 w + 3]); one block per (domain, parameter), fixed summation order.
__global__ void __launch_bounds__(256) k_domain_sums(const double* __restrict__ part, long nw,
                                                     double* __restrict__ sums)
{
    int d = blockIdx.x / 3, q = blockIdx.x % 3;
    double a = 0.0;
    for (long w = threadIdx.x; w < nw; w += blockDim.x)
        if ((int)part[4 * w + 3] == d)
            a += part[4 * w + q];
    __shared__ double red[256];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        sums[blockIdx.x] += red[0];
}

void LaunchHistoryGrad(const ElemGeo* geo, const ElemPoints* pts, const NodeRefs* refs, PairGrid grid,
                       const long* coff, const long* off, const int* kb, const int* len, const unsigned char* cls,
                       long npairs, long chunks, const PwWave* waves, long nwaves, const unsigned char* alive,
                       const double* C, const double* J, const double* dT2, double* part, hipStream_t s)
{
    const long step = 1L << 20;
    for (long c0 = 0; c0 < chunks; c0 += step) {
        long c1 = std::min(chunks, c0 + step);
        k_hgrad_reg<<<(unsigned)(c1 - c0), 64, 0, s>>>(geo, pts, refs, grid, coff, off, kb, len, cls, npairs, c0, c1,
                                                       alive, C, J, dT2, part);
    }
    for (long w0 = 0; w0 < nwaves; w0 += step) {
        long w1 = std::min(nwaves, w0 + step);
        k_hgrad_pw<<<(unsigned)(w1 - w0), 64, 0, s>>>(geo, pts, refs, grid, waves, w0, w1, off, kb, len, cls, alive,
                                                      C, J, dT2, part + 4 * chunks);
    }
}

void LaunchDomainSums(const double* part, long nw, int nd, double* sums, hipStream_t s)
{
    if (nw > 0)
        k_domain_sums<<<3 * nd, 256, 0, s>>>(part, nw, sums);
}


// ---------------------------------------------------------------------------
// Step-0 coefficient blocks (DBEM1/dsquareelement.cpp:2756-2868,3281-3394,
// 4215-4550,5157-5491,5922-6078,6476-6632) for FlagDynaMethod 2:
// U: 4 U(0) + U(1);  T: 4 T1(0) + T2(1) + T1(1);  then +-K * R^T.
// ---------------------------------------------------------------------------
namespace dev {

__device__ int Step0Coef(const DynaMat& m, const ElemGeo& fe, const ElemPoints& P, bool self, const double* u1,
                         const double* t1, const double* t2, const double* src, bool unknownPass, int bcid,
                         double* coef, int frozen)
{
    bool useU = unknownPass ? (bcid == 123) : (bcid == 456);
    bool plus = unknownPass ? (bcid != 123) : (bcid == 456);
    double K[9], tmp[9];
    for (int k = 0; k < 9; ++k) {
        K[k] = 0.0;
        coef[k] = 0.0;
    }
    int fsum;
    if (self) {
        if (useU) {
            for (int k = 0; k < 9; ++k)
                K[k] += 4.0 * u1[k];
            IntUPW(m, fe, P, src, 1.0, tmp);
        } else {
            for (int k = 0; k < 9; ++k) {
                K[k] += 4.0 * t1[k];
                K[k] += t2[k];
            }
            IntTPW(m, fe, P, 1, src, 1.0, tmp);
        }
        for (int k = 0; k < 9; ++k)
            K[k] += tmp[k];
        fsum = 1;
    } else {
        int f0 = frozen >= 0 ? (frozen & 3) : NeedCal(m, fe, 1, 0.0, src);
        fsum = f0;
        if (f0) {
            if (useU) {
                if (f0 == 1)
                    IntUPW(m, fe, P, src, 0.0, tmp);
                else
                    IntUReg(m, P, src, 0.0, tmp);
                for (int k = 0; k < 9; ++k)
                    K[k] += 4.0 * tmp[k];
            } else {
                if (f0 == 1)
                    IntTPW(m, fe, P, 1, src, 0.0, tmp);
                else
                    IntTReg(m, P, src, 0.0, tmp);
                for (int k = 0; k < 9; ++k)
                    K[k] += 4.0 * tmp[k];
                if (f0 == 1)
                    IntTPW(m, fe, P, 2, src, 1.0, tmp);
                else
                    IntTReg(m, P, src, 1.0, tmp);
                for (int k = 0; k < 9; ++k)
                    K[k] += tmp[k];
            }
        }
        int f1 = frozen >= 0 ? ((frozen >> 2) & 3) : NeedCal(m, fe, 1, 1.0, src);
        fsum += f1;
        if (f1) {
            if (useU) {
                if (f1 == 1)
                    IntUPW(m, fe, P, src, 1.0, tmp);
                else
                    IntUReg(m, P, src, 1.0, tmp);
            } else {
                if (f1 == 1)
                    IntTPW(m, fe, P, 1, src, 1.0, tmp);
                else
                    IntTReg(m, P, src, 1.0, tmp);
            }
            for (int k = 0; k < 9; ++k)
                K[k] += tmp[k];
        }
        if (!fsum)
            return 0;
    }
    const double* R = fe.trans;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                if (plus)
                    coef[3 * j + i] += K[3 * k + i] * R[3 * j + k];
                else
                    coef[3 * j + i] -= K[3 * k + i] * R[3 * j + k];
            }
    return fsum;
}

// TransformColumnBlock9 (DBEM1/multidomain.cpp:1668-1680)
__device__ void TransformColumnBlock(const double* b, const double* T, double* out)
{
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) {
            double v = 0.0;
            for (int c = 0; c < 3; ++c)
                v += b[c * 3 + r] * T[c * 3 + q];
            out[q * 3 + r] = v;
        }
}

}  // namespace dev

__global__ void __launch_bounds__(64) k_step0(const ElemGeo* __restrict__ geo, const ElemPoints* __restrict__ pts,
                                              const int* __restrict__ dom, const NodeRefs* __restrict__ refs,
                                              const double* __restrict__ su1, const double* __restrict__ st1,
                                              const double* __restrict__ st2, PairGrid grid,
                                              const long* __restrict__ pairIdx, long count,
                                              const unsigned char* __restrict__ cls0, double* out, int* flags)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    int s, f, d;
    dev::DecodePair(grid, pairIdx[i], s, f, d);
    const DynaMat& m = grid.mats[d];
    const ElemGeo& fe = geo[f];
    const ElemPoints& P = pts[f];
    const double* src = geo[s].center;
    bool self = s == f;
    const double* u1 = su1 + 9 * (long)f;
    const double* t1 = st1 + 9 * (long)f;
    const double* t2 = st2 + 9 * (long)f;
    double b0[9], b1[9];
    int fl0, fl1;
    const NodeRefs& r = refs[f];
    int fz = cls0 ? (int)cls0[i] : -1;
    if (r.surface == 1) {
        double c[9];
        fl0 = dev::Step0Coef(m, fe, P, self, u1, t1, t2, src, true, 456, c, fz);
        dev::TransformColumnBlock(c, r.itf_T, b0);
        fl1 = dev::Step0Coef(m, fe, P, self, u1, t1, t2, src, true, 123, c, fz);
        dev::TransformColumnBlock(c, r.itf_T, b1);
    } else {
        fl0 = dev::Step0Coef(m, fe, P, self, u1, t1, t2, src, true, r.bcid, b0, fz);
        fl1 = dev::Step0Coef(m, fe, P, self, u1, t1, t2, src, false, r.bcid, b1, fz);
    }
    for (int k = 0; k < 9; ++k) {
        out[18 * i + k] = b0[k];
        out[18 * i + 9 + k] = b1[k];
    }
    flags[2 * i] = fl0;
    flags[2 * i + 1] = fl1;
}

// Step-0 judges (lag 0 | lag 1 << 2) of the listed near pairs with the
// current materials.
__global__ void __launch_bounds__(256) k_classify0(const ElemGeo* __restrict__ geo, PairGrid grid,
                                                   const long* __restrict__ pairIdx, long count,
                                                   unsigned char* __restrict__ cls0)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    int s, f, d;
    dev::DecodePair(grid, pairIdx[i], s, f, d);
    const DynaMat& m = grid.mats[d];
    const double* src = geo[s].center;
    int f0 = dev::NeedCal(m, geo[f], 1, 0.0, src);
    int f1 = dev::NeedCal(m, geo[f], 1, 1.0, src);
    cls0[i] = (unsigned char)(f0 | (f1 << 2));
}

void LaunchClassify0(const ElemGeo* geo, PairGrid grid, const long* pairIdx, long count, unsigned char* cls0,
                     hipStream_t s)
{
    if (count > 0)
        k_classify0<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(geo, grid, pairIdx, count, cls0);
}

void LaunchStep0(const ElemGeo* geo, const ElemPoints* pts, const int* dom, const NodeRefs* refs,
                 const double* selfU1, const double* selfT1, const double* selfT2, PairGrid grid,
                 const long* pairIdx, long count, const unsigned char* cls0, double* out, int* flags, hipStream_t s)
{
    if (count <= 0)
        return;
    k_step0<<<(unsigned)((count + 63) / 64), 64, 0, s>>>(geo, pts, dom, refs, selfU1, selfT1, selfT2, grid, pairIdx,
                                                         count, cls0, out, flags);
}

// ---------------------------------------------------------------------------
// History convolution (DBEM1/multidomain.cpp:3201-3225,3849-3866):
//   rhs_K += sum_j [ MGS_j (T_{K-j+1} + 2T_{K-j} + T_{K-j-1})
//                  - MTS_j (U_{K-j+1} + 2U_{K-j} + U_{K-j-1}) ]
// with state indices restricted to (0, K).  One workgroup per source row;
// a wavefront walks one pair's band with one lag per lane, so band blocks
// and the per-column state time series are both read contiguously.
// State history layout: g[(col * nstore + m) * 3 + c].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_conv(PairGrid grid, const NodeRefs* __restrict__ refs,
                                              const long* __restrict__ off, const int* __restrict__ kb,
                                              const int* __restrict__ len, const double* __restrict__ mts,
                                              const double* __restrict__ mgs, const double* __restrict__ gU,
                                              const double* __restrict__ gT, int nstore, int K, int maxn,
                                              double* __restrict__ rhs)
{
    int s = blockIdx.x;
    int d = refs[s].domain;
    int nd = grid.eleCount[d];
    long rowStart = grid.pairBase[d] + (long)(s - grid.eleBegin[d]) * nd;
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int jmax = K < maxn ? K : maxn;
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    for (int q = wave; q < nd; q += 4) {
        long p = rowStart + q;
        int L = len[p];
        if (L == 0)
            continue;
        int k0 = kb[p];
        int f = grid.eleBegin[d] + q;
        const double* su = gU + (long)refs[f].hu_block * nstore * 3;
        const double* st = gT + (long)refs[f].ht_block * nstore * 3;
        long base = off[p];
        for (int t = lane; t < L; t += 64) {
            int j = k0 + t;
            if (j > jmax)
                break;
            long e = base + t;
            int mm[3] = {K - j + 1, K - j, K - j - 1};
            const double co[3] = {1.0, 2.0, 1.0};
            double cu0 = 0.0, cu1 = 0.0, cu2 = 0.0, ct0 = 0.0, ct1 = 0.0, ct2 = 0.0;
            for (int z = 0; z < 3; ++z) {
                int mz = mm[z];
                if (mz > 0 && mz < K) {
                    const double* a = su + 3 * (long)mz;
                    const double* b = st + 3 * (long)mz;
                    cu0 += co[z] * a[0];
                    cu1 += co[z] * a[1];
                    cu2 += co[z] * a[2];
                    ct0 += co[z] * b[0];
                    ct1 += co[z] * b[1];
                    ct2 += co[z] * b[2];
                }
            }
            const double* M = mts + 9 * e;
            const double* G = mgs + 6 * e;
            y0 += G[0] * ct0 + G[1] * ct1 + G[2] * ct2;
            y1 += G[1] * ct0 + G[3] * ct1 + G[4] * ct2;
            y2 += G[2] * ct0 + G[4] * ct1 + G[5] * ct2;
            y0 -= M[0] * cu0 + M[1] * cu1 + M[2] * cu2;
            y1 -= M[3] * cu0 + M[4] * cu1 + M[5] * cu2;
            y2 -= M[6] * cu0 + M[7] * cu1 + M[8] * cu2;
        }
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = y0;
    red[1][threadIdx.x] = y1;
    red[2][threadIdx.x] = y2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rhs[3 * (long)s] += red[0][0];
        rhs[3 * (long)s + 1] += red[1][0];
        rhs[3 * (long)s + 2] += red[2][0];
    }
}

void LaunchConvolution(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                       const double* mts, const double* mgs, const double* gU, const double* gT, int nstore, int K,
                       int maxn, int ne, double* rhs, hipStream_t s)
{
    k_conv<<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, gU, gT, nstore, K, maxn, rhs);
}

// ---------------------------------------------------------------------------
// Time-blocked history convolution.
//
// Writing the history of step K as a causal convolution over solved states,
//   h_K = sum_{m=1}^{K-1} sum_j [j, m] (MGS_j T_m - MTS_j U_m),
// with weight 1, 2, 1 for m = K-j+1, K-j, K-j-1, the steps K0 .. K0+T-1 of a
// tile split exactly into
//   far_K  : states m <= K0-1 (all solved when the tile starts), and
//   near_K : states K0 <= m <= K-1, which only reach lags j <= T+1.
// k_conv_far reads every band block once per tile and applies it to the T
// steps of the tile (states not yet solved are still zero in the history
// buffers), so the block traffic per time step drops by T.  The arithmetic
// is reassociated relative to the per-step sum, so results agree with it to
// rounding, not bitwise.
// ---------------------------------------------------------------------------
namespace dev {

__device__ __forceinline__ void BlockApply(const double* M, const double* G, const double* u, const double* t,
                                           double* v)
{
    v[0] = (G[0] * t[0] + G[1] * t[1] + G[2] * t[2]) - (M[0] * u[0] + M[1] * u[1] + M[2] * u[2]);
    v[1] = (G[1] * t[0] + G[3] * t[1] + G[4] * t[2]) - (M[3] * u[0] + M[4] * u[1] + M[5] * u[2]);
    v[2] = (G[2] * t[0] + G[4] * t[1] + G[5] * t[2]) - (M[6] * u[0] + M[7] * u[1] + M[8] * u[2]);
}

__device__ __forceinline__ double WaveSum(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace dev

namespace dev {

// Orders LDS writes and reads of one wavefront's lanes (LDS operations of a
// wavefront complete in issue order).
__device__ __forceinline__ void WaveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace dev

// One block of four wavefronts per source row; each wavefront takes every
// fourth pair of the row and 64 lags of it at a time.  The 1-2-1 combination
// of the three states a lag touches is applied to the state series first:
// the wave stages the solved states (m <= K0-1) its lags reach in LDS and
// forms the filtered series u~[x] = u[x-1] + 2 u[x] + u[x+1] (t~ likewise),
// so each (entry, step) costs one block product G t~ - M u~, formed as 18
// fused multiply-adds into the step's accumulator.
template <int T, typename TB>
__global__ void __launch_bounds__(256) k_conv_far(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                  const long* __restrict__ off, const int* __restrict__ kb,
                                                  const int* __restrict__ len, const TB* __restrict__ mts,
                                                  const TB* __restrict__ mgs, const unsigned short* __restrict__ mtsLo,
                                                  const unsigned short* __restrict__ mgsLo,
                                                  const double* __restrict__ gU, const double* __restrict__ gT,
                                                  int nstore, int K0, int Tn, int maxn, int ne,
                                                  double* __restrict__ hfar)
{
    constexpr int kRaw = 64 + T + 1;  // states a 64-lag chunk reaches
    __shared__ double raw[4][kRaw * 6];
    __shared__ double filt[4][kRaw * 6];
    int s = blockIdx.x;
    int d = refs[s].domain;
    int nd = grid.eleCount[d];
    int eb = grid.eleBegin[d];
    long rowStart = grid.pairBase[d] + (long)(s - eb) * nd;
    int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int lane = threadIdx.x & 63;
    int Kmax = K0 + Tn - 1;
    int jmax = Kmax < maxn ? Kmax : maxn;
    double* rw = raw[wave];
    double* fw = filt[wave];
    double acc[T][3];
#pragma unroll
    for (int i = 0; i < T; ++i)
        acc[i][0] = acc[i][1] = acc[i][2] = 0.0;
    // per-pair metadata, loaded one pair ahead
    int nL = 0, nk0 = 0, nhb = 0, ntb = 0;
    long nbase = 0;
    auto meta = [&](int q) {
        long p = rowStart + q;
        nL = len[p];
        nk0 = kb[p];
        nbase = off[p];
        nhb = refs[eb + q].hu_block;
        ntb = refs[eb + q].ht_block;
    };
    if (wave < nd)
        meta(wave);
    for (int q = wave; q < nd; q += 4) {
        int L = nL, k0 = nk0, hb = nhb, tb = ntb;
        long base = nbase;
        if (q + 4 < nd)
            meta(q + 4);
        int Lj = jmax - k0 + 1 < L ? jmax - k0 + 1 : L;  // lags j <= jmax
        if (Lj <= 0)
            continue;
        const double* su = gU + (long)hb * nstore * 3;
        const double* st = gT + (long)tb * nstore * 3;
        for (int t0 = 0; t0 < Lj; t0 += 64) {
            int nl = Lj - t0 < 64 ? Lj - t0 : 64;
            int t = t0 + lane;
            bool own = lane < nl;
            // the lane's blocks, issued together with the state staging loads
            double M[9], G[6];
            long e = base + (own ? t : 0);
#pragma unroll
            for (int i = 0; i < 9; ++i)
                M[i] = (double)mts[9 * e + i];
#pragma unroll
            for (int i = 0; i < 6; ++i)
                G[i] = (double)mgs[6 * e + i];
            if (mtsLo) {
#pragma unroll
                for (int i = 0; i < 9; ++i)
                    M[i] += dev::Bf16ToDouble(mtsLo[9 * e + i]);
#pragma unroll
                for (int i = 0; i < 6; ++i)
                    G[i] += dev::Bf16ToDouble(mgsLo[6 * e + i]);
            }
            // x = K - j for the chunk's lags and the tile's steps
            int xlo = K0 - (k0 + t0 + nl - 1), xhi = Kmax - (k0 + t0);
            int nx = xhi - xlo + 1;  // <= nl + Tn - 1
            int mlo = xlo - 1;
            for (int idx = lane; idx < (nx + 2) * 6; idx += 64) {
                int r = idx / 6, c = idx - 6 * r;
                int mm = mlo + r;
                double v = 0.0;
                if (mm >= 1 && mm <= K0 - 1)
                    v = c < 3 ? su[3 * (long)mm + c] : st[3 * (long)mm + (c - 3)];
                rw[idx] = v;
            }
            dev::WaveSync();
            for (int idx = lane; idx < nx * 6; idx += 64)
                fw[idx] = (rw[idx] + 2.0 * rw[idx + 6]) + rw[idx + 12];
            dev::WaveSync();
            if (own) {
                int x0 = K0 - (k0 + t) - xlo;  // filtered index of step K0
#pragma unroll
                for (int i = 0; i < T; ++i) {
                    {
                        // acc += MGS t~ - MTS u~, fused multiply-adds straight into the accumulator;
                        // all T steps in one block (the staged series is zero past the tile, and
                        // steps past Tn are never stored)
                        const double* w = fw + 6 * (x0 + i);
                        double u0 = w[0], u1 = w[1], u2 = w[2], t0v = w[3], t1v = w[4], t2v = w[5];
                        double a0 = acc[i][0], a1 = acc[i][1], a2 = acc[i][2];
                        a0 = fma(G[0], t0v, a0); a0 = fma(G[1], t1v, a0); a0 = fma(G[2], t2v, a0);
                        a1 = fma(G[1], t0v, a1); a1 = fma(G[3], t1v, a1); a1 = fma(G[4], t2v, a1);
                        a2 = fma(G[2], t0v, a2); a2 = fma(G[4], t1v, a2); a2 = fma(G[5], t2v, a2);
                        a0 = fma(-M[0], u0, a0); a0 = fma(-M[1], u1, a0); a0 = fma(-M[2], u2, a0);
                        a1 = fma(-M[3], u0, a1); a1 = fma(-M[4], u1, a1); a1 = fma(-M[5], u2, a1);
                        a2 = fma(-M[6], u0, a2); a2 = fma(-M[7], u1, a2); a2 = fma(-M[8], u2, a2);
                        acc[i][0] = a0;
                        acc[i][1] = a1;
                        acc[i][2] = a2;
                    }
                }
            }
            dev::WaveSync();
        }
    }
    __shared__ double red[4][T * 3];
#pragma unroll
    for (int i = 0; i < T; ++i)
        for (int c = 0; c < 3; ++c) {
            double v = dev::WaveSum(acc[i][c]);
            if (lane == 0)
                red[wave][i * 3 + c] = v;
        }
    __syncthreads();
    for (int x = threadIdx.x; x < Tn * 3; x += blockDim.x) {
        double v = ((red[0][x] + red[1][x]) + red[2][x]) + red[3][x];
        int i = x / 3, c = x % 3;
        hfar[(long)i * 3 * ne + 3 * (long)s + c] = v;
    }
}

// Filtered state series of the tile starting at K0, per field element:
// F[f][x] = (u~, t~), u~ = (gU_{x-1} + 2 gU_x) + gU_{x+1} over the solved
// steps 1 .. K0-1 (other states count as zero), for x in [x0, x1].
__global__ void __launch_bounds__(256) k_filter_tile(const NodeRefs* __restrict__ refs, const double* __restrict__ gU,
                                                     const double* __restrict__ gT, int nstore, int ne, int K0,
                                                     int x0, int x1, double* __restrict__ F)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int nx = x1 - x0 + 1;
    if (i >= (long)ne * nx)
        return;
    int f = (int)(i / nx);
    int x = x0 + (int)(i - (long)f * nx);
    const double* su = gU + (long)refs[f].hu_block * nstore * 3;
    const double* st = gT + (long)refs[f].ht_block * nstore * 3;
    double v[6];
    for (int c = 0; c < 3; ++c) {
        double a = (x - 1 >= 1 && x - 1 <= K0 - 1) ? su[3 * (long)(x - 1) + c] : 0.0;
        double b = (x >= 1 && x <= K0 - 1) ? su[3 * (long)x + c] : 0.0;
        double d = (x + 1 >= 1 && x + 1 <= K0 - 1) ? su[3 * (long)(x + 1) + c] : 0.0;
        v[c] = (a + 2.0 * b) + d;
        a = (x - 1 >= 1 && x - 1 <= K0 - 1) ? st[3 * (long)(x - 1) + c] : 0.0;
        b = (x >= 1 && x <= K0 - 1) ? st[3 * (long)x + c] : 0.0;
        d = (x + 1 >= 1 && x + 1 <= K0 - 1) ? st[3 * (long)(x + 1) + c] : 0.0;
        v[3 + c] = (a + 2.0 * b) + d;
    }
    double* o = F + ((long)f * (nstore + 1) + x) * 6;
    for (int c = 0; c < 6; ++c)
        o[c] = v[c];
}

void LaunchFilterTile(const NodeRefs* refs, const double* gU, const double* gT, int nstore, int ne, int K0, int x0,
                      int x1, double* F, hipStream_t s)
{
    long n = (long)ne * (x1 - x0 + 1);
    if (n > 0)
        k_filter_tile<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(refs, gU, gT, nstore, ne, K0, x0, x1, F);
}

namespace dev {

// One (pair, chunk of up to 64 lags) work item of the far convolution with
// its band blocks and filtered-state rows held in registers, so the next
// item's loads are in flight while the current one is computed.
struct FarItem {
    int q, t0, Lj, k0, nl, xlo, nx;
    long base;
    bool valid;
    double M[9], G[6], r0[6], r1[6];
};

}  // namespace dev

// The far part reading a pre-filtered state series (k_filter_tile) and
// prefetching the next work item: rows of F for the chunk's x range go
// straight into LDS, and each wavefront loads item n+1 while it computes
// item n.
template <int T, typename TB>
__global__ void __launch_bounds__(256) k_conv_far_p(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                    const long* __restrict__ off, const int* __restrict__ kb,
                                                    const int* __restrict__ len, const TB* __restrict__ mts,
                                                    const TB* __restrict__ mgs,
                                                    const unsigned short* __restrict__ mtsLo,
                                                    const unsigned short* __restrict__ mgsLo,
                                                    const double* __restrict__ F, int fstride, int K0, int Tn,
                                                    int maxn, int ne, double* __restrict__ hfar)
{
    constexpr int kRows = 64 + T;
    __shared__ double fwin[4][kRows * 6];
    int s = blockIdx.x;
    int d = refs[s].domain;
    int nd = grid.eleCount[d];
    int eb = grid.eleBegin[d];
    long rowStart = grid.pairBase[d] + (long)(s - eb) * nd;
    int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int lane = threadIdx.x & 63;
    int Kmax = K0 + Tn - 1;
    int jmax = Kmax < maxn ? Kmax : maxn;
    double* fw = fwin[wave];
    double acc[T][3];
#pragma unroll
    for (int i = 0; i < T; ++i)
        acc[i][0] = acc[i][1] = acc[i][2] = 0.0;
    // next item after (q, t0) of this wave: the next 64 lags of the pair, or
    // the first chunk of the next pair with lags <= jmax
    auto advance = [&](dev::FarItem& it, int q, int t0, int Lj, int k0, long base) {
        t0 += 64;
        while (q < nd && t0 >= Lj) {
            q += 4;
            t0 = 0;
            if (q < nd) {
                long p = rowStart + q;
                int L = len[p];
                k0 = kb[p];
                base = off[p];
                Lj = jmax - k0 + 1 < L ? jmax - k0 + 1 : L;
            }
        }
        it.valid = q < nd;
        it.q = q;
        it.t0 = t0;
        it.Lj = Lj;
        it.k0 = k0;
        it.base = base;
        if (!it.valid)
            return;
        it.nl = Lj - t0 < 64 ? Lj - t0 : 64;
        it.xlo = K0 - (k0 + t0 + it.nl - 1);
        it.nx = Kmax - (k0 + t0) - it.xlo + 1;
        bool own = lane < it.nl;
        long e = base + (own ? t0 + lane : 0);
#pragma unroll
        for (int i = 0; i < 9; ++i)
            it.M[i] = (double)mts[9 * e + i];
#pragma unroll
        for (int i = 0; i < 6; ++i)
            it.G[i] = (double)mgs[6 * e + i];
        if (mtsLo) {
#pragma unroll
            for (int i = 0; i < 9; ++i)
                it.M[i] += dev::Bf16ToDouble(mtsLo[9 * e + i]);
#pragma unroll
            for (int i = 0; i < 6; ++i)
                it.G[i] += dev::Bf16ToDouble(mgsLo[6 * e + i]);
        }
        const double* Fb = F + (long)(eb + q) * fstride * 6;
        int x = it.xlo + lane;
        bool in0 = lane < it.nx && x >= 0 && x <= K0;
        const double* a = Fb + 6 * (long)(in0 ? x : 0);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            it.r0[c] = in0 ? a[c] : 0.0;
        x += 64;
        bool in1 = lane + 64 < it.nx && x >= 0 && x <= K0;
        const double* b = Fb + 6 * (long)(in1 ? x : 0);
#pragma unroll
        for (int c = 0; c < 6; ++c)
            it.r1[c] = in1 ? b[c] : 0.0;
    };
    auto body = [&](dev::FarItem& cur, dev::FarItem& nxt) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
            fw[6 * lane + c] = cur.r0[c];
        if (lane + 64 < cur.nx)
#pragma unroll
            for (int c = 0; c < 6; ++c)
                fw[6 * (lane + 64) + c] = cur.r1[c];
        dev::WaveSync();
        advance(nxt, cur.q, cur.t0, cur.Lj, cur.k0, cur.base);
        if (lane < cur.nl) {
            int x0 = K0 - (cur.k0 + cur.t0 + lane) - cur.xlo;
            const double* M = cur.M;
            const double* G = cur.G;
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const double* w = fw + 6 * (x0 + i);
                double u0 = w[0], u1 = w[1], u2 = w[2], t0v = w[3], t1v = w[4], t2v = w[5];
                double a0 = acc[i][0], a1 = acc[i][1], a2 = acc[i][2];
                a0 = fma(G[0], t0v, a0); a0 = fma(G[1], t1v, a0); a0 = fma(G[2], t2v, a0);
                a1 = fma(G[1], t0v, a1); a1 = fma(G[3], t1v, a1); a1 = fma(G[4], t2v, a1);
                a2 = fma(G[2], t0v, a2); a2 = fma(G[4], t1v, a2); a2 = fma(G[5], t2v, a2);
                a0 = fma(-M[0], u0, a0); a0 = fma(-M[1], u1, a0); a0 = fma(-M[2], u2, a0);
                a1 = fma(-M[3], u0, a1); a1 = fma(-M[4], u1, a1); a1 = fma(-M[5], u2, a1);
                a2 = fma(-M[6], u0, a2); a2 = fma(-M[7], u1, a2); a2 = fma(-M[8], u2, a2);
                acc[i][0] = a0;
                acc[i][1] = a1;
                acc[i][2] = a2;
            }
        }
        dev::WaveSync();
    };
    dev::FarItem A, B;
    advance(A, wave - 4, 0, 0, 0, 0);
    while (A.valid) {
        body(A, B);
        if (!B.valid)
            break;
        body(B, A);
    }
    __shared__ double red[4][T * 3];
#pragma unroll
    for (int i = 0; i < T; ++i)
        for (int c = 0; c < 3; ++c) {
            double v = dev::WaveSum(acc[i][c]);
            if (lane == 0)
                red[wave][i * 3 + c] = v;
        }
    __syncthreads();
    for (int x = threadIdx.x; x < Tn * 3; x += blockDim.x) {
        double v = ((red[0][x] + red[1][x]) + red[2][x]) + red[3][x];
        int i = x / 3, c = x % 3;
        hfar[(long)i * 3 * ne + 3 * (long)s + c] = v;
    }
}

namespace dev {

constexpr int kFarGroup = 32;  // pairs per group table of a wavefront
constexpr int kFarPairs = 4;   // pairs one packed item may span

// One packed work item of the far convolution: up to 64 consecutive band
// entries of a wavefront's group of pairs, one entry per lane, crossing pair
// boundaries (at most kFarPairs pairs).  The item's pairs own consecutive
// windows of nI + T - 1 filtered-state rows in LDS (nI = the pair's entries
// in the item), ascending in time; lane l stages window slots l and l + 64.
struct FarPack {
    bool own;  // the lane holds an entry
    int row;   // LDS row of the lane's step-K0 state
    double M[9], G[6], r0[6], r1[6];
};

__device__ __forceinline__ int WaveInclusiveScan(int v, int lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(v, o, 64);
        if (lane >= o)
            v += u;
    }
    return v;
}

__device__ __forceinline__ int Uniform(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

}  // namespace dev

// The far part with entries packed across pairs: every wavefront takes
// groups of 32 consecutive pairs of the row (groups wave, wave+4, ...), lists
// the entries with lag <= jmax of the group in pair order and computes them
// 64 at a time, so a lane idles only at the end of a group instead of at the
// end of every band.  Per entry the arithmetic is the one of k_conv_far_p
// (sixteen steps of MGS t~ - MTS u~ fused into the step accumulators); only
// the assignment of entries to lanes differs, and with it the order of the
// final sums.
template <int T, typename TB>
__global__ void __launch_bounds__(64) k_conv_far_q(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                   const long* __restrict__ off, const int* __restrict__ kb,
                                                   const int* __restrict__ len, const TB* __restrict__ mts,
                                                   const TB* __restrict__ mgs,
                                                   const unsigned short* __restrict__ mtsLo,
                                                   const unsigned short* __restrict__ mgsLo,
                                                   const double* __restrict__ F, int fstride, int K0, int Tn,
                                                   int maxn, int ne, double* __restrict__ hfar, long partStride,
                                                   int parts, int rowBegin, long ebase)
{
    constexpr int P = dev::kFarPairs;
    constexpr int kRows = 64 + P * (T - 1);
    static_assert(kRows <= 128, "two window slots per lane");
    // window rows, then T zero rows read by the lanes without an entry (their
    // FMAs then add exact zeros, so the FMA block needs no branch)
    __shared__ double fw[(kRows + T) * 6];
    __shared__ int ex[dev::kFarGroup + 1];  // exclusive prefix of the clipped band lengths, then the total
    __shared__ int gk[dev::kFarGroup];
    __shared__ long go[dev::kFarGroup];
    int s = rowBegin + (int)blockIdx.x / parts;  // source row
    int wave = blockIdx.x % parts;               // part of the row: pair groups wave, wave + parts, ...
    int d = refs[s].domain;
    int nd = grid.eleCount[d];
    int eb = grid.eleBegin[d];
    long rowStart = grid.pairBase[d] + (long)(s - eb) * nd;
    int lane = threadIdx.x;
    int Kmax = K0 + Tn - 1;
    int jmax = Kmax < maxn ? Kmax : maxn;
    for (int z = lane; z < T * 6; z += 64)
        fw[kRows * 6 + z] = 0.0;
    double acc[T][3];
#pragma unroll
    for (int i = 0; i < T; ++i)
        acc[i][0] = acc[i][1] = acc[i][2] = 0.0;
    // wavefront-uniform cursor: group g, next entry E of Gt, first pair i0 not yet finished
    int g = wave - parts, E = 0, Gt = 0, i0 = 0;
    auto loadGroup = [&]() -> bool {
        for (;;) {
            g += parts;
            if (g * dev::kFarGroup >= nd)
                return false;
            int q = g * dev::kFarGroup + lane;
            int L = 0, k0 = 0;
            long o = 0;
            if (lane < dev::kFarGroup && q < nd) {
                long p = rowStart + q;
                k0 = kb[p];
                o = off[p] - ebase;
                int l = len[p];
                int lj = jmax - k0 + 1;
                L = lj < l ? lj : l;
                L = L > 0 ? L : 0;
            }
            int incl = dev::WaveInclusiveScan(L, lane);
            if (lane < dev::kFarGroup) {
                ex[lane] = incl - L;
                gk[lane] = k0;
                go[lane] = o;
            }
            if (lane == dev::kFarGroup - 1)
                ex[dev::kFarGroup] = incl;
            Gt = dev::Uniform(__shfl(incl, dev::kFarGroup - 1, 64));
            dev::WaveSync();
            E = 0;
            i0 = 0;
            if (Gt > 0)
                return true;
        }
    };
    // Issues the same loads whether or not work is left (with no work left the
    // item is empty: no lane owns an entry, every window slot reads the zero
    // row), so the number of loads in flight after an advance is the same on
    // every path and the wait before the current item's FMAs covers only the
    // loads of the item before it.
    auto advance = [&](dev::FarPack& it) -> bool {
        const bool live = !(E >= Gt && !loadGroup());
        if (live)
            while (ex[i0 + 1] <= E)
                ++i0;
        // the item's candidate pairs i0 .. i0+P-1, wavefront-uniform
        int cEx[P + 1], cKb[P];
        long cOff[P];
#pragma unroll
        for (int k = 0; k <= P; ++k) {
            int c = i0 + k;
            cEx[k] = dev::Uniform(ex[c <= dev::kFarGroup ? c : dev::kFarGroup]);
        }
#pragma unroll
        for (int k = 0; k < P; ++k) {
            int c = i0 + k < dev::kFarGroup ? i0 + k : dev::kFarGroup - 1;
            cKb[k] = dev::Uniform(gk[c]);
            long o = go[c];
            cOff[k] = ((long)(unsigned)dev::Uniform((int)(o >> 32)) << 32) | (unsigned)dev::Uniform((int)o);
        }
        int end = live ? (E + 64 < cEx[P] ? E + 64 : cEx[P]) : E;
        // windows: pair k holds entries [first, last] of the item and rows wb[k] .. wb[k] + n + T - 2
        int wb[P], tb[P], xl[P];
        int rows = 0;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            int first = cEx[k] > E ? cEx[k] : E;
            int last = (cEx[k + 1] < end ? cEx[k + 1] : end) - 1;
            int n = last - first + 1;
            wb[k] = rows;
            tb[k] = last - cEx[k];
            xl[k] = K0 - cKb[k] - tb[k];
            rows += n > 0 ? n + T - 1 : 0;
        }
        // the lane's entry e: the last pair starting at or before e
        int e = E + lane;
        it.own = e < end;
        int ke = 0;
#pragma unroll
        for (int k = 1; k < P; ++k)
            if (cEx[k] <= e)
                ke = k;
        int wbE = wb[0], tbE = tb[0], exE = cEx[0];
        long offE = cOff[0];
#pragma unroll
        for (int k = 1; k < P; ++k)
            if (ke == k) {
                wbE = wb[k];
                tbE = tb[k];
                exE = cEx[k];
                offE = cOff[k];
            }
        int t = e - exE;
        it.row = it.own ? wbE + tbE - t : kRows;  // lanes without an entry: the zero rows
        long en = it.own ? offE + t : 0;  // lanes without an entry read entry 0 (times zero states)
#pragma unroll
        for (int c = 0; c < 9; ++c)
            it.M[c] = (double)mts[9 * en + c];
#pragma unroll
        for (int c = 0; c < 6; ++c)
            it.G[c] = (double)mgs[6 * en + c];
        if (mtsLo) {
#pragma unroll
            for (int c = 0; c < 9; ++c)
                it.M[c] += dev::Bf16ToDouble(mtsLo[9 * en + c]);
#pragma unroll
            for (int c = 0; c < 6; ++c)
                it.G[c] += dev::Bf16ToDouble(mgsLo[6 * en + c]);
        }
        // window slots lane and lane + 64: the last pair whose window starts at or before the slot
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int sl = lane + 64 * h;
            int ks = 0;
#pragma unroll
            for (int k = 1; k < P; ++k)
                if (wb[k] <= sl)
                    ks = k;
            int wbS = wb[0], xlS = xl[0];
#pragma unroll
            for (int k = 1; k < P; ++k)
                if (ks == k) {
                    wbS = wb[k];
                    xlS = xl[k];
                }
            int x = xlS + (sl - wbS);
            bool in = sl < rows && x >= 0 && x <= K0;
            // slots outside the windows read row fstride - 1 of element 0,
            // which the filter never writes (zero): the loads stay
            // unconditional, so the wait for them falls at their first use
            // in the next body and not before the current item's FMAs
            const double* a =
                F + ((long)(in ? eb + g * dev::kFarGroup + i0 + ks : 0) * fstride + (in ? x : fstride - 1)) * 6;
            double* r = h == 0 ? it.r0 : it.r1;
#pragma unroll
            for (int c = 0; c < 6; ++c)
                r[c] = a[c];
        }
        E = end;
        return live;
    };
    auto body = [&](dev::FarPack& cur, dev::FarPack& nxt) -> bool {
#pragma unroll
        for (int c = 0; c < 6; ++c)
            fw[6 * lane + c] = cur.r0[c];
        if (lane + 64 < kRows)
#pragma unroll
            for (int c = 0; c < 6; ++c)
                fw[6 * (lane + 64) + c] = cur.r1[c];
        dev::WaveSync();
        bool more = advance(nxt);
        {
            const double* M = cur.M;
            const double* G = cur.G;
            const double* wr = fw + 6 * cur.row;
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const double* w = wr + 6 * i;
                double u0 = w[0], u1 = w[1], u2 = w[2], t0v = w[3], t1v = w[4], t2v = w[5];
                double a0 = acc[i][0], a1 = acc[i][1], a2 = acc[i][2];
                a0 = fma(G[0], t0v, a0); a0 = fma(G[1], t1v, a0); a0 = fma(G[2], t2v, a0);
                a1 = fma(G[1], t0v, a1); a1 = fma(G[3], t1v, a1); a1 = fma(G[4], t2v, a1);
                a2 = fma(G[2], t0v, a2); a2 = fma(G[4], t1v, a2); a2 = fma(G[5], t2v, a2);
                a0 = fma(-M[0], u0, a0); a0 = fma(-M[1], u1, a0); a0 = fma(-M[2], u2, a0);
                a1 = fma(-M[3], u0, a1); a1 = fma(-M[4], u1, a1); a1 = fma(-M[5], u2, a1);
                a2 = fma(-M[6], u0, a2); a2 = fma(-M[7], u1, a2); a2 = fma(-M[8], u2, a2);
                acc[i][0] = a0;
                acc[i][1] = a1;
                acc[i][2] = a2;
            }
        }
        dev::WaveSync();
        return more;
    };
    dev::FarPack A, B;
    bool more = advance(A);
    while (more) {
        if (!body(A, B))
            break;
        more = body(B, A);
    }
    // the part's partial sums; the near kernel adds the parts in order
    double* out = hfar + wave * partStride + 3 * (long)s;
#pragma unroll
    for (int i = 0; i < T; ++i)
        for (int c = 0; c < 3; ++c) {
            double v = dev::WaveSum(acc[i][c]);
            if (lane == 0 && i < Tn)
                out[(long)i * 3 * ne + c] = v;
        }
}

template <typename TB>
void LaunchConvFarF(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                    const TB* mts, const TB* mgs, const unsigned short* mtsLo, const unsigned short* mgsLo,
                    const double* F, int fstride, int K0, int Tn, int maxn, int ne, double* hfar, int parts,
                    hipStream_t s, int rowBegin, int rowCount, long ebase)
{
    if (rowCount < 0)
        rowCount = ne - rowBegin;
    if (rowCount <= 0)
        return;
    if (parts > 0)
        k_conv_far_q<16, TB><<<(unsigned)(parts * rowCount), 64, 0, s>>>(
            grid, refs, off, kb, len, mts, mgs, mtsLo, mgsLo, F, fstride, K0, Tn, maxn, ne, hfar,
            (long)kConvTileMax * 3 * ne, parts, rowBegin, ebase);
    else if (rowBegin == 0 && rowCount == ne && ebase == 0)
        k_conv_far_p<16, TB><<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, mtsLo, mgsLo, F, fstride, K0,
                                                Tn, maxn, ne, hfar);
    else
        throw std::runtime_error("the one-pair-per-item far kernel runs over every row of resident bands only");
}
template void LaunchConvFarF<double>(PairGrid, const NodeRefs*, const long*, const int*, const int*, const double*,
                                     const double*, const unsigned short*, const unsigned short*, const double*, int,
                                     int, int, int, int, double*, int, hipStream_t, int, int, long);
template void LaunchConvFarF<float>(PairGrid, const NodeRefs*, const long*, const int*, const int*, const float*,
                                    const float*, const unsigned short*, const unsigned short*, const double*, int,
                                    int, int, int, int, double*, int, hipStream_t, int, int, long);

// Near part of step K in the tile starting at K0 plus the far part:
// rhs_K += hfar[K-K0] + sum over near pairs (band start <= K-K0+1) of the
// taps with K0 <= m <= K-1.  One wavefront per source row, one near pair per
// lane (each pair has at most K-K0+1 near lags).
template <typename TB>
__global__ void __launch_bounds__(64) k_conv_near(const NodeRefs* __restrict__ refs, const long* __restrict__ nearPtr,
                                                  const long* __restrict__ nearPair, const int* __restrict__ nearField,
                                                  const long* __restrict__ off, const int* __restrict__ kb,
                                                  const int* __restrict__ len, const TB* __restrict__ mts,
                                                  const TB* __restrict__ mgs, const unsigned short* __restrict__ mtsLo,
                                                  const unsigned short* __restrict__ mgsLo,
                                                  const double* __restrict__ gU,
                                                  const double* __restrict__ gT, int nstore, int K, int K0, int maxn,
                                                  int ne, const double* __restrict__ hfar, int parts,
                                                  long partStride, double* __restrict__ rhs)
{
    int s = blockIdx.x;
    int lane = threadIdx.x;
    int jlim = K - K0 + 1;
    if (jlim > maxn)
        jlim = maxn;
    double y[3] = {0.0, 0.0, 0.0};
    for (long q = nearPtr[s] + lane; q < nearPtr[s + 1]; q += 64) {
        long p = nearPair[q];
        int k0 = kb[p], L = len[p];
        int jmax = k0 + L - 1 < jlim ? k0 + L - 1 : jlim;
        if (jmax < k0)
            continue;
        int f = nearField[q];
        const double* su = gU + (long)refs[f].hu_block * nstore * 3;
        const double* st = gT + (long)refs[f].ht_block * nstore * 3;
        long e0 = off[p] - k0;
        for (int j = k0; j <= jmax; ++j) {
            long e = e0 + j;
            double M[9], G[6];
            for (int i = 0; i < 9; ++i)
                M[i] = (double)mts[9 * e + i] + (mtsLo ? dev::Bf16ToDouble(mtsLo[9 * e + i]) : 0.0);
            for (int i = 0; i < 6; ++i)
                G[i] = (double)mgs[6 * e + i] + (mgsLo ? dev::Bf16ToDouble(mgsLo[6 * e + i]) : 0.0);
            const int mm[3] = {K - j + 1, K - j, K - j - 1};
            const double co[3] = {1.0, 2.0, 1.0};
            double cu[3] = {0.0, 0.0, 0.0}, ct[3] = {0.0, 0.0, 0.0};
            for (int z = 0; z < 3; ++z) {
                int m = mm[z];
                if (m >= K0 && m <= K - 1 && m >= 1) {
                    for (int c = 0; c < 3; ++c) {
                        cu[c] += co[z] * su[3 * (long)m + c];
                        ct[c] += co[z] * st[3 * (long)m + c];
                    }
                }
            }
            double v[3];
            dev::BlockApply(M, G, cu, ct, v);
            y[0] += v[0];
            y[1] += v[1];
            y[2] += v[2];
        }
    }
    for (int c = 0; c < 3; ++c)
        y[c] = dev::WaveSum(y[c]);
    if (lane == 0) {
        const double* hf = hfar + (long)(K - K0) * 3 * ne + 3 * (long)s;
        for (int c = 0; c < 3; ++c) {
            double h = hf[c];
            for (int q = 1; q < parts; ++q)
                h += hf[q * partStride + c];
            rhs[3 * (long)s + c] += h + y[c];
        }
    }
}

// The same near part with one (pair, lag) entry per lane: four wavefronts per
// source row take its near pairs in turn, a wavefront holds 64 / L pairs at a
// time with L = the power of two >= the step's lag limit, and the rows' four
// wavefront sums are added in order with the far parts.
template <typename TB>
__global__ void __launch_bounds__(256) k_conv_near_e(const NodeRefs* __restrict__ refs,
                                                     const long* __restrict__ nearPtr,
                                                     const long* __restrict__ nearPair,
                                                     const int* __restrict__ nearField, const long* __restrict__ off,
                                                     const int* __restrict__ kb, const int* __restrict__ len,
                                                     const TB* __restrict__ mts, const TB* __restrict__ mgs,
                                                     const unsigned short* __restrict__ mtsLo,
                                                     const unsigned short* __restrict__ mgsLo,
                                                     const double* __restrict__ gU, const double* __restrict__ gT,
                                                     int nstore, int K, int K0, int maxn, int ne,
                                                     const double* __restrict__ hfar, int parts, long partStride,
                                                     double* __restrict__ rhs)
{
    __shared__ double red[4][3];
    int s = blockIdx.x;
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int jlim = K - K0 + 1;
    if (jlim > maxn)
        jlim = maxn;
    int L = 1;
    while (L < jlim)
        L <<= 1;
    const int per = 64 / L;  // pairs per wavefront pass
    int slot = lane / L, t = lane - slot * L;
    double y[3] = {0.0, 0.0, 0.0};
    long q0 = nearPtr[s], q1 = nearPtr[s + 1];
    for (long qb = q0 + (long)wave * per; qb < q1; qb += 4L * per) {
        // the list is ordered by band start: past the lag limit nothing is left
        if (__builtin_amdgcn_readfirstlane(kb[nearPair[qb]]) > jlim)
            break;
        long q = qb + slot;
        if (q >= q1)
            continue;
        long p = nearPair[q];
        int k0 = kb[p], Lp = len[p];
        int j = k0 + t;
        if (j > jlim || t >= Lp)
            continue;
        int f = nearField[q];
        const double* su = gU + (long)refs[f].hu_block * nstore * 3;
        const double* st = gT + (long)refs[f].ht_block * nstore * 3;
        long e = off[p] + t;
        double M[9], G[6];
        for (int i = 0; i < 9; ++i)
            M[i] = (double)mts[9 * e + i] + (mtsLo ? dev::Bf16ToDouble(mtsLo[9 * e + i]) : 0.0);
        for (int i = 0; i < 6; ++i)
            G[i] = (double)mgs[6 * e + i] + (mgsLo ? dev::Bf16ToDouble(mgsLo[6 * e + i]) : 0.0);
        const int mm[3] = {K - j + 1, K - j, K - j - 1};
        const double co[3] = {1.0, 2.0, 1.0};
        double cu[3] = {0.0, 0.0, 0.0}, ct[3] = {0.0, 0.0, 0.0};
        for (int z = 0; z < 3; ++z) {
            int m = mm[z];
            if (m >= K0 && m <= K - 1 && m >= 1) {
                for (int c = 0; c < 3; ++c) {
                    cu[c] += co[z] * su[3 * (long)m + c];
                    ct[c] += co[z] * st[3 * (long)m + c];
                }
            }
        }
        double v[3];
        dev::BlockApply(M, G, cu, ct, v);
        y[0] += v[0];
        y[1] += v[1];
        y[2] += v[2];
    }
    for (int c = 0; c < 3; ++c) {
        double w = dev::WaveSum(y[c]);
        if (lane == 0)
            red[wave][c] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double* hf = hfar + (long)(K - K0) * 3 * ne + 3 * (long)s;
        for (int c = 0; c < 3; ++c) {
            double h = hf[c];
            for (int qq = 1; qq < parts; ++qq)
                h += hf[qq * partStride + c];
            double yn = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
            rhs[3 * (long)s + c] += h + yn;
        }
    }
}

template <typename TB>
void LaunchConvFar(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len, const TB* mts,
                   const TB* mgs, const unsigned short* mtsLo, const unsigned short* mgsLo, const double* gU,
                   const double* gT, int nstore, int K0, int Tn, int tile, int maxn, int ne, double* hfar, hipStream_t s)
{
    if (tile == 16)
        k_conv_far<16, TB><<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, mtsLo, mgsLo, gU, gT, nstore, K0, Tn,
                                              maxn, ne, hfar);
    else
        k_conv_far<8, TB><<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, mtsLo, mgsLo, gU, gT, nstore, K0, Tn,
                                             maxn, ne, hfar);
}

template <typename TB>
void LaunchConvNear(const NodeRefs* refs, const long* nearPtr, const long* nearPair, const int* nearField,
                    const long* off, const int* kb, const int* len, const TB* mts, const TB* mgs,
                    const unsigned short* mtsLo, const unsigned short* mgsLo, const double* gU, const double* gT,
                    int nstore, int K, int K0, int maxn, int ne, const double* hfar, int parts, double* rhs,
                    bool perEntry, hipStream_t s)
{
    if (perEntry)
        k_conv_near_e<TB><<<ne, 256, 0, s>>>(refs, nearPtr, nearPair, nearField, off, kb, len, mts, mgs, mtsLo, mgsLo,
                                             gU, gT, nstore, K, K0, maxn, ne, hfar, parts,
                                             (long)kConvTileMax * 3 * ne, rhs);
    else
        k_conv_near<TB><<<ne, 64, 0, s>>>(refs, nearPtr, nearPair, nearField, off, kb, len, mts, mgs, mtsLo, mgsLo, gU,
                                          gT, nstore, K, K0, maxn, ne, hfar, parts, (long)kConvTileMax * 3 * ne, rhs);
}

#define TDBEM_CONV_INSTANTIATE(TB)                                                                                   \
    template void LaunchConvFar<TB>(PairGrid, const NodeRefs*, const long*, const int*, const int*, const TB*,        \
                                    const TB*, const unsigned short*, const unsigned short*, const double*,           \
                                    const double*, int, int, int, int, int, int, double*, hipStream_t);               \
    template void LaunchConvNear<TB>(const NodeRefs*, const long*, const long*, const int*, const long*, const int*,  \
                                     const int*, const TB*, const TB*, const unsigned short*, const unsigned short*,  \
                                     const double*, const double*, int, int, int, int, int, const double*, int,       \
                                     double*, bool, hipStream_t);
TDBEM_CONV_INSTANTIATE(double)
TDBEM_CONV_INSTANTIATE(float)
#undef TDBEM_CONV_INSTANTIATE

// ---------------------------------------------------------------------------
// Linear algebra for GMRES
// ---------------------------------------------------------------------------
// Storage types: TA (matrix / inverse values), TX (input vector), TY (output
// vector); every product and sum is formed in FP64.  With all types double
// these are the FP64 kernels of the parity path; float storage serves the
// inner solves of the mixed-precision mode.
// One thread per (block row, component): each component's sum over the
// row's blocks in CCSR order, three products per block in the reference
// order (the same bits as one thread per row computing all three).
template <typename TA, typename TX, typename TY>
__global__ void __launch_bounds__(256) k_spmv(long nrows, const int* __restrict__ rp, const int* __restrict__ col,
                                              const TA* __restrict__ val, const TX* __restrict__ x,
                                              TY* __restrict__ y)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nrows)
        return;
    const long i = t / 3;
    const int c = (int)(t - 3 * i);
    double a = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const TA* v = val + 9 * (long)j + 3 * c;
        const TX* xc = x + 3 * (long)col[j];
        a += (double)v[0] * (double)xc[0] + (double)v[1] * (double)xc[1] + (double)v[2] * (double)xc[2];
    }
    y[t] = (TY)a;
}

template <typename TA, typename TX, typename TY>
void LaunchSpmv(long nrows, const int* rp, const int* col, const TA* val, const TX* x, TY* y, hipStream_t s)
{
    k_spmv<TA, TX, TY><<<(unsigned)((3 * nrows + 255) / 256), 256, 0, s>>>(nrows, rp, col, val, x, y);
}

// GMRES_M_X (DBEM1/pre_gmres.cpp:837-895): each output scalar is owned by
// one lane and summed in the reference order.
template <typename TA, typename TX, typename TY>
__global__ void __launch_bounds__(256) k_precond(long n, const long* __restrict__ posLeaf,
                                                 const long* __restrict__ leafBegin, const long* __restrict__ leafSize,
                                                 const long* __restrict__ matOff, const TA* __restrict__ inv,
                                                 const long* __restrict__ repid, const long* __restrict__ outBlock,
                                                 const TX* __restrict__ x, TY* __restrict__ y)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n)
        return;
    long i = t / 3;
    int mrow = (int)(t % 3);
    long leaf = posLeaf[i];
    long b = leafBegin[leaf], c = leafSize[leaf], jl = i - b, n3 = 3 * c;
    const TA* P = inv + matOff[leaf] + (3 * jl + mrow) * n3;
    double acc = 0.0;
    for (long k = 0; k < c; ++k) {
        const TX* xc = x + 3 * repid[b + k];
        acc += (double)P[3 * k] * (double)xc[0];
        acc += (double)P[3 * k + 1] * (double)xc[1];
        acc += (double)P[3 * k + 2] * (double)xc[2];
    }
    y[3 * outBlock[i] + mrow] = (TY)acc;
}

template <typename TA, typename TX, typename TY>
void LaunchPrecond(long n, const long* posLeaf, const long* leafBegin, const long* leafSize, const long* matOff,
                   const TA* inv, const long* repid, const long* outBlock, const TX* x, TY* y, hipStream_t s)
{
    k_precond<TA, TX, TY><<<(unsigned)((3 * n + 255) / 256), 256, 0, s>>>(n, posLeaf, leafBegin, leafSize, matOff,
                                                                          inv, repid, outBlock, x, y);
}

// Deterministic two-stage reductions (fixed partition and order).
template <typename TV, typename TU>
__global__ void __launch_bounds__(256) k_multidot(const TV* __restrict__ V, long ldv, int nv,
                                                  const TU* __restrict__ u, long n, double* __restrict__ partial)
{
    __shared__ double red[256];
    for (int k = 0; k < nv; ++k) {
        double a = 0.0;
        for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
            a += (double)V[k * ldv + r] * (double)u[r];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
                red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partial[(long)blockIdx.x * nv + k] = red[0];
        __syncthreads();
    }
}

// h[k] = sum of the block partials in block order; the loads go out eight at
// a time ahead of the adds.
__global__ void k_finish(const double* __restrict__ partial, int nblocks, int nv, double* __restrict__ h, int doSqrt)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nv)
        return;
    double a = 0.0;
    int b = 0;
    for (; b + 8 <= nblocks; b += 8) {
        double t[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            t[i] = partial[(long)(b + i) * nv + k];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            a += t[i];
    }
    for (; b < nblocks; ++b)
        a += partial[(long)b * nv + k];
    h[k] = doSqrt ? sqrt(a) : a;
}

template <typename TV, typename TU>
void LaunchMultiDot(const TV* V, long ldv, int nv, const TU* u, long n, double* partial, int nblocks, double* h,
                    hipStream_t s)
{
    k_multidot<TV, TU><<<nblocks, 256, 0, s>>>(V, ldv, nv, u, n, partial);
    k_finish<<<(nv + 63) / 64, 64, 0, s>>>(partial, nblocks, nv, h, 0);
}

template <typename T>
void LaunchNorm(const T* x, long n, double* partial, int nblocks, double* out, hipStream_t s)
{
    k_multidot<T, T><<<nblocks, 256, 0, s>>>(x, n, 1, x, n, partial);
    k_finish<<<1, 64, 0, s>>>(partial, nblocks, 1, out, 1);
}

template <typename TV, typename TU>
__global__ void __launch_bounds__(256) k_multiaxpy(const TV* __restrict__ V, long ldv, int nv,
                                                   const double* __restrict__ h, TU* __restrict__ u, long n,
                                                   double* htot)
{
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (htot && r < nv)
        htot[r] += h[r];
    if (r >= n)
        return;
    double a = (double)u[r];
    for (int k = 0; k < nv; ++k)
        a -= h[k] * (double)V[k * ldv + r];
    u[r] = (TU)a;
}

template <typename TV, typename TU>
void LaunchMultiAxpy(const TV* V, long ldv, int nv, const double* h, TU* u, long n, double* htot, hipStream_t s)
{
    long g = n > nv ? n : nv;
    k_multiaxpy<TV, TU><<<(unsigned)((g + 255) / 256), 256, 0, s>>>(V, ldv, nv, h, u, n, htot);
}

// y = x / a[0] (invert) or y = x * a[0]; a zero or tiny denominator (<= 1e-12)
// gives y = 0, as the reference does for a lucky breakdown.
template <typename TX, typename TY>
__global__ void __launch_bounds__(256) k_scale(const TX* __restrict__ x, const double* __restrict__ a, int invert,
                                               TY* __restrict__ y, long n)
{
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    double s = a[0];
    if (invert)
        y[r] = (TY)(s <= 1.0e-12 ? 0.0 : (double)x[r] / s);
    else
        y[r] = (TY)((double)x[r] * s);
}

template <typename TX, typename TY>
void LaunchScaleCopy(const TX* x, const double* a, int invert, TY* y, long n, hipStream_t s)
{
    k_scale<TX, TY><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(x, a, invert, y, n);
}

__global__ void __launch_bounds__(256) k_residual(const double* __restrict__ b, double* __restrict__ y, long n)
{
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        y[r] = b[r] - y[r];
}

void LaunchResidual(const double* b, double* y, long n, hipStream_t s)
{
    k_residual<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(b, y, n);
}

// x += y (FP64), the refinement update of the mixed-precision solve.
__global__ void __launch_bounds__(256) k_axpy1(const double* __restrict__ y, double* __restrict__ x, long n)
{
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        x[r] += y[r];
}

void LaunchAdd(const double* y, double* x, long n, hipStream_t s)
{
    k_axpy1<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(y, x, n);
}

// float copy of an FP64 array (mixed-precision matrices).
__global__ void __launch_bounds__(256) k_tofloat(const double* __restrict__ x, float* __restrict__ y, long n)
{
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        y[r] = (float)x[r];
}

void LaunchToFloat(const double* x, float* y, long n, hipStream_t s)
{
    if (n > 0)
        k_tofloat<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(x, y, n);
}

// Givens rotation in QT form for column j (DBEM1/pre_gmres.cpp:1492-1533),
// FP64 on the device.  h is 1-based (h[1..j+1]); QT is (restart+1)^2 and R is
// restart^2, both row-major 1-based via the macros below.
__global__ void __launch_bounds__(256) k_givens(double* h, double* QT, double* R, double* f, double* phibar,
                                                int* stag, int j, int restart)
{
    const int lq = restart + 1;
#define QQ(a, b) QT[((a)-1) * lq + ((b)-1)]
#define RR(a, b) R[((a)-1) * restart + ((b)-1)]
    for (int k = 1 + (int)threadIdx.x; k <= j; k += blockDim.x) {
        double v = 0.0;
        for (int l = 1; l <= j; ++l)
            v += QQ(k, l) * h[l];
        RR(k, j) = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double rt = RR(j, j), c, s, temp;
        double hj1 = h[j + 1];
        if (hj1 == 0.0) {
            c = 1.0;
            s = 0.0;
        } else if (hj1 * hj1 > rt * rt) {
            temp = rt / hj1;
            s = 1.0 / sqrt(1.0 + fabs(temp) * fabs(temp));
            c = -temp * s;
        } else {
            temp = hj1 / rt;
            c = 1.0 / sqrt(1.0 + fabs(temp) * fabs(temp));
            s = -temp * c;
        }
        RR(j, j) = c * rt - s * hj1;
        for (int k = 1; k <= j; ++k) {
            double qk = QQ(j, k);
            QQ(j, k) = c * qk;
            QQ(j + 1, k) = s * qk;
        }
        QQ(j, j + 1) = -s;
        QQ(j + 1, j + 1) = c;
        f[j] = c * phibar[0];
        phibar[0] = s * phibar[0];
        if (j < restart && f[j] == 0.0)
            stag[0] = 1;
    }
#undef QQ
#undef RR
}

void LaunchGivens(double* h, double* QT, double* R, double* f, double* phibar, int* stag, int j, int restart,
                  hipStream_t s)
{
    k_givens<<<1, 256, 0, s>>>(h, QT, R, f, phibar, stag, j, restart);
}

// W(:,j) = (V(:,j) - sum_{l<j} W(:,l) R(l,j)) / R(j,j);  x += f_j W(:,j)
// (DBEM1/pre_gmres.cpp:1535-1555).  V and W column-major with ld = n.
template <typename TV>
__global__ void __launch_bounds__(256) k_wupdate(const TV* __restrict__ V, TV* __restrict__ W, long n,
                                                 const double* __restrict__ R, int j, int restart,
                                                 const double* __restrict__ f, double* __restrict__ x)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    double rjj = R[(j - 1) * restart + (j - 1)];
    double w = (double)V[(long)(j - 1) * n + k] / rjj;
    for (int l = 1; l <= j - 1; ++l)
        w -= (double)W[(long)(l - 1) * n + k] * R[(l - 1) * restart + (j - 1)] / rjj;
    W[(long)(j - 1) * n + k] = (TV)w;
    x[k] += f[j] * w;
}

template <typename TV>
void LaunchWUpdate(const TV* V, TV* W, long n, const double* R, int j, int restart, const double* f, double* x,
                   hipStream_t s)
{
    k_wupdate<TV><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(V, W, n, R, j, restart, f, x);
}

#define TDBEM_LA_INSTANTIATE(T)                                                                                      \
    template void LaunchSpmv<T, T, T>(long, const int*, const int*, const T*, const T*, T*, hipStream_t);           \
    template void LaunchPrecond<T, T, T>(long, const long*, const long*, const long*, const long*, const T*,        \
                                         const long*, const long*, const T*, T*, hipStream_t);                      \
    template void LaunchMultiDot<T, T>(const T*, long, int, const T*, long, double*, int, double*, hipStream_t);    \
    template void LaunchNorm<T>(const T*, long, double*, int, double*, hipStream_t);                                \
    template void LaunchMultiAxpy<T, T>(const T*, long, int, const double*, T*, long, double*, hipStream_t);        \
    template void LaunchScaleCopy<T, T>(const T*, const double*, int, T*, long, hipStream_t);                       \
    template void LaunchWUpdate<T>(const T*, T*, long, const double*, int, int, const double*, double*, hipStream_t);
TDBEM_LA_INSTANTIATE(double)
TDBEM_LA_INSTANTIATE(float)
#undef TDBEM_LA_INSTANTIATE
template void LaunchSpmv<float, double, double>(long, const int*, const int*, const float*, const double*, double*,
                                                hipStream_t);
template void LaunchPrecond<float, double, float>(long, const long*, const long*, const long*, const long*,
                                                  const float*, const long*, const long*, const double*, float*,
                                                  hipStream_t);

// ---------------------------------------------------------------------------
// State plumbing (DBEM1/multidomain.cpp:3047-3200; dsquareelement.h:88-122)
// ---------------------------------------------------------------------------
__global__ void k_known(const NodeRefs* __restrict__ refs, const double* __restrict__ bd, int ne,
                        double* __restrict__ known)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne)
        return;
    const NodeRefs& r = refs[i];
    if (r.surface == 1)
        return;
    int blk = r.bcid == 123 ? r.u_block : r.t_block;
    int kn = r.bcid == 123 ? r.u_known : r.t_known;
    double sg = r.bcid == 123 ? r.u_sign : r.t_sign;
    if (!kn)
        return;
    for (int c = 0; c < 3; ++c)
        known[3 * (long)blk + c] = sg * bd[3 * (long)i + c];
}

void LaunchKnown(const NodeRefs* refs, const double* bd, int ne, double* known, long nknown, hipStream_t s)
{
    (void)nknown;
    k_known<<<(ne + 255) / 256, 256, 0, s>>>(refs, bd, ne, known);
}

__global__ void k_scatter(const NodeRefs* __restrict__ refs, const ElemGeo* __restrict__ geo,
                          const double* __restrict__ bd, const double* __restrict__ x, int ne, double* __restrict__ lu,
                          double* __restrict__ lt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne)
        return;
    const NodeRefs& r = refs[i];
    const double* T = r.itf_T;
    double u[3], t[3];
    const double* in = bd + 3 * (long)i;
    if (!r.u_known) {
        const double* v = x + 3 * (long)r.u_block;
        for (int q = 0; q < 3; ++q)
            u[q] = r.u_sign * (T[q * 3 + 0] * v[0] + T[q * 3 + 1] * v[1] + T[q * 3 + 2] * v[2]);
    } else {
        for (int q = 0; q < 3; ++q)
            u[q] = r.u_sign * in[q];
    }
    if (!r.t_known) {
        const double* v = x + 3 * (long)r.t_block;
        for (int q = 0; q < 3; ++q)
            t[q] = r.t_sign * (T[q * 3 + 0] * v[0] + T[q * 3 + 1] * v[1] + T[q * 3 + 2] * v[2]);
    } else {
        for (int q = 0; q < 3; ++q)
            t[q] = r.t_sign * in[q];
    }
    const double* F = geo[i].trans;
    double* ou = lu + 3 * (long)i;
    double* ot = lt + 3 * (long)i;
    ou[0] = F[0] * u[0] + F[3] * u[1] + F[6] * u[2];
    ou[1] = F[1] * u[0] + F[4] * u[1] + F[7] * u[2];
    ou[2] = F[2] * u[0] + F[5] * u[1] + F[8] * u[2];
    ot[0] = F[0] * t[0] + F[3] * t[1] + F[6] * t[2];
    ot[1] = F[1] * t[0] + F[4] * t[1] + F[7] * t[2];
    ot[2] = F[2] * t[0] + F[5] * t[1] + F[8] * t[2];
}

void LaunchScatter(const NodeRefs* refs, const ElemGeo* geo, const double* bd, const double* x, int ne, double* lu,
                   double* lt, hipStream_t s)
{
    k_scatter<<<(ne + 255) / 256, 256, 0, s>>>(refs, geo, bd, x, ne, lu, lt);
}

__global__ void k_state(const NodeRefs* __restrict__ refs, const double* __restrict__ lu,
                        const double* __restrict__ lt, int ne, int K, int nstore, double* __restrict__ gU,
                        double* __restrict__ gT)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne)
        return;
    const NodeRefs& r = refs[i];
    if (r.surface == 1 && !r.ref_side)
        return;
    double* du = gU + ((long)r.hu_block * nstore + K) * 3;
    double* dt = gT + ((long)r.ht_block * nstore + K) * 3;
    for (int c = 0; c < 3; ++c) {
        du[c] = r.hu_sign * lu[3 * (long)i + c];
        dt[c] = r.ht_sign * lt[3 * (long)i + c];
    }
}

void LaunchState(const NodeRefs* refs, const double* lu, const double* lt, int ne, int K, int nstore, double* gU,
                 double* gT, hipStream_t s)
{
    k_state<<<(ne + 255) / 256, 256, 0, s>>>(refs, lu, lt, ne, K, nstore, gU, gT);
}

__global__ void k_zero(double* x, long n)
{
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        x[r] = 0.0;
}

void LaunchZero(double* x, long n, hipStream_t s)
{
    if (n > 0)
        k_zero<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(x, n);
}

}  // namespace tdbem
