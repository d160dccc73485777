// Quadrature tables of the constant square element, built on the host once
// per gap value and uploaded to the device.
//
// Values follow DBEM1/gaussnumematrix.cpp:4-85 (Gauss-Legendre nodes and
// weights of orders 3 and 4, read by the reference from the 15-decimal text
// tables DBEM1/gaussposition and DBEM1/gaussweight) and DBEM1/quadinfo.cpp
// (patch layout, N0 regular weights, constant-element polar tables).
#include "common.hpp"

#include <cstdlib>
#include <cstring>

#pragma clang fp contract(off)

namespace tdbem {

namespace {

// sin and cos are evaluated by separate libm calls, never fused into one
// sincos call, so every angle gets exactly the values separate calls give.
__attribute__((noinline)) double Cos(double x) { return cos(x); }
__attribute__((noinline)) double Sin(double x) { return sin(x); }

// Orders 3 and 4 of the reference's Gauss tables, as the decimal strings the
// reference parses (w3 = 0.555555555555555 is deliberately not 5/9).
const char* const kGaussPos3[3] = {"-0.774596669241484", "0.000000000000000", "0.774596669241484"};
const char* const kGaussW3[3] = {"0.555555555555555", "0.888888888888889", "0.555555555555555"};
const char* const kGaussPos4[4] = {"-0.861136311594052", "-0.339981043584856", "0.339981043584856",
                                   "0.861136311594052"};
const char* const kGaussW4[4] = {"0.347854845137454", "0.652145154862546", "0.652145154862546",
                                 "0.347854845137454"};

// Serendipity shape function N0 on nodes at +-gap (DBEM1/quadinfo.h:149-172).
double ShapeN0(const double (*lc)[2], double s1, double s2)
{
    double lx = lc[1][0] - lc[0][0];
    double ly = lc[3][1] - lc[0][1];
    double lxly = lx * ly;
    double hx = 0.25 * lx * lxly;
    double hy = 0.25 * ly * lxly;
    double n0 = (s1 - lc[1][0]) * (s2 - lc[3][1]) / lxly;
    double n4 = (s1 - lc[0][0]) * (s1 - lc[1][0]) * (s2 - lc[6][1]) / (hx);
    double n7 = (s1 - lc[5][0]) * (s2 - lc[0][1]) * (s2 - lc[3][1]) / (hy);
    n0 -= 0.5 * (n4 + n7);
    return n0;
}

// Constant-element polar sectors around the centre (DBEM1/quadinfo.cpp:5-48).
double SectorStart(int a)
{
    switch (a) {
    case 0: return -0.75 * kPi;
    case 1: return -0.25 * kPi;
    case 2: return 0.25 * kPi;
    case 3: return 0.75 * kPi;
    default: return 0.0;
    }
}
double SectorSpan() { return 0.5 * kPi; }
double SectorAngle(double localTheta, int a) { return SectorStart(a) + 0.5 * (localTheta + 1.0) * SectorSpan(); }
double SectorEdge(double theta, int a)
{
    switch (a) {
    case 0: return -1.0 / Sin(theta);
    case 1: return 1.0 / Cos(theta);
    case 2: return 1.0 / Sin(theta);
    case 3: return -1.0 / Cos(theta);
    default: return 0.0;
    }
}
double SectorRadius(double localRho, double theta, int a) { return 0.5 * (localRho + 1.0) * SectorEdge(theta, a); }

}  // namespace

void BuildQuadTables(QuadTables& q, double gap)
{
    std::memset(&q, 0, sizeof(q));
    double p4[4], w4[4], p3[3], w3[3];
    for (int i = 0; i < 4; ++i) {
        p4[i] = std::strtod(kGaussPos4[i], nullptr);
        w4[i] = std::strtod(kGaussW4[i], nullptr);
    }
    for (int i = 0; i < 3; ++i) {
        p3[i] = std::strtod(kGaussPos3[i], nullptr);
        w3[i] = std::strtod(kGaussW3[i], nullptr);
    }
    q.gap = gap;
    const double lcv[8][2] = {{-gap, -gap}, {gap, -gap}, {gap, gap}, {-gap, gap},
                              {0.0, -gap},  {gap, 0.0},  {0.0, gap}, {-gap, 0.0}};
    std::memcpy(q.lc, lcv, sizeof(lcv));

    // 4x4 regular rule: point m = i*4 + j at (p[j], p[i]), weight w[j]*w[i].
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            int m = i * 4 + j;
            q.rp[m][0] = p4[j];
            q.rp[m][1] = p4[i];
            double ra = w4[j] * w4[i];
            q.nrgv0[m] = ShapeN0(q.lc, q.rp[m][0], q.rp[m][1]) * ra;
        }

    // 6x6 patches with 3x3 points each.
    double pwrp[9][2], pwra[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int m = i * 3 + j;
            pwrp[m][0] = p3[j];
            pwrp[m][1] = p3[i];
            pwra[m] = w3[j] * w3[i];
        }
    {
        double L = 2.0 / kNPW;
        for (int i = 0; i < kNPW; ++i)
            for (int j = 0; j < kNPW; ++j) {
                int p = i * kNPW + j;
                q.cenpw[p][0] = -1.0 + L / 2.0 + L * j;
                q.cenpw[p][1] = -1.0 + L / 2.0 + L * i;
                for (int m = 0; m < kNPWG2; ++m) {
                    q.lpospw[p][m][0] = q.cenpw[p][0] + pwrp[m][0] / kNPW;
                    q.lpospw[p][m][1] = q.cenpw[p][1] + pwrp[m][1] / kNPW;
                }
            }
        for (int p = 0; p < kNPW2; ++p)
            for (int m = 0; m < kNPWG2; ++m)
                q.pwgv[p][m] = pwra[m] * (L / 2.0) * (L / 2.0);
    }

    // Static strong-singular part: 4 sectors x 4x4 polar points.
    double sa[16], sp[16][2];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            int m = i * 4 + j;
            sp[m][0] = p4[j];
            sp[m][1] = p4[i];
            sa[m] = w4[j] * w4[i];
        }
    for (int a = 0; a < 4; ++a) {
        for (int n = 0; n < kNSG2; ++n) {
            double th = SectorAngle(sp[n][0], a);
            double fa = 0.25 * SectorEdge(th, a) * SectorSpan();
            double rho = SectorRadius(sp[n][1], th, a);
            q.secconstpos[a][0][n] = rho * Cos(th);
            q.secconstpos[a][1][n] = rho * Sin(th);
            q.secconstval[a][n] = rho * fa * sa[n];
            q.constfm1secval[a][n] = (rho * fa * sa[n]) / rho / rho;
        }
        for (int n = 0; n < kNSG; ++n) {
            double th = SectorAngle(p4[n], a);
            double rho = SectorRadius(1.0, th, a);
            q.constfm1lineval[a][n] = 0.5 * SectorSpan() * log(rho) * w4[n];
            q.linecos[a][n] = Cos(th);
            q.linesin[a][n] = Sin(th);
        }
    }

    // Dynamic singular part: 4 sectors x 20x20 polar patches x 3x3 points.
    {
        double L = 2.0 / 20;
        double wpw[9];
        for (int m = 0; m < 9; ++m)
            wpw[m] = pwra[m] * (L / 2.0) * (L / 2.0);
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 20; ++i)
                for (int j = 0; j < 20; ++j) {
                    int p = i * 20 + j;
                    double c0 = -1.0 + L / 2.0 + L * j;
                    double c1 = -1.0 + L / 2.0 + L * i;
                    for (int m = 0; m < 9; ++m) {
                        double lt = c0 + pwrp[m][0] / 20;
                        double lr = c1 + pwrp[m][1] / 20;
                        double th = SectorAngle(lt, a);
                        double fa = 0.25 * SectorEdge(th, a) * SectorSpan();
                        double rho = SectorRadius(lr, th, a);
                        q.secconstpospw[a][0][p][m] = rho * Cos(th);
                        q.secconstpospw[a][1][p][m] = rho * Sin(th);
                        q.secconstvalpw[a][p][m] = rho * fa * wpw[m];
                    }
                }
    }
}

}  // namespace tdbem
