// A caller of the reference's multi-domain API written against the C++
// drop-in header (include/tdbem_multidomain.hpp): it reads a reference case
// (positional datacard + multi-domain tail, .model, .bd) the way the
// reference's main does (DBEM1/DBEM.cpp:637-1085,1232), builds the DOF map and
// calls DynaGMRESSolverMultiDomainCCSR.
//
//   rod_dropin DATACARD MODEL BD [--global-E E] [--global-rho RHO] [--host-only]
//
// --host-only stops after GlobalDofMap::Build and a MultiDomainCCSRBuilder
// round trip (no device work).  Otherwise it prints the solver's result and,
// for the last step, the mean x displacement of the BCID-456 elements on the
// end face of largest x (the quantity the reference's analytic validation
// reports, DBEM1/analytic_validation.txt).
#include "tdbem_multidomain.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s DATACARD MODEL BD [--global-E E] [--global-rho RHO] [--host-only]\n", argv[0]);
        return 2;
    }
    double overrideE = -1.0, overrideRho = -1.0;
    bool hostOnly = false;
    for (int i = 4; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--global-E") && i + 1 < argc)
            overrideE = std::atof(argv[++i]);
        else if (!std::strcmp(argv[i], "--global-rho") && i + 1 < argc)
            overrideRho = std::atof(argv[++i]);
        else if (!std::strcmp(argv[i], "--host-only"))
            hostOnly = true;
    }
    // positional datacard values (DBEM1/DBEM.cpp:637-700)
    FILE* card = std::fopen(argv[1], "r");
    if (!card) {
        std::printf("cannot open %s\n", argv[1]);
        return 1;
    }
    char title[256];
    int threads = 0, flagDyna = 0, solver = 0, inMode = 0, batch = 0;
    long nstep = 0, iterations = 0, premax = 0;
    double beta = 0, gap = 0, E = 0, v = 0, rho = 0, amplitude = 0, error = 0;
    if (std::fscanf(card, "%255s %d %d %lf %ld %lf %lf %lf %lf %lf %d %ld %lf %ld %d %d", title, &threads, &flagDyna,
                    &beta, &nstep, &gap, &E, &v, &rho, &amplitude, &solver, &iterations, &error, &premax, &inMode,
                    &batch) != 16) {
        std::printf("datacard %s: positional values missing\n", argv[1]);
        return 1;
    }
    MultiDomainInputConfig cfg;
    int tail = ReadOptionalMultiDomainInput(card, cfg);
    std::fclose(card);
    if (tail < 0)
        return 1;
    if (overrideE > 0)
        E = overrideE;
    if (overrideRho > 0)
        rho = overrideRho;

    long np = 0, ne = 0, ninf = 0;
    if (tdbem_model_counts(argv[2], inMode, &np, &ne, &ninf) != 0) {
        std::printf("%s\n", tdbem_last_error());
        return 1;
    }
    std::vector<double> xyz(3 * (size_t)np);
    std::vector<int> pid(8 * (size_t)ne), bcid((size_t)ne);
    std::vector<double> known(3 * (size_t)ne * (size_t)(nstep + 1));
    if (tdbem_read_model(argv[2], inMode, xyz.data(), pid.data(), nullptr) != 0 ||
        tdbem_read_bd(argv[3], inMode, ne, (int)nstep, nullptr, bcid.data(), known.data()) != 0) {
        std::printf("%s\n", tdbem_last_error());
        return 1;
    }
    // dt = Beta * AveEleSize / c1 of the global material line (DBEM1/DBEM.cpp:912-914,2383-2396)
    double ave = 0.0;
    for (long e = 0; e < ne; ++e) {
        const double* a = &xyz[3 * (size_t)pid[8 * (size_t)e]];
        const double* b = &xyz[3 * (size_t)pid[8 * (size_t)e + 1]];
        ave += std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
    }
    ave /= (double)ne;
    const double G = E / 2.0 / (1.0 + v);
    const double c1 = std::sqrt((2.0 - 2.0 * v) / (1.0 - 2.0 * v) * G / rho);
    const double dt = beta * ave / c1;

    MultiDomainModel model;
    if (!BuildMultiDomainModel(cfg, ne, ne, E, v, rho, dt, model))
        return 1;
    GlobalDofMap dof;
    if (!dof.Build(model, bcid.data(), ne))
        return 1;
    std::printf("dofmap unknown=%ld known=%ld equation=%ld historyU=%ld historyT=%ld\n", dof.unknownBlockCount,
                dof.knownBlockCount, dof.equationBlockCount, dof.historyUBlockCount, dof.historyTBlockCount);
    if (hostOnly) {
        // a builder round trip: the column-major block is stored transposed times the sign
        MultiDomainCCSRBuilder b(dof.equationBlockCount, dof.unknownBlockCount);
        const double blk[9] = {1, 2, 3, 4, 5, 6, 7, 8, 9};
        b.AddBlock(0, 0, -1.0, blk);
        CCSRMat m;
        b.Build(m);
        std::printf("builder nnz=%ld first=%g,%g,%g\n", m.NonZeroBlocks(), m.value[0], m.value[1], m.value[2]);
        return 0;
    }
    std::vector<BoundaryValue> bd((size_t)nstep + 1);
    for (long k = 0; k <= nstep; ++k) {
        bd[(size_t)k].init(ne);
        std::memcpy(bd[(size_t)k].lu.data(), &known[3 * (size_t)ne * (size_t)k], sizeof(double) * 3 * (size_t)ne);
    }
    MultiDomainMesh mesh{np, xyz.data(), ne, pid.data(), bcid.data(), gap};
    const double MaxLe = 1000.0;  // DBEM1/DBEM.cpp:1195
    int result = DynaGMRESSolverMultiDomainCCSR(mesh, bd.data(), model, iterations, error, premax, nstep, MaxLe);
    // loaded end: BCID 456 elements whose centre has the largest x
    double xmax = -1e300;
    std::vector<double> cx((size_t)ne);
    for (long e = 0; e < ne; ++e) {
        double s = 0.0;
        for (int j = 0; j < 4; ++j)
            s += xyz[3 * (size_t)pid[8 * (size_t)e + (size_t)j]];
        cx[(size_t)e] = 0.25 * s;
        if (cx[(size_t)e] > xmax)
            xmax = cx[(size_t)e];
    }
    double sum = 0.0;
    int cnt = 0;
    for (long e = 0; e < ne; ++e)
        if (bcid[(size_t)e] == 456 && std::fabs(cx[(size_t)e] - xmax) < 1e-8) {
            sum += bd[(size_t)nstep].lu[3 * (size_t)e];
            ++cnt;
        }
    std::printf("result=%d elements=%d mean_ux_last=%.17g\n", result, cnt, cnt ? sum / cnt : 0.0);
    return result == 1 ? 0 : 3;
}
