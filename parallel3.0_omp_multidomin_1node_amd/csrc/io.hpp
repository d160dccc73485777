// Reference input / output formats around the path (host side).
#pragma once

#include "model.hpp"

#include <string>
#include <vector>

namespace tdbem {

struct ModelData {
    std::vector<double> points;  // 3 per point
    std::vector<long> elePid;    // 8 per element, 0-based
    std::vector<long> infPid;    // 4 per infinite element, 0-based
};

// .model (text, 1-based ids) or its DBEMBIN twin path + ".bin"; mode 0 text,
// 1 binary required, 2 binary when it exists.
ModelData ReadModelFile(const std::string& path, int mode);
// .bd: BCIDs then (nstep + 1) x ne triplets; expectedBcid (may be null) must
// match when given.
void ReadBdFile(const std::string& path, int mode, long ne, int nstep, const int* expectedBcid, int* bcid,
                double* lu);
void WriteModelBin(const std::string& path, long np, const double* points, long ne, const long* elePid, long ninf,
                   const long* infPid);
void WriteBdBin(const std::string& path, long ne, int nstep, const int* bcid, const double* lu);

struct StructureStats {
    long zeroRows = 0, zeroCols = 0, nonFinite = 0, diagonal = 0, invertibleDiagonal = 0;
    bool valid = false;
};
StructureStats DiagnoseStructure(const CCSRMat& M);
long InterfaceDofFailures(const GlobalDofMap& dof, const MultiDomainModel& model);

struct ValidationInput {
    std::string dir;
    long ne = 0, nstep = 0, solved = 0;
    double dt = 0.0, assemblyTime = 0.0, gmresTime = 0.0, avgMTSBlocks = 0.0, avgMGSBlocks = 0.0;
    int solver = 2;
    const double* nodes = nullptr;  // 3 per element (collocation node = centre)
    const double* lu = nullptr;     // (nstep + 1) x 3 ne, global frame
    const double* lt = nullptr;
    const int* iters = nullptr;     // per solved step
    const int* flags = nullptr;
    std::vector<int> domainId, surface, bcid;
    const MultiDomainModel* model = nullptr;
    const GlobalDofMap* dof = nullptr;
    const CCSRMat* A = nullptr;
};
void WriteValidationFiles(const ValidationInput& in);

}  // namespace tdbem
