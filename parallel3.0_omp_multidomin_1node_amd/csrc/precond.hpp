// Leaf block preconditioner of the multi-domain GMRES (host build).
//
// Octree over element centres with at most maxLeaf points per leaf
// (DBEM1/tree.cpp:12-165,242-323; DBEM1/geo.cpp:8-40,82-113); each leaf holds
// the dense block A(row = leaf point, col = preferred unknown of the leaf
// point) inverted by full-pivot Gauss-Jordan (DBEM1/lin_comp.cpp:83-190),
// as BuildMappedLeafPreConditioner does (DBEM1/multidomain.cpp:3493-3565).
#pragma once

#include "model.hpp"

#include <vector>

namespace tdbem {

struct LeafPreconditioner {
    long leafCount = 0;
    std::vector<long> leafSize;   // points per leaf
    std::vector<long> leafBegin;  // first position in repid
    std::vector<long> repid;      // point ids in leaf order
    std::vector<long> matOffset;  // offset of each leaf's (3c x 3c) row-major inverse
    std::vector<double> inv;      // concatenated inverses
};

// Full-pivot Gauss-Jordan inversion in place of an n x n row-major matrix.
// Returns -1 when a pivot is <= 1e-10 (matrix left partially reduced, as the
// reference leaves it).
int InvMat(double* A, long n);

// With groups (contiguous point ranges [groups[g], groups[g+1]) of
// independent models), one octree per group over that group's points, the
// leaves of group g after those of the groups before it; without groups (or
// with one) one octree over all points.
bool BuildMappedLeafPreConditioner(const double* centres, long pointCount, long maxLeaf, const CCSRMat& A,
                                   const GlobalDofMap& dof, LeafPreconditioner& pre,
                                   const std::vector<long>* groups = nullptr);

}  // namespace tdbem
