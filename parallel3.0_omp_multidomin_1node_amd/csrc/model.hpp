// Host-side mirror of the reference's multi-domain model, DOF map, block
// builder and block-CCSR containers (DBEM1/multidomain.h:14-160,
// DBEM1/mat_vec.h:220-298), with the same names, argument meaning and error
// behaviour (bool / int status, printf diagnostics).
#pragma once

#include "common.hpp"

#include <array>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

namespace tdbem {

struct Domain {
    int id = 0;
    DynaMat mat{};
    double E = 0.0, v = 0.0, Rou = 0.0;  // the material mat was built from
    long eleBegin = 0;
    long eleCount = 0;
    long nodeBegin = 0;
    long nodeCount = 0;
    std::vector<long> elementIds;
};

enum SurfaceType { SurfaceOuter = 0, SurfaceInterface = 1 };

struct InterfacePair {
    int domainA = 0;
    int domainB = 0;
    long eleA = 0;
    long eleB = 0;
    int normalSign = -1;
};

struct DomainMaterialInput {
    int id;
    double E, v, Rou;
};

// Multi-domain tail of the datacard (DBEM1/multidomain.h:58-67).
struct MultiDomainInputConfig {
    bool enabled = false;
    int domainCount = 1;
    bool hasExplicitInterfaces = false;
    std::vector<DomainMaterialInput> materials;
    std::vector<InterfacePair> interfaces;
};

class MultiDomainModel {
public:
    bool enabled = false;
    std::vector<Domain> domains;
    std::vector<InterfacePair> interfaces;
    long elementCount = 0;
    long nodeCount = 0;

    bool IsActive() const { return enabled; }
    int DomainCount() const { return (int)domains.size(); }
    bool Validate() const;
};

// BuildMultiDomainModel (DBEM1/multidomain.cpp:1364-1480): contiguous equal
// element split, per-domain DynaMat with the global dt, id normalisation.
int BuildMultiDomainModel(const MultiDomainInputConfig& config, long eleNum, long nodeNum, double defaultE,
                          double defaultV, double defaultRou, double dt, MultiDomainModel& model);

struct VariableRef {
    long globalBlock = -1;
    double sign = 1.0;
    bool known = true;
    VariableRef() = default;
    VariableRef(long b, double s, bool k) : globalBlock(b), sign(s), known(k) {}
};

// GlobalDofMap (DBEM1/multidomain.h:103-137, DBEM1/multidomain.cpp:777-1026).
class GlobalDofMap {
public:
    long unknownBlockCount = 0;
    long knownBlockCount = 0;
    long equationBlockCount = 0;
    long outerDisplacementUnknownBlocks = 0;
    long outerTractionUnknownBlocks = 0;
    long interfaceDisplacementUnknownBlocks = 0;
    long interfaceTractionUnknownBlocks = 0;
    long historyUBlockCount = 0;
    long historyTBlockCount = 0;

    // bcid[e] of every element; only 123 and 456 are supported.  With
    // groups (contiguous element ranges [groups[g], groups[g+1]) of
    // independent models placed one after another), the blocks are allocated
    // group by group, each group's interface pairs (by eleA) first and then
    // its other elements, so a group's unknown, known and history blocks are
    // contiguous and in the order the group alone would give them; without
    // groups (or with one) this is the reference's allocation order.
    bool Build(const MultiDomainModel& model, const int* bcid, long eleNum,
               const std::vector<long>* groups = nullptr);
    VariableRef GetU(int domainId, long node) const { return At(m_uRefs, node); }
    VariableRef GetT(int domainId, long node) const { return At(m_tRefs, node); }
    VariableRef GetHistoryU(int domainId, long node) const { return At(m_historyURefs, node); }
    VariableRef GetHistoryT(int domainId, long node) const { return At(m_historyTRefs, node); }
    long GetEquationRow(int domainId, long node) const
    {
        return node >= 0 && node < (long)m_equationRows.size() ? m_equationRows[(size_t)node] : -1;
    }
    long GetPreferredUnknownBlockForRow(long row) const
    {
        return row >= 0 && row < (long)m_preferred.size() ? m_preferred[(size_t)row] : -1;
    }
    bool Validate() const;

private:
    static VariableRef At(const std::vector<VariableRef>& v, long i)
    {
        return i >= 0 && i < (long)v.size() ? v[(size_t)i] : VariableRef();
    }
    std::vector<VariableRef> m_uRefs, m_tRefs, m_historyURefs, m_historyTRefs;
    std::vector<long> m_equationRows, m_preferred;
};

// Block CCSR with row-major 3x3 blocks (DBEM1/mat_vec.h:220-240).
struct CCSRMat {
    long n = 0;
    std::vector<int> row_ptr;
    std::vector<int> col_index;
    std::vector<double> value;  // 9 per block
    long NonZeroBlocks() const { return (long)col_index.size(); }
};

// Symmetric block CCSR storing (b00 b01 b02 b11 b12 b22) (DBEM1/mat_vec.h:278-298).
struct SymCCSRMat {
    long n = 0;
    std::vector<int> row_ptr;
    std::vector<int> col_index;
    std::vector<double> value;  // 6 per block
    long NonZeroBlocks() const { return (long)col_index.size(); }
};

// MultiDomainCCSRBuilder (DBEM1/multidomain.cpp:1101-1303): AddBlock stores
// the transpose of a column-major block times sign, drops blocks whose
// max |.| <= 1e-14 or that are not finite, and erases accumulated entries
// that fall to <= 1e-14.
class MultiDomainCCSRBuilder {
public:
    explicit MultiDomainCCSRBuilder(long blockRows = 0, long blockCols = 0) { Reset(blockRows, blockCols); }
    void Reset(long blockRows, long blockCols);
    void AddBlock(long rowBlock, long colBlock, double sign, const double block9[9]);
    void MergeFrom(const MultiDomainCCSRBuilder& other);
    void Build(CCSRMat& out) const;
    void BuildSym(SymCCSRMat& out, double tolerance, bool* symmetric) const;
    long NonZeroBlocks() const;
    long BlockRows() const { return m_blockRows; }
    long BlockCols() const { return m_blockCols; }
    // blocks AddBlock skipped for a NaN or infinite value since the last Reset
    long SkippedNonFiniteBlocks() const { return m_skippedNonFinite; }

private:
    struct Entry {
        std::array<double, 9> v;
        bool alive;
    };
    long m_blockRows = 0, m_blockCols = 0;
    long m_skippedNonFinite = 0;
    std::unordered_map<uint64_t, Entry> m_blocks;
    void Sorted(std::vector<std::pair<uint64_t, const Entry*>>& out) const;
};

// SparseMul (DBEM1/mat_vec.cpp:2273-2300,2407-2434), 0-based vectors.
void SparseMul(const CCSRMat& M, const double* x, double* b);
void SparseMul(const SymCCSRMat& M, const double* x, double* b);

double BlockMaxAbs9(const double* b);
bool IsFiniteBlock9(const double* b);

}  // namespace tdbem
