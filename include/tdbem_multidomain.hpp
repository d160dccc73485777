/*
 * tdbem_multidomain.hpp -- C++ drop-in for the reference's multi-domain API,
 * header-only over the C-ABI of tdbem_capi.h (link lib/libtdbem.so).
 *
 * Same names, argument meaning and error behaviour as the reference's
 * DBEM1/multidomain.h:14-223 and DBEM1/mat_vec.h:220-298,343-344 for the
 * types and functions of the GMRES hot path:
 *   MultiDomainInputConfig, DomainMaterialInput, InterfacePair
 *                               (DBEM1/multidomain.h:30-67)
 *   MultiDomainModel, BuildMultiDomainModel (multidomain.h:75-91,179-186)
 *   VariableRef, GlobalDofMap   (multidomain.h:93-137)
 *   MultiDomainCCSRBuilder      (multidomain.h:139-160)
 *   CCSRMat, SymCCSRMat, SparseMul (mat_vec.h:220-298,343-344; 0-based vectors)
 *   BoundaryValue               (dsquareelement.h:64-80; std::vector storage)
 *   DynaGMRESSolverMultiDomainCCSR (multidomain.h:212-221)
 *
 * One difference is forced by the boundary: the reference hands the solver
 * its DSquareElement array (geometry, BCID) and the element point ids; here
 * the geometry travels as the point list and element point ids the reference
 * read from the .model file (MultiDomainMesh), and GlobalDofMap::Build takes
 * the per-element BCIDs.  The element frames, singular tables and every
 * matrix are built by the library from those.
 *
 * Status values follow the reference: bool for Build, 1 / -1 for the solver
 * (1 when the last solved step's GMRES flag is 0), diagnostics via printf.
 */
#ifndef TDBEM_MULTIDOMAIN_HPP
#define TDBEM_MULTIDOMAIN_HPP

#include "tdbem_capi.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>

struct DomainMaterialInput {
    int id;
    double E;
    double v;
    double Rou;
};

struct InterfacePair {
    int domainA;
    int domainB;
    long eleA;
    long eleB;
    int normalSign;
    InterfacePair() : domainA(0), domainB(0), eleA(0), eleB(0), normalSign(-1) {}
};

/* Multi-domain tail of the datacard (ReadOptionalMultiDomainInput fills it). */
struct MultiDomainInputConfig {
    bool enabled;
    int domainCount;
    bool hasExplicitInterfaces;
    std::vector<DomainMaterialInput> materials;
    std::vector<InterfacePair> interfaces;
    MultiDomainInputConfig() : enabled(false), domainCount(1), hasExplicitInterfaces(false) {}
};

/* ReadOptionalMultiDomainInput (DBEM1/multidomain.cpp:1305-1362): reads the
 * datacard tail after the positional values.  Returns 0 when there is no
 * tail (one domain), -1 for an invalid DomainCount, material line,
 * InterfaceCount or interface line, 1 otherwise; ids are kept as written
 * (0- or 1-based, the library normalises them). */
inline int ReadOptionalMultiDomainInput(FILE* input, MultiDomainInputConfig& config)
{
    config = MultiDomainInputConfig();
    if (!input)
        return 0;
    int domainCount = 1;
    if (std::fscanf(input, "%d", &domainCount) != 1)
        return 0;
    if (domainCount <= 0) {
        std::printf("Invalid DomainCount %d.\n", domainCount);
        return -1;
    }
    config.enabled = true;
    config.domainCount = domainCount;
    config.materials.resize((size_t)domainCount);
    for (int i = 0; i < domainCount; ++i) {
        DomainMaterialInput& m = config.materials[(size_t)i];
        if (std::fscanf(input, "%d %lf %lf %lf", &m.id, &m.E, &m.v, &m.Rou) != 4) {
            std::printf("Invalid material line for domain %d.\n", i);
            return -1;
        }
    }
    int interfaceCount = 0;
    if (std::fscanf(input, "%d", &interfaceCount) == 1) {
        if (interfaceCount < 0) {
            std::printf("Invalid InterfaceCount %d.\n", interfaceCount);
            return -1;
        }
        config.hasExplicitInterfaces = true;
        config.interfaces.resize((size_t)interfaceCount);
        for (int i = 0; i < interfaceCount; ++i) {
            InterfacePair& f = config.interfaces[(size_t)i];
            if (std::fscanf(input, "%d %ld %d %ld %d", &f.domainA, &f.eleA, &f.domainB, &f.eleB, &f.normalSign) != 5) {
                std::printf("Invalid interface line %d.\n", i);
                return -1;
            }
        }
    }
    return 1;
}

/* The case description the library builds its model from: the datacard tail,
 * the global material line and dt (DBEM1/DBEM.cpp:912-914). */
class MultiDomainModel {
public:
    bool enabled;
    MultiDomainInputConfig config;
    double defaultE, defaultV, defaultRou, dt;
    long elementCount;
    long nodeCount;

    MultiDomainModel() : enabled(false), defaultE(0), defaultV(0), defaultRou(0), dt(0), elementCount(0), nodeCount(0) {}
    bool IsActive() const { return enabled; }
    int DomainCount() const { return config.enabled ? config.domainCount : 1; }

    /* flat rows of the C-ABI: (id, E, v, rho) per domain, (dA, eA, dB, eB, sign) per interface */
    void Rows(std::vector<double>& mats, std::vector<long>& itf) const
    {
        mats.clear();
        itf.clear();
        for (const DomainMaterialInput& m : config.materials) {
            mats.push_back(m.id);
            mats.push_back(m.E);
            mats.push_back(m.v);
            mats.push_back(m.Rou);
        }
        for (const InterfacePair& f : config.interfaces) {
            itf.push_back(f.domainA);
            itf.push_back(f.eleA);
            itf.push_back(f.domainB);
            itf.push_back(f.eleB);
            itf.push_back(f.normalSign);
        }
        if (mats.empty())
            mats.push_back(0.0);
        if (itf.empty())
            itf.push_back(0);
    }
};

/* BuildMultiDomainModel (DBEM1/multidomain.cpp:1364-1480).  Records the
 * inputs; the library validates them (domain split, ids, interface pairs) when
 * the DOF map or the solver context is built.  Returns 1, or 0 for a domain
 * count the tail does not cover. */
inline int BuildMultiDomainModel(const MultiDomainInputConfig& config, long eleNum, long nodeNum, double defaultE,
                                 double defaultV, double defaultRou, double dt, MultiDomainModel& model)
{
    if (eleNum <= 0 || (config.enabled && (config.domainCount <= 0 || config.domainCount > (int)eleNum))) {
        std::printf("BuildMultiDomainModel failed: %d domains for %ld elements.\n", config.domainCount, eleNum);
        return 0;
    }
    model.enabled = true;
    model.config = config;
    model.defaultE = defaultE;
    model.defaultV = defaultV;
    model.defaultRou = defaultRou;
    model.dt = dt;
    model.elementCount = eleNum;
    model.nodeCount = nodeNum;
    return 1;
}

struct VariableRef {
    long globalBlock;
    double sign;
    bool known;
    VariableRef() : globalBlock(-1), sign(1.0), known(true) {}
    VariableRef(long block, double s, bool isKnown) : globalBlock(block), sign(s), known(isKnown) {}
};

/* GlobalDofMap (DBEM1/multidomain.h:103-137): the host DOF map of the
 * library (tdbem_dofmap_build).  Nodes are elements (one collocation node per
 * element), so localNodeOrEle is the global element id. */
class GlobalDofMap {
public:
    long unknownBlockCount;
    long knownBlockCount;
    long equationBlockCount;
    long outerDisplacementUnknownBlocks;
    long outerTractionUnknownBlocks;
    long interfaceDisplacementUnknownBlocks;
    long interfaceTractionUnknownBlocks;
    long historyUBlockCount;
    long historyTBlockCount;

    GlobalDofMap()
        : unknownBlockCount(0), knownBlockCount(0), equationBlockCount(0), outerDisplacementUnknownBlocks(0),
          outerTractionUnknownBlocks(0), interfaceDisplacementUnknownBlocks(0), interfaceTractionUnknownBlocks(0),
          historyUBlockCount(0), historyTBlockCount(0)
    {
    }

    bool Build(const MultiDomainModel& model, const int* bcid, long eleNum)
    {
        std::vector<double> mats;
        std::vector<long> itf;
        model.Rows(mats, itf);
        const size_t n = (size_t)(eleNum > 0 ? eleNum : 0);
        std::vector<long> ub(n), tb(n), hub(n), htb(n), pref(n);
        std::vector<double> us(n), ts(n), hus(n), hts(n);
        std::vector<int> uk(n), tk(n);
        long c[9];
        const MultiDomainInputConfig& cfg = model.config;
        int rc = tdbem_dofmap_build(eleNum, bcid, cfg.enabled ? cfg.domainCount : 0, mats.data(),
                                    cfg.hasExplicitInterfaces ? (int)cfg.interfaces.size() : -1, itf.data(), c,
                                    ub.data(), us.data(), uk.data(), tb.data(), ts.data(), tk.data(), hub.data(),
                                    hus.data(), htb.data(), hts.data(), pref.data());
        if (rc != 0) {
            std::printf("%s\n", tdbem_last_error());
            return false;
        }
        unknownBlockCount = c[0];
        knownBlockCount = c[1];
        equationBlockCount = c[2];
        outerDisplacementUnknownBlocks = c[3];
        outerTractionUnknownBlocks = c[4];
        interfaceDisplacementUnknownBlocks = c[5];
        interfaceTractionUnknownBlocks = c[6];
        historyUBlockCount = c[7];
        historyTBlockCount = c[8];
        m_u.resize(n);
        m_t.resize(n);
        m_hu.resize(n);
        m_ht.resize(n);
        for (size_t i = 0; i < n; ++i) {
            m_u[i] = VariableRef(ub[i], us[i], uk[i] != 0);
            m_t[i] = VariableRef(tb[i], ts[i], tk[i] != 0);
            m_hu[i] = VariableRef(hub[i], hus[i], false);
            m_ht[i] = VariableRef(htb[i], hts[i], false);
        }
        m_pref = pref;
        return true;
    }
    VariableRef GetU(int, long node) const { return At(m_u, node); }
    VariableRef GetT(int, long node) const { return At(m_t, node); }
    VariableRef GetHistoryU(int, long node) const { return At(m_hu, node); }
    VariableRef GetHistoryT(int, long node) const { return At(m_ht, node); }
    long GetEquationRow(int, long node) const { return node >= 0 && node < (long)m_u.size() ? node : -1; }
    long GetPreferredUnknownBlockForRow(long row) const
    {
        return row >= 0 && row < (long)m_pref.size() ? m_pref[(size_t)row] : -1;
    }

private:
    static VariableRef At(const std::vector<VariableRef>& v, long i)
    {
        return i >= 0 && i < (long)v.size() ? v[(size_t)i] : VariableRef();
    }
    std::vector<VariableRef> m_u, m_t, m_hu, m_ht;
    std::vector<long> m_pref;
};

/* Block CCSR, row-major 3x3 blocks (DBEM1/mat_vec.h:220-240). */
struct CCSRMat {
    long n;
    std::vector<int> row_ptr;
    std::vector<int> col_index;
    std::vector<double> value;
    CCSRMat() : n(0) {}
    long NonZeroBlocks() const { return (long)col_index.size(); }
};

/* Symmetric block CCSR storing (b00 b01 b02 b11 b12 b22) (DBEM1/mat_vec.h:278-298). */
struct SymCCSRMat {
    long n;
    std::vector<int> row_ptr;
    std::vector<int> col_index;
    std::vector<double> value;
    SymCCSRMat() : n(0) {}
    long NonZeroBlocks() const { return (long)col_index.size(); }
};

/* SparseMul (DBEM1/mat_vec.cpp:2273-2300,2407-2434): b = M x, 0-based. */
inline void SparseMul(const CCSRMat& M, const double* x, double* b)
{
    for (long i = 0; i < M.n; ++i) {
        double b0 = 0.0, b1 = 0.0, b2 = 0.0;
        for (int j = M.row_ptr[(size_t)i]; j < M.row_ptr[(size_t)i + 1]; ++j) {
            const double* v = &M.value[(size_t)j * 9];
            const double* xc = &x[3 * (size_t)M.col_index[(size_t)j]];
            b0 += v[0] * xc[0] + v[1] * xc[1] + v[2] * xc[2];
            b1 += v[3] * xc[0] + v[4] * xc[1] + v[5] * xc[2];
            b2 += v[6] * xc[0] + v[7] * xc[1] + v[8] * xc[2];
        }
        b[3 * i] = b0;
        b[3 * i + 1] = b1;
        b[3 * i + 2] = b2;
    }
}

inline void SparseMul(const SymCCSRMat& M, const double* x, double* b)
{
    for (long i = 0; i < M.n; ++i) {
        double b0 = 0.0, b1 = 0.0, b2 = 0.0;
        for (int j = M.row_ptr[(size_t)i]; j < M.row_ptr[(size_t)i + 1]; ++j) {
            const double* v = &M.value[(size_t)j * 6];
            const double* xc = &x[3 * (size_t)M.col_index[(size_t)j]];
            b0 += v[0] * xc[0] + v[1] * xc[1] + v[2] * xc[2];
            b1 += v[1] * xc[0] + v[3] * xc[1] + v[4] * xc[2];
            b2 += v[2] * xc[0] + v[4] * xc[1] + v[5] * xc[2];
        }
        b[3 * i] = b0;
        b[3 * i + 1] = b1;
        b[3 * i + 2] = b2;
    }
}

/* MultiDomainCCSRBuilder (DBEM1/multidomain.h:139-160): the library's host
 * builder (tdbem_builder_*), owned here. */
class MultiDomainCCSRBuilder {
public:
    explicit MultiDomainCCSRBuilder(long blockRows = 0, long blockCols = 0) : m_b(nullptr), m_rows(0), m_cols(0)
    {
        Reset(blockRows, blockCols);
    }
    ~MultiDomainCCSRBuilder() { tdbem_builder_destroy(m_b); }
    MultiDomainCCSRBuilder(const MultiDomainCCSRBuilder&) = delete;
    MultiDomainCCSRBuilder& operator=(const MultiDomainCCSRBuilder&) = delete;

    void Reset(long blockRows, long blockCols)
    {
        tdbem_builder_destroy(m_b);
        m_b = nullptr;
        m_rows = blockRows;
        m_cols = blockCols;
        if (tdbem_builder_create(&m_b, blockRows, blockCols) != 0)
            std::printf("%s\n", tdbem_last_error());
    }
    void AddBlock(long rowBlock, long colBlock, double sign, const double block9[9])
    {
        tdbem_builder_add_block(m_b, rowBlock, colBlock, sign, block9);
    }
    void MergeFrom(const MultiDomainCCSRBuilder& other) { tdbem_builder_merge(m_b, other.m_b); }
    void Build(CCSRMat& out) const
    {
        long nnz = NonZeroBlocks();
        out.n = m_rows;
        out.row_ptr.assign((size_t)m_rows + 1, 0);
        out.col_index.assign((size_t)nnz, 0);
        out.value.assign(9 * (size_t)nnz, 0.0);
        tdbem_builder_build(m_b, out.row_ptr.data(), out.col_index.data(), out.value.data());
    }
    void BuildSym(SymCCSRMat& out, double tolerance, bool* symmetric) const
    {
        long nnz = NonZeroBlocks();
        int sym = 0;
        out.n = m_rows;
        out.row_ptr.assign((size_t)m_rows + 1, 0);
        out.col_index.assign((size_t)nnz, 0);
        out.value.assign(6 * (size_t)nnz, 0.0);
        tdbem_builder_build_sym(m_b, tolerance, out.row_ptr.data(), out.col_index.data(), out.value.data(), &sym);
        if (symmetric)
            *symmetric = sym != 0;
    }
    long NonZeroBlocks() const
    {
        long nnz = 0;
        tdbem_builder_info(m_b, &nnz, nullptr);
        return nnz;
    }
    long SkippedNonFiniteBlocks() const
    {
        long skipped = 0;
        tdbem_builder_info(m_b, nullptr, &skipped);
        return skipped;
    }
    long BlockRows() const { return m_rows; }
    long BlockCols() const { return m_cols; }

private:
    tdbem_builder* m_b;
    long m_rows, m_cols;
};

/* Per-step boundary values, 3 per element: lu in = the .bd triplets (element
 * frame), out = displacement; lt out = traction (global frame, as
 * AllLocToAbs leaves them). */
class BoundaryValue {
public:
    std::vector<double> lu;
    std::vector<double> lt;
    int init(long NodeNum)
    {
        lu.assign(3 * (size_t)NodeNum, 0.0);
        lt.assign(3 * (size_t)NodeNum, 0.0);
        return 1;
    }
};

/* What the reference's DSquareElement array and element point ids carry into
 * the solver: points[3 * pointCount], elePid[8 * eleCount] (0-based, corners
 * then midsides), bcid[eleCount] (123 / 456), and the datacard's gap. */
struct MultiDomainMesh {
    long pointCount;
    const double* points;
    long eleCount;
    const int* elePid;
    const int* bcid;
    double gap;
};

/* DynaGMRESSolverMultiDomainCCSR (DBEM1/multidomain.cpp:3664-3926): step-0
 * assembly, history assembly for MaxN steps (MaxN = NStep unless the slowest
 * domain's S front passes MaxLength first, 3718-3739), then the time loop with
 * GMRES per step; bd[0..NStep] in: lu triplets, out: lu / lt.  Returns 1 when
 * the last solved step's GMRES flag is 0, -1 otherwise or on failure. */
inline int DynaGMRESSolverMultiDomainCCSR(const MultiDomainMesh& mesh, BoundaryValue* bd, const MultiDomainModel& model,
                                          long iterations, double error, long maxleafpointnum, long NStep,
                                          double MaxLength)
{
    if (!model.IsActive() || !bd || NStep < 1 || !mesh.points || !mesh.elePid || !mesh.bcid) {
        std::printf("MultiDomain solver called with inactive model.\n");
        return -1;
    }
    const long ne = mesh.eleCount;
    // MaxN from the smallest C2 * dt over the domains (defaults for a domain without a material row)
    double minC2Dt = std::numeric_limits<double>::max();
    const int nd = model.DomainCount();
    for (int d = 0; d < nd; ++d) {
        double E = model.defaultE, v = model.defaultV, rho = model.defaultRou;
        if (d < (int)model.config.materials.size()) {
            E = model.config.materials[(size_t)d].E;
            v = model.config.materials[(size_t)d].v;
            rho = model.config.materials[(size_t)d].Rou;
        }
        double c2dt = std::sqrt(E / (2.0 * (1.0 + v)) / rho) * model.dt;
        if (c2dt > 0.0 && c2dt < minC2Dt)
            minC2Dt = c2dt;
    }
    if (minC2Dt == std::numeric_limits<double>::max()) {
        std::printf("MultiDomain MaxN failed: no valid domain C2Dt.\n");
        return -1;
    }
    long MaxN = NStep;
    for (long i = 0; i <= NStep; ++i)
        if (minC2Dt * (double)i > MaxLength) {
            MaxN = i;
            break;
        }
    if (MaxN > NStep)
        MaxN = NStep;
    if (MaxN < 1)
        MaxN = 1;
    std::vector<double> mats;
    std::vector<long> itf;
    model.Rows(mats, itf);
    const MultiDomainInputConfig& cfg = model.config;
    tdbem_ctx* ctx = nullptr;
    if (tdbem_create(&ctx, mesh.pointCount, mesh.points, ne, mesh.elePid, mesh.bcid, cfg.enabled ? cfg.domainCount : 0,
                     mats.data(), cfg.hasExplicitInterfaces ? (int)cfg.interfaces.size() : -1, itf.data(),
                     model.defaultE, model.defaultV, model.defaultRou, model.dt, mesh.gap) != 0) {
        std::printf("%s\n", tdbem_last_error());
        return -1;
    }
    const size_t n3 = 3 * (size_t)ne, total = n3 * (size_t)(NStep + 1);
    std::vector<double> known(total), lu(total), lt(total);
    for (long k = 0; k <= NStep; ++k) {
        if (bd[k].lu.size() < n3) {
            std::printf("MultiDomain solver: boundary values of step %ld are not initialised.\n", k);
            tdbem_destroy(ctx);
            return -1;
        }
        std::memcpy(&known[n3 * (size_t)k], bd[k].lu.data(), sizeof(double) * n3);
    }
    std::vector<int> iters((size_t)NStep), flags((size_t)NStep);
    int solved = 0;
    int rc = tdbem_assemble_step0(ctx, maxleafpointnum);
    if (rc == 0)
        rc = tdbem_assemble_history(ctx, (int)MaxN, nullptr);
    if (rc == 0)
        rc = tdbem_solve(ctx, (int)NStep, known.data(), iterations, error, lu.data(), lt.data(), iters.data(),
                         flags.data(), &solved, nullptr);
    if (rc != 0) {
        std::printf("%s\n", tdbem_last_error());
        tdbem_destroy(ctx);
        return -1;
    }
    for (long k = 0; k <= NStep; ++k) {
        bd[k].lu.assign(lu.begin() + (long)(n3 * (size_t)k), lu.begin() + (long)(n3 * (size_t)(k + 1)));
        bd[k].lt.assign(lt.begin() + (long)(n3 * (size_t)k), lt.begin() + (long)(n3 * (size_t)(k + 1)));
    }
    tdbem_destroy(ctx);
    return solved > 0 && flags[(size_t)solved - 1] == 0 ? 1 : -1;
}

#endif
