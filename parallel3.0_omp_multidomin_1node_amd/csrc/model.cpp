// Multi-domain model, DOF map and block builder on the host.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

#pragma clang fp contract(off)

namespace tdbem {

// ---------------------------------------------------------------------------
// Model (DBEM1/multidomain.cpp:500-523,631-706,1364-1480)
// ---------------------------------------------------------------------------
namespace {
int NormalizeDomainId(int id, int count, bool zeroBased)
{
    if (zeroBased && id >= 0 && id < count)
        return id;
    if (id >= 1 && id <= count)
        return id - 1;
    return -1;
}
long NormalizeElementId(long id, long ne, bool zeroBased)
{
    if (zeroBased)
        return (id >= 0 && id < ne) ? id : -1;
    return (id >= 1 && id <= ne) ? id - 1 : -1;
}
}  // namespace

bool MultiDomainModel::Validate() const
{
    if (domains.empty()) {
        printf("MultiDomain validation failed: no domains.\n");
        return false;
    }
    std::vector<int> owner((size_t)elementCount, -1);
    for (size_t i = 0; i < domains.size(); ++i) {
        const Domain& d = domains[i];
        if (d.id != (int)i || d.eleBegin < 0 || d.eleCount <= 0 || d.eleBegin + d.eleCount > elementCount) {
            printf("MultiDomain validation failed: invalid domain %d.\n", d.id);
            return false;
        }
        for (long e : d.elementIds) {
            if (e < 0 || e >= elementCount || owner[(size_t)e] != -1) {
                printf("MultiDomain validation failed: element %ld.\n", e);
                return false;
            }
            owner[(size_t)e] = d.id;
        }
    }
    for (long e = 0; e < elementCount; ++e)
        if (owner[(size_t)e] == -1) {
            printf("MultiDomain validation failed: element %ld has no domain.\n", e);
            return false;
        }
    std::vector<int> use((size_t)elementCount, 0);
    for (size_t i = 0; i < interfaces.size(); ++i) {
        const InterfacePair& f = interfaces[i];
        if (f.domainA < 0 || f.domainA >= DomainCount() || f.domainB < 0 || f.domainB >= DomainCount() ||
            f.domainA == f.domainB || f.eleA < 0 || f.eleA >= elementCount || f.eleB < 0 ||
            f.eleB >= elementCount || owner[(size_t)f.eleA] != f.domainA || owner[(size_t)f.eleB] != f.domainB) {
            printf("MultiDomain validation failed: invalid interface %ld.\n", (long)i);
            return false;
        }
        if (++use[(size_t)f.eleA] > 1 || ++use[(size_t)f.eleB] > 1) {
            printf("MultiDomain validation failed: repeated interface element in interface %ld.\n", (long)i);
            return false;
        }
    }
    return true;
}

int BuildMultiDomainModel(const MultiDomainInputConfig& config, long eleNum, long nodeNum, double defaultE,
                          double defaultV, double defaultRou, double dt, MultiDomainModel& model)
{
    model = MultiDomainModel();
    model.enabled = config.enabled;
    model.elementCount = eleNum;
    model.nodeCount = nodeNum;
    int dc = config.enabled ? config.domainCount : 1;
    if (dc <= 0 || dc > kMaxDomains || eleNum <= 0)
        return 0;
    model.domains.resize((size_t)dc);
    bool zb = false;
    for (const auto& m : config.materials)
        if (m.id == 0)
            zb = true;
    std::vector<DomainMaterialInput> mats((size_t)dc);
    for (int i = 0; i < dc; ++i)
        mats[(size_t)i] = DomainMaterialInput{i, defaultE, defaultV, defaultRou};
    if (config.enabled)
        for (const auto& m : config.materials) {
            int id = NormalizeDomainId(m.id, dc, zb);
            if (id < 0) {
                printf("Invalid domain material id %d.\n", m.id);
                return 0;
            }
            mats[(size_t)id] = m;
            mats[(size_t)id].id = id;
        }
    long b = 0;
    for (int d = 0; d < dc; ++d) {
        long base = eleNum / dc, extra = eleNum % dc;
        long cnt = base + (d < extra ? 1 : 0);
        Domain& D = model.domains[(size_t)d];
        D.id = d;
        D.eleBegin = b;
        D.eleCount = cnt;
        D.nodeBegin = b;
        D.nodeCount = cnt;
        double G = mats[(size_t)d].E / 2.0 / (1.0 + mats[(size_t)d].v);
        D.mat = BuildDynaMat(mats[(size_t)d].v, G, mats[(size_t)d].Rou, dt);
        D.E = mats[(size_t)d].E;
        D.v = mats[(size_t)d].v;
        D.Rou = mats[(size_t)d].Rou;
        for (long e = 0; e < cnt; ++e)
            D.elementIds.push_back(b + e);
        b += cnt;
    }
    if (config.enabled && config.hasExplicitInterfaces) {
        bool zbe = false, zbd = false;
        for (const auto& f : config.interfaces) {
            if (f.eleA == 0 || f.eleB == 0)
                zbe = true;
            if (f.domainA == 0 || f.domainB == 0)
                zbd = true;
        }
        for (InterfacePair f : config.interfaces) {
            f.domainA = NormalizeDomainId(f.domainA, dc, zbd);
            f.domainB = NormalizeDomainId(f.domainB, dc, zbd);
            f.eleA = NormalizeElementId(f.eleA, eleNum, zbe);
            f.eleB = NormalizeElementId(f.eleB, eleNum, zbe);
            if (f.normalSign == 0)
                f.normalSign = -1;
            model.interfaces.push_back(f);
        }
    } else if (dc > 1) {
        for (int d = 0; d < dc - 1; ++d) {
            InterfacePair f;
            f.domainA = d;
            f.domainB = d + 1;
            f.eleA = model.domains[(size_t)d].eleBegin + model.domains[(size_t)d].eleCount - 1;
            f.eleB = model.domains[(size_t)d + 1].eleBegin;
            f.normalSign = -1;
            model.interfaces.push_back(f);
        }
    }
    return model.Validate() ? 1 : 0;
}

// ---------------------------------------------------------------------------
// GlobalDofMap (DBEM1/multidomain.cpp:777-1026)
// ---------------------------------------------------------------------------
bool GlobalDofMap::Build(const MultiDomainModel& model, const int* bcid, long eleNum, const std::vector<long>* groups)
{
    const long N = model.nodeCount;
    unknownBlockCount = knownBlockCount = 0;
    equationBlockCount = N;
    outerDisplacementUnknownBlocks = outerTractionUnknownBlocks = 0;
    interfaceDisplacementUnknownBlocks = interfaceTractionUnknownBlocks = 0;
    historyUBlockCount = historyTBlockCount = 0;
    m_uRefs.assign((size_t)N, VariableRef());
    m_tRefs.assign((size_t)N, VariableRef());
    m_historyURefs.assign((size_t)N, VariableRef());
    m_historyTRefs.assign((size_t)N, VariableRef());
    m_equationRows.assign((size_t)N, -1);
    m_preferred.assign((size_t)N, -1);
    for (long i = 0; i < N; ++i)
        m_equationRows[(size_t)i] = i;
    std::vector<long> bounds = {0, eleNum};
    if (groups && groups->size() >= 2)
        bounds = *groups;
    if (bounds.front() != 0 || bounds.back() != eleNum)
        return false;
    long unknown = 0;
    std::vector<char> side((size_t)N, 0);
    for (size_t g = 0; g + 1 < bounds.size(); ++g) {
        const long e0 = bounds[g], e1 = bounds[g + 1];
        for (const InterfacePair& f : model.interfaces) {
            if (f.eleA < e0 || f.eleA >= e1)
                continue;
            if (f.eleB < e0 || f.eleB >= e1)
                return false;  // an interface joins two groups
            double sB = (double)(f.normalSign == 0 ? -1 : f.normalSign);
            long ub = unknown++, tb = unknown++;
            m_uRefs[(size_t)f.eleA] = VariableRef(ub, 1.0, false);
            m_uRefs[(size_t)f.eleB] = VariableRef(ub, 1.0, false);
            m_tRefs[(size_t)f.eleA] = VariableRef(tb, 1.0, false);
            m_tRefs[(size_t)f.eleB] = VariableRef(tb, sB, false);
            m_preferred[(size_t)f.eleA] = ub;
            m_preferred[(size_t)f.eleB] = tb;
            long hu = historyUBlockCount++, ht = historyTBlockCount++;
            m_historyURefs[(size_t)f.eleA] = VariableRef(hu, 1.0, false);
            m_historyURefs[(size_t)f.eleB] = VariableRef(hu, 1.0, false);
            m_historyTRefs[(size_t)f.eleA] = VariableRef(ht, 1.0, false);
            m_historyTRefs[(size_t)f.eleB] = VariableRef(ht, sB, false);
            side[(size_t)f.eleA] = side[(size_t)f.eleB] = 1;
            ++interfaceDisplacementUnknownBlocks;
            ++interfaceTractionUnknownBlocks;
        }
        for (long e = e0; e < e1; ++e) {
            if (bcid[e] != 123 && bcid[e] != 456) {
                printf("MultiDomain dof map failed: element %ld has mixed BCID %d. Only 123 and 456 are supported.\n",
                       e, bcid[e]);
                return false;
            }
            if (side[(size_t)e])
                continue;
            if (bcid[e] == 123) {
                long tb = unknown++;
                m_uRefs[(size_t)e] = VariableRef(knownBlockCount++, 1.0, true);
                m_tRefs[(size_t)e] = VariableRef(tb, 1.0, false);
                m_preferred[(size_t)e] = tb;
                ++outerTractionUnknownBlocks;
            } else {
                long ub = unknown++;
                m_uRefs[(size_t)e] = VariableRef(ub, 1.0, false);
                m_tRefs[(size_t)e] = VariableRef(knownBlockCount++, 1.0, true);
                m_preferred[(size_t)e] = ub;
                ++outerDisplacementUnknownBlocks;
            }
            m_historyURefs[(size_t)e] = VariableRef(historyUBlockCount++, 1.0, false);
            m_historyTRefs[(size_t)e] = VariableRef(historyTBlockCount++, 1.0, false);
        }
    }
    unknownBlockCount = unknown;
    if (unknownBlockCount != equationBlockCount) {
        printf("MultiDomain dof map failed: unknown blocks %ld != equation blocks %ld.\n", unknownBlockCount,
               equationBlockCount);
        return false;
    }
    return Validate();
}

bool GlobalDofMap::Validate() const
{
    if (equationBlockCount <= 0 || unknownBlockCount != equationBlockCount)
        return false;
    std::vector<int> used((size_t)unknownBlockCount, 0), pref((size_t)unknownBlockCount, 0);
    for (size_t i = 0; i < m_uRefs.size(); ++i) {
        const VariableRef &u = m_uRefs[i], &t = m_tRefs[i];
        if (m_equationRows[i] < 0 || u.globalBlock < 0 || t.globalBlock < 0)
            return false;
        if ((!u.known && u.globalBlock >= unknownBlockCount) || (!t.known && t.globalBlock >= unknownBlockCount))
            return false;
        if ((u.known && u.globalBlock >= knownBlockCount) || (t.known && t.globalBlock >= knownBlockCount))
            return false;
        if (!u.known)
            used[(size_t)u.globalBlock] = 1;
        if (!t.known)
            used[(size_t)t.globalBlock] = 1;
        if (m_historyURefs[i].globalBlock < 0 || m_historyURefs[i].globalBlock >= historyUBlockCount ||
            m_historyTRefs[i].globalBlock < 0 || m_historyTRefs[i].globalBlock >= historyTBlockCount)
            return false;
        if (m_preferred[i] < 0 || m_preferred[i] >= unknownBlockCount)
            return false;
        ++pref[(size_t)m_preferred[i]];
    }
    for (long b = 0; b < unknownBlockCount; ++b)
        if (pref[(size_t)b] != 1 || !used[(size_t)b]) {
            printf("MultiDomain dof map validation failed at unknown block %ld.\n", b);
            return false;
        }
    return true;
}

// ---------------------------------------------------------------------------
// Builder (DBEM1/multidomain.cpp:1101-1269)
// ---------------------------------------------------------------------------
double BlockMaxAbs9(const double* b)
{
    double m = 0.0;
    for (int i = 0; i < 9; ++i) {
        double a = std::fabs(b[i]);
        if (a > m)
            m = a;
    }
    return m;
}

bool IsFiniteBlock9(const double* b)
{
    for (int i = 0; i < 9; ++i)
        if (!(b[i] == b[i] && std::fabs(b[i]) <= std::numeric_limits<double>::max()))
            return false;
    return true;
}

static inline uint64_t Key(long r, long c) { return ((uint64_t)r << 32) | (uint64_t)(uint32_t)c; }

void MultiDomainCCSRBuilder::Reset(long blockRows, long blockCols)
{
    m_blockRows = blockRows;
    m_blockCols = blockCols;
    m_blocks.clear();
    m_skippedNonFinite = 0;
}

void MultiDomainCCSRBuilder::AddBlock(long row, long col, double sign, const double block9[9])
{
    if (row < 0 || row >= m_blockRows || col < 0 || col >= m_blockCols)
        return;
    if (!block9 || sign == 0.0)
        return;
    bool zero = true;
    for (int i = 0; i < 9; ++i)
        if (block9[i] != 0.0)
            zero = false;
    if (zero || BlockMaxAbs9(block9) <= kAbsTol)
        return;
    if (!IsFiniteBlock9(block9)) {
        // skipped and reported, the first 20 per builder (DBEM1/multidomain.cpp:1126-1133)
        if (m_skippedNonFinite < 20)
            std::fprintf(stderr, "MultiDomain skipped non-finite block row=%ld col=%ld\n", row, col);
        ++m_skippedNonFinite;
        return;
    }
    Entry& e = m_blocks[Key(row, col)];
    if (!e.alive) {
        e.v.fill(0.0);
        e.alive = true;
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            e.v[(size_t)(r * 3 + c)] += sign * block9[c * 3 + r];
    if (BlockMaxAbs9(e.v.data()) <= kAbsTol)
        e.alive = false;
}

void MultiDomainCCSRBuilder::MergeFrom(const MultiDomainCCSRBuilder& other)
{
    if (other.m_blockRows != m_blockRows || other.m_blockCols != m_blockCols)
        return;
    std::vector<std::pair<uint64_t, const Entry*>> list;
    other.Sorted(list);
    for (const auto& kv : list) {
        Entry& e = m_blocks[kv.first];
        if (!e.alive) {
            e.v.fill(0.0);
            e.alive = true;
        }
        for (int k = 0; k < 9; ++k)
            e.v[(size_t)k] += kv.second->v[(size_t)k];
        if (BlockMaxAbs9(e.v.data()) <= kAbsTol)
            e.alive = false;
    }
}

void MultiDomainCCSRBuilder::Sorted(std::vector<std::pair<uint64_t, const Entry*>>& out) const
{
    out.clear();
    out.reserve(m_blocks.size());
    for (const auto& kv : m_blocks)
        if (kv.second.alive)
            out.emplace_back(kv.first, &kv.second);
    std::sort(out.begin(), out.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
}

long MultiDomainCCSRBuilder::NonZeroBlocks() const
{
    long n = 0;
    for (const auto& kv : m_blocks)
        n += kv.second.alive ? 1 : 0;
    return n;
}

void MultiDomainCCSRBuilder::Build(CCSRMat& out) const
{
    std::vector<std::pair<uint64_t, const Entry*>> list;
    Sorted(list);
    out.n = m_blockRows;
    out.row_ptr.assign((size_t)m_blockRows + 1, 0);
    out.col_index.resize(list.size());
    out.value.resize(list.size() * 9);
    for (size_t i = 0; i < list.size(); ++i) {
        long row = (long)(list[i].first >> 32);
        out.row_ptr[(size_t)row + 1]++;
        out.col_index[i] = (int)(uint32_t)list[i].first;
        for (int k = 0; k < 9; ++k)
            out.value[i * 9 + (size_t)k] = list[i].second->v[(size_t)k];
    }
    for (long r = 0; r < m_blockRows; ++r)
        out.row_ptr[(size_t)r + 1] += out.row_ptr[(size_t)r];
}

void MultiDomainCCSRBuilder::BuildSym(SymCCSRMat& out, double tolerance, bool* symmetric) const
{
    std::vector<std::pair<uint64_t, const Entry*>> list;
    Sorted(list);
    bool sym = true;
    out.n = m_blockRows;
    out.row_ptr.assign((size_t)m_blockRows + 1, 0);
    out.col_index.resize(list.size());
    out.value.resize(list.size() * 6);
    for (size_t i = 0; i < list.size(); ++i) {
        const double* b = list[i].second->v.data();
        double th = tolerance, mixed = kAbsTol + kRelTol * BlockMaxAbs9(b);
        if (mixed > th)
            th = mixed;
        if (std::fabs(b[1] - b[3]) > th || std::fabs(b[2] - b[6]) > th || std::fabs(b[5] - b[7]) > th)
            sym = false;
        long row = (long)(list[i].first >> 32);
        out.row_ptr[(size_t)row + 1]++;
        out.col_index[i] = (int)(uint32_t)list[i].first;
        const int pick[6] = {0, 1, 2, 4, 5, 8};
        for (int k = 0; k < 6; ++k)
            out.value[i * 6 + (size_t)k] = b[pick[k]];
    }
    for (long r = 0; r < m_blockRows; ++r)
        out.row_ptr[(size_t)r + 1] += out.row_ptr[(size_t)r];
    if (symmetric)
        *symmetric = sym;
}

void SparseMul(const CCSRMat& M, const double* x, double* b)
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

void SparseMul(const SymCCSRMat& M, const double* x, double* b)
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

}  // namespace tdbem
