// Input and output formats of the reference around the path:
//  - DBEMBIN v1 binary model / boundary files (DBEM1/DBEM.cpp:89-330; the
//    reference's converter convert_inputs_to_binary.py writes them) and the
//    text .model / .bd files they mirror (DBEM1/DBEM.cpp:849-900,1010-1085),
//    with the reference's binary-mode choice (0 text, 1 binary required,
//    2 binary when the ".bin" file exists, DBEM1/DBEM.cpp:347-362,828-848);
//  - validation_state.csv and validation_metrics.txt
//    (WriteValidationOutputs, DBEM1/DBEM.cpp:434-561, with
//    WriteMultiDomainValidationMetrics, DBEM1/multidomain.cpp:388-416, and
//    the structure / interface diagnostics of DBEM1/multidomain.cpp:1027-1070,
//    3289-3420).
#include "io.hpp"

#include "precond.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace tdbem {

namespace {

const char kMagic[8] = {'D', 'B', 'E', 'M', 'B', 'I', 'N', '\0'};
const uint32_t kVersion = 1, kKindModel = 1, kKindBd = 2, kEndian = 0x01020304u;

struct FileCloser {
    void operator()(FILE* f) const
    {
        if (f)
            std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

File OpenFile(const std::string& path, const char* mode)
{
    FILE* f = std::fopen(path.c_str(), mode);
    return File(f);
}

bool Exists(const std::string& path)
{
    return (bool)OpenFile(path, "rb");
}

template <typename T>
void ReadRaw(FILE* f, T* out, size_t n, const std::string& what, const std::string& path)
{
    if (n && std::fread(out, sizeof(T), n, f) != n)
        throw std::runtime_error("invalid " + what + " in binary file " + path);
}

void ReadHeader(FILE* f, uint32_t kind, const std::string& path)
{
    char magic[8];
    uint32_t h[4];
    if (std::fread(magic, 1, 8, f) != 8 || std::fread(h, sizeof(uint32_t), 4, f) != 4)
        throw std::runtime_error("invalid binary header in " + path);
    if (std::memcmp(magic, kMagic, 8) != 0 || h[0] != kVersion || h[1] != kind || h[2] != kEndian)
        throw std::runtime_error("binary header mismatch in " + path);
}

void WriteHeader(FILE* f, uint32_t kind)
{
    uint32_t h[4] = {kVersion, kind, kEndian, 0u};
    std::fwrite(kMagic, 1, 8, f);
    std::fwrite(h, sizeof(uint32_t), 4, f);
}

// whitespace-separated tokens of a text file
class Tokens {
public:
    explicit Tokens(const std::string& path) : in_(path)
    {
        if (!in_)
            throw std::runtime_error("cannot open " + path);
        path_ = path;
    }
    long Long()
    {
        long v;
        if (!(in_ >> v))
            throw std::runtime_error("unexpected end or bad integer in " + path_);
        return v;
    }
    double Double()
    {
        double v;
        if (!(in_ >> v))
            throw std::runtime_error("unexpected end or bad number in " + path_);
        return v;
    }
    bool TryLong(long& v) { return (bool)(in_ >> v); }

private:
    std::ifstream in_;
    std::string path_;
};

bool UseBinary(const std::string& text, int mode)
{
    if (mode < 0 || mode > 2)
        throw std::runtime_error("binary mode must be 0, 1 or 2");
    if (mode == 1) {
        if (!Exists(text + ".bin"))
            throw std::runtime_error("missing required binary file: " + text + ".bin");
        return true;
    }
    return mode == 2 && Exists(text + ".bin");
}

}  // namespace

ModelData ReadModelFile(const std::string& path, int mode)
{
    ModelData m;
    if (UseBinary(path, mode)) {
        std::string bin = path + ".bin";
        File f = OpenFile(bin, "rb");
        if (!f)
            throw std::runtime_error("cannot open binary model file: " + bin);
        ReadHeader(f.get(), kKindModel, bin);
        int64_t n = 0;
        ReadRaw(f.get(), &n, 1, "PointNum", bin);
        if (n <= 0)
            throw std::runtime_error("invalid PointNum in binary model file: " + bin);
        m.points.resize(3 * (size_t)n);
        ReadRaw(f.get(), m.points.data(), m.points.size(), "point data", bin);
        ReadRaw(f.get(), &n, 1, "EleNum", bin);
        if (n <= 0)
            throw std::runtime_error("invalid EleNum in binary model file: " + bin);
        std::vector<int64_t> ids(8 * (size_t)n);
        ReadRaw(f.get(), ids.data(), ids.size(), "element data", bin);
        m.elePid.assign(ids.begin(), ids.end());
        ReadRaw(f.get(), &n, 1, "InfEleNum", bin);
        if (n < 0)
            throw std::runtime_error("invalid InfEleNum in binary model file: " + bin);
        std::vector<int64_t> inf(4 * (size_t)n);
        ReadRaw(f.get(), inf.data(), inf.size(), "infinite element data", bin);
        m.infPid.assign(inf.begin(), inf.end());
        return m;
    }
    Tokens t(path);
    long np = t.Long();
    if (np <= 0)
        throw std::runtime_error("invalid PointNum in " + path);
    m.points.resize(3 * (size_t)np);
    for (double& v : m.points)
        v = t.Double();
    long ne = t.Long();
    if (ne <= 0)
        throw std::runtime_error("invalid EleNum in " + path);
    m.elePid.resize(8 * (size_t)ne);
    for (long& v : m.elePid)
        v = t.Long() - 1;  // 1-based in the text file
    long ninf = 0;
    if (t.TryLong(ninf) && ninf > 0) {
        m.infPid.resize(4 * (size_t)ninf);
        for (long& v : m.infPid)
            v = t.Long() - 1;
    }
    return m;
}

void ReadBdFile(const std::string& path, int mode, long ne, int nstep, const int* expectedBcid, int* bcid,
                double* lu)
{
    if (ne <= 0 || nstep < 0)
        throw std::runtime_error("invalid element or step count");
    if (UseBinary(path, mode)) {
        std::string bin = path + ".bin";
        File f = OpenFile(bin, "rb");
        if (!f)
            throw std::runtime_error("cannot open binary boundary file: " + bin);
        ReadHeader(f.get(), kKindBd, bin);
        int64_t fe = 0, fs = 0;
        ReadRaw(f.get(), &fe, 1, "EleNum", bin);
        ReadRaw(f.get(), &fs, 1, "NStep", bin);
        if (fe != ne || fs != nstep)
            throw std::runtime_error("boundary binary size mismatch in " + bin);
        std::vector<int32_t> ids((size_t)ne);
        ReadRaw(f.get(), ids.data(), ids.size(), "boundary type section", bin);
        for (long i = 0; i < ne; ++i) {
            if (expectedBcid && ids[(size_t)i] != expectedBcid[i])
                throw std::runtime_error("boundary type changed in " + bin);
            bcid[i] = ids[(size_t)i];
        }
        ReadRaw(f.get(), lu, (size_t)(nstep + 1) * (size_t)ne * 3, "boundary value section", bin);
        return;
    }
    Tokens t(path);
    for (long i = 0; i < ne; ++i) {
        long v = t.Long();
        if (expectedBcid && v != expectedBcid[i])
            throw std::runtime_error("boundary type changed in " + path);
        bcid[i] = (int)v;
    }
    size_t n = (size_t)(nstep + 1) * (size_t)ne * 3;
    for (size_t i = 0; i < n; ++i)
        lu[i] = t.Double();
}

void WriteModelBin(const std::string& path, long np, const double* points, long ne, const long* elePid, long ninf,
                   const long* infPid)
{
    File f = OpenFile(path, "wb");
    if (!f)
        throw std::runtime_error("cannot write " + path);
    WriteHeader(f.get(), kKindModel);
    int64_t n = np;
    std::fwrite(&n, sizeof(n), 1, f.get());
    std::fwrite(points, sizeof(double), 3 * (size_t)np, f.get());
    n = ne;
    std::fwrite(&n, sizeof(n), 1, f.get());
    std::vector<int64_t> ids(elePid, elePid + 8 * (size_t)ne);
    std::fwrite(ids.data(), sizeof(int64_t), ids.size(), f.get());
    n = ninf;
    std::fwrite(&n, sizeof(n), 1, f.get());
    if (ninf > 0) {
        std::vector<int64_t> inf(infPid, infPid + 4 * (size_t)ninf);
        std::fwrite(inf.data(), sizeof(int64_t), inf.size(), f.get());
    }
    if (std::ferror(f.get()))
        throw std::runtime_error("write error on " + path);
}

void WriteBdBin(const std::string& path, long ne, int nstep, const int* bcid, const double* lu)
{
    File f = OpenFile(path, "wb");
    if (!f)
        throw std::runtime_error("cannot write " + path);
    WriteHeader(f.get(), kKindBd);
    int64_t h[2] = {ne, nstep};
    std::fwrite(h, sizeof(int64_t), 2, f.get());
    std::vector<int32_t> ids(bcid, bcid + ne);
    std::fwrite(ids.data(), sizeof(int32_t), ids.size(), f.get());
    std::fwrite(lu, sizeof(double), (size_t)(nstep + 1) * (size_t)ne * 3, f.get());
    if (std::ferror(f.get()))
        throw std::runtime_error("write error on " + path);
}

namespace {

bool FiniteBlock(const double* b)
{
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(b[i]))
            return false;
    return true;
}

bool SameBlockAndSign(const VariableRef& a, const VariableRef& b, double sign)
{
    return a.globalBlock == b.globalBlock && a.sign == sign * b.sign;
}

}  // namespace

StructureStats DiagnoseStructure(const CCSRMat& M)
{
    StructureStats s;
    if (M.n <= 0)
        return s;
    std::vector<long> colCount((size_t)M.n, 0);
    for (long r = 0; r < M.n; ++r) {
        int b = M.row_ptr[(size_t)r], e = M.row_ptr[(size_t)r + 1];
        if (b == e)
            ++s.zeroRows;
        for (int p = b; p < e; ++p) {
            int c = M.col_index[(size_t)p];
            if (c < 0 || c >= M.n) {
                ++s.nonFinite;
                continue;
            }
            ++colCount[(size_t)c];
            const double* blk = &M.value[9 * (size_t)p];
            if (!FiniteBlock(blk))
                ++s.nonFinite;
            if (c == r) {
                ++s.diagonal;
                double inv[9];
                std::memcpy(inv, blk, sizeof(inv));
                if (FiniteBlock(inv) && InvMat(inv, 3) >= 0 && FiniteBlock(inv))
                    ++s.invertibleDiagonal;
            }
        }
    }
    for (long c = 0; c < M.n; ++c)
        if (!colCount[(size_t)c])
            ++s.zeroCols;
    s.valid = s.zeroRows == 0 && s.zeroCols == 0 && s.nonFinite == 0;
    return s;
}

long InterfaceDofFailures(const GlobalDofMap& dof, const MultiDomainModel& model)
{
    long failures = 0;
    for (const InterfacePair& itf : model.interfaces) {
        double tSignB = itf.normalSign == 0 ? -1.0 : (double)itf.normalSign;
        long a = itf.eleA, b = itf.eleB;  // constant elements: node == element
        VariableRef uA = dof.GetU(itf.domainA, a), uB = dof.GetU(itf.domainB, b);
        VariableRef tA = dof.GetT(itf.domainA, a), tB = dof.GetT(itf.domainB, b);
        VariableRef huA = dof.GetHistoryU(itf.domainA, a), huB = dof.GetHistoryU(itf.domainB, b);
        VariableRef htA = dof.GetHistoryT(itf.domainA, a), htB = dof.GetHistoryT(itf.domainB, b);
        bool ok = !uA.known && !uB.known && SameBlockAndSign(uA, uB, 1.0);
        ok = ok && !tA.known && !tB.known && SameBlockAndSign(tA, tB, tSignB);
        ok = ok && SameBlockAndSign(huA, huB, 1.0);
        ok = ok && SameBlockAndSign(htA, htB, tSignB);
        ok = ok && dof.GetPreferredUnknownBlockForRow(a) == uA.globalBlock;
        ok = ok && dof.GetPreferredUnknownBlockForRow(b) == tA.globalBlock;
        if (!ok)
            ++failures;
    }
    return failures;
}

void WriteValidationFiles(const ValidationInput& in)
{
    const long ne = in.ne;
    const long n3 = 3 * ne;
    {
        std::string path = in.dir + "/validation_state.csv";
        File f = OpenFile(path, "w");
        if (!f)
            throw std::runtime_error("cannot open " + path + " for writing");
        std::fprintf(f.get(), "step,element,localNode,nodeId,domain,surfaceType,bcid,x,y,z,ux,uy,uz,tx,ty,tz\n");
        for (long step = 0; step <= in.nstep; ++step)
            for (long e = 0; e < ne; ++e) {
                const double* u = in.lu + (size_t)step * n3 + 3 * (size_t)e;
                const double* t = in.lt + (size_t)step * n3 + 3 * (size_t)e;
                const double* x = in.nodes + 3 * (size_t)e;
                std::fprintf(f.get(),
                             "%ld,%ld,%d,%ld,%d,%d,%d,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", step,
                             e, 0, e, in.domainId[(size_t)e], in.surface[(size_t)e], in.bcid[(size_t)e], x[0], x[1],
                             x[2], u[0], u[1], u[2], t[0], t[1], t[2]);
            }
        if (std::ferror(f.get()))
            throw std::runtime_error("write error on " + path);
    }
    std::string path = in.dir + "/validation_metrics.txt";
    File f = OpenFile(path, "w");
    if (!f)
        throw std::runtime_error("cannot open " + path + " for writing");
    FILE* m = f.get();
    const MultiDomainModel& model = *in.model;
    std::fprintf(m, "ValidationOutput=1\n");
    std::fprintf(m, "Solver=%d\n", in.solver);
    std::fprintf(m, "NStep=%ld\n", in.nstep);
    std::fprintf(m, "Dt=%.17g\n", in.dt);
    std::fprintf(m, "ElementCount=%ld\n", ne);
    std::fprintf(m, "NodeCount=%ld\n", ne);
    std::fprintf(m, "MultiDomainActive=%d\n", model.IsActive() ? 1 : 0);
    std::fprintf(m, "DomainCount=%d\n", model.DomainCount());
    std::fprintf(m, "InterfaceCount=%ld\n", (long)model.interfaces.size());
    for (const Domain& d : model.domains)
        std::fprintf(m, "DomainMaterial[%d]=C1 %.17g C2 %.17g C1Dt %.17g C2Dt %.17g\n", d.id, d.mat.C1, d.mat.C2,
                     d.mat.C1Dt, d.mat.C2Dt);
    double overallU = 0.0, overallT = 0.0;
    for (long step = 0; step <= in.nstep; ++step) {
        double maxU = 0.0, maxT = 0.0;
        for (const InterfacePair& itf : model.interfaces) {
            double tSignB = itf.normalSign == 0 ? -1.0 : (double)itf.normalSign;
            const double* ua = in.lu + (size_t)step * n3 + 3 * (size_t)itf.eleA;
            const double* ub = in.lu + (size_t)step * n3 + 3 * (size_t)itf.eleB;
            const double* ta = in.lt + (size_t)step * n3 + 3 * (size_t)itf.eleA;
            const double* tb = in.lt + (size_t)step * n3 + 3 * (size_t)itf.eleB;
            for (int k = 0; k < 3; ++k) {
                double du = std::fabs(ua[k] - ub[k]);
                double bt = std::fabs(tb[k] - tSignB * ta[k]);
                if (du > maxU)
                    maxU = du;
                if (bt > maxT)
                    maxT = bt;
            }
        }
        if (maxU > overallU)
            overallU = maxU;
        if (maxT > overallT)
            overallT = maxT;
        std::fprintf(m, "InterfaceStep[%ld].MaxAbsU=%.17g\n", step, maxU);
        std::fprintf(m, "InterfaceStep[%ld].MaxAbsTBalance=%.17g\n", step, maxT);
    }
    std::fprintf(m, "InterfaceOverallMaxAbsU=%.17g\n", overallU);
    std::fprintf(m, "InterfaceOverallMaxAbsTBalance=%.17g\n", overallT);
    if (!model.IsActive()) {
        std::fprintf(m, "MultiDomainMetricsValid=0\n");
        return;
    }
    // GMRES with maxit = 1 reports one outer iteration per step, none when the
    // right-hand side is zero (inner count 0; DBEM1/pre_gmres.cpp:1403-1411)
    long inner = 0, outer = 0;
    for (long k = 0; k < in.solved; ++k) {
        inner += in.iters[k];
        outer += in.iters[k] > 0 ? 1 : 0;
    }
    int finalFlag = in.solved > 0 ? in.flags[in.solved - 1] : -999;
    double ns = (double)in.solved;
    std::fprintf(m, "MultiDomainMetricsValid=%d\n", 1);
    std::fprintf(m, "MultiDomainFinalFlag=%d\n", finalFlag);
    std::fprintf(m, "MultiDomainSolvedSteps=%ld\n", in.solved);
    std::fprintf(m, "MultiDomainTotalOuterIterations=%ld\n", outer);
    std::fprintf(m, "MultiDomainTotalInnerIterations=%ld\n", inner);
    std::fprintf(m, "MultiDomainAverageOuterIterations=%lf\n", in.solved > 0 ? (double)outer / ns : 0.0);
    std::fprintf(m, "MultiDomainAverageInnerIterations=%lf\n", in.solved > 0 ? (double)inner / ns : 0.0);
    std::fprintf(m, "MultiDomainDomainCount=%d\n", model.DomainCount());
    std::fprintf(m, "MultiDomainGlobalUnknownBlocks=%ld\n", in.dof->unknownBlockCount);
    std::fprintf(m, "MultiDomainMTS0NonZeroBlocks=%ld\n", in.A->NonZeroBlocks());
    std::fprintf(m, "MultiDomainMTSAverageBlocks=%lf\n", in.avgMTSBlocks);
    std::fprintf(m, "MultiDomainMGSAverageBlocks=%lf\n", in.avgMGSBlocks);
    std::fprintf(m, "MultiDomainMatrixAssemblyTime=%lf\n", in.assemblyTime);
    std::fprintf(m, "MultiDomainGMRESTime=%lf\n", in.gmresTime);
    StructureStats st = DiagnoseStructure(*in.A);
    std::fprintf(m, "MultiDomainMatrixStructureValid=%d\n", st.valid ? 1 : 0);
    std::fprintf(m, "MultiDomainMatrixZeroRows=%ld\n", st.zeroRows);
    std::fprintf(m, "MultiDomainMatrixZeroCols=%ld\n", st.zeroCols);
    std::fprintf(m, "MultiDomainMatrixNonFiniteBlocks=%ld\n", st.nonFinite);
    std::fprintf(m, "MultiDomainMatrixDiagonalBlocks=%ld\n", st.diagonal);
    std::fprintf(m, "MultiDomainMatrixInvertibleDiagonalBlocks=%ld\n", st.invertibleDiagonal);
    long fails = InterfaceDofFailures(*in.dof, model);
    std::fprintf(m, "MultiDomainInterfaceDofInvariantValid=%d\n", fails == 0 ? 1 : 0);
    std::fprintf(m, "MultiDomainInterfaceDofInvariantFailures=%ld\n", fails);
    if (std::ferror(m))
        throw std::runtime_error("write error on " + path);
}

}  // namespace tdbem
