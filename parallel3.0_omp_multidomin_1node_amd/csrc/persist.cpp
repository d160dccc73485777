// History-matrix persistence: the assembled history bands of a case written
// to a file and loaded into another case of the same model, the analogue of
// the reference's CCSR matrix cache and of its in-process reuse across load
// cases (DBEM1/pre_gmres.cpp:16-27,199-544; DBEM1/mat_vec.cpp:1690-1730).
// The file keeps the bands as the device holds them (per pair: first lag and
// band length; per entry: the MTS and MGS blocks and the presence bits), so a
// load is one streamed upload.  A load re-runs the causality cull of the
// loading case and requires the same band structure, materials, dt and gap.
#include "solver.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

namespace tdbem {

namespace {

const char kHistMagic[8] = {'T', 'D', 'B', 'E', 'M', 'H', 'S', 'T'};
const uint32_t kHistVersion = 1;
const size_t kStage = (size_t)64 << 20;  // bytes per streamed chunk

struct Closer {
    void operator()(FILE* f) const
    {
        if (f)
            std::fclose(f);
    }
};

void Put(FILE* f, const void* p, size_t bytes, const std::string& path)
{
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes)
        throw std::runtime_error("write error on " + path);
}

void Get(FILE* f, void* p, size_t bytes, const std::string& path)
{
    if (bytes && std::fread(p, 1, bytes, f) != bytes)
        throw std::runtime_error("truncated history file " + path);
}

// device array -> file through one host staging buffer
void PutDevice(FILE* f, const void* dev, size_t bytes, std::vector<char>& stage, const std::string& path,
               hipStream_t stream)
{
    const char* d = (const char*)dev;
    for (size_t o = 0; o < bytes; o += kStage) {
        size_t n = std::min(kStage, bytes - o);
        if (CopyOnStream(stage.data(), d + o, n, hipMemcpyDeviceToHost, stream) != hipSuccess)
            throw std::runtime_error("device copy failed while saving the history");
        Put(f, stage.data(), n, path);
    }
}

void GetDevice(FILE* f, void* dev, size_t bytes, std::vector<char>& stage, const std::string& path,
               hipStream_t stream)
{
    char* d = (char*)dev;
    for (size_t o = 0; o < bytes; o += kStage) {
        size_t n = std::min(kStage, bytes - o);
        Get(f, stage.data(), n, path);
        if (CopyOnStream(d + o, stage.data(), n, hipMemcpyHostToDevice, stream) != hipSuccess)
            throw std::runtime_error("device copy failed while loading the history");
    }
}

}  // namespace

void Case::SaveHistory(const std::string& path)
{
    RefuseStreamed("saving the history");
    if (!historyDone)
        throw std::runtime_error("no assembled history to save");
    std::unique_ptr<FILE, Closer> f(std::fopen(path.c_str(), "wb"));
    if (!f)
        throw std::runtime_error("cannot write " + path);
    uint32_t head[2] = {kHistVersion, (uint32_t)(historyMixed ? (historySplit ? 2 : 1) : 0)};
    int64_t counts[4] = {ne, npairs, entries, maxn};
    double td[2] = {dt, tables.gap};
    int32_t nd[2] = {(int32_t)model.domains.size(), 0};
    Put(f.get(), kHistMagic, 8, path);
    Put(f.get(), head, sizeof(head), path);
    Put(f.get(), counts, sizeof(counts), path);
    Put(f.get(), td, sizeof(td), path);
    Put(f.get(), nd, sizeof(nd), path);
    for (const Domain& d : model.domains) {
        double evr[3] = {d.E, d.v, d.Rou};
        Put(f.get(), evr, sizeof(evr), path);
    }
    HostBands();
    Put(f.get(), h_kb.data(), sizeof(int) * (size_t)npairs, path);
    Put(f.get(), h_len.data(), sizeof(int) * (size_t)npairs, path);
    std::vector<char> stage(std::min(kStage, (size_t)std::max<long>(1, entries) * 72));
    size_t e = (size_t)entries;
    if (historyMixed) {
        PutDevice(f.get(), d_mts32.p, 9 * e * sizeof(float), stage, path, stream);
        PutDevice(f.get(), d_mgs32.p, 6 * e * sizeof(float), stage, path, stream);
        if (historySplit) {
            PutDevice(f.get(), d_mtsLo.p, 9 * e * sizeof(unsigned short), stage, path, stream);
            PutDevice(f.get(), d_mgsLo.p, 6 * e * sizeof(unsigned short), stage, path, stream);
        }
    } else {
        PutDevice(f.get(), d_mts.p, 9 * e * sizeof(double), stage, path, stream);
        PutDevice(f.get(), d_mgs.p, 6 * e * sizeof(double), stage, path, stream);
    }
    PutDevice(f.get(), d_alive.p, e, stage, path, stream);
    if (std::fflush(f.get()) != 0)
        throw std::runtime_error("write error on " + path);
}

void Case::LoadHistory(const std::string& path)
{
    if (frozen)
        throw std::runtime_error("release the frozen structure before loading a history");
    std::unique_ptr<FILE, Closer> f(std::fopen(path.c_str(), "rb"));
    if (!f)
        throw std::runtime_error("cannot open " + path);
    char magic[8];
    uint32_t head[2];
    int64_t counts[4];
    double td[2];
    int32_t nd[2];
    Get(f.get(), magic, 8, path);
    Get(f.get(), head, sizeof(head), path);
    if (std::memcmp(magic, kHistMagic, 8) != 0 || head[0] != kHistVersion || head[1] > 2)
        throw std::runtime_error(path + " is not a history file of this version");
    Get(f.get(), counts, sizeof(counts), path);
    Get(f.get(), td, sizeof(td), path);
    Get(f.get(), nd, sizeof(nd), path);
    if (counts[0] != ne || counts[1] != npairs)
        throw std::runtime_error("history file is for another model (element or pair count)");
    if (td[0] != dt || td[1] != tables.gap)
        throw std::runtime_error("history file has another dt or gap");
    if (nd[0] != (int32_t)model.domains.size())
        throw std::runtime_error("history file has another domain count");
    for (const Domain& d : model.domains) {
        double evr[3];
        Get(f.get(), evr, sizeof(evr), path);
        if (evr[0] != d.E || evr[1] != d.v || evr[2] != d.Rou)
            throw std::runtime_error("history file has other materials");
    }
    const int fileMaxn = (int)counts[3];
    if (fileMaxn < 1)
        throw std::runtime_error("history file holds no steps");
    std::vector<int> kb((size_t)npairs), len((size_t)npairs);
    Get(f.get(), kb.data(), sizeof(int) * kb.size(), path);
    Get(f.get(), len.data(), sizeof(int) * len.size(), path);
    // The loading case's own cull must give the same bands.  The cull
    // replaces the band offsets the resident blocks were indexed by, so from
    // here on any refusal or read error also drops those blocks: the context
    // is left with no history rather than with blocks of another band layout.
    ReleaseHistory();
    precull = false;
    try {
        Cull(fileMaxn);
        HostBands();
        if (kb != h_kb || len != h_len || entries != counts[2])
            throw std::runtime_error("history file bands differ from this model's causality cull");
        EnsureSelfTables();
        LoadHistoryBlocks(f.get(), path, head[1], fileMaxn);
    } catch (...) {
        ReleaseHistory();
        throw;
    }
}

void Case::LoadHistoryBlocks(FILE* f, const std::string& path, uint32_t precision, int fileMaxn)
{
    maxn = fileMaxn;
    historyMixed = precision != 0;
    historySplit = precision == 2;
    size_t e = (size_t)entries;
    std::vector<char> stage(std::min(kStage, std::max<size_t>(1, e) * 72));
    if (historyMixed) {
        d_mts32.alloc(9 * e);
        d_mgs32.alloc(6 * e);
        GetDevice(f, d_mts32.p, 9 * e * sizeof(float), stage, path, stream);
        GetDevice(f, d_mgs32.p, 6 * e * sizeof(float), stage, path, stream);
        if (historySplit) {
            d_mtsLo.alloc(9 * e);
            d_mgsLo.alloc(6 * e);
            GetDevice(f, d_mtsLo.p, 9 * e * sizeof(unsigned short), stage, path, stream);
            GetDevice(f, d_mgsLo.p, 6 * e * sizeof(unsigned short), stage, path, stream);
        }
    } else {
        d_mts.alloc(9 * e);
        d_mgs.alloc(6 * e);
        GetDevice(f, d_mts.p, 9 * e * sizeof(double), stage, path, stream);
        GetDevice(f, d_mgs.p, 6 * e * sizeof(double), stage, path, stream);
    }
    d_alive.alloc(std::max<size_t>(1, e));
    GetDevice(f, d_alive.p, e, stage, path, stream);
    char extra;
    if (std::fread(&extra, 1, 1, f) == 1)
        throw std::runtime_error("history file " + path + " is longer than its header says");
    historyDone = true;
}

}  // namespace tdbem
