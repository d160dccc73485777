// Context of one multi-domain case: host model, device buffers, assembly
// and the time loop.  This file is the host side of the drop-in
// DynaGMRESSolverMultiDomainCCSR (DBEM1/multidomain.cpp:3664-3926).
#include "solver.hpp"
#include "batched.hpp"

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <unordered_map>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>

#pragma clang fp contract(off)

namespace tdbem {

namespace {

void Check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

double Now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {

// Events of one solve: created on demand, timed after the stream has drained,
// destroyed with the pool.
class EventPool {
public:
    ~EventPool()
    {
        for (hipEvent_t e : all)
            (void)hipEventDestroy(e);
    }
    hipEvent_t Get()
    {
        hipEvent_t e = nullptr;
        Check(hipEventCreate(&e), "event");
        all.push_back(e);
        return e;
    }
    double Elapsed(hipEvent_t a, hipEvent_t b)
    {
        float ms = 0.f;
        Check(hipEventSynchronize(b), "event");
        Check(hipEventElapsedTime(&ms, a, b), "elapsed");
        return ms * 1e-3;
    }

private:
    std::vector<hipEvent_t> all;
};

}  // namespace

}  // namespace

// Large device buffers (history bands and their like) are kept when freed
// and handed to the next request they fit (up to twice its size): mapping
// and unmapping hundreds of GB costs seconds per call on this device, and a
// process that runs many cases or models in turn would pay it every time.  A
// failed allocation first returns the cached blocks to the device.
namespace {

constexpr size_t kPoolMin = (size_t)1 << 30;
std::mutex g_poolMu;
std::multimap<size_t, void*> g_pool;          // cached blocks by capacity
std::unordered_map<void*, size_t> g_pooled;   // every pooled block and its capacity
size_t g_poolCached = 0;

void FlushPoolLocked()
{
    for (auto& b : g_pool) {
        g_pooled.erase(b.second);
        (void)hipFree(b.second);
    }
    g_pool.clear();
    g_poolCached = 0;
}

hipError_t MallocOrFlush(void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess)
        return e;
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(g_poolMu);
    FlushPoolLocked();
    return hipMalloc(p, bytes);
}

void* PoolAcquire(size_t bytes)
{
    {
        std::lock_guard<std::mutex> lk(g_poolMu);
        auto it = g_pool.lower_bound(bytes);
        if (it != g_pool.end() && it->first / 2 <= bytes) {
            void* p = it->second;
            g_poolCached -= it->first;
            g_pool.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    Check(MallocOrFlush(&p, bytes), "hipMalloc");
    std::lock_guard<std::mutex> lk(g_poolMu);
    g_pooled[p] = bytes;
    return p;
}

// true when p is a pooled block (now cached); the device is drained first so
// no queued kernel still reads the block when it is handed out again
bool PoolRelease(void* p)
{
    {
        std::lock_guard<std::mutex> lk(g_poolMu);
        if (!g_pooled.count(p))
            return false;
    }
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(g_poolMu);
    size_t cap = g_pooled[p];
    g_pool.emplace(cap, p);
    g_poolCached += cap;
    return true;
}

}  // namespace

size_t PooledBytesCached()
{
    std::lock_guard<std::mutex> lk(g_poolMu);
    return g_poolCached;
}

void ReleaseCachedDeviceMemory()
{
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(g_poolMu);
    FlushPoolLocked();
}

namespace {
thread_local ScratchArena* g_arena = nullptr;
}

void SetScratchArena(ScratchArena* a) { g_arena = a; }

template <typename T>
void DevBuf<T>::alloc(size_t count)
{
    free();
    if (count == 0)
        count = 1;
    size_t bytes = count * sizeof(T);
    if (g_arena) {  // carve from the caller's device memory; nothing is allocated
        size_t at = (g_arena->used + 255) & ~(size_t)255;
        if (at + bytes > g_arena->size)
            throw std::runtime_error("scratch of " + std::to_string(g_arena->size) + " bytes exhausted (" +
                                     std::to_string(at + bytes) + " needed)");
        p = reinterpret_cast<T*>(g_arena->base + at);
        g_arena->used = at + bytes;
        n = count;
        return;
    }
    if (bytes >= kPoolMin)
        p = (T*)PoolAcquire(bytes);
    else
        Check(MallocOrFlush((void**)&p, bytes), "hipMalloc");
    n = count;
}

template <typename T>
void DevBuf<T>::free()
{
    if (p && g_arena && (char*)p >= g_arena->base && (char*)p < g_arena->base + g_arena->size) {
        p = nullptr;  // a view into the arena: nothing to return
        n = 0;
        return;
    }
    if (p && !PoolRelease(p))
        (void)hipFree(p);
    p = nullptr;
    n = 0;
}

template <typename T>
void DevBuf<T>::upload(const T* h, size_t count, hipStream_t s)
{
    alloc(count);
    if (count)
        Check(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

template struct DevBuf<double>;
template struct DevBuf<int>;
template struct DevBuf<long>;
template struct DevBuf<unsigned char>;
template struct DevBuf<ElemGeo>;
template struct DevBuf<ElemPoints>;
template struct DevBuf<NodeRefs>;
template struct DevBuf<QuadTables>;
template struct DevBuf<PwWave>;
template struct DevBuf<float>;
template struct DevBuf<DynaMat>;
template struct DevBuf<unsigned short>;
template struct DevBuf<GmresGroupTable>;

void Case::GroupWork::release()
{
    if (gBegin)
        (void)hipGraphExecDestroy(gBegin);
    if (gIter)
        (void)hipGraphExecDestroy(gIter);
    gBegin = gIter = nullptr;
    graphB = nullptr;
    if (hst)
        (void)hipHostFree(hst);
    hst = nullptr;
    for (DevBuf<double>* b : {&V, &W, &u, &u2, &vh, &h, &h2, &QT, &R, &f, &phibar, &part, &st})
        b->free();
    stag.free();
    active.free();
    tab.free();
    restart = -1;
}

Case::~Case()
{
    if (stream)
        (void)hipStreamSynchronize(stream);
    gw.release();
    bw.release();
    ReleaseGraphs();
    if (h_gst)
        (void)hipHostFree(h_gst);
    if (h_pin)
        (void)hipHostFree(h_pin);
    if (stream)
        (void)hipStreamDestroy(stream);
}

void Case::Create(long npoints, const double* points, long nele, const int* ele_pid, const int* bcid_in,
                  const MultiDomainInputConfig& cfg, double E, double v, double rho, double dt_in, double gap)
{
    (void)npoints;
    if (nele <= 0)
        throw std::runtime_error("no elements");
    ne = nele;
    dt = dt_in;
    // non-blocking: work of other contexts on the null stream never orders
    // against this one (every copy of the case is ordered on its own stream)
    Check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!BuildMultiDomainModel(cfg, ne, ne, E, v, rho, dt, model))
        throw std::runtime_error("BuildMultiDomainModel failed");
    if (model.DomainCount() > kMaxDomains)
        throw std::runtime_error("too many domains");
    bcid.assign(bcid_in, bcid_in + ne);
    if (!dof.Build(model, bcid.data(), ne))
        throw std::runtime_error("GlobalDofMap::Build failed");

    BuildQuadTables(tables, gap);
    // element geometry (GeoInit / GetNodeInfo / PthPhyInit frame)
    geo.resize((size_t)ne);
    for (long i = 0; i < ne; ++i) {
        ElemGeo& g = geo[(size_t)i];
        double coord[3][8];
        for (int k = 0; k < 3; ++k) {
            for (int j = 0; j < 8; ++j)
                coord[k][j] = points[3 * (long)ele_pid[8 * i + j] + k];
            GetCoef(g.cc[k], coord[k]);
        }
        for (int k = 0; k < kNPW2; ++k)
            FromLocal(g.cc, tables.cenpw[k][0], tables.cenpw[k][1], g.rcenpw[k]);
        FromLocal(g.cc, 0.0, 0.0, g.center);
        g.radius = -1.0;
        for (int j = 0; j < 8; ++j) {
            double dd = Dist3(g.center, &points[3 * (long)ele_pid[8 * i + j]]);
            if (g.radius < dd)
                g.radius = dd;
        }
        g.radiusPW = g.radius / kNPW;
        TranslateMatrix(g.cc, tables.lc[0][0], tables.lc[0][1], g.trans);
    }
    FinishCreate();
}

// Models stacked as groups: each model's domains, interfaces, elements and
// geometry placed after those of the models before it (ids offset), the DOF
// map allocated model by model, one octree per model for the preconditioner
// and the options of the first model.  Every model keeps its own materials
// and time step (each domain's DynaMat carries its dt); the pairs of a domain
// never leave it, so the stacked system is block diagonal and each model's
// rows see exactly the blocks of its own job.
void Case::CreateGroup(const std::vector<const Case*>& ms)
{
    if (ms.empty())
        throw std::runtime_error("no models to group");
    if ((int)ms.size() > kGmGroupsMax)
        throw std::runtime_error("at most " + std::to_string(kGmGroupsMax) + " models per group");
    long total = 0;
    int nd = 0;
    for (const Case* m : ms) {
        if (!m || m->ne <= 0)
            throw std::runtime_error("a grouped model has no elements");
        if (m->tables.gap != ms[0]->tables.gap)
            throw std::runtime_error("grouped models need the same gap (quadrature tables)");
        total += m->ne;
        nd += m->model.DomainCount();
    }
    if (nd > kMaxDomains)
        throw std::runtime_error("grouped models hold more than " + std::to_string(kMaxDomains) + " domains");
    const Case& m0 = *ms[0];
    ne = total;
    dt = m0.dt;
    quadStaged = m0.quadStaged;
    nearPerEntry = m0.nearPerEntry;
    farParts = m0.farParts;
    convPacked = m0.convPacked;
    convFiltered = m0.convFiltered;
    convTiled = m0.convTiled;
    convTile = m0.convTile;
    gmBlocks = m0.gmBlocks;
    gmresDevice = m0.gmresDevice;
    Check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    model = MultiDomainModel();
    model.enabled = true;
    model.elementCount = model.nodeCount = ne;
    groupEle.assign(1, 0);
    bcid.clear();
    geo.clear();
    long eb = 0;
    int db = 0;
    for (const Case* m : ms) {
        for (Domain D : m->model.domains) {
            D.id += db;
            D.eleBegin += eb;
            D.nodeBegin += eb;
            for (long& e : D.elementIds)
                e += eb;
            model.domains.push_back(std::move(D));
        }
        for (InterfacePair f : m->model.interfaces) {
            f.domainA += db;
            f.domainB += db;
            f.eleA += eb;
            f.eleB += eb;
            model.interfaces.push_back(f);
        }
        bcid.insert(bcid.end(), m->bcid.begin(), m->bcid.end());
        geo.insert(geo.end(), m->geo.begin(), m->geo.end());
        eb += m->ne;
        db += m->model.DomainCount();
        groupEle.push_back(eb);
    }
    if (!model.Validate())
        throw std::runtime_error("grouped model validation failed");
    if (!dof.Build(model, bcid.data(), ne, &groupEle))
        throw std::runtime_error("GlobalDofMap::Build failed for the grouped models");
    tables = m0.tables;
    FinishCreate();
}

// Domain and surface flags, per-node references, the pair grid, the device
// copies and the per-element set-up kernels of a created case.
void Case::FinishCreate()
{
    domOf.assign((size_t)ne, 0);
    surface.assign((size_t)ne, 0);
    for (const Domain& d : model.domains)
        for (long e : d.elementIds)
            domOf[(size_t)e] = d.id;
    for (const InterfacePair& f : model.interfaces) {
        surface[(size_t)f.eleA] = 1;
        surface[(size_t)f.eleB] = 1;
    }
    // per-node references and interface transforms
    refs.resize((size_t)ne);
    for (long i = 0; i < ne; ++i) {
        NodeRefs& r = refs[(size_t)i];
        int d = domOf[(size_t)i];
        VariableRef u = dof.GetU(d, i), t = dof.GetT(d, i), hu = dof.GetHistoryU(d, i), ht = dof.GetHistoryT(d, i);
        r.u_block = (int)u.globalBlock;
        r.t_block = (int)t.globalBlock;
        r.hu_block = (int)hu.globalBlock;
        r.ht_block = (int)ht.globalBlock;
        r.u_sign = u.sign;
        r.t_sign = t.sign;
        r.hu_sign = hu.sign;
        r.ht_sign = ht.sign;
        r.u_known = u.known ? 1 : 0;
        r.t_known = t.known ? 1 : 0;
        r.bcid = bcid[(size_t)i];
        r.domain = d;
        r.surface = surface[(size_t)i];
        r.ref_side = 1;
        for (int k = 0; k < 9; ++k)
            r.itf_T[k] = (k % 4 == 0) ? 1.0 : 0.0;
        // GetElementLocalTransformFromInterfaceReference / IsInterfaceReferenceSide
        // (DBEM1/multidomain.cpp:1619-1658): first matching interface decides.
        for (const InterfacePair& f : model.interfaces) {
            if (f.domainA == d && f.eleA == i)
                break;
            if (f.domainB == d && f.eleB == i) {
                r.ref_side = 0;
                const double* side = geo[(size_t)f.eleB].trans;
                const double* ref = geo[(size_t)f.eleA].trans;
                for (int sc = 0; sc < 3; ++sc)
                    for (int rc = 0; rc < 3; ++rc) {
                        double dot = 0.0;
                        for (int ax = 0; ax < 3; ++ax)
                            dot += side[sc * 3 + ax] * ref[rc * 3 + ax];
                        r.itf_T[sc * 3 + rc] = dot;
                    }
                break;
            }
        }
    }
    // pair grid
    grid.nd = model.DomainCount();
    long pb = 0;
    for (int d = 0; d < grid.nd; ++d) {
        grid.pairBase[d] = pb;
        grid.eleBegin[d] = (int)model.domains[(size_t)d].eleBegin;
        grid.eleCount[d] = (int)model.domains[(size_t)d].eleCount;
        pb += (long)grid.eleCount[d] * grid.eleCount[d];
    }
    grid.pairBase[grid.nd] = pb;
    npairs = pb;

    // device upload and per-element setup kernels
    std::vector<DynaMat> mats;
    for (const Domain& d : model.domains)
        mats.push_back(d.mat);
    d_mats.upload(mats.data(), mats.size(), stream);
    grid.mats = d_mats.p;
    d_tables.upload(&tables, 1, stream);
    d_geo.upload(geo.data(), geo.size(), stream);
    d_dom.upload(domOf.data(), domOf.size(), stream);
    d_refs.upload(refs.data(), refs.size(), stream);
    (void)ScratchOnDevice();  // the no-allocation solve's scratch address on this device
    // GMRES host-visible state, set up once per context (no solve allocates it)
    if (!h_pin)
        Check(hipHostMalloc((void**)&h_pin, 4 * sizeof(double), hipHostMallocDefault), "pinned");
    if (!d_gst.p) {
        d_gst.alloc((sizeof(GmState) + sizeof(double) - 1) / sizeof(double));
        Check(hipHostMalloc((void**)&h_gst, sizeof(GmState), hipHostMallocDefault), "pinned");
    }
    d_pts.alloc((size_t)ne);
    d_u1.alloc(9 * (size_t)ne);
    d_t1.alloc(9 * (size_t)ne);
    d_t2.alloc(9 * (size_t)ne);
    double t0 = Now();
    LaunchPoints(d_geo.p, d_tables.p, d_pts.p, (int)ne, stream);
    LaunchSingular(d_geo.p, d_dom.p, d_mats.p, d_tables.p, d_u1.p, d_t1.p, d_t2.p, (int)ne, !quadStaged, stream);
    selfMode = quadStaged ? 1 : 0;
    Check(hipStreamSynchronize(stream), "setup kernels");
    timing[0] = Now() - t0;
    Check(hipGetLastError(), "setup launch");
}

void Case::EnsureSelfTables()
{
    int want = quadStaged ? 1 : 0;
    if (selfMode == want)
        return;
    double t0 = Now();
    LaunchSingular(d_geo.p, d_dom.p, d_mats.p, d_tables.p, d_u1.p, d_t1.p, d_t2.p, (int)ne, !quadStaged, stream);
    Check(hipStreamSynchronize(stream), "self tables");
    selfMode = want;
    timing[0] = Now() - t0;
}

void Case::Cull(int maxn)
{
    d_kb.alloc((size_t)npairs);
    d_len.alloc((size_t)npairs);
    d_near.alloc((size_t)npairs);
    double t0 = Now();
    LaunchCull(d_geo.p, d_dom.p, grid, npairs, maxn, d_kb.p, d_len.p, d_near.p, stream);
    Check(hipGetLastError(), "cull launch");
    // entry offsets and 64-lag chunk offsets of the bands, on the device
    d_off.alloc((size_t)npairs + 1);
    d_coff.alloc((size_t)npairs + 1);
    DevBuf<long> work;
    work.alloc((size_t)BandOffsetsWork(npairs));
    LaunchBandOffsets(d_len.p, npairs, work.p, d_off.p, d_coff.p, stream);
    long tot[2] = {0, 0};
    Check(CopyOnStream(&tot[0], d_off.p + npairs, sizeof(long), hipMemcpyDeviceToHost, stream), "entries");
    Check(CopyOnStream(&tot[1], d_coff.p + npairs, sizeof(long), hipMemcpyDeviceToHost, stream), "chunks");
    entries = tot[0];
    chunks = tot[1];
    hostBands = false;
    BuildNearLists(convTile);
    culledFor = maxn;
    timing[1] = Now() - t0;
}

// Host copies of the band starts, lengths and offsets (for export, statistics
// and persistence), fetched when first needed after a cull.
void Case::HostBands() const
{
    if (hostBands)
        return;
    h_len.resize((size_t)npairs);
    h_kb.resize((size_t)npairs);
    h_off.resize((size_t)npairs + 1);
    Check(CopyOnStream(h_len.data(), d_len.p, sizeof(int) * (size_t)npairs, hipMemcpyDeviceToHost, stream), "len");
    Check(CopyOnStream(h_kb.data(), d_kb.p, sizeof(int) * (size_t)npairs, hipMemcpyDeviceToHost, stream), "kb");
    Check(CopyOnStream(h_off.data(), d_off.p, sizeof(long) * ((size_t)npairs + 1), hipMemcpyDeviceToHost, stream),
          "off");
    hostBands = true;
}

// Per source row, the pairs whose band starts at a lag <= tile + 1: the only
// ones the near part of the time-blocked convolution touches (field order).
void Case::BuildNearLists(int tile)
{
    DevBuf<int> cnt;
    cnt.alloc((size_t)ne);
    LaunchNearLists(grid, d_refs.p, d_kb.p, d_len.p, tile, (int)ne, nullptr, cnt.p, nullptr, nullptr, stream);
    d_nearPtr.alloc((size_t)ne + 1);
    DevBuf<long> work, unused;
    work.alloc((size_t)BandOffsetsWork(ne));
    unused.alloc((size_t)ne + 1);
    LaunchBandOffsets(cnt.p, ne, work.p, d_nearPtr.p, unused.p, stream);
    long total = 0;
    Check(CopyOnStream(&total, d_nearPtr.p + ne, sizeof(long), hipMemcpyDeviceToHost, stream), "near count");
    d_nearPair.alloc((size_t)std::max<long>(1, total));
    d_nearField.alloc((size_t)std::max<long>(1, total));
    LaunchNearLists(grid, d_refs.p, d_kb.p, d_len.p, tile, (int)ne, d_nearPtr.p, nullptr, d_nearPair.p,
                    d_nearField.p, stream);
    Check(hipStreamSynchronize(stream), "near lists");
    nearTile = tile;
}

// The pairs that have step-0 blocks (near flag of the cull), in pair order.
std::vector<long> Case::NearPairs()
{
    std::vector<unsigned char> nearFlag((size_t)npairs);
    Check(CopyOnStream(nearFlag.data(), d_near.p, (size_t)npairs, hipMemcpyDeviceToHost, stream), "near");
    std::vector<long> list;
    for (long p = 0; p < npairs; ++p)
        if (nearFlag[(size_t)p])
            list.push_back(p);
    return list;
}

// Freeze (or release) the band structure and the judges of every history
// entry and step-0 pair at the current materials: later quadratures at other
// materials keep the same bands and take the same patch-wise / regular paths.
void Case::FreezeStructure(bool on)
{
    if (on)
        RefuseStreamed("freezing the structure");
    pwPacked = false;
    if (!on) {
        frozen = false;
        d_cls.free();
        d_cls0.free();
        return;
    }
    if (!historyDone || !step0Done)
        throw std::runtime_error("freezing needs assembled step 0 and history");
    d_cls.alloc((size_t)std::max<long>(1, entries));
    LaunchClassify(d_geo.p, grid, d_kb.p, d_len.p, d_off.p, npairs, d_cls.p, stream);
    std::vector<long> list = NearPairs();
    DevBuf<long> d_list;
    d_list.upload(list.data(), list.size(), stream);
    d_cls0.alloc(std::max<size_t>(1, list.size()));
    LaunchClassify0(d_geo.p, grid, d_list.p, (long)list.size(), d_cls0.p, stream);
    Check(hipStreamSynchronize(stream), "freeze");
    frozen = true;
}

// Step-0 matrices A (unknown) and KnownM (known) with the current materials
// (DBEM1/multidomain.cpp:2256-2327): device blocks of the near pairs, then the
// AddBlock builders on the host.
void Case::ComputeStep0(CCSRMat& Aout, CCSRMat& Kout)
{
    EnsureSelfTables();
    std::vector<long> list = NearPairs();
    DevBuf<long> d_list;
    d_list.upload(list.data(), list.size(), stream);
    DevBuf<double> d_out;
    DevBuf<int> d_flags;
    d_out.alloc(18 * list.size());
    d_flags.alloc(2 * list.size());
    LaunchStep0(d_geo.p, d_pts.p, d_dom.p, d_refs.p, d_u1.p, d_t1.p, d_t2.p, grid, d_list.p, (long)list.size(),
                frozen ? d_cls0.p : nullptr, d_out.p, d_flags.p, stream);
    std::vector<double> out(18 * list.size());
    std::vector<int> flags(2 * list.size());
    Check(hipStreamSynchronize(stream), "step0");
    Check(hipGetLastError(), "step0 launch");
    if (!list.empty()) {
        Check(CopyOnStream(out.data(), d_out.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost, stream), "step0 out");
        Check(CopyOnStream(flags.data(), d_flags.p, sizeof(int) * flags.size(), hipMemcpyDeviceToHost, stream), "step0 flags");
    }
    // AddBlock into the unknown / known builders (DBEM1/multidomain.cpp:1907-1937,2286-2321)
    MultiDomainCCSRBuilder bA(dof.equationBlockCount, dof.unknownBlockCount);
    MultiDomainCCSRBuilder bK(dof.equationBlockCount, dof.knownBlockCount);
    for (size_t i = 0; i < list.size(); ++i) {
        long p = list[i];
        int d = 0;
        while (d + 1 < grid.nd && p >= grid.pairBase[d + 1])
            ++d;
        long local = p - grid.pairBase[d];
        long s = grid.eleBegin[d] + local / grid.eleCount[d];
        long f = grid.eleBegin[d] + local % grid.eleCount[d];
        const double* b0 = &out[18 * i];
        const double* b1 = &out[18 * i + 9];
        if (surface[(size_t)f]) {
            VariableRef u = dof.GetU(d, f), t = dof.GetT(d, f);
            if (flags[2 * i])
                bA.AddBlock(s, u.globalBlock, u.sign, b0);
            if (flags[2 * i + 1])
                bA.AddBlock(s, t.globalBlock, t.sign, b1);
        } else {
            if (flags[2 * i]) {
                VariableRef r = bcid[(size_t)f] == 123 ? dof.GetT(d, f) : dof.GetU(d, f);
                if (!r.known)
                    bA.AddBlock(s, r.globalBlock, r.sign, b0);
            }
            if (flags[2 * i + 1]) {
                VariableRef r = bcid[(size_t)f] == 123 ? dof.GetU(d, f) : dof.GetT(d, f);
                if (r.known)
                    bK.AddBlock(s, r.globalBlock, r.sign, b1);
            }
        }
    }
    bA.Build(Aout);
    bK.Build(Kout);
}

void Case::AssembleStep0(long maxleaf)
{
    if (culledFor < 1)
        Cull(1);
    double t0 = Now();
    ComputeStep0(A, KnownM);
    if (!StructureValid())
        throw std::runtime_error("MultiDomain matrix structure check failed (zero rows/cols or non-finite blocks)");
    d_Arp.upload(A.row_ptr.data(), A.row_ptr.size(), stream);
    d_Acol.upload(A.col_index.data(), A.col_index.size(), stream);
    d_Aval.upload(A.value.data(), A.value.size(), stream);
    d_Krp.upload(KnownM.row_ptr.data(), KnownM.row_ptr.size(), stream);
    d_Kcol.upload(KnownM.col_index.data(), KnownM.col_index.size(), stream);
    d_Kval.upload(KnownM.value.data(), KnownM.value.size(), stream);
    // preconditioner over element centres
    std::vector<double> centres(3 * (size_t)ne);
    for (long i = 0; i < ne; ++i)
        for (int k = 0; k < 3; ++k)
            centres[3 * (size_t)i + (size_t)k] = geo[(size_t)i].center[k];
    if (!BuildMappedLeafPreConditioner(centres.data(), ne, maxleaf, A, dof, pre, Grouped() ? &groupEle : nullptr))
        throw std::runtime_error("preconditioner build failed");
    std::vector<long> posLeaf((size_t)ne), outBlock((size_t)ne);
    for (long l = 0; l < pre.leafCount; ++l)
        for (long j = 0; j < pre.leafSize[(size_t)l]; ++j)
            posLeaf[(size_t)(pre.leafBegin[(size_t)l] + j)] = l;
    for (long i = 0; i < ne; ++i)
        outBlock[(size_t)i] = dof.GetPreferredUnknownBlockForRow(pre.repid[(size_t)i]);
    d_posLeaf.upload(posLeaf.data(), posLeaf.size(), stream);
    d_leafBegin.upload(pre.leafBegin.data(), pre.leafBegin.size(), stream);
    d_leafSize.upload(pre.leafSize.data(), pre.leafSize.size(), stream);
    d_matOff.upload(pre.matOffset.data(), pre.matOffset.size(), stream);
    d_inv.upload(pre.inv.data(), pre.inv.size(), stream);
    d_repid.upload(pre.repid.data(), pre.repid.size(), stream);
    d_outBlock.upload(outBlock.data(), outBlock.size(), stream);
    Check(hipStreamSynchronize(stream), "step0 upload");
    step0Done = true;
    timing[3] = Now() - t0;
}

bool Case::StructureValid() const
{
    std::vector<long> cc((size_t)A.n, 0);
    for (long r = 0; r < A.n; ++r) {
        if (A.row_ptr[(size_t)r] == A.row_ptr[(size_t)r + 1])
            return false;
        for (int p = A.row_ptr[(size_t)r]; p < A.row_ptr[(size_t)r + 1]; ++p) {
            int c = A.col_index[(size_t)p];
            if (c < 0 || c >= A.n || !IsFiniteBlock9(&A.value[(size_t)p * 9]))
                return false;
            cc[(size_t)c]++;
        }
    }
    for (long c = 0; c < A.n; ++c)
        if (!cc[(size_t)c])
            return false;
    return true;
}

// The causality cull for a history assembly of maxn steps ahead of it, and
// the device bytes that assembly will allocate for the bands.
double Case::PrepareHistory(int maxn_in)
{
    if (maxn_in < 1)
        throw std::runtime_error("maxn must be >= 1");
    if (frozen)
        throw std::runtime_error("the frozen structure holds the bands of the last history assembly");
    EnsureSelfTables();
    Cull(maxn_in);
    precull = true;
    const double vbytes = mixed != 0 ? (mixed == 2 ? 6.0 : 4.0) : 8.0;
    return (double)entries * (15.0 * vbytes + 1.0);
}

// Frees the history bands; the step-0 matrices and the cull stay.
void Case::ReleaseHistory()
{
    h_mts.clear();
    h_mgs.clear();
    h_alive.clear();
    d_mts.free();
    d_mgs.free();
    d_mts32.free();
    d_mgs32.free();
    d_mtsLo.free();
    d_mgsLo.free();
    d_alive.free();
    ReleaseStreamed();
    historyDone = false;
}

void Case::AssembleHistory(int maxn_in)
{
    if (maxn_in < 1)
        throw std::runtime_error("maxn must be >= 1");
    if (frozen && maxn_in != maxn)
        throw std::runtime_error("the frozen structure holds the bands of the last history assembly");
    EnsureSelfTables();
    maxn = maxn_in;
    h_mts.clear();
    h_mgs.clear();
    h_alive.clear();
    d_mts.free();
    d_mgs.free();
    d_mts32.free();
    d_mgs32.free();
    d_mtsLo.free();
    d_mgsLo.free();
    d_alive.free();
    ReleaseStreamed();
    historyDone = false;
    historyMixed = mixed != 0;
    historySplit = mixed == 2;
    if (historyMixed && !quadStaged)
        throw std::runtime_error("the mixed-precision mode needs the staged history quadrature");
    if (!frozen && !(precull && culledFor == maxn))
        Cull(maxn);
    precull = false;
    size_t freeB = 0, totalB = 0;
    Check(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
    freeB += PooledBytesCached();
    const double vbytes = historyMixed ? (historySplit ? 6.0 : 4.0) : 8.0;
    double need = (double)entries * (15.0 * vbytes + 1.0);
    const double budget = historyBudget > 0.0 ? std::min(historyBudget, 0.92 * (double)freeB) : 0.92 * (double)freeB;
    if (historyStream == 1 || (historyStream == -1 && need > budget)) {
        AssembleStreamed(budget);
        return;
    }
    if (need > budget)
        throw std::runtime_error("history bands need " + std::to_string(need / 1e9) + " GB but only " +
                                 std::to_string(budget / 1e9) + " GB of device memory are available to them "
                                 "(option history_stream -1 or 1 regenerates them per tile instead)");
    if (historyMixed) {
        d_mts32.alloc(9 * (size_t)entries);
        d_mgs32.alloc(6 * (size_t)entries);
        if (historySplit) {
            d_mtsLo.alloc(9 * (size_t)entries);
            d_mgsLo.alloc(6 * (size_t)entries);
        }
    } else {
        d_mts.alloc(9 * (size_t)entries);
        d_mgs.alloc(6 * (size_t)entries);
    }
    d_alive.alloc((size_t)entries);
    hipEvent_t a, b;
    Check(hipEventCreate(&a), "event");
    Check(hipEventCreate(&b), "event");
    Check(hipEventRecord(a, stream), "event");
    if (quadStaged) {
        if (historyMixed)
            StagedQuadrature(nullptr, nullptr, d_mts32.p, d_mgs32.p, d_alive.p);
        else
            StagedQuadrature(d_mts.p, d_mgs.p, nullptr, nullptr, d_alive.p);
    } else {
        const long chunk = 1L << 24;
        for (long e0 = 0; e0 < entries; e0 += chunk) {
            long e1 = std::min(entries, e0 + chunk);
            LaunchHistory(d_geo.p, d_pts.p, d_dom.p, d_refs.p, d_t2.p, grid, d_off.p, d_kb.p, npairs, e0, e1,
                          d_mts.p, d_mgs.p, d_alive.p, stream);
        }
    }
    Check(hipEventRecord(b, stream), "event");
    Check(hipEventSynchronize(b), "history");
    Check(hipGetLastError(), "history launch");
    float ms = 0.f;
    Check(hipEventElapsedTime(&ms, a, b), "elapsed");
    timing[2] = ms * 1e-3;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    historyDone = true;
}

// ---------------------------------------------------------------------------
// Streamed history.  The reference holds every MTS[k] / MGS[k]
// (DBEM1/multidomain.cpp:3718-3790) and reads them every step (3840-3893).
// When the bands exceed the device budget they are regenerated instead: the
// cull (first lag, length and offset of every pair) stays resident, and so do
// the blocks of lags <= kStreamNearLag in a compact array, which the near part
// of every step reads; the far part of each 16-step tile recomputes the bands
// of one slab of source rows at a time (a slab's entries are contiguous,
// since a row's pairs are) into one slab buffer and convolves those rows.
// Regeneration runs the lane-per-entry quadrature, so the blocks equal the
// resident ones of that kernel (option quad_staged 0) bit for bit.
// ---------------------------------------------------------------------------
constexpr int kStreamNearLag = kConvTileMax;  // the near part of a 16-step tile reads lags <= 16

void Case::RefuseStreamed(const char* what) const
{
    if (historyStreamed)
        throw std::runtime_error(std::string(what) + " needs resident history bands; this case streams them "
                                 "(assemble with history_stream 0 and a budget that holds them)");
}

void Case::ReleaseStreamed()
{
    d_nearOff.free();
    d_nmts.free();
    d_nmgs.free();
    d_nalive.free();
    d_smts.free();
    d_smgs.free();
    d_salive.free();
    slabRow0.clear();
    slabRow1.clear();
    slabE0.clear();
    slabE1.clear();
    slabC0.clear();
    slabC1.clear();
    slabW0.clear();
    slabW1.clear();
    d_slabWaves.free();
    historyStreamed = false;
    timing[20] = 0.0;
}

void Case::StreamSlab(size_t i, int maxLag)
{
    const long e0 = slabE0[i], e1 = slabE1[i];
    if (streamStaged) {
        const long step = 1L << 20;  // wavefronts per launch
        for (long c = slabC0[i]; c < slabC1[i]; c += step)
            LaunchHistoryReg(d_geo.p, d_pts.p, d_refs.p, d_t2.p, grid, d_coff.p, d_off.p, d_kb.p, d_len.p,
                             (const unsigned char*)nullptr, npairs, c, std::min(slabC1[i], c + step), d_smts.p,
                             d_smgs.p, (unsigned short*)nullptr, (unsigned short*)nullptr, d_salive.p, stream, e0,
                             maxLag);
        for (long w = slabW0[i]; w < slabW1[i]; w += step)
            LaunchHistoryPw(d_geo.p, d_pts.p, d_refs.p, d_t2.p, grid, d_slabWaves.p, w, std::min(slabW1[i], w + step),
                            d_off.p, d_kb.p, d_len.p, (const unsigned char*)nullptr, d_smts.p, d_smgs.p,
                            (unsigned short*)nullptr, (unsigned short*)nullptr, d_salive.p, stream, e0, maxLag);
        return;
    }
    const long chunk = 1L << 24;
    for (long a = e0; a < e1; a += chunk)
        LaunchHistory(d_geo.p, d_pts.p, d_dom.p, d_refs.p, d_t2.p, grid, d_off.p, d_kb.p, npairs, a,
                      std::min(e1, a + chunk), d_smts.p, d_smgs.p, d_salive.p, stream, e0, maxLag);
}

void Case::AssembleStreamed(double budget)
{
    if (historyMixed)
        throw std::runtime_error("streamed history bands are FP64 (mixed 0)");
    if (frozen)
        throw std::runtime_error("the frozen structure holds resident bands");
    HostBands();
    // near blocks: lags kb .. min(kb + len - 1, kStreamNearLag) of every pair
    std::vector<long> nearOff((size_t)npairs + 1, 0);
    for (long p = 0; p < npairs; ++p) {
        const int k0 = h_kb[(size_t)p], L = h_len[(size_t)p];
        const long c = (L > 0 && k0 <= kStreamNearLag) ? std::min<long>(L, kStreamNearLag - k0 + 1) : 0;
        nearOff[(size_t)p + 1] = nearOff[(size_t)p] + c;
    }
    const long nearN = nearOff.back();
    const double perEntry = 15.0 * 8.0 + 1.0;
    // row slabs: greedy, each within what the budget leaves after the near blocks
    std::vector<long> rowPair((size_t)ne + 1, npairs);
    for (long r = 0; r < ne; ++r) {
        const int d = refs[(size_t)r].domain;
        rowPair[(size_t)r] = grid.pairBase[d] + (r - grid.eleBegin[d]) * (long)grid.eleCount[d];
    }
    long maxRow = 1;
    for (long r = 0; r < ne; ++r)
        maxRow = std::max(maxRow, h_off[(size_t)rowPair[(size_t)r + 1]] - h_off[(size_t)rowPair[(size_t)r]]);
    const double room = budget - perEntry * (double)nearN - 8.0 * (double)(npairs + 1);
    const long cap = std::max(maxRow, (long)std::min(room / perEntry, (double)entries));
    if (room < perEntry * (double)maxRow)
        throw std::runtime_error("the history budget does not hold the near blocks and one source row's bands (" +
                                 std::to_string((perEntry * (double)(nearN + maxRow)) / 1e9) + " GB)");
    for (long r = 0; r < ne;) {
        const long e0 = h_off[(size_t)rowPair[(size_t)r]];
        long r1 = r + 1;
        while (r1 < ne && h_off[(size_t)rowPair[(size_t)r1 + 1]] - e0 <= cap)
            ++r1;
        slabRow0.push_back(r);
        slabRow1.push_back(r1);
        slabE0.push_back(e0);
        slabE1.push_back(h_off[(size_t)rowPair[(size_t)r1]]);
        r = r1;
    }
    long slabMax = 1;
    for (size_t i = 0; i < slabE0.size(); ++i)
        slabMax = std::max(slabMax, slabE1[i] - slabE0[i]);
    streamStaged = quadStaged;
    if (streamStaged) {
        // regular chunks of a slab: its pairs' chunks, contiguous like the pairs
        std::vector<long> coff((size_t)npairs + 1);
        Check(CopyOnStream(coff.data(), d_coff.p, sizeof(long) * coff.size(), hipMemcpyDeviceToHost, stream), "coff");
        // patch-wise waves of a slab: each field element's lags of the slab's sources only
        DevBuf<int> d_pwcnt;
        d_pwcnt.alloc((size_t)npairs);
        LaunchPwCount(d_geo.p, grid, d_kb.p, d_len.p, d_off.p, nullptr, npairs, d_pwcnt.p, stream);
        DevBuf<long> d_nw, d_woff;
        d_nw.alloc((size_t)ne);
        std::vector<long> nw((size_t)ne);
        std::vector<std::vector<long>> woff(slabE0.size(), std::vector<long>((size_t)ne));
        long total = 0;
        for (size_t i = 0; i < slabE0.size(); ++i) {
            slabC0.push_back(coff[(size_t)rowPair[(size_t)slabRow0[i]]]);
            slabC1.push_back(coff[(size_t)rowPair[(size_t)slabRow1[i]]]);
            LaunchPwPack(grid, d_refs.p, d_pwcnt.p, (int)ne, nullptr, d_nw.p, nullptr, stream, slabRow0[i],
                         slabRow1[i]);
            Check(CopyOnStream(nw.data(), d_nw.p, sizeof(long) * nw.size(), hipMemcpyDeviceToHost, stream), "waves");
            slabW0.push_back(total);
            for (long f = 0; f < ne; ++f) {
                woff[i][(size_t)f] = total;
                total += nw[(size_t)f];
            }
            slabW1.push_back(total);
        }
        d_slabWaves.alloc((size_t)std::max<long>(1, total));
        d_woff.alloc((size_t)ne);
        for (size_t i = 0; i < slabE0.size(); ++i) {
            Check(hipMemcpyAsync(d_woff.p, woff[i].data(), sizeof(long) * (size_t)ne, hipMemcpyHostToDevice, stream),
                  "waves");
            LaunchPwPack(grid, d_refs.p, d_pwcnt.p, (int)ne, d_woff.p, nullptr, d_slabWaves.p, stream, slabRow0[i],
                         slabRow1[i]);
            Check(hipStreamSynchronize(stream), "waves");  // woff[i] is read before the next copy
        }
    }
    d_nearOff.upload(nearOff.data(), nearOff.size(), stream);
    d_nmts.alloc(9 * (size_t)std::max<long>(1, nearN));
    d_nmgs.alloc(6 * (size_t)std::max<long>(1, nearN));
    d_nalive.alloc((size_t)std::max<long>(1, nearN));
    d_smts.alloc(9 * (size_t)slabMax);
    d_smgs.alloc(6 * (size_t)slabMax);
    d_salive.alloc((size_t)slabMax);
    hipEvent_t a, b;
    Check(hipEventCreate(&a), "event");
    Check(hipEventCreate(&b), "event");
    Check(hipEventRecord(a, stream), "event");
    for (size_t i = 0; i < slabE0.size(); ++i) {
        StreamSlab(i, kStreamNearLag);
        LaunchNearGather(d_off.p, d_kb.p, d_nearOff.p, rowPair[(size_t)slabRow0[i]], rowPair[(size_t)slabRow1[i]],
                         slabE0[i], d_smts.p, d_smgs.p, d_salive.p, d_nmts.p, d_nmgs.p, d_nalive.p, stream);
    }
    Check(hipEventRecord(b, stream), "event");
    Check(hipEventSynchronize(b), "streamed history");
    Check(hipGetLastError(), "streamed history launch");
    float ms = 0.f;
    Check(hipEventElapsedTime(&ms, a, b), "elapsed");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    timing[2] = ms * 1e-3;
    timing[8] = ms * 1e-3;
    timing[20] = (double)slabE0.size();
    historyStreamed = true;
    historyDone = true;
}

void Case::PackPwWaves()
{
    if (frozen && pwPacked)
        return;
    const unsigned char* cls = frozen ? d_cls.p : nullptr;
    DevBuf<int> d_pwcnt;
    d_pwcnt.alloc((size_t)npairs);
    LaunchPwCount(d_geo.p, grid, d_kb.p, d_len.p, d_off.p, cls, npairs, d_pwcnt.p, stream);
    DevBuf<long> d_nw;
    d_nw.alloc((size_t)ne);
    LaunchPwPack(grid, d_refs.p, d_pwcnt.p, (int)ne, nullptr, d_nw.p, nullptr, stream);
    std::vector<long> nw((size_t)ne), woff((size_t)ne);
    Check(hipMemcpyAsync(nw.data(), d_nw.p, sizeof(long) * (size_t)ne, hipMemcpyDeviceToHost, stream), "pw pack");
    Check(hipStreamSynchronize(stream), "pw pack");
    // waves in field-element order (domains are contiguous element ranges)
    long total = 0;
    for (long f = 0; f < ne; ++f) {
        woff[(size_t)f] = total;
        total += nw[(size_t)f];
    }
    pwWaves = total;
    DevBuf<long> d_woff;
    d_woff.upload(woff.data(), woff.size(), stream);
    d_pwWaves.alloc((size_t)std::max<long>(1, pwWaves));
    LaunchPwPack(grid, d_refs.p, d_pwcnt.p, (int)ne, d_woff.p, nullptr, d_pwWaves.p, stream);
    Check(hipStreamSynchronize(stream), "pw pack");
    if (quadCount) {  // lane occupancy of the patch-wise waves: entries / (64 waves)
        std::vector<int> cnt((size_t)npairs);
        Check(CopyOnStream(cnt.data(), d_pwcnt.p, sizeof(int) * (size_t)npairs, hipMemcpyDeviceToHost, stream),
              "pw counts");
        double tot = 0.0;
        for (int c : cnt)
            tot += c;
        timing[43] = tot;
        timing[44] = (double)pwWaves;
    }
    pwPacked = frozen;
}

// Staged history quadrature of the current bands with the current materials
// into mts / mgs (FP64) or mts32 / mgs32 (FP32, plus the bf16 residuals of the
// split mode); alive receives the presence bits.
void Case::StagedQuadrature(double* mts, double* mgs, float* mts32, float* mgs32, unsigned char* alive)
{
    // patch-wise lags per pair, then waves that pack the patch-wise lags
    // of consecutive sources with the same field element
    const unsigned char* cls = frozen ? d_cls.p : nullptr;
    PackPwWaves();
    hipEvent_t qa, qb;
    Check(hipEventCreate(&qa), "event");
    Check(hipEventCreate(&qb), "event");
    Check(hipEventRecord(qa, stream), "event");
    const long step = 1L << 20;  // wavefronts per launch
    DevBuf<unsigned long long> d_qc;
    if (quadCount) {
        d_qc.alloc(4);
        LaunchZero(reinterpret_cast<double*>(d_qc.p), 4, stream);
    }
    auto launch = [&](auto* mts, auto* mgs) {
        for (long c0 = 0; c0 < chunks; c0 += step)
            LaunchHistoryReg(d_geo.p, d_pts.p, d_refs.p, d_t2.p, grid, d_coff.p, d_off.p, d_kb.p, d_len.p, cls,
                             npairs, c0, std::min(chunks, c0 + step), mts, mgs, d_mtsLo.p, d_mgsLo.p, alive, stream,
                             0, 1 << 30, d_qc.p);
        for (long w0 = 0; w0 < pwWaves; w0 += step)
            LaunchHistoryPw(d_geo.p, d_pts.p, d_refs.p, d_t2.p, grid, d_pwWaves.p, w0, std::min(pwWaves, w0 + step),
                            d_off.p, d_kb.p, d_len.p, cls, mts, mgs, d_mtsLo.p, d_mgsLo.p, alive, stream, 0, 1 << 30,
                            d_qc.p);
    };
    if (mts32)
        launch(mts32, mgs32);
    else
        launch(mts, mgs);
    Check(hipEventRecord(qb, stream), "event");
    Check(hipEventSynchronize(qb), "history");
    float ms = 0.f;
    Check(hipEventElapsedTime(&ms, qa, qb), "elapsed");
    timing[8] = ms * 1e-3;
    for (int i = 0; i < 4; ++i)
        timing[22 + i] = 0.0;
    if (quadCount) {
        unsigned long long qc[4];
        Check(CopyOnStream(qc, d_qc.p, sizeof(qc), hipMemcpyDeviceToHost, stream), "evaluation counts");
        for (int i = 0; i < 4; ++i)
            timing[22 + i] = (double)qc[i];
    }
    (void)hipEventDestroy(qa);
    (void)hipEventDestroy(qb);
}

void Case::HistoryStats(long* counts) const
{
    if (!historyDone)
        throw std::runtime_error("history not assembled");
    HostBands();
    for (int j = 0; j < maxn; ++j)
        counts[j] = 0;
    for (long p = 0; p < npairs; ++p) {
        int k0 = h_kb[(size_t)p], L = h_len[(size_t)p];
        for (int t = 0; t < L; ++t)
            counts[k0 + t - 1]++;
    }
}

void Case::ExportMatrix(int which, int step, std::vector<int>& rp, std::vector<int>& col, std::vector<double>& val)
{
    if (which == 0 || which == 1) {
        const CCSRMat& M = which == 0 ? A : KnownM;
        if (!step0Done)
            throw std::runtime_error("step 0 not assembled");
        rp = M.row_ptr;
        col = M.col_index;
        val = M.value;
        return;
    }
    RefuseStreamed("exporting history blocks");
    if (!historyDone || step < 1 || step > maxn)
        throw std::runtime_error("history step not assembled");
    if (h_mts.empty()) {
        h_mts.resize(9 * (size_t)entries);
        h_mgs.resize(6 * (size_t)entries);
        h_alive.resize((size_t)entries);
        if (historyMixed) {
            std::vector<float> a(h_mts.size()), b(h_mgs.size());
            Check(CopyOnStream(a.data(), d_mts32.p, sizeof(float) * a.size(), hipMemcpyDeviceToHost, stream), "mts");
            Check(CopyOnStream(b.data(), d_mgs32.p, sizeof(float) * b.size(), hipMemcpyDeviceToHost, stream), "mgs");
            std::copy(a.begin(), a.end(), h_mts.begin());
            std::copy(b.begin(), b.end(), h_mgs.begin());
            if (historySplit) {  // value = hi + bf16 residual
                std::vector<unsigned short> la(h_mts.size()), lb(h_mgs.size());
                Check(CopyOnStream(la.data(), d_mtsLo.p, 2 * la.size(), hipMemcpyDeviceToHost, stream), "mts lo");
                Check(CopyOnStream(lb.data(), d_mgsLo.p, 2 * lb.size(), hipMemcpyDeviceToHost, stream), "mgs lo");
                auto lo = [](unsigned short v) {
                    unsigned int bits = (unsigned int)v << 16;
                    float f;
                    std::memcpy(&f, &bits, 4);
                    return (double)f;
                };
                for (size_t i = 0; i < la.size(); ++i)
                    h_mts[i] += lo(la[i]);
                for (size_t i = 0; i < lb.size(); ++i)
                    h_mgs[i] += lo(lb[i]);
            }
        } else {
            Check(CopyOnStream(h_mts.data(), d_mts.p, sizeof(double) * h_mts.size(), hipMemcpyDeviceToHost, stream), "mts");
            Check(CopyOnStream(h_mgs.data(), d_mgs.p, sizeof(double) * h_mgs.size(), hipMemcpyDeviceToHost, stream), "mgs");
        }
        Check(CopyOnStream(h_alive.data(), d_alive.p, h_alive.size(), hipMemcpyDeviceToHost, stream), "alive");
    }
    HostBands();
    struct Item {
        long row, col;
        long e;
    };
    std::vector<Item> items;
    for (int d = 0; d < grid.nd; ++d) {
        long n = grid.eleCount[d];
        for (long q = 0; q < n * n; ++q) {
            long p = grid.pairBase[d] + q;
            int L = h_len[(size_t)p], k0 = h_kb[(size_t)p];
            if (step < k0 || step >= k0 + L)
                continue;
            long e = h_off[(size_t)p] + (step - k0);
            unsigned char bit = which == 2 ? 1 : 2;
            if (!(h_alive[(size_t)e] & bit))
                continue;
            long s = grid.eleBegin[d] + q / n, f = grid.eleBegin[d] + q % n;
            long c = which == 2 ? refs[(size_t)f].hu_block : refs[(size_t)f].ht_block;
            items.push_back(Item{s, c, e});
        }
    }
    std::sort(items.begin(), items.end(),
              [](const Item& x, const Item& y) { return x.row != y.row ? x.row < y.row : x.col < y.col; });
    rp.assign((size_t)ne + 1, 0);
    col.resize(items.size());
    val.resize(items.size() * 9);
    for (size_t i = 0; i < items.size(); ++i) {
        rp[(size_t)items[i].row + 1]++;
        col[i] = (int)items[i].col;
        if (which == 2) {
            for (int k = 0; k < 9; ++k)
                val[i * 9 + (size_t)k] = h_mts[9 * (size_t)items[i].e + (size_t)k];
        } else {
            const double* g = &h_mgs[6 * (size_t)items[i].e];
            const double full[9] = {g[0], g[1], g[2], g[1], g[3], g[4], g[2], g[4], g[5]};
            for (int k = 0; k < 9; ++k)
                val[i * 9 + (size_t)k] = full[k];
        }
    }
    for (long r = 0; r < ne; ++r)
        rp[(size_t)r + 1] += rp[(size_t)r];
}

// MTS[step] / MGS[step] restricted to source rows row0 .. row1-1: only the
// band slice of those rows' pairs is copied from the device (a source row's
// pairs, and so its band entries, are contiguous), so sampled rows of a
// large case export without moving the whole history to the host.
void Case::ExportRows(int which, int step, long row0, long row1, std::vector<int>& rp, std::vector<int>& col,
                      std::vector<double>& val)
{
    if (which != 2 && which != 3)
        throw std::runtime_error("row export covers the history matrices (2, 3)");
    if (which >= 2)
        RefuseStreamed("exporting history blocks");
    if (!historyDone || step < 1 || step > maxn)
        throw std::runtime_error("history step not assembled");
    if (row0 < 0 || row1 > ne || row0 >= row1)
        throw std::runtime_error("row range out of bounds");
    HostBands();
    rp.assign((size_t)(row1 - row0) + 1, 0);
    col.clear();
    val.clear();
    for (long s = row0; s < row1; ++s) {
        int d = 0;
        while (d + 1 < grid.nd && s >= grid.eleBegin[d + 1])
            ++d;
        const long n = grid.eleCount[d];
        const long p0 = grid.pairBase[d] + (s - grid.eleBegin[d]) * n, p1 = p0 + n;
        const long e0 = h_off[(size_t)p0], e1 = h_off[(size_t)p1], m = e1 - e0;
        std::vector<double> mts(9 * (size_t)m), mgs(6 * (size_t)m);
        std::vector<unsigned char> alive((size_t)m);
        if (m > 0) {
            if (historyMixed) {
                std::vector<float> a(mts.size()), b(mgs.size());
                Check(CopyOnStream(a.data(), d_mts32.p + 9 * e0, sizeof(float) * a.size(), hipMemcpyDeviceToHost,
                                   stream), "mts rows");
                Check(CopyOnStream(b.data(), d_mgs32.p + 6 * e0, sizeof(float) * b.size(), hipMemcpyDeviceToHost,
                                   stream), "mgs rows");
                std::copy(a.begin(), a.end(), mts.begin());
                std::copy(b.begin(), b.end(), mgs.begin());
                if (historySplit) {
                    std::vector<unsigned short> la(mts.size()), lb(mgs.size());
                    Check(CopyOnStream(la.data(), d_mtsLo.p + 9 * e0, 2 * la.size(), hipMemcpyDeviceToHost, stream),
                          "mts lo rows");
                    Check(CopyOnStream(lb.data(), d_mgsLo.p + 6 * e0, 2 * lb.size(), hipMemcpyDeviceToHost, stream),
                          "mgs lo rows");
                    auto lo = [](unsigned short v) {
                        unsigned int bits = (unsigned int)v << 16;
                        float f;
                        std::memcpy(&f, &bits, 4);
                        return (double)f;
                    };
                    for (size_t i = 0; i < la.size(); ++i)
                        mts[i] += lo(la[i]);
                    for (size_t i = 0; i < lb.size(); ++i)
                        mgs[i] += lo(lb[i]);
                }
            } else {
                Check(CopyOnStream(mts.data(), d_mts.p + 9 * e0, sizeof(double) * mts.size(), hipMemcpyDeviceToHost,
                                   stream), "mts rows");
                Check(CopyOnStream(mgs.data(), d_mgs.p + 6 * e0, sizeof(double) * mgs.size(), hipMemcpyDeviceToHost,
                                   stream), "mgs rows");
            }
            Check(CopyOnStream(alive.data(), d_alive.p + e0, (size_t)m, hipMemcpyDeviceToHost, stream), "alive rows");
        }
        std::vector<std::pair<long, long>> items;  // (column block, entry - e0)
        for (long p = p0; p < p1; ++p) {
            const int L = h_len[(size_t)p], k0 = h_kb[(size_t)p];
            if (step < k0 || step >= k0 + L)
                continue;
            const long e = h_off[(size_t)p] + (step - k0) - e0;
            if (!(alive[(size_t)e] & (which == 2 ? 1 : 2)))
                continue;
            const long f = grid.eleBegin[d] + (p - p0);
            items.emplace_back(which == 2 ? refs[(size_t)f].hu_block : refs[(size_t)f].ht_block, e);
        }
        std::sort(items.begin(), items.end());
        for (const auto& it : items) {
            col.push_back((int)it.first);
            if (which == 2) {
                for (int k = 0; k < 9; ++k)
                    val.push_back(mts[9 * (size_t)it.second + (size_t)k]);
            } else {
                const double* g = &mgs[6 * (size_t)it.second];
                const double full[9] = {g[0], g[1], g[2], g[1], g[3], g[4], g[2], g[4], g[5]};
                val.insert(val.end(), full, full + 9);
            }
        }
        rp[(size_t)(s - row0) + 1] = (int)col.size();
    }
}

// ---------------------------------------------------------------------------
// GMRES (DBEM1/pre_gmres.cpp:1369-1605) with maxit = 1 and x0 = 0.  The
// Arnoldi step uses classical Gram-Schmidt applied twice (two fused
// multi-dot / multi-axpy passes instead of j sequential MGS passes); the
// Hessenberg/Givens update, W = V R^-1 and the true residual test
// ||b - A x|| <= tol ||b|| follow the reference.
// ---------------------------------------------------------------------------
void Case::PrepareGmres(long restart)
{
    long n = 3 * ne;
    if (gmBlocks < 1 || gmBlocks > kGmBlocksMax)
        throw std::runtime_error("GMRES reduction blocks must be 1 .. " + std::to_string(kGmBlocksMax));
    // reduction partition: few blocks keep the in-order finish short
    gm_blocks = (int)std::min<long>(gmBlocks, std::max<long>(1, (n + 255) / 256));
    if (gm_restart == restart && gm_n == n)
        return;
    gm_restart = restart;
    gm_n = n;
    d_V.alloc((size_t)n * (size_t)(restart + 1));
    d_W.alloc((size_t)n * (size_t)restart);
    d_u.alloc((size_t)n);
    d_u2.alloc((size_t)n);
    d_vh.alloc((size_t)n);
    d_h.alloc((size_t)restart + 2);
    d_h2.alloc((size_t)restart + 2);
    d_QT.alloc((size_t)(restart + 1) * (size_t)(restart + 1));
    d_R.alloc((size_t)restart * (size_t)restart);
    d_f.alloc((size_t)restart + 2);
    d_scal.alloc(4);
    d_stag.alloc(1);
    d_part.alloc((size_t)kGmBlocksMax * (size_t)(restart + 1));
    if (!h_pin)
        Check(hipHostMalloc((void**)&h_pin, 4 * sizeof(double), hipHostMallocDefault), "pinned");
    gm_mixed = false;
}

// FP32 copies of A and the leaf inverses and the FP32 Krylov workspace of the
// mixed-precision inner solves.
void Case::PrepareMixed(long restart)
{
    if (gm_mixed)
        return;
    long n = 3 * ne;
    d_V32.alloc((size_t)n * (size_t)(restart + 1));
    d_W32.alloc((size_t)n * (size_t)restart);
    d_u32.alloc((size_t)n);
    d_u232.alloc((size_t)n);
    d_r.alloc((size_t)n);
    d_dx.alloc((size_t)n);
    d_Aval32.alloc(d_Aval.n);
    LaunchToFloat(d_Aval.p, d_Aval32.p, (long)d_Aval.n, stream);
    d_inv32.alloc(d_inv.n);
    LaunchToFloat(d_inv.p, d_inv32.p, (long)d_inv.n, stream);
    Check(hipStreamSynchronize(stream), "mixed setup");
    gm_mixed = true;
}

// Left-preconditioned GMRES (DBEM1/pre_gmres.cpp:1369-1605) with the matrix,
// inverses and Krylov vectors stored as T; reductions, Hessenberg, Givens and
// the solution x are FP64.  T = double is the parity path.
template <typename T>
int Case::GmresCore(const double* d_b, double* d_x, long restart, double tol, int* iters, const GmresOp<T>& op, T* V,
                    T* W, T* u, T* u2)
{
    const long n = 3 * ne;
    auto applyPre = [&](auto in, T* out) {
        if constexpr (std::is_same<T, double>::value) {
            if (op.prp) {
                LaunchSpmv(ne, op.prp, op.pcol, op.pval, in, out, stream);
                return;
            }
        }
        LaunchPrecond(ne, d_posLeaf.p, d_leafBegin.p, d_leafSize.p, d_matOff.p, op.inv, d_repid.p, d_outBlock.p, in,
                      out, stream);
    };
    double* normr = d_scal.p;      // [0]
    double* bnorm = d_scal.p + 1;  // [1]
    double* phibar = d_scal.p + 2; // [2]
    *iters = 0;
    LaunchNorm(d_b, n, d_part.p, gm_blocks, bnorm, stream);
    Check(hipMemcpyAsync(h_pin, bnorm, sizeof(double), hipMemcpyDeviceToHost, stream), "d2h");
    Check(hipStreamSynchronize(stream), "gmres");
    double n2b = h_pin[0];
    LaunchZero(d_x, n, stream);
    if (n2b == 0.0)
        return 0;
    double tolb = tol * n2b;
    // r = b - A x0 (x0 = 0)
    LaunchSpmv(ne, op.rp, op.col, op.val, (const double*)d_x, d_vh.p, stream);
    LaunchResidual(d_b, d_vh.p, n, stream);
    LaunchNorm((const double*)d_vh.p, n, d_part.p, gm_blocks, normr, stream);
    Check(hipMemcpyAsync(h_pin, normr, sizeof(double), hipMemcpyDeviceToHost, stream), "d2h");
    Check(hipStreamSynchronize(stream), "gmres");
    if (h_pin[0] <= tolb)
        return 0;
    Check(hipMemsetAsync(d_QT.p, 0, sizeof(double) * d_QT.n, stream), "memset");
    Check(hipMemsetAsync(d_R.p, 0, sizeof(double) * d_R.n, stream), "memset");
    Check(hipMemsetAsync(d_stag.p, 0, sizeof(int), stream), "memset");
    applyPre((const double*)d_vh.p, u);
    LaunchNorm((const T*)u, n, d_part.p, gm_blocks, d_h.p + 1, stream);
    LaunchScaleCopy((const T*)u, d_h.p + 1, 1, V, n, stream);
    const double one = 1.0;
    Check(hipMemcpyAsync(d_QT.p, &one, sizeof(double), hipMemcpyHostToDevice, stream), "qt");
    Check(hipMemcpyAsync(phibar, d_h.p + 1, sizeof(double), hipMemcpyDeviceToDevice, stream), "phibar");
    for (int j = 1; j <= restart; ++j) {
        T* Vj = V + (size_t)(j - 1) * (size_t)n;
        LaunchSpmv(ne, op.rp, op.col, op.val, (const T*)Vj, u2, stream);
        applyPre((const T*)u2, u);
        LaunchMultiDot((const T*)V, n, j, (const T*)u, n, d_part.p, gm_blocks, d_h.p + 1, stream);
        LaunchMultiAxpy((const T*)V, n, j, d_h.p + 1, u, n, nullptr, stream);
        LaunchMultiDot((const T*)V, n, j, (const T*)u, n, d_part.p, gm_blocks, d_h2.p, stream);
        LaunchMultiAxpy((const T*)V, n, j, d_h2.p, u, n, d_h.p + 1, stream);
        LaunchNorm((const T*)u, n, d_part.p, gm_blocks, d_h.p + j + 1, stream);
        LaunchScaleCopy((const T*)u, d_h.p + j + 1, 1, V + (size_t)j * (size_t)n, n, stream);
        LaunchGivens(d_h.p, d_QT.p, d_R.p, d_f.p, phibar, d_stag.p, j, (int)restart, stream);
        LaunchWUpdate((const T*)V, W, n, d_R.p, j, (int)restart, d_f.p, d_x, stream);
        LaunchSpmv(ne, op.rp, op.col, op.val, (const double*)d_x, d_vh.p, stream);
        LaunchResidual(d_b, d_vh.p, n, stream);
        LaunchNorm((const double*)d_vh.p, n, d_part.p, gm_blocks, normr, stream);
        Check(hipMemcpyAsync(h_pin, normr, sizeof(double), hipMemcpyDeviceToHost, stream), "d2h");
        Check(hipMemcpyAsync(h_pin + 3, d_stag.p, sizeof(int), hipMemcpyDeviceToHost, stream), "d2h");
        Check(hipStreamSynchronize(stream), "gmres iteration");
        *iters = j;
        if (h_pin[0] <= tolb)
            return 0;
        int stag = 0;
        std::memcpy(&stag, h_pin + 3, sizeof(int));
        if (stag)
            return 3;
    }
    return 1;
}

template int Case::GmresCore<double>(const double*, double*, long, double, int*, const GmresOp<double>&, double*,
                                     double*, double*, double*);
template int Case::GmresCore<float>(const double*, double*, long, double, int*, const GmresOp<float>&, float*, float*,
                                    float*, float*);

int Case::Gmres(const double* d_b, double* d_x, long restart, double tol, int* iters)
{
    GmresOp<double> op{d_Arp.p, d_Acol.p, d_Aval.p, d_inv.p, nullptr, nullptr, nullptr};
    return GmresCore<double>(d_b, d_x, restart, tol, iters, op, d_V.p, d_W.p, d_u.p, d_u2.p);
}

// Field by field (the struct has padding): equal when the captured kernels
// would be given the same operands.
static bool SameGmresArgs(const GmresDevArgs& p, const GmresDevArgs& q)
{
    return p.ne == q.ne && p.restart == q.restart && p.nblocks == q.nblocks && p.rp == q.rp && p.col == q.col &&
           p.val == q.val && p.posLeaf == q.posLeaf && p.leafBegin == q.leafBegin && p.leafSize == q.leafSize &&
           p.matOff == q.matOff && p.repid == q.repid && p.outBlock == q.outBlock && p.inv == q.inv &&
           p.prp == q.prp && p.pcol == q.pcol && p.pval == q.pval && p.V == q.V && p.W == q.W && p.u == q.u &&
           p.u2 == q.u2 && p.vh == q.vh && p.h == q.h && p.h2 == q.h2 && p.QT == q.QT && p.R == q.R && p.f == q.f &&
           p.phibar == q.phibar && p.part == q.part && p.x == q.x && p.nQT == q.nQT && p.nR == q.nR &&
           p.stag == q.stag && p.st == q.st;
}

// The same solve driven from the device (gmres.hip): the opening kernels and
// one iteration are captured once as hipGraphs (gmresDevice 1) or launched
// directly (2); iterations are queued ahead, as many as the previous step
// needed, and the host reads the device state once they have run, queueing
// two more at a time until the state says done.
int Case::GmresDevice(const double* d_b, double* d_x, long restart, double tol, int* iters,
                      const GmresOp<double>* op)
{
    if (!d_gst.p) {
        d_gst.alloc((sizeof(GmState) + sizeof(double) - 1) / sizeof(double));
        Check(hipHostMalloc((void**)&h_gst, sizeof(GmState), hipHostMallocDefault), "pinned");
    }
    GmresDevArgs a{};
    a.ne = ne;
    a.restart = restart;
    a.nblocks = gm_blocks;
    a.rp = op ? op->rp : d_Arp.p;
    a.col = op ? op->col : d_Acol.p;
    a.val = op ? op->val : d_Aval.p;
    a.prp = op ? op->prp : nullptr;
    a.pcol = op ? op->pcol : nullptr;
    a.pval = op ? op->pval : nullptr;
    a.posLeaf = d_posLeaf.p;
    a.leafBegin = d_leafBegin.p;
    a.leafSize = d_leafSize.p;
    a.matOff = d_matOff.p;
    a.repid = d_repid.p;
    a.outBlock = d_outBlock.p;
    a.inv = op && op->inv ? op->inv : d_inv.p;
    a.V = d_V.p;
    a.W = d_W.p;
    a.u = d_u.p;
    a.u2 = d_u2.p;
    a.vh = d_vh.p;
    a.h = d_h.p;
    a.h2 = d_h2.p;
    a.QT = d_QT.p;
    a.R = d_R.p;
    a.f = d_f.p;
    a.phibar = d_scal.p + 2;
    a.part = d_part.p;
    a.x = d_x;
    a.nQT = (long)d_QT.n;
    a.nR = (long)d_R.n;
    a.stag = d_stag.p;
    a.st = reinterpret_cast<GmState*>(d_gst.p);
    const bool useGraph = gmresDevice == 1;
    if (useGraph && !(graphValid && graphB == d_b && graphTol == tol && SameGmresArgs(graphArgs, a))) {
        ReleaseGraphs();
        hipGraph_t g = nullptr;
        Check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
        LaunchGmresBegin(a, d_b, tol, stream);
        Check(hipStreamEndCapture(stream, &g), "capture");
        Check(hipGraphInstantiate(&graphBegin, g, nullptr, nullptr, 0), "instantiate");
        Check(hipGraphDestroy(g), "graph");
        Check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
        LaunchGmresIteration(a, d_b, stream);
        Check(hipStreamEndCapture(stream, &g), "capture");
        Check(hipGraphInstantiate(&graphIter, g, nullptr, nullptr, 0), "instantiate");
        Check(hipGraphDestroy(g), "graph");
        graphB = d_b;
        graphTol = tol;
        graphArgs = a;
        graphValid = true;
    }
    if (useGraph)
        Check(hipGraphLaunch(graphBegin, stream), "graph launch");
    else
        LaunchGmresBegin(a, d_b, tol, stream);
    long queued = 0;
    long ahead = std::max(1, gm_lastIters);
    for (;;) {
        long k = std::min(ahead, restart - queued);
        for (long i = 0; i < k; ++i) {
            if (useGraph)
                Check(hipGraphLaunch(graphIter, stream), "graph launch");
            else
                LaunchGmresIteration(a, d_b, stream);
        }
        queued += k;
        Check(hipMemcpyAsync(h_gst, d_gst.p, sizeof(GmState), hipMemcpyDeviceToHost, stream), "d2h");
        Check(hipStreamSynchronize(stream), "gmres");
        if (h_gst->done || queued >= restart)
            break;
        ahead = 2;
    }
    if (!h_gst->done)
        throw std::runtime_error("device GMRES ended without a stopping state");
    *iters = h_gst->j;
    gm_lastIters = h_gst->j;
    return h_gst->flag;
}

void Case::ReleaseGraphs()
{
    if (graphBegin)
        (void)hipGraphExecDestroy(graphBegin);
    if (graphIter)
        (void)hipGraphExecDestroy(graphIter);
    graphBegin = graphIter = nullptr;
    graphB = nullptr;
    graphValid = false;
}

// Mixed precision: FP32-stored inner GMRES (GmresCore<float>) on A d = r
// under an FP64 iterative-refinement residual r = b - A x with the FP64 A,
// until ||r|| <= tol ||b|| (the reference's stopping test,
// DBEM1/pre_gmres.cpp:1572).  Returns 0, 1 (no convergence within kMaxRefine
// rounds) or 3 (stagnation); iters = total inner iterations.
int Case::GmresMixed(const double* d_b, double* d_x, long restart, double tol, int* iters)
{
    constexpr int kMaxRefine = 12;
    constexpr double kInnerTol = 1e-6;
    const long n = 3 * ne;
    *iters = 0;
    double* scal = d_scal.p + 3;
    LaunchNorm(d_b, n, d_part.p, gm_blocks, scal, stream);
    Check(hipMemcpyAsync(h_pin, scal, sizeof(double), hipMemcpyDeviceToHost, stream), "d2h");
    Check(hipStreamSynchronize(stream), "gmres");
    double n2b = h_pin[0];
    LaunchZero(d_x, n, stream);
    if (n2b == 0.0)
        return 0;
    Check(hipMemcpyAsync(d_r.p, d_b, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, stream), "r");
    double normr = n2b;
    int total = 0;
    for (int round = 0; round < kMaxRefine; ++round) {
        double itol = std::max(kInnerTol, tol * n2b / normr);
        int it = 0;
        GmresOp<float> op{d_Arp.p, d_Acol.p, d_Aval32.p, d_inv32.p, nullptr, nullptr, nullptr};
        int f = GmresCore<float>(d_r.p, d_dx.p, restart, itol, &it, op, d_V32.p, d_W32.p, d_u32.p, d_u232.p);
        total += it;
        LaunchAdd(d_dx.p, d_x, n, stream);
        LaunchSpmv(ne, d_Arp.p, d_Acol.p, (const double*)d_Aval.p, (const double*)d_x, d_r.p, stream);
        LaunchResidual(d_b, d_r.p, n, stream);
        LaunchNorm((const double*)d_r.p, n, d_part.p, gm_blocks, scal, stream);
        Check(hipMemcpyAsync(h_pin, scal, sizeof(double), hipMemcpyDeviceToHost, stream), "d2h");
        Check(hipStreamSynchronize(stream), "refinement");
        double prev = normr;
        normr = h_pin[0];
        *iters = total;
        if (normr <= tol * n2b)
            return 0;
        if (f == 3 || normr >= 0.9 * prev)
            return 3;
    }
    return 1;
}

void Case::Solve(int nstep, const double* known_lu, long restart, double tol, double* lu, double* lt, int* iters,
                 int* flags, int* solved, bool deviceIO)
{
    if (deviceIO && keep)
        throw std::runtime_error("device-resident inputs are not taken by the adjoint's forward solve");
    if (!step0Done)
        AssembleStep0(10);
    if (!historyDone || maxn < 1)
        throw std::runtime_error("history not assembled");
    PrepareGmres(restart);
    // steps the loop does not reach keep zero counts and flags
    std::fill_n(iters, nstep, 0);
    std::fill_n(flags, nstep, 0);
    if (!noAllocMode)
        ReleaseGraphs();  // captured per solve: the operands may have moved since the last one
    gm_lastIters = 1;
    if (historyMixed)
        PrepareMixed(restart);
    if (historyMixed && !convTiled)
        throw std::runtime_error("the mixed-precision mode needs the time-blocked convolution");
    if (convTile != 8 && convTile != 16)
        throw std::runtime_error("convolution tile must be 8 or 16");
    if (convTiled && nearTile != convTile)
        BuildNearLists(convTile);
    if (historyStreamed && (!convTiled || convTile != 16 || !convFiltered || !convPacked || keep))
        throw std::runtime_error("streamed history bands need the packed, filtered 16-step convolution "
                                 "(conv_tiled 1, conv_tile 16, conv_filtered 1, conv_packed 1) and no adjoint");
    const long n3 = 3 * ne;
    const int nstore = nstep + 1;
    if (keep) {
        keep->X.alloc((size_t)nstore * (size_t)n3);
        LaunchZero(keep->X.p, (long)keep->X.n, stream);
    }
    DevBuf<double> d_bd, d_lu, d_lt, d_gU, d_gT, d_known, d_rhs, d_x, d_hfar;
    const size_t outN = (size_t)nstore * (size_t)n3;
    const double* bd = known_lu;
    double* olu = lu;
    double* olt = lt;
    if (!deviceIO) {
        d_bd.upload(known_lu, outN, stream);
        d_lu.alloc(outN);
        d_lt.alloc(outN);
        bd = d_bd.p;
        olu = d_lu.p;
        olt = d_lt.p;
    }
    d_gU.alloc((size_t)dof.historyUBlockCount * 3 * (size_t)nstore);
    d_gT.alloc((size_t)dof.historyTBlockCount * 3 * (size_t)nstore);
    d_known.alloc(3 * (size_t)std::max<long>(1, dof.knownBlockCount));
    d_rhs.alloc((size_t)n3);
    d_x.alloc((size_t)n3);
    d_hfar.alloc((size_t)kFarPartsMax * (size_t)kConvTileMax * (size_t)n3);  // partial sums per row (packed far part)
    DevBuf<double> d_F;
    const bool farF = convTiled && convFiltered && convTile == 16;
    if (farParts < 1 || farParts > kFarPartsMax)
        throw std::runtime_error("far-part wavefronts per row must be 1 .. " + std::to_string(kFarPartsMax));
    const int hfarParts = farF && convPacked ? farParts : 1;
    if (farF) {
        d_F.alloc((size_t)ne * (size_t)(nstore + 1) * 6);
        LaunchZero(d_F.p, (long)d_F.n, stream);
    }
    LaunchZero(d_gU.p, (long)d_gU.n, stream);
    LaunchZero(d_gT.p, (long)d_gT.n, stream);
    LaunchZero(d_known.p, (long)d_known.n, stream);
    LaunchZero(olu, (long)outN, stream);
    LaunchZero(olt, (long)outN, stream);
    // per-step events (convolution start, GMRES start, GMRES end) and per-tile
    // far-kernel events, read once after the loop
    EventPool ev;
    std::vector<std::array<hipEvent_t, 3>> stepEv;
    std::vector<std::array<hipEvent_t, 2>> farEv;
    stepEv.reserve((size_t)nstep);
    *solved = 0;
    int K0 = 2;
    for (int K = 1; K <= nstep; ++K) {
        LaunchKnown(d_refs.p, bd + (size_t)K * (size_t)n3, (int)ne, d_known.p, dof.knownBlockCount, stream);
        LaunchSpmv(ne, d_Krp.p, d_Kcol.p, d_Kval.p, d_known.p, d_rhs.p, stream);
        // no events without allocation: the no-allocation variant keeps no timings
        std::array<hipEvent_t, 3> se{};
        if (!noAllocMode) {
            se = {ev.Get(), ev.Get(), ev.Get()};
            Check(hipEventRecord(se[0], stream), "event");
        }
        if (K > 1) {
            if (convTiled) {
                if (K == 2 || K - K0 >= convTile) {
                    K0 = K;
                    int Tn = std::min(convTile, nstep - K0 + 1);
                    std::array<hipEvent_t, 2> fe{};
                    if (!noAllocMode) {
                        fe = {ev.Get(), ev.Get()};
                        Check(hipEventRecord(fe[0], stream), "event");
                    }
                    if (farF) {
                        LaunchFilterTile(d_refs.p, d_gU.p, d_gT.p, nstore, (int)ne, K0, std::max(0, K0 - convTile - 1),
                                         K0, d_F.p, stream);
                        if (historyStreamed)
                            for (size_t i = 0; i < slabE0.size(); ++i) {
                                StreamSlab(i, std::min(K0 + Tn - 1, maxn));  // the lags this tile reads
                                LaunchConvFarF(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const double*)d_smts.p,
                                               (const double*)d_smgs.p, nullptr, nullptr, d_F.p, nstore + 1, K0, Tn,
                                               maxn, (int)ne, d_hfar.p, farParts, stream, (int)slabRow0[i],
                                               (int)(slabRow1[i] - slabRow0[i]), slabE0[i]);
                            }
                        else if (historyMixed)
                            LaunchConvFarF(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const float*)d_mts32.p,
                                           (const float*)d_mgs32.p, (const unsigned short*)d_mtsLo.p,
                                           (const unsigned short*)d_mgsLo.p, d_F.p, nstore + 1, K0, Tn, maxn,
                                           (int)ne, d_hfar.p, convPacked ? farParts : 0, stream);
                        else
                            LaunchConvFarF(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const double*)d_mts.p,
                                           (const double*)d_mgs.p, nullptr, nullptr, d_F.p, nstore + 1, K0, Tn, maxn,
                                           (int)ne, d_hfar.p, convPacked ? farParts : 0, stream);
                    } else if (historyMixed)
                        LaunchConvFar(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const float*)d_mts32.p,
                                      (const float*)d_mgs32.p, (const unsigned short*)d_mtsLo.p,
                                      (const unsigned short*)d_mgsLo.p, d_gU.p, d_gT.p, nstore, K0, Tn, convTile,
                                      maxn, (int)ne, d_hfar.p, stream);
                    else
                        LaunchConvFar(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const double*)d_mts.p,
                                      (const double*)d_mgs.p, nullptr, nullptr, d_gU.p, d_gT.p, nstore, K0, Tn,
                                      convTile, maxn, (int)ne, d_hfar.p, stream);
                    if (!noAllocMode) {
                        Check(hipEventRecord(fe[1], stream), "event");
                        farEv.push_back(fe);
                    }
                }
                if (historyMixed)
                    LaunchConvNear(d_refs.p, d_nearPtr.p, d_nearPair.p, d_nearField.p, d_off.p, d_kb.p, d_len.p,
                                   (const float*)d_mts32.p, (const float*)d_mgs32.p, (const unsigned short*)d_mtsLo.p,
                                   (const unsigned short*)d_mgsLo.p, d_gU.p, d_gT.p, nstore, K, K0, maxn, (int)ne,
                                   d_hfar.p, hfarParts, d_rhs.p, nearPerEntry, stream);
                else if (historyStreamed)  // the resident near blocks, pair p's lags at nearOff[p]
                    LaunchConvNear(d_refs.p, d_nearPtr.p, d_nearPair.p, d_nearField.p, d_nearOff.p, d_kb.p, d_len.p,
                                   (const double*)d_nmts.p, (const double*)d_nmgs.p, nullptr, nullptr, d_gU.p, d_gT.p,
                                   nstore, K, K0, maxn, (int)ne, d_hfar.p, hfarParts, d_rhs.p, nearPerEntry, stream);
                else
                    LaunchConvNear(d_refs.p, d_nearPtr.p, d_nearPair.p, d_nearField.p, d_off.p, d_kb.p, d_len.p,
                                   (const double*)d_mts.p, (const double*)d_mgs.p, nullptr, nullptr, d_gU.p, d_gT.p,
                                   nstore, K, K0, maxn, (int)ne, d_hfar.p, hfarParts, d_rhs.p, nearPerEntry, stream);
            } else {
                LaunchConvolution(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, d_mts.p, d_mgs.p, d_gU.p, d_gT.p, nstore,
                                  K, maxn, (int)ne, d_rhs.p, stream);
            }
        }
        if (!noAllocMode)
            Check(hipEventRecord(se[1], stream), "event");
        int it = 0;
        int flag = historyMixed ? GmresMixed(d_rhs.p, d_x.p, restart, tol, &it)
                   : gmresDevice && restart <= kGmDeviceMaxRestart ? GmresDevice(d_rhs.p, d_x.p, restart, tol, &it)
                                                                   : Gmres(d_rhs.p, d_x.p, restart, tol, &it);
        if (!noAllocMode) {
            Check(hipEventRecord(se[2], stream), "event");
            stepEv.push_back(se);
        }
        iters[K - 1] = it;
        flags[K - 1] = flag;
        *solved = K;
        if (flag != 0)
            break;
        double* slu = olu + (size_t)K * (size_t)n3;
        double* slt = olt + (size_t)K * (size_t)n3;
        LaunchScatter(d_refs.p, d_geo.p, bd + (size_t)K * (size_t)n3, d_x.p, (int)ne, slu, slt, stream);
        LaunchState(d_refs.p, slu, slt, (int)ne, K, nstore, d_gU.p, d_gT.p, stream);
        if (keep)
            Check(hipMemcpyAsync(keep->X.p + (size_t)K * (size_t)n3, d_x.p, sizeof(double) * (size_t)n3,
                                 hipMemcpyDeviceToDevice, stream),
                  "keep x");
    }
    Check(hipStreamSynchronize(stream), "solve");
    if (keep) {  // hand the forward history to the adjoint
        std::swap(keep->gU.p, d_gU.p);
        std::swap(keep->gU.n, d_gU.n);
        std::swap(keep->gT.p, d_gT.p);
        std::swap(keep->gT.n, d_gT.n);
        std::swap(keep->bd.p, d_bd.p);
        std::swap(keep->bd.n, d_bd.n);
    }
    Check(hipGetLastError(), "solve launch");
    if (deviceIO) {
        LaunchStep0Out(d_refs.p, d_geo.p, bd, (int)ne, nstore, 1, olu, olt, stream);
        Check(hipStreamSynchronize(stream), "solve");
    } else {
        Check(CopyOnStream(lu, d_lu.p, sizeof(double) * d_lu.n, hipMemcpyDeviceToHost, stream), "lu");
        Check(CopyOnStream(lt, d_lt.p, sizeof(double) * d_lt.n, hipMemcpyDeviceToHost, stream), "lt");
    }
    // step 0: lt = 0, Convert (lu <-> lt for BCID 456), then LocToAbs
    // (DBEM1/multidomain.cpp:3819-3822)
    for (long i = 0; i < ne && !deviceIO; ++i) {
        double u[3], t[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = known_lu[3 * i + k];
            t[k] = 0.0;
        }
        if (bcid[(size_t)i] == 456)
            for (int k = 0; k < 3; ++k)
                std::swap(u[k], t[k]);
        const double* F = geo[(size_t)i].trans;
        lu[3 * i + 0] = F[0] * u[0] + F[3] * u[1] + F[6] * u[2];
        lu[3 * i + 1] = F[1] * u[0] + F[4] * u[1] + F[7] * u[2];
        lu[3 * i + 2] = F[2] * u[0] + F[5] * u[1] + F[8] * u[2];
        lt[3 * i + 0] = F[0] * t[0] + F[3] * t[1] + F[6] * t[2];
        lt[3 * i + 1] = F[1] * t[0] + F[4] * t[1] + F[7] * t[2];
        lt[3 * i + 2] = F[2] * t[0] + F[5] * t[1] + F[8] * t[2];
    }
    double convTime = 0.0, gmresTime = 0.0, farTime = 0.0;
    for (const auto& e : stepEv) {
        convTime += ev.Elapsed(e[0], e[1]);
        gmresTime += ev.Elapsed(e[1], e[2]);
    }
    for (const auto& e : farEv)
        farTime += ev.Elapsed(e[0], e[1]);
    timing[4] = convTime;
    timing[5] = gmresTime;
    timing[6] = farTime;
    timing[7] = (double)farEv.size();
    timing[21] = historyStreamed ? farTime : 0.0;  // far part with the slabs' regeneration
}

// ---------------------------------------------------------------------------
// Grouped models (CreateGroup): one time loop for all of them.
// ---------------------------------------------------------------------------
void Case::GmresDeviceGroups(const double* d_b, double* d_x, long restart, double tol, const int* active, int* its,
                             int* fls)
{
    const int G = GroupCount();
    GmresGroupArgs a{};
    a.ne = ne;
    a.restart = restart;
    a.rp = d_Arp.p;
    a.col = d_Acol.p;
    a.val = d_Aval.p;
    a.posLeaf = d_posLeaf.p;
    a.leafBegin = d_leafBegin.p;
    a.leafSize = d_leafSize.p;
    a.matOff = d_matOff.p;
    a.repid = d_repid.p;
    a.outBlock = d_outBlock.p;
    a.inv = d_inv.p;
    a.V = gw.V.p;
    a.W = gw.W.p;
    a.u = gw.u.p;
    a.u2 = gw.u2.p;
    a.vh = gw.vh.p;
    a.x = d_x;
    a.h = gw.h.p;
    a.h2 = gw.h2.p;
    a.QT = gw.QT.p;
    a.R = gw.R.p;
    a.f = gw.f.p;
    a.phibar = gw.phibar.p;
    a.part = gw.part.p;
    a.hs = restart + 2;
    a.qs = (restart + 1) * (restart + 1);
    a.rs = restart * restart;
    a.ps = (long)kGmBlocksMax * (restart + 1);
    a.stag = gw.stag.p;
    a.active = gw.active.p;
    a.st = reinterpret_cast<GmState*>(gw.st.p);
    Check(hipMemcpyAsync(gw.active.p, active, sizeof(int) * (size_t)G, hipMemcpyHostToDevice, stream), "active");
    // the operands stay fixed for a solve, so one capture serves every step
    const bool useGraph = gmresDevice == 1;
    if (useGraph && !(gw.gBegin && gw.graphB == d_b && gw.graphTol == tol)) {
        if (gw.gBegin)
            Check(hipGraphExecDestroy(gw.gBegin), "graph");
        if (gw.gIter)
            Check(hipGraphExecDestroy(gw.gIter), "graph");
        gw.gBegin = gw.gIter = nullptr;
        hipGraph_t g = nullptr;
        Check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
        LaunchGmresBeginGroups(a, gw.t, gw.tab.p, d_b, tol, stream);
        Check(hipStreamEndCapture(stream, &g), "capture");
        Check(hipGraphInstantiate(&gw.gBegin, g, nullptr, nullptr, 0), "instantiate");
        Check(hipGraphDestroy(g), "graph");
        Check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
        LaunchGmresIterationGroups(a, gw.t, gw.tab.p, d_b, stream);
        Check(hipStreamEndCapture(stream, &g), "capture");
        Check(hipGraphInstantiate(&gw.gIter, g, nullptr, nullptr, 0), "instantiate");
        Check(hipGraphDestroy(g), "graph");
        gw.graphB = d_b;
        gw.graphTol = tol;
    }
    if (useGraph)
        Check(hipGraphLaunch(gw.gBegin, stream), "graph launch");
    else
        LaunchGmresBeginGroups(a, gw.t, gw.tab.p, d_b, tol, stream);
    long queued = 0;
    long ahead = std::max(1, gw.lastIters);
    bool all = false;
    for (;;) {
        long k = std::min(ahead, restart - queued);
        for (long i = 0; i < k; ++i) {
            if (useGraph)
                Check(hipGraphLaunch(gw.gIter, stream), "graph launch");
            else
                LaunchGmresIterationGroups(a, gw.t, gw.tab.p, d_b, stream);
        }
        queued += k;
        Check(hipMemcpyAsync(gw.hst, gw.st.p, sizeof(GmState) * (size_t)G, hipMemcpyDeviceToHost, stream), "d2h");
        Check(hipStreamSynchronize(stream), "gmres");
        all = true;
        for (int g = 0; g < G; ++g)
            all = all && gw.hst[g].done;
        if (all || queued >= restart)
            break;
        ahead = 2;
    }
    if (!all)
        throw std::runtime_error("grouped device GMRES ended without a stopping state");
    int most = 1;
    for (int g = 0; g < G; ++g) {
        its[g] = gw.hst[g].j;
        fls[g] = gw.hst[g].flag;
        if (active[g])
            most = std::max(most, gw.hst[g].j);
    }
    gw.lastIters = most;
}

void Case::SolveGroups(int nstep, const double* known_lu, long restart, double tol, double* lu, double* lt,
                       int* iters, int* flags, int* solved)
{
    const int G = GroupCount();
    if (!step0Done)
        AssembleStep0(10);
    if (!historyDone || maxn < 1)
        throw std::runtime_error("history not assembled");
    if (historyMixed || historyStreamed)
        throw std::runtime_error("grouped models run over FP64 bands held on the device");
    if (!convTiled || convTile != 16 || !convFiltered || !gmresDevice || restart > kGmDeviceMaxRestart)
        throw std::runtime_error("grouped models run the filtered 16-step convolution and the device-driven GMRES");
    if (nearTile != convTile)
        BuildNearLists(convTile);
    const long n3 = 3 * ne;
    const int nstore = nstep + 1;
    // GMRES workspace: vectors for all groups, scalars per group
    if (gw.restart != restart) {
        gw.release();
        gw.restart = restart;
        gw.V.alloc((size_t)n3 * (size_t)(restart + 1));
        gw.W.alloc((size_t)n3 * (size_t)restart);
        gw.u.alloc((size_t)n3);
        gw.u2.alloc((size_t)n3);
        gw.vh.alloc((size_t)n3);
        gw.h.alloc((size_t)G * (size_t)(restart + 2));
        gw.h2.alloc((size_t)G * (size_t)(restart + 2));
        gw.f.alloc((size_t)G * (size_t)(restart + 2));
        gw.QT.alloc((size_t)G * (size_t)(restart + 1) * (size_t)(restart + 1));
        gw.R.alloc((size_t)G * (size_t)restart * (size_t)restart);
        gw.part.alloc((size_t)G * (size_t)kGmBlocksMax * (size_t)(restart + 1));
        gw.phibar.alloc((size_t)G);
        gw.st.alloc((size_t)G * ((sizeof(GmState) + sizeof(double) - 1) / sizeof(double)));
        gw.stag.alloc((size_t)G);
        gw.active.alloc((size_t)G);
        Check(hipHostMalloc((void**)&gw.hst, sizeof(GmState) * (size_t)G, hipHostMallocDefault), "pinned");
        // block table: each group's reduction partition is the one of its own job
        GmresGroupTable& t = gw.t;
        t = GmresGroupTable{};
        t.G = G;
        if (gmBlocks < 1 || gmBlocks > kGmBlocksMax)
            throw std::runtime_error("GMRES reduction blocks must be 1 .. " + std::to_string(kGmBlocksMax));
        for (int g = 0; g <= G; ++g)
            t.ele[g] = groupEle[(size_t)g];
        for (int g = 0; g < G; ++g) {
            const long neg = t.ele[g + 1] - t.ele[g];
            t.nb[g] = (int)std::min<long>(gmBlocks, std::max<long>(1, (3 * neg + 255) / 256));
            t.nbPre[g + 1] = t.nbPre[g] + t.nb[g];
            t.rbPre[g + 1] = t.rbPre[g] + (int)((3 * neg + 255) / 256);
        }
        gw.tab.upload(&t, 1, stream);
    }
    gw.lastIters = 1;
    DevBuf<double> d_bd, d_lu, d_lt, d_gU, d_gT, d_known, d_rhs, d_x, d_hfar, d_F;
    const size_t outN = (size_t)nstore * (size_t)n3;
    d_bd.upload(known_lu, outN, stream);
    d_lu.alloc(outN);
    d_lt.alloc(outN);
    d_gU.alloc((size_t)dof.historyUBlockCount * 3 * (size_t)nstore);
    d_gT.alloc((size_t)dof.historyTBlockCount * 3 * (size_t)nstore);
    d_known.alloc(3 * (size_t)std::max<long>(1, dof.knownBlockCount));
    d_rhs.alloc((size_t)n3);
    d_x.alloc((size_t)n3);
    if (farParts < 1 || farParts > kFarPartsMax)
        throw std::runtime_error("far-part wavefronts per row must be 1 .. " + std::to_string(kFarPartsMax));
    const int hfarParts = convPacked ? farParts : 1;
    d_hfar.alloc((size_t)kFarPartsMax * (size_t)kConvTileMax * (size_t)n3);
    d_F.alloc((size_t)ne * (size_t)(nstore + 1) * 6);
    LaunchZero(d_F.p, (long)d_F.n, stream);
    LaunchZero(d_gU.p, (long)d_gU.n, stream);
    LaunchZero(d_gT.p, (long)d_gT.n, stream);
    LaunchZero(d_known.p, (long)d_known.n, stream);
    LaunchZero(d_lu.p, (long)outN, stream);
    LaunchZero(d_lt.p, (long)outN, stream);
    EventPool ev;
    std::vector<std::array<hipEvent_t, 3>> stepEv;
    std::vector<std::array<hipEvent_t, 2>> farEv;
    stepEv.reserve((size_t)nstep);
    std::vector<int> active((size_t)G, 1), it((size_t)G), fl((size_t)G);
    for (int g = 0; g < G; ++g)
        solved[g] = 0;
    int K0 = 2;
    for (int K = 1; K <= nstep; ++K) {
        LaunchKnown(d_refs.p, d_bd.p + (size_t)K * (size_t)n3, (int)ne, d_known.p, dof.knownBlockCount, stream);
        LaunchSpmv(ne, d_Krp.p, d_Kcol.p, d_Kval.p, d_known.p, d_rhs.p, stream);
        std::array<hipEvent_t, 3> se = {ev.Get(), ev.Get(), ev.Get()};
        Check(hipEventRecord(se[0], stream), "event");
        if (K > 1) {
            if (K == 2 || K - K0 >= convTile) {
                K0 = K;
                const int Tn = std::min(convTile, nstep - K0 + 1);
                std::array<hipEvent_t, 2> fe = {ev.Get(), ev.Get()};
                Check(hipEventRecord(fe[0], stream), "event");
                LaunchFilterTile(d_refs.p, d_gU.p, d_gT.p, nstore, (int)ne, K0, std::max(0, K0 - convTile - 1), K0,
                                 d_F.p, stream);
                LaunchConvFarF(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const double*)d_mts.p,
                               (const double*)d_mgs.p, nullptr, nullptr, d_F.p, nstore + 1, K0, Tn, maxn, (int)ne,
                               d_hfar.p, convPacked ? farParts : 0, stream);
                Check(hipEventRecord(fe[1], stream), "event");
                farEv.push_back(fe);
            }
            LaunchConvNear(d_refs.p, d_nearPtr.p, d_nearPair.p, d_nearField.p, d_off.p, d_kb.p, d_len.p,
                           (const double*)d_mts.p, (const double*)d_mgs.p, nullptr, nullptr, d_gU.p, d_gT.p, nstore,
                           K, K0, maxn, (int)ne, d_hfar.p, hfarParts, d_rhs.p, nearPerEntry, stream);
        }
        Check(hipEventRecord(se[1], stream), "event");
        GmresDeviceGroups(d_rhs.p, d_x.p, restart, tol, active.data(), it.data(), fl.data());
        Check(hipEventRecord(se[2], stream), "event");
        stepEv.push_back(se);
        bool any = false;
        for (int g = 0; g < G; ++g) {
            if (!active[(size_t)g])
                continue;
            iters[(size_t)g * nstep + (K - 1)] = it[(size_t)g];
            flags[(size_t)g * nstep + (K - 1)] = fl[(size_t)g];
            solved[g] = K;
            if (fl[(size_t)g] != 0)
                active[(size_t)g] = 0;
            any = any || active[(size_t)g];
        }
        if (!any)
            break;
        double* slu = d_lu.p + (size_t)K * (size_t)n3;
        double* slt = d_lt.p + (size_t)K * (size_t)n3;
        LaunchScatter(d_refs.p, d_geo.p, d_bd.p + (size_t)K * (size_t)n3, d_x.p, (int)ne, slu, slt, stream);
        LaunchState(d_refs.p, slu, slt, (int)ne, K, nstore, d_gU.p, d_gT.p, stream);
    }
    Check(hipStreamSynchronize(stream), "grouped solve");
    Check(hipGetLastError(), "grouped solve launch");
    Check(CopyOnStream(lu, d_lu.p, sizeof(double) * outN, hipMemcpyDeviceToHost, stream), "lu");
    Check(CopyOnStream(lt, d_lt.p, sizeof(double) * outN, hipMemcpyDeviceToHost, stream), "lt");
    // step 0: lt = 0, Convert (lu <-> lt for BCID 456), then LocToAbs
    // (DBEM1/multidomain.cpp:3819-3822)
    for (long i = 0; i < ne; ++i) {
        double u[3], t[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = known_lu[3 * i + k];
            t[k] = 0.0;
        }
        if (bcid[(size_t)i] == 456)
            for (int k = 0; k < 3; ++k)
                std::swap(u[k], t[k]);
        const double* F = geo[(size_t)i].trans;
        lu[3 * i + 0] = F[0] * u[0] + F[3] * u[1] + F[6] * u[2];
        lu[3 * i + 1] = F[1] * u[0] + F[4] * u[1] + F[7] * u[2];
        lu[3 * i + 2] = F[2] * u[0] + F[5] * u[1] + F[8] * u[2];
        lt[3 * i + 0] = F[0] * t[0] + F[3] * t[1] + F[6] * t[2];
        lt[3 * i + 1] = F[1] * t[0] + F[4] * t[1] + F[7] * t[2];
        lt[3 * i + 2] = F[2] * t[0] + F[5] * t[1] + F[8] * t[2];
    }
    double convTime = 0.0, gmresTime = 0.0, farTime = 0.0;
    for (const auto& e : stepEv) {
        convTime += ev.Elapsed(e[0], e[1]);
        gmresTime += ev.Elapsed(e[1], e[2]);
    }
    for (const auto& e : farEv)
        farTime += ev.Elapsed(e[0], e[1]);
    timing[4] = convTime;
    timing[5] = gmresTime;
    timing[6] = farTime;
    timing[7] = (double)farEv.size();
    gw.release();
}

}  // namespace tdbem
