// Batched load cases on one context: B right-hand-side histories share the
// history bands, the step-0 matrices and the preconditioner; every load case
// follows the reference's time loop (DBEM1/multidomain.cpp:3793-3926) with
// its own GMRES (DBEM1/pre_gmres.cpp:1369-1605) run in lockstep with the
// others.
#include "batched.hpp"
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

namespace tdbem {

namespace {

void CheckB(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Host buffers of a batched solve: the load histories go to the device from
// a second thread on a stream of its own, chunk by chunk ahead of the time
// loop (chunk c: steps c * kSteps .. c * kSteps + kSteps - 1); Need(K, s)
// makes the solving stream s wait for the chunk of step K.
// One chunk of steps of every load case between host and device arrays laid
// out [case][step][n]: a 2D copy (one row per case), or one copy per case
// when the row pitch is past what a 2D copy takes.
hipError_t CopyStepChunk(double* dst, const double* src, size_t pitch, size_t width, int B, hipMemcpyKind kind,
                         hipStream_t s)
{
    if (pitch <= (size_t)0x7fffffff)
        return hipMemcpy2DAsync(dst, pitch, src, pitch, width, (size_t)B, kind, s);
    for (int b = 0; b < B; ++b) {
        const size_t at = (size_t)b * (pitch / sizeof(double));
        hipError_t e = hipMemcpyAsync(dst + at, src + at, width, kind, s);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

class InputFeed {
public:
    static constexpr int kSteps = 16;
    InputFeed(int nstore, long n, int B, const double* h, double* d)
        : nch_((nstore + kSteps - 1) / kSteps), nstore_(nstore), n_(n), B_(B), h_(h), d_(d),
          ev_((size_t)nch_, nullptr)
    {
        for (hipEvent_t& e : ev_)
            CheckB(hipEventCreateWithFlags(&e, hipEventDisableTiming), "feed event");
        CheckB(hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking), "feed stream");
        th_ = std::thread([this]() { Run(); });
    }
    ~InputFeed()
    {
        if (th_.joinable()) {
            {
                std::lock_guard<std::mutex> lk(m_);
                stop_ = true;
            }
            th_.join();
        }
        if (cs_) {
            (void)hipStreamSynchronize(cs_);
            (void)hipStreamDestroy(cs_);
        }
        for (hipEvent_t e : ev_)
            if (e)
                (void)hipEventDestroy(e);
    }
    void Need(int K, hipStream_t s)
    {
        const int c = K / kSteps;
        if (c < waited_)
            return;
        {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [&]() { return queued_ > c || !err_.empty(); });
            if (!err_.empty())
                throw std::runtime_error("batched solve: copying the load histories in: " + err_);
        }
        for (int q = waited_; q <= c; ++q)
            CheckB(hipStreamWaitEvent(s, ev_[(size_t)q], 0), "feed wait");
        waited_ = c + 1;
    }

private:
    void Run()
    {
        for (int c = 0; c < nch_; ++c) {
            {
                std::lock_guard<std::mutex> lk(m_);
                if (stop_)
                    return;
            }
            const int K0 = c * kSteps, K1 = std::min(nstore_ - 1, K0 + kSteps - 1);
            const size_t pitch = (size_t)nstore_ * (size_t)n_ * sizeof(double);
            const size_t width = (size_t)(K1 - K0 + 1) * (size_t)n_ * sizeof(double);
            const size_t off = (size_t)K0 * (size_t)n_;
            hipError_t e = CopyStepChunk(d_ + off, h_ + off, pitch, width, B_, hipMemcpyHostToDevice, cs_);
            if (e == hipSuccess)
                e = hipEventRecord(ev_[(size_t)c], cs_);
            {
                std::lock_guard<std::mutex> lk(m_);
                if (e != hipSuccess)
                    err_ = hipGetErrorString(e);
                else
                    queued_ = c + 1;
            }
            cv_.notify_all();
            if (e != hipSuccess)
                return;
        }
    }
    int nch_, nstore_;
    long n_;
    int B_;
    const double* h_;
    double* d_;
    std::vector<hipEvent_t> ev_;
    hipStream_t cs_ = nullptr;
    std::thread th_;
    std::mutex m_;
    std::condition_variable cv_;
    int queued_ = 0, waited_ = 0;
    bool stop_ = false;
    std::string err_;
};

// Host buffers of a batched solve: the solved steps of lu / lt are copied to
// the caller's arrays by a second thread on a stream of its own, chunk by
// chunk, while the time loop runs on.  Chunk c (steps 1 + c * kSteps ..) is
// copied after the event Ready(c) records on the solving stream; Stop ends
// the thread (after the chunks marked ready, or at once on an error path).
class OutputDrain {
public:
    static constexpr int kSteps = 16;
    OutputDrain(int nstep, int nstore, long n, int B, const double* dlu, const double* dlt, double* lu, double* lt)
        : nch_((nstep + kSteps - 1) / kSteps), nstep_(nstep), nstore_(nstore), n_(n), B_(B), dlu_(dlu), dlt_(dlt),
          lu_(lu), lt_(lt), ev_((size_t)std::max(0, nch_), nullptr)
    {
        for (hipEvent_t& e : ev_)
            CheckB(hipEventCreateWithFlags(&e, hipEventDisableTiming), "drain event");
        CheckB(hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking), "drain stream");
        th_ = std::thread([this]() { Run(); });
    }
    ~OutputDrain()
    {
        Stop(false);
        for (hipEvent_t e : ev_)
            if (e)
                (void)hipEventDestroy(e);
        if (cs_)
            (void)hipStreamDestroy(cs_);
    }
    // chunks below `upTo` are complete once the work queued on s so far is
    int Chunks() const { return nch_; }
    void Ready(int upTo, hipStream_t s)
    {
        upTo = std::min(upTo, nch_);
        if (upTo <= marked_)
            return;
        for (int c = marked_; c < upTo; ++c)
            CheckB(hipEventRecord(ev_[(size_t)c], s), "drain event");
        marked_ = upTo;
        {
            std::lock_guard<std::mutex> lk(m_);
            ready_ = upTo;
        }
        cv_.notify_one();
    }
    // waits for the copies of the chunks marked ready (all of them with
    // finish), then ends the thread; throws what a copy failed with
    void Stop(bool finish)
    {
        if (th_.joinable()) {
            {
                std::lock_guard<std::mutex> lk(m_);
                stop_ = true;
                drainAll_ = finish;
            }
            cv_.notify_one();
            th_.join();
        }
        if (finish && !err_.empty())
            throw std::runtime_error("batched solve: copying the solution out: " + err_);
    }

private:
    void Run()
    {
        for (int c = 0; c < nch_; ++c) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&]() { return ready_ > c || stop_; });
                if (ready_ <= c || (stop_ && !drainAll_) || !err_.empty())
                    return;
            }
            const int K0 = 1 + c * kSteps, K1 = std::min(nstep_, K0 + kSteps - 1);
            const size_t pitch = (size_t)nstore_ * (size_t)n_ * sizeof(double);
            const size_t width = (size_t)(K1 - K0 + 1) * (size_t)n_ * sizeof(double);
            const size_t off = (size_t)K0 * (size_t)n_;
            hipError_t e = hipStreamWaitEvent(cs_, ev_[(size_t)c], 0);
            if (e == hipSuccess)
                e = CopyStepChunk(lu_ + off, dlu_ + off, pitch, width, B_, hipMemcpyDeviceToHost, cs_);
            if (e == hipSuccess)
                e = CopyStepChunk(lt_ + off, dlt_ + off, pitch, width, B_, hipMemcpyDeviceToHost, cs_);
            if (e == hipSuccess)
                e = hipStreamSynchronize(cs_);
            if (e != hipSuccess) {
                std::lock_guard<std::mutex> lk(m_);
                err_ = hipGetErrorString(e);
                return;
            }
        }
    }
    int nch_, nstep_, nstore_;
    long n_;
    int B_;
    const double* dlu_;
    const double* dlt_;
    double* lu_;
    double* lt_;
    std::vector<hipEvent_t> ev_;
    hipStream_t cs_ = nullptr;
    std::thread th_;
    std::mutex m_;
    std::condition_variable cv_;
    int ready_ = 0, marked_ = 0;
    bool stop_ = false, drainAll_ = false;
    std::string err_;
};

double NowB()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// Lockstep GMRES for the B systems A x_b = rhs_b.  Each case stops at its own
// convergence (x frozen, iteration count and flag recorded); the loop ends
// when no case is active.
void Case::GmresBatched(int B, const double* rhs, double* x, long restart, double tol, int* its, int* flg)
{
    const long n = 3 * ne;
    // reduction blocks of the dot products: 16 wavefronts of 64 load cases
    // per block, at least 4 rows per wavefront (one workgroup per CU on the
    // 8192-element box)
    const int nblk = (int)std::min<long>(kBatchRedBlocks, std::max<long>(1, (n + 63) / 64));
    // the same for the preconditioner's pass over elements (three rows each,
    // at least one element per wavefront; 256 blocks on the 8192-element box)
    const int nblkE = (int)std::min<long>(kBatchRedBlocks, std::max<long>(1, (ne + 15) / 16));
    BatchWork& w = bw;
    std::vector<double> bn((size_t)B), nr((size_t)B);
    std::vector<int> act((size_t)B), stag((size_t)B);
    for (int b = 0; b < B; ++b) {
        its[b] = 0;
        flg[b] = 0;
    }
    LaunchNormB(rhs, n, B, w.part.p, nblk, w.scal.p, stream);
    CheckB(hipMemcpyAsync(bn.data(), w.scal.p, sizeof(double) * (size_t)B, hipMemcpyDeviceToHost, stream), "d2h");
    LaunchZero(x, n * B, stream);
    // r = b - A x0 with x0 = 0
    LaunchSpmmB(ne, d_Arp.p, d_Acol.p, d_Aval.p, x, w.vh.p, B, stream);
    LaunchPassB(nullptr, n, B, w.vh.p, rhs, nullptr, 0, 0, true, w.part.p, nblk, nullptr, nullptr, w.scal.p + B,
                stream);
    CheckB(hipMemcpyAsync(nr.data(), w.scal.p + B, sizeof(double) * (size_t)B, hipMemcpyDeviceToHost, stream), "d2h");
    CheckB(hipStreamSynchronize(stream), "gmres");
    int nact = 0;
    for (int b = 0; b < B; ++b) {
        act[(size_t)b] = (bn[(size_t)b] != 0.0 && !(nr[(size_t)b] <= tol * bn[(size_t)b])) ? 1 : 0;
        nact += act[(size_t)b];
    }
    if (nact == 0)
        return;
    CheckB(hipMemcpyAsync(w.active.p, act.data(), sizeof(int) * (size_t)B, hipMemcpyHostToDevice, stream), "active");
    CheckB(hipMemsetAsync(w.QT.p, 0, sizeof(double) * w.QT.n, stream), "memset");
    CheckB(hipMemsetAsync(w.R.p, 0, sizeof(double) * w.R.n, stream), "memset");
    CheckB(hipMemsetAsync(w.stag.p, 0, sizeof(int) * (size_t)B, stream), "memset");
    LaunchPrecondB(ne, d_posLeaf.p, d_leafBegin.p, d_leafSize.p, d_matOff.p, d_inv.p, d_repid.p, d_outBlock.p,
                   w.vh.p, w.u.p, B, stream);
    double* h = w.h.p;  // h[k * B + b], k = 0 .. restart + 1
    LaunchNormB(w.u.p, n, B, w.part.p, nblk, h + B, stream);
    LaunchScaleB(w.u.p, h + B, w.V.p, n, B, stream);
    {
        std::vector<double> one((size_t)B, 1.0);
        CheckB(hipMemcpyAsync(w.QT.p, one.data(), sizeof(double) * (size_t)B, hipMemcpyHostToDevice, stream), "qt");
        CheckB(hipMemcpyAsync(w.phibar.p, h + B, sizeof(double) * (size_t)B, hipMemcpyDeviceToDevice, stream),
               "phibar");
        CheckB(hipStreamSynchronize(stream), "gmres");
    }
    // the stopping test runs on the device after every iteration (the cases
    // still iterating in w.left, which every iteration kernel checks first);
    // the host reads the count after as many iterations as the last solve
    // took and then every other one, so it does not wait on every iteration
    CheckB(hipMemcpyAsync(w.its.p, its, sizeof(int) * (size_t)B, hipMemcpyHostToDevice, stream), "its");
    CheckB(hipMemcpyAsync(w.flg.p, flg, sizeof(int) * (size_t)B, hipMemcpyHostToDevice, stream), "flags");
    *w.hLeft = nact;
    CheckB(hipMemcpyAsync(w.left.p, w.hLeft, sizeof(int), hipMemcpyHostToDevice, stream), "left");
    const int* left = w.left.p;
    int ahead = std::max(1, w.lastIters), since = 0;
    int j = 1;
    for (; j <= restart; ++j) {
        const double* Vj = w.V.p + (size_t)(j - 1) * (size_t)n * (size_t)B;
        LaunchSpmmB(ne, d_Arp.p, d_Acol.p, d_Aval.p, Vj, w.u2.p, B, stream, left);
        // classical Gram-Schmidt twice: h = V'u, u -= V h; h2 = V'u, u -= V h2,
        // h += h2; h(j+1) = ||u||.  The first dot products share their pass
        // with the preconditioner, each update with the next dot products or
        // norm.
        LaunchPrecondDotB(ne, d_posLeaf.p, d_leafBegin.p, d_leafSize.p, d_matOff.p, d_inv.p, d_repid.p,
                          d_outBlock.p, w.u2.p, w.u.p, B, w.V.p, j, w.part.p, nblkE, h + B, stream, left);
        LaunchPassB(w.V.p, n, B, w.u.p, nullptr, h + B, j, j, false, w.part.p, nblk, w.h2.p, h + B, nullptr,
                    stream, left);
        LaunchPassB(w.V.p, n, B, w.u.p, nullptr, w.h2.p, j, 0, true, w.part.p, nblk, nullptr, nullptr,
                    h + (size_t)(j + 1) * B, stream, left);
        LaunchGivensB(h, w.QT.p, w.R.p, w.f.p, w.phibar.p, w.stag.p, w.active.p, j, (int)restart, B, stream, left);
        LaunchWUpdateB(w.V.p, w.W.p, n, w.R.p, j, (int)restart, w.f.p, x, w.active.p, B, stream, w.u.p,
                       h + (size_t)(j + 1) * B, w.V.p + (size_t)j * (size_t)n * (size_t)B, left);
        LaunchSpmmB(ne, d_Arp.p, d_Acol.p, d_Aval.p, x, w.vh.p, B, stream, left);
        LaunchPassB(nullptr, n, B, w.vh.p, rhs, nullptr, 0, 0, true, w.part.p, nblk, nullptr, nullptr, w.scal.p + B,
                    stream, left);
        LaunchCheckB(w.scal.p, tol, w.stag.p, w.active.p, w.its.p, w.flg.p, j, B, w.left.p, stream);
        if (++since >= ahead || j == restart) {
            CheckB(hipMemcpyAsync(w.hLeft, w.left.p, sizeof(int), hipMemcpyDeviceToHost, stream), "d2h");
            CheckB(hipStreamSynchronize(stream), "gmres iteration");
            if (*w.hLeft == 0)
                break;
            since = 0;
            ahead = 2;
        }
    }
    CheckB(hipMemcpyAsync(its, w.its.p, sizeof(int) * (size_t)B, hipMemcpyDeviceToHost, stream), "its");
    CheckB(hipMemcpyAsync(flg, w.flg.p, sizeof(int) * (size_t)B, hipMemcpyDeviceToHost, stream), "flags");
    CheckB(hipMemcpyAsync(act.data(), w.active.p, sizeof(int) * (size_t)B, hipMemcpyDeviceToHost, stream), "active");
    CheckB(hipStreamSynchronize(stream), "gmres");
    int most = 1;
    for (int b = 0; b < B; ++b)
        most = std::max(most, its[b]);
    w.lastIters = most;
    for (int b = 0; b < B; ++b)
        if (act[(size_t)b])
            flg[b] = 1;
}

void Case::SolveBatched(int nstep, int B, const double* known_lu, long restart, double tol, double* lu, double* lt,
                        int* iters, int* flags, int* solved, bool deviceIO)
{
    if (B < 1)
        throw std::runtime_error("batch size must be >= 1");
    if (!step0Done)
        AssembleStep0(10);
    RefuseStreamed("the batched solve");
    if (!historyDone || maxn < 1)
        throw std::runtime_error("history not assembled");
    // far part: the lag-by-lag kernels (8-step tiles) take every lag, or, in
    // the spectral form (panels of 16 or more load cases unless the option
    // says otherwise, and when its filter spectra fit the device), the raw
    // lags <= 16 while batched_fft.hip takes the rest over 16-step tiles
    if (nearTile != kBatchTile)
        BuildNearLists(kBatchTile);
    bool fft = batchedFft > 0 || (batchedFft < 0 && B >= 16);
    const long n = 3 * ne;
    const int nstore = nstep + 1;
    const long histU = dof.historyUBlockCount, histT = dof.historyTBlockCount;
    // spectral part: row groups (shared by the levels); per level its (field,
    // sub-filter) items, filter spectra and window-spectra ring
    DevBuf<int> d_grpDom, d_grpRowList, d_grpRows;
    DevBuf<double> d_Y, d_hfar16;
    struct FftLevel {
        int L = kFftTile, lagLo = kFftLagMin, lagHi = 2 * kFftTile, nwr = 1, qcap = 1, K0 = 0;
        long nitems = 0, hvals = 0, strideU = 0, strideT = 0;
        std::vector<long> ioff, hoff, qItemsUpTo, qRowsUpTo;
        DevBuf<long> itemOff, hOff, items;
        DevBuf<int> order;  // row groups by their summed item rows, most first (products launch order)
        DevBuf<double> H, WU, WT, hfar;
        double prodTime = 0.0, flops = 0.0, bytes = 0.0;
        int launches = 0;
        bool ran = false;
    } lv[kFftMaxLevels];
    int ngroups = 0, nl = 0;
    for (int i = 26; i < 43; ++i)
        timing[i] = 0.0;
    timing[16] = 0.0;
    if (fft) {
        // levels: sub-filter length 16 << l, each with lags of its own
        for (int l = 0; l < std::min(batchedFftLevels, kFftMaxLevels); ++l)
            if ((kFftTile << l) + 1 <= maxn)
                nl = l + 1;
        fft = nl > 0;
    }
    if (fft) {
        const double fs0 = NowB();
        // groups of kFftRows source rows of one domain in the Z-order (Morton)
        // of their element centres, so a group's rows lie close together and
        // share more of their fields' sub-filters (fewer, fuller items)
        std::vector<int> gd, gl, gr;
        for (int dd = 0; dd < grid.nd; ++dd) {
            const int eb = grid.eleBegin[dd], nd = grid.eleCount[dd];
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int r = eb; r < eb + nd; ++r)
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], geo[(size_t)r].center[a]);
                    hi[a] = std::max(hi[a], geo[(size_t)r].center[a]);
                }
            std::vector<std::pair<unsigned long long, int>> key((size_t)nd);
            for (int i = 0; i < nd; ++i) {
                unsigned long long code = 0;
                for (int a = 0; a < 3; ++a) {
                    const double span = hi[a] - lo[a];
                    const double x = span > 0.0 ? (geo[(size_t)(eb + i)].center[a] - lo[a]) / span : 0.0;
                    const unsigned v = (unsigned)std::min(1023.0, x * 1023.0);
                    for (int bit = 0; bit < 10; ++bit)
                        code |= (unsigned long long)((v >> bit) & 1u) << (3 * bit + a);
                }
                key[(size_t)i] = {code, eb + i};
            }
            std::sort(key.begin(), key.end());  // ties by element index: deterministic
            for (int r0 = 0; r0 < nd; r0 += kFftRows) {
                gd.push_back(dd);
                const int cnt = std::min(kFftRows, nd - r0);
                gr.push_back(cnt);
                for (int r = 0; r < kFftRows; ++r)
                    gl.push_back(r < cnt ? key[(size_t)(r0 + r)].second : key[(size_t)r0].second);
            }
        }
        ngroups = (int)gd.size();
        d_grpDom.upload(gd.data(), gd.size(), stream);
        d_grpRowList.upload(gl.data(), gl.size(), stream);
        d_grpRows.upload(gr.data(), gr.size(), stream);
        HostBands();
        double bytes = 0.0;
        for (int l = 0; l < nl; ++l) {
            FftLevel& v = lv[l];
            v.L = kFftTile << l;
            v.lagLo = v.L + 1;
            v.lagHi = l == nl - 1 ? maxn : 2 * v.L;
            // per group: items, rows; per sub-filter q <= qcap: items, rows
            v.qcap = (maxn - 1) / v.L + 1;
            const size_t ncnt = 2 * (size_t)ngroups + 2 * (size_t)(v.qcap + 1);
            DevBuf<long> d_cnt;
            d_cnt.alloc(ncnt);
            LaunchZero(reinterpret_cast<double*>(d_cnt.p), (long)ncnt, stream);
            LaunchFftItems(grid, d_refs.p, d_kb.p, d_len.p, maxn, v.L, v.lagLo, v.lagHi, ngroups, d_grpDom.p,
                           d_grpRowList.p, d_grpRows.p, nullptr, nullptr, d_cnt.p, d_cnt.p + ngroups,
                           d_cnt.p + 2 * ngroups, v.qcap, nullptr, stream);
            std::vector<long> cnt(ncnt);
            CheckB(CopyOnStream(cnt.data(), d_cnt.p, sizeof(long) * cnt.size(), hipMemcpyDeviceToHost, stream),
                   "items");
            v.ioff.assign((size_t)ngroups + 1, 0);
            v.hoff.assign((size_t)ngroups + 1, 0);
            for (int gi = 0; gi < ngroups; ++gi) {
                v.ioff[(size_t)gi + 1] = v.ioff[(size_t)gi] + cnt[(size_t)gi];
                v.hoff[(size_t)gi + 1] = v.hoff[(size_t)gi] + cnt[(size_t)ngroups + gi] * v.L * 30;
            }
            // items and rows with sub-filter <= q, for the per-launch work of tile t (q <= t + 1)
            v.qItemsUpTo.assign((size_t)v.qcap + 1, 0);
            v.qRowsUpTo.assign((size_t)v.qcap + 1, 0);
            for (int q = 0; q <= v.qcap; ++q) {
                const size_t i = 2 * (size_t)ngroups + (size_t)q, r = i + (size_t)v.qcap + 1;
                v.qItemsUpTo[(size_t)q] = (q ? v.qItemsUpTo[(size_t)q - 1] : 0) + cnt[i];
                v.qRowsUpTo[(size_t)q] = (q ? v.qRowsUpTo[(size_t)q - 1] : 0) + cnt[r];
            }
            v.nitems = v.ioff.back();
            v.hvals = v.hoff.back();
            // window ring: the level's largest sub-filter over the bands
            int qmax = 1;
            for (long pp = 0; pp < npairs; ++pp) {
                const int len = h_len[(size_t)pp], k0 = h_kb[(size_t)pp];
                const int jhi = std::min(std::min(k0 + len - 1, maxn), v.lagHi);
                if (len > 0 && std::max(k0, v.lagLo) <= jhi)
                    qmax = std::max(qmax, (jhi - 1) / v.L);
            }
            v.nwr = qmax;
            v.strideU = histU * 6 * (long)v.L * B;
            v.strideT = histT * 6 * (long)v.L * B;
            bytes += 8.0 * ((double)v.hvals + kFftHPad + (double)(v.strideU + v.strideT) * v.nwr) +
                     32.0 * (double)v.nitems + (l > 0 ? 8.0 * (double)v.L * (double)n * B : 0.0);
        }
        const int Lmax = kFftTile << (nl - 1);
        // the spectral part's buffers, and what the time loop allocates after them
        const double nB = (double)n * B, ns = (double)nstep + 1;
        bytes += 8.0 * ((double)ne * 6 * Lmax * B + (double)kFftTile * nB);
        const double rest = 8.0 * (nB * (2.0 * restart + 6) + (double)(histU + histT) * 3 * B * (2 * ns + 1) +
                                   3.0 * std::max<long>(1, dof.knownBlockCount) * B +
                                   std::max(1, batchedMfma ? batchedFarParts : 0) * kBatchTile * nB +
                                   (deviceIO ? 0.0 : 3.0 * B * ns * n));
        size_t freeB = 0, totalB = 0;
        CheckB(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
        freeB += PooledBytesCached();
        if (bytes + rest > 0.92 * (double)freeB) {
            fft = false;  // the filter spectra do not fit: every lag lag-by-lag
            nl = 0;
        } else {
            for (int l = 0; l < nl; ++l) {
                FftLevel& v = lv[l];
                v.itemOff.upload(v.ioff.data(), v.ioff.size(), stream);
                v.hOff.upload(v.hoff.data(), v.hoff.size(), stream);
                {
                    // the groups with the most item rows start first, so the
                    // launch does not end on a long one
                    std::vector<long> rowsOf((size_t)ngroups);
                    for (int gi = 0; gi < ngroups; ++gi)
                        rowsOf[(size_t)gi] = (v.hoff[(size_t)gi + 1] - v.hoff[(size_t)gi]) / ((long)v.L * 30);
                    std::vector<int> ord((size_t)ngroups);
                    for (int gi = 0; gi < ngroups; ++gi)
                        ord[(size_t)gi] = gi;
                    std::stable_sort(ord.begin(), ord.end(),
                                     [&](int a, int b2) { return rowsOf[(size_t)a] > rowsOf[(size_t)b2]; });
                    v.order.upload(ord.data(), ord.size(), stream);
                }
                v.items.alloc(4 * (size_t)std::max<long>(1, v.nitems));  // 32 bytes per item
                FftItem* items = reinterpret_cast<FftItem*>(v.items.p);
                LaunchFftItems(grid, d_refs.p, d_kb.p, d_len.p, maxn, v.L, v.lagLo, v.lagHi, ngroups, d_grpDom.p,
                               d_grpRowList.p, d_grpRows.p, v.itemOff.p, v.hOff.p, nullptr, nullptr, nullptr, 0, items,
                               stream);
                v.H.alloc((size_t)v.hvals + kFftHPad);
                LaunchZero(v.H.p + v.hvals, kFftHPad, stream);
                if (historyMixed)
                    LaunchFftFilters(grid, d_off.p, d_kb.p, d_len.p, (const float*)d_mts32.p, (const float*)d_mgs32.p,
                                     historySplit ? d_mtsLo.p : nullptr, historySplit ? d_mgsLo.p : nullptr, ngroups,
                                     d_grpDom.p, d_grpRowList.p, v.itemOff.p, v.nitems, items, maxn, v.L, v.lagLo,
                                     v.lagHi, v.H.p, stream);
                else
                    LaunchFftFilters(grid, d_off.p, d_kb.p, d_len.p, (const double*)d_mts.p, (const double*)d_mgs.p,
                                     nullptr, nullptr, ngroups, d_grpDom.p, d_grpRowList.p, v.itemOff.p, v.nitems,
                                     items, maxn, v.L, v.lagLo, v.lagHi, v.H.p, stream);
                v.WU.alloc((size_t)v.strideU * v.nwr);
                v.WT.alloc((size_t)v.strideT * v.nwr);
                if (l > 0)
                    v.hfar.alloc((size_t)v.L * (size_t)n * B);
            }
            d_Y.alloc((size_t)ne * 6 * Lmax * B);
            d_hfar16.alloc((size_t)kFftTile * (size_t)n * B);
            CheckB(hipStreamSynchronize(stream), "spectral set-up");
        }
        timing[16] = NowB() - fs0;
    }
    const int lagMax = fft ? std::min(maxn, kFftLagMin - 1) : maxn;  // the lag-by-lag kernels' last lag
    double t0 = NowB();
    BatchWork& w = bw;
    w.V.alloc((size_t)n * (size_t)(restart + 1) * (size_t)B);
    w.W.alloc((size_t)n * (size_t)restart * (size_t)B);
    w.u.alloc((size_t)n * B);
    w.u2.alloc((size_t)n * B);
    w.vh.alloc((size_t)n * B);
    w.h.alloc((size_t)(restart + 2) * B);
    w.h2.alloc((size_t)(restart + 2) * B);
    w.QT.alloc((size_t)(restart + 1) * (size_t)(restart + 1) * B);
    w.R.alloc((size_t)restart * (size_t)restart * B);
    w.f.alloc((size_t)(restart + 2) * B);
    w.phibar.alloc((size_t)B);
    w.scal.alloc(2 * (size_t)B);
    w.stag.alloc((size_t)B);
    w.active.alloc((size_t)B);
    w.its.alloc((size_t)B);
    w.flg.alloc((size_t)B);
    w.left.alloc(1);
    if (!w.hLeft)
        CheckB(hipHostMalloc((void**)&w.hLeft, sizeof(int), hipHostMallocDefault), "pinned");
    w.lastIters = 1;
    w.part.alloc((size_t)kBatchRedBlocks * (size_t)(restart + 1) * B);
    DevBuf<double> d_bd, d_lu, d_lt, d_gU, d_gT, d_FU, d_FT, d_known, d_rhs, d_x, d_hfar;
    const size_t outN = (size_t)B * nstore * (size_t)n;
    const double* bd = known_lu;
    double* olu = lu;
    double* olt = lt;
    std::unique_ptr<InputFeed> feed;
    if (!deviceIO) {
        d_bd.alloc(outN);
        feed.reset(new InputFeed(nstore, n, B, known_lu, d_bd.p));
        d_lu.alloc(outN);
        d_lt.alloc(outN);
        bd = d_bd.p;
        olu = d_lu.p;
        olt = d_lt.p;
    }
    d_gU.alloc((size_t)histU * 3 * nstore * B);
    d_gT.alloc((size_t)histT * 3 * nstore * B);
    // filtered series: nstore + 1 rows per column, the last never written (zero)
    d_FU.alloc((size_t)histU * 3 * (nstore + 1) * B);
    d_FT.alloc((size_t)histT * 3 * (nstore + 1) * B);
    d_known.alloc(3 * (size_t)std::max<long>(1, dof.knownBlockCount) * B);
    d_rhs.alloc((size_t)n * B);
    d_x.alloc((size_t)n * B);
    const int farParts = batchedMfma ? batchedFarParts : 0;
    const int hfarParts = batchedMfma ? std::max(1, farParts) : kFarBWaves;  // partial sums the near part adds
    d_hfar.alloc((size_t)hfarParts * kBatchTile * (size_t)n * B);
    for (DevBuf<double>* z : {&d_gU, &d_gT, &d_FU, &d_FT, &d_known})
        LaunchZero(z->p, (long)z->n, stream);
    LaunchZero(olu, (long)outN, stream);
    LaunchZero(olt, (long)outN, stream);
    const unsigned short* mLo = historySplit ? d_mtsLo.p : nullptr;
    const unsigned short* gLo = historySplit ? d_mgsLo.p : nullptr;
    std::vector<int> it((size_t)B), fl((size_t)B);
    std::vector<char> alive((size_t)B, 1);
    for (int b = 0; b < B; ++b)
        solved[b] = 0;
    // steps a case does not reach keep zero counts and flags
    std::fill_n(iters, (size_t)B * nstep, 0);
    std::fill_n(flags, (size_t)B * nstep, 0);
    double convTime = 0.0, gmresTime = 0.0, farTime = 0.0, fftTime = 0.0, fftProdTime = 0.0;
    int farLaunches = 0, fftLaunches = 0, K016 = 0;
    // timings of the spectral part: totals and per level
    auto fftTimings = [&]() {
        double fl = 0.0, by = 0.0;
        for (int l = 0; l < nl; ++l) {
            fl += lv[l].flops;
            by += lv[l].bytes;
            timing[26 + l] = lv[l].prodTime;
            timing[30 + l] = lv[l].flops;
            timing[34 + l] = lv[l].launches;
            timing[38 + l] = lv[l].bytes;
        }
        timing[13] = fftTime;
        timing[14] = fftLaunches;
        timing[15] = fft ? 1.0 : 0.0;
        timing[17] = fftProdTime;
        timing[18] = fl;
        timing[19] = by;
        timing[42] = nl;
    };
    struct EvGuard {
        hipEvent_t e[4 + 2 * kFftMaxLevels] = {};
        ~EvGuard()
        {
            for (hipEvent_t x : e)
                if (x)
                    (void)hipEventDestroy(x);
        }
    } evg;
    for (hipEvent_t& x : evg.e)
        CheckB(hipEventCreate(&x), "event");
    hipEvent_t fa = evg.e[0], fb = evg.e[1];
    // host buffers: the solved steps go out while the loop runs on
    std::unique_ptr<OutputDrain> drain;
    if (!deviceIO && nstep > 0)
        drain.reset(new OutputDrain(nstep, nstore, n, B, d_lu.p, d_lt.p, lu, lt));
    int K0 = 0;
    for (int K = 1; K <= nstep; ++K) {
        if (feed)
            feed->Need(K, stream);
        LaunchKnownB(d_refs.p, bd, (int)ne, nstore, K, B, d_known.p, stream);
        LaunchSpmmB(ne, d_Krp.p, d_Kcol.p, d_Kval.p, d_known.p, d_rhs.p, B, stream);
        if (K > 1) {
            double c0 = NowB();
            if (fft && (K == 2 || K - K016 >= kFftTile)) {
                // every level whose tile starts here (the longest first), then
                // level 0, whose inverse adds the other levels' samples of
                // these 16 steps
                K016 = K;
                CheckB(hipEventRecord(fa, stream), "event");
                FftAdds adds{};
                adds.n = 0;
                int ran = 0;
                for (int l = nl - 1; l >= 0; --l) {
                    FftLevel& v = lv[l];
                    v.ran = false;
                    if ((K - 2) % v.L == 0) {
                        v.K0 = K;
                        v.ran = true;
                        const int t = (K - 2) / v.L;
                        const int Tn = std::min(v.L, nstep - K + 1);
                        LaunchFftWindows(d_gU.p, histU, nstore, B, v.L, t, v.WU.p + (size_t)(t % v.nwr) * v.strideU,
                                         stream);
                        LaunchFftWindows(d_gT.p, histT, nstore, B, v.L, t, v.WT.p + (size_t)(t % v.nwr) * v.strideT,
                                         stream);
                        CheckB(hipEventRecord(evg.e[4 + 2 * l], stream), "event");
                        LaunchConvFarFft(ngroups, v.order.p, d_grpRowList.p, d_grpRows.p, v.itemOff.p,
                                         reinterpret_cast<const FftItem*>(v.items.p), v.H.p, v.WU.p, v.WT.p, v.strideU,
                                         v.strideT, v.nwr, v.L, t, B, d_Y.p, batchedFftMfma != 0, batchedFftPace,
                                         stream);
                        CheckB(hipEventRecord(evg.e[5 + 2 * l], stream), "event");
                        const FftAdds none{};
                        if (l > 0)
                            LaunchFftInverse(d_Y.p, (int)ne, B, v.L, Tn, v.hfar.p, none, stream);
                        else
                            LaunchFftInverse(d_Y.p, (int)ne, B, v.L, Tn, d_hfar16.p, adds, stream);
                        ++ran;
                        // algorithmic work of the products: per row of an active
                        // item (sub-filter <= t + 1) and load case, L - 1 slots of
                        // 18 complex multiply-adds (8 flops each, in 4 FMAs) and
                        // slot 0's 36 real ones; bytes: its filter spectra (L x
                        // 30 values) once, the window spectra (L slots x 12
                        // values per load case) once per item and the summed
                        // spectra written once
                        const size_t qa = (size_t)std::min(t + 1, v.qcap);
                        const double rowsA = (double)v.qRowsUpTo[qa], itemsA = (double)v.qItemsUpTo[qa];
                        v.flops += rowsA * ((v.L - 1) * 144.0 + 72.0) * B;
                        v.bytes += rowsA * v.L * 30 * 8 + itemsA * v.L * 12 * 8.0 * B + (double)ne * 6 * v.L * B * 8;
                        ++v.launches;
                    }
                    if (l > 0)
                        adds.p[adds.n++] = v.hfar.p + (size_t)(K - v.K0) * (size_t)n * B;
                }
                CheckB(hipEventRecord(fb, stream), "event");
                CheckB(hipEventSynchronize(fb), "far part");
                float ms = 0.f;
                CheckB(hipEventElapsedTime(&ms, fa, fb), "elapsed");
                fftTime += ms * 1e-3;
                for (int l = 0; l < nl; ++l)
                    if (lv[l].ran) {
                        float msp = 0.f;
                        CheckB(hipEventElapsedTime(&msp, evg.e[4 + 2 * l], evg.e[5 + 2 * l]), "elapsed");
                        lv[l].prodTime += msp * 1e-3;
                        fftProdTime += msp * 1e-3;
                    }
                fftLaunches += ran;
            }
            if (K == 2 || K - K0 >= kBatchTile) {
                int prev = K0;
                K0 = K;
                int x0 = prev == 0 ? 0 : prev - 1;
                LaunchFilterB(d_gU.p, histU, nstore, B, x0, K0, K0, d_FU.p, stream);
                LaunchFilterB(d_gT.p, histT, nstore, B, x0, K0, K0, d_FT.p, stream);
                int Tn = std::min(kBatchTile, nstep - K0 + 1);
                CheckB(hipEventRecord(fa, stream), "event");
                if (historyMixed)
                    LaunchConvFarB(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const float*)d_mts32.p,
                                   (const float*)d_mgs32.p, mLo, gLo, d_FU.p, d_FT.p, nstore, K0, Tn, lagMax, (int)ne,
                                   B, d_hfar.p, farParts, stream);
                else
                    LaunchConvFarB(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const double*)d_mts.p,
                                   (const double*)d_mgs.p, nullptr, nullptr, d_FU.p, d_FT.p, nstore, K0, Tn, lagMax,
                                   (int)ne, B, d_hfar.p, farParts, stream);
                CheckB(hipEventRecord(fb, stream), "event");
                CheckB(hipEventSynchronize(fb), "far part");
                float ms = 0.f;
                CheckB(hipEventElapsedTime(&ms, fa, fb), "elapsed");
                farTime += ms * 1e-3;
                ++farLaunches;
            }
            if (historyMixed)
                LaunchConvNearB(d_refs.p, d_nearPtr.p, d_nearPair.p, d_nearField.p, d_off.p, d_kb.p, d_len.p,
                                (const float*)d_mts32.p, (const float*)d_mgs32.p, mLo, gLo, d_gU.p, d_gT.p, nstore, K,
                                K0, lagMax, (int)ne, B, d_hfar.p, hfarParts, d_rhs.p, stream,
                                fft ? d_hfar16.p : nullptr, K016);
            else
                LaunchConvNearB(d_refs.p, d_nearPtr.p, d_nearPair.p, d_nearField.p, d_off.p, d_kb.p, d_len.p,
                                (const double*)d_mts.p, (const double*)d_mgs.p, nullptr, nullptr, d_gU.p, d_gT.p,
                                nstore, K, K0, lagMax, (int)ne, B, d_hfar.p, hfarParts, d_rhs.p, stream,
                                fft ? d_hfar16.p : nullptr, K016);
            CheckB(hipStreamSynchronize(stream), "batched convolution");
            convTime += NowB() - c0;
        }
        double g0 = NowB();
        GmresBatched(B, d_rhs.p, d_x.p, restart, tol, it.data(), fl.data());
        gmresTime += NowB() - g0;
        bool any = false;
        for (int b = 0; b < B; ++b) {
            if (!alive[(size_t)b])
                continue;
            iters[(size_t)b * nstep + (K - 1)] = it[(size_t)b];
            flags[(size_t)b * nstep + (K - 1)] = fl[(size_t)b];
            solved[b] = K;
            if (fl[(size_t)b] != 0)
                alive[(size_t)b] = 0;
            else
                any = true;
        }
        // step K's outputs and its state series in one pass
        LaunchScatterB(d_refs.p, d_geo.p, bd, d_x.p, (int)ne, nstore, K, B, olu, olt, stream, d_gU.p, d_gT.p);
        if (drain && (K % OutputDrain::kSteps == 0 || K == nstep))
            drain->Ready((K - 1) / OutputDrain::kSteps + 1, stream);
        if (!any)
            break;
    }
    if (drain) {
        drain->Ready(drain->Chunks(), stream);
        drain->Stop(true);
    }
    // the last step whose solution a case keeps: a step whose GMRES stopped
    // with a flag keeps no solution (the reference leaves the loop before
    // storing it, multidomain.cpp:3884-3888), nor does any step after it
    std::vector<int> kept((size_t)B);
    for (int b = 0; b < B; ++b)
        kept[(size_t)b] = solved[b] - (solved[b] > 0 && flags[(size_t)b * nstep + (solved[b] - 1)] != 0 ? 1 : 0);
    if (deviceIO) {
        // those steps are left zero; step 0 in the global frame
        for (int b = 0; b < B; ++b)
            if (kept[(size_t)b] < nstep) {
                size_t from = ((size_t)b * nstore + (size_t)kept[(size_t)b] + 1) * (size_t)n;
                size_t cnt = (size_t)(nstep - kept[(size_t)b]) * (size_t)n;
                LaunchZero(olu + from, (long)cnt, stream);
                LaunchZero(olt + from, (long)cnt, stream);
            }
        LaunchStep0Out(d_refs.p, d_geo.p, bd, (int)ne, nstore, B, olu, olt, stream);
        CheckB(hipStreamSynchronize(stream), "batched solve");
        CheckB(hipGetLastError(), "batched solve launch");
        timing[4] = convTime;
        timing[5] = gmresTime;
        timing[6] = farTime;
        timing[7] = farLaunches;
        fftTimings();
        bw.release();
        return;
    }
    CheckB(hipStreamSynchronize(stream), "batched solve");
    CheckB(hipGetLastError(), "batched solve launch");
    // the steps without a kept solution are left zero, as a single solve leaves them
    for (int b = 0; b < B; ++b)
        for (int K = kept[(size_t)b] + 1; K <= nstep; ++K) {
            std::fill_n(lu + ((size_t)b * nstore + K) * n, n, 0.0);
            std::fill_n(lt + ((size_t)b * nstore + K) * n, n, 0.0);
        }
    // step 0 of each case: lt = 0, Convert (lu <-> lt for BCID 456), LocToAbs
    for (int b = 0; b < B; ++b) {
        const double* in = known_lu + (size_t)b * nstore * n;
        double* ou = lu + (size_t)b * nstore * n;
        double* ot = lt + (size_t)b * nstore * n;
        for (long i = 0; i < ne; ++i) {
            double u[3], t[3];
            for (int k = 0; k < 3; ++k) {
                u[k] = in[3 * i + k];
                t[k] = 0.0;
            }
            if (bcid[(size_t)i] == 456)
                for (int k = 0; k < 3; ++k)
                    std::swap(u[k], t[k]);
            const double* F = geo[(size_t)i].trans;
            for (int r = 0; r < 3; ++r) {
                ou[3 * i + r] = F[r] * u[0] + F[3 + r] * u[1] + F[6 + r] * u[2];
                ot[3 * i + r] = F[r] * t[0] + F[3 + r] * t[1] + F[6 + r] * t[2];
            }
        }
    }
    timing[4] = convTime;
    timing[5] = gmresTime;
    timing[6] = farTime;
    timing[7] = farLaunches;
    fftTimings();
    (void)t0;
    bw.release();
}

}  // namespace tdbem
