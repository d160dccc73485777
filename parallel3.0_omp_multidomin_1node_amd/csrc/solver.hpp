// One multi-domain case on one GPU: host model and device buffers.
#pragma once

#include "gmres.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "precond.hpp"

#include <string>
#include <vector>

namespace tdbem {

// Device bytes held by the cache of large freed buffers, and its release.
size_t PooledBytesCached();
void ReleaseCachedDeviceMemory();

// A copy ordered after the work already queued on stream s, complete on
// return (no use of the null stream, so contexts on other streams and
// threads are not synchronized by it).
inline hipError_t CopyOnStream(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    if (e != hipSuccess)
        return e;
    return hipStreamSynchronize(s);
}

// Caller-provided device memory that DevBuf::alloc carves from (256-byte
// aligned bump allocation) while it is set on the calling thread; DevBuf::free
// of memory inside it only forgets the pointer.
struct ScratchArena {
    char* base;
    size_t size;
    size_t used;
};
void SetScratchArena(ScratchArena* a);  // nullptr ends it
// Address of the no-allocation solve's scratch array on the current device
// (looked up once per device) and its size.
void* ScratchOnDevice();
size_t ScratchBytes();

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { free(); }
    void alloc(size_t count);
    void free();
    void upload(const T* h, size_t count, hipStream_t s);
};

// Operator of one GMRES solve: block CCSR matrix (rp, col, val) and either the
// leaf preconditioner inverses (inv) or an explicit block CCSR preconditioner
// (prp, pcol, pval; FP64 only).
template <typename T>
struct GmresOp {
    const int* rp;
    const int* col;
    const T* val;
    const T* inv;
    const int* prp;
    const int* pcol;
    const T* pval;
};

// Forward history a solve hands to the adjoint: unknown vectors x_K
// (K = 0 .. NStep), state columns and the device copy of the known inputs.
struct AdjointKeep {
    DevBuf<double> X, gU, gT, bd;
};

// Workspace of the lockstep GMRES of a batch of load cases: the per-case
// iteration counts, flags and activity on the device, the count of cases
// still iterating (device and pinned host copy) and the last solve's
// iteration count (how many iterations go before the host first looks).
struct BatchWork {
    DevBuf<double> V, W, u, u2, vh, h, h2, QT, R, f, phibar, scal, part;
    DevBuf<int> stag, active, its, flg, left;
    int* hLeft = nullptr;
    int lastIters = 1;
    void release()
    {
        for (DevBuf<double>* b : {&V, &W, &u, &u2, &vh, &h, &h2, &QT, &R, &f, &phibar, &scal, &part})
            b->free();
        for (DevBuf<int>* b : {&stag, &active, &its, &flg, &left})
            b->free();
        if (hLeft)
            (void)hipHostFree(hLeft);
        hLeft = nullptr;
        lastIters = 1;
    }
};

constexpr int kGmBlocksMax = 128;

// Launch shapes picked by Case::Autotune: far-part wavefronts per row (0 =
// one pair's lags per item) and GMRES reduction blocks, with the timings of
// the chosen shapes (0 when taken from the cache).
struct TuneChoice {
    int farParts, gmBlocks;
    double farMs, gmUs;
    bool cached;
};

class Case {
public:
    ~Case();
    void Create(long npoints, const double* points, long ne, const int* ele_pid, const int* bcid,
                const MultiDomainInputConfig& cfg, double E, double v, double rho, double dt, double gap);
    // The given (created) models stacked as the groups of one case: one
    // launch per kernel for all of them, each model's results those of its
    // own job (SolveGroups).
    void CreateGroup(const std::vector<const Case*>& models);
    bool Grouped() const { return groupEle.size() > 2; }
    int GroupCount() const { return groupEle.size() >= 2 ? (int)groupEle.size() - 1 : 1; }
    // The time loop of grouped models over FP64 resident bands: known_lu, lu
    // and lt stacked per step ([K][3 ne], group g's elements at groupEle[g]);
    // iters / flags [g][nstep], solved[g].  A group whose GMRES stops with a
    // flag other than 0 takes no further steps (the others go on); its lu / lt
    // from that step on are not its job's (the caller drops them).
    void SolveGroups(int nstep, const double* known_lu, long restart, double tol, double* lu, double* lt,
                     int* iters, int* flags, int* solved);
    void AssembleStep0(long maxleaf);
    void AssembleHistory(int maxn);
    double PrepareHistory(int maxn);  // cull ahead of AssembleHistory; returns the band bytes it will allocate
    TuneChoice Autotune(const std::string& cachePath);
    std::string TuneKey() const;
    void ReleaseHistory();
    // deviceIO: known_lu, lu and lt are device pointers (inputs already in HBM,
    // outputs left there); iters / flags / solved stay host arrays.
    void Solve(int nstep, const double* known_lu, long restart, double tol, double* lu, double* lt, int* iters,
               int* flags, int* solved, bool deviceIO = false);
    // Solve with device-resident I/O whose every device buffer is carved from
    // the file-scope scratch array of noalloc.hip (no allocation, no event, no
    // graph); FP64 bands only.  Sizes above the array are refused.
    void SolveNoAlloc(int nstep, const double* d_known, long restart, double tol, double* d_lu, double* d_lt,
                      int* iters, int* flags, int* solved);
    // B load cases over the same bands: known_lu, lu, lt are [b][NStep+1][3N],
    // iters / flags [b][NStep], solved [b].
    void SolveBatched(int nstep, int B, const double* known_lu, long restart, double tol, double* lu, double* lt,
                      int* iters, int* flags, int* solved, bool deviceIO = false);
    void HistoryStats(long* counts) const;
    // Reverse-time adjoint of the FP64 time loop for the misfit
    // J = sum_{K=1..NStep} sum_r |lu_K[r] - obs[K][r]|^2 over the receiver
    // elements r (global-frame displacements).  Returns J; grad[3 d + p] =
    // dJ/d(E, v, rho)_p of domain d; lu (may be null) receives the forward
    // displacements.  fdStep: relative step of the central differences of
    // the operators over the materials.
    double Adjoint(int nstep, const double* known_lu, const int* receivers, int nrec, const double* obs, long restart,
                   double tol, double fdStep, double* grad, double* lu);
    // Per-domain materials (E, v, rho) x domains; setting them rebuilds the
    // DynaMats and the singular self tables (step 0 and history must be
    // assembled again before a solve).
    void GetMaterials(double* evr) const;
    void FreezeStructure(bool on);
    // (Re)computes the singular self tables when they are missing or were made
    // by the other kernel than quadStaged asks for (reference order when false).
    void EnsureSelfTables();
    void SetMaterials(const double* evr);
    // Patch-wise waves of the current (or frozen) judges (cached while frozen).
    void PackPwWaves();
    void ExportMatrix(int which, int step, std::vector<int>& rp, std::vector<int>& col, std::vector<double>& val);
    void ExportRows(int which, int step, long row0, long row1, std::vector<int>& rp, std::vector<int>& col,
                    std::vector<double>& val);
    bool HistoryReady() const { return historyDone; }
    // Assembled history bands to a file / from a file written for the same
    // model (same cull bands, materials, dt and gap); the loaded history
    // replaces any assembled one.
    void SaveHistory(const std::string& path);
    void LoadHistory(const std::string& path);
    void LoadHistoryBlocks(FILE* f, const std::string& path, uint32_t precision, int fileMaxn);

    long ne = 0;
    double dt = 0.0;
    std::vector<long> groupEle;  // element bounds of the stacked models (CreateGroup); empty otherwise
    MultiDomainModel model;
    GlobalDofMap dof;
    std::vector<int> bcid, domOf, surface;
    QuadTables tables;
    std::vector<ElemGeo> geo;
    std::vector<NodeRefs> refs;
    PairGrid grid{};
    long npairs = 0;
    long entries = 0;
    int maxn = 0;
    CCSRMat A, KnownM;
    LeafPreconditioner pre;
    // 0 singular, 1 cull, 2 history, 3 step 0, 4 convolution, 5 GMRES, 6 far
    // convolution kernels, 7 their launches, 8 history quadrature kernels;
    // adjoint: 9 forward solve, 10 backward sweep, 11 step-0 and self-table
    // differences with the correlation store, 12 history-block derivatives
    // streamed history: 20 band slabs, 21 their regeneration in the time loop (device events);
    // 22-25 the staged quadrature's evaluation counts (with quadCount); 26-42
    // spectral levels; 43, 44 the patch-wise entries and waves of the last
    // staged quadrature (with quadCount)
    double timing[45] = {0};
    bool convTiled = true;   // time-blocked convolution (false: one full pass per step)
    int convTile = 16;       // steps per far-part tile (8 or 16); fixes the near lists built by the cull
    bool adjointAnalytic = true;  // history-block derivatives analytic (false: operator differences)
    bool convFiltered = true;  // 16-step far tiles read a pre-filtered state series and prefetch work items
    bool convPacked = true;    // filtered far tiles pack 64 band entries per work item across pairs
    int farParts = 4;          // wavefronts (and partial sums) per source row of the packed far part
    bool nearPerEntry = true;  // near part: one (pair, lag) entry per lane (false: one pair per lane)
    int batchedFftLevels = 1;  // levels of the spectral part (sub-filter lengths 16, 32, ... 16 << (levels - 1))
    int batchedFftPace = 0;  // items between the pacing barriers of the spectral products (0: none)
    int batchedFftMfma = 1;  // the spectral products on the FP64 matrix cores (1) or with vector FMAs (0)
    int batchedFft = -1;  // batched far part in the frequency domain: 1 on, 0 off, -1 for 16 or more load cases
    bool batchedMfma = false;  // batched far part on the FP64 matrix cores (false: vector FMAs)
    int batchedFarParts = 16;  // column parts (partial sums) of the matrix-core batched far part
    int gmBlocks = 32;         // reduction blocks of the GMRES dot products (at most; <= kGmBlocksMax)
    bool noAllocMode = false;  // SolveNoAlloc in progress
    int gmresDevice = 1;       // FP64 GMRES: 0 host-driven, 1 device-driven hipGraphs, 2 device-driven launches
    int selfMode = -1;       // kernel of the current self tables: 1 wave per element, 0 reference order
    bool quadCount = false;  // count the staged quadrature's evaluations (timing[22..25]; costs a little)
    bool quadStaged = true;  // wave-per-pair quadrature with LDS-staged point geometry (false: lane per entry)
    int mixed = 0;  // for the next history assembly: 0 FP64, 1 FP32 blocks, 2 FP32 + bf16 residual blocks
    // Streamed history (bands larger than the device budget): 1 always, 0
    // never (an assembly that does not fit is refused), -1 when they do not fit.
    int historyStream = -1;
    double historyBudget = 0.0;  // device bytes the history may take (0: 92 % of the free memory)
    bool historyStreamed = false;  // the bands are regenerated per far tile, slab by slab
    bool historyMixed = false, historySplit = false;
    hipStream_t stream = nullptr;

    DevBuf<double> d_u1, d_t1, d_t2;

private:
    void FinishCreate();
    // grouped GMRES of one step: per group iterations and flags
    void GmresDeviceGroups(const double* d_b, double* d_x, long restart, double tol, const int* active, int* its,
                           int* fls);
    struct GroupWork {
        long restart = -1;
        DevBuf<double> V, W, u, u2, vh, h, h2, QT, R, f, phibar, part, st;
        DevBuf<int> stag, active;
        DevBuf<GmresGroupTable> tab;
        GmresGroupTable t{};
        GmState* hst = nullptr;  // pinned copies of the group states
        hipGraphExec_t gBegin = nullptr, gIter = nullptr;
        const double* graphB = nullptr;
        double graphTol = -1.0;
        int lastIters = 1;
        void release();
    } gw;
    void Cull(int maxn);
    void BuildNearLists(int tile);
    void StagedQuadrature(double* mts, double* mgs, float* mts32, float* mgs32, unsigned char* alive);
    void ComputeStep0(CCSRMat& Aout, CCSRMat& Kout);
    std::vector<long> NearPairs();
    bool frozen = false;
    DevBuf<unsigned char> d_cls, d_cls0;
    AdjointKeep* keep = nullptr;
    bool StructureValid() const;
    void PrepareGmres(long restart);
    void PrepareMixed(long restart);
    int Gmres(const double* d_b, double* d_x, long restart, double tol, int* iters);
    template <typename T>
    int GmresCore(const double* d_b, double* d_x, long restart, double tol, int* iters, const GmresOp<T>& op, T* V,
                  T* W, T* u, T* u2);
    int GmresMixed(const double* d_b, double* d_x, long restart, double tol, int* iters);
    int GmresDevice(const double* d_b, double* d_x, long restart, double tol, int* iters,
                    const GmresOp<double>* op = nullptr);
    void ReleaseGraphs();
    bool precull = false;  // the last cull was run for the next history assembly
    void GmresBatched(int B, const double* rhs, double* x, long restart, double tol, int* its, int* flg);
    BatchWork bw;

    bool step0Done = false, historyDone = false;
    int culledFor = 0;
    DevBuf<QuadTables> d_tables;
    DevBuf<DynaMat> d_mats;  // per-domain materials, read through PairGrid::mats
    DevBuf<ElemGeo> d_geo;
    DevBuf<ElemPoints> d_pts;
    DevBuf<int> d_dom;
    DevBuf<NodeRefs> d_refs;
    DevBuf<int> d_kb, d_len;
    DevBuf<unsigned char> d_near;
    DevBuf<long> d_off, d_coff;
    long chunks = 0;
    long pwWaves = 0;  // patch-wise waves of the last history assembly
    DevBuf<PwWave> d_pwWaves;
    bool pwPacked = false;  // d_pwWaves holds the packing of the frozen structure
    DevBuf<long> d_nearPtr, d_nearPair;
    DevBuf<int> d_nearField;
    int nearTile = 0;  // tile the near lists were built for
    void HostBands() const;
    // streamed history: row slabs [slabRow0, slabRow1) with band entries
    // [slabE0, slabE1); the blocks of lags <= kStreamNearLag resident in a
    // compact array (pair p's at nearOff[p]); one slab buffer
    void AssembleStreamed(double budget);
    void StreamSlab(size_t i, int maxLag);  // the lags <= maxLag of slab i (and the rest of their chunks / waves)
    void ReleaseStreamed();
    void RefuseStreamed(const char* what) const;
    std::vector<long> slabRow0, slabRow1, slabE0, slabE1;
    std::vector<long> slabC0, slabC1, slabW0, slabW1;  // staged quadrature: regular chunks, patch-wise waves
    bool streamStaged = true;  // the slabs are regenerated by the staged kernels (else lane per entry)
    DevBuf<PwWave> d_slabWaves;
    DevBuf<long> d_nearOff;
    DevBuf<double> d_nmts, d_nmgs, d_smts, d_smgs;
    DevBuf<unsigned char> d_nalive, d_salive;
    mutable bool hostBands = false;  // h_kb, h_len, h_off hold the current cull
    mutable std::vector<int> h_kb, h_len;
    mutable std::vector<long> h_off;
    DevBuf<double> d_mts, d_mgs;
    DevBuf<float> d_mts32, d_mgs32;  // mixed precision: FP32 band blocks
    DevBuf<unsigned short> d_mtsLo, d_mgsLo;  // and their bf16 residuals (mixed == 2)
    DevBuf<unsigned char> d_alive;
    std::vector<double> h_mts, h_mgs;
    std::vector<unsigned char> h_alive;
    DevBuf<int> d_Arp, d_Acol, d_Krp, d_Kcol;
    DevBuf<double> d_Aval, d_Kval;
    DevBuf<long> d_posLeaf, d_leafBegin, d_leafSize, d_matOff, d_repid, d_outBlock;
    DevBuf<double> d_inv;
    // GMRES workspace
    long gm_restart = -1, gm_n = -1;
    int gm_blocks = 1;
    DevBuf<double> d_V, d_W, d_u, d_u2, d_vh, d_h, d_h2, d_QT, d_R, d_f, d_scal, d_part;
    DevBuf<int> d_stag;
    bool gm_mixed = false;
    DevBuf<float> d_V32, d_W32, d_u32, d_u232, d_Aval32, d_inv32;
    DevBuf<double> d_r, d_dx;
    double* h_pin = nullptr;
    // device-driven GMRES: state, its pinned copy, captured graphs and the
    // operands they were captured for
    DevBuf<double> d_gst;
    GmState* h_gst = nullptr;
    hipGraphExec_t graphBegin = nullptr, graphIter = nullptr;
    // what the captured kernels were given: every operand and size, the
    // right-hand side and the tolerance (a difference in any re-captures)
    GmresDevArgs graphArgs{};
    const double* graphB = nullptr;
    double graphTol = -1.0;
    bool graphValid = false;
    int gm_lastIters = 1;
};

}  // namespace tdbem
