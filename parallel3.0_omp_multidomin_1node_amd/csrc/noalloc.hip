// The time loop without any device allocation, for callers that cannot
// allocate (code under graph capture, tight serving loops): every working
// buffer of one tdbem_solve_device is carved from `scratch`, a file-scope
// device array of kScratchBytes held by each GPU the library runs on until
// the process ends.  One copy exists per device; its address on the current
// device is taken once, at context set-up, with hipGetSymbolAddress (the
// array's name in host code is a host-side stand-in, never a device address).
// Calls that overlap on one GPU share the array and race.
#include "solver.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

namespace tdbem {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratchPtr[kMaxDevices] = {};

void CheckNA(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device bytes one SolveNoAlloc of this shape carves (each buffer rounded up
// to the arena's 256-byte alignment): the solve's state histories, filtered
// series, right-hand side, unknowns and far-part partial sums, and the GMRES
// workspace.
size_t NeedBytes(long ne, long histU, long histT, long known, int nstep, long restart, bool filtered)
{
    const size_t n3 = 3 * (size_t)ne, nstore = (size_t)nstep + 1;
    auto r = [](size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255; };
    size_t b = 0;
    b += r(8 * (size_t)histU * 3 * nstore) + r(8 * (size_t)histT * 3 * nstore);
    b += r(8 * 3 * (size_t)std::max<long>(1, known)) + 2 * r(8 * n3);
    b += r(8 * (size_t)kFarPartsMax * (size_t)kConvTileMax * n3);
    if (filtered)
        b += r(8 * (size_t)ne * (nstore + 1) * 6);
    const size_t R = (size_t)restart;
    b += r(8 * n3 * (R + 1)) + r(8 * n3 * R) + 3 * r(8 * n3);
    b += 2 * r(8 * (R + 2)) + r(8 * (R + 1) * (R + 1)) + r(8 * R * R) + r(8 * (R + 2)) + r(8 * 4) + r(4);
    b += r(8 * (size_t)kGmBlocksMax * (R + 1));
    return b;
}

}  // namespace

// The scratch array's address on the current device, looked up once per
// device (Case::Create calls it, so a solve never does the lookup).
void* ScratchOnDevice()
{
    int dev = 0;
    CheckNA(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
        throw std::runtime_error("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        CheckNA(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratchPtr[dev] = p;
    }
    return g_scratchPtr[dev];
}

size_t ScratchBytes() { return kScratchBytes; }

void Case::SolveNoAlloc(int nstep, const double* d_known, long restart, double tol, double* d_lu, double* d_lt,
                        int* iters, int* flags, int* solved)
{
    if (!step0Done || !historyDone || maxn < 1)
        throw std::runtime_error("step 0 and the history must be assembled before a solve without allocation");
    if (historyMixed)
        throw std::runtime_error("the solve without allocation takes FP64 history bands (mixed 0)");
    RefuseStreamed("the solve without allocation");
    if (!convTiled || (nearTile != 8 && nearTile != 16))
        throw std::runtime_error("the solve without allocation needs the time-blocked convolution and its near lists");
    if (!d_gst.p || !h_gst || !h_pin)
        throw std::runtime_error("GMRES state missing (set up by tdbem_create)");
    const int tile = nearTile;  // the tile whose near lists are built: none are rebuilt here
    const bool filtered = convFiltered && tile == 16;
    const size_t need = NeedBytes(ne, dof.historyUBlockCount, dof.historyTBlockCount, dof.knownBlockCount, nstep,
                                  restart, filtered);
    if (need > kScratchBytes)
        throw std::runtime_error("this solve needs " + std::to_string(need) + " bytes of scratch; the solve without "
                                 "allocation holds at most " + std::to_string(kScratchBytes) +
                                 " (fewer steps, elements or a smaller restart)");
    int dev = 0;
    CheckNA(hipGetDevice(&dev), "hipGetDevice");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratchMu);
        base = dev >= 0 && dev < kMaxDevices ? g_scratchPtr[dev] : nullptr;
    }
    if (!base)
        throw std::runtime_error("scratch address of this device not set up (create a context on it first)");

    // The GMRES workspace of the context moves aside (no free) and is carved
    // from the scratch for this call, then put back; the step's tile, the
    // GMRES driver (direct launches: no graph) and the mode are restored.
    DevBuf<double>* ws[] = {&d_V, &d_W, &d_u, &d_u2, &d_vh, &d_h, &d_h2, &d_QT, &d_R, &d_f, &d_scal, &d_part};
    struct Saved {
        double* p;
        size_t n;
    } saved[12];
    for (int i = 0; i < 12; ++i) {
        saved[i] = {ws[i]->p, ws[i]->n};
        ws[i]->p = nullptr;
        ws[i]->n = 0;
    }
    int* savedStag = d_stag.p;
    size_t savedStagN = d_stag.n;
    d_stag.p = nullptr;
    d_stag.n = 0;
    const long sRestart = gm_restart, sN = gm_n;
    const int sDevice = gmresDevice, sTile = convTile, sLast = gm_lastIters;
    gm_restart = -1;
    gmresDevice = 2;
    convTile = tile;
    noAllocMode = true;
    ScratchArena arena{static_cast<char*>(base), kScratchBytes, 0};
    SetScratchArena(&arena);
    auto restore = [&]() {
        for (DevBuf<double>* b : ws)
            b->free();  // views into the arena: forgotten, not freed
        d_stag.free();
        SetScratchArena(nullptr);
        for (int i = 0; i < 12; ++i) {
            ws[i]->p = saved[i].p;
            ws[i]->n = saved[i].n;
        }
        d_stag.p = savedStag;
        d_stag.n = savedStagN;
        gm_restart = sRestart;
        gm_n = sN;
        gmresDevice = sDevice;
        convTile = sTile;
        gm_lastIters = sLast;
        noAllocMode = false;
    };
    try {
        Solve(nstep, d_known, restart, tol, d_lu, d_lt, iters, flags, solved, true);
    } catch (...) {
        restore();
        throw;
    }
    restore();
}

}  // namespace tdbem
