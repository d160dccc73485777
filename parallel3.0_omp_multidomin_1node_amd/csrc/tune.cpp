// Launch-shape autotuner.  For an assembled case it times, with HIP events
// on the case's stream, the candidate shapes of the two kernels whose best
// shape depends on the case:
//   - the far part of the time-blocked convolution: packed items with 2, 4
//     or 8 wavefronts per source row, or one pair's lags per item, launched
//     on the last 16-step tile (every band entry inside the tile, so the
//     launch reads the whole band set); the filtered series is zero, which
//     does not change the work;
//   - the GMRES reductions: 8 .. 128 blocks for a four-column multi-dot and a
//     norm with their in-order finishes.
// The fastest shape of each is applied to the case.  Choices can be kept in
// a text cache keyed by the case's shape (elements, pairs, band entries,
// steps, domains, storage precision), one line per shape; a cached choice is
// applied without timing.  Different shapes reassociate the convolution and
// reduction sums, so solves before and after tuning agree to rounding.
#include "solver.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace tdbem {

namespace {

void CheckT(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// minimum over rounds of the event time of launch(), in ms
template <typename F>
double TimeMin(hipStream_t s, int rounds, F launch)
{
    hipEvent_t a, b;
    CheckT(hipEventCreate(&a), "event");
    CheckT(hipEventCreate(&b), "event");
    launch();  // warm
    double best = 1e30;
    for (int r = 0; r < rounds; ++r) {
        CheckT(hipEventRecord(a, s), "event");
        launch();
        CheckT(hipEventRecord(b, s), "event");
        CheckT(hipEventSynchronize(b), "event");
        float ms = 0.f;
        CheckT(hipEventElapsedTime(&ms, a, b), "elapsed");
        best = std::min(best, (double)ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

}  // namespace

std::string Case::TuneKey() const
{
    std::ostringstream k;
    k << "tdbem-tune-v1 " << ne << ' ' << npairs << ' ' << entries << ' ' << maxn << ' ' << grid.nd << ' '
      << (historyMixed ? (historySplit ? 2 : 1) : 0);
    return k.str();
}

TuneChoice Case::Autotune(const std::string& cachePath)
{
    RefuseStreamed("the autotuner");
    if (!historyDone)
        throw std::runtime_error("autotune needs the assembled history");
    TuneChoice c{convPacked ? farParts : 0, gmBlocks, 0.0, 0.0, false};
    const std::string key = TuneKey();
    if (!cachePath.empty()) {
        std::ifstream in(cachePath);
        std::string line;
        while (std::getline(in, line)) {
            size_t sep = line.find(" : ");
            if (sep == std::string::npos || line.substr(0, sep) != key)
                continue;
            std::istringstream v(line.substr(sep + 3));
            int fp = -1, gb = -1;
            if (v >> fp >> gb && fp >= 0 && fp <= kFarPartsMax && gb >= 1 && gb <= kGmBlocksMax) {
                c.farParts = fp;
                c.gmBlocks = gb;
                c.cached = true;
            }
        }
    }
    const long n3 = 3 * ne;
    if (!c.cached) {
        // far part on the last tile
        if (convTiled && convFiltered && convTile == 16) {
            const int nstore = maxn + 1;
            const int K0 = std::max(2, maxn - 15);
            const int Tn = std::min(16, maxn - K0 + 1);
            DevBuf<double> F, hfar;
            F.alloc((size_t)ne * (size_t)(nstore + 1) * 6);
            LaunchZero(F.p, (long)F.n, stream);
            hfar.alloc((size_t)kFarPartsMax * (size_t)kConvTileMax * (size_t)n3);
            auto far = [&](int parts) {
                if (historyMixed)
                    LaunchConvFarF(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const float*)d_mts32.p,
                                   (const float*)d_mgs32.p, (const unsigned short*)d_mtsLo.p,
                                   (const unsigned short*)d_mgsLo.p, F.p, nstore + 1, K0, Tn, maxn, (int)ne, hfar.p,
                                   parts, stream);
                else
                    LaunchConvFarF(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, (const double*)d_mts.p,
                                   (const double*)d_mgs.p, nullptr, nullptr, F.p, nstore + 1, K0, Tn, maxn, (int)ne,
                                   hfar.p, parts, stream);
            };
            double best = 1e30;
            for (int parts : {0, 2, 4, 8}) {
                double ms = TimeMin(stream, 3, [&] { far(parts); });
                if (ms < best) {
                    best = ms;
                    c.farParts = parts;
                }
            }
            c.farMs = best;
        }
        // GMRES reductions
        PrepareGmres(gm_restart > 0 ? gm_restart : 200);
        LaunchZero(d_V.p, std::min<long>((long)d_V.n, 5 * n3), stream);
        double best = 1e30;
        for (int nb : {8, 16, 32, 64, 128}) {
            int blocks = (int)std::min<long>(nb, std::max<long>(1, (n3 + 255) / 256));
            if (blocks != nb && nb != 8)
                continue;  // a larger count gives the same partition on a small case
            double ms = TimeMin(stream, 3, [&] {
                for (int r = 0; r < 10; ++r) {
                    LaunchMultiDot((const double*)d_V.p, n3, 4, (const double*)(d_V.p + 4 * n3), n3, d_part.p,
                                   blocks, d_h.p + 1, stream);
                    LaunchNorm((const double*)d_V.p, n3, d_part.p, blocks, d_h.p + 5, stream);
                }
            });
            if (ms < best) {
                best = ms;
                c.gmBlocks = nb;
            }
        }
        c.gmUs = best * 100.0;  // per (multi-dot + norm), microseconds
        if (!cachePath.empty()) {
            std::ofstream out(cachePath, std::ios::app);
            if (!out)
                throw std::runtime_error("cannot write the tuning cache " + cachePath);
            out << key << " : " << c.farParts << ' ' << c.gmBlocks << '\n';
        }
    }
    convPacked = c.farParts > 0;
    if (c.farParts > 0)
        farParts = c.farParts;
    gmBlocks = c.gmBlocks;
    CheckT(hipStreamSynchronize(stream), "autotune");
    return c;
}

}  // namespace tdbem
