// Batched load cases: B right-hand-side histories solved together over the
// same history bands, step-0 matrices and preconditioner.  Device vectors of
// a batch keep the load case fastest: element (r, b) of an n-vector batch is
// at r * B + b, so the 64 lanes of a wavefront read 64 load cases of one
// scalar with one coalesced access while matrix blocks are wave-uniform.
#pragma once

#include "kernels.hpp"

namespace tdbem {

constexpr int kBatchTile = 8;  // steps per far-part tile of the batched convolution
constexpr int kFarBWaves = 4;  // partial sums per row of the lag-by-lag batched far part (vector FMAs)
constexpr int kBatchFarPartsMax = 16;  // column parts of the matrix-core far part
constexpr int kBatchRedBlocks = 256;  // most reduction blocks (16 wavefronts each) of the batched dot products
constexpr int kFftTile = 16;           // steps per tile (and lags per sub-filter) of the spectral part's level 0
constexpr int kFftRows = 64;           // source rows per group of the spectral far part (8 per wavefront)
constexpr int kFftHPad = 256;          // values past the filter spectra that the products kernel may read
constexpr int kFftLagMin = 17;         // the spectral part takes the raw lags >= 17, the lag-by-lag kernels the rest

// Per-step plumbing.  bd / lu / lt are [b][K][3 ne] (host layout).
void LaunchKnownB(const NodeRefs* refs, const double* bd, int ne, int nstore, int K, int B, double* known,
                  hipStream_t s);
void LaunchStep0Out(const NodeRefs* refs, const ElemGeo* geo, const double* bd, int ne, int nstore, int B, double* lu,
                    double* lt, hipStream_t s);
// Step K's lu / lt from the solution x and the known values; with gU / gT
// also step K of the raw state series (as LaunchStateB makes it).
void LaunchScatterB(const NodeRefs* refs, const ElemGeo* geo, const double* bd, const double* x, int ne, int nstore,
                    int K, int B, double* lu, double* lt, hipStream_t s, double* gU = nullptr, double* gT = nullptr);
// raw state series g[((col * nstore + K) * 3 + c) * B + b]
void LaunchStateB(const NodeRefs* refs, const double* lu, const double* lt, int ne, int nstore, int K, int B,
                  double* gU, double* gT, hipStream_t s);
// 1-2-1 filtered state series for x in [x0, x1] with the solved states m in
// [1, K0 - 1]: F[x] = (g[x-1] + 2 g[x]) + g[x+1], F[((col * (nstore + 1) + x)
// * 3 + c) * B + b] (row nstore of each column is left zero).
void LaunchFilterB(const double* g, long cols, int nstore, int B, int x0, int x1, int K0, double* F, hipStream_t s);

// y = A x for B vectors (block CCSR, row-major 3x3 blocks).  The GMRES
// kernels below take an optional device counter `left` of the cases still
// iterating and return at once when it is 0.
void LaunchSpmmB(long nrows, const int* rp, const int* col, const double* val, const double* x, double* y, int B,
                 hipStream_t s, const int* left = nullptr);

// Far part of the tile [K0, K0+Tn) for B load cases into hfar[(i*3ne + r)*B + b]
// as partial sums (part p at hfar + p * kBatchTile * 3ne * B): on the FP64
// matrix cores with the columns of each row split into parts > 0 parts, or
// (parts = 0) with vector FMAs in kFarBWaves parts, one per wavefront of a
// row.
template <typename TB>
void LaunchConvFarB(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                    const TB* mts, const TB* mgs, const unsigned short* mtsLo, const unsigned short* mgsLo,
                    const double* FU, const double* FT, int nstore, int K0, int Tn, int maxn, int ne, int B,
                    double* hfar, int parts, hipStream_t s);
// Spectral part (batched_fft.hip): the raw lags >= kFftLagMin of every pair,
// split into levels l = 0 .. levels - 1 of sub-filter length L = 16 << l
// (level l < last: raw lags L + 1 .. 2L; last level: L + 1 .. maxn; sub-
// filter q of a level: raw lags L q + 1 .. L q + L).  Row groups: up to
// kFftRows source rows of one domain, near each other in space (grpDom; rows
// grpRowList[g * kFftRows + r], r < grpRows[g]).  Items of a level: per
// group, the (field, sub-filter q >= 1) pairs with the mask of rows whose
// band has that sub-filter in the level's range and the offset of their
// filter spectra, H[hoff + (slot * rows + k) * 30 + ch * 2 + ri] for the
// k-th row of the mask, slot < L (ch: the 9 values of -MTS, then the 6 of
// MGS).
struct FftItem {
    int f;                    // field element within the domain
    int q;                    // sub-filter (raw lags L q + 1 .. L q + L)
    int cu, ct;               // the field's displacement / traction history columns
    unsigned long long mask;  // rows of the group (bit r: row grpRowList[g * kFftRows + r])
    long hoff;                // offset of the item's filter spectra in H
};
static_assert(sizeof(FftItem) == 32, "FftItem is 32 bytes");
constexpr int kFftMaxLevels = 3;  // sub-filter lengths 16, 32, 64
// Count mode (items null): count[g] items and rowCount[g] rows summed over
// them; qHist[q] += items and qHist[qcap + 1 + q] += rows of sub-filter q.
void LaunchFftItems(PairGrid grid, const NodeRefs* refs, const int* kb, const int* len, int maxn, int L, int lagLo,
                    int lagHi, int ngroups, const int* grpDom, const int* grpRowList, const int* grpRows,
                    const long* itemOff, const long* hOff, long* count, long* rowCount, long* qHist, int qcap,
                    FftItem* items, hipStream_t s);
template <typename TB>
void LaunchFftFilters(PairGrid grid, const long* off, const int* kb, const int* len, const TB* mts, const TB* mgs,
                      const unsigned short* mtsLo, const unsigned short* mgsLo, int ngroups, const int* grpDom,
                      const int* grpRowList, const long* itemOff, long nitems, const FftItem* items, int maxn, int L,
                      int lagLo, int lagHi, double* H, hipStream_t s);
// Spectra of the 1-2-1 combined state window beta of a level (F_x for x =
// L beta - 2L + 1 .. L beta, from the raw states g[((col * nstore + m) * 3 +
// c) * B + b], zero outside 1 <= m < nstore) of every column:
// W[(((col * L + slot) * 3 + c) * 2 + ri) * B + b].
void LaunchFftWindows(const double* g, long cols, int nstore, int B, int L, int beta, double* W, hipStream_t s);
// Spectral products of a level's tile t (steps 2 + L t + i) summed per row
// and slot into Y[((row * L + slot) * 6 + c * 2 + ri) * B + b]; window beta
// at WU / WT + (beta % nwr) * ringStride, every sub-filter q <= nwr.  H holds
// kFftHPad values past its last item's spectra (read, never used).  mfma: on
// the FP64 matrix cores, else with vector FMAs; pace > 0: a workgroup's
// wavefronts meet at a barrier after every `pace` items of its list.  On the
// matrix cores without pacing the row groups run in the order order[0 ..
// ngroups - 1].
void LaunchConvFarFft(int ngroups, const int* order, const int* grpRowList, const int* grpRows, const long* itemOff,
                      const FftItem* items,
                      const double* H, const double* WU, const double* WT, long ringStrideU, long ringStrideT,
                      int nwr, int L, int t, int B, double* Y, bool mfma, int pace, hipStream_t s);
// Buffers of the inverse transform's layout whose values it adds.
struct FftAdds {
    const double* p[kFftMaxLevels - 1];
    int n;
};
// Samples L .. L + Tn - 1 of the inverse transforms of a level's Y into
// hfar[(i*3ne + r)*B + b], plus the same index of every add buffer.
void LaunchFftInverse(const double* Y, int ne, int B, int L, int Tn, double* hfar, FftAdds add, hipStream_t s);

// Near part of step K plus the tile's far part (the sum of its parts, in
// order) and, with hfar2, the spectral share of the 16-step tile starting at
// K02, into rhs.
template <typename TB>
void LaunchConvNearB(const NodeRefs* refs, const long* nearPtr, const long* nearPair, const int* nearField,
                     const long* off, const int* kb, const int* len, const TB* mts, const TB* mgs,
                     const unsigned short* mtsLo, const unsigned short* mgsLo, const double* gU, const double* gT,
                     int nstore, int K, int K0, int maxn, int ne, int B, const double* hfar, int parts,
                     double* rhs, hipStream_t s, const double* hfar2 = nullptr, int K02 = 0);

// Batched GMRES linear algebra (every reduction per load case, FP64).
void LaunchPrecondB(long n, const long* posLeaf, const long* leafBegin, const long* leafSize, const long* matOff,
                    const double* inv, const long* repid, const long* outBlock, const double* x, double* y, int B,
                    hipStream_t s);
// y = the preconditioner applied to x (as LaunchPrecondB) and, fused with it,
// h[k*B + b] = <V_k(:, b), y(:, b)> for k < nvD (V_k at V + k*3n*B; n: block
// rows); partial holds nblocks * nvD * B doubles.
void LaunchPrecondDotB(long n, const long* posLeaf, const long* leafBegin, const long* leafSize, const long* matOff,
                       const double* inv, const long* repid, const long* outBlock, const double* x, double* y, int B,
                       const double* V, int nvD, double* partial, int nblocks, double* h, hipStream_t s,
                       const int* left = nullptr);
// One pass over the rows of a batch: v = u (rhs - u when rhs is given), v -=
// sum_{k < nvA} hA[k] V_k, stored back into u when it changed; then per load
// case out[k] = <V_k, v> for k < nvD (addTo[k] += out[k] when given) and,
// with norm, normOut = ||v||.  V_k at V + k*n*B; partial holds nblocks *
// (nvD + 1) * B doubles.  Each sum is taken in a fixed order (blocks of rows,
// then blocks), so results do not depend on timing.
void LaunchPassB(const double* V, long n, int B, double* u, const double* rhs, const double* hA, int nvA, int nvD,
                 bool norm, double* partial, int nblocks, double* out, double* addTo, double* normOut, hipStream_t s,
                 const int* left = nullptr);
// h[k*B + b] = <V_k(:, b), u(:, b)> for k < nv
void LaunchMultiDotB(const double* V, long n, int nv, const double* u, int B, double* partial, int nblocks,
                     double* h, hipStream_t s);
void LaunchNormB(const double* x, long n, int B, double* partial, int nblocks, double* out, hipStream_t s);
// y(:, b) = x(:, b) / a[b] (a <= 1e-12 gives 0)
void LaunchScaleB(const double* x, const double* a, double* y, long n, int B, hipStream_t s);
void LaunchGivensB(double* h, double* QT, double* R, double* f, double* phibar, int* stag, const int* active, int j,
                   int restart, int B, hipStream_t s, const int* left = nullptr);
// The stopping test of iteration j on the device: every active case records
// its = j and leaves when ||r|| <= tol ||b|| (scal[B + b] and scal[b]) or,
// with flg 3, on stagnation; left receives the count of cases still active.
void LaunchCheckB(const double* scal, double tol, const int* stag, int* active, int* its, int* flg, int j, int B,
                  int* left, hipStream_t s);
// W / x update of column j for the active cases; with Vj, also Vj = u / hs
// (the scale of LaunchScaleB) for every case
void LaunchWUpdateB(const double* V, double* W, long n, const double* R, int j, int restart, const double* f,
                    double* x, const int* active, int B, hipStream_t s, const double* u = nullptr,
                    const double* hs = nullptr, double* Vj = nullptr, const int* left = nullptr);

}  // namespace tdbem
