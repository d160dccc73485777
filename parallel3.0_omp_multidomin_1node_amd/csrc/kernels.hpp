// Device data layout and launch wrappers of the HIP kernels.
#pragma once

#include "common.hpp"

namespace tdbem {

// Quadrature points of one field element, precomputed once: position,
// normal and weight*Jacobian of the 36x9 patch points and of the 4x4
// regular points (recomputing them per triple would give the same values).
struct ElemPoints {
    double pw_pos[kNPW2 * kNPWG2][3];
    double pw_nor[kNPW2 * kNPWG2][3];
    double pw_w[kNPW2 * kNPWG2];
    double rg_pos[kNG2][3];
    double rg_nor[kNG2][3];
    double rg_w[kNG2];
};

// Per-node variable references of the DOF map, flattened for the device.
struct NodeRefs {
    int u_block, t_block, hu_block, ht_block;
    double u_sign, t_sign, hu_sign, ht_sign;
    int u_known, t_known;
    int bcid, domain;
    int surface;   // 0 outer, 1 interface
    int ref_side;  // 1 unless this is the B side of an interface
    double itf_T[9];  // interface transform to the A frame (identity elsewhere)
};

// Pair grid: the pairs (source s, field f) of one domain d are numbered
// pairBase[d] + (s - eleBegin[d]) * n_d + (f - eleBegin[d]), so the pairs of a
// source row are contiguous.
struct PairGrid {
    int nd;
    long pairBase[kMaxDomains + 1];
    int eleBegin[kMaxDomains];
    int eleCount[kMaxDomains];
    const DynaMat* mats;  // per-domain materials of the case (device memory)
};

// Pair p of the grid -> source s, field f, domain d (pairs of a domain are
// numbered source-major).
__device__ __forceinline__ void DecodePairOf(const PairGrid& g, long p, int& s, int& f, int& d)
{
    d = 0;
    while (d + 1 < g.nd && p >= g.pairBase[d + 1])
        ++d;
    long local = p - g.pairBase[d];
    long nd = g.eleCount[d];
    s = g.eleBegin[d] + (int)(local / nd);
    f = g.eleBegin[d] + (int)(local % nd);
}

void LaunchPoints(const ElemGeo* geo, const QuadTables* q, ElemPoints* pts, int ne, hipStream_t s);
// Singular self tables; exact = one lane per element in the reference's
// summation order, else one wavefront per element (same sums, tree order).
void LaunchSingular(const ElemGeo* geo, const int* dom, const DynaMat* mats, const QuadTables* q, double* u1,
                    double* t1, double* t2, int ne, bool exact, hipStream_t s);

// Causality cull: band [kb, kb+len) of each pair and near-field flag.
// off[p] / coff[p] = exclusive prefix sums of len and of its 64-lag chunk
// counts over n pairs, with the totals at off[n] / coff[n]; work holds
// BandOffsetsWork(n) values.
void LaunchBandOffsets(const int* len, long n, long* work, long* off, long* coff, hipStream_t s);
long BandOffsetsWork(long n);
// Near pairs per source row: with ptr null, cnt[s] = their number; else they
// are written from pairOut[ptr[s]] (pair) and fieldOut[ptr[s]] (field element).
void LaunchNearLists(PairGrid grid, const NodeRefs* refs, const int* kb, const int* len, int tile, int ne,
                     const long* ptr, int* cnt, long* pairOut, int* fieldOut, hipStream_t s);
void LaunchCull(const ElemGeo* geo, const int* dom, PairGrid grid, long npairs, int maxn, int* kb, int* len,
                unsigned char* nearFlag, hipStream_t s);

// History quadrature for all band entries [e0, e1); entry e lands at place
// e - ebase of mts / mgs / alive (a slab of the bands when ebase > 0).
// Entries of lags > maxLag are skipped (left as they were).
void LaunchHistory(const ElemGeo* geo, const ElemPoints* pts, const int* dom, const NodeRefs* refs,
                   const double* selfT2, PairGrid grid, const long* off, const int* kb, long npairs, long e0, long e1,
                   double* mts, double* mgs, unsigned char* alive, hipStream_t s, long ebase = 0,
                   int maxLag = 1 << 30);
// Copies the blocks of lags <= lagMax of the pairs [p0, p1) from a slab of the
// bands (entry e at place e - ebase) to a compact array (pair p's lags from
// kb[p] at nearOff[p]).
void LaunchNearGather(const long* off, const int* kb, const long* nearOff, long p0, long p1, long ebase,
                      const double* mts, const double* mgs, const unsigned char* alive, double* nmts, double* nmgs,
                      unsigned char* nalive, hipStream_t s);

// Staged history quadrature in two kernels that together write every band
// entry once (entry e at place e - ebase of the outputs; the chunks / waves
// whose lags all exceed maxLag are skipped, the rest computed whole).  Regular entries:
// one wavefront per (pair, 64-lag chunk), chunk offsets coff[p] (npairs+1
// entries), chunks [c0, c1).  Patch-wise entries
// (a judge picking the 36-patch path): waves [w0, w1) of PwWave, each packing
// the patch-wise lags of up to kPwSlots pairs with the same field element.
constexpr int kPwSlots = 4;
struct PwWave {
    long pair[kPwSlots];
    int rank0[kPwSlots];  // first patch-wise lag (rank within the pair) of each slot
    int cnt[kPwSlots];    // lanes of each slot
    int nseg;
    int pad;
};
// cls (nullable): frozen judges of every band entry (f2 | f1 << 2) used
// instead of judging with the current materials.
// Material-derivative pass of the history blocks contracted with the
// per-entry correlation C (15 per entry): part[4 w + p] (and the domain in
// part[4 w + 3]), w over the regular chunks then the patch-wise waves; J[9 d + 3 q + p] = d q / d (E, v, rho)_p
// with q = (C1, C2, rho); dT2[27 f + 9 p + i] the self table derivatives.
void LaunchHistoryGrad(const ElemGeo* geo, const ElemPoints* pts, const NodeRefs* refs, PairGrid grid,
                       const long* coff, const long* off, const int* kb, const int* len, const unsigned char* cls,
                       long npairs, long chunks, const PwWave* waves, long nwaves, const unsigned char* alive,
                       const double* C, const double* J, const double* dT2, double* part, hipStream_t s);
// sums[3 d + p] += the partials of the waves of domain d.
void LaunchDomainSums(const double* part, long nw, int nd, double* sums, hipStream_t s);
// PwWave packing per field element: count pass (waves null) -> nw[f], write
// pass -> waves[woff[f] ..].
// (waves of the sources in [rowBegin, rowEnd) only; rowEnd < 0: every source)
void LaunchPwPack(PairGrid grid, const NodeRefs* refs, const int* cnt, int ne, const long* woff, long* nw,
                  PwWave* waves, hipStream_t s, long rowBegin = 0, long rowEnd = -1);
void LaunchPwCount(const ElemGeo* geo, PairGrid grid, const int* kb, const int* len, const long* off,
                   const unsigned char* cls, long npairs, int* cnt, hipStream_t s);
void LaunchClassify(const ElemGeo* geo, PairGrid grid, const int* kb, const int* len, const long* off, long npairs,
                    unsigned char* cls, hipStream_t s);
template <typename TB>  // TB: stored block type (double, or float in the mixed-precision mode)
void LaunchHistoryReg(const ElemGeo* geo, const ElemPoints* pts, const NodeRefs* refs, const double* selfT2,
                      PairGrid grid, const long* coff, const long* off, const int* kb, const int* len,
                      const unsigned char* cls, long npairs, long c0, long c1, TB* mts, TB* mgs,
                      unsigned short* mtsLo, unsigned short* mgsLo, unsigned char* alive, hipStream_t s,
                      long ebase = 0, int maxLag = 1 << 30, unsigned long long* qcount = nullptr);
template <typename TB>
void LaunchHistoryPw(const ElemGeo* geo, const ElemPoints* pts, const NodeRefs* refs, const double* selfT2,
                     PairGrid grid, const PwWave* waves, long w0, long w1, const long* off, const int* kb,
                     const int* len, const unsigned char* cls, TB* mts, TB* mgs, unsigned short* mtsLo,
                     unsigned short* mgsLo, unsigned char* alive, hipStream_t s, long ebase = 0,
                     int maxLag = 1 << 30, unsigned long long* qcount = nullptr);
// qcount (nullable, 4 counters): evaluations of the two kernels, [0] patch-wise
// T2 (entry, point), [1] patch-wise T1 + U (entry, point), [2] regular
// (entry, point), [3] staged (source, point) records.

// Step-0 blocks of the near pairs listed in pairIdx (column-major, ready for
// AddBlock): out[2*i] and out[2*i+1] with flags.
// cls0 (nullable): frozen step-0 judges (lag 0 | lag 1 << 2) of the listed pairs.
void LaunchStep0(const ElemGeo* geo, const ElemPoints* pts, const int* dom, const NodeRefs* refs,
                 const double* selfU1, const double* selfT1, const double* selfT2, PairGrid grid,
                 const long* pairIdx, long count, const unsigned char* cls0, double* out, int* flags,
                 hipStream_t s);
void LaunchClassify0(const ElemGeo* geo, PairGrid grid, const long* pairIdx, long count, unsigned char* cls0,
                     hipStream_t s);

// History convolution for step K into rhs (rhs += hist).
void LaunchConvolution(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                       const double* mts, const double* mgs, const double* gU, const double* gT, int nstore, int K,
                       int maxn, int ne, double* rhs, hipStream_t s);

// Time-blocked convolution: far part of the tile [K0, K0+Tn) (tile 8 or 16
// steps) into hfar[i][3*ne] (i < Tn), then per step the near part plus hfar
// into rhs.
constexpr int kConvTileMax = 16;
// mtsLo / mgsLo (nullable): bf16 residuals of the split FP32 storage.
// Filtered state series of the tile starting at K0 for x in [x0, x1]
// (F[f][x][6], x = 0 .. nstore).
void LaunchFilterTile(const NodeRefs* refs, const double* gU, const double* gT, int nstore, int ne, int K0, int x0,
                      int x1, double* F, hipStream_t s);
// Far part of a 16-step tile from the filtered series, next work item
// prefetched.  parts > 0: 64 band entries per item across pair boundaries,
// parts wavefronts per row, one partial sum each (hfar holds parts blocks of
// kConvTileMax * 3 * ne values, at most kFarPartsMax); parts = 0: 64 lags of
// one pair per item and one sum per row.  With parts > 0 the rows may be
// limited to [rowBegin, rowBegin + rowCount) and the bands to a slab whose
// first entry is ebase (pair p's lags at off[p] - ebase).
constexpr int kFarPartsMax = 8;
template <typename TB>
void LaunchConvFarF(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                    const TB* mts, const TB* mgs, const unsigned short* mtsLo, const unsigned short* mgsLo,
                    const double* F, int fstride, int K0, int Tn, int maxn, int ne, double* hfar, int parts,
                    hipStream_t s, int rowBegin = 0, int rowCount = -1, long ebase = 0);
template <typename TB>
void LaunchConvFar(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len, const TB* mts,
                   const TB* mgs, const unsigned short* mtsLo, const unsigned short* mgsLo, const double* gU,
                   const double* gT, int nstore, int K0, int Tn, int tile, int maxn, int ne, double* hfar,
                   hipStream_t s);
template <typename TB>
void LaunchConvNear(const NodeRefs* refs, const long* nearPtr, const long* nearPair, const int* nearField,
                    const long* off, const int* kb, const int* len, const TB* mts, const TB* mgs,
                    const unsigned short* mtsLo, const unsigned short* mgsLo, const double* gU, const double* gT,
                    int nstore, int K, int K0, int maxn, int ne, const double* hfar, int parts, double* rhs,
                    bool perEntry, hipStream_t s);

// GMRES linear algebra.  TA: matrix / inverse storage, TX / TY: vector
// storage (double, or float for the inner solves of the mixed mode); every
// product and sum is formed in FP64.
// Block CCSR SpMV y = A x (row-major 3x3 blocks).
template <typename TA, typename TX, typename TY>
void LaunchSpmv(long nrows, const int* rp, const int* col, const TA* val, const TX* x, TY* y, hipStream_t s);
// y = M^{-1} x with the leaf preconditioner.
template <typename TA, typename TX, typename TY>
void LaunchPrecond(long n, const long* posLeaf, const long* leafBegin, const long* leafSize, const long* matOff,
                   const TA* inv, const long* repid, const long* outBlock, const TX* x, TY* y, hipStream_t s);
// Reductions: partial sums of dot products of u with columns 0..nv-1 of V
// (column stride ldv), finished into h[0..nv).
template <typename TV, typename TU>
void LaunchMultiDot(const TV* V, long ldv, int nv, const TU* u, long n, double* partial, int nblocks, double* h,
                    hipStream_t s);
// u -= V[:, 0:nv] h ; accumulate htot += h when htot != null
template <typename TV, typename TU>
void LaunchMultiAxpy(const TV* V, long ldv, int nv, const double* h, TU* u, long n, double* htot, hipStream_t s);
template <typename T>
void LaunchNorm(const T* x, long n, double* partial, int nblocks, double* out, hipStream_t s);
// y = a*x   and   y = b - y   and   x += y
template <typename TX, typename TY>
void LaunchScaleCopy(const TX* x, const double* a, int invert, TY* y, long n, hipStream_t s);
void LaunchResidual(const double* b, double* y, long n, hipStream_t s);
void LaunchAdd(const double* y, double* x, long n, hipStream_t s);
void LaunchToFloat(const double* x, float* y, long n, hipStream_t s);
// GMRES Givens step j (1-based) on device scalars.
void LaunchGivens(double* h, double* QT, double* R, double* f, double* phibar, int* stag, int j, int restart,
                  hipStream_t s);
// W_j and x += f_j W_j.
template <typename TV>
void LaunchWUpdate(const TV* V, TV* W, long n, const double* R, int j, int restart, const double* f, double* x,
                   hipStream_t s);

// Per-step state plumbing.
void LaunchKnown(const NodeRefs* refs, const double* bd, int ne, double* known, long nknown, hipStream_t s);
void LaunchScatter(const NodeRefs* refs, const ElemGeo* geo, const double* bd, const double* x, int ne, double* lu,
                   double* lt, hipStream_t s);
void LaunchState(const NodeRefs* refs, const double* lu, const double* lt, int ne, int K, int nstore, double* gU,
                 double* gT, hipStream_t s);
void LaunchZero(double* x, long n, hipStream_t s);

}  // namespace tdbem
