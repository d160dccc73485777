// Frequency-domain part of the batched history convolution (gfx950).
//
// The convolution of step k sums, for every pair (s, f), the history blocks
// times the reference's 1-2-1 combination of states
// (DBEM1/multidomain.cpp:3201-3225,3849-3866).  The raw lags j <= 16 stay
// with the lag-by-lag kernels (8-step tiles, batched.hip).  For the raw lags
// j >= 17 the 1-2-1 combination folds into the blocks: with
//   B'_d = B_{d-1} + 2 B_d + B_{d+1}   (B_j = 0 for j < 17, outside the band
//                                        or past maxn; so B'_d = 0 for d < 16)
// their share is one linear convolution in time with the raw states g_m,
// sum_{d >= 16} B'_d g_{k-d}, and every state it needs was solved before the
// 16-step tile of k (k - d <= K0 - 1).  Cut into 16-lag sub-filters q >= 1
// (lags 16q .. 16q+15) and 16-state blocks beta (states 16 beta - 14 ..
// 16 beta + 1), it is the uniformly partitioned overlap-save convolution:
// with 32-point spectra H_q of the zero-padded sub-filters and W_beta of the
// state windows (blocks beta - 1, beta), tile t (steps 2 + 16 t + i, i < 16)
// gets samples 16 .. 31 of the inverse transform of Y = sum_q H_q W_{t+1-q}.
// Per (pair, sub-filter, tile) that is 17 spectral bins x 18 complex products
// instead of 16 lags x 16 steps x 18 block products: about a third of the
// multiply-adds on the workloads' bands, in FP64 (the transforms round at
// 1e-16).  The 17 real-input bins are held as 16 complex slots: slot 0 packs
// the real bins 0 and 16 as (re, im).
//
// Kernels: the per row-group list of (field, sub-filter) items and the
// sub-filter spectra of every item (once per solve); the state-window
// spectra of every history column and load case (once per tile); the
// spectral products, where a workgroup of eight wavefronts takes one slot,
// 64 source rows of one domain (eight per wavefront) and 64 load cases, so
// a window spectrum fetched for an item serves every row of the group; and
// the inverse transforms of the summed spectra (once per tile).
#include "batched.hpp"

namespace tdbem {

namespace {

// cos / sin of 2 pi j / 32, j = 0 .. 31 (first-quadrant values, signs by symmetry)
#define TDBEM_C1 0.98078528040323044913
#define TDBEM_C2 0.92387953251128675613
#define TDBEM_C3 0.83146961230254523708
#define TDBEM_C4 0.70710678118654752440
#define TDBEM_C5 0.55557023301960222474
#define TDBEM_C6 0.38268343236508977173
#define TDBEM_C7 0.19509032201612826785
__constant__ double c_cos32[32] = {
    1.0,       TDBEM_C1,  TDBEM_C2,  TDBEM_C3,  TDBEM_C4,  TDBEM_C5,  TDBEM_C6,  TDBEM_C7,
    0.0,       -TDBEM_C7, -TDBEM_C6, -TDBEM_C5, -TDBEM_C4, -TDBEM_C3, -TDBEM_C2, -TDBEM_C1,
    -1.0,      -TDBEM_C1, -TDBEM_C2, -TDBEM_C3, -TDBEM_C4, -TDBEM_C5, -TDBEM_C6, -TDBEM_C7,
    0.0,       TDBEM_C7,  TDBEM_C6,  TDBEM_C5,  TDBEM_C4,  TDBEM_C3,  TDBEM_C2,  TDBEM_C1};
__constant__ double c_sin32[32] = {
    0.0,       TDBEM_C7,  TDBEM_C6,  TDBEM_C5,  TDBEM_C4,  TDBEM_C3,  TDBEM_C2,  TDBEM_C1,
    1.0,       TDBEM_C1,  TDBEM_C2,  TDBEM_C3,  TDBEM_C4,  TDBEM_C5,  TDBEM_C6,  TDBEM_C7,
    0.0,       -TDBEM_C7, -TDBEM_C6, -TDBEM_C5, -TDBEM_C4, -TDBEM_C3, -TDBEM_C2, -TDBEM_C1,
    -1.0,      -TDBEM_C1, -TDBEM_C2, -TDBEM_C3, -TDBEM_C4, -TDBEM_C5, -TDBEM_C6, -TDBEM_C7};
#undef TDBEM_C1
#undef TDBEM_C2
#undef TDBEM_C3
#undef TDBEM_C4
#undef TDBEM_C5
#undef TDBEM_C6
#undef TDBEM_C7

__device__ __forceinline__ double Bf16LoF(unsigned short lo) { return (double)__uint_as_float(((unsigned int)lo) << 16); }

__device__ __forceinline__ int WaveIdx() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ void WaveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sub-filter range [qlo, qhi] of a pair's band [k0, k0 + L) restricted to the
// raw lags j >= kFftLagMin (the spectral part's share): the filtered blocks
// are nonzero on lags jlo - 1 .. jhi + 1 with jlo = max(k0, kFftLagMin),
// jhi = min(k0 + L - 1, maxn); every such lag is >= 16, so q >= 1.
__device__ __forceinline__ bool QRange(int k0, int L, int maxn, int& qlo, int& qhi)
{
    const int jlo = k0 > kFftLagMin ? k0 : kFftLagMin;
    const int jhi = k0 + L - 1 < maxn ? k0 + L - 1 : maxn;
    if (L <= 0 || jlo > jhi)
        return false;
    qlo = (jlo - 1) / kFftTile;
    qhi = (jhi + 1) / kFftTile;
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Items: one wavefront per row group; lanes over the domain's field elements.
// With items null, count[g] = number of (field, sub-filter) items of group g
// and rowCount[g] the sum of their row counts; else they are written from
// items[itemOff[g]] in field order, sub-filters ascending, each with the
// offset of its filter spectra (16 slots x its rows x 30 values) from
// hOff[g] on.  Count mode also adds the items and rows of each sub-filter q
// to qHist[q] and qHist[qcap + 1 + q] (q <= qcap).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_fft_items(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                  const int* __restrict__ kb, const int* __restrict__ len, int maxn,
                                                  const int* __restrict__ grpDom, const int* __restrict__ grpRowList,
                                                  const int* __restrict__ grpRows, const long* __restrict__ itemOff,
                                                  const long* __restrict__ hOff, long* __restrict__ count,
                                                  long* __restrict__ rowCount, long* __restrict__ qHist, int qcap,
                                                  FftItem* __restrict__ items)
{
    __shared__ long pre[65], prr[65];
    const int g = blockIdx.x, lane = threadIdx.x;
    const int d = grpDom[g], nr = grpRows[g];
    const int* rows = grpRowList + (long)g * kFftRows;
    const int eb = grid.eleBegin[d], n = grid.eleCount[d];
    const long domPair0 = grid.pairBase[d];
    const long base = items ? itemOff[g] : 0, hbase = items ? hOff[g] : 0;
    long total = 0, rtotal = 0;
    for (int f0 = 0; f0 < n; f0 += 64) {
        const int f = f0 + lane;
        int lo = 1 << 30, hi = -1;
        if (f < n)
            for (int r = 0; r < nr; ++r) {
                const long p = domPair0 + (long)(rows[r] - eb) * n + f;
                int a, z;
                if (QRange(kb[p], len[p], maxn, a, z)) {
                    lo = a < lo ? a : lo;
                    hi = z > hi ? z : hi;
                }
            }
        // the mask of rows whose band has sub-filter q
        auto maskOf = [&](int q) {
            unsigned long long m = 0;
            for (int r = 0; r < nr; ++r) {
                const long p = domPair0 + (long)(rows[r] - eb) * n + f;
                int a, z;
                if (QRange(kb[p], len[p], maxn, a, z) && q >= a && q <= z)
                    m |= 1ull << r;
            }
            return m;
        };
        long cnt = 0, rows = 0;  // sub-filters that some row has (the rows' ranges may leave gaps)
        for (int q = lo; q <= hi; ++q) {
            const int pc = __popcll(maskOf(q));
            cnt += pc > 0;
            rows += pc;
            if (!items && pc && q <= qcap) {
                atomicAdd(reinterpret_cast<unsigned long long*>(qHist + q), 1ull);
                atomicAdd(reinterpret_cast<unsigned long long*>(qHist + qcap + 1 + q), (unsigned long long)pc);
            }
        }
        pre[lane + 1] = cnt;
        prr[lane + 1] = rows;
        __syncthreads();
        if (lane == 0) {
            pre[0] = prr[0] = 0;
            for (int i = 1; i <= 64; ++i) {
                pre[i] += pre[i - 1];
                prr[i] += prr[i - 1];
            }
        }
        __syncthreads();
        if (items && cnt) {
            long at = base + total + pre[lane];
            long ho = hbase + (rtotal + prr[lane]) * 16 * 30;
            const int cu = (int)refs[eb + f].hu_block, ct = (int)refs[eb + f].ht_block;
            for (int q = lo; q <= hi; ++q) {
                const unsigned long long m = maskOf(q);
                if (!m)
                    continue;
                FftItem it;
                it.f = f;
                it.q = q;
                it.cu = cu;
                it.ct = ct;
                it.mask = m;
                it.hoff = ho;
                items[at++] = it;
                ho += (long)__popcll(m) * 16 * 30;
            }
        }
        total += pre[64];
        rtotal += prr[64];
        __syncthreads();
    }
    if (!items && lane == 0) {
        count[g] = total;
        rowCount[g] = rtotal;
    }
}

void LaunchFftItems(PairGrid grid, const NodeRefs* refs, const int* kb, const int* len, int maxn, int ngroups,
                    const int* grpDom, const int* grpRowList, const int* grpRows, const long* itemOff, const long* hOff,
                    long* count, long* rowCount, long* qHist, int qcap, FftItem* items, hipStream_t s)
{
    k_fft_items<<<(unsigned)ngroups, 64, 0, s>>>(grid, refs, kb, len, maxn, grpDom, grpRowList, grpRows, itemOff, hOff,
                                                 count, rowCount, qHist, qcap, items);
}

// ---------------------------------------------------------------------------
// Window spectra: one thread per (column, component, load case).  Window
// beta holds the raw states m = 16 beta - 30 + n, n = 0 .. 31 (m >= 1 and
// m < nstore; with `partial` the second half, the tile's own states, zero).
// W[(((col * 16 + slot) * 3 + c) * 2 + ri) * B + b].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fft_windows(const double* __restrict__ g, long cols, int nstore, int B,
                                                     int beta, int partial, double* __restrict__ W)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= cols * 3 * B)
        return;
    const long col = idx / (3 * B);
    const int rem = (int)(idx - col * 3 * B);
    const int c = rem / B, b = rem - c * B;
    double x[32];
#pragma unroll
    for (int n = 0; n < 32; ++n) {
        const int m = 16 * beta - 30 + n;
        const bool ok = m >= 1 && m < nstore && (!partial || n < 16);
        x[n] = ok ? g[((col * nstore + m) * 3 + c) * B + b] : 0.0;
    }
    double* out = W + (col * 16 * 3 + c) * 2 * (long)B + b;
    {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int n = 0; n < 32; ++n) {
            re += x[n];
            im += (n & 1) ? -x[n] : x[n];
        }
        out[0] = re;
        out[(long)B] = im;
    }
    for (int s = 1; s < 16; ++s) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int n = 0; n < 32; ++n) {
            const int j = (s * n) & 31;
            re = fma(x[n], c_cos32[j], re);
            im = fma(-x[n], c_sin32[j], im);
        }
        double* o = out + (long)s * 3 * 2 * B;
        o[0] = re;
        o[(long)B] = im;
    }
}

void LaunchFftWindows(const double* g, long cols, int nstore, int B, int beta, bool partial, double* W, hipStream_t s)
{
    long n = cols * 3 * B;
    if (n <= 0)
        return;
    k_fft_windows<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, cols, nstore, B, beta, partial ? 1 : 0, W);
}

// ---------------------------------------------------------------------------
// Sub-filter spectra of every item: one wavefront per item (its row group
// found in itemOff); the item's rows four at a time, their band slices (raw
// lags 16q - 1 .. 16q + 16 that are >= kFftLagMin, -M and G, 15 values per
// lag) staged in the wavefront's LDS, filtered (B'_d = B_{d-1} + 2 B_d +
// B_{d+1}) and transformed.  H[hoff + (slot * rows + k) * 30 + ch * 2 + ri]
// for the item's k-th row.
// ---------------------------------------------------------------------------
template <typename TB>
__global__ void __launch_bounds__(256) k_fft_filters(
    PairGrid grid, const long* __restrict__ off, const int* __restrict__ kb, const int* __restrict__ len,
    const TB* __restrict__ mts, const TB* __restrict__ mgs, const unsigned short* __restrict__ mtsLo,
    const unsigned short* __restrict__ mgsLo, const int* __restrict__ grpDom, const int* __restrict__ grpRowList,
    int ngroups, const long* __restrict__ itemOff, const FftItem* __restrict__ items, int maxn,
    double* __restrict__ H)
{
    constexpr int NS = 18 * 15;
    __shared__ double lds[4][4 * NS];
    __shared__ int rowOf[4][64];
    const int wave = WaveIdx(), lane = threadIdx.x & 63;
    const long it = (long)blockIdx.x * 4 + wave;
    if (it >= itemOff[ngroups])
        return;  // no workgroup barriers follow
    int glo = 0, ghi = ngroups - 1;  // the group g with itemOff[g] <= it < itemOff[g + 1]
    while (glo < ghi) {
        const int mid = (glo + ghi + 1) >> 1;
        if (itemOff[mid] <= it)
            glo = mid;
        else
            ghi = mid - 1;
    }
    const int g = glo;
    double* S = lds[wave];
    int* rw = rowOf[wave];
    const int d = grpDom[g];
    const int* rows = grpRowList + (long)g * kFftRows;
    const int eb = grid.eleBegin[d], n = grid.eleCount[d];
    const long domPair0 = grid.pairBase[d];
    {
        const FftItem item = items[it];
        const int f = item.f, q = item.q;
        const unsigned long long mask = item.mask;
        const int pc = __popcll(mask);
        if ((mask >> lane) & 1)
            rw[__popcll(mask & ((1ull << lane) - 1))] = lane;
        WaveSync();
        for (int k0 = 0; k0 < pc; k0 += 4) {
            for (int idx = lane; idx < 4 * NS; idx += 64) {
                const int kk = idx / NS, rem = idx - kk * NS;
                const int i = rem / 15, v = rem - i * 15;
                const bool row = k0 + kk < pc;
                const long p = domPair0 + (long)(rows[row ? rw[k0 + kk] : 0] - eb) * n + f;
                const int k0b = kb[p], L = len[p];
                const int j = kFftTile * q - 1 + i;
                const bool ok = row && j >= kFftLagMin && j <= maxn && j >= k0b && j < k0b + L;
                const long e = ok ? off[p] + (j - k0b) : 0;
                double val;
                if (v < 9) {
                    val = (double)mts[9 * e + v];
                    if (mtsLo)
                        val += Bf16LoF(mtsLo[9 * e + v]);
                    val = -val;
                } else {
                    val = (double)mgs[6 * e + (v - 9)];
                    if (mgsLo)
                        val += Bf16LoF(mgsLo[6 * e + (v - 9)]);
                }
                S[idx] = ok ? val : 0.0;
            }
            WaveSync();
            const int kk = lane / 15, ch = lane - kk * 15;
            if (lane < 60 && k0 + kk < pc) {
                const double* sr = S + kk * NS + ch;
                double sv[18];
#pragma unroll
                for (int i = 0; i < 18; ++i)
                    sv[i] = sr[15 * i];
                double hf[16];
#pragma unroll
                for (int m = 0; m < 16; ++m)
                    hf[m] = (sv[m] + 2.0 * sv[m + 1]) + sv[m + 2];
                double* out = H + item.hoff + (long)(k0 + kk) * 30 + ch * 2;
                {
                    double re = 0.0, im = 0.0;
#pragma unroll
                    for (int m = 0; m < 16; ++m) {
                        re += hf[m];
                        im += (m & 1) ? -hf[m] : hf[m];
                    }
                    out[0] = re;
                    out[1] = im;
                }
                for (int slot = 1; slot < 16; ++slot) {
                    double re = 0.0, im = 0.0;
#pragma unroll
                    for (int m = 0; m < 16; ++m) {
                        const int jj = (slot * m) & 31;
                        re = fma(hf[m], c_cos32[jj], re);
                        im = fma(-hf[m], c_sin32[jj], im);
                    }
                    double* o = out + (long)slot * pc * 30;
                    o[0] = re;
                    o[1] = im;
                }
            }
            WaveSync();
        }
    }
}

template <typename TB>
void LaunchFftFilters(PairGrid grid, const long* off, const int* kb, const int* len, const TB* mts, const TB* mgs,
                      const unsigned short* mtsLo, const unsigned short* mgsLo, int ngroups, const int* grpDom,
                      const int* grpRowList, const long* itemOff, long nitems, const FftItem* items, int maxn,
                      double* H, hipStream_t s)
{
    if (nitems <= 0)
        return;
    k_fft_filters<TB><<<(unsigned)((nitems + 3) / 4), 256, 0, s>>>(grid, off, kb, len, mts, mgs, mtsLo, mgsLo, grpDom,
                                                                   grpRowList, ngroups, itemOff, items, maxn, H);
}

// ---------------------------------------------------------------------------
// Spectral products: one workgroup of eight wavefronts per (row group, slot,
// 64 load cases); wavefront w owns rows 8w .. 8w + 7 of the group and skips
// the items that none of them has.  Items are loaded two ahead into a ring of
// three register sets (the window spectra of the lane's load case, column f,
// window t + 1 - q and this slot; the filter spectra of the wavefront's rows,
// at most 240 values, four per lane, read past the item's end into H's
// padding on the lanes that hold none); an item's filter spectra are written
// to the wavefront's LDS slice before its products, which read them at
// wave-uniform addresses.  The summed spectra go to
// Y[((row * 16 + slot) * 6 + c * 2 + ri) * B + b].  Slot 0 packs the real
// bins 0 and 16, so its products are two real ones per value.
// ---------------------------------------------------------------------------
struct FftRegs {
    double xu[3][2], xt[3][2];
    double2 h0, h1;
};

template <bool PACKED>
__device__ __forceinline__ void FftProducts(const FftItem* __restrict__ items, long i0, long i1, int t, int tRing,
                                            int nwr, int pace, int sh, int slot, int lane, int B, int bb,
                                            const double* __restrict__ H, const double* __restrict__ WU,
                                            const double* __restrict__ WT, long ringStrideU, long ringStrideT,
                                            double2* hs2, double (&acc)[8][3][2])
{
    const double* hs = reinterpret_cast<const double*>(hs2);
    // next item at or after `it` that has one of this wavefront's rows and
    // whose window exists in this tile (q <= t + 1)
    auto nextItem = [&](long it) {
        while (it < i1 && (items[it].q > t + 1 || ((items[it].mask >> sh) & 0xffull) == 0))
            ++it;
        return it;
    };
    auto load = [&](long it, FftRegs& x) {
        const FftItem item = items[it];
        // ring slot of window t + 1 - q: (t + 1 - q) mod nwr with 1 <= q <= nwr
        const int beta = tRing - item.q + (tRing < item.q ? nwr : 0);
        const double* wu = WU + beta * ringStrideU + ((long)item.cu * 96 + slot * 6) * B + bb;
        const double* wt = WT + beta * ringStrideT + ((long)item.ct * 96 + slot * 6) * B + bb;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int ri = 0; ri < 2; ++ri) {
                x.xu[c][ri] = wu[(long)(c * 2 + ri) * B];
                x.xt[c][ri] = wt[(long)(c * 2 + ri) * B];
            }
        const unsigned long long m = item.mask;
        const int pc = __popcll(m), below = __popcll(m & ((1ull << sh) - 1));
        const int n2 = 15 * __popcll((m >> sh) & 0xffull);  // double2 values of the wavefront's rows
        const double2* hp = reinterpret_cast<const double2*>(H + item.hoff + ((long)slot * pc + below) * 30);
        // lanes past the rows keep what they held: the products never read those values
        if (lane < n2)
            x.h0 = hp[lane];
        if (lane + 64 < n2)
            x.h1 = hp[lane + 64];
    };
    // the item's filter spectra into the wavefront's LDS slice; the wave's
    // LDS accesses run in order, so the previous item's products have read
    // theirs, and these writes land before the next products read them (the
    // wavefront-scope fences keep the compiler from moving accesses across)
    auto stage = [&](const FftRegs& x) {
        WaveSync();
        hs2[lane] = x.h0;
        hs2[lane + 64] = x.h1;
        WaveSync();
    };
    auto mac = [&](long it, const FftRegs& x) {
        const unsigned sub = (unsigned)((items[it].mask >> sh) & 0xffull);
        const int gi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (!((sub >> r) & 1u))
                continue;
            const double* h = hs + __popc(sub & ((1u << r) - 1u)) * 30;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double re = acc[r][c][0], im = acc[r][c][1];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const double mr = h[2 * (3 * c + cc)], mi = h[2 * (3 * c + cc) + 1];
                    const double gr = h[18 + 2 * gi[c][cc]], gim = h[18 + 2 * gi[c][cc] + 1];
                    const double ur = x.xu[cc][0], ui = x.xu[cc][1];
                    const double tr = x.xt[cc][0], ti = x.xt[cc][1];
                    if (PACKED) {
                        re = fma(mr, ur, re);
                        im = fma(mi, ui, im);
                        re = fma(gr, tr, re);
                        im = fma(gim, ti, im);
                    } else {
                        re = fma(mr, ur, re);
                        re = fma(-mi, ui, re);
                        im = fma(mr, ui, im);
                        im = fma(mi, ur, im);
                        re = fma(gr, tr, re);
                        re = fma(-gim, ti, re);
                        im = fma(gr, ti, im);
                        im = fma(gim, tr, im);
                    }
                }
                acc[r][c][0] = re;
                acc[r][c][1] = im;
            }
        }
    };
    // The wavefronts of the workgroup share each item's window spectra
    // through L2: a barrier after every `pace` items of the list keeps them
    // within that many items of each other.  Every wavefront passes the same
    // nPace barriers (those of the chunks it skips on its way, and the rest at
    // the end); they order nothing in memory, so no fence goes with them.
    const int nPace = pace > 0 ? (int)((i1 - i0 + pace - 1) / pace) : 0;
    int paced = 0;
    auto paceTo = [&](long it) {
        const int chunk = pace > 0 ? (int)((it - i0) / pace) : 0;
        while (paced < chunk) {
            __builtin_amdgcn_s_barrier();
            ++paced;
        }
    };
    long cur = nextItem(i0);
    if (cur < i1) {
        FftRegs x0, x1, x2;
        long n1 = nextItem(cur + 1), n2;
        load(cur, x0);
        load(n1 < i1 ? n1 : cur, x1);
        stage(x0);
        // x0 / x1 / x2 rotate: the current item's set is staged, the next
        // one's loads are in flight and the one after is loaded before the
        // products; past the list's end the current item is loaded again
        for (;;) {
            n2 = n1 < i1 ? nextItem(n1 + 1) : i1;
            load(n2 < i1 ? n2 : cur, x2);
            paceTo(cur);
            mac(cur, x0);
            if (n1 >= i1)
                break;
            stage(x1);
            cur = n1;
            n1 = n2;
            n2 = n1 < i1 ? nextItem(n1 + 1) : i1;
            load(n2 < i1 ? n2 : cur, x0);
            paceTo(cur);
            mac(cur, x1);
            if (n1 >= i1)
                break;
            stage(x2);
            cur = n1;
            n1 = n2;
            n2 = n1 < i1 ? nextItem(n1 + 1) : i1;
            load(n2 < i1 ? n2 : cur, x1);
            paceTo(cur);
            mac(cur, x2);
            if (n1 >= i1)
                break;
            stage(x0);
            cur = n1;
            n1 = n2;
        }
    }
    while (paced < nPace) {
        __builtin_amdgcn_s_barrier();
        ++paced;
    }
}

__global__ void __launch_bounds__(512) k_conv_far_fft(
    const int* __restrict__ grpRowList, const int* __restrict__ grpRows, const long* __restrict__ itemOff,
    const FftItem* __restrict__ items, const double* __restrict__ H, const double* __restrict__ WU,
    const double* __restrict__ WT, long ringStrideU, long ringStrideT, int nwr, int t, int B, int pace,
    double* __restrict__ Y)
{
    static_assert(kFftRows == 64, "eight rows per wavefront: the filter spectra of a wavefront fill at most 240 values");
    __shared__ __align__(16) double2 Hs[8][128];
    const int slot = blockIdx.x & 15, g = blockIdx.x >> 4;
    const int wave = WaveIdx(), lane = threadIdx.x & 63;
    const int sh = 8 * wave;
    const int b = blockIdx.y * 64 + lane;
    const bool live = b < B;
    const int bb = live ? b : B - 1;
    const int* rows = grpRowList + (long)g * kFftRows;
    const int nr = grpRows[g];
    const long i0 = itemOff[g], i1 = itemOff[g + 1];
    const int tRing = (t + 1) % nwr;
    double acc[8][3][2];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            acc[r][c][0] = acc[r][c][1] = 0.0;
    if (slot == 0)
        FftProducts<true>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, bb, H, WU, WT, ringStrideU, ringStrideT,
                          Hs[wave], acc);
    else
        FftProducts<false>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, bb, H, WU, WT, ringStrideU, ringStrideT,
                           Hs[wave], acc);
    if (!live)
        return;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (sh + r >= nr)
            break;
        double* y = Y + (((long)rows[sh + r] * 16 + slot) * 6) * B + b;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int ri = 0; ri < 2; ++ri)
                y[(long)(c * 2 + ri) * B] = acc[r][c][ri];
    }
}

// ---------------------------------------------------------------------------
// The same products on the FP64 matrix cores (v_mfma_f64_16x16x4f64; A (16 x
// 4): lane l holds A[l % 16][l / 16]; B (4 x 16): lane l holds
// B[l / 16][l % 16]; D (16 x 16): lane l, register q holds
// D[l / 16 + 4 q][l % 16]).  Per wavefront and item, the complex products of
// its 8 rows and 64 load cases are one real product
//   D (48 x 64) += A (48 x 12) B (12 x 64)
// with output o = 16 c + 8 ri + r (component c, real / imaginary ri, row r)
// and k = 6 field + 2 cc + wri (field U / T, component cc of the window,
// real / imaginary wri): A holds the filter values re / -im / im / re (slot
// 0: its two real bins, re . re and im . im), zero for the rows the item does
// not have; B the window spectra.  3 (c) x 3 (k steps) x 4 (16-case tiles)
// MFMAs per item; the 12 accumulator tiles stay in registers.  Each lane
// loads its own A entries (nine values of its row's filter spectra) and B
// entries one item ahead into registers, so no LDS stands between the loads
// and the MFMAs.  Items, windows and pacing as in
// k_conv_far_fft.
// ---------------------------------------------------------------------------
typedef double FftD4 __attribute__((ext_vector_type(4)));

template <bool PACKED>
__device__ __forceinline__ void FftProductsMf(const FftItem* __restrict__ items, long i0, long i1, int t, int tRing,
                                              int nwr, int pace, int sh, int slot, int lane, int B, int b0,
                                              const double* __restrict__ H, const double* __restrict__ WU,
                                              const double* __restrict__ WT, long ringStrideU, long ringStrideT,
                                              FftD4 (&acc)[3][4])
{
    const int li = lane & 15, lj = lane >> 4;
    const int ri = li >> 3, r = li & 7;
    // this lane's A entries: row r, output (component c, ri), k = 4 kt + lj:
    // the value at aoff[c][kt] of the row's filter spectra, negated for bit
    // 3 c + kt of neg, zero for that bit of zero
    int aoff[3][3], neg = 0, zero = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int kt = 0; kt < 3; ++kt) {
            const int k = 4 * kt + lj, field = k / 6, cc = (k % 6) >> 1, wri = k & 1;
            const int gi = c == cc ? (c == 0 ? 0 : (c == 1 ? 3 : 5)) : (c + cc == 1 ? 1 : (c + cc == 2 ? 2 : 4));
            const int base = field == 0 ? 2 * (3 * c + cc) : 18 + 2 * gi;
            const int bit = 1 << (3 * c + kt);
            if (PACKED) {
                aoff[c][kt] = base + ri;
                zero |= ri == wri ? 0 : bit;
            } else {
                aoff[c][kt] = base + (ri ^ wri);
                neg |= (ri == 0 && wri == 1) ? bit : 0;
            }
        }
    // this lane's B entries: k = 4 kt + lj -> field, value row kk of the slot
    long boff[3];
    bool bT[3];
#pragma unroll
    for (int kt = 0; kt < 3; ++kt) {
        const int k = 4 * kt + lj;
        bT[kt] = k >= 6;
        boff[kt] = ((long)slot * 6 + (k % 6)) * B;
    }
    auto nextItem = [&](long it) {
        while (it < i1 && (items[it].q > t + 1 || ((items[it].mask >> sh) & 0xffull) == 0))
            ++it;
        return it;
    };
    // window spectra (B operands) of item `it` (L2 hits: the workgroup's
    // wavefronts share them)
    auto loadW = [&](long it, double (&bw)[3][4]) {
        const FftItem item = items[it];
        const int beta = tRing - item.q + (tRing < item.q ? nwr : 0);
        const double* wu = WU + beta * ringStrideU + (long)item.cu * 96 * B;
        const double* wt = WT + beta * ringStrideT + (long)item.ct * 96 * B;
#pragma unroll
        for (int kt = 0; kt < 3; ++kt) {
            const double* w = (bT[kt] ? wt : wu) + boff[kt];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                int bn = b0 + nt * 16 + li;
                bn = bn < B ? bn : B - 1;
                bw[kt][nt] = w[bn];
            }
        }
    };
    // the lane's row of the item's filter spectra (A operands, read once from
    // HBM); a row the item lacks reads the wavefront's first row and its A
    // entries are zeroed
    auto loadH = [&](long it, double (&ha)[3][3]) {
        const FftItem item = items[it];
        const unsigned long long m = item.mask;
        const unsigned sub = (unsigned)((m >> sh) & 0xffull);
        const int pc = __popcll(m), below = __popcll(m & ((1ull << sh) - 1));
        const int rank = ((sub >> r) & 1u) ? __popc(sub & ((1u << r) - 1u)) : 0;
        const double* hr = H + item.hoff + ((long)slot * pc + below + rank) * 30;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int kt = 0; kt < 3; ++kt)
                ha[c][kt] = hr[aoff[c][kt]];
    };
    auto mac = [&](long it, const double (&bw)[3][4], const double (&ha)[3][3]) {
        const unsigned sub = (unsigned)((items[it].mask >> sh) & 0xffull);
        const bool have = (sub >> r) & 1u;
#pragma unroll
        for (int kt = 0; kt < 3; ++kt) {
            double a[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int bit = 1 << (3 * c + kt);
                const double v = ha[c][kt];
                a[c] = (!have || (zero & bit)) ? 0.0 : ((neg & bit) ? -v : v);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[c][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c], bw[kt][nt], acc[c][nt], 0, 0, 0);
        }
    };
    const int nPace = pace > 0 ? (int)((i1 - i0 + pace - 1) / pace) : 0;
    int paced = 0;
    auto paceTo = [&](long it) {
        const int chunk = pace > 0 ? (int)((it - i0) / pace) : 0;
        while (paced < chunk) {
            __builtin_amdgcn_s_barrier();
            ++paced;
        }
    };
    long cur = nextItem(i0);
    if (cur < i1) {
        // two register sets: the next item's window spectra and filter values
        // are in flight while the current item's products run (36 MFMAs of
        // 64 cycles); a deeper ring does not fit the registers of two
        // wavefronts per SIMD without spilling
        double w0[3][4], w1[3][4], h0[3][3], h1[3][3];
        loadW(cur, w0);
        loadH(cur, h0);
        for (;;) {
            long n1 = nextItem(cur + 1);
            loadW(n1 < i1 ? n1 : cur, w1);
            loadH(n1 < i1 ? n1 : cur, h1);
            paceTo(cur);
            mac(cur, w0, h0);
            if (n1 >= i1)
                break;
            cur = n1;
            n1 = nextItem(cur + 1);
            loadW(n1 < i1 ? n1 : cur, w0);
            loadH(n1 < i1 ? n1 : cur, h0);
            paceTo(cur);
            mac(cur, w1, h1);
            if (n1 >= i1)
                break;
            cur = n1;
        }
    }
    while (paced < nPace) {
        __builtin_amdgcn_s_barrier();
        ++paced;
    }
}

__global__ void __launch_bounds__(512) k_conv_far_fft_mf(
    const int* __restrict__ grpRowList, const int* __restrict__ grpRows, const long* __restrict__ itemOff,
    const FftItem* __restrict__ items, const double* __restrict__ H, const double* __restrict__ WU,
    const double* __restrict__ WT, long ringStrideU, long ringStrideT, int nwr, int t, int B, int pace,
    double* __restrict__ Y)
{
    const int slot = blockIdx.x & 15, g = blockIdx.x >> 4;
    const int wave = WaveIdx(), lane = threadIdx.x & 63;
    const int sh = 8 * wave;
    const int b0 = blockIdx.y * 64;
    const int* rows = grpRowList + (long)g * kFftRows;
    const int nr = grpRows[g];
    const long i0 = itemOff[g], i1 = itemOff[g + 1];
    const int tRing = (t + 1) % nwr;
    FftD4 acc[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            acc[c][nt] = FftD4{0.0, 0.0, 0.0, 0.0};
    if (slot == 0)
        FftProductsMf<true>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, b0, H, WU, WT, ringStrideU, ringStrideT,
                            acc);
    else
        FftProductsMf<false>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, b0, H, WU, WT, ringStrideU,
                             ringStrideT, acc);
    // D[o][n] with o = l / 16 + 4 q = 8 ri + r of component c, case 16 nt + l % 16
    const int li = lane & 15, lj = lane >> 4;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int bn = b0 + nt * 16 + li;
            if (bn >= B)
                continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int o = lj + 4 * q, ri = o >> 3, r = o & 7;
                if (sh + r < nr)
                    Y[(((long)rows[sh + r] * 16 + slot) * 6 + c * 2 + ri) * B + bn] = acc[c][nt][q];
            }
        }
}

void LaunchConvFarFft(int ngroups, const int* grpRowList, const int* grpRows, const long* itemOff, const FftItem* items,
                      const double* H, const double* WU, const double* WT, long ringStrideU, long ringStrideT,
                      int nwr, int t, int B, double* Y, bool mfma, int pace, hipStream_t s)
{
    dim3 g((unsigned)ngroups * 16, (unsigned)((B + 63) / 64));
    if (mfma)
        k_conv_far_fft_mf<<<g, 512, 0, s>>>(grpRowList, grpRows, itemOff, items, H, WU, WT, ringStrideU, ringStrideT, nwr,
                                            t, B, pace, Y);
    else
        k_conv_far_fft<<<g, 512, 0, s>>>(grpRowList, grpRows, itemOff, items, H, WU, WT, ringStrideU, ringStrideT, nwr, t,
                                         B, pace, Y);
}

// ---------------------------------------------------------------------------
// Inverse transforms: one thread per (source row, component, load case);
// samples 16 .. 16 + Tn - 1 of the 32-point inverse transform of the row's
// summed spectra into hfar[(i * 3ne + 3 row + c) * B + b].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fft_inverse(const double* __restrict__ Y, int ne, int B, int Tn,
                                                     double* __restrict__ hfar)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ne * 3 * B)
        return;
    const long row = idx / (3 * B);
    const int rem = (int)(idx - row * 3 * B);
    const int c = rem / B, b = rem - c * B;
    const double* y = Y + (row * 16 * 6 + c * 2) * (long)B + b;
    double yr[16], yi[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        yr[s] = y[(long)s * 6 * B];
        yi[s] = y[(long)s * 6 * B + B];
    }
    for (int i = 0; i < Tn; ++i) {
        const int nn = 16 + i;
        double v = yr[0] + ((nn & 1) ? -yi[0] : yi[0]);
        double sum = 0.0;
#pragma unroll
        for (int s = 1; s < 16; ++s) {
            const int j = (s * nn) & 31;
            sum = fma(yr[s], c_cos32[j], sum);
            sum = fma(-yi[s], c_sin32[j], sum);
        }
        v = (v + 2.0 * sum) * (1.0 / 32.0);
        hfar[((long)i * 3 * ne + 3 * row + c) * B + b] = v;
    }
}

void LaunchFftInverse(const double* Y, int ne, int B, int Tn, double* hfar, hipStream_t s)
{
    const long n = (long)ne * 3 * B;
    if (n <= 0)
        return;
    k_fft_inverse<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(Y, ne, B, Tn, hfar);
}

template void LaunchFftFilters<double>(PairGrid, const long*, const int*, const int*, const double*, const double*,
                                       const unsigned short*, const unsigned short*, int, const int*, const int*,
                                       const long*, long, const FftItem*, int, double*, hipStream_t);
template void LaunchFftFilters<float>(PairGrid, const long*, const int*, const int*, const float*, const float*,
                                      const unsigned short*, const unsigned short*, int, const int*, const int*,
                                      const long*, long, const FftItem*, int, double*, hipStream_t);

}  // namespace tdbem
