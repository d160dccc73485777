// HIP kernels of the batched load-case solve for gfx950.  The load case is
// the lane index: a wavefront carries up to 64 load cases of one scalar, so
// every matrix block, band entry and preconditioner value is wave-uniform
// (scalar loads, SGPR operands) and each lane streams only its own vectors.
#include "batched.hpp"

#include <type_traits>

#pragma clang fp contract(off)

namespace tdbem {

namespace {

__device__ __forceinline__ double Bf16Lo(unsigned short lo)
{
    return (double)__uint_as_float(((unsigned int)lo) << 16);
}

__device__ __forceinline__ int WaveIndex() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

}  // namespace

// ---------------------------------------------------------------------------
// Per-step plumbing (DBEM1/multidomain.cpp:3047-3200 for each load case)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_known_b(const NodeRefs* __restrict__ refs, const double* __restrict__ bd,
                                                 int ne, int nstore, int K, int B, double* __restrict__ known)
{
    // element fastest: the reads of bd, laid out per load case, are
    // contiguous across a wavefront
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ne * B)
        return;
    int b = (int)(idx / ne), i = (int)(idx % ne);
    const NodeRefs& r = refs[i];
    if (r.surface == 1)
        return;
    int blk = r.bcid == 123 ? r.u_block : r.t_block;
    int kn = r.bcid == 123 ? r.u_known : r.t_known;
    double sg = r.bcid == 123 ? r.u_sign : r.t_sign;
    if (!kn)
        return;
    const double* in = bd + ((long)b * nstore + K) * 3 * ne + 3 * (long)i;
    for (int c = 0; c < 3; ++c)
        known[(3 * (long)blk + c) * B + b] = sg * in[c];
}

void LaunchKnownB(const NodeRefs* refs, const double* bd, int ne, int nstore, int K, int B, double* known,
                  hipStream_t s)
{
    long n = (long)ne * B;
    k_known_b<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(refs, bd, ne, nstore, K, B, known);
}

__global__ void __launch_bounds__(256) k_scatter_b(const NodeRefs* __restrict__ refs, const ElemGeo* __restrict__ geo,
                                                   const double* __restrict__ bd, const double* __restrict__ x,
                                                   int ne, int nstore, int K, int B, double* __restrict__ lu,
                                                   double* __restrict__ lt, double* __restrict__ gU,
                                                   double* __restrict__ gT)
{
    // element fastest: the writes of lu / lt (and the reads of bd), laid out
    // per load case, are contiguous across a wavefront; the reads of x
    // (load case fastest) are not
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ne * B)
        return;
    int b = (int)(idx / ne), i = (int)(idx % ne);
    const NodeRefs& r = refs[i];
    const double* T = r.itf_T;
    long slice = ((long)b * nstore + K) * 3 * ne;
    const double* in = bd + slice + 3 * (long)i;
    double u[3], t[3], v[3];
    if (!r.u_known) {
        for (int q = 0; q < 3; ++q)
            v[q] = x[(3 * (long)r.u_block + q) * B + b];
        for (int q = 0; q < 3; ++q)
            u[q] = r.u_sign * (T[q * 3 + 0] * v[0] + T[q * 3 + 1] * v[1] + T[q * 3 + 2] * v[2]);
    } else {
        for (int q = 0; q < 3; ++q)
            u[q] = r.u_sign * in[q];
    }
    if (!r.t_known) {
        for (int q = 0; q < 3; ++q)
            v[q] = x[(3 * (long)r.t_block + q) * B + b];
        for (int q = 0; q < 3; ++q)
            t[q] = r.t_sign * (T[q * 3 + 0] * v[0] + T[q * 3 + 1] * v[1] + T[q * 3 + 2] * v[2]);
    } else {
        for (int q = 0; q < 3; ++q)
            t[q] = r.t_sign * in[q];
    }
    const double* F = geo[i].trans;
    double* ou = lu + slice + 3 * (long)i;
    double* ot = lt + slice + 3 * (long)i;
    double gu[3], gt[3];
    gu[0] = F[0] * u[0] + F[3] * u[1] + F[6] * u[2];
    gu[1] = F[1] * u[0] + F[4] * u[1] + F[7] * u[2];
    gu[2] = F[2] * u[0] + F[5] * u[1] + F[8] * u[2];
    gt[0] = F[0] * t[0] + F[3] * t[1] + F[6] * t[2];
    gt[1] = F[1] * t[0] + F[4] * t[1] + F[7] * t[2];
    gt[2] = F[2] * t[0] + F[5] * t[1] + F[8] * t[2];
    for (int c = 0; c < 3; ++c) {
        ou[c] = gu[c];
        ot[c] = gt[c];
    }
    // with gU / gT, also the state series of step K (as k_state_b makes them)
    if (gU && !(r.surface == 1 && !r.ref_side))
        for (int c = 0; c < 3; ++c) {
            gU[(((long)r.hu_block * nstore + K) * 3 + c) * B + b] = r.hu_sign * gu[c];
            gT[(((long)r.ht_block * nstore + K) * 3 + c) * B + b] = r.ht_sign * gt[c];
        }
}

void LaunchScatterB(const NodeRefs* refs, const ElemGeo* geo, const double* bd, const double* x, int ne, int nstore,
                    int K, int B, double* lu, double* lt, hipStream_t s, double* gU, double* gT)
{
    long n = (long)ne * B;
    k_scatter_b<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(refs, geo, bd, x, ne, nstore, K, B, lu, lt, gU, gT);
}

__global__ void __launch_bounds__(256) k_state_b(const NodeRefs* __restrict__ refs, const double* __restrict__ lu,
                                                 const double* __restrict__ lt, int ne, int nstore, int K, int B,
                                                 double* __restrict__ gU, double* __restrict__ gT)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ne * B)
        return;
    int i = (int)(idx / B), b = (int)(idx % B);
    const NodeRefs& r = refs[i];
    if (r.surface == 1 && !r.ref_side)
        return;
    long slice = ((long)b * nstore + K) * 3 * ne + 3 * (long)i;
    for (int c = 0; c < 3; ++c) {
        gU[(((long)r.hu_block * nstore + K) * 3 + c) * B + b] = r.hu_sign * lu[slice + c];
        gT[(((long)r.ht_block * nstore + K) * 3 + c) * B + b] = r.ht_sign * lt[slice + c];
    }
}

void LaunchStateB(const NodeRefs* refs, const double* lu, const double* lt, int ne, int nstore, int K, int B,
                  double* gU, double* gT, hipStream_t s)
{
    long n = (long)ne * B;
    k_state_b<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(refs, lu, lt, ne, nstore, K, B, gU, gT);
}

// Step 0 of every case in the global frame (DBEM1/multidomain.cpp:3819-3822):
// lt = 0, the given triplets moved to lt for BCID 456 (Convert), then
// LocToAbs; the outputs of device-resident solves.
__global__ void __launch_bounds__(256) k_step0_out(const NodeRefs* __restrict__ refs, const ElemGeo* __restrict__ geo,
                                                   const double* __restrict__ bd, int ne, int nstore, int B,
                                                   double* __restrict__ lu, double* __restrict__ lt)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ne * B)
        return;
    int b = (int)(idx / ne), i = (int)(idx % ne);
    long slice = (long)b * nstore * 3 * ne + 3 * (long)i;
    double u[3], t[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = bd[slice + k];
        t[k] = 0.0;
    }
    if (refs[i].bcid == 456)
        for (int k = 0; k < 3; ++k) {
            t[k] = u[k];
            u[k] = 0.0;
        }
    const double* F = geo[i].trans;
    for (int r = 0; r < 3; ++r) {
        lu[slice + r] = F[r] * u[0] + F[3 + r] * u[1] + F[6 + r] * u[2];
        lt[slice + r] = F[r] * t[0] + F[3 + r] * t[1] + F[6 + r] * t[2];
    }
}

void LaunchStep0Out(const NodeRefs* refs, const ElemGeo* geo, const double* bd, int ne, int nstore, int B, double* lu,
                    double* lt, hipStream_t s)
{
    long n = (long)ne * B;
    k_step0_out<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(refs, geo, bd, ne, nstore, B, lu, lt);
}

__global__ void __launch_bounds__(256) k_filter_b(const double* __restrict__ g, long cols, int nstore, int B, int x0,
                                                  int nx, int K0, double* __restrict__ F)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long per = (long)nx * 3 * B;
    if (idx >= cols * per)
        return;
    long col = idx / per;
    long rem = idx - col * per;
    int x = x0 + (int)(rem / (3 * B));
    long cb = rem % (3 * B);  // c * B + b
    auto raw = [&](int m) {
        return (m >= 1 && m <= K0 - 1 && m < nstore) ? g[(col * nstore + m) * 3 * B + cb] : 0.0;
    };
    F[(col * (nstore + 1) + x) * 3 * B + cb] = (raw(x - 1) + 2.0 * raw(x)) + raw(x + 1);
}

void LaunchFilterB(const double* g, long cols, int nstore, int B, int x0, int x1, int K0, double* F, hipStream_t s)
{
    if (x0 < 0)
        x0 = 0;
    if (x1 > nstore - 1)
        x1 = nstore - 1;
    int nx = x1 - x0 + 1;
    if (nx <= 0 || cols <= 0)
        return;
    long n = cols * nx * 3 * B;
    k_filter_b<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, cols, nstore, B, x0, nx, K0, F);
}

// ---------------------------------------------------------------------------
// y = A x for B vectors: one wavefront per block row, one lane per load case,
// the reference's per-block summation (DBEM1/mat_vec.cpp:2273-2300).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_spmm_b(long nrows, const int* __restrict__ rp, const int* __restrict__ col,
                                                const double* __restrict__ val, const double* __restrict__ x,
                                                double* __restrict__ y, int B, const int* __restrict__ left)
{
    if (left && *left == 0)
        return;
    long row = (long)blockIdx.x * 4 + WaveIndex();
    int b = blockIdx.y * 64 + (threadIdx.x & 63);
    if (row >= nrows || b >= B)
        return;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    auto add = [&](const double* v, double x0, double x1, double x2) {
        b0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
        b1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
        b2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    };
    int j = rp[row];
    const int je = rp[row + 1];
    // four blocks at a time: their twelve state loads issued together, the
    // sums in block order (as one block at a time)
    for (; j + 4 <= je; j += 4) {
        double xs[4][3];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const long c3 = 3 * (long)col[j + t];
            xs[t][0] = x[c3 * B + b];
            xs[t][1] = x[(c3 + 1) * B + b];
            xs[t][2] = x[(c3 + 2) * B + b];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            add(val + 9 * (long)(j + t), xs[t][0], xs[t][1], xs[t][2]);
    }
    for (; j < je; ++j) {
        const long c3 = 3 * (long)col[j];
        add(val + 9 * (long)j, x[c3 * B + b], x[(c3 + 1) * B + b], x[(c3 + 2) * B + b]);
    }
    y[(3 * row) * B + b] = b0;
    y[(3 * row + 1) * B + b] = b1;
    y[(3 * row + 2) * B + b] = b2;
}

void LaunchSpmmB(long nrows, const int* rp, const int* col, const double* val, const double* x, double* y, int B,
                 hipStream_t s, const int* left)
{
    dim3 grid((unsigned)((nrows + 3) / 4), (unsigned)((B + 63) / 64));
    k_spmm_b<<<grid, 256, 0, s>>>(nrows, rp, col, val, x, y, B, left);
}

// ---------------------------------------------------------------------------
// Batched time-blocked convolution.  Far part: one workgroup of four
// wavefronts per (source row, 64 load cases); a wavefront takes every fourth
// pair of the row.  The blocks
// of up to 64 lags of a pair are staged in LDS with coalesced loads and read
// back as wave-uniform broadcasts.  Each lane keeps the filtered states u~, t~
// of x = K0 - j .. K0 - j + T - 1 in a register window whose slots rotate with
// the lag (the slot of x is fixed per lag within a group of T lags), so a new
// lag needs one state (6 values, loaded two lags ahead: the loads are latency-
// bound, and a second lag in flight took 14 % off the launch) and applies its block
// to all T steps (18 FMA per step).
// ---------------------------------------------------------------------------
template <int T, typename TB>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_conv_far_b(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                    const long* __restrict__ off, const int* __restrict__ kb,
                                                    const int* __restrict__ len, const TB* __restrict__ mts,
                                                    const TB* __restrict__ mgs,
                                                    const unsigned short* __restrict__ mtsLo,
                                                    const unsigned short* __restrict__ mgsLo,
                                                    const double* __restrict__ FU, const double* __restrict__ FT,
                                                    int nstore, int K0, int Tn, int maxn, int ne, int B,
                                                    double* __restrict__ hfar, long partStride)
{
    constexpr int D = 3;          // lags between a state's load and its first use
    constexpr int R = T + D - 1;  // state ring slots
    constexpr int CH = 6 * R;     // lags per staged chunk (a multiple of R)
    __shared__ double sblk[CH * 15];
    // four one-wavefront workgroups per source row: wavefront `wave` takes
    // every fourth pair and writes its own partial sums (part `wave` of hfar)
    int wave = (int)(blockIdx.x & 3u);
    int s = (int)(blockIdx.x >> 2);
    int lane = threadIdx.x & 63;
    int b = blockIdx.y * 64 + lane;
    bool live = b < B;
    int bb = live ? b : B - 1;
    int d = refs[s].domain;
    int nd = grid.eleCount[d];
    int eb = grid.eleBegin[d];
    long rowStart = grid.pairBase[d] + (long)(s - eb) * nd;
    double* sb = sblk;
    int Kmax = K0 + Tn - 1;
    int jmax = Kmax < maxn ? Kmax : maxn;
    double acc[T][3];
#pragma unroll
    for (int i = 0; i < T; ++i)
        acc[i][0] = acc[i][1] = acc[i][2] = 0.0;
    for (int q = wave; q < nd; q += 4) {
        long p = rowStart + q;
        int L = len[p];
        int k0 = kb[p];
        int jend = k0 + L - 1 < jmax ? k0 + L - 1 : jmax;
        if (L == 0 || k0 > jend)
            continue;
        int f = eb + q;
        const double* fu = FU + (long)refs[f].hu_block * (nstore + 1) * 3 * B + bb;
        const double* ft = FT + (long)refs[f].ht_block * (nstore + 1) * 3 * B + bb;
        long base = off[p];
        // the state of x; x outside 0 .. nstore-1 reads row nstore of the
        // column, which the filter never writes (zero): a wave-uniform row
        // choice, the loads issued on every path and no per-value select
        auto loadx = [&](int x, double* w) {
            const int xr = x >= 0 && x < nstore ? x : nstore;
            const double* a = fu + (long)xr * 3 * B;
            const double* c = ft + (long)xr * 3 * B;
            w[0] = a[0];
            w[1] = a[B];
            w[2] = a[2 * B];
            w[3] = c[0];
            w[4] = c[B];
            w[5] = c[2 * B];
        };
        // ring of R = T + D - 1 states: slot (x - (K0 - k0)) mod R holds the
        // state of x; at lag j (rotation u = (j - k0) mod R) the steps read
        // slots (i - u) mod R and the state lag j + D adds is loaded into the
        // slot step T - 1 has just used, D lags before its first use, with
        // no register copies
        double ring[R][6];
#pragma unroll
        for (int qq = 0; qq < T; ++qq)
            loadx(K0 - k0 + qq, ring[qq]);
#pragma unroll
        for (int dd = 1; dd < D; ++dd)
            loadx(K0 - k0 - dd, ring[R - dd]);
        for (int c0 = k0; c0 <= jend; c0 += CH) {
            int c1 = c0 + CH - 1 < jend ? c0 + CH - 1 : jend;
            int nlag = c1 - c0 + 1;
            long e0 = base + (c0 - k0);
            // the chunk's loads issued together per block type (9 M values,
            // then 6 G values per lane), each lane's values converted and
            // stored after its loads; indexes past the chunk read its first
            // entry and are not stored
            auto stage = [&](const TB* src, const unsigned short* lo, int nv, int off9, auto nper) {
                constexpr int NP = decltype(nper)::value;
                TB rv[NP];
                unsigned short rl[NP];
#pragma unroll
                for (int v = 0; v < NP; ++v) {
                    int idx = lane + 64 * v;
                    rv[v] = src[(long)nv * e0 + (idx < nlag * nv ? idx : 0)];
                }
                if (lo) {
#pragma unroll
                    for (int v = 0; v < NP; ++v) {
                        int idx = lane + 64 * v;
                        rl[v] = lo[(long)nv * e0 + (idx < nlag * nv ? idx : 0)];
                    }
                }
#pragma unroll
                for (int v = 0; v < NP; ++v) {
                    int idx = lane + 64 * v;
                    if (idx < nlag * nv) {
                        int r = idx / nv, k = idx - nv * r;
                        double x = (double)rv[v];
                        if (lo)
                            x += Bf16Lo(rl[v]);
                        sb[r * 15 + off9 + k] = x;
                    }
                }
            };
            stage(mts, mtsLo, 9, 0, std::integral_constant<int, (CH * 9 + 63) / 64>{});
            stage(mgs, mgsLo, 6, 9, std::integral_constant<int, (CH * 6 + 63) / 64>{});
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int jg = c0; jg <= c1; jg += R) {
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    int j = jg + u;
                    if (j <= c1) {
                        const double* bl = sb + (j - c0) * 15;
                        double M[9], G[6];
#pragma unroll
                        for (int k = 0; k < 9; ++k)
                            M[k] = bl[k];
#pragma unroll
                        for (int k = 0; k < 6; ++k)
                            G[k] = bl[9 + k];
                        // all T steps (one basic block; steps past Tn are never stored)
#pragma unroll
                        for (int i = 0; i < T; ++i) {
                            const double* w = ring[(i - u + R) % R];
                            double a0 = acc[i][0], a1 = acc[i][1], a2 = acc[i][2];
                            a0 = fma(G[0], w[3], a0); a0 = fma(G[1], w[4], a0); a0 = fma(G[2], w[5], a0);
                            a1 = fma(G[1], w[3], a1); a1 = fma(G[3], w[4], a1); a1 = fma(G[4], w[5], a1);
                            a2 = fma(G[2], w[3], a2); a2 = fma(G[4], w[4], a2); a2 = fma(G[5], w[5], a2);
                            a0 = fma(-M[0], w[0], a0); a0 = fma(-M[1], w[1], a0); a0 = fma(-M[2], w[2], a0);
                            a1 = fma(-M[3], w[0], a1); a1 = fma(-M[4], w[1], a1); a1 = fma(-M[5], w[2], a1);
                            a2 = fma(-M[6], w[0], a2); a2 = fma(-M[7], w[1], a2); a2 = fma(-M[8], w[2], a2);
                            acc[i][0] = a0;
                            acc[i][1] = a1;
                            acc[i][2] = a2;
                        }
                        loadx(K0 - j - D, ring[(T - 1 - u + R) % R]);  // the state lag j + D adds
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
#pragma unroll
    for (int i = 0; i < T; ++i)
        for (int c = 0; c < 3; ++c)
            if (live && i < Tn)
                hfar[wave * partStride + ((long)i * 3 * ne + 3 * (long)s + c) * B + b] = acc[i][c];
}

// ---------------------------------------------------------------------------
// The same far part on the FP64 matrix cores.  For one field element f and
// kMfRows consecutive source rows s of its domain, the tile's far sums are
//   h[s, i, c, b] += sum_x sum_c' A_i[(s, c), (x, c')] F[f, x, c', b],
//   A_i[(s, c), (x, c')] = block of pair (s, f) at lag j = K0 + i - x,
// with the state x fixed per MFMA k-step (four consecutive states) and the
// step i selecting which lags of the staged window form the A operand.  So
// each state value loaded for a k-step feeds T steps x 3 components x
// kMfRows rows of 16 load cases, and each lag block is read from LDS by
// address arithmetic alone (no Toeplitz expansion in memory).  M = 3 x
// kMfRows = 15 of the 16 rows of v_mfma_f64_16x16x4f64 (row 15 reads a zero
// LDS row), N = 16 load cases per wavefront, K = 4 states of one component.
// Lane layout of v_mfma_f64_16x16x4f64 (64 lanes): A (16 x 4) lane l holds
// A[l % 16][l / 16]; B (4 x 16) lane l holds B[l / 16][l % 16]; D (16 x 16)
// lane l, register r holds D[l / 16 + 4 r][l % 16].
// One workgroup = one row group x 64 load cases (a wavefront per 16 cases);
// the lag window of each k-step is staged once for the four wavefronts, in
// two LDS buffers (the next window is fetched into registers while the
// current k-step runs).  The sums differ from k_conv_far_b by rounding only.
// ---------------------------------------------------------------------------
constexpr int kMfRows = 5;

typedef double MfD4 __attribute__((ext_vector_type(4)));

template <int T, typename TB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_conv_far_bm(
    PairGrid grid, const NodeRefs* __restrict__ refs, const long* __restrict__ off, const int* __restrict__ kb,
    const int* __restrict__ len, const TB* __restrict__ mts, const TB* __restrict__ mgs,
    const unsigned short* __restrict__ mtsLo, const unsigned short* __restrict__ mgsLo,
    const double* __restrict__ FU, const double* __restrict__ FT, int nstore, int K0, int Tn, int maxn, int ne,
    int B, int parts, long partStride, double* __restrict__ hfar)
{
    static_assert(T == 8, "two groups of four steps per k-step");
    constexpr int W = T + 3;        // lags of one k-step's window
    constexpr int VL = 18;          // staged values per lag: -M (9), G as a full 3 x 3 (9)
    constexpr int RS = W * VL + 3;  // LDS doubles per row of a window (A reads at most 2-way bank conflicts)
    constexpr int TPR = 256 / kMfRows;   // staging threads per row
    constexpr int PER = (W * VL + TPR - 1) / TPR;  // staged values per thread
    __shared__ double sw[2][(kMfRows + 1) * RS];
    // workgroup -> (row group, part of the columns); row group -> domain, first local row
    const int part = (int)blockIdx.x % parts;
    int g = (int)blockIdx.x / parts, d = 0;
    for (;;) {
        int ng = (grid.eleCount[d] + kMfRows - 1) / kMfRows;
        if (g < ng || d + 1 >= grid.nd)
            break;
        g -= ng;
        ++d;
    }
    const int nd = grid.eleCount[d], eb = grid.eleBegin[d];
    const int r0 = g * kMfRows;
    if (r0 >= nd)
        return;
    const int nrows = nd - r0 < kMfRows ? nd - r0 : kMfRows;
    const int tid = threadIdx.x, lane = tid & 63, wave = WaveIndex();
    const int kq = lane >> 4;
    const int b = blockIdx.y * 64 + wave * 16 + (lane & 15);
    const bool live = b < B;
    const bool waveLive = blockIdx.y * 64 + wave * 16 < B;
    const int bb = live ? b : 0;
    // this lane's A row m = 3 sl + c (row 15 -> the zero row kMfRows)
    const int m = lane & 15, sl = m / 3, c = m - 3 * (m / 3);
    const int abase = sl * RS + kq * VL + 3 * c;  // + component offset + VL * step
    const int Kmax = K0 + Tn - 1;
    const int jmax = Kmax < maxn ? Kmax : maxn;
    MfD4 acc[T];
#pragma unroll
    for (int i = 0; i < T; ++i)
        acc[i] = MfD4{0.0, 0.0, 0.0, 0.0};
    for (int idx = tid; idx < RS; idx += 256) {
        sw[0][kMfRows * RS + idx] = 0.0;
        sw[1][kMfRows * RS + idx] = 0.0;
    }
    const long pair0 = grid.pairBase[d] + (long)r0 * nd;  // pair (row r0, column 0)
    const int sr = tid / TPR, st = tid - TPR * (tid / TPR);  // staging row (kMfRows: idle) and slot
    constexpr int kG9 = 0 | 1 << 3 | 2 << 6 | 1 << 9 | 3 << 12 | 4 << 15 | 2 << 18 | 4 << 21 | 5 << 24;
    const int q0 = (int)((long)nd * part / parts), q1 = (int)((long)nd * (part + 1) / parts);
    for (int q = q0; q < q1; ++q) {
        // band union of the group's rows (the same in every wavefront)
        int jlo = 0x7fffffff, jhi = -1;
        for (int l = 0; l < nrows; ++l) {
            long p = pair0 + (long)l * nd + q;
            int L = len[p], k0 = kb[p];
            int je = k0 + L - 1 < jmax ? k0 + L - 1 : jmax;
            if (L > 0 && k0 <= je) {
                jlo = k0 < jlo ? k0 : jlo;
                jhi = je > jhi ? je : jhi;
            }
        }
        if (jhi < 0)
            continue;
        // this thread stages values t, t + TPR, ... (window lag u = t / VL,
        // value k = t % VL) of row sr: first lag rk0, last lag rje, entry roff
        int rk0 = 1, rje = 0;
        long roff = 0;
        if (sr < nrows) {
            long p = pair0 + (long)sr * nd + q;
            int L = len[p];
            rk0 = kb[p];
            rje = L > 0 ? (rk0 + L - 1 < jmax ? rk0 + L - 1 : jmax) : rk0 - 1;
            roff = off[p];
        }
        TB raw[PER];
        unsigned short rawLo[PER];
        auto fetch = [&](int jw) {
#pragma unroll
            for (int v = 0; v < PER; ++v) {
                int t = st + TPR * v;
                int u = t / VL, k = t - VL * (t / VL);
                int j = jw + u;
                bool ok = j >= rk0 && j <= rje;
                long e = ok ? roff + (j - rk0) : 0;
                // G (00 01 02 11 12 22) expanded row by row: 0 1 2 / 1 3 4 / 2 4 5
                long at = k < 9 ? 9 * e + k : 6 * e + ((kG9 >> (3 * (k - 9))) & 7);
                raw[v] = (k < 9 ? mts : mgs)[at];
                if (mtsLo)
                    rawLo[v] = (k < 9 ? mtsLo : mgsLo)[at];
            }
        };
        auto store = [&](int buf, int jw) {
#pragma unroll
            for (int v = 0; v < PER; ++v) {
                int t = st + TPR * v;
                int u = t / VL, k = t - VL * (t / VL);
                int j = jw + u;
                bool ok = j >= rk0 && j <= rje;
                double x = (double)raw[v];
                if (mtsLo)
                    x += Bf16Lo(rawLo[v]);
                x = ok ? (k < 9 ? -x : x) : 0.0;
                if (sr < kMfRows && t < W * VL)
                    sw[buf][sr * RS + t] = x;
            }
        };
        const int f = eb + q;
        const double* fu = FU + (long)refs[f].hu_block * (nstore + 1) * 3 * B + bb;
        const double* ft = FT + (long)refs[f].ht_block * (nstore + 1) * 3 * B + bb;
        // k-steps: states xb - 3 .. xb, xb = xTop, xTop - 4, ... while xb >= K0 - jhi
        const int xTop = K0 + Tn - 1 - jlo;
        const int nk = (xTop - (K0 - jhi)) / 4 + 1;
        // B operand of the k-step with top state xb: F[x = xb - kq][c'][b]
        auto loadB = [&](int xb, double* bv) {
            const int x = xb - kq;
            const bool xin = live && x >= 0 && x < nstore;
            const long xo = xin ? (long)x * 3 * B : 0;
#pragma unroll
            for (int cp = 0; cp < 6; ++cp) {
                double v = (cp < 3 ? fu : ft)[xo + (long)(cp % 3) * B];
                bv[cp] = xin ? v : 0.0;
            }
        };
        fetch(K0 - xTop);
        double bv[6];
        loadB(xTop, bv);
        store(0, K0 - xTop);
        __syncthreads();
        for (int ks = 0; ks < nk; ++ks) {
            const int xb = xTop - 4 * ks;
            const int jw = K0 - xb;  // window lags jw .. jw + W - 1
            const bool more = ks + 1 < nk;
            if (more)
                fetch(jw + 4);
            double bn[6];
            if (more)
                loadB(xb - 4, bn);
            if (waveLive) {
                // steps i with a lag of the union in this k-step: [ia, ib]
                const bool lo4 = jlo - 3 - jw <= 3;  // some i in 0..3
                const bool hi4 = jhi - jw >= 4;      // some i in 4..7
                const double* A = sw[ks & 1] + abase;
#pragma unroll
                for (int cp = 0; cp < 6; ++cp) {
                    const int co = cp < 3 ? cp : cp + 6;
                    if (lo4) {
                        double a[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            a[i] = A[co + VL * i];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[cp], acc[i], 0, 0, 0);
                    }
                    if (hi4) {
                        double a[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            a[i] = A[co + VL * (i + 4)];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            acc[i + 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[cp], acc[i + 4], 0, 0, 0);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            if (more) {
                store((ks + 1) & 1, jw + 4);
#pragma unroll
                for (int cp = 0; cp < 6; ++cp)
                    bv[cp] = bn[cp];
            }
            __syncthreads();
        }
    }
    if (!live)
        return;
#pragma unroll
    for (int i = 0; i < T; ++i) {
        if (i >= Tn)
            break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int row = kq + 4 * r;
            int rs = row / 3, rc = row - 3 * (row / 3);
            if (row < 3 * kMfRows && rs < nrows)
                hfar[part * partStride + ((long)i * 3 * ne + 3 * (long)(eb + r0 + rs) + rc) * B + b] = acc[i][r];
        }
    }
}

template <typename TB>
void LaunchConvFarB(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                    const TB* mts, const TB* mgs, const unsigned short* mtsLo, const unsigned short* mgsLo,
                    const double* FU, const double* FT, int nstore, int K0, int Tn, int maxn, int ne, int B,
                    double* hfar, int parts, hipStream_t s)
{
    if (parts > 0) {
        long groups = 0;
        for (int d = 0; d < grid.nd; ++d)
            groups += (grid.eleCount[d] + kMfRows - 1) / kMfRows;
        dim3 g((unsigned)(groups * parts), (unsigned)((B + 63) / 64));
        k_conv_far_bm<kBatchTile, TB><<<g, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, mtsLo, mgsLo, FU, FT,
                                                        nstore, K0, Tn, maxn, ne, B, parts,
                                                        (long)kBatchTile * 3 * ne * B, hfar);
        return;
    }
    dim3 g((unsigned)ne * kFarBWaves, (unsigned)((B + 63) / 64));
    k_conv_far_b<kBatchTile, TB><<<g, 64, 0, s>>>(grid, refs, off, kb, len, mts, mgs, mtsLo, mgsLo, FU, FT, nstore,
                                                  K0, Tn, maxn, ne, B, hfar, (long)kBatchTile * 3 * ne * B);
}

// Near part of step K (states K0 .. K-1 inside the tile) plus the tile's far
// part: one workgroup of four wavefronts per (source row, 64 load cases),
// each wavefront summing every fourth near pair's lags per lane in order.
template <typename TB>
__global__ void __launch_bounds__(256) k_conv_near_b(const NodeRefs* __restrict__ refs,
                                                    const long* __restrict__ nearPtr,
                                                    const long* __restrict__ nearPair,
                                                    const int* __restrict__ nearField, const long* __restrict__ off,
                                                    const int* __restrict__ kb, const int* __restrict__ len,
                                                    const TB* __restrict__ mts, const TB* __restrict__ mgs,
                                                    const unsigned short* __restrict__ mtsLo,
                                                    const unsigned short* __restrict__ mgsLo,
                                                    const double* __restrict__ gU, const double* __restrict__ gT,
                                                    int nstore, int K, int K0, int maxn, int ne, int B,
                                                    const double* __restrict__ hfar, int parts, long partStride,
                                                    double* __restrict__ rhs, const double* __restrict__ hfar2,
                                                    int K02)
{
    __shared__ double red[4][3][64];
    int s = blockIdx.x;
    int lane = threadIdx.x & 63, wave = WaveIndex();
    int b = blockIdx.y * 64 + lane;
    bool live = b < B;
    int bb = live ? b : B - 1;
    int jlim = K - K0 + 1;
    if (jlim > maxn)
        jlim = maxn;
    double y[3] = {0.0, 0.0, 0.0};
    // four wavefronts per source row, each every fourth near pair; partial
    // sums added in wavefront order
    for (long q = nearPtr[s] + wave; q < nearPtr[s + 1]; q += 4) {
        long p = nearPair[q];
        int k0 = kb[p], L = len[p];
        if (k0 > jlim)
            break;  // the list is ordered by band start: nothing further is within the lag limit
        int f = nearField[q];
        const double* su = gU + (long)refs[f].hu_block * nstore * 3 * B + bb;
        const double* st = gT + (long)refs[f].ht_block * nstore * 3 * B + bb;
        for (int t = 0; t < L; ++t) {
            int j = k0 + t;
            if (j > jlim)
                break;
            long e = off[p] + t;
            double M[9], G[6];
            for (int i = 0; i < 9; ++i)
                M[i] = (double)mts[9 * e + i] + (mtsLo ? Bf16Lo(mtsLo[9 * e + i]) : 0.0);
            for (int i = 0; i < 6; ++i)
                G[i] = (double)mgs[6 * e + i] + (mgsLo ? Bf16Lo(mgsLo[6 * e + i]) : 0.0);
            const int mm[3] = {K - j + 1, K - j, K - j - 1};
            const double co[3] = {1.0, 2.0, 1.0};
            double cu[3] = {0.0, 0.0, 0.0}, ct[3] = {0.0, 0.0, 0.0};
            for (int z = 0; z < 3; ++z) {
                int m = mm[z];
                if (m >= K0 && m <= K - 1 && m >= 1) {
                    for (int c = 0; c < 3; ++c) {
                        cu[c] += co[z] * su[((long)m * 3 + c) * B];
                        ct[c] += co[z] * st[((long)m * 3 + c) * B];
                    }
                }
            }
            y[0] += (G[0] * ct[0] + G[1] * ct[1] + G[2] * ct[2]) - (M[0] * cu[0] + M[1] * cu[1] + M[2] * cu[2]);
            y[1] += (G[1] * ct[0] + G[3] * ct[1] + G[4] * ct[2]) - (M[3] * cu[0] + M[4] * cu[1] + M[5] * cu[2]);
            y[2] += (G[2] * ct[0] + G[4] * ct[1] + G[5] * ct[2]) - (M[6] * cu[0] + M[7] * cu[1] + M[8] * cu[2]);
        }
    }
    for (int c = 0; c < 3; ++c)
        red[wave][c][lane] = y[c];
    __syncthreads();
    if (wave != 0 || !live)
        return;
    for (int c = 0; c < 3; ++c)
        y[c] = ((red[0][c][lane] + red[1][c][lane]) + red[2][c][lane]) + red[3][c][lane];
    const double* hf = hfar + ((long)(K - K0) * 3 * ne + 3 * (long)s) * B + b;
    for (int c = 0; c < 3; ++c) {
        double v = hf[(long)c * B];
        for (int p = 1; p < parts; ++p)
            v += hf[p * partStride + (long)c * B];
        if (hfar2)
            v += hfar2[((long)(K - K02) * 3 * ne + 3 * (long)s + c) * B + b];
        rhs[(3 * (long)s + c) * B + b] += v + y[c];
    }
}

template <typename TB>
void LaunchConvNearB(const NodeRefs* refs, const long* nearPtr, const long* nearPair, const int* nearField,
                     const long* off, const int* kb, const int* len, const TB* mts, const TB* mgs,
                     const unsigned short* mtsLo, const unsigned short* mgsLo, const double* gU, const double* gT,
                     int nstore, int K, int K0, int maxn, int ne, int B, const double* hfar, int parts,
                     double* rhs, hipStream_t s, const double* hfar2, int K02)
{
    dim3 g((unsigned)ne, (unsigned)((B + 63) / 64));
    k_conv_near_b<TB><<<g, 256, 0, s>>>(refs, nearPtr, nearPair, nearField, off, kb, len, mts, mgs, mtsLo, mgsLo, gU,
                                       gT, nstore, K, K0, maxn, ne, B, hfar, parts < 1 ? 1 : parts,
                                       (long)kBatchTile * 3 * ne * B, rhs, hfar2, K02);
}

#define TDBEM_BATCH_INSTANTIATE(TB)                                                                                  \
    template void LaunchConvFarB<TB>(PairGrid, const NodeRefs*, const long*, const int*, const int*, const TB*,       \
                                     const TB*, const unsigned short*, const unsigned short*, const double*,          \
                                     const double*, int, int, int, int, int, int, double*, int, hipStream_t);         \
    template void LaunchConvNearB<TB>(const NodeRefs*, const long*, const long*, const int*, const long*, const int*, \
                                      const int*, const TB*, const TB*, const unsigned short*,                        \
                                      const unsigned short*, const double*, const double*, int, int, int, int, int,   \
                                      int, const double*, int, double*, hipStream_t, const double*, int);
TDBEM_BATCH_INSTANTIATE(double)
TDBEM_BATCH_INSTANTIATE(float)
#undef TDBEM_BATCH_INSTANTIATE

// ---------------------------------------------------------------------------
// Batched GMRES linear algebra.  Vectors of a batch: (r, b) at r * B + b.
// Per-load-case scalars (Hessenberg h, Givens QT / R, f, phibar, flags):
// entry k of case b at k * B + b.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_precond_b(long n, const long* __restrict__ posLeaf,
                                                   const long* __restrict__ leafBegin,
                                                   const long* __restrict__ leafSize, const long* __restrict__ matOff,
                                                   const double* __restrict__ inv, const long* __restrict__ repid,
                                                   const long* __restrict__ outBlock, const double* __restrict__ x,
                                                   double* __restrict__ y, int B)
{
    // one wavefront per element: its three rows share every x load
    long i = (long)blockIdx.x * 4 + WaveIndex();
    int b = blockIdx.y * 64 + (threadIdx.x & 63);
    if (i >= n || b >= B)
        return;
    long leaf = posLeaf[i];
    long lb = leafBegin[leaf], c = leafSize[leaf], jl = i - lb, n3 = 3 * c;
    const double* P0 = inv + matOff[leaf] + 3 * jl * n3;
    const double* P1 = P0 + n3;
    const double* P2 = P1 + n3;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (long k = 0; k < c; ++k) {
        long r3 = 3 * repid[lb + k];
        double x0 = x[r3 * B + b], x1 = x[(r3 + 1) * B + b], x2 = x[(r3 + 2) * B + b];
        a0 += P0[3 * k] * x0;
        a0 += P0[3 * k + 1] * x1;
        a0 += P0[3 * k + 2] * x2;
        a1 += P1[3 * k] * x0;
        a1 += P1[3 * k + 1] * x1;
        a1 += P1[3 * k + 2] * x2;
        a2 += P2[3 * k] * x0;
        a2 += P2[3 * k + 1] * x1;
        a2 += P2[3 * k + 2] * x2;
    }
    long o = 3 * outBlock[i];
    y[o * B + b] = a0;
    y[(o + 1) * B + b] = a1;
    y[(o + 2) * B + b] = a2;
}

void LaunchPrecondB(long n, const long* posLeaf, const long* leafBegin, const long* leafSize, const long* matOff,
                    const double* inv, const long* repid, const long* outBlock, const double* x, double* y, int B,
                    hipStream_t s)
{
    dim3 g((unsigned)((n + 3) / 4), (unsigned)((B + 63) / 64));
    k_precond_b<<<g, 256, 0, s>>>(n, posLeaf, leafBegin, leafSize, matOff, inv, repid, outBlock, x, y, B);
}

// The preconditioner applied as in k_precond_b (same per-row sums), with the
// elements taken grid-stride (workgroup x of kPassWaves wavefronts,
// wavefront w: elements kPassWaves x + w, then every kPassWaves gridDim.x),
// fused with the partial sums of V_k . y for k < nvD over the rows each
// wavefront wrote (partial[(x * nvD + k) * B + b]).
constexpr int kPassColsP = 8;
constexpr int kPassWaves = 16;  // wavefronts per workgroup of the passes (partials are per workgroup)

__global__ void __launch_bounds__(64 * kPassWaves) k_precond_dot_b(long n, const long* __restrict__ posLeaf,
                                                       const long* __restrict__ leafBegin,
                                                       const long* __restrict__ leafSize,
                                                       const long* __restrict__ matOff, const double* __restrict__ inv,
                                                       const long* __restrict__ repid,
                                                       const long* __restrict__ outBlock, const double* __restrict__ x,
                                                       double* y, int B, const double* __restrict__ V, int nvD,
                                                       double* __restrict__ partial, const int* __restrict__ left)
{
    if (left && *left == 0)
        return;
    __shared__ double red[kPassWaves][64];
    const int w = WaveIndex(), lane = threadIdx.x & 63;
    const int b = blockIdx.y * 64 + lane;
    const bool live = b < B;
    const int bb = live ? b : B - 1;
    const long stride = (long)gridDim.x * kPassWaves;
    const long nB = 3 * n * B;
    const int groups = nvD > 0 ? (nvD + kPassColsP - 1) / kPassColsP : 1;
    for (int g = 0; g < groups; ++g) {
        const int c0 = g * kPassColsP;
        const int cn = nvD - c0 < kPassColsP ? nvD - c0 : kPassColsP;
        double acc[kPassColsP];
#pragma unroll
        for (int c = 0; c < kPassColsP; ++c)
            acc[c] = 0.0;
        for (long i = (long)blockIdx.x * kPassWaves + w; i < n; i += stride) {
            const long o = 3 * outBlock[i];
            double a[3];
            if (g == 0) {
                long leaf = posLeaf[i];
                long lb = leafBegin[leaf], c = leafSize[leaf], jl = i - lb, n3 = 3 * c;
                const double* P0 = inv + matOff[leaf] + 3 * jl * n3;
                const double* P1 = P0 + n3;
                const double* P2 = P1 + n3;
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                auto add = [&](long k, double x0, double x1, double x2) {
                    a0 += P0[3 * k] * x0;
                    a0 += P0[3 * k + 1] * x1;
                    a0 += P0[3 * k + 2] * x2;
                    a1 += P1[3 * k] * x0;
                    a1 += P1[3 * k + 1] * x1;
                    a1 += P1[3 * k + 2] * x2;
                    a2 += P2[3 * k] * x0;
                    a2 += P2[3 * k + 1] * x1;
                    a2 += P2[3 * k + 2] * x2;
                };
                // four leaf elements at a time: their state loads issued
                // together, the sums in element order
                long k = 0;
                for (; k + 4 <= c; k += 4) {
                    double xs[4][3];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const long r3 = 3 * repid[lb + k + t];
                        xs[t][0] = x[r3 * B + bb];
                        xs[t][1] = x[(r3 + 1) * B + bb];
                        xs[t][2] = x[(r3 + 2) * B + bb];
                    }
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        add(k + t, xs[t][0], xs[t][1], xs[t][2]);
                }
                for (; k < c; ++k) {
                    const long r3 = 3 * repid[lb + k];
                    add(k, x[r3 * B + bb], x[(r3 + 1) * B + bb], x[(r3 + 2) * B + bb]);
                }
                a[0] = a0;
                a[1] = a1;
                a[2] = a2;
                if (live) {
                    y[o * B + b] = a0;
                    y[(o + 1) * B + b] = a1;
                    y[(o + 2) * B + b] = a2;
                }
            } else {
                for (int m = 0; m < 3; ++m)
                    a[m] = y[(o + m) * B + bb];
            }
            double vv[kPassColsP][3];
#pragma unroll
            for (int c = 0; c < kPassColsP; ++c)
                if (c < cn) {
                    const double* Vc = V + (long)(c0 + c) * nB;
                    for (int m = 0; m < 3; ++m)
                        vv[c][m] = Vc[(o + m) * B + bb];
                }
#pragma unroll
            for (int c = 0; c < kPassColsP; ++c)
                if (c < cn)
                    for (int m = 0; m < 3; ++m)
                        acc[c] += vv[c][m] * a[m];
        }
        for (int c = 0; c < cn; ++c) {
            red[w][lane] = acc[c];
            __syncthreads();
            if (w == 0 && live) {
                double t = 0.0;
#pragma unroll
                for (int q = 0; q < kPassWaves; ++q)
                    t += red[q][lane];
                partial[((long)blockIdx.x * nvD + c0 + c) * B + b] = t;
            }
            __syncthreads();
        }
    }
}

// One pass over the rows of a batch (workgroup x of kPassWaves wavefronts,
// wavefront w: rows kPassWaves x + w, then every kPassWaves gridDim.x):
//   v = u, or rhs - u when rhs is given;  v -= hA[k] V_k for k < nvA;
//   u = v when it changed;
// and the workgroup's partial sums of V_k . v (columns k < nvD) and, with
// norm, of v . v (column nvD), at partial[(x * nout + k) * B + b].  Columns
// are summed kPassCols at a time; the first group shares the update's loop.
constexpr int kPassCols = 8;

__global__ void __launch_bounds__(64 * kPassWaves) k_pass_b(const double* __restrict__ V, long n, int B, double* u,
                                                const double* __restrict__ rhs, const double* __restrict__ hA,
                                                int nvA, int nvD, int norm, double* __restrict__ partial,
                                                const int* __restrict__ left)
{
    if (left && *left == 0)
        return;
    __shared__ double red[kPassWaves][64];
    const int w = WaveIndex(), lane = threadIdx.x & 63;
    const int b = blockIdx.y * 64 + lane;
    const bool live = b < B;
    const int bb = live ? b : B - 1;
    const long stride = (long)gridDim.x * kPassWaves;
    const long nB = n * B;
    const int nout = nvD + (norm ? 1 : 0);
    const bool update = rhs || nvA > 0;
    const int groups = nout > 0 ? (nout + kPassCols - 1) / kPassCols : 1;
    for (int g = 0; g < groups; ++g) {
        const int c0 = g * kPassCols;
        const int cn = nout - c0 < kPassCols ? nout - c0 : kPassCols;
        double acc[kPassCols];
#pragma unroll
        for (int c = 0; c < kPassCols; ++c)
            acc[c] = 0.0;
        // rows r and r + stride together, all loads of both issued before
        // any store, the sums in row order (as one row at a time)
        for (long r = (long)blockIdx.x * kPassWaves + w; r < n; r += 2 * stride) {
            const bool two = r + stride < n;
            const long i = r * B + bb, i2 = (two ? r + stride : r) * B + bb;
            double v = u[i], v2 = u[i2];
            if (g == 0 && update) {
                if (rhs) {
                    v = rhs[i] - v;
                    v2 = rhs[i2] - v2;
                }
                for (int k = 0; k < nvA; ++k) {
                    const double h = hA[(long)k * B + bb];
                    v -= h * V[(long)k * nB + i];
                    v2 -= h * V[(long)k * nB + i2];
                }
                if (live) {
                    u[i] = v;
                    if (two)
                        u[i2] = v2;
                }
            }
            double x[kPassCols], x2[kPassCols];
#pragma unroll
            for (int c = 0; c < kPassCols; ++c)
                if (c < cn) {
                    const int col = c0 + c;
                    x[c] = col < nvD ? V[(long)col * nB + i] : v;
                    x2[c] = col < nvD ? V[(long)col * nB + i2] : v2;
                }
#pragma unroll
            for (int c = 0; c < kPassCols; ++c)
                if (c < cn) {
                    acc[c] += x[c] * v;
                    if (two)
                        acc[c] += x2[c] * v2;
                }
        }
        for (int c = 0; c < cn; ++c) {
            red[w][lane] = acc[c];
            __syncthreads();
            if (w == 0 && live) {
                double t = 0.0;
#pragma unroll
                for (int q = 0; q < kPassWaves; ++q)
                    t += red[q][lane];
                partial[((long)blockIdx.x * nout + c0 + c) * B + b] = t;
            }
            __syncthreads();
        }
    }
}

// Sums of the pass partials (at most kBatchRedBlocks = 16 x 16 blocks),
// column k per workgroup: wavefront w adds blocks w, w + 16, ... with all its
// loads issued together, then the 16 sums are added in wavefront order;
// out[k * B + b] for k < nvD (and addTo[k * B + b] += it when given), the
// square root of the norm column into normOut[b].
static_assert(kBatchRedBlocks <= 16 * 16, "the finish reads at most 16 partials per wavefront");

__global__ void __launch_bounds__(1024) k_finish_b(const double* __restrict__ partial, int nblocks, int nvD, int nout,
                                                   int B, double* __restrict__ out, double* __restrict__ addTo,
                                                   double* __restrict__ normOut, const int* __restrict__ left)
{
    if (left && *left == 0)
        return;
    __shared__ double red[16][64];
    const int k = blockIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.y * 64 + lane;
    const bool live = b < B;
    const int bb = live ? b : B - 1;
    double t[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int x = w + 16 * q;
        t[q] = x < nblocks ? partial[((long)x * nout + k) * B + bb] : 0.0;
    }
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q)
        a += t[q];
    red[w][lane] = a;
    __syncthreads();
    if (w != 0 || !live)
        return;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        s += red[i][lane];
    if (k < nvD) {
        out[(long)k * B + b] = s;
        if (addTo)
            addTo[(long)k * B + b] += s;
    } else {
        normOut[b] = sqrt(s);
    }
}

void LaunchPassB(const double* V, long n, int B, double* u, const double* rhs, const double* hA, int nvA, int nvD,
                 bool norm, double* partial, int nblocks, double* out, double* addTo, double* normOut, hipStream_t s,
                 const int* left)
{
    const unsigned by = (unsigned)((B + 63) / 64);
    k_pass_b<<<dim3((unsigned)nblocks, by), 64 * kPassWaves, 0, s>>>(V, n, B, u, rhs, hA, nvA, nvD, norm ? 1 : 0, partial,
                                                                     left);
    const int nout = nvD + (norm ? 1 : 0);
    if (nout > 0)
        k_finish_b<<<dim3((unsigned)nout, by), 1024, 0, s>>>(partial, nblocks, nvD, nout, B, out, addTo, normOut,
                                                             left);
}

void LaunchPrecondDotB(long n, const long* posLeaf, const long* leafBegin, const long* leafSize, const long* matOff,
                       const double* inv, const long* repid, const long* outBlock, const double* x, double* y, int B,
                       const double* V, int nvD, double* partial, int nblocks, double* h, hipStream_t s,
                       const int* left)
{
    const unsigned by = (unsigned)((B + 63) / 64);
    k_precond_dot_b<<<dim3((unsigned)nblocks, by), 64 * kPassWaves, 0, s>>>(n, posLeaf, leafBegin, leafSize, matOff, inv, repid,
                                                                outBlock, x, y, B, V, nvD, partial, left);
    if (nvD > 0)
        k_finish_b<<<dim3((unsigned)nvD, by), 1024, 0, s>>>(partial, nblocks, nvD, nvD, B, h, nullptr, nullptr, left);
}

void LaunchMultiDotB(const double* V, long n, int nv, const double* u, int B, double* partial, int nblocks,
                     double* h, hipStream_t s)
{
    LaunchPassB(V, n, B, const_cast<double*>(u), nullptr, nullptr, 0, nv, false, partial, nblocks, h, nullptr,
                nullptr, s);
}

void LaunchNormB(const double* x, long n, int B, double* partial, int nblocks, double* out, hipStream_t s)
{
    LaunchPassB(nullptr, n, B, const_cast<double*>(x), nullptr, nullptr, 0, 0, true, partial, nblocks, nullptr,
                nullptr, out, s);
}

__global__ void __launch_bounds__(256) k_scale_b(const double* __restrict__ x, const double* __restrict__ a,
                                                 double* __restrict__ y, long n, int B)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * B)
        return;
    double s = a[idx % B];
    y[idx] = s <= 1.0e-12 ? 0.0 : x[idx] / s;
}

void LaunchScaleB(const double* x, const double* a, double* y, long n, int B, hipStream_t s)
{
    k_scale_b<<<(unsigned)((n * B + 255) / 256), 256, 0, s>>>(x, a, y, n, B);
}

// Givens rotation in QT form for column j of every active load case
// (DBEM1/pre_gmres.cpp:1492-1533); one lane per case.
__global__ void __launch_bounds__(64) k_givens_b(const double* __restrict__ h, double* __restrict__ QT,
                                                 double* __restrict__ R, double* __restrict__ f,
                                                 double* __restrict__ phibar, int* __restrict__ stag,
                                                 const int* __restrict__ active, int j, int restart, int B,
                                                 const int* __restrict__ left)
{
    if (left && *left == 0)
        return;
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || !active[b])
        return;
    const int lq = restart + 1;
#define QQ(a, c) QT[((long)((a)-1) * lq + ((c)-1)) * B + b]
#define RR(a, c) R[((long)((a)-1) * restart + ((c)-1)) * B + b]
#define HH(a) h[(long)(a) * B + b]
    // R(1:j, j) = QT(1:j, 1:j) h(1:j), the products of each row summed in
    // column order; h(1:j) is read once, eight values at a time in
    // registers, so the loads of a row are independent of its sums
    for (int k = 1; k <= j; ++k) {
        double v = 0.0;
        for (int l0 = 1; l0 <= j; l0 += 8) {
            double q[8], hv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int l = l0 + t <= j ? l0 + t : j;
                q[t] = QQ(k, l);
                hv[t] = HH(l);
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (l0 + t <= j)
                    v += q[t] * hv[t];
        }
        RR(k, j) = v;
    }
    double rt = RR(j, j), c, s, temp;
    double hj1 = HH(j + 1);
    if (hj1 == 0.0) {
        c = 1.0;
        s = 0.0;
    } else if (hj1 * hj1 > rt * rt) {
        temp = rt / hj1;
        s = 1.0 / sqrt(1.0 + fabs(temp) * fabs(temp));
        c = -temp * s;
    } else {
        temp = hj1 / rt;
        c = 1.0 / sqrt(1.0 + fabs(temp) * fabs(temp));
        s = -temp * c;
    }
    RR(j, j) = c * rt - s * hj1;
    for (int k = 1; k <= j; ++k) {
        double qk = QQ(j, k);
        QQ(j, k) = c * qk;
        QQ(j + 1, k) = s * qk;
    }
    QQ(j, j + 1) = -s;
    QQ(j + 1, j + 1) = c;
    f[(long)j * B + b] = c * phibar[b];
    phibar[b] = s * phibar[b];
    if (j < restart && f[(long)j * B + b] == 0.0)
        stag[b] = 1;
#undef QQ
#undef RR
#undef HH
}

void LaunchGivensB(double* h, double* QT, double* R, double* f, double* phibar, int* stag, const int* active, int j,
                   int restart, int B, hipStream_t s, const int* left)
{
    k_givens_b<<<(unsigned)((B + 63) / 64), 64, 0, s>>>(h, QT, R, f, phibar, stag, active, j, restart, B, left);
}

// The stopping test of iteration j for every active case (DBEM1/pre_gmres.cpp
// true-residual test, as the host loop made it): its = j; a case leaves when
// its residual norm (scal[B + b]) is <= tol times ||b|| (scal[b]), or with
// flag 3 when the Givens step found stagnation; left = the cases still
// active.  One workgroup.
__global__ void __launch_bounds__(256) k_check_b(const double* __restrict__ scal, double tol,
                                                 const int* __restrict__ stag, int* active, int* its, int* flg,
                                                 int j, int B, int* left)
{
    __shared__ int cnt;
    if (threadIdx.x == 0)
        cnt = 0;
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        if (!active[b])
            continue;
        its[b] = j;
        if (scal[B + b] <= tol * scal[b]) {
            active[b] = 0;
        } else if (stag[b]) {
            active[b] = 0;
            flg[b] = 3;
        } else {
            atomicAdd(&cnt, 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        *left = cnt;
}

void LaunchCheckB(const double* scal, double tol, const int* stag, int* active, int* its, int* flg, int j, int B,
                  int* left, hipStream_t s)
{
    k_check_b<<<1, 256, 0, s>>>(scal, tol, stag, active, its, flg, j, B, left);
}

// W_{j-1} = (V_{j-1} - sum_{l<j} W_{l-1} R(l, j)) / R(j, j) and x += f_j W_{j-1}
// for the active load cases; with Vj, also the next basis vector of every
// case, Vj = u / hs (0 where hs <= 1e-12), which this update does not read.
__global__ void __launch_bounds__(256) k_wupdate_b(const double* __restrict__ V, double* __restrict__ W, long n,
                                                   const double* __restrict__ R, int j, int restart,
                                                   const double* __restrict__ f, double* __restrict__ x,
                                                   const int* __restrict__ active, int B, const double* __restrict__ u,
                                                   const double* __restrict__ hs, double* Vj,
                                                   const int* __restrict__ left)
{
    if (left && *left == 0)
        return;
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * B)
        return;
    int b = (int)(idx % B);
    if (Vj) {
        double sc = hs[b];
        Vj[idx] = sc <= 1.0e-12 ? 0.0 : u[idx] / sc;
    }
    if (!active[b])
        return;
    double rjj = R[((long)(j - 1) * restart + (j - 1)) * B + b];
    double w = V[(long)(j - 1) * n * B + idx] / rjj;
    for (int l = 1; l <= j - 1; ++l)
        w -= W[(long)(l - 1) * n * B + idx] * R[((long)(l - 1) * restart + (j - 1)) * B + b] / rjj;
    W[(long)(j - 1) * n * B + idx] = w;
    x[idx] += f[(long)j * B + b] * w;
}

void LaunchWUpdateB(const double* V, double* W, long n, const double* R, int j, int restart, const double* f,
                    double* x, const int* active, int B, hipStream_t s, const double* u, const double* hs, double* Vj,
                    const int* left)
{
    k_wupdate_b<<<(unsigned)((n * B + 255) / 256), 256, 0, s>>>(V, W, n, R, j, restart, f, x, active, B, u, hs, Vj,
                                                                left);
}

}  // namespace tdbem
