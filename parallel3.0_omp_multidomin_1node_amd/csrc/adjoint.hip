// HIP kernels of the reverse-time adjoint of the FP64 time loop.
//
// The forward step K solves A x_K = KnownM k_K + h_K with the history term
// h_K = sum_j sum_z c_z [MGS_j gT_m - MTS_j gU_m], m = K - j + 1 - z in
// [1, K - 1], c = (1, 2, 1) (DBEM1/multidomain.cpp:3201-3225), and the state
// columns gU, gT are linear in x_m.  The adjoint lambda_K solves
// A^T lambda_K = dJ/dx_K + sum_{K' > K} (dh_K' / dx_K)^T lambda_K', whose
// second term is the transposed history convolution below followed by the
// transposed state map.
#include "adjoint.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace tdbem {

// z[6 f .. 6 f + 5] = ( -sum MTS_j(s, f)^T lt , sum MGS_j(s, f) lt ) over the
// sources s of f's domain and the lags j of each pair, lt = sum_z c_z
// lambda_{m + j - 1 + z}[s] over m + 1 <= m + j - 1 + z <= N.  One workgroup
// per field element, four wavefronts over the sources, one lag per lane.
__global__ void __launch_bounds__(256) k_conv_t(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                const long* __restrict__ off, const int* __restrict__ kb,
                                                const int* __restrict__ len, const double* __restrict__ mts,
                                                const double* __restrict__ mgs, const double* __restrict__ lam,
                                                long n3, int N, int m, int jmax, const double* __restrict__ zadd,
                                                double* __restrict__ z)
{
    int f = blockIdx.x;
    int d = refs[f].domain;
    int nd = grid.eleCount[d], eb = grid.eleBegin[d];
    int fl = f - eb;
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = wave; q < nd; q += 4) {
        long p = grid.pairBase[d] + (long)q * nd + fl;
        int L = len[p];
        if (L == 0)
            continue;
        int k0 = kb[p];
        long s = eb + q;
        long base = off[p];
        for (int t = lane; t < L; t += 64) {
            int j = k0 + t;
            if (m + j - 1 > N || j > jmax)
                break;
            double lt0 = 0.0, lt1 = 0.0, lt2 = 0.0;
            for (int zz = 0; zz < 3; ++zz) {
                int Kp = m + j - 1 + zz;
                if (Kp >= m + 1 && Kp <= N) {
                    double w = zz == 1 ? 2.0 : 1.0;
                    const double* l = lam + (long)Kp * n3 + 3 * s;
                    lt0 += w * l[0];
                    lt1 += w * l[1];
                    lt2 += w * l[2];
                }
            }
            long e = base + t;
            const double* M = mts + 9 * e;
            const double* G = mgs + 6 * e;
            a[0] -= M[0] * lt0 + M[3] * lt1 + M[6] * lt2;
            a[1] -= M[1] * lt0 + M[4] * lt1 + M[7] * lt2;
            a[2] -= M[2] * lt0 + M[5] * lt1 + M[8] * lt2;
            a[3] += G[0] * lt0 + G[1] * lt1 + G[2] * lt2;
            a[4] += G[1] * lt0 + G[3] * lt1 + G[4] * lt2;
            a[5] += G[2] * lt0 + G[4] * lt1 + G[5] * lt2;
        }
    }
    __shared__ double red[6][256];
    for (int k = 0; k < 6; ++k)
        red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k)
                red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k)
            z[6 * (long)f + k] = red[k][0] + (zadd ? zadd[6 * (long)f + k] : 0.0);
}

void LaunchConvT(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                 const double* mts, const double* mgs, const double* lam, long n3, int N, int m, int jmax,
                 const double* zadd, int ne, double* z, hipStream_t s)
{
    k_conv_t<<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, lam, n3, N, m, jmax, zadd, z);
}

// Short-lag part (lags j <= jmax) of the transposed history term of step m,
// plus zadd: one workgroup per field element, one source pair per thread
// (each pair has at most jmax short lags).
__global__ void __launch_bounds__(256) k_conv_t_near(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                     const long* __restrict__ off, const int* __restrict__ kb,
                                                     const int* __restrict__ len, const double* __restrict__ mts,
                                                     const double* __restrict__ mgs, const double* __restrict__ lam,
                                                     long n3, int N, int m, int jmax,
                                                     const double* __restrict__ zadd, double* __restrict__ z)
{
    int f = blockIdx.x;
    int d = refs[f].domain;
    int nd = grid.eleCount[d], eb = grid.eleBegin[d];
    int fl = f - eb;
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nd; q += 256) {
        long p = grid.pairBase[d] + (long)q * nd + fl;
        int L = len[p], k0 = kb[p];
        int j1 = k0 + L - 1;
        if (j1 > jmax)
            j1 = jmax;
        if (j1 > N - m + 1)
            j1 = N - m + 1;
        long s = eb + q;
        for (int j = k0; j <= j1; ++j) {
            double lt0 = 0.0, lt1 = 0.0, lt2 = 0.0;
            for (int zz = 0; zz < 3; ++zz) {
                int Kp = m + j - 1 + zz;
                if (Kp >= m + 1 && Kp <= N) {
                    double w = zz == 1 ? 2.0 : 1.0;
                    const double* l = lam + (long)Kp * n3 + 3 * s;
                    lt0 += w * l[0];
                    lt1 += w * l[1];
                    lt2 += w * l[2];
                }
            }
            long e = off[p] + (j - k0);
            const double* M = mts + 9 * e;
            const double* G = mgs + 6 * e;
            a[0] -= M[0] * lt0 + M[3] * lt1 + M[6] * lt2;
            a[1] -= M[1] * lt0 + M[4] * lt1 + M[7] * lt2;
            a[2] -= M[2] * lt0 + M[5] * lt1 + M[8] * lt2;
            a[3] += G[0] * lt0 + G[1] * lt1 + G[2] * lt2;
            a[4] += G[1] * lt0 + G[3] * lt1 + G[4] * lt2;
            a[5] += G[2] * lt0 + G[4] * lt1 + G[5] * lt2;
        }
    }
    __shared__ double red[4][6];
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int k = 0; k < 6; ++k) {
        double v = a[k];
        for (int w = 32; w > 0; w >>= 1)
            v += __shfl_xor(v, w);
        if (lane == 0)
            red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        z[6 * (long)f + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k] + (zadd ? zadd[6 * (long)f + k] : 0.0);
    }
}

void LaunchConvTNear(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                     const double* mts, const double* mgs, const double* lam, long n3, int N, int m, int jmax,
                     const double* zadd, int ne, double* z, hipStream_t s)
{
    k_conv_t_near<<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, lam, n3, N, m, jmax, zadd, z);
}

// Filtered adjoint series of every source row: Phi[s][K] = lambda_{K-1} +
// 2 lambda_K + lambda_{K+1} (lambda_K = 0 outside 1 .. N) for K in [K0, K1],
// so the transposed history term of lag j at step m reads Phi[s][m + j]
// whenever m + j - 1 > m (every lag j >= 2).
__global__ void __launch_bounds__(256) k_phi(const double* __restrict__ lam, long n3, int ne, int N, int K0, int K1,
                                             int stride, double* __restrict__ phi)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int nk = K1 - K0 + 1;
    if (i >= (long)ne * nk * 3)
        return;
    int s = (int)(i / ((long)nk * 3));
    int r = (int)(i - (long)s * nk * 3);
    int K = K0 + r / 3, c = r % 3;
    double v = 0.0;
    if (K - 1 >= 1 && K - 1 <= N)
        v += lam[(long)(K - 1) * n3 + 3 * (long)s + c];
    if (K >= 1 && K <= N)
        v += 2.0 * lam[(long)K * n3 + 3 * (long)s + c];
    if (K + 1 >= 1 && K + 1 <= N)
        v += lam[(long)(K + 1) * n3 + 3 * (long)s + c];
    phi[((long)s * stride + K) * 3 + c] = v;
}

void LaunchPhi(const double* lam, long n3, int ne, int N, int K0, int K1, int stride, double* phi, hipStream_t s)
{
    long n = (long)ne * (K1 - K0 + 1) * 3;
    if (n > 0)
        k_phi<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(lam, n3, ne, N, K0, K1, stride, phi);
}

// Far part of the transposed history term for a block of T backward steps
// m = mlo .. mlo + T - 1: every lag j > T of every pair reads
// Phi[s][m + j] >= m + T + 1 > the block, all known when the block starts.
// zf[i][6 f .. 6 f + 5] = ( -sum MTS^T Phi , sum MGS Phi ) for m = mlo + i.
// One workgroup per field element, four wavefronts over its sources, one lag
// per lane; each band entry is read once per block.
template <int T>
__global__ void __launch_bounds__(256) k_conv_t_far(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                    const long* __restrict__ off, const int* __restrict__ kb,
                                                    const int* __restrict__ len, const double* __restrict__ mts,
                                                    const double* __restrict__ mgs, const double* __restrict__ phi,
                                                    int stride, int N, int mlo, long ne, double* __restrict__ zf)
{
    int f = blockIdx.x;
    int d = refs[f].domain;
    int nd = grid.eleCount[d], eb = grid.eleBegin[d];
    int fl = f - eb;
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double a[T][6];
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int k = 0; k < 6; ++k)
            a[i][k] = 0.0;
    for (int q = wave; q < nd; q += 4) {
        long p = grid.pairBase[d] + (long)q * nd + fl;
        int L = len[p];
        if (L == 0)
            continue;
        int k0 = kb[p];
        long base = off[p];
        const double* ph = phi + (long)(eb + q) * stride * 3;
        int t = lane + (k0 > T ? 0 : T + 1 - k0);
        for (; t < L; t += 64) {
            int j = k0 + t;
            if (mlo + j > N + 1)
                break;
            long e = base + t;
            const double* M = mts + 9 * e;
            const double* G = mgs + 6 * e;
            double m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3], m4 = M[4], m5 = M[5], m6 = M[6], m7 = M[7], m8 = M[8];
            double g0 = G[0], g1 = G[1], g2 = G[2], g3 = G[3], g4 = G[4], g5 = G[5];
            const double* pw = ph + (long)(mlo + j) * 3;
#pragma unroll
            for (int i = 0; i < T; ++i) {
                double l0 = pw[3 * i], l1 = pw[3 * i + 1], l2 = pw[3 * i + 2];
                a[i][0] -= fma(m0, l0, fma(m3, l1, m6 * l2));
                a[i][1] -= fma(m1, l0, fma(m4, l1, m7 * l2));
                a[i][2] -= fma(m2, l0, fma(m5, l1, m8 * l2));
                a[i][3] += fma(g0, l0, fma(g1, l1, g2 * l2));
                a[i][4] += fma(g1, l0, fma(g3, l1, g4 * l2));
                a[i][5] += fma(g2, l0, fma(g4, l1, g5 * l2));
            }
        }
    }
    __shared__ double red[4][T * 6];
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double v = a[i][k];
            for (int w = 32; w > 0; w >>= 1)
                v += __shfl_xor(v, w);
            if (lane == 0)
                red[wave][i * 6 + k] = v;
        }
    __syncthreads();
    if ((int)threadIdx.x < T * 6) {
        int i = threadIdx.x / 6, k = threadIdx.x - 6 * i;
        zf[(long)i * 6 * ne + 6 * (long)f + k] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    }
}

void LaunchConvTFar(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                    const double* mts, const double* mgs, const double* phi, int stride, int N, int mlo, int ne,
                    double* zf, hipStream_t s)
{
    k_conv_t_far<kConvTTile><<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, phi, stride, N, mlo, ne, zf);
}

// Transposed state map: the column sums of z go back through
// gU = hu_sign * F^T (u_sign T x_u) (and gT alike) to the unknown blocks
// (transpose of k_state / k_scatter).  y += ...
__global__ void __launch_bounds__(256) k_state_t(const NodeRefs* __restrict__ refs, const ElemGeo* __restrict__ geo,
                                                 const long* __restrict__ colU, const int* __restrict__ colUf,
                                                 const long* __restrict__ colT, const int* __restrict__ colTf,
                                                 const double* __restrict__ z, int ne, double* __restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne)
        return;
    const NodeRefs& r = refs[i];
    if (r.surface == 1 && !r.ref_side)
        return;
    const double* F = geo[i].trans;
    const double* T = r.itf_T;
    for (int pass = 0; pass < 2; ++pass) {
        bool known = pass == 0 ? r.u_known : r.t_known;
        if (known)
            continue;
        int col = pass == 0 ? r.hu_block : r.ht_block;
        const long* ptr = pass == 0 ? colU : colT;
        const int* mem = pass == 0 ? colUf : colTf;
        double hs = pass == 0 ? r.hu_sign : r.ht_sign;
        double vs = pass == 0 ? r.u_sign : r.t_sign;
        int blk = pass == 0 ? r.u_block : r.t_block;
        double Z[3] = {0.0, 0.0, 0.0};
        for (long k = ptr[col]; k < ptr[col + 1]; ++k) {
            const double* zf = z + 6 * (long)mem[k] + 3 * pass;
            Z[0] += zf[0];
            Z[1] += zf[1];
            Z[2] += zf[2];
        }
        double l[3] = {hs * Z[0], hs * Z[1], hs * Z[2]};
        double ub[3];
        for (int rr = 0; rr < 3; ++rr)
            ub[rr] = F[rr * 3 + 0] * l[0] + F[rr * 3 + 1] * l[1] + F[rr * 3 + 2] * l[2];
        for (int k = 0; k < 3; ++k)
            y[3 * (long)blk + k] += vs * (T[0 * 3 + k] * ub[0] + T[1 * 3 + k] * ub[1] + T[2 * 3 + k] * ub[2]);
    }
}

void LaunchStateT(const NodeRefs* refs, const ElemGeo* geo, const long* colU, const int* colUf, const long* colT,
                  const int* colTf, const double* z, int ne, double* y, hipStream_t s)
{
    k_state_t<<<(ne + 255) / 256, 256, 0, s>>>(refs, geo, colU, colUf, colT, colTf, z, ne, y);
}

// dB = (dB - B) * scale over [e0, e1) values (central difference of two
// quadratures at perturbed materials).
__global__ void __launch_bounds__(256) k_fd_band(double* __restrict__ dB, const double* __restrict__ B, long e0,
                                                 long e1, double scale)
{
    long e = e0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < e1)
        dB[e] = (dB[e] - B[e]) * scale;
}

void LaunchFdBand(double* dB, const double* B, long e0, long e1, double scale, hipStream_t s)
{
    long n = e1 - e0;
    if (n > 0)
        k_fd_band<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dB, B, e0, e1, scale);
}

// acc[s] += lambda[s] . v[s] per block row.
__global__ void __launch_bounds__(256) k_rowdot_acc(const double* __restrict__ lam, const double* __restrict__ v,
                                                    long nrows, double* __restrict__ acc)
{
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nrows)
        acc[s] += lam[3 * s] * v[3 * s] + lam[3 * s + 1] * v[3 * s + 1] + lam[3 * s + 2] * v[3 * s + 2];
}

void LaunchRowDotAcc(const double* lam, const double* v, long nrows, double* acc, hipStream_t s)
{
    k_rowdot_acc<<<(unsigned)((nrows + 255) / 256), 256, 0, s>>>(lam, v, nrows, acc);
}

// v += a - b
__global__ void __launch_bounds__(256) k_addsub(double* __restrict__ v, const double* __restrict__ a,
                                                const double* __restrict__ b, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        v[i] += a[i] - b[i];
}

void LaunchAddSub(double* v, const double* a, const double* b, long n, hipStream_t s)
{
    k_addsub<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(v, a, b, n);
}

// 1-2-1 filtered state series of every history column: F[c][t] = g_{t+1} +
// 2 g_t + g_{t-1} over the stored steps 1 .. N (zero outside), t = 0 .. N,
// so the forward history term of lag j at step K reads F[c][K - j] (minus
// g_K when j = 1: the forward sum only reads states before step K).
__global__ void __launch_bounds__(256) k_filter_series(const double* __restrict__ g, long ncol, int nstore, int N,
                                                       double* __restrict__ F)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long per = (long)(N + 1) * 3;
    if (i >= ncol * per)
        return;
    long c = i / per;
    int r = (int)(i - c * per), t = r / 3, k = r - 3 * t;
    const double* sg = g + c * (long)nstore * 3;
    double v = 0.0;
    if (t + 1 <= N)
        v += sg[3 * (t + 1) + k];
    if (t >= 1)
        v += 2.0 * sg[3 * t + k];
    if (t >= 2)
        v += sg[3 * (t - 1) + k];
    F[i] = v;
}

void LaunchFilterSeries(const double* g, long ncol, int nstore, int N, double* F, hipStream_t s)
{
    long n = ncol * (long)(N + 1) * 3;
    if (n > 0)
        k_filter_series<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, ncol, nstore, N, F);
}

// History part of the material gradient for one parameter: for every band
// entry e = (s, f, j) of the operator difference dM = (M+ - M-) scale[d],
//   sum_{K >= j} lambda_K[s]^T (dMGS_e cT_{K-j} - dMTS_e cU_{K-j})
// with c the filtered series above; this is sum_K lambda_K . h_K(dM) of the
// forward history term, summed entry by entry instead of step by step.  One
// wavefront per (pair, chunk of 64 lags), one lag per lane; lambda_K[s] is
// the same for every lane.  The wave's sum goes to part[chunk].
__global__ void __launch_bounds__(64) k_corr_fd(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                const long* __restrict__ coff, const long* __restrict__ off,
                                                const int* __restrict__ kb, const int* __restrict__ len, long npairs,
                                                long c0, long c1, const double* __restrict__ mp,
                                                const double* __restrict__ gp, const double* __restrict__ mm,
                                                const double* __restrict__ gm, const double* __restrict__ scale,
                                                const double* __restrict__ lam, long n3, const double* __restrict__ FU,
                                                const double* __restrict__ FT, const double* __restrict__ gU,
                                                const double* __restrict__ gT, int nstore, int N,
                                                double* __restrict__ part)
{
    long wv = c0 + (long)blockIdx.x;
    if (wv >= c1)
        return;
    int lane = threadIdx.x;
    long lo = 0, hi = npairs;
    while (hi - lo > 1) {
        long mid = (lo + hi) >> 1;
        if (coff[mid] <= wv)
            lo = mid;
        else
            hi = mid;
    }
    long p = lo;
    int s, f, d;
    DecodePairOf(grid, p, s, f, d);
    int t0 = (int)(wv - coff[p]) * 64;
    int t = t0 + lane;
    bool in = t < len[p];
    int j = kb[p] + t, jlo = kb[p] + t0;
    const double* fu = FU + (long)refs[f].hu_block * (N + 1) * 3;
    const double* ft = FT + (long)refs[f].ht_block * (N + 1) * 3;
    double cu[9], ct[9];
    for (int i = 0; i < 9; ++i) {
        cu[i] = 0.0;
        ct[i] = 0.0;
    }
    for (int K = jlo; K <= N; ++K) {
        const double* l = lam + (long)K * n3 + 3 * (long)s;
        double l0 = l[0], l1 = l[1], l2 = l[2];
        if (in && K >= j) {
            int tt = K - j;
            double u[3] = {fu[3 * tt], fu[3 * tt + 1], fu[3 * tt + 2]};
            double v[3] = {ft[3 * tt], ft[3 * tt + 1], ft[3 * tt + 2]};
            if (j == 1) {
                const double* gu = gU + ((long)refs[f].hu_block * nstore + K) * 3;
                const double* gt = gT + ((long)refs[f].ht_block * nstore + K) * 3;
                for (int c = 0; c < 3; ++c) {
                    u[c] -= gu[c];
                    v[c] -= gt[c];
                }
            }
            for (int c = 0; c < 3; ++c) {
                cu[c] = fma(l0, u[c], cu[c]);
                cu[3 + c] = fma(l1, u[c], cu[3 + c]);
                cu[6 + c] = fma(l2, u[c], cu[6 + c]);
                ct[c] = fma(l0, v[c], ct[c]);
                ct[3 + c] = fma(l1, v[c], ct[3 + c]);
                ct[6 + c] = fma(l2, v[c], ct[6 + c]);
            }
        }
    }
    double val = 0.0;
    if (in) {
        long e = off[p] + t;
        const double* Mp = mp + 9 * e;
        const double* Mm = mm + 9 * e;
        const double* Gp = gp + 6 * e;
        const double* Gm = gm + 6 * e;
        double a = 0.0;
        for (int i = 0; i < 9; ++i)
            a = fma(Mp[i] - Mm[i], cu[i], a);
        double dg[6];
        for (int i = 0; i < 6; ++i)
            dg[i] = Gp[i] - Gm[i];
        double b = dg[0] * ct[0] + dg[3] * ct[4] + dg[5] * ct[8] + dg[1] * (ct[1] + ct[3]) +
                   dg[2] * (ct[2] + ct[6]) + dg[4] * (ct[5] + ct[7]);
        val = (b - a) * scale[d];
    }
    for (int w = 32; w > 0; w >>= 1)
        val += __shfl_xor(val, w);
    if (lane == 0)
        part[wv] = val;
}

void LaunchCorrFd(PairGrid grid, const NodeRefs* refs, const long* coff, const long* off, const int* kb,
                  const int* len, long npairs, long chunks, const double* mp, const double* gp, const double* mm,
                  const double* gm, const double* scale, const double* lam, long n3, const double* FU,
                  const double* FT, const double* gU, const double* gT, int nstore, int N, double* part,
                  hipStream_t s)
{
    const long step = 1L << 20;
    for (long c0 = 0; c0 < chunks; c0 += step) {
        long c1 = std::min(chunks, c0 + step);
        k_corr_fd<<<(unsigned)(c1 - c0), 64, 0, s>>>(grid, refs, coff, off, kb, len, npairs, c0, c1, mp, gp, mm, gm,
                                                    scale, lam, n3, FU, FT, gU, gT, nstore, N, part);
    }
}

// Per band entry e = (s, f, j) the correlation of the adjoint with the
// filtered states the entry multiplies in the forward history term:
//   CU[r][c] = sum_{K >= j} lambda_K[s][r] cU_{K-j}[c],  CT likewise,
// stored as C[15 e ..] = (CU transposed: C[3a+b] = CU[b][a], then the
// symmetric sums of CT: 00, 01+10, 02+20, 11, 12+21, 22), so that
// sum_K lambda_K . h_K(dM) = sum_e (dMGS_e . CT_e - dMTS_e : CU_e).
__global__ void __launch_bounds__(64) k_corr_store(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                   const long* __restrict__ coff, const long* __restrict__ off,
                                                   const int* __restrict__ kb, const int* __restrict__ len,
                                                   long npairs, long c0, long c1, const double* __restrict__ lam,
                                                   long n3, const double* __restrict__ FU,
                                                   const double* __restrict__ FT, const double* __restrict__ gU,
                                                   const double* __restrict__ gT, int nstore, int N,
                                                   double* __restrict__ C)
{
    // steps K are taken kCorrSeg at a time: lambda_K[s] of the segment and
    // the filtered-state rows its lags read are staged in LDS
    constexpr int kSeg = 256;
    constexpr int kRowsMax = kSeg + 63;
    __shared__ double sl[kSeg * 3];
    __shared__ double sx[kRowsMax * 6];
    long wv = c0 + (long)blockIdx.x;
    if (wv >= c1)
        return;
    int lane = threadIdx.x;
    long lo = 0, hi = npairs;
    while (hi - lo > 1) {
        long mid = (lo + hi) >> 1;
        if (coff[mid] <= wv)
            lo = mid;
        else
            hi = mid;
    }
    long p = lo;
    int s, f, d;
    DecodePairOf(grid, p, s, f, d);
    int t0 = (int)(wv - coff[p]) * 64;
    int L = len[p];
    int t = t0 + lane;
    bool in = t < L;
    int j = kb[p] + t, jlo = kb[p] + t0;
    int jhi = kb[p] + (t0 + 63 < L - 1 ? t0 + 63 : L - 1);
    const double* fu = FU + (long)refs[f].hu_block * (N + 1) * 3;
    const double* ft = FT + (long)refs[f].ht_block * (N + 1) * 3;
    const double* gu1 = gU + (long)refs[f].hu_block * nstore * 3;
    const double* gt1 = gT + (long)refs[f].ht_block * nstore * 3;
    double cu[9], ct[9];
    for (int i = 0; i < 9; ++i) {
        cu[i] = 0.0;
        ct[i] = 0.0;
    }
    for (int K1 = jlo; K1 <= N; K1 += kSeg) {
        int K2 = K1 + kSeg - 1 < N ? K1 + kSeg - 1 : N;
        // lambda_K[s], K = K1 .. K2
        for (int idx = lane; idx < (K2 - K1 + 1) * 3; idx += 64)
            sl[idx] = lam[(long)(K1 + idx / 3) * n3 + 3 * (long)s + idx % 3];
        // rows tt = K - j of the chunk's lags: rlo = K1 - jhi .. K2 - jlo, clipped to >= 0
        int rlo = K1 - jhi, rhi = K2 - jlo;
        int r0 = rlo > 0 ? rlo : 0;
        for (int idx = lane; idx < (rhi - r0 + 1) * 6; idx += 64) {
            int rr = r0 + idx / 6, c = idx % 6;
            sx[idx] = c < 3 ? fu[3 * (long)rr + c] : ft[3 * (long)rr + (c - 3)];
        }
        __syncthreads();
        if (in) {
            int Kb = j > K1 ? j : K1;
            for (int K = Kb; K <= K2; ++K) {
                const double* l = sl + 3 * (K - K1);
                const double* x = sx + 6 * (K - j - r0);
                double l0 = l[0], l1 = l[1], l2 = l[2];
                double u[3] = {x[0], x[1], x[2]};
                double v[3] = {x[3], x[4], x[5]};
                if (j == 1) {
                    for (int c = 0; c < 3; ++c) {
                        u[c] -= gu1[3 * (long)K + c];
                        v[c] -= gt1[3 * (long)K + c];
                    }
                }
                for (int c = 0; c < 3; ++c) {
                    cu[c] = fma(l0, u[c], cu[c]);
                    cu[3 + c] = fma(l1, u[c], cu[3 + c]);
                    cu[6 + c] = fma(l2, u[c], cu[6 + c]);
                    ct[c] = fma(l0, v[c], ct[c]);
                    ct[3 + c] = fma(l1, v[c], ct[3 + c]);
                    ct[6 + c] = fma(l2, v[c], ct[6 + c]);
                }
            }
        }
        __syncthreads();
    }
    if (!in)
        return;
    double* o = C + 15 * (off[p] + t);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            o[3 * a + b] = cu[3 * b + a];
    o[9] = ct[0];
    o[10] = ct[1] + ct[3];
    o[11] = ct[2] + ct[6];
    o[12] = ct[4];
    o[13] = ct[5] + ct[7];
    o[14] = ct[8];
}

void LaunchCorrStore(PairGrid grid, const NodeRefs* refs, const long* coff, const long* off, const int* kb,
                     const int* len, long npairs, long chunks, const double* lam, long n3, const double* FU,
                     const double* FT, const double* gU, const double* gT, int nstore, int N, double* C,
                     hipStream_t s)
{
    const long step = 1L << 20;
    for (long c0 = 0; c0 < chunks; c0 += step) {
        long c1 = std::min(chunks, c0 + step);
        k_corr_store<<<(unsigned)(c1 - c0), 64, 0, s>>>(grid, refs, coff, off, kb, len, npairs, c0, c1, lam, n3, FU,
                                                        FT, gU, gT, nstore, N, C);
    }
}

// Known vectors of every step: kn[K][3 blk + c] = sign * bd[K][3 i + c]
// (the map of k_known for all K at once).
__global__ void __launch_bounds__(256) k_known_all(const NodeRefs* __restrict__ refs, const double* __restrict__ bd,
                                                   int ne, int nstore, long nk3, double* __restrict__ kn)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ne * nstore)
        return;
    int K = (int)(idx / ne), i = (int)(idx - (long)K * ne);
    const NodeRefs& r = refs[i];
    if (r.surface == 1)
        return;
    int blk = r.bcid == 123 ? r.u_block : r.t_block;
    int known = r.bcid == 123 ? r.u_known : r.t_known;
    double sg = r.bcid == 123 ? r.u_sign : r.t_sign;
    if (!known)
        return;
    for (int c = 0; c < 3; ++c)
        kn[(long)K * nk3 + 3 * (long)blk + c] = sg * bd[((long)K * ne + i) * 3 + c];
}

void LaunchKnownAll(const NodeRefs* refs, const double* bd, int ne, int nstore, long nk3, double* kn, hipStream_t s)
{
    long n = (long)ne * nstore;
    k_known_all<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(refs, bd, ne, nstore, nk3, kn);
}

// Step-0 part of one parameter's gradient per near pair i: the difference of
// the two perturbed step-0 passes (outP - outM, scaled by inv2[s]) placed as
// the unknown / known builders place the blocks (stored transposed, times the
// DOF sign; DBEM1/multidomain.cpp:1907-1937,2286-2321), contracted with
// sum_K lambda_K[s] (x) x_K[col] (unknown columns, -dA x_K) or k_K[col]
// (known columns, +dKnownM k_K).  A block counts when either pass has it.
__global__ void __launch_bounds__(256) k_step0_grad(const long* __restrict__ list, long n, PairGrid grid,
                                                    const NodeRefs* __restrict__ refs,
                                                    const double* __restrict__ outP, const int* __restrict__ flagP,
                                                    const double* __restrict__ outM, const int* __restrict__ flagM,
                                                    const double* __restrict__ inv2, const double* __restrict__ lam,
                                                    const double* __restrict__ X, long n3,
                                                    const double* __restrict__ kn, long nk3, int N,
                                                    double* __restrict__ gpair)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int s, f, d;
    DecodePairOf(grid, list[i], s, f, d);
    const NodeRefs& r = refs[f];
    double g = 0.0;
    for (int which = 0; which < 2; ++which) {
        if (!(flagP[2 * i + which] || flagM[2 * i + which]))
            continue;
        int col;
        double sgn;
        bool unknownCol;
        if (r.surface) {
            col = which == 0 ? r.u_block : r.t_block;
            sgn = which == 0 ? r.u_sign : r.t_sign;
            unknownCol = true;
        } else if (which == 0) {
            bool useT = r.bcid == 123;
            if (useT ? r.t_known : r.u_known)
                continue;
            col = useT ? r.t_block : r.u_block;
            sgn = useT ? r.t_sign : r.u_sign;
            unknownCol = true;
        } else {
            bool useU = r.bcid == 123;
            if (!(useU ? r.u_known : r.t_known))
                continue;
            col = useU ? r.u_block : r.t_block;
            sgn = useU ? r.u_sign : r.t_sign;
            unknownCol = false;
        }
        double db[9];
        for (int q = 0; q < 9; ++q)
            db[q] = (outP[18 * i + 9 * which + q] - outM[18 * i + 9 * which + q]) * inv2[s];
        // C[r][c] = sum_K lambda_K[s][r] v_K[col][c]
        double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int K = 1; K <= N; ++K) {
            const double* l = lam + (long)K * n3 + 3 * (long)s;
            const double* v = unknownCol ? X + (long)K * n3 + 3 * (long)col : kn + (long)K * nk3 + 3 * (long)col;
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c)
                    C[3 * a + c] = fma(l[a], v[c], C[3 * a + c]);
        }
        // stored block M[a][c] = sgn * db[3 c + a]
        double t = 0.0;
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c)
                t += db[3 * c + a] * C[3 * a + c];
        g += (unknownCol ? -sgn : sgn) * t;
    }
    gpair[i] = g;
}

// acc[s] += sum of gpair over the near pairs of row s (rowPtr, in list order)
__global__ void __launch_bounds__(256) k_pair_rows(const double* __restrict__ gpair, const long* __restrict__ rowPtr,
                                                   int ne, double* __restrict__ acc)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ne)
        return;
    double a = 0.0;
    for (long k = rowPtr[s]; k < rowPtr[s + 1]; ++k)
        a += gpair[k];
    acc[s] += a;
}

void LaunchStep0Grad(const long* list, long n, PairGrid grid, const NodeRefs* refs, const double* outP,
                     const int* flagP, const double* outM, const int* flagM, const double* inv2, const double* lam,
                     const double* X, long n3, const double* kn, long nk3, int N, double* gpair,
                     const long* rowPtr, int ne, double* acc, hipStream_t s)
{
    if (n > 0)
        k_step0_grad<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(list, n, grid, refs, outP, flagP, outM, flagM, inv2,
                                                                  lam, X, n3, kn, nk3, N, gpair);
    k_pair_rows<<<(ne + 255) / 256, 256, 0, s>>>(gpair, rowPtr, ne, acc);
}

// acc[s] += sum of the chunk partials of row s (its pairs are contiguous).
__global__ void __launch_bounds__(256) k_row_chunks(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                    const long* __restrict__ coff, const double* __restrict__ part,
                                                    int ne, double* __restrict__ acc)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ne)
        return;
    int d = refs[s].domain;
    long nd = grid.eleCount[d];
    long p0 = grid.pairBase[d] + (long)(s - grid.eleBegin[d]) * nd;
    double a = 0.0;
    for (long c = coff[p0]; c < coff[p0 + nd]; ++c)
        a += part[c];
    acc[s] += a;
}

void LaunchRowChunks(PairGrid grid, const NodeRefs* refs, const long* coff, const double* part, int ne, double* acc,
                     hipStream_t s)
{
    k_row_chunks<<<(ne + 255) / 256, 256, 0, s>>>(grid, refs, coff, part, ne, acc);
}

}  // namespace tdbem
