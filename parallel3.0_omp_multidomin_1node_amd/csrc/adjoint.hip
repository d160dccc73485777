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
                                                long n3, int N, int m, double* __restrict__ z)
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
            if (m + j - 1 > N)
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
            z[6 * (long)f + k] = red[k][0];
}

void LaunchConvT(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                 const double* mts, const double* mgs, const double* lam, long n3, int N, int m, int ne, double* z,
                 hipStream_t s)
{
    k_conv_t<<<ne, 256, 0, s>>>(grid, refs, off, kb, len, mts, mgs, lam, n3, N, m, z);
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

}  // namespace tdbem
