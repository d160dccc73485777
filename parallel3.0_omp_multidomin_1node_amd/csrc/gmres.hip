// Device-driven GMRES of one time step (DBEM1/pre_gmres.cpp:1369-1605 with
// maxit = 1 and x0 = 0), FP64.  The iteration counter, the convergence test
// and the stopping flags live in device memory (GmState), so one iteration is
// a fixed sequence of kernels with fixed arguments: it is captured once in a
// hipGraph and replayed, and the host reads the state back once per time
// step.  Every kernel returns at once when the state says the solve is done,
// so iterations launched past convergence do no work.  The arithmetic, the
// reduction partition and the summation orders are those of the host-driven
// kernels in kernels.hip, so both drivers give the same iterates bit for bit.
#include "gmres.hpp"

namespace tdbem {

namespace {

// a null state never reads as done (the kernels that open a solve)
__device__ __forceinline__ bool Done(const GmState* st)
{
    return st && __builtin_amdgcn_readfirstlane(st->done) != 0;
}

// the next iteration's column count j (1-based)
__device__ __forceinline__ int NextJ(const GmState* st)
{
    return __builtin_amdgcn_readfirstlane(st->j) + 1;
}

// sum of the block partials of column k in block order, loads eight at a time
// (the order of k_finish)
__device__ double SumPartials(const double* __restrict__ partial, int nblocks, int nv, int k)
{
    double a = 0.0;
    int b = 0;
    for (; b + 8 <= nblocks; b += 8) {
        double t[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            t[i] = partial[(long)(b + i) * nv + k];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            a += t[i];
    }
    for (; b < nblocks; ++b)
        a += partial[(long)b * nv + k];
    return a;
}

}  // namespace

// bnorm = ||b|| from the partials of b.b; x0 = 0 so r = b and ||r|| = bnorm.
// Done at once for b = 0 or when bnorm <= tol * bnorm.
__global__ void k_gm_begin(const double* __restrict__ partial, int nblocks, double tol, GmState* st, int* stag)
{
    if (threadIdx.x != 0)
        return;
    double bn = sqrt(SumPartials(partial, nblocks, 1, 0));
    st->bnorm = bn;
    st->normr = bn;
    st->tolb = tol * bn;
    st->j = 0;
    st->flag = 0;
    st->done = (bn == 0.0 || bn <= tol * bn) ? 1 : 0;
    stag[0] = 0;
}

__global__ void __launch_bounds__(256) k_gm_fill(const GmState* st, double* __restrict__ a, long n, double* __restrict__ b,
                                                 long m)
{
    if (Done(st))
        return;
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        a[r] = 0.0;
    if (r < m)
        b[r] = 0.0;
}

// Component c of block row i of y = A x: the sum over the row's blocks in
// CCSR order, each block's three products in the order of the host-driven
// SpMV, so a thread per (row, component) gives its bits exactly.
__device__ __forceinline__ double RowComponent(const int* __restrict__ rp, const int* __restrict__ col,
                                               const double* __restrict__ val, const double* __restrict__ x, long i,
                                               int c)
{
    double a = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const double* v = val + 9 * (long)j + 3 * c;
        const double* xc = x + 3 * (long)col[j];
        a += v[0] * xc[0] + v[1] * xc[1] + v[2] * xc[2];
    }
    return a;
}

// y = A x with x = V(:, st->j + colOff) (colOff < 0: x itself), or the
// residual y = b - A x when b is given; one thread per (block row, component)
__global__ void __launch_bounds__(256) k_gm_spmv(const GmState* st, long nrows, const int* __restrict__ rp,
                                                 const int* __restrict__ col, const double* __restrict__ val,
                                                 const double* __restrict__ x, long ldx, int colOff,
                                                 const double* __restrict__ b, double* __restrict__ y)
{
    if (Done(st))
        return;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nrows)
        return;
    if (colOff >= 0)
        x += (long)(NextJ(st) - 1 + colOff) * ldx;
    const long i = t / 3;
    double a = RowComponent(rp, col, val, x, i, (int)(t - 3 * i));
    if (b)
        a = b[t] - a;
    y[t] = a;
}

__global__ void __launch_bounds__(256) k_gm_precond(const GmState* st, long n, const long* __restrict__ posLeaf,
                                                    const long* __restrict__ leafBegin,
                                                    const long* __restrict__ leafSize, const long* __restrict__ matOff,
                                                    const double* __restrict__ inv, const long* __restrict__ repid,
                                                    const long* __restrict__ outBlock, const double* __restrict__ x,
                                                    double* __restrict__ y)
{
    if (Done(st))
        return;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n)
        return;
    long i = t / 3;
    int mrow = (int)(t % 3);
    long leaf = posLeaf[i];
    long b = leafBegin[leaf], c = leafSize[leaf], jl = i - b, n3 = 3 * c;
    const double* P = inv + matOff[leaf] + (3 * jl + mrow) * n3;
    double acc = 0.0;
    for (long k = 0; k < c; ++k) {
        const double* xc = x + 3 * repid[b + k];
        acc += P[3 * k] * xc[0];
        acc += P[3 * k + 1] * xc[1];
        acc += P[3 * k + 2] * xc[2];
    }
    y[3 * outBlock[i] + mrow] = acc;
}

// block partials of V(:, 0 .. nv-1) . u with nv = j (all) or 1 (V = u: norm)
__global__ void __launch_bounds__(256) k_gm_multidot(const GmState* st, const double* __restrict__ V, long ldv,
                                                     int all, const double* __restrict__ u, long n,
                                                     double* __restrict__ partial)
{
    if (Done(st))
        return;
    int nv = all ? NextJ(st) : 1;
    __shared__ double red[256];
    for (int k = 0; k < nv; ++k) {
        double a = 0.0;
        for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
            a += V[k * ldv + r] * u[r];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
                red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partial[(long)blockIdx.x * nv + k] = red[0];
        __syncthreads();
    }
}

// u -= V(:, 0 .. j-1) h(0 .. j-1), the coefficients h(k) summed from the
// multi-dot partials by every block (the order of k_finish); block 0 stores
// them at hout and, when given, adds them to htot
__global__ void __launch_bounds__(256) k_gm_multiaxpy(const GmState* st, const double* __restrict__ V, long ldv,
                                                      const double* __restrict__ partial, int nblocks,
                                                      double* __restrict__ hout, double* __restrict__ u, long n,
                                                      double* htot)
{
    if (Done(st))
        return;
    __shared__ double h[256];
    int nv = NextJ(st);
    int tid = threadIdx.x;
    if (tid < nv) {
        double v = SumPartials(partial, nblocks, nv, tid);
        h[tid] = v;
        if (blockIdx.x == 0) {
            hout[tid] = v;
            if (htot)
                htot[tid] += v;
        }
    }
    __syncthreads();
    long r = (long)blockIdx.x * blockDim.x + tid;
    if (r >= n)
        return;
    double a = u[r];
    for (int k = 0; k < nv; ++k)
        a -= h[k] * V[k * ldv + r];
    u[r] = a;
}

// h[1] = ||u|| from the partials (every block, the order of k_finish),
// V(:,0) = u / h[1]; block 0 also starts the Givens sequence: QT(1,1) = 1,
// phibar = h[1]
__global__ void __launch_bounds__(256) k_gm_open(const GmState* st, const double* __restrict__ partial, int nblocks,
                                                 const double* __restrict__ u, double* __restrict__ V, long n,
                                                 double* h, double* QT, double* phibar)
{
    if (Done(st))
        return;
    __shared__ double hs;
    if (threadIdx.x == 0)
        hs = sqrt(SumPartials(partial, nblocks, 1, 0));
    __syncthreads();
    double s = hs;
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        V[r] = s <= 1.0e-12 ? 0.0 : u[r] / s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        h[1] = s;
        QT[0] = 1.0;
        phibar[0] = s;
    }
}

// Givens rotation in QT form for column j (DBEM1/pre_gmres.cpp:1492-1533)
__global__ void __launch_bounds__(256) k_gm_givens(const GmState* st, double* h, double* QT, double* R, double* f,
                                                   double* phibar, int* stag, int restart,
                                                   const double* __restrict__ partial, int nblocks)
{
    if (Done(st))
        return;
    const int j = NextJ(st);
    // h[j+1] = ||u|| from the partials of u.u
    if (threadIdx.x == 0)
        h[j + 1] = sqrt(SumPartials(partial, nblocks, 1, 0));
    __syncthreads();
    const int lq = restart + 1;
#define QQ(a, b) QT[((a)-1) * lq + ((b)-1)]
#define RR(a, b) R[((a)-1) * restart + ((b)-1)]
    for (int k = 1 + (int)threadIdx.x; k <= j; k += blockDim.x) {
        double v = 0.0;
        for (int l = 1; l <= j; ++l)
            v += QQ(k, l) * h[l];
        RR(k, j) = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double rt = RR(j, j), c, s, temp;
        double hj1 = h[j + 1];
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
        f[j] = c * phibar[0];
        phibar[0] = s * phibar[0];
        if (j < restart && f[j] == 0.0)
            stag[0] = 1;
    }
#undef QQ
#undef RR
}

// W(:,j) = (V(:,j) - sum_{l<j} W(:,l) R(l,j)) / R(j,j);  x += f_j W(:,j)
// (1-based columns), together with the next Krylov vector
// V(:, j+1) = u / h[j+1] (tiny denominators give 0, as the opening's scale)
__global__ void __launch_bounds__(256) k_gm_wupdate(const GmState* st, double* __restrict__ V,
                                                    double* __restrict__ W, long n, const double* __restrict__ R,
                                                    int restart, const double* __restrict__ f, double* __restrict__ x,
                                                    const double* __restrict__ u, const double* __restrict__ h)
{
    if (Done(st))
        return;
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    int j = NextJ(st);
    double sc = h[j + 1];
    V[(long)j * n + k] = sc <= 1.0e-12 ? 0.0 : u[k] / sc;
    double rjj = R[(j - 1) * restart + (j - 1)];
    double w = V[(long)(j - 1) * n + k] / rjj;
    for (int l = 1; l <= j - 1; ++l)
        w -= W[(long)(l - 1) * n + k] * R[(l - 1) * restart + (j - 1)] / rjj;
    W[(long)(j - 1) * n + k] = w;
    x[k] += f[j] * w;
}

// ||r|| of iteration j from the partials of r.r, then the reference's tests:
// converged (flag 0), stagnation (3) or the restart length reached (1)
__global__ void k_gm_check(GmState* st, const double* __restrict__ partial, int nblocks, const int* stag,
                           int restart)
{
    if (Done(st) || threadIdx.x != 0)
        return;
    int j = st->j + 1;
    double nr = sqrt(SumPartials(partial, nblocks, 1, 0));
    st->normr = nr;
    st->j = j;
    if (nr <= st->tolb) {
        st->flag = 0;
        st->done = 1;
    } else if (stag[0]) {
        st->flag = 3;
        st->done = 1;
    } else if (j >= restart) {
        st->flag = 1;
        st->done = 1;
    }
}

namespace {

unsigned Blocks(long n)
{
    return (unsigned)((n + 255) / 256);
}

// y = M^-1 x: the leaf inverses, or the explicit preconditioner matrix
void LaunchPre(const GmresDevArgs& a, const double* x, double* y, hipStream_t s)
{
    if (a.prp)
        k_gm_spmv<<<Blocks(3 * a.ne), 256, 0, s>>>(a.st, a.ne, a.prp, a.pcol, a.pval, x, 3 * a.ne, -1, nullptr, y);
    else
        k_gm_precond<<<Blocks(3 * a.ne), 256, 0, s>>>(a.st, a.ne, a.posLeaf, a.leafBegin, a.leafSize, a.matOff,
                                                      a.inv, a.repid, a.outBlock, x, y);
}

}  // namespace

void LaunchGmresBegin(const GmresDevArgs& a, const double* b, double tol, hipStream_t s)
{
    const long n = 3 * a.ne;
    // ||b|| partials (the partition of LaunchNorm), x = 0
    k_gm_multidot<<<a.nblocks, 256, 0, s>>>(nullptr, b, n, 0, b, n, a.part);
    k_gm_fill<<<Blocks(n), 256, 0, s>>>(nullptr, a.x, n, a.x, 0);
    k_gm_begin<<<1, 64, 0, s>>>(a.part, a.nblocks, tol, a.st, a.stag);
    k_gm_fill<<<Blocks(std::max(a.nQT, a.nR)), 256, 0, s>>>(a.st, a.QT, a.nQT, a.R, a.nR);
    // u = M^-1 r with r = b; h[1] = ||u||; V(:,0) = u / h[1]
    LaunchPre(a, b, a.u, s);
    k_gm_multidot<<<a.nblocks, 256, 0, s>>>(a.st, a.u, n, 0, a.u, n, a.part);
    k_gm_open<<<Blocks(n), 256, 0, s>>>(a.st, a.part, a.nblocks, a.u, a.V, n, a.h, a.QT, a.phibar);
}

void LaunchGmresIteration(const GmresDevArgs& a, const double* b, hipStream_t s)
{
    const long n = 3 * a.ne;
    k_gm_spmv<<<Blocks(3 * a.ne), 256, 0, s>>>(a.st, a.ne, a.rp, a.col, a.val, a.V, n, 0, nullptr, a.u2);
    LaunchPre(a, a.u2, a.u, s);
    // classical Gram-Schmidt twice
    k_gm_multidot<<<a.nblocks, 256, 0, s>>>(a.st, a.V, n, 1, a.u, n, a.part);
    k_gm_multiaxpy<<<Blocks(n), 256, 0, s>>>(a.st, a.V, n, a.part, a.nblocks, a.h + 1, a.u, n, nullptr);
    k_gm_multidot<<<a.nblocks, 256, 0, s>>>(a.st, a.V, n, 1, a.u, n, a.part);
    k_gm_multiaxpy<<<Blocks(n), 256, 0, s>>>(a.st, a.V, n, a.part, a.nblocks, a.h2, a.u, n, a.h + 1);
    // h[j+1] = ||u|| with the Givens rotation; V(:,j) = u / h[j+1] with the W update and x
    k_gm_multidot<<<a.nblocks, 256, 0, s>>>(a.st, a.u, n, 0, a.u, n, a.part);
    k_gm_givens<<<1, 256, 0, s>>>(a.st, a.h, a.QT, a.R, a.f, a.phibar, a.stag, (int)a.restart, a.part, a.nblocks);
    k_gm_wupdate<<<Blocks(n), 256, 0, s>>>(a.st, a.V, a.W, n, a.R, (int)a.restart, a.f, a.x, a.u, a.h);
    // true residual b - A x and its norm
    k_gm_spmv<<<Blocks(3 * a.ne), 256, 0, s>>>(a.st, a.ne, a.rp, a.col, a.val, a.x, n, -1, b, a.vh);
    k_gm_multidot<<<a.nblocks, 256, 0, s>>>(a.st, a.vh, n, 0, a.vh, n, a.part);
    k_gm_check<<<1, 64, 0, s>>>(a.st, a.part, a.nblocks, a.stag, (int)a.restart);
}

// ---------------------------------------------------------------------------
// Grouped solve (GmresGroupTable): the kernels above with group g's operands.
// Every grid gives each group its own run of blocks, so a block's group is
// uniform and each group's rows, partials and reductions are exactly those of
// its own job (row r of group g is row r - 3 ele[g] there).
// ---------------------------------------------------------------------------
namespace {

// group of this block in a grid where group g owns blocks [pre[g], pre[g+1])
__device__ __forceinline__ int BlockGroup(const int* __restrict__ pre, int G)
{
    const int b = (int)blockIdx.x;
    int lo = 0, hi = G;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (__builtin_amdgcn_readfirstlane(pre[mid]) <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

}  // namespace

__global__ void __launch_bounds__(256) k_gg_multidot(const GmState* st, const GmresGroupTable* __restrict__ tab,
                                                     const double* __restrict__ V, long ldv, int all,
                                                     const double* __restrict__ u, double* __restrict__ partial,
                                                     long ps)
{
    const int g = BlockGroup(tab->nbPre, tab->G);
    const GmState* sg = st ? st + g : nullptr;
    if (Done(sg))
        return;
    const int nb = tab->nb[g], lb = (int)blockIdx.x - tab->nbPre[g];
    const long r0 = 3 * tab->ele[g], n = 3 * (tab->ele[g + 1] - tab->ele[g]);
    const int nv = all ? NextJ(sg) : 1;
    double* pg = partial + (long)g * ps;
    __shared__ double red[256];
    for (int k = 0; k < nv; ++k) {
        double a = 0.0;
        for (long r = (long)lb * blockDim.x + threadIdx.x; r < n; r += (long)nb * blockDim.x)
            a += V[k * ldv + r0 + r] * u[r0 + r];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
                red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            pg[(long)lb * nv + k] = red[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_gg_fillqr(const GmState* st, double* __restrict__ QT, long qs,
                                                   double* __restrict__ R, long rs)
{
    const int g = (int)blockIdx.y;
    if (Done(st + g))
        return;
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < qs)
        QT[(long)g * qs + r] = 0.0;
    if (r < rs)
        R[(long)g * rs + r] = 0.0;
}

// the opening of every group (k_gm_begin); a group not active opens done
__global__ void k_gg_begin(const double* __restrict__ partial, long ps, const GmresGroupTable* __restrict__ tab,
                           double tol, GmState* st, int* stag, const int* __restrict__ active)
{
    const int g = (int)blockIdx.x;
    if (threadIdx.x != 0)
        return;
    double bn = sqrt(SumPartials(partial + (long)g * ps, tab->nb[g], 1, 0));
    GmState* sg = st + g;
    sg->bnorm = bn;
    sg->normr = bn;
    sg->tolb = tol * bn;
    sg->j = 0;
    sg->flag = 0;
    sg->done = (!active[g] || bn == 0.0 || bn <= tol * bn) ? 1 : 0;
    stag[g] = 0;
}

__global__ void __launch_bounds__(256) k_gg_spmv(const GmState* st, const GmresGroupTable* __restrict__ tab,
                                                 const int* __restrict__ rp, const int* __restrict__ col,
                                                 const double* __restrict__ val, const double* __restrict__ x,
                                                 long ldx, int colOff, const double* __restrict__ b,
                                                 double* __restrict__ y)
{
    const int g = BlockGroup(tab->rbPre, tab->G);
    if (Done(st + g))
        return;
    long t = 3 * tab->ele[g] + (long)((int)blockIdx.x - tab->rbPre[g]) * blockDim.x + threadIdx.x;
    if (t >= 3 * tab->ele[g + 1])
        return;
    if (colOff >= 0)
        x += (long)(NextJ(st + g) - 1 + colOff) * ldx;
    const long i = t / 3;
    double a = RowComponent(rp, col, val, x, i, (int)(t - 3 * i));
    if (b)
        a = b[t] - a;
    y[t] = a;
}

__global__ void __launch_bounds__(256) k_gg_precond(const GmState* st, const GmresGroupTable* __restrict__ tab,
                                                    const long* __restrict__ posLeaf,
                                                    const long* __restrict__ leafBegin,
                                                    const long* __restrict__ leafSize, const long* __restrict__ matOff,
                                                    const double* __restrict__ inv, const long* __restrict__ repid,
                                                    const long* __restrict__ outBlock, const double* __restrict__ x,
                                                    double* __restrict__ y)
{
    const int g = BlockGroup(tab->rbPre, tab->G);
    if (Done(st + g))
        return;
    long t = 3 * tab->ele[g] + (long)((int)blockIdx.x - tab->rbPre[g]) * blockDim.x + threadIdx.x;
    if (t >= 3 * tab->ele[g + 1])
        return;
    long i = t / 3;
    int mrow = (int)(t % 3);
    long leaf = posLeaf[i];
    long b = leafBegin[leaf], c = leafSize[leaf], jl = i - b, n3 = 3 * c;
    const double* P = inv + matOff[leaf] + (3 * jl + mrow) * n3;
    double acc = 0.0;
    for (long k = 0; k < c; ++k) {
        const double* xc = x + 3 * repid[b + k];
        acc += P[3 * k] * xc[0];
        acc += P[3 * k + 1] * xc[1];
        acc += P[3 * k + 2] * xc[2];
    }
    y[3 * outBlock[i] + mrow] = acc;
}

__global__ void __launch_bounds__(256) k_gg_multiaxpy(const GmState* st, const GmresGroupTable* __restrict__ tab,
                                                      const double* __restrict__ V, long ldv,
                                                      const double* __restrict__ partial, long ps,
                                                      double* __restrict__ hout, long hs, double* __restrict__ u,
                                                      double* htot)
{
    const int g = BlockGroup(tab->rbPre, tab->G);
    if (Done(st + g))
        return;
    __shared__ double h[256];
    const int nv = NextJ(st + g);
    const int tid = threadIdx.x, lb = (int)blockIdx.x - tab->rbPre[g];
    if (tid < nv) {
        double v = SumPartials(partial + (long)g * ps, tab->nb[g], nv, tid);
        h[tid] = v;
        if (lb == 0) {
            hout[(long)g * hs + tid] = v;
            if (htot)
                htot[(long)g * hs + tid] += v;
        }
    }
    __syncthreads();
    long r = 3 * tab->ele[g] + (long)lb * blockDim.x + tid;
    if (r >= 3 * tab->ele[g + 1])
        return;
    double a = u[r];
    for (int k = 0; k < nv; ++k)
        a -= h[k] * V[k * ldv + r];
    u[r] = a;
}

__global__ void __launch_bounds__(256) k_gg_open(const GmState* st, const GmresGroupTable* __restrict__ tab,
                                                 const double* __restrict__ partial, long ps,
                                                 const double* __restrict__ u, double* __restrict__ V, double* h,
                                                 long hs, double* QT, long qs, double* phibar)
{
    const int g = BlockGroup(tab->rbPre, tab->G);
    if (Done(st + g))
        return;
    __shared__ double hsh;
    if (threadIdx.x == 0)
        hsh = sqrt(SumPartials(partial + (long)g * ps, tab->nb[g], 1, 0));
    __syncthreads();
    double s = hsh;
    const int lb = (int)blockIdx.x - tab->rbPre[g];
    long r = 3 * tab->ele[g] + (long)lb * blockDim.x + threadIdx.x;
    if (r < 3 * tab->ele[g + 1])
        V[r] = s <= 1.0e-12 ? 0.0 : u[r] / s;
    if (lb == 0 && threadIdx.x == 0) {
        h[(long)g * hs + 1] = s;
        QT[(long)g * qs] = 1.0;
        phibar[g] = s;
    }
}

__global__ void __launch_bounds__(256) k_gg_givens(const GmState* st, double* h, long hs, double* QT, long qs,
                                                   double* R, long rs, double* f, double* phibar, int* stag,
                                                   int restart, const double* __restrict__ partial, long ps,
                                                   const GmresGroupTable* __restrict__ tab)
{
    const int g = (int)blockIdx.x;
    if (Done(st + g))
        return;
    const int j = NextJ(st + g);
    h += (long)g * hs;
    QT += (long)g * qs;
    R += (long)g * rs;
    f += (long)g * hs;
    if (threadIdx.x == 0)
        h[j + 1] = sqrt(SumPartials(partial + (long)g * ps, tab->nb[g], 1, 0));
    __syncthreads();
    const int lq = restart + 1;
#define QQ(a, b) QT[((a)-1) * lq + ((b)-1)]
#define RR(a, b) R[((a)-1) * restart + ((b)-1)]
    for (int k = 1 + (int)threadIdx.x; k <= j; k += blockDim.x) {
        double v = 0.0;
        for (int l = 1; l <= j; ++l)
            v += QQ(k, l) * h[l];
        RR(k, j) = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double rt = RR(j, j), c, s, temp;
        double hj1 = h[j + 1];
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
        f[j] = c * phibar[g];
        phibar[g] = s * phibar[g];
        if (j < restart && f[j] == 0.0)
            stag[g] = 1;
    }
#undef QQ
#undef RR
}

__global__ void __launch_bounds__(256) k_gg_wupdate(const GmState* st, const GmresGroupTable* __restrict__ tab,
                                                    double* __restrict__ V, double* __restrict__ W, long n,
                                                    const double* __restrict__ R, long rs, int restart,
                                                    const double* __restrict__ f, long hs, double* __restrict__ x,
                                                    const double* __restrict__ u, const double* __restrict__ h)
{
    const int g = BlockGroup(tab->rbPre, tab->G);
    if (Done(st + g))
        return;
    long k = 3 * tab->ele[g] + (long)((int)blockIdx.x - tab->rbPre[g]) * blockDim.x + threadIdx.x;
    if (k >= 3 * tab->ele[g + 1])
        return;
    R += (long)g * rs;
    f += (long)g * hs;
    h += (long)g * hs;
    int j = NextJ(st + g);
    double sc = h[j + 1];
    V[(long)j * n + k] = sc <= 1.0e-12 ? 0.0 : u[k] / sc;
    double rjj = R[(j - 1) * restart + (j - 1)];
    double w = V[(long)(j - 1) * n + k] / rjj;
    for (int l = 1; l <= j - 1; ++l)
        w -= W[(long)(l - 1) * n + k] * R[(l - 1) * restart + (j - 1)] / rjj;
    W[(long)(j - 1) * n + k] = w;
    x[k] += f[j] * w;
}

__global__ void k_gg_check(GmState* st, const double* __restrict__ partial, long ps,
                           const GmresGroupTable* __restrict__ tab, const int* stag, int restart)
{
    const int g = (int)blockIdx.x;
    if (Done(st + g) || threadIdx.x != 0)
        return;
    GmState* sg = st + g;
    int j = sg->j + 1;
    double nr = sqrt(SumPartials(partial + (long)g * ps, tab->nb[g], 1, 0));
    sg->normr = nr;
    sg->j = j;
    if (nr <= sg->tolb) {
        sg->flag = 0;
        sg->done = 1;
    } else if (stag[g]) {
        sg->flag = 3;
        sg->done = 1;
    } else if (j >= restart) {
        sg->flag = 1;
        sg->done = 1;
    }
}

void LaunchGmresBeginGroups(const GmresGroupArgs& a, const GmresGroupTable& t, const GmresGroupTable* dt,
                            const double* b, double tol, hipStream_t s)
{
    const long n = 3 * a.ne;
    const int G = t.G;
    const unsigned nbT = (unsigned)t.nbPre[G], rbT = (unsigned)t.rbPre[G];
    // ||b|| partials of every group, x = 0
    k_gg_multidot<<<nbT, 256, 0, s>>>(nullptr, dt, b, n, 0, b, a.part, a.ps);
    k_gm_fill<<<Blocks(n), 256, 0, s>>>(nullptr, a.x, n, a.x, 0);
    k_gg_begin<<<(unsigned)G, 64, 0, s>>>(a.part, a.ps, dt, tol, a.st, a.stag, a.active);
    k_gg_fillqr<<<dim3(Blocks(std::max(a.qs, a.rs)), (unsigned)G), 256, 0, s>>>(a.st, a.QT, a.qs, a.R, a.rs);
    // u = M^-1 b; h[1] = ||u||; V(:,0) = u / h[1]
    k_gg_precond<<<rbT, 256, 0, s>>>(a.st, dt, a.posLeaf, a.leafBegin, a.leafSize, a.matOff, a.inv, a.repid,
                                     a.outBlock, b, a.u);
    k_gg_multidot<<<nbT, 256, 0, s>>>(a.st, dt, a.u, n, 0, a.u, a.part, a.ps);
    k_gg_open<<<rbT, 256, 0, s>>>(a.st, dt, a.part, a.ps, a.u, a.V, a.h, a.hs, a.QT, a.qs, a.phibar);
}

void LaunchGmresIterationGroups(const GmresGroupArgs& a, const GmresGroupTable& t, const GmresGroupTable* dt,
                                const double* b, hipStream_t s)
{
    const long n = 3 * a.ne;
    const int G = t.G;
    const unsigned nbT = (unsigned)t.nbPre[G], rbT = (unsigned)t.rbPre[G];
    k_gg_spmv<<<rbT, 256, 0, s>>>(a.st, dt, a.rp, a.col, a.val, a.V, n, 0, nullptr, a.u2);
    k_gg_precond<<<rbT, 256, 0, s>>>(a.st, dt, a.posLeaf, a.leafBegin, a.leafSize, a.matOff, a.inv, a.repid,
                                     a.outBlock, a.u2, a.u);
    // classical Gram-Schmidt twice
    k_gg_multidot<<<nbT, 256, 0, s>>>(a.st, dt, a.V, n, 1, a.u, a.part, a.ps);
    k_gg_multiaxpy<<<rbT, 256, 0, s>>>(a.st, dt, a.V, n, a.part, a.ps, a.h + 1, a.hs, a.u, nullptr);
    k_gg_multidot<<<nbT, 256, 0, s>>>(a.st, dt, a.V, n, 1, a.u, a.part, a.ps);
    k_gg_multiaxpy<<<rbT, 256, 0, s>>>(a.st, dt, a.V, n, a.part, a.ps, a.h2, a.hs, a.u, a.h + 1);
    // h[j+1] = ||u|| with the Givens rotation; V(:,j) = u / h[j+1] with the W update and x
    k_gg_multidot<<<nbT, 256, 0, s>>>(a.st, dt, a.u, n, 0, a.u, a.part, a.ps);
    k_gg_givens<<<(unsigned)G, 256, 0, s>>>(a.st, a.h, a.hs, a.QT, a.qs, a.R, a.rs, a.f, a.phibar, a.stag,
                                            (int)a.restart, a.part, a.ps, dt);
    k_gg_wupdate<<<rbT, 256, 0, s>>>(a.st, dt, a.V, a.W, n, a.R, a.rs, (int)a.restart, a.f, a.hs, a.x, a.u, a.h);
    // true residual b - A x and its norm
    k_gg_spmv<<<rbT, 256, 0, s>>>(a.st, dt, a.rp, a.col, a.val, a.x, n, -1, b, a.vh);
    k_gg_multidot<<<nbT, 256, 0, s>>>(a.st, dt, a.vh, n, 0, a.vh, a.part, a.ps);
    k_gg_check<<<(unsigned)G, 64, 0, s>>>(a.st, a.part, a.ps, dt, a.stag, (int)a.restart);
}

}  // namespace tdbem
