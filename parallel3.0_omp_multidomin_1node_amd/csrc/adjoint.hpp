// Launch wrappers of the reverse-time adjoint kernels.
#pragma once

#include "kernels.hpp"

namespace tdbem {

// Steps per block of the time-blocked transposed history convolution.
constexpr int kConvTTile = 8;

// Transposed history convolution for the adjoint state of step m over the
// lags j <= jmax; lam holds lambda_K at lam + K * n3 (K = 1 .. N); z receives
// 6 values per element (plus zadd when given).
void LaunchConvT(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                 const double* mts, const double* mgs, const double* lam, long n3, int N, int m, int jmax,
                 const double* zadd, int ne, double* z, hipStream_t s);
// The lags j <= jmax of the same term, one source pair per thread (for the
// short lags of the time-blocked sweep).
void LaunchConvTNear(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                     const double* mts, const double* mgs, const double* lam, long n3, int N, int m, int jmax,
                     const double* zadd, int ne, double* z, hipStream_t s);
// Phi[s][K] = lambda_{K-1} + 2 lambda_K + lambda_{K+1} for K in [K0, K1]
// (stride values of K per row).
void LaunchPhi(const double* lam, long n3, int ne, int N, int K0, int K1, int stride, double* phi, hipStream_t s);
// Far part (lags > kConvTTile) of the transposed history term for the steps
// mlo .. mlo + kConvTTile - 1, into zf[i][6 ne].
void LaunchConvTFar(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                    const double* mts, const double* mgs, const double* phi, int stride, int N, int mlo, int ne,
                    double* zf, hipStream_t s);
// Transposed state map into the unknown vector y (y += ...); colU / colT are
// CSR lists of the elements reading each history column.
void LaunchStateT(const NodeRefs* refs, const ElemGeo* geo, const long* colU, const int* colUf, const long* colT,
                  const int* colTf, const double* z, int ne, double* y, hipStream_t s);
void LaunchFdBand(double* dB, const double* B, long e0, long e1, double scale, hipStream_t s);
void LaunchRowDotAcc(const double* lam, const double* v, long nrows, double* acc, hipStream_t s);
void LaunchAddSub(double* v, const double* a, const double* b, long n, hipStream_t s);
// F[c][t] = g_{t+1} + 2 g_t + g_{t-1} per history column (t = 0 .. N).
void LaunchFilterSeries(const double* g, long ncol, int nstore, int N, double* F, hipStream_t s);
// History part of one parameter's gradient, per (pair, chunk) into part;
// then acc[s] += the partials of row s.
void LaunchCorrFd(PairGrid grid, const NodeRefs* refs, const long* coff, const long* off, const int* kb,
                  const int* len, long npairs, long chunks, const double* mp, const double* gp, const double* mm,
                  const double* gm, const double* scale, const double* lam, long n3, const double* FU,
                  const double* FT, const double* gU, const double* gT, int nstore, int N, double* part,
                  hipStream_t s);
// Per band entry: (CU transposed, symmetric sums of CT), 15 doubles.
void LaunchCorrStore(PairGrid grid, const NodeRefs* refs, const long* coff, const long* off, const int* kb,
                     const int* len, long npairs, long chunks, const double* lam, long n3, const double* FU,
                     const double* FT, const double* gU, const double* gT, int nstore, int N, double* C,
                     hipStream_t s);
void LaunchRowChunks(PairGrid grid, const NodeRefs* refs, const long* coff, const double* part, int ne, double* acc,
                     hipStream_t s);
// kn[K][...] = the known vector of step K for K = 0 .. nstore - 1 (nk3 values each).
void LaunchKnownAll(const NodeRefs* refs, const double* bd, int ne, int nstore, long nk3, double* kn, hipStream_t s);
// Step-0 part of one parameter's gradient from the two perturbed step-0
// passes over the near pairs in list: acc[s] += per row s (rowPtr) the
// contraction with the adjoint and forward histories.
void LaunchStep0Grad(const long* list, long n, PairGrid grid, const NodeRefs* refs, const double* outP,
                     const int* flagP, const double* outM, const int* flagM, const double* inv2, const double* lam,
                     const double* X, long n3, const double* kn, long nk3, int N, double* gpair,
                     const long* rowPtr, int ne, double* acc, hipStream_t s);

}  // namespace tdbem
