// Launch wrappers of the reverse-time adjoint kernels.
#pragma once

#include "kernels.hpp"

namespace tdbem {

// Transposed history convolution for the adjoint state of step m; lam holds
// lambda_K at lam + K * n3 (K = 1 .. N); z receives 6 values per element.
void LaunchConvT(PairGrid grid, const NodeRefs* refs, const long* off, const int* kb, const int* len,
                 const double* mts, const double* mgs, const double* lam, long n3, int N, int m, int ne, double* z,
                 hipStream_t s);
// Transposed state map into the unknown vector y (y += ...); colU / colT are
// CSR lists of the elements reading each history column.
void LaunchStateT(const NodeRefs* refs, const ElemGeo* geo, const long* colU, const int* colUf, const long* colT,
                  const int* colTf, const double* z, int ne, double* y, hipStream_t s);
void LaunchFdBand(double* dB, const double* B, long e0, long e1, double scale, hipStream_t s);
void LaunchRowDotAcc(const double* lam, const double* v, long nrows, double* acc, hipStream_t s);
void LaunchAddSub(double* v, const double* a, const double* b, long n, hipStream_t s);

}  // namespace tdbem
