// Device-driven FP64 GMRES of one time step (gmres.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

namespace tdbem {

// Solve state in device memory: written by the opening kernel and the
// per-iteration convergence check, read by every kernel of an iteration.
struct GmState {
    double bnorm, tolb, normr;
    int j;     // iterations completed
    int done;  // 1 once the solve has stopped
    int flag;  // 0 converged, 1 restart length reached, 3 stagnation
    int pad;
};

// Longest restart the device-driven solve takes (coefficients of one Arnoldi
// step are held in LDS); longer restarts use the host-driven loop.
constexpr long kGmDeviceMaxRestart = 256;

// Operands and workspace of one solve (all device pointers, 1-based h).
struct GmresDevArgs {
    long ne, restart;
    int nblocks;  // reduction blocks (the partition of the host-driven kernels)
    const int *rp, *col;
    const double* val;
    const long *posLeaf, *leafBegin, *leafSize, *matOff, *repid, *outBlock;
    const double* inv;
    // explicit block CCSR preconditioner (used instead of the leaf inverses when prp is set)
    const int *prp, *pcol;
    const double* pval;
    double *V, *W, *u, *u2, *vh, *h, *h2, *QT, *R, *f, *phibar, *part, *x;
    long nQT, nR;
    int* stag;
    GmState* st;
};

// ||b||, x = 0 and the first Krylov vector; the state says done when b = 0 or
// the zero start already meets the tolerance.
void LaunchGmresBegin(const GmresDevArgs& a, const double* b, double tol, hipStream_t s);
// One Arnoldi / Givens / W-update / true-residual iteration with its
// convergence test; no work once the state is done.
void LaunchGmresIteration(const GmresDevArgs& a, const double* b, hipStream_t s);

}  // namespace tdbem
