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

// Grouped solve: G independent models stacked in one system (block-diagonal
// operator, model g on element rows [ele[g], ele[g+1])), each with the GMRES
// of its own job: its own state, Hessenberg, Givens and reduction partials,
// and the reduction partition its own job uses (nb[g] blocks), so each model
// follows its own iterates bit for bit while every kernel of an iteration is
// one launch for all models.  Every grid gives group g its own run of blocks
// (the prefixes below, read from device memory).
constexpr int kGmGroupsMax = 64;

struct GmresGroupTable {
    int G;
    long ele[kGmGroupsMax + 1];  // element row bounds
    int nb[kGmGroupsMax];        // reduction blocks per group
    int nbPre[kGmGroupsMax + 1];  // prefix of nb
    int rbPre[kGmGroupsMax + 1];  // prefix of ceil(3 ne_g / 256): blocks over scalar rows
};

struct GmresGroupArgs {
    long ne, restart;
    const int *rp, *col;
    const double* val;
    const long *posLeaf, *leafBegin, *leafSize, *matOff, *repid, *outBlock;
    const double* inv;
    // global vectors (leading dimension 3 ne)
    double *V, *W, *u, *u2, *vh, *x;
    // per group g at g times the stride: h, h2, f (hs = restart + 2), QT
    // (qs = (restart + 1)^2), R (rs = restart^2), part (ps = kGmBlocksMax *
    // (restart + 1)), phibar, stag, st, active (1 each)
    double *h, *h2, *QT, *R, *f, *phibar, *part;
    long hs, qs, rs, ps;
    int* stag;
    const int* active;  // 0: the group takes no step (its state opens done)
    GmState* st;
};

// t: the host copy of the table, dt: its device copy (grid sizes come from t)
void LaunchGmresBeginGroups(const GmresGroupArgs& a, const GmresGroupTable& t, const GmresGroupTable* dt,
                            const double* b, double tol, hipStream_t s);
void LaunchGmresIterationGroups(const GmresGroupArgs& a, const GmresGroupTable& t, const GmresGroupTable* dt,
                                const double* b, hipStream_t s);

}  // namespace tdbem
