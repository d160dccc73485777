// Reverse-time adjoint of the FP64 multi-domain time loop and the gradient of
// a receiver misfit with respect to the per-domain materials (E, v, rho).
//
// Forward (DBEM1/multidomain.cpp:3815-3893): A x_K = KnownM k_K + h_K(x_<K).
// Adjoint: A^T lambda_K = dJ/dx_K + sum_{K' > K} (dh_K'/dx_K)^T lambda_K',
// solved backward in time by GMRES on A^T with the transposed leaf
// preconditioner.  Gradient: dJ/dtheta = sum_K lambda_K^T (dKnownM k_K
// - dA x_K + dh_K), where the operator derivatives are central differences
// over the materials of the same quadrature kernels with the band structure
// held fixed (every block of domain d depends only on domain d's material,
// so all domains are perturbed at once, one parameter at a time).
#include "adjoint.hpp"

#include <chrono>
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace tdbem {

namespace {

double NowA()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void CheckA(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Block transpose of a square block CCSR matrix (rows sorted by column).
CCSRMat TransposeBlocks(const CCSRMat& M, long n)
{
    CCSRMat T;
    T.n = n;
    T.row_ptr.assign((size_t)n + 1, 0);
    for (int c : M.col_index)
        T.row_ptr[(size_t)c + 1]++;
    for (long r = 0; r < n; ++r)
        T.row_ptr[(size_t)r + 1] += T.row_ptr[(size_t)r];
    std::vector<int> fill(T.row_ptr.begin(), T.row_ptr.end() - 1);
    T.col_index.resize(M.col_index.size());
    T.value.resize(M.value.size());
    for (long r = 0; r < (long)M.row_ptr.size() - 1; ++r)
        for (int k = M.row_ptr[(size_t)r]; k < M.row_ptr[(size_t)r + 1]; ++k) {
            int c = M.col_index[(size_t)k];
            int dst = fill[(size_t)c]++;
            T.col_index[(size_t)dst] = (int)r;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    T.value[9 * (size_t)dst + 3 * a + b] = M.value[9 * (size_t)k + 3 * b + a];
        }
    return T;
}

struct DevCCSR {
    DevBuf<int> rp, col;
    DevBuf<double> val;
    void Upload(const CCSRMat& M, hipStream_t s)
    {
        rp.upload(M.row_ptr.data(), M.row_ptr.size(), s);
        col.upload(M.col_index.data(), std::max<size_t>(1, M.col_index.size()), s);
        val.upload(M.value.data(), std::max<size_t>(9, M.value.size()), s);
    }
};

}  // namespace

void Case::GetMaterials(double* evr) const
{
    for (size_t d = 0; d < model.domains.size(); ++d) {
        evr[3 * d] = model.domains[d].E;
        evr[3 * d + 1] = model.domains[d].v;
        evr[3 * d + 2] = model.domains[d].Rou;
    }
}

void Case::SetMaterials(const double* evr)
{
    std::vector<DynaMat> mats;
    for (size_t d = 0; d < model.domains.size(); ++d) {
        Domain& D = model.domains[d];
        D.E = evr[3 * d];
        D.v = evr[3 * d + 1];
        D.Rou = evr[3 * d + 2];
        double G = D.E / 2.0 / (1.0 + D.v);
        D.mat = BuildDynaMat(D.v, G, D.Rou, dt);
        mats.push_back(D.mat);
    }
    d_mats.upload(mats.data(), mats.size(), stream);
    grid.mats = d_mats.p;
    LaunchSingular(d_geo.p, d_dom.p, d_mats.p, d_tables.p, d_u1.p, d_t1.p, d_t2.p, (int)ne, !quadStaged, stream);
    selfMode = quadStaged ? 1 : 0;
    CheckA(hipStreamSynchronize(stream), "materials");
}

double Case::Adjoint(int nstep, const double* known_lu, const int* recv, int nrec, const double* obs, long restart,
                     double tol, double fdStep, double* grad, double* lu_out)
{
    RefuseStreamed("the adjoint");
    if (!historyDone || maxn < nstep)
        throw std::runtime_error("history not assembled for the requested steps");
    if (historyMixed)
        throw std::runtime_error("the adjoint needs FP64 history blocks (mixed 0)");
    if (!quadStaged)
        throw std::runtime_error("the adjoint needs the staged history quadrature");
    for (int r = 0; r < nrec; ++r)
        if (recv[r] < 0 || recv[r] >= ne)
            throw std::runtime_error("receiver element out of range");
    const long n3 = 3 * ne;
    const int nstore = nstep + 1;
    const int nd = grid.nd;

    // 1. forward solve, keeping x_K, the state columns and the known inputs
    double tA = NowA();
    AdjointKeep kp;
    std::vector<double> lu((size_t)nstore * n3), lt((size_t)nstore * n3);
    std::vector<int> its((size_t)nstep), fls((size_t)nstep);
    int solved = 0;
    keep = &kp;
    try {
        Solve(nstep, known_lu, restart, tol, lu.data(), lt.data(), its.data(), fls.data(), &solved);
    } catch (...) {
        keep = nullptr;
        throw;
    }
    keep = nullptr;
    if (solved != nstep || fls[(size_t)nstep - 1] != 0)
        throw std::runtime_error("forward solve did not converge");
    timing[9] = NowA() - tA;
    tA = NowA();

    // 2. misfit and dJ/dx_K through lu = F^T (u_sign T x_u)
    double J = 0.0;
    std::vector<double> dJ((size_t)nstore * n3, 0.0);
    for (int K = 1; K <= nstep; ++K)
        for (int r = 0; r < nrec; ++r) {
            long i = recv[r];
            double l[3];
            for (int c = 0; c < 3; ++c) {
                double diff = lu[((size_t)K * ne + i) * 3 + c] - obs[((size_t)K * nrec + r) * 3 + c];
                J += diff * diff;
                l[c] = 2.0 * diff;
            }
            const NodeRefs& R = refs[(size_t)i];
            if (R.u_known)
                continue;
            const double* F = geo[(size_t)i].trans;
            double ub[3];
            for (int rr = 0; rr < 3; ++rr)
                ub[rr] = F[rr * 3] * l[0] + F[rr * 3 + 1] * l[1] + F[rr * 3 + 2] * l[2];
            for (int k = 0; k < 3; ++k)
                dJ[(size_t)K * n3 + 3 * (size_t)R.u_block + k] +=
                    R.u_sign * (R.itf_T[k] * ub[0] + R.itf_T[3 + k] * ub[1] + R.itf_T[6 + k] * ub[2]);
        }

    // 3. A^T and the transposed leaf preconditioner
    CCSRMat At = TransposeBlocks(A, ne);
    std::vector<long> outBlock((size_t)ne);
    CheckA(CopyOnStream(outBlock.data(), d_outBlock.p, sizeof(long) * (size_t)ne, hipMemcpyDeviceToHost, stream), "outBlock");
    struct Tri {
        int r, c;
        double v[9];
    };
    std::vector<Tri> tri;
    for (long l = 0; l < pre.leafCount; ++l) {
        long b = pre.leafBegin[(size_t)l], c = pre.leafSize[(size_t)l], n3l = 3 * c;
        const double* inv = &pre.inv[(size_t)pre.matOffset[(size_t)l]];
        for (long jl = 0; jl < c; ++jl)
            for (long k = 0; k < c; ++k) {
                Tri t;
                t.r = (int)pre.repid[(size_t)(b + k)];
                t.c = (int)outBlock[(size_t)(b + jl)];
                for (int mrow = 0; mrow < 3; ++mrow)
                    for (int cc = 0; cc < 3; ++cc)
                        t.v[3 * cc + mrow] = inv[(3 * jl + mrow) * n3l + 3 * k + cc];
                tri.push_back(t);
            }
    }
    std::sort(tri.begin(), tri.end(), [](const Tri& a, const Tri& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
    CCSRMat Mt;
    Mt.n = ne;
    Mt.row_ptr.assign((size_t)ne + 1, 0);
    for (const Tri& t : tri) {  // (row, col) pairs are distinct: each row lies in one leaf
        Mt.row_ptr[(size_t)t.r + 1]++;
        Mt.col_index.push_back(t.c);
        Mt.value.insert(Mt.value.end(), t.v, t.v + 9);
    }
    for (long r = 0; r < ne; ++r)
        Mt.row_ptr[(size_t)r + 1] += Mt.row_ptr[(size_t)r];
    DevCCSR dAt, dMt;
    dAt.Upload(At, stream);
    dMt.Upload(Mt, stream);

    // history columns -> the elements that read them
    auto columns = [&](bool u, std::vector<long>& ptr, std::vector<int>& mem) {
        long ncol = u ? dof.historyUBlockCount : dof.historyTBlockCount;
        ptr.assign((size_t)ncol + 1, 0);
        for (long f = 0; f < ne; ++f)
            ptr[(size_t)(u ? refs[(size_t)f].hu_block : refs[(size_t)f].ht_block) + 1]++;
        for (long c = 0; c < ncol; ++c)
            ptr[(size_t)c + 1] += ptr[(size_t)c];
        mem.assign((size_t)ne, 0);
        std::vector<long> fill(ptr.begin(), ptr.end() - 1);
        for (long f = 0; f < ne; ++f)
            mem[(size_t)fill[(size_t)(u ? refs[(size_t)f].hu_block : refs[(size_t)f].ht_block)]++] = (int)f;
    };
    std::vector<long> cuPtr, ctPtr;
    std::vector<int> cuMem, ctMem;
    columns(true, cuPtr, cuMem);
    columns(false, ctPtr, ctMem);
    DevBuf<long> d_cuPtr, d_ctPtr;
    DevBuf<int> d_cuMem, d_ctMem;
    d_cuPtr.upload(cuPtr.data(), cuPtr.size(), stream);
    d_ctPtr.upload(ctPtr.data(), ctPtr.size(), stream);
    d_cuMem.upload(cuMem.data(), cuMem.size(), stream);
    d_ctMem.upload(ctMem.data(), ctMem.size(), stream);

    // 4. backward sweep
    DevBuf<double> d_lam, d_y, d_z, d_dj, d_xs;
    d_lam.alloc((size_t)nstore * n3);
    d_xs.alloc((size_t)n3);
    LaunchZero(d_lam.p, (long)d_lam.n, stream);
    d_y.alloc((size_t)n3);
    d_z.alloc(6 * (size_t)ne);
    d_dj.upload(dJ.data(), dJ.size(), stream);
    PrepareGmres(restart);
    GmresOp<double> opT{dAt.rp.p, dAt.col.p, dAt.val.p, nullptr, dMt.rp.p, dMt.col.p, dMt.val.p};
    // time-blocked transposed history term: per block of kConvTTile backward
    // steps the lags > kConvTTile in one pass over the bands, the shorter lags
    // per step
    const int TT = convTiled ? kConvTTile : 0;
    const int pstride = nstep + kConvTTile + 2;
    DevBuf<double> d_phi, d_zf;
    if (TT) {
        d_phi.alloc((size_t)ne * (size_t)pstride * 3);
        LaunchZero(d_phi.p, (long)d_phi.n, stream);
        d_zf.alloc((size_t)TT * 6 * (size_t)ne);
    }
    for (int m1 = nstep; m1 >= 1; m1 -= std::max(TT, 1)) {
        int mlo = TT ? std::max(1, m1 - TT + 1) : m1;
        if (TT) {
            LaunchPhi(d_lam.p, n3, (int)ne, nstep, m1 + 2, std::min(m1 + TT + 1, nstep + 1), pstride, d_phi.p,
                      stream);
            LaunchConvTFar(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, d_mts.p, d_mgs.p, d_phi.p, pstride, nstep, mlo,
                           (int)ne, d_zf.p, stream);
        }
        for (int K = m1; K >= mlo; --K) {
            CheckA(hipMemcpyAsync(d_y.p, d_dj.p + (size_t)K * n3, sizeof(double) * (size_t)n3,
                                  hipMemcpyDeviceToDevice, stream),
                   "adjoint rhs");
            if (K < nstep) {
                if (TT)
                    LaunchConvTNear(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, d_mts.p, d_mgs.p, d_lam.p, n3, nstep,
                                    K, TT, d_zf.p + (size_t)(K - mlo) * 6 * (size_t)ne, (int)ne, d_z.p, stream);
                else
                    LaunchConvT(grid, d_refs.p, d_off.p, d_kb.p, d_len.p, d_mts.p, d_mgs.p, d_lam.p, n3, nstep, K,
                                nstep + 1, nullptr, (int)ne, d_z.p, stream);
                LaunchStateT(d_refs.p, d_geo.p, d_cuPtr.p, d_cuMem.p, d_ctPtr.p, d_ctMem.p, d_z.p, (int)ne, d_y.p,
                             stream);
            }
            int it = 0;
            int flag;
            if (gmresDevice && restart <= kGmDeviceMaxRestart) {
                // device-driven; solved into a fixed buffer so the captured graphs stay valid
                flag = GmresDevice(d_y.p, d_xs.p, restart, tol, &it, &opT);
                CheckA(hipMemcpyAsync(d_lam.p + (size_t)K * n3, d_xs.p, sizeof(double) * (size_t)n3,
                                      hipMemcpyDeviceToDevice, stream),
                       "adjoint solution");
            } else {
                flag = GmresCore<double>(d_y.p, d_lam.p + (size_t)K * n3, restart, tol, &it, opT, d_V.p, d_W.p,
                                         d_u.p, d_u2.p);
            }
            if (flag != 0)
                throw std::runtime_error("adjoint GMRES did not converge at step " + std::to_string(K));
        }
    }

    CheckA(hipStreamSynchronize(stream), "backward sweep");
    timing[10] = NowA() - tA;
    tA = NowA();
    // 5. gradient, with the bands and judges frozen at the current materials.
    // History blocks: analytic material derivatives contracted with the
    // per-entry correlation of lambda and the filtered states (or, with
    // adjointAnalytic off, central differences of two quadratures per
    // parameter).  Step-0 blocks and the singular self tables: central
    // differences (few blocks, reference-order kernels).
    const bool analytic = adjointAnalytic;
    bool wasFrozen = frozen;
    if (!wasFrozen)
        FreezeStructure(true);
    std::vector<double> theta((size_t)nd * 3);
    GetMaterials(theta.data());
    DevBuf<double> d_dM, d_dG, d_acc, d_C;
    DevBuf<unsigned char> d_aliveS;
    d_acc.alloc(3 * (size_t)ne);
    LaunchZero(d_acc.p, (long)d_acc.n, stream);
    std::vector<double> delta((size_t)nd), inv2row((size_t)ne), scaleDom((size_t)nd);
    // filtered state series for the entry-wise history contraction
    DevBuf<double> d_FU, d_FT, d_part, d_scale;
    d_FU.alloc(std::max<size_t>(1, (size_t)dof.historyUBlockCount * 3 * (size_t)nstore));
    d_FT.alloc(std::max<size_t>(1, (size_t)dof.historyTBlockCount * 3 * (size_t)nstore));
    LaunchFilterSeries(kp.gU.p, dof.historyUBlockCount, nstore, nstep, d_FU.p, stream);
    LaunchFilterSeries(kp.gT.p, dof.historyTBlockCount, nstore, nstep, d_FT.p, stream);
    std::vector<double> dT2;
    if (analytic) {
        d_C.alloc(15 * (size_t)std::max<long>(1, entries));
        LaunchCorrStore(grid, d_refs.p, d_coff.p, d_off.p, d_kb.p, d_len.p, npairs, chunks, d_lam.p, n3, d_FU.p,
                        d_FT.p, kp.gU.p, kp.gT.p, nstore, nstep, d_C.p, stream);
        dT2.assign(27 * (size_t)ne, 0.0);
    } else {
        d_dM.alloc(9 * (size_t)entries);
        d_dG.alloc(6 * (size_t)entries);
        d_aliveS.alloc((size_t)entries);
        d_part.alloc(std::max<size_t>(1, (size_t)chunks));
        d_scale.alloc((size_t)nd);
    }
    // step-0 passes at the perturbed materials over the near pairs of the
    // frozen structure, contracted on the device (LaunchStep0Grad)
    std::vector<long> nlist = NearPairs();
    const long nnear = (long)nlist.size();
    std::vector<long> nearRow((size_t)ne + 1, 0);
    for (long i = 0; i < nnear; ++i) {
        int s0, f0, d0;
        long p0 = nlist[(size_t)i];
        d0 = 0;
        while (d0 + 1 < grid.nd && p0 >= grid.pairBase[d0 + 1])
            ++d0;
        s0 = grid.eleBegin[d0] + (int)((p0 - grid.pairBase[d0]) / grid.eleCount[d0]);
        (void)f0;
        nearRow[(size_t)s0 + 1]++;
    }
    for (long r = 0; r < ne; ++r)
        nearRow[(size_t)r + 1] += nearRow[(size_t)r];
    DevBuf<long> d_nlist, d_nrow;
    DevBuf<double> d_outP, d_outM, d_gpair, d_inv2, d_knAll;
    DevBuf<int> d_flagP, d_flagM;
    d_nlist.upload(nlist.data(), std::max<size_t>(1, nlist.size()), stream);
    d_nrow.upload(nearRow.data(), nearRow.size(), stream);
    d_outP.alloc(18 * (size_t)std::max<long>(1, nnear));
    d_outM.alloc(18 * (size_t)std::max<long>(1, nnear));
    d_flagP.alloc(2 * (size_t)std::max<long>(1, nnear));
    d_flagM.alloc(2 * (size_t)std::max<long>(1, nnear));
    d_gpair.alloc((size_t)std::max<long>(1, nnear));
    const long nk3 = 3 * std::max<long>(1, dof.knownBlockCount);
    d_knAll.alloc((size_t)nstore * (size_t)nk3);
    LaunchZero(d_knAll.p, (long)d_knAll.n, stream);
    LaunchKnownAll(d_refs.p, kp.bd.p, (int)ne, nstore, nk3, d_knAll.p, stream);
    auto step0Pass = [&](double* out, int* flags) {
        EnsureSelfTables();
        LaunchStep0(d_geo.p, d_pts.p, d_dom.p, d_refs.p, d_u1.p, d_t1.p, d_t2.p, grid, d_nlist.p, nnear,
                    frozen ? d_cls0.p : nullptr, out, flags, stream);
    };
    std::vector<double> t2p, t2m;
    for (int p = 0; p < 3; ++p) {
        std::vector<double> tp = theta, tm = theta;
        for (int d = 0; d < nd; ++d) {
            double base = theta[(size_t)d * 3 + p];
            delta[(size_t)d] = fdStep * std::max(std::fabs(base), 0.1);
            tp[(size_t)d * 3 + p] = base + delta[(size_t)d];
            tm[(size_t)d * 3 + p] = base - delta[(size_t)d];
        }
        for (long s = 0; s < ne; ++s)
            inv2row[(size_t)s] = 1.0 / (2.0 * delta[(size_t)domOf[(size_t)s]]);
        SetMaterials(tp.data());
        step0Pass(d_outP.p, d_flagP.p);
        if (analytic) {
            t2p.resize(9 * (size_t)ne);
            CheckA(CopyOnStream(t2p.data(), d_t2.p, sizeof(double) * t2p.size(),
                             hipMemcpyDeviceToHost, stream),
                   "self tables");
        } else {
            StagedQuadrature(d_dM.p, d_dG.p, nullptr, nullptr, d_aliveS.p);
        }
        SetMaterials(tm.data());
        step0Pass(d_outM.p, d_flagM.p);
        if (analytic) {
            t2m.resize(9 * (size_t)ne);
            CheckA(CopyOnStream(t2m.data(), d_t2.p, sizeof(double) * t2m.size(), hipMemcpyDeviceToHost, stream),
                   "self tables");
            for (long f = 0; f < ne; ++f)
                for (int i = 0; i < 9; ++i)
                    dT2[27 * (size_t)f + 9 * (size_t)p + i] = (t2p[9 * (size_t)f + i] - t2m[9 * (size_t)f + i]) *
                                                               inv2row[(size_t)f];
        } else {
            StagedQuadrature(d_mts.p, d_mgs.p, nullptr, nullptr, d_aliveS.p);
            // history blocks: sum_K lambda_K . h_K(dM), entry by entry
            for (int d = 0; d < nd; ++d)
                scaleDom[(size_t)d] = 1.0 / (2.0 * delta[(size_t)d]);
            d_scale.upload(scaleDom.data(), scaleDom.size(), stream);
            LaunchCorrFd(grid, d_refs.p, d_coff.p, d_off.p, d_kb.p, d_len.p, npairs, chunks, d_dM.p, d_dG.p, d_mts.p,
                         d_mgs.p, d_scale.p, d_lam.p, n3, d_FU.p, d_FT.p, kp.gU.p, kp.gT.p, nstore, nstep, d_part.p,
                         stream);
            LaunchRowChunks(grid, d_refs.p, d_coff.p, d_part.p, (int)ne, d_acc.p + (size_t)p * ne, stream);
        }
        // step-0 blocks: lambda_K . (dKnownM k_K - dA x_K), summed over K per near pair
        d_inv2.upload(inv2row.data(), inv2row.size(), stream);
        LaunchStep0Grad(d_nlist.p, nnear, grid, d_refs.p, d_outP.p, d_flagP.p, d_outM.p, d_flagM.p, d_inv2.p,
                        d_lam.p, kp.X.p, n3, d_knAll.p, nk3, nstep, d_gpair.p, d_nrow.p, (int)ne,
                        d_acc.p + (size_t)p * ne, stream);
        CheckA(hipStreamSynchronize(stream), "gradient");
    }
    // restore the materials and self tables (and, after differences, the bands)
    SetMaterials(theta.data());
    timing[11] = NowA() - tA;
    tA = NowA();
    std::vector<double> hsum((size_t)nd * 3, 0.0);
    if (analytic) {
        // d(C1, C2, rho) / d(E, v, rho) per domain
        std::vector<double> J((size_t)nd * 9, 0.0);
        for (int d = 0; d < nd; ++d) {
            const Domain& D = model.domains[(size_t)d];
            double E = D.E, v = D.v, rho = D.Rou, C1 = D.mat.C1, C2 = D.mat.C2;
            double* j = &J[(size_t)d * 9];
            j[0] = C1 / (2.0 * E);
            j[1] = 0.5 * C1 * (1.0 / ((1.0 - 2.0 * v) * (1.0 - v)) - 1.0 / (1.0 + v));
            j[2] = -C1 / (2.0 * rho);
            j[3] = C2 / (2.0 * E);
            j[4] = -C2 / (2.0 * (1.0 + v));
            j[5] = -C2 / (2.0 * rho);
            j[8] = 1.0;
        }
        DevBuf<double> d_J, d_dT2, d_gpart, d_gsum;
        d_J.upload(J.data(), J.size(), stream);
        d_dT2.upload(dT2.data(), dT2.size(), stream);
        PackPwWaves();
        long nw = chunks + pwWaves;
        d_gpart.alloc(4 * (size_t)std::max<long>(1, nw));
        d_gsum.alloc((size_t)nd * 3);
        LaunchZero(d_gsum.p, (long)d_gsum.n, stream);
        LaunchHistoryGrad(d_geo.p, d_pts.p, d_refs.p, grid, d_coff.p, d_off.p, d_kb.p, d_len.p, d_cls.p, npairs,
                          chunks, d_pwWaves.p, pwWaves, d_alive.p, d_C.p, d_J.p, d_dT2.p, d_gpart.p, stream);
        LaunchDomainSums(d_gpart.p, nw, nd, d_gsum.p, stream);
        CheckA(CopyOnStream(hsum.data(), d_gsum.p, sizeof(double) * hsum.size(), hipMemcpyDeviceToHost, stream),
               "history gradient");
    } else {
        StagedQuadrature(d_mts.p, d_mgs.p, nullptr, nullptr, d_alive.p);
    }
    if (!wasFrozen)
        FreezeStructure(false);
    timing[12] = NowA() - tA;
    std::vector<double> acc(3 * (size_t)ne);
    CheckA(CopyOnStream(acc.data(), d_acc.p, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, stream), "gradient");
    for (int i = 0; i < 3 * nd; ++i)
        grad[i] = hsum[(size_t)i];
    for (int p = 0; p < 3; ++p)
        for (long s = 0; s < ne; ++s)
            grad[3 * domOf[(size_t)s] + p] += acc[(size_t)p * ne + s];
    if (lu_out)
        std::memcpy(lu_out, lu.data(), sizeof(double) * lu.size());
    return J;
}

}  // namespace tdbem
