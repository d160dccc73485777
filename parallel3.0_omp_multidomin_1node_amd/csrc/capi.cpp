// extern "C" boundary of the MI355X TDBEM path (declared in include/tdbem_capi.h).
#include "../../include/tdbem_capi.h"

#include "batched.hpp"
#include "io.hpp"
#include "solver.hpp"

#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

struct tdbem_ctx {
    tdbem::Case c;
};

namespace {
thread_local std::string g_err;

int Fail(const char* where, const std::exception& e)
{
    g_err = std::string(where) + ": " + e.what();
    return -1;
}

int Fail(const char* msg)
{
    g_err = msg;
    return -2;
}

int Fail(const std::string& msg)
{
    return Fail(msg.c_str());
}

// The multi-domain datacard tail from the C-ABI's flat rows.
tdbem::MultiDomainInputConfig MakeConfig(int domain_count, const double* mat_rows, int interface_count,
                                         const long* itf_rows)
{
    tdbem::MultiDomainInputConfig cfg;
    if (domain_count > 0) {
        cfg.enabled = true;
        cfg.domainCount = domain_count;
        for (int d = 0; d < domain_count; ++d)
            cfg.materials.push_back(tdbem::DomainMaterialInput{(int)mat_rows[4 * d], mat_rows[4 * d + 1],
                                                               mat_rows[4 * d + 2], mat_rows[4 * d + 3]});
        if (interface_count >= 0) {
            cfg.hasExplicitInterfaces = true;
            for (int i = 0; i < interface_count; ++i) {
                tdbem::InterfacePair f;
                f.domainA = (int)itf_rows[5 * i];
                f.eleA = itf_rows[5 * i + 1];
                f.domainB = (int)itf_rows[5 * i + 2];
                f.eleB = itf_rows[5 * i + 3];
                f.normalSign = (int)itf_rows[5 * i + 4];
                cfg.interfaces.push_back(f);
            }
        }
    }
    return cfg;
}

void ExportDofMap(const tdbem::GlobalDofMap& d, const std::vector<int>& domOf, long ne, long* counts, long* u_block,
                  double* u_sign, int* u_known, long* t_block, double* t_sign, int* t_known, long* hu_block,
                  double* hu_sign, long* ht_block, double* ht_sign, long* preferred)
{
    counts[0] = d.unknownBlockCount;
    counts[1] = d.knownBlockCount;
    counts[2] = d.equationBlockCount;
    counts[3] = d.outerDisplacementUnknownBlocks;
    counts[4] = d.outerTractionUnknownBlocks;
    counts[5] = d.interfaceDisplacementUnknownBlocks;
    counts[6] = d.interfaceTractionUnknownBlocks;
    counts[7] = d.historyUBlockCount;
    counts[8] = d.historyTBlockCount;
    for (long i = 0; i < ne; ++i) {
        int dm = domOf[(size_t)i];
        tdbem::VariableRef u = d.GetU(dm, i), t = d.GetT(dm, i), hu = d.GetHistoryU(dm, i), ht = d.GetHistoryT(dm, i);
        if (u_block) {
            u_block[i] = u.globalBlock;
            u_sign[i] = u.sign;
            u_known[i] = u.known;
            t_block[i] = t.globalBlock;
            t_sign[i] = t.sign;
            t_known[i] = t.known;
            hu_block[i] = hu.globalBlock;
            hu_sign[i] = hu.sign;
            ht_block[i] = ht.globalBlock;
            ht_sign[i] = ht.sign;
        }
        if (preferred)
            preferred[i] = d.GetPreferredUnknownBlockForRow(i);
    }
}
}  // namespace

struct tdbem_builder {
    tdbem::MultiDomainCCSRBuilder b;
};

extern "C" {

const char* tdbem_last_error(void) { return g_err.c_str(); }

int tdbem_create(tdbem_ctx** out, long npoints, const double* points, long ne, const int* ele_pid, const int* bcid,
                 int domain_count, const double* mat_rows, int interface_count, const long* itf_rows,
                 double default_E, double default_v, double default_rho, double dt, double gap)
{
    if (!out || !points || !ele_pid || !bcid || ne <= 0 || npoints <= 0)
        return Fail("tdbem_create: invalid arguments");
    for (long i = 0; i < 8 * ne; ++i)
        if (ele_pid[i] < 0 || ele_pid[i] >= npoints)
            return Fail("tdbem_create: element point id out of range");
    if ((domain_count > 0 && !mat_rows) || (domain_count > 0 && interface_count > 0 && !itf_rows))
        return Fail("tdbem_create: invalid arguments");
    tdbem::MultiDomainInputConfig cfg = MakeConfig(domain_count, mat_rows, interface_count, itf_rows);
    std::unique_ptr<tdbem_ctx> ctx(new tdbem_ctx());
    try {
        ctx->c.Create(npoints, points, ne, ele_pid, bcid, cfg, default_E, default_v, default_rho, dt, gap);
    } catch (const std::exception& e) {
        return Fail("tdbem_create", e);
    }
    *out = ctx.release();
    return 0;
}

void tdbem_destroy(tdbem_ctx* ctx) { delete ctx; }

int tdbem_dofmap_export(tdbem_ctx* ctx, long* counts, long* u_block, double* u_sign, int* u_known, long* t_block,
                        double* t_sign, int* t_known, long* hu_block, double* hu_sign, long* ht_block,
                        double* ht_sign, long* preferred)
{
    if (!ctx || !counts || (u_block && (!u_sign || !u_known || !t_block || !t_sign || !t_known || !hu_block ||
                                        !hu_sign || !ht_block || !ht_sign)))
        return Fail("tdbem_dofmap_export: invalid arguments");
    ExportDofMap(ctx->c.dof, ctx->c.domOf, ctx->c.ne, counts, u_block, u_sign, u_known, t_block, t_sign, t_known,
                 hu_block, hu_sign, ht_block, ht_sign, preferred);
    return 0;
}

int tdbem_dofmap_build(long ne, const int* bcid, int domain_count, const double* mat_rows, int interface_count,
                       const long* itf_rows, long* counts, long* u_block, double* u_sign, int* u_known,
                       long* t_block, double* t_sign, int* t_known, long* hu_block, double* hu_sign,
                       long* ht_block, double* ht_sign, long* preferred)
{
    if (ne <= 0 || !bcid || !counts || (domain_count > 0 && !mat_rows) ||
        (domain_count > 0 && interface_count > 0 && !itf_rows) ||
        (u_block && (!u_sign || !u_known || !t_block || !t_sign || !t_known || !hu_block || !hu_sign || !ht_block ||
                     !ht_sign)))
        return Fail("tdbem_dofmap_build: invalid arguments");
    try {
        tdbem::MultiDomainInputConfig cfg = MakeConfig(domain_count, mat_rows, interface_count, itf_rows);
        tdbem::MultiDomainModel model;
        // the materials do not enter the DOF map; unit values stand in for the global line
        if (!tdbem::BuildMultiDomainModel(cfg, ne, ne, 1.0, 0.0, 1.0, 1.0, model))
            return Fail("tdbem_dofmap_build: BuildMultiDomainModel failed");
        tdbem::GlobalDofMap dof;
        if (!dof.Build(model, bcid, ne))
            return Fail("tdbem_dofmap_build: GlobalDofMap::Build failed");
        std::vector<int> domOf((size_t)ne, 0);
        for (const tdbem::Domain& d : model.domains)
            for (long e : d.elementIds)
                domOf[(size_t)e] = d.id;
        ExportDofMap(dof, domOf, ne, counts, u_block, u_sign, u_known, t_block, t_sign, t_known, hu_block, hu_sign,
                     ht_block, ht_sign, preferred);
    } catch (const std::exception& e) {
        return Fail("tdbem_dofmap_build", e);
    }
    return 0;
}

int tdbem_builder_create(tdbem_builder** out, long block_rows, long block_cols)
{
    if (!out || block_rows < 0 || block_cols < 0)
        return Fail("tdbem_builder_create: invalid arguments");
    try {
        tdbem_builder* b = new tdbem_builder();
        b->b.Reset(block_rows, block_cols);
        *out = b;
    } catch (const std::exception& e) {
        return Fail("tdbem_builder_create", e);
    }
    return 0;
}

void tdbem_builder_destroy(tdbem_builder* b) { delete b; }

int tdbem_builder_add_block(tdbem_builder* b, long row, long col, double sign, const double* block9)
{
    if (!b)
        return Fail("tdbem_builder_add_block: invalid arguments");
    b->b.AddBlock(row, col, sign, block9);
    return 0;
}

int tdbem_builder_merge(tdbem_builder* b, const tdbem_builder* other)
{
    if (!b || !other)
        return Fail("tdbem_builder_merge: invalid arguments");
    b->b.MergeFrom(other->b);
    return 0;
}

int tdbem_builder_info(const tdbem_builder* b, long* nnz, long* skipped_nonfinite)
{
    if (!b)
        return Fail("tdbem_builder_info: invalid arguments");
    if (nnz)
        *nnz = b->b.NonZeroBlocks();
    if (skipped_nonfinite)
        *skipped_nonfinite = b->b.SkippedNonFiniteBlocks();
    return 0;
}

int tdbem_builder_build(const tdbem_builder* b, int* row_ptr, int* col, double* val)
{
    if (!b || !row_ptr)
        return Fail("tdbem_builder_build: invalid arguments");
    tdbem::CCSRMat m;
    b->b.Build(m);
    std::memcpy(row_ptr, m.row_ptr.data(), sizeof(int) * m.row_ptr.size());
    if (col && !m.col_index.empty())
        std::memcpy(col, m.col_index.data(), sizeof(int) * m.col_index.size());
    if (val && !m.value.empty())
        std::memcpy(val, m.value.data(), sizeof(double) * m.value.size());
    return 0;
}

int tdbem_builder_build_sym(const tdbem_builder* b, double tolerance, int* row_ptr, int* col, double* val6,
                            int* symmetric)
{
    if (!b || !row_ptr)
        return Fail("tdbem_builder_build_sym: invalid arguments");
    tdbem::SymCCSRMat m;
    bool sym = false;
    b->b.BuildSym(m, tolerance, &sym);
    std::memcpy(row_ptr, m.row_ptr.data(), sizeof(int) * m.row_ptr.size());
    if (col && !m.col_index.empty())
        std::memcpy(col, m.col_index.data(), sizeof(int) * m.col_index.size());
    if (val6 && !m.value.empty())
        std::memcpy(val6, m.value.data(), sizeof(double) * m.value.size());
    if (symmetric)
        *symmetric = sym ? 1 : 0;
    return 0;
}

int tdbem_assemble_step0(tdbem_ctx* ctx, long maxleaf)
{
    if (!ctx || maxleaf <= 0)
        return Fail("tdbem_assemble_step0: invalid arguments");
    try {
        ctx->c.AssembleStep0(maxleaf);
    } catch (const std::exception& e) {
        return Fail("tdbem_assemble_step0", e);
    }
    return 0;
}

int tdbem_assemble_history(tdbem_ctx* ctx, int maxn, long* entries)
{
    if (!ctx || maxn < 1)
        return Fail("tdbem_assemble_history: invalid arguments");
    try {
        ctx->c.AssembleHistory(maxn);
    } catch (const std::exception& e) {
        return Fail("tdbem_assemble_history", e);
    }
    if (entries)
        *entries = ctx->c.entries;
    return 0;
}

int tdbem_ccsr_export(tdbem_ctx* ctx, int which, int step, long* nrows, long* nnz, int* row_ptr, int* col,
                      double* val)
{
    if (!ctx || !nrows || !nnz || which < 0 || which > 3)
        return Fail("tdbem_ccsr_export: invalid arguments");
    std::vector<int> rp, c;
    std::vector<double> v;
    try {
        ctx->c.ExportMatrix(which, step, rp, c, v);
    } catch (const std::exception& e) {
        return Fail("tdbem_ccsr_export", e);
    }
    *nrows = (long)rp.size() - 1;
    *nnz = (long)c.size();
    if (row_ptr)
        std::memcpy(row_ptr, rp.data(), sizeof(int) * rp.size());
    if (col && !c.empty())
        std::memcpy(col, c.data(), sizeof(int) * c.size());
    if (val && !v.empty())
        std::memcpy(val, v.data(), sizeof(double) * v.size());
    return 0;
}

int tdbem_ccsr_export_rows(tdbem_ctx* ctx, int which, int step, long row0, long row1, long* nnz, int* row_ptr,
                           int* col, double* val)
{
    if (!ctx || !nnz || which < 2 || which > 3)
        return Fail("tdbem_ccsr_export_rows: invalid arguments");
    std::vector<int> rp, c;
    std::vector<double> v;
    try {
        ctx->c.ExportRows(which, step, row0, row1, rp, c, v);
    } catch (const std::exception& e) {
        return Fail("tdbem_ccsr_export_rows", e);
    }
    *nnz = (long)c.size();
    if (row_ptr)
        std::memcpy(row_ptr, rp.data(), sizeof(int) * rp.size());
    if (col && !c.empty())
        std::memcpy(col, c.data(), sizeof(int) * c.size());
    if (val && !v.empty())
        std::memcpy(val, v.data(), sizeof(double) * v.size());
    return 0;
}

int tdbem_self_tables(tdbem_ctx* ctx, double* u1, double* t1, double* t2)
{
    if (!ctx || !u1 || !t1 || !t2)
        return Fail("tdbem_self_tables: invalid arguments");
    try {
        ctx->c.EnsureSelfTables();
    } catch (const std::exception& e) {
        return Fail("tdbem_self_tables", e);
    }
    size_t n = 9 * (size_t)ctx->c.ne;
    hipStream_t s = ctx->c.stream;
    if (tdbem::CopyOnStream(u1, ctx->c.d_u1.p, n * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        tdbem::CopyOnStream(t1, ctx->c.d_t1.p, n * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        tdbem::CopyOnStream(t2, ctx->c.d_t2.p, n * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        return Fail("tdbem_self_tables: device copy failed");
    return 0;
}

int tdbem_solve(tdbem_ctx* ctx, int nstep, const double* known_lu, long restart, double tol, double* lu, double* lt,
                int* iters, int* flags, int* solved, double* times)
{
    if (!ctx || nstep < 1 || !known_lu || !lu || !lt || !iters || !flags || !solved || restart < 1 || !(tol > 0.0))
        return Fail("tdbem_solve: invalid arguments");
    try {
        ctx->c.Solve(nstep, known_lu, restart, tol, lu, lt, iters, flags, solved);
    } catch (const std::exception& e) {
        return Fail("tdbem_solve", e);
    }
    if (times) {
        times[0] = ctx->c.timing[4];
        times[1] = ctx->c.timing[5];
        times[2] = ctx->c.timing[4] + ctx->c.timing[5];
    }
    return 0;
}

int tdbem_run_models(int n, tdbem_ctx* const* ctx, long maxleaf, const int* nstep, const double* const* known_lu,
                     long restart, double tol, double budget_bytes, double* const* lu, double* const* lt,
                     int* const* iters, int* const* flags, int* solved)
{
    if (n < 1 || !ctx || maxleaf <= 0 || !nstep || !known_lu || !lu || !lt || !iters || !flags || !solved ||
        restart < 1 || !(tol > 0.0))
        return Fail("tdbem_run_models: invalid arguments");
    for (int i = 0; i < n; ++i) {
        if (!ctx[i] || nstep[i] < 1 || !known_lu[i] || !lu[i] || !lt[i] || !iters[i] || !flags[i])
            return Fail("tdbem_run_models: invalid arguments for model " + std::to_string(i));
        for (int k = 0; k < i; ++k)
            if (ctx[k] == ctx[i])
                return Fail("tdbem_run_models: model " + std::to_string(i) + " repeats a context");
    }
    if (!(budget_bytes > 0.0)) {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) != hipSuccess)
            return Fail("tdbem_run_models: hipMemGetInfo failed");
        budget_bytes = 0.8 * (double)(freeB + tdbem::PooledBytesCached());
    }
    // history bands held at once stay within the budget: a model whose bands
    // do not fit next to the running ones waits for them to finish (a model
    // larger than the budget runs alone)
    std::mutex mu;
    std::condition_variable cv;
    double held = 0.0;
    std::vector<std::string> errs((size_t)n);
    std::vector<char> restoreGraphs((size_t)n, 0);
    std::vector<std::thread> pool;
    pool.reserve((size_t)n);
    for (int i = 0; i < n; ++i)
        pool.emplace_back([&, i] {
            double need = 0.0;
            bool holding = false;
            try {
                tdbem::Case& c = ctx[i]->c;
                // device-driven GMRES with direct launches: no stream capture
                // while other threads allocate and free
                restoreGraphs[(size_t)i] = c.gmresDevice == 1;
                if (restoreGraphs[(size_t)i])
                    c.gmresDevice = 2;
                c.AssembleStep0(maxleaf);
                need = c.PrepareHistory(nstep[i]);
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return held == 0.0 || held + need <= budget_bytes; });
                    held += need;
                    holding = true;
                }
                c.AssembleHistory(nstep[i]);
                c.Solve(nstep[i], known_lu[i], restart, tol, lu[i], lt[i], iters[i], flags[i], &solved[i]);
                c.ReleaseHistory();
            } catch (const std::exception& e) {
                errs[(size_t)i] = e.what();
                try {
                    ctx[i]->c.ReleaseHistory();
                } catch (...) {
                }
            }
            if (restoreGraphs[(size_t)i])
                ctx[i]->c.gmresDevice = 1;
            if (holding) {
                std::lock_guard<std::mutex> lk(mu);
                held -= need;
                if (held < 1.0)
                    held = 0.0;
            }
            cv.notify_all();
        });
    for (std::thread& t : pool)
        t.join();
    for (int i = 0; i < n; ++i)
        if (!errs[(size_t)i].empty())
        {
            g_err = "tdbem_run_models: model " + std::to_string(i) + ": " + errs[(size_t)i];
            return -1;
        }
    return 0;
}

int tdbem_run_models_grouped(int n, tdbem_ctx* const* ctx, long maxleaf, const int* nstep,
                             const double* const* known_lu, long restart, double tol, double budget_bytes,
                             double* const* lu, double* const* lt, int* const* iters, int* const* flags, int* solved,
                             int* rounds)
{
    if (n < 1 || !ctx || maxleaf <= 0 || !nstep || !known_lu || !lu || !lt || !iters || !flags || !solved ||
        restart < 1 || !(tol > 0.0))
        return Fail("tdbem_run_models_grouped: invalid arguments");
    for (int i = 0; i < n; ++i) {
        if (!ctx[i] || nstep[i] < 1 || !known_lu[i] || !lu[i] || !lt[i] || !iters[i] || !flags[i])
            return Fail("tdbem_run_models_grouped: invalid arguments for model " + std::to_string(i));
        for (int k = 0; k < i; ++k)
            if (ctx[k] == ctx[i])
                return Fail("tdbem_run_models_grouped: model " + std::to_string(i) + " repeats a context");
        const tdbem::Case& c = ctx[i]->c;
        const tdbem::Case& c0 = ctx[0]->c;
        if (c.mixed != 0)
            return Fail("tdbem_run_models_grouped: model " + std::to_string(i) + " asks for mixed-precision bands; "
                        "grouped models run FP64 bands");
        if (c.quadStaged != c0.quadStaged || c.nearPerEntry != c0.nearPerEntry || c.farParts != c0.farParts ||
            c.convPacked != c0.convPacked || c.convFiltered != c0.convFiltered || c.convTiled != c0.convTiled ||
            c.convTile != c0.convTile || c.gmBlocks != c0.gmBlocks || c.gmresDevice != c0.gmresDevice)
            return Fail("tdbem_run_models_grouped: model " + std::to_string(i) +
                        " has other solver options than model 0 (grouped models share one set)");
    }
    if (!(budget_bytes > 0.0)) {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) != hipSuccess)
            return Fail("tdbem_run_models_grouped: hipMemGetInfo failed");
        budget_bytes = 0.8 * (double)(freeB + tdbem::PooledBytesCached());
    }
    // band bytes of every model from its own cull, then rounds: models taken
    // in order while their bands fit the budget together, share the step count
    // and stay within the group and domain limits; a model larger than the
    // budget makes a round of its own
    std::vector<double> need((size_t)n);
    try {
        for (int i = 0; i < n; ++i) {
            need[(size_t)i] = ctx[i]->c.PrepareHistory(nstep[i]);
            ctx[i]->c.ReleaseHistory();
        }
    } catch (const std::exception& e) {
        return Fail("tdbem_run_models_grouped: sizing", e);
    }
    std::vector<std::vector<int>> plan;
    {
        double held = 0.0;
        int doms = 0;
        for (int i = 0; i < n; ++i) {
            const int d = ctx[i]->c.model.DomainCount();
            bool fits = !plan.empty() && !plan.back().empty() && held + need[(size_t)i] <= budget_bytes &&
                        nstep[plan.back()[0]] == nstep[i] && doms + d <= tdbem::kMaxDomains &&
                        (int)plan.back().size() < tdbem::kGmGroupsMax;
            if (!fits) {
                plan.emplace_back();
                held = 0.0;
                doms = 0;
            }
            plan.back().push_back(i);
            held += need[(size_t)i];
            doms += d;
        }
    }
    if (rounds)
        *rounds = (int)plan.size();
    for (const std::vector<int>& R : plan) {
        const int G = (int)R.size();
        const int ns = nstep[R[0]];
        try {
            tdbem::Case m;
            std::vector<const tdbem::Case*> ms;
            for (int i : R)
                ms.push_back(&ctx[i]->c);
            m.CreateGroup(ms);
            m.historyStream = 0;  // the round's bands were sized to fit the device
            m.AssembleStep0(maxleaf);
            m.AssembleHistory(ns);
            const size_t n3 = 3 * (size_t)m.ne, steps = (size_t)ns + 1;
            std::vector<double> known(steps * n3), mlu(steps * n3), mlt(steps * n3);
            for (size_t K = 0; K < steps; ++K)
                for (int g = 0; g < G; ++g) {
                    const size_t ng = 3 * (size_t)ctx[R[(size_t)g]]->c.ne, at = 3 * (size_t)m.groupEle[(size_t)g];
                    std::memcpy(&known[K * n3 + at], known_lu[R[(size_t)g]] + K * ng, sizeof(double) * ng);
                }
            std::vector<int> its((size_t)G * (size_t)ns, 0), fls((size_t)G * (size_t)ns, 0), sol((size_t)G, 0);
            m.SolveGroups(ns, known.data(), restart, tol, mlu.data(), mlt.data(), its.data(), fls.data(), sol.data());
            // each model's outputs as its own job leaves them: the steps it
            // solved (not the step whose GMRES stopped with a flag), zeros after
            for (int g = 0; g < G; ++g) {
                const int i = R[(size_t)g];
                const size_t ng = 3 * (size_t)ctx[i]->c.ne, at = 3 * (size_t)m.groupEle[(size_t)g];
                const int sv = sol[(size_t)g];
                const int good = sv > 0 && fls[(size_t)g * ns + (sv - 1)] != 0 ? sv - 1 : sv;
                for (size_t K = 0; K < steps; ++K) {
                    if ((int)K <= good) {
                        std::memcpy(lu[i] + K * ng, &mlu[K * n3 + at], sizeof(double) * ng);
                        std::memcpy(lt[i] + K * ng, &mlt[K * n3 + at], sizeof(double) * ng);
                    } else {
                        std::memset(lu[i] + K * ng, 0, sizeof(double) * ng);
                        std::memset(lt[i] + K * ng, 0, sizeof(double) * ng);
                    }
                }
                for (int K = 0; K < ns; ++K) {  // zero past the solved steps
                    iters[i][K] = K < sv ? its[(size_t)g * ns + K] : 0;
                    flags[i][K] = K < sv ? fls[(size_t)g * ns + K] : 0;
                }
                solved[i] = sv;
            }
        } catch (const std::exception& e) {
            std::string who;
            for (int i : R)
                who += (who.empty() ? "" : ",") + std::to_string(i);
            return Fail(std::string("tdbem_run_models_grouped: round of models ") + who + ": " + e.what());
        }
    }
    return 0;
}

int tdbem_solve_batched(tdbem_ctx* ctx, int nstep, int nrhs, const double* known_lu, long restart, double tol,
                        double* lu, double* lt, int* iters, int* flags, int* solved, double* times)
{
    if (!ctx || nstep < 1 || nrhs < 1 || !known_lu || !lu || !lt || !iters || !flags || !solved || restart < 1 ||
        !(tol > 0.0))
        return Fail("tdbem_solve_batched: invalid arguments");
    try {
        ctx->c.SolveBatched(nstep, nrhs, known_lu, restart, tol, lu, lt, iters, flags, solved);
    } catch (const std::exception& e) {
        return Fail("tdbem_solve_batched", e);
    }
    if (times) {
        times[0] = ctx->c.timing[4];
        times[1] = ctx->c.timing[5];
        times[2] = ctx->c.timing[4] + ctx->c.timing[5];
    }
    return 0;
}

int tdbem_solve_device(tdbem_ctx* ctx, int nstep, const double* d_known_lu, long restart, double tol, double* d_lu,
                       double* d_lt, int* iters, int* flags, int* solved, double* times)
{
    if (!ctx || nstep < 1 || !d_known_lu || !d_lu || !d_lt || !iters || !flags || !solved || restart < 1 ||
        !(tol > 0.0))
        return Fail("tdbem_solve_device: invalid arguments");
    try {
        ctx->c.Solve(nstep, d_known_lu, restart, tol, d_lu, d_lt, iters, flags, solved, true);
    } catch (const std::exception& e) {
        return Fail("tdbem_solve_device", e);
    }
    if (times) {
        times[0] = ctx->c.timing[4];
        times[1] = ctx->c.timing[5];
        times[2] = ctx->c.timing[4] + ctx->c.timing[5];
    }
    return 0;
}

int tdbem_solve_batched_device(tdbem_ctx* ctx, int nstep, int nrhs, const double* d_known_lu, long restart,
                               double tol, double* d_lu, double* d_lt, int* iters, int* flags, int* solved,
                               double* times)
{
    if (!ctx || nstep < 1 || nrhs < 1 || !d_known_lu || !d_lu || !d_lt || !iters || !flags || !solved ||
        restart < 1 || !(tol > 0.0))
        return Fail("tdbem_solve_batched_device: invalid arguments");
    try {
        ctx->c.SolveBatched(nstep, nrhs, d_known_lu, restart, tol, d_lu, d_lt, iters, flags, solved, true);
    } catch (const std::exception& e) {
        return Fail("tdbem_solve_batched_device", e);
    }
    if (times) {
        times[0] = ctx->c.timing[4];
        times[1] = ctx->c.timing[5];
        times[2] = ctx->c.timing[4] + ctx->c.timing[5];
    }
    return 0;
}

int tdbem_solve_device_noalloc(tdbem_ctx* ctx, int nstep, const double* d_known_lu, long restart, double tol,
                               double* d_lu, double* d_lt, int* iters, int* flags, int* solved)
{
    if (!ctx || nstep < 1 || !d_known_lu || !d_lu || !d_lt || !iters || !flags || !solved || restart < 1 ||
        !(tol > 0.0))
        return Fail("tdbem_solve_device_noalloc: invalid arguments");
    try {
        ctx->c.SolveNoAlloc(nstep, d_known_lu, restart, tol, d_lu, d_lt, iters, flags, solved);
    } catch (const std::exception& e) {
        return Fail("tdbem_solve_device_noalloc", e);
    }
    return 0;
}

int tdbem_device_alloc(size_t bytes, void** out)
{
    if (!out || bytes == 0)
        return Fail("tdbem_device_alloc: invalid arguments");
    *out = nullptr;
    if (hipMalloc(out, bytes) != hipSuccess) {
        *out = nullptr;
        (void)hipGetLastError();
        return Fail("tdbem_device_alloc: hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return 0;
}

int tdbem_device_free(void* p)
{
    if (p && hipFree(p) != hipSuccess)
        return Fail("tdbem_device_free: hipFree failed");
    return 0;
}

int tdbem_memcpy(void* dst, const void* src, size_t bytes, int kind)
{
    static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    if ((!dst || !src) && bytes)
        return Fail("tdbem_memcpy: invalid arguments");
    if (kind < 0 || kind > 2)
        return Fail("tdbem_memcpy: kind must be 0 (host to device), 1 (device to host) or 2 (device to device)");
    if (bytes && hipMemcpy(dst, src, bytes, kinds[kind]) != hipSuccess)
        return Fail("tdbem_memcpy: hipMemcpy failed");
    return 0;
}

int tdbem_adjoint(tdbem_ctx* ctx, int nstep, const double* known_lu, int nrec, const int* receivers,
                  const double* obs, long restart, double tol, double fd_step, double* J, double* grad, double* lu)
{
    if (!ctx || nstep < 1 || !known_lu || nrec < 1 || !receivers || !obs || restart < 1 || !(tol > 0.0) ||
        !(fd_step > 0.0) || !J || !grad)
        return Fail("tdbem_adjoint: invalid arguments");
    try {
        *J = ctx->c.Adjoint(nstep, known_lu, receivers, nrec, obs, restart, tol, fd_step, grad, lu);
    } catch (const std::exception& e) {
        return Fail("tdbem_adjoint", e);
    }
    return 0;
}

int tdbem_get_materials(tdbem_ctx* ctx, int capacity, int* ndomains, double* evr)
{
    if (!ctx || !evr || !ndomains)
        return Fail("tdbem_get_materials: invalid arguments");
    int nd = (int)ctx->c.model.domains.size();
    *ndomains = nd;
    if (capacity < nd)
        return Fail("tdbem_get_materials: capacity below the domain count");
    ctx->c.GetMaterials(evr);
    return 0;
}

int tdbem_set_materials(tdbem_ctx* ctx, const double* evr)
{
    if (!ctx || !evr)
        return Fail("tdbem_set_materials: invalid arguments");
    try {
        ctx->c.SetMaterials(evr);
    } catch (const std::exception& e) {
        return Fail("tdbem_set_materials", e);
    }
    return 0;
}

int tdbem_dyna_gmres_solver(tdbem_ctx* ctx, long iterations, double error, long maxleafpointnum, int nstep,
                            const double* known_lu, double* lu, double* lt, int* result)
{
    if (!ctx || !result || !known_lu || !lu || !lt || nstep < 1 || iterations < 1 || maxleafpointnum < 1)
        return Fail("tdbem_dyna_gmres_solver: invalid arguments");
    std::vector<int> it((size_t)nstep), fl((size_t)nstep);
    int solved = 0;
    try {
        ctx->c.AssembleStep0(maxleafpointnum);
        ctx->c.AssembleHistory(nstep);
        ctx->c.Solve(nstep, known_lu, iterations, error, lu, lt, it.data(), fl.data(), &solved);
    } catch (const std::exception& e) {
        return Fail("tdbem_dyna_gmres_solver", e);
    }
    *result = (solved > 0 && fl[(size_t)solved - 1] == 0) ? 1 : -1;
    return 0;
}

int tdbem_timings(tdbem_ctx* ctx, double* t8)
{
    if (!ctx || !t8)
        return Fail("tdbem_timings: invalid arguments");
    for (int i = 0; i < 8; ++i)
        t8[i] = ctx->c.timing[i];
    return 0;
}

int tdbem_timings_n(tdbem_ctx* ctx, int n, double* t)
{
    if (!ctx || !t || n < 0)
        return Fail("tdbem_timings_n: invalid arguments");
    for (int i = 0; i < n; ++i)
        t[i] = i < 45 ? ctx->c.timing[i] : 0.0;
    return 0;
}

int tdbem_history_stats(tdbem_ctx* ctx, long* counts, long* npairs)
{
    if (!ctx || !counts || !npairs)
        return Fail("tdbem_history_stats: invalid arguments");
    try {
        ctx->c.HistoryStats(counts);
    } catch (const std::exception& e) {
        return Fail("tdbem_history_stats", e);
    }
    *npairs = ctx->c.npairs;
    return 0;
}

int tdbem_set_option(tdbem_ctx* ctx, const char* name, double value)
{
    if (!ctx || !name)
        return Fail("tdbem_set_option: invalid arguments");
    std::string n(name);
    if (n == "conv_tiled") {
        ctx->c.convTiled = value != 0.0;
        return 0;
    }
    if (n == "conv_tile") {
        if (value != 8.0 && value != 16.0)
            return Fail("conv_tile must be 8 or 16");
        ctx->c.convTile = (int)value;  // near lists are rebuilt by the next solve
        return 0;
    }
    if (n == "adjoint_analytic") {
        ctx->c.adjointAnalytic = value != 0.0;
        return 0;
    }
    if (n == "conv_filtered") {
        ctx->c.convFiltered = value != 0.0;
        return 0;
    }
    if (n == "conv_packed") {
        ctx->c.convPacked = value != 0.0;
        return 0;
    }
    if (n == "near_per_entry") {
        ctx->c.nearPerEntry = value != 0.0;
        return 0;
    }
    if (n == "quad_count") {
        ctx->c.quadCount = value != 0.0;
        return 0;
    }
    if (n == "history_stream") {
        if (value != -1.0 && value != 0.0 && value != 1.0)
            return Fail("tdbem_set_option: history_stream must be -1, 0 or 1");
        ctx->c.historyStream = (int)value;
        return 0;
    }
    if (n == "history_budget_gb") {
        if (!(value >= 0.0))
            return Fail("tdbem_set_option: history_budget_gb must be >= 0");
        ctx->c.historyBudget = value * 1e9;
        return 0;
    }
    if (n == "batched_fft_levels") {
        if (!(value >= 1.0 && value <= (double)tdbem::kFftMaxLevels && value == (double)(int)value))
            return Fail("tdbem_set_option: batched_fft_levels must be 1 .. " + std::to_string(tdbem::kFftMaxLevels));
        ctx->c.batchedFftLevels = (int)value;
        return 0;
    }
    if (n == "batched_fft_pace") {
        if (!(value >= 0.0 && value <= 1e6))
            return Fail("tdbem_set_option: batched_fft_pace must be 0 .. 1e6");
        ctx->c.batchedFftPace = (int)value;
        return 0;
    }
    if (n == "batched_fft_mfma") {
        if (value != 0.0 && value != 1.0)
            return Fail("tdbem_set_option: batched_fft_mfma must be 0 or 1");
        ctx->c.batchedFftMfma = (int)value;
        return 0;
    }
    if (n == "batched_fft") {
        if (value != -1.0 && value != 0.0 && value != 1.0)
            return Fail("tdbem_set_option: batched_fft must be -1, 0 or 1");
        ctx->c.batchedFft = (int)value;
        return 0;
    }
    if (n == "batched_mfma") {
        ctx->c.batchedMfma = value != 0.0;
        return 0;
    }
    if (n == "batched_far_parts") {
        if (value < 1.0 || value > (double)tdbem::kBatchFarPartsMax || value != (double)(int)value)
            return Fail("batched_far_parts must be a whole number from 1 to 16");
        ctx->c.batchedFarParts = (int)value;
        return 0;
    }
    if (n == "far_parts") {
        if (value < 1.0 || value > (double)tdbem::kFarPartsMax || value != (double)(int)value)
            return Fail("far_parts must be a whole number from 1 to 8");
        ctx->c.farParts = (int)value;
        return 0;
    }
    if (n == "gmres_blocks") {
        if (value < 1.0 || value > (double)tdbem::kGmBlocksMax || value != (double)(int)value)
            return Fail("gmres_blocks must be a whole number from 1 to 128");
        ctx->c.gmBlocks = (int)value;
        return 0;
    }
    if (n == "gmres_device") {
        if (value != 0.0 && value != 1.0 && value != 2.0)
            return Fail("gmres_device must be 0, 1 or 2");
        ctx->c.gmresDevice = (int)value;
        return 0;
    }
    if (n == "mixed") {
        if (value != 0.0 && value != 1.0 && value != 2.0)
            return Fail("mixed must be 0, 1 or 2");
        ctx->c.mixed = (int)value;
        return 0;
    }
    if (n == "freeze_structure") {
        try {
            ctx->c.FreezeStructure(value != 0.0);
        } catch (const std::exception& e) {
            return Fail("tdbem_set_option freeze_structure", e);
        }
        return 0;
    }
    if (n == "quad_staged") {
        ctx->c.quadStaged = value != 0.0;
        return 0;
    }
    return Fail("tdbem_set_option: unknown option");
}

int tdbem_model_counts(const char* path, int binary_mode, long* npoints, long* ne, long* ninf)
{
    if (!path || !npoints || !ne || !ninf)
        return Fail("tdbem_model_counts: invalid arguments");
    try {
        tdbem::ModelData m = tdbem::ReadModelFile(path, binary_mode);
        *npoints = (long)(m.points.size() / 3);
        *ne = (long)(m.elePid.size() / 8);
        *ninf = (long)(m.infPid.size() / 4);
    } catch (const std::exception& e) {
        return Fail("tdbem_model_counts", e);
    }
    return 0;
}

int tdbem_read_model(const char* path, int binary_mode, double* points, int* ele_pid, long* inf_pid)
{
    if (!path || !points || !ele_pid)
        return Fail("tdbem_read_model: invalid arguments");
    try {
        tdbem::ModelData m = tdbem::ReadModelFile(path, binary_mode);
        std::memcpy(points, m.points.data(), sizeof(double) * m.points.size());
        for (size_t i = 0; i < m.elePid.size(); ++i)
            ele_pid[i] = (int)m.elePid[i];
        if (inf_pid)
            for (size_t i = 0; i < m.infPid.size(); ++i)
                inf_pid[i] = m.infPid[i];
    } catch (const std::exception& e) {
        return Fail("tdbem_read_model", e);
    }
    return 0;
}

int tdbem_read_bd(const char* path, int binary_mode, long ne, int nstep, const int* expected_bcid, int* bcid,
                  double* lu)
{
    if (!path || !bcid || !lu || ne <= 0 || nstep < 0)
        return Fail("tdbem_read_bd: invalid arguments");
    try {
        tdbem::ReadBdFile(path, binary_mode, ne, nstep, expected_bcid, bcid, lu);
    } catch (const std::exception& e) {
        return Fail("tdbem_read_bd", e);
    }
    return 0;
}

int tdbem_write_model_bin(const char* path, long npoints, const double* points, long ne, const int* ele_pid,
                          long ninf, const long* inf_pid)
{
    if (!path || !points || !ele_pid || npoints <= 0 || ne <= 0 || ninf < 0 || (ninf > 0 && !inf_pid))
        return Fail("tdbem_write_model_bin: invalid arguments");
    try {
        std::vector<long> ids(ele_pid, ele_pid + 8 * (size_t)ne);
        tdbem::WriteModelBin(path, npoints, points, ne, ids.data(), ninf, inf_pid);
    } catch (const std::exception& e) {
        return Fail("tdbem_write_model_bin", e);
    }
    return 0;
}

int tdbem_write_bd_bin(const char* path, long ne, int nstep, const int* bcid, const double* lu)
{
    if (!path || !bcid || !lu || ne <= 0 || nstep < 0)
        return Fail("tdbem_write_bd_bin: invalid arguments");
    try {
        tdbem::WriteBdBin(path, ne, nstep, bcid, lu);
    } catch (const std::exception& e) {
        return Fail("tdbem_write_bd_bin", e);
    }
    return 0;
}

int tdbem_write_validation(tdbem_ctx* ctx, const char* dir, int nstep, const double* lu, const double* lt,
                           const int* iters, const int* flags, int solved, int solver)
{
    if (!ctx || !dir || !lu || !lt || nstep < 0 || solved < 0 || solved > nstep || (solved > 0 && (!iters || !flags)))
        return Fail("tdbem_write_validation: invalid arguments");
    try {
        tdbem::Case& c = ctx->c;
        tdbem::ValidationInput in;
        in.dir = dir;
        in.ne = c.ne;
        in.nstep = nstep;
        in.solved = solved;
        in.dt = c.dt;
        in.solver = solver;
        in.lu = lu;
        in.lt = lt;
        in.iters = iters;
        in.flags = flags;
        std::vector<double> nodes(3 * (size_t)c.ne);
        for (long e = 0; e < c.ne; ++e)
            for (int k = 0; k < 3; ++k)
                nodes[3 * (size_t)e + k] = c.geo[(size_t)e].center[k];
        in.nodes = nodes.data();
        in.domainId.resize((size_t)c.ne);
        for (long e = 0; e < c.ne; ++e)
            in.domainId[(size_t)e] = c.model.domains[(size_t)c.domOf[(size_t)e]].id;
        in.surface = c.surface;
        in.bcid = c.bcid;
        in.model = &c.model;
        in.dof = &c.dof;
        in.A = &c.A;
        in.assemblyTime = c.timing[1] + c.timing[2] + c.timing[3];
        in.gmresTime = c.timing[5];
        if (c.HistoryReady() && c.maxn > 0) {
            double mts = 0.0, mgs = 0.0;
            std::vector<int> rp, col;
            std::vector<double> val;
            for (int k = 1; k <= c.maxn; ++k) {
                c.ExportMatrix(2, k, rp, col, val);
                mts += (double)col.size();
                c.ExportMatrix(3, k, rp, col, val);
                mgs += (double)col.size();
            }
            in.avgMTSBlocks = mts / c.maxn;
            in.avgMGSBlocks = mgs / c.maxn;
        }
        tdbem::WriteValidationFiles(in);
    } catch (const std::exception& e) {
        return Fail("tdbem_write_validation", e);
    }
    return 0;
}

int tdbem_history_save(tdbem_ctx* ctx, const char* path)
{
    if (!ctx || !path)
        return Fail("tdbem_history_save: invalid arguments");
    try {
        ctx->c.SaveHistory(path);
    } catch (const std::exception& e) {
        return Fail("tdbem_history_save", e);
    }
    return 0;
}

int tdbem_history_load(tdbem_ctx* ctx, const char* path, int* maxn)
{
    if (!ctx || !path)
        return Fail("tdbem_history_load: invalid arguments");
    try {
        ctx->c.LoadHistory(path);
        if (maxn)
            *maxn = ctx->c.maxn;
    } catch (const std::exception& e) {
        return Fail("tdbem_history_load", e);
    }
    return 0;
}

int tdbem_autotune(tdbem_ctx* ctx, const char* cache_path, int* far_parts, int* gmres_blocks, double* far_ms)
{
    if (!ctx)
        return Fail("tdbem_autotune: invalid arguments");
    try {
        tdbem::TuneChoice c = ctx->c.Autotune(cache_path ? std::string(cache_path) : std::string());
        if (far_parts)
            *far_parts = c.farParts;
        if (gmres_blocks)
            *gmres_blocks = c.gmBlocks;
        if (far_ms)
            *far_ms = c.farMs;
    } catch (const std::exception& e) {
        return Fail("tdbem_autotune", e);
    }
    return 0;
}

int tdbem_release_cached_memory(void)
{
    tdbem::ReleaseCachedDeviceMemory();
    return 0;
}

int tdbem_device_synchronize(void)
{
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess)
        return Fail(hipGetErrorString(e));
    return 0;
}

}  // extern "C"
