/*
 * tdbem_capi.h -- C-ABI of the MI355X multi-domain TDBEM GMRES path.
 *
 * Plain pointers and sizes only.  One context per host thread; the context
 * owns its device memory.  Every function returns 0 on success and a
 * negative code on failure, with a message in tdbem_last_error().
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference tree Ting-Ji/parallel3.0_OMP_MultiDomin_1node).
 */
#ifndef TDBEM_CAPI_H
#define TDBEM_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdbem_ctx tdbem_ctx;

/* Message of the last failing call on this thread. */
const char* tdbem_last_error(void);

/* Build the case: mesh, per-element BCID (123 / 456), the multi-domain
 * datacard tail and the global dt.  Replaces ReadOptionalMultiDomainInput's
 * consumer BuildMultiDomainModel + ApplyMultiDomainModelToElements +
 * element GeoInit/PthPhyInit (DBEM1/multidomain.cpp:1364-1531,
 * DBEM1/DBEM.cpp:1091-1160).
 *   points[3*npoints]; ele_pid[8*ne] 0-based point ids (corners then midsides)
 *   domain_count <= 0 disables the multi-domain tail (one domain);
 *   mat_rows[4*domain_count] = (id, E, v, rho) with 0- or 1-based ids;
 *   interface_count < 0 means "no InterfaceCount line" (implicit interfaces),
 *   itf_rows[5*interface_count] = (domainA, eleA, domainB, eleB, normalSign)
 *   with 0- or 1-based ids, normalised as the reference does.
 *   default_E/v/rho: global material line (used for domains without a row). */
int tdbem_create(tdbem_ctx** out, long npoints, const double* points, long ne, const int* ele_pid,
                 const int* bcid, int domain_count, const double* mat_rows, int interface_count,
                 const long* itf_rows, double default_E, double default_v, double default_rho, double dt,
                 double gap);
void tdbem_destroy(tdbem_ctx* ctx);

/* GlobalDofMap export (DBEM1/multidomain.h:103-137).  counts[9] = unknown,
 * known, equation, outerU, outerT, interfaceU, interfaceT, historyU,
 * historyT.  Arrays of length ne (node == element); signs as doubles. */
int tdbem_dofmap_export(tdbem_ctx* ctx, long* counts, long* u_block, double* u_sign, int* u_known,
                        long* t_block, double* t_sign, int* t_known, long* hu_block, double* hu_sign,
                        long* ht_block, double* ht_sign, long* preferred);

/* GlobalDofMap::Build on the host alone (DBEM1/multidomain.cpp:777-1098):
 * the same map tdbem_create builds, from the element BCIDs and the
 * multi-domain tail (arguments as for tdbem_create), without a context or a
 * device.  Outputs as for tdbem_dofmap_export; the per-node arrays may be
 * null (counts only).  Fails, as the reference's Build returns false, for a
 * BCID other than 123 / 456 or a mismatch of unknown and equation blocks. */
int tdbem_dofmap_build(long ne, const int* bcid, int domain_count, const double* mat_rows, int interface_count,
                       const long* itf_rows, long* counts, long* u_block, double* u_sign, int* u_known,
                       long* t_block, double* t_sign, int* t_known, long* hu_block, double* hu_sign,
                       long* ht_block, double* ht_sign, long* preferred);

/* MultiDomainCCSRBuilder (DBEM1/multidomain.h:139-160, DBEM1/multidomain.cpp:
 * 1101-1303) on the host: add_block stores the transpose of a column-major
 * 3x3 block times sign, skips null / zero-sign / all-zero / max|.| <= 1e-14
 * blocks and non-finite ones (counted, the first 20 reported on stderr) and
 * drops an accumulated entry that falls to max|.| <= 1e-14; merge adds another
 * builder's entries in (row, col) order (MergeFrom); build writes block CCSR
 * with rows and columns ascending and row-major blocks (Build), build_sym the
 * upper triangle (b00 b01 b02 b11 b12 b22) with *symmetric = 0 when a block
 * misses the mixed tolerance max(tolerance, 1e-14 + 1e-10 * max|block|)
 * (BuildSym).  row_ptr needs block_rows + 1 ints, col / val the nnz that
 * tdbem_builder_info reports. */
typedef struct tdbem_builder tdbem_builder;
int tdbem_builder_create(tdbem_builder** out, long block_rows, long block_cols);
void tdbem_builder_destroy(tdbem_builder* b);
int tdbem_builder_add_block(tdbem_builder* b, long row, long col, double sign, const double* block9);
int tdbem_builder_merge(tdbem_builder* b, const tdbem_builder* other);
int tdbem_builder_info(const tdbem_builder* b, long* nnz, long* skipped_nonfinite);
int tdbem_builder_build(const tdbem_builder* b, int* row_ptr, int* col, double* val);
int tdbem_builder_build_sym(const tdbem_builder* b, double tolerance, int* row_ptr, int* col, double* val6,
                            int* symmetric);

/* Step-0 assembly of A and KnownM plus the leaf preconditioner
 * (AssembleMultiDomainStep0, DBEM1/multidomain.cpp:2256-2327,2658-2693;
 * BuildMappedLeafPreConditioner 3493-3565). */
int tdbem_assemble_step0(tdbem_ctx* ctx, long maxleaf);

/* History assembly MTS[k], MGS[k] for k = 1..maxn on the device: causality
 * cull, compaction and quadrature (AssembleMultiDomainHistoryStep,
 * DBEM1/multidomain.cpp:2329-2378,2695-2722). Reports the stored band entry
 * count through *entries (may be null). */
int tdbem_assemble_history(tdbem_ctx* ctx, int maxn, long* entries);

/* Export one assembled matrix as block CCSR with row-major 3x3 blocks
 * (CCSRMat, DBEM1/mat_vec.h:220-240): which 0 = A, 1 = KnownM, 2 = MTS[step],
 * 3 = MGS[step] (symmetric blocks expanded to 9 values).  Call with null
 * arrays to get *nrows and *nnz; row_ptr needs nrows+1 ints. */
int tdbem_ccsr_export(tdbem_ctx* ctx, int which, int step, long* nrows, long* nnz, int* row_ptr, int* col,
                      double* val);

/* The rows row0 .. row1-1 of MTS[step] (which 2) or MGS[step] (which 3) in
 * the same form, copying only those rows' bands from the device (for checks
 * of sampled rows of large cases); row_ptr has row1-row0+1 entries starting
 * at 0.  Call with null arrays to get *nnz. */
int tdbem_ccsr_export_rows(tdbem_ctx* ctx, int which, int step, long row0, long row1, long* nnz, int* row_ptr,
                           int* col, double* val);

/* Singular self tables of every element (U1(0), T1(0), T2(1); column-major
 * 9 values each), for checks (DBEM1/dsquareelementdyna.cpp:239-490). */
int tdbem_self_tables(tdbem_ctx* ctx, double* u1, double* t1, double* t2);

/* Time loop with GMRES per step (DynaGMRESSolverMultiDomainCCSR time loop,
 * DBEM1/multidomain.cpp:3815-3893; gmres DBEM1/pre_gmres.cpp:1369-1605).
 *   known_lu[(nstep+1)*ne*3]: the .bd triplets (element frame);
 *   lu, lt [(nstep+1)*ne*3]: displacement / traction, global frame;
 *   iters[nstep], flags[nstep]: GMRES inner iterations and flag per step
 *     (zero for the steps after the last solved one);
 *   times[3] (may be null): seconds spent in convolution, GMRES, total.
 * Returns through *solved the number of solved steps; a GMRES flag != 0
 * stops the loop as the reference does (the call still returns 0): that
 * step counts as solved but keeps no solution, so its lu / lt and those of
 * the later steps are zero. */
int tdbem_solve(tdbem_ctx* ctx, int nstep, const double* known_lu, long restart, double tol, double* lu,
                double* lt, int* iters, int* flags, int* solved, double* times);

/* Reverse-time adjoint (no counterpart in the reference): gradient of the
 * misfit J = sum_{K=1..nstep} sum_r |lu_K[r] - obs[K][r]|^2 over the receiver
 * elements receivers[nrec] (0-based; lu in the global frame as tdbem_solve
 * returns it; obs[(nstep+1)][nrec][3]) with respect to each domain's
 * (E, v, rho): grad[3*d + p].  The forward solve, the backward sweep
 * (A^T solves with the transposed leaf preconditioner and the transposed
 * history convolution) and the gradient contraction all run on the device;
 * the history blocks' material derivatives are analytic (option
 * "adjoint_analytic"), those of the step-0 blocks and the singular self
 * tables central differences with relative step fd_step, all with the band
 * structure and judges held fixed.  Needs FP64 history
 * blocks assembled for at least nstep steps; lu (may be null) receives the
 * forward displacements [(nstep+1)][ne][3]. */
int tdbem_adjoint(tdbem_ctx* ctx, int nstep, const double* known_lu, int nrec, const int* receivers,
                  const double* obs, long restart, double tol, double fd_step, double* J, double* grad, double* lu);

/* Per-domain materials (E, v, rho), 3 per domain (no counterpart in the
 * reference, whose materials are fixed per run); get fills at most capacity
 * domains and reports the domain count.  Setting them rebuilds the
 * dynamic constants and the singular self tables; assemble step 0 and the
 * history again before solving. */
int tdbem_get_materials(tdbem_ctx* ctx, int capacity, int* ndomains, double* evr);
int tdbem_set_materials(tdbem_ctx* ctx, const double* evr);

/* Batched load cases over the same bands (no counterpart in the reference,
 * which solves one load history per run; each case follows the tdbem_solve
 * time loop above): nrhs histories known_lu[nrhs][nstep+1][ne*3] give
 * lu, lt [nrhs][nstep+1][ne*3], iters / flags [nrhs][nstep], solved[nrhs].
 * The load histories are copied in and the solutions out in 16-step chunks
 * while the time loop runs (the arrays must stay valid and unchanged until
 * the call returns, as for any call).
 * The GMRES of all cases run in lockstep; a case whose GMRES flag is != 0
 * stops there (that step and the later ones are zero, as tdbem_solve leaves
 * them) while the others continue.
 * times[3] as for tdbem_solve. */
int tdbem_solve_batched(tdbem_ctx* ctx, int nstep, int nrhs, const double* known_lu, long restart, double tol,
                        double* lu, double* lt, int* iters, int* flags, int* solved, double* times);

/* tdbem_solve / tdbem_solve_batched with the load histories already in
 * device memory and the outputs left there (no counterpart in the reference,
 * whose arrays are host memory): d_known_lu, d_lu, d_lt are device pointers
 * of the same layouts (for example from tdbem_device_alloc); iters, flags,
 * solved and times stay host arrays.  Same results as the host-buffer calls. */
int tdbem_solve_device(tdbem_ctx* ctx, int nstep, const double* d_known_lu, long restart, double tol, double* d_lu,
                       double* d_lt, int* iters, int* flags, int* solved, double* times);
int tdbem_solve_batched_device(tdbem_ctx* ctx, int nstep, int nrhs, const double* d_known_lu, long restart,
                               double tol, double* d_lu, double* d_lt, int* iters, int* flags, int* solved,
                               double* times);
/* tdbem_solve_device without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate, such as code under HIP graph
 * capture and tight serving loops.  Every working buffer of the time loop
 * (state histories, right-hand sides, the GMRES workspace) is carved from a
 * file-scope device array of 64 MiB per GPU, held until the process ends; the
 * call makes no hipMalloc / hipFree, creates no event and no graph (GMRES is
 * device-driven with direct launches).  Needs FP64 history bands assembled
 * (mixed 0) and the time-blocked convolution; a shape whose buffers exceed
 * the array is refused with an error naming the bytes it needs.  Calls that
 * overlap on one GPU share the array and race; the array keeps the last
 * call's contents (the solve clears what it needs). */
int tdbem_solve_device_noalloc(tdbem_ctx* ctx, int nstep, const double* d_known_lu, long restart, double tol,
                               double* d_lu, double* d_lt, int* iters, int* flags, int* solved);
/* Device memory for the calls above: allocate, free, and copy (kind 0 host to
 * device, 1 device to host, 2 device to device; synchronous). */
int tdbem_device_alloc(size_t bytes, void** out);
int tdbem_device_free(void* p);
int tdbem_memcpy(void* dst, const void* src, size_t bytes, int kind);

/* Independent models run together: n contexts, each on its own stream and
 * driven by its own host thread, do their whole job concurrently -- step-0
 * assembly (maxleaf), history assembly for nstep[i] steps and the time loop
 * (known_lu[i], outputs lu[i], lt[i], iters[i], flags[i], solved[i] as for
 * tdbem_solve).  The device interleaves the models' kernels, which keeps it
 * busy where one model's launches are too small to fill it.  History bands
 * held at once stay within budget_bytes (<= 0: 80 % of the free device
 * memory): a model waits, after its cull, until its bands fit next to those
 * of the running models, and releases them after its time loop (the
 * contexts keep their step-0 matrices, not their bands).  Each model's
 * results equal those of its own tdbem_assemble_step0 / _history / _solve.
 * No counterpart in the reference (one model per process there); the
 * per-model operations are those of DBEM1/multidomain.cpp:3664-3926.  On
 * failure the message names the first failing model. */
int tdbem_run_models(int n, tdbem_ctx* const* ctx, long maxleaf, const int* nstep, const double* const* known_lu,
                     long restart, double tol, double budget_bytes, double* const* lu, double* const* lt,
                     int* const* iters, int* const* flags, int* solved);

/* The same jobs as tdbem_run_models, with the models stacked: models whose
 * FP64 history bands fit budget_bytes together (0: 80 % of the free device
 * memory) and that share the step count form one round, up to 64 models and
 * 64 domains; a round is one case whose every kernel (cull, quadrature, step
 * 0, convolution, every GMRES kernel) is one launch for all its models.  Each
 * model keeps its own materials, time step, DOF numbering, preconditioner
 * leaves and GMRES (state, Hessenberg, Givens, reduction partition), so its
 * outputs are those of its own job with the same options.  The contexts'
 * options must agree and ask for FP64 bands (mixed 0).  *rounds (may be null)
 * receives the number of rounds.  No counterpart in the reference (one model
 * per process there); the per-model operations are those of
 * DBEM1/multidomain.cpp:3664-3926. */
int tdbem_run_models_grouped(int n, tdbem_ctx* const* ctx, long maxleaf, const int* nstep,
                             const double* const* known_lu, long restart, double tol, double budget_bytes,
                             double* const* lu, double* const* lt, int* const* iters, int* const* flags, int* solved,
                             int* rounds);

/* Convenience: full DynaGMRESSolverMultiDomainCCSR (DBEM1/multidomain.cpp:
 * 3664-3926): step 0, history for maxn = nstep, time loop.  Returns 1 when
 * the final GMRES flag is 0 and -1 otherwise through *result. */
int tdbem_dyna_gmres_solver(tdbem_ctx* ctx, long iterations, double error, long maxleafpointnum, int nstep,
                            const double* known_lu, double* lu, double* lt, int* result);

/* Timing of the last assembly and solve, seconds (device events where the
 * work is a kernel): t[0] singular tables, t[1] cull, t[2] history
 * quadrature, t[3] step 0, t[4] convolution (sum over steps), t[5] GMRES
 * (sum over steps), t[6] far-part convolution kernels, t[7] their count. */
int tdbem_timings(tdbem_ctx* ctx, double* t8);
/* The same timings and more, n values: t[8] the history quadrature kernels
 * alone (device events; t[2] also holds the cull and the wave packing);
 * t[9..12] the adjoint's parts (forward, backward, step-0 differences,
 * history derivatives); for tdbem_solve_batched t[13] the spectral far
 * part's kernels (window spectra, products, inverse transforms), t[14] its
 * products launches and t[15] 1 when the spectral form ran (t[6], t[7] then
 * cover the lag-by-lag far kernel of the lags <= 16), t[16] its set-up (item
 * lists and filter spectra, once per solve), t[17] its products kernels
 * alone (device events) with t[18] their algorithmic FP64 flops and t[19]
 * their algorithmic bytes, summed over the launches; for streamed history
 * bands t[20] the slab count and t[21] the far part with the slabs'
 * regeneration; with option "quad_count" 1 the staged quadrature's
 * evaluations: t[22] patch-wise T2 (entry, point), t[23] patch-wise T1 + U,
 * t[24] regular (entry, point), t[25] staged (source, point) records; per
 * level l < 4 of the spectral part t[26 + l] its products seconds, t[30 + l]
 * their flops, t[34 + l] their launches, t[38 + l] their bytes, and t[42]
 * the number of levels; with "quad_count" 1, t[43] the patch-wise entries
 * and t[44] the patch-wise waves of the last staged quadrature.  Values past the known ones are 0. */
int tdbem_timings_n(tdbem_ctx* ctx, int n, double* t);

/* Number of stored history band entries per lag j = 1..maxn
 * (counts[j-1]) and the pair count, for byte accounting of the convolution.
 * Build-side helper with no counterpart in the reference. */
int tdbem_history_stats(tdbem_ctx* ctx, long* counts, long* npairs);

/* Tuning options (no counterpart in the reference):
 *   "conv_tiled" 1 (default) time-blocked history convolution, 0 one full pass
 *   over the history bands per step (the reference's per-step summation);
 *   "conv_tile" 16 (default) or 8 steps per far-part tile of the
 *   time-blocked convolution;
 *   "adjoint_analytic" 1 (default) tdbem_adjoint takes the history blocks'
 *   material derivatives analytically (the derivative of the closed-form
 *   quadrature with the judges, patch masks and every point's front / window
 *   case held at the current materials), 0 by central differences of two
 *   quadratures per parameter; step-0 blocks and the singular self tables
 *   are differenced in both modes;
 *   "conv_filtered" 1 (default) 16-step far tiles read a pre-filtered state
 *   series and prefetch the next (pair, lags) item, 0 each item stages and
 *   filters the raw states itself (same sums in the same order);
 *   "mixed" 0 (default) FP64 throughout; 1 or 2 mixed precision for the
 *   next tdbem_assemble_history and the solves that use it: blocks computed
 *   in FP64 and stored in FP32 (1) or as FP32 plus a bf16 residual (2, about
 *   32 significant bits, 90 B per band entry instead of 120), FP32-stored
 *   inner GMRES under an FP64 iterative-refinement residual (singular
 *   tables, step-0 matrices, Hessenberg / Givens and all reductions stay
 *   FP64);
 *   "freeze_structure" 1 freezes the band structure and the judges
 *   (patch-wise / regular path of every history entry and step-0 pair) at
 *   the current materials, so assemblies after tdbem_set_materials stay on
 *   the same piece of the piecewise-smooth operator (0 releases them);
 *   "quad_staged" 1 (default) wave-per-pair history quadrature with the
 *   point terms staged in LDS as basis tensors (same band pattern, blocks
 *   within 1e-10 relative of the reference's arithmetic), 0 one lane per
 *   band entry in the reference's operation order (bit-identical to it);
 *   "batched_mfma" 0 (default) the batched far convolution of
 *   tdbem_solve_batched on FP64 vector FMAs, 1 on the FP64 matrix cores
 *   (same sums to rounding), with "batched_far_parts" 1-16 (default 16)
 *   column parts summed in order by the near part;
 *   "batched_fft" -1 (default) tdbem_solve_batched takes the raw lags >= 17
 *   of its far part in the frequency domain (16-lag sub-filters, 32-point
 *   spectra, 16-step tiles) from 16 load cases on, 1 always, 0 never;
 *   "batched_fft_levels" 1 (default) .. 3 splits those lags into levels
 *   of sub-filter length 16, 32, 64 (lags 17-32 at 16, 33-64 at 32, the
 *   last level the rest): fewer multiply-adds for the long lags, more
 *   rounding in each product (a 2L-point spectrum sums more samples);
 *   "batched_fft_mfma" 1 (default) its spectral products on the FP64
 *   matrix cores, 0 with vector FMAs; "batched_fft_pace" n > 0 makes a
 *   workgroup's wavefronts meet at a barrier every n items (0, the default:
 *   never);
 *   "history_stream" -1 (default) tdbem_assemble_history keeps the FP64
 *   bands resident when they fit the budget and otherwise streams them:
 *   the blocks of lags <= 16 stay resident and the far part of every
 *   16-step tile regenerates the rest slab by slab (source-row ranges) with
 *   the quadrature "quad_staged" selects (bit-identical to the resident
 *   bands of the same option); 1 always streams, 0 refuses bands that do
 *   not fit; streamed cases solve with tdbem_solve / tdbem_solve_device and
 *   the default convolution only; "history_budget_gb" the device memory the
 *   history may take (0, the default: 92 % of the free memory). */
int tdbem_set_option(tdbem_ctx* ctx, const char* name, double value);

/* Reference input files (host only, no device work).  binary_mode follows the
 * datacard's InputBinaryMode (DBEM1/DBEM.cpp:686-700,828-848,1023-1040):
 * 0 the text file at path, 1 the DBEMBIN file path + ".bin" (required),
 * 2 the DBEMBIN file when it exists, else the text file.
 *   tdbem_model_counts / tdbem_read_model: the .model mesh (text:
 *   DBEM1/DBEM.cpp:849-900, 1-based ids; binary: ReadModelBinary,
 *   DBEM1/DBEM.cpp:133-225); ele_pid[8 ne] and inf_pid[4 ninf] (may be null)
 *   come back 0-based;
 *   tdbem_read_bd: BCIDs and (nstep + 1) x ne triplets of the .bd file (text:
 *   DBEM1/DBEM.cpp:1010-1085; binary: ReadBoundaryBinary, DBEM1/DBEM.cpp:
 *   227-291, whose header must carry the same ne and nstep); expected_bcid
 *   (may be null) must match, as for the reference's batch boundary files. */
int tdbem_model_counts(const char* path, int binary_mode, long* npoints, long* ne, long* ninf);
int tdbem_read_model(const char* path, int binary_mode, double* points, int* ele_pid, long* inf_pid);
int tdbem_read_bd(const char* path, int binary_mode, long ne, int nstep, const int* expected_bcid, int* bcid,
                  double* lu);

/* DBEMBIN v1 writers: the layout convert_inputs_to_binary.py writes and
 * DBEM1/DBEM.cpp:89-330 reads (magic "DBEMBIN\0", version 1, kind 1 model /
 * 2 boundary, endian mark 0x01020304, reserved 0; little-endian int64 counts,
 * float64 points, int64 0-based ids; int32 BCIDs and float64 triplets). */
int tdbem_write_model_bin(const char* path, long npoints, const double* points, long ne, const int* ele_pid,
                          long ninf, const long* inf_pid);
int tdbem_write_bd_bin(const char* path, long ne, int nstep, const int* bcid, const double* lu);

/* validation_state.csv and validation_metrics.txt into directory dir
 * (WriteValidationOutputs, DBEM1/DBEM.cpp:434-561, and
 * WriteMultiDomainValidationMetrics, DBEM1/multidomain.cpp:388-416) for a
 * solve of this context: lu, lt [(nstep+1)][ne][3] in the global frame as
 * tdbem_solve returns them, its iters / flags and solved count, the
 * datacard's solver number.  The average MTS / MGS block counts are those of
 * the assembled CCSR history matrices (the bands are copied to the host). */
int tdbem_write_validation(tdbem_ctx* ctx, const char* dir, int nstep, const double* lu, const double* lt,
                           const int* iters, const int* flags, int solved, int solver);

/* History-matrix persistence (the analogue of the reference's CCSR matrix
 * cache, DBEM1/pre_gmres.cpp:16-27,199-544, and of its in-process reuse
 * across load cases): tdbem_history_save writes the assembled history bands
 * (per pair: first lag and band length; per entry: the MTS / MGS blocks in the
 * storage precision and the presence bits) with the element / pair counts,
 * dt, gap and per-domain materials; tdbem_history_load checks those against
 * the context, re-runs the causality cull and requires the same bands, then
 * uploads the blocks, so tdbem_solve / tdbem_solve_batched run without a
 * history assembly (*maxn, may be null, receives the stored step count).
 * Within a process every solve of a context already reuses its bands. */
int tdbem_history_save(tdbem_ctx* ctx, const char* path);
int tdbem_history_load(tdbem_ctx* ctx, const char* path, int* maxn);

/* Launch-shape autotuner (after tdbem_assemble_history): times, with device
 * events, the far part of the time-blocked convolution on the last 16-step
 * tile with 2, 4 or 8 wavefronts per source row or one pair's lags per work
 * item, and the GMRES dot-product reductions with 8 .. 128 blocks, then
 * applies the fastest of each to the context (options "conv_packed",
 * "far_parts", "gmres_blocks").  With cache_path (may be null) a choice
 * stored for the same case shape (elements, pairs, band entries, steps,
 * domains, storage precision) is applied without timing, and a new one is
 * appended.  Outputs (may be null): wavefronts per row (0 = one pair per
 * item), reduction blocks, time of the chosen far launch in ms (0 when
 * cached).  The shapes reassociate sums, so results change only by rounding.
 * No counterpart in the reference. */
int tdbem_autotune(tdbem_ctx* ctx, const char* cache_path, int* far_parts, int* gmres_blocks, double* far_ms);

/* Freed device buffers of 1 GB or more (history bands and the like) are
 * kept by the library and reused by the next allocation they fit (mapping
 * hundreds of GB costs seconds per call); a failed allocation returns them
 * first.  This returns every cached buffer to the device now.  No
 * counterpart in the reference (host memory there). */
int tdbem_release_cached_memory(void);

/* Wait for all device work of this process (hipDeviceSynchronize). */
int tdbem_device_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
