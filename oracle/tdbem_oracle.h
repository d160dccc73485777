/*
 * tdbem_oracle.h -- CPU FP64 restatement of the reference multi-domain TDBEM
 * GMRES path (Ting-Ji/parallel3.0_OMP_MultiDomin_1node, DBEM1/).
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * All functions return 0 on success and a negative code on failure.
 */
#ifndef TDBEM_ORACLE_H
#define TDBEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case orc_case;

/* Build the case: geometry (GeoInit/GetNodeInfo), per-domain DynaMat, the
 * contiguous domain split, element frames and the singular self integrals.
 *   points[3*np], ele_pid[8*ne] (0-based point ids), bcid[ne] (123 or 456),
 *   mats[3*nd] = (E, v, rho) per domain (0-based domain order),
 *   itf[5*ni]  = (domainA, eleA, domainB, eleB, normalSign), 0-based.
 * dt is the global time step (the reference derives it from the global
 * material line), gap the serendipity node offset used by the regular path. */
orc_case* orc_create(int np, const double* points, int ne, const int* ele_pid,
                     const int* bcid, int nd, const double* mats, int ni,
                     const int* itf, double dt, double gap);
void orc_destroy(orc_case* c);
double orc_reference_dt(const double* points, const int* ele_pid, int ne, double beta, double E, double v,
                        double rho);
int orc_num_elements(const orc_case* c);

/* Element geometry exported for checks: centre[3], radius, frame[9]. */
int orc_element_info(const orc_case* c, int e, double* centre, double* radius, double* frame);
/* Singular self tables of element e (column-major 9 each, reference layout). */
int orc_self_tables(const orc_case* c, int e, double* u1, double* t1, double* t2);
/* DynaMat of domain d, 24 doubles in declaration order of the reference class. */
int orc_dynamat(const orc_case* c, int d, double* out24);

/* GlobalDofMap export.  counts[9] = unknown, known, equation, outerU, outerT,
 * itfU, itfT, histU, histT.  Per node arrays of length ne:
 *   u_block/u_sign/u_known, t_block/t_sign/t_known, hu_block/hu_sign,
 *   ht_block/ht_sign, preferred. */
int orc_dofmap(orc_case* c, long* counts, long* u_block, double* u_sign, int* u_known,
               long* t_block, double* t_sign, int* t_known, long* hu_block, double* hu_sign,
               long* ht_block, double* ht_sign, long* preferred);

/* Kernel probes (a5): evaluate GetDynaUij / GetDynaT1ij / GetDynaT2ij of
 * domain d at distance vector ri[3] with normal nor[3] and step n. */
int orc_kernel(orc_case* c, int d, int which, double n, const double* ri, const double* nor, double* out9);
/* Regular-path table (a3, quirk Q1): pos32 = 16 (s1, s2) points of the 4x4
 * rule, ra16 their weights, nrgv0_16 = N0(s1, s2; gap) * weight. */
int orc_regular_table(const orc_case* c, double* pos32, double* ra16, double* nrgv0_16);
/* NeedCal (a4) of field element f seen from the node of element s. */
int orc_needcal(orc_case* c, int field, int source, int type, double n);
/* Judged quadrature (a6-a8) of field f, source node s, step n:
 * which 0 = U (IntDynaUijJudge), 1 = T1, 2 = T2 (IntDynaTijJudge). Returns the
 * judge flag (0/1/2) and the column-major raw block. */
int orc_quadrature(orc_case* c, int field, int source, int which, double n, double* out9);

/* Assembly.  Step 0 builds A (unknown) and KnownM; history builds MTS[k],
 * MGS[k] for k = 1..maxn.  Each matrix is exported as block CCSR with
 * row-major 3x3 blocks; MGS is exported in full 9-value form.  Call with
 * null arrays to query nnz (returned through *nnz). */
int orc_assemble(orc_case* c, int maxn);
int orc_ccsr(orc_case* c, int which, int step, long* nrows, long* nnz,
             int* row_ptr, int* col, double* val);
/* which: 0 = A, 1 = KnownM, 2 = MTS[step], 3 = MGS[step] (full 9 values) */

/* History rows without a full assembly: MTS[k] (which 2) or MGS[k] (which 3,
 * full 9 values) restricted to source rows row0 .. row1-1, assembled exactly
 * as orc_assemble does for those rows.  row_ptr has row1-row0+1 entries
 * (starting at 0); call with null arrays for *nnz. */
int orc_history_rows(orc_case* c, int k, long row0, long row1, int which, long* nnz, int* row_ptr, int* col,
                     double* val);

/* Preconditioner export: leaf count, RePID order and leaf sizes. */
int orc_precond_info(orc_case* c, long* leaf_count, long* leaf_size, long* repid);

/* Full time loop (CS3).  known_lu[(nstep+1)*ne*3] are the .bd triplets
 * (element frame).  Outputs lu/lt[(nstep+1)*ne*3] in the global frame,
 * iters[nstep] inner GMRES iterations and flags[nstep]. Returns the number
 * of solved steps, or a negative code. */
int orc_solve(orc_case* c, int nstep, int maxn, const double* known_lu, long restart,
              double tol, long maxleaf, double* lu, double* lt, int* iters, int* flags);

/* Dense GMRES probe on a block CCSR matrix with a given leaf preconditioner
 * (used to test the solver in isolation). */
int orc_gmres_ccsr(long nrows, const int* row_ptr, const int* col, const double* val,
                   const double* b, long restart, double tol, double* x, int* iters);

#ifdef __cplusplus
}
#endif
#endif
