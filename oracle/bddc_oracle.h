/*
 * bddc_oracle.h -- CPU FP64 restatement of the reference BDDC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product library (bddcml_amd) never links or calls it.
 *
 * The functions mirror the reference's per-subdomain algorithm of
 * src/module_levels.f90 / src/module_dd.f90 / src/module_krylov.f90 with the
 * explicit-Schur variant of the local problems (module_dd.f90:4175-4296 with
 * explicit_schurs, dense aug factorisation module_dd.f90:5330-5341).
 * Indices passed in follow `numbase` exactly like bddcml_init_c.
 */
#ifndef BDDC_ORACLE_H
#define BDDC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef void (*orc_coarse_fn)(void *ctx, const double *rc, double *uc);
void *orc_create(int nsub, int nnod_global, int ndof_global, int ndim,
                 int meshdim, int numbase);
void orc_destroy(void *h);

/* Same argument meaning as bddcml_upload_subdomain_data_c
 * (reference src/module_bddcml.f90:186-356). Returns 0 on success. */
int orc_upload_subdomain(void *h, int isub, int nelems, int nnods, int ndofs,
                         const int *inet, int linet, const int *nnet,
                         const int *nndf, const int *isngn, const int *isvgvn,
                         const double *xyz, int lxyz1, int lxyz2,
                         const int *ifix, const double *fixv,
                         const double *rhs, int is_rhs_complete,
                         const double *sol, int matrixtype,
                         const int *i_sparse, const int *j_sparse,
                         const double *a_sparse, int la, int is_assembled,
                         const double *dof_data, int ldof_data,
                         int find_components, int use_dual_mesh_graph,
                         int neighbouring);

/* bddcml_setup_preconditioner semantics (module_levels.f90:905-1069). */
int orc_setup(void *h, int use_corner_constraints,
              int use_arithmetic_constraints, int weights_type);

/* bddcml_solve semantics: method 0 = PCG, 1 = BiCGSTAB. */
int orc_solve(void *h, int method, double tol, int maxit, int ndecrmax,
              int *num_iter, int *converged_reason, double *condition_number);

int orc_download_local_solution(void *h, int isub, double *sols, int lsols);
int orc_dotprod_subdomain(void *h, int isub, const double *v1,
                          const double *v2, int len, double *out);

/* ---- probes used by the parity tests ---------------------------------- */
/* global counts: out[0]=ncorner out[1]=nedge out[2]=nface out[3]=ndofc
 * out[4]=number of unique interface dofs */
void orc_global_counts(void *h, int *out);
/* sorted unique global dof ids (0-based) of the interface */
void orc_interface_dofs(void *h, int *out);
/* sorted global node ids (0-based) of all corners */
void orc_corner_nodes(void *h, int *out);
/* per subdomain (isub given in numbase):
 * out[0]=ndofs out[1]=ndofi out[2]=ndofo out[3]=nconstr out[4]=nnodi
 * out[5]=ncorner out[6]=nedge out[7]=nface */
void orc_sub_sizes(void *h, int isub, int *out);
/* global dof (0-based) of each local interface dof, reference order */
void orc_sub_interface_gdofs(void *h, int isub, int *out);
void orc_sub_weights(void *h, int isub, double *out);
void orc_sub_schur(void *h, int isub, double *out);      /* ndofi^2 row-major */
void orc_sub_constraints(void *h, int isub, double *C, int *indrowc); /* nconstr x ndofi */
void orc_sub_phi(void *h, int isub, double *out);        /* ndofi x nconstr row-major */
void orc_sub_coarse(void *h, int isub, double *out);     /* nconstr^2 row-major */
/* unique-interface operators (vectors ordered as orc_interface_dofs) */
void orc_apply_schur(void *h, const double *x, double *y);
void orc_apply_pc(void *h, const double *r, double *z);
void orc_reduced_rhs(void *h, double *g);
/* last solve's interface solution (unique order) */
void orc_interface_solution(void *h, double *u);
/* residual history of the last solve: n values (relres per iteration) */
int orc_residual_history(void *h, double *out, int maxn);

/* ---- three levels ------------------------------------------------------ */
/* coarse nodes of the setup: nndfc (nnodc) and their coordinates xyzc
 * (nnodc x ndim, column major): a corner's node, the mean of a glob's nodes
 * summed in ascending global node order (src/module_dd.f90:4941-4992) */
void orc_coarse_nodes(void *h, int *nndfc, double *xyzc);
/* global coarse node of each coarse node of subdomain isub (corners, then
 * globs); returns their number (out may be NULL to query it) */
int orc_sub_cnodes(void *h, int isub, int *out);
/* uc = (coarse solve)(rc) replaces the Cholesky solve of the coarse problem
 * in the preconditioner (fn NULL restores it) */
void orc_set_coarse_solver(void *h, orc_coarse_fn fn, void *ctx);
/* one application of this level's BDDC to a whole-level vector, as the
 * reference's levels >= 2 do (src/module_levels.f90:3849-4026, 4064-4249):
 * f (ndof_global, a complete right-hand side) -> x (ndof_global):
 * x_G = M (f_G - A21 A11^{-1} f_o) assembled, x_o = A11^{-1}(f_o - A12 x_G).
 * Returns 0, or -1 where Dirichlet data would be needed. */
int orc_apply_level(void *h, const double *f, double *x);
/* replace the constraint rows of one subdomain by nc dense rows over its
 * interface dofs (row major, the oracle's interface order) with global
 * coarse dofs indrowc (0-based) of a coarse problem of ndofc dofs; then
 * orc_refresh_coarse rebuilds every Phi_i, Ac_i and the coarse factor
 * (adaptive constraints, module_dd.f90:4386-4457, 4637-4802) */
int orc_set_constraints(void *h, int isub, int nc, const double *C, const int *indrowc,
                        int ndofc);
int orc_refresh_coarse(void *h);

#ifdef __cplusplus
}
#endif
#endif
