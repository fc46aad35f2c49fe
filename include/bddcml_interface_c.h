/*
 * bddcml_interface_c.h -- drop-in C interface of the MI355X BDDC library.
 *
 * Every symbol keeps the name, argument list and argument meaning of the
 * reference interface it replaces (src/bddcml_interface_c.h:6-60, bound in
 * src/module_bddcml_c_wrapper.f90:37-577, semantics in
 * src/module_bddcml.f90:36-770).  All arguments are passed by pointer, as
 * the Fortran bind(c) wrapper expects.  `numbase` (0 for C, 1 for Fortran)
 * shifts every index array.
 *
 * Differences from the reference are documented per function:
 *   - one process drives one GPU; all subdomains are local to it
 *     (nsub_loc_1 is returned as nsublev[0]); comm_* handles are ignored;
 *   - matrixtype 1 (symmetric positive definite, upper triangle) is the
 *     supported matrix type of the device path;
 *   - errors print "ERROR in <ROUTINE>: <message>" to stderr and abort(),
 *     as the reference does (src/module_utils.f90:265-299).
 */
#ifndef bddcml_h
#define bddcml_h

#ifdef __cplusplus
extern "C" {
#endif

/* reference src/bddcml_interface_c.h:6-7, src/module_bddcml.f90:36-101.
 * nl >= 2 levels, nsublev[0..nl-1] non-increasing, nsublev[nl-1] = 1. */
void bddcml_init_c(int *nl, int *nsublev, int *lnsublev, int *nsub_loc_1, int *comm_init,
                   int *verbose_level, int *numbase, int *just_direct_solve_int);

/* reference src/bddcml_interface_c.h:9-13, src/module_bddcml.f90:104-183,
 * src/module_levels.f90:422-552.  The mesh is divided into nsublev[0]
 * subdomains here: with load_division_int the file partition_l1.ES of the
 * working directory, as the reference reads it; otherwise recursive
 * coordinate bisection of the element centroids (where the reference calls
 * METIS).  idelm: this library has no Fortran runtime, so it is the POSIX
 * file descriptor of the ELM file (a Fortran caller passes FNUM(unit));
 * the records are gfortran unformatted sequential ones (4-byte markers),
 * read from the start as the reference rewinds the unit.  Element
 * matrices in symmetric storage (matrixtype 1). */
void bddcml_upload_global_data_c(int *nelem, int *nnod, int *ndof, int *ndim, int *meshdim,
                                 int *inet, int *linet, int *nnet, int *lnnet, int *nndf,
                                 int *lnndf, double *xyz, int *lxyz1, int *lxyz2, int *ifix,
                                 int *lifix, double *fixv, int *lfixv, double *rhs, int *lrhs,
                                 double *sol, int *lsol, int *idelm, int *neighbouring,
                                 int *load_division_int);

/* reference src/bddcml_interface_c.h:15-27, src/module_bddcml.f90:186-356 */
void bddcml_upload_subdomain_data_c(
    int *nelem, int *nnod, int *ndof, int *ndim, int *meshdim, int *isub, int *nelems,
    int *nnods, int *ndofs, int *inet, int *linet, int *nnet, int *lnnet, int *nndf, int *lnndf,
    int *isngn, int *lisngn, int *isvgvn, int *lisvgvn, int *isegn, int *lisegn, double *xyz,
    int *lxyz1, int *lxyz2, int *ifix, int *lifix, double *fixv, int *lfixv, double *rhs,
    int *lrhs, int *is_rhs_complete, double *sol, int *lsol, int *matrixtype, int *i_sparse,
    int *j_sparse, double *a_sparse, int *la, int *is_assembled_int, double *user_constraints,
    int *luser_constraints1, int *luser_constraints2, double *element_data,
    int *lelement_data1, int *lelement_data2, double *dof_data, int *ldof_data,
    int *find_components_int, int *use_dual_mesh_graph_int, int *neighbouring);

/* reference src/bddcml_interface_c.h:29-36, src/module_bddcml.f90:359-455.
 * Weights types 0 (cardinality), 1 (diagonal stiffness) and 3 (dof data). */
void bddcml_setup_preconditioner_c(int *matrixtype, int *use_defaults_int,
                                   int *parallel_division_int, int *use_corner_constraints_int,
                                   int *use_arithmetic_constraints_int,
                                   int *use_adaptive_constraints_int,
                                   int *use_user_constraints_int, int *weights_type);

/* reference src/bddcml_interface_c.h:38-40, src/module_bddcml.f90:459-569.
 * method -1 (defaults), 0 (PCG) and 1 (BiCGSTAB). */
void bddcml_solve_c(int *comm_all, int *method, double *tol, int *maxit, int *ndecrmax,
                    int *recycling_int, int *max_number_of_stored_vectors, int *num_iter,
                    int *converged_reason, double *condition_number);

/* reference src/bddcml_interface_c.h:42, src/module_bddcml.f90:572-589 */
void bddcml_download_local_solution_c(int *isub, double *sols, int *lsols);

/* reference src/bddcml_interface_c.h:44, src/module_bddcml.f90:592-605 */
void bddcml_download_global_solution_c(double *sol, int *lsol);

/* reference src/bddcml_interface_c.h:46, src/module_bddcml.f90:608-625;
 * computation src/module_dd.f90:1258-1380: rea = A_f u (entries coupling to
 * fixed dofs, as uploaded), summed over subdomains at the interface, minus
 * the right-hand side at fixed dofs. */
void bddcml_download_local_reactions_c(int *isub, double *reas, int *lreas);

/* reference src/bddcml_interface_c.h:48, src/module_bddcml.f90:628-641,
 * src/module_levels.f90:4785-4900: rea(isvgvn) = local reactions */
void bddcml_download_global_reactions_c(double *rea, int *lrea);

/* reference src/bddcml_interface_c.h:50, src/module_bddcml.f90:644-679,
 * src/module_levels.f90:555-612: after global loading, new Dirichlet
 * values, right-hand side and initial solution (IFIX unchanged); then
 * bddcml_setup_new_data_c. */
void bddcml_change_global_data_c(int *ifix, int *lifix, double *fixv, int *lfixv, double *rhs,
                                 int *lrhs, double *sol, int *lsol);

/* reference src/bddcml_interface_c.h:52-55, src/module_bddcml.f90:682-718 */
void bddcml_change_subdomain_data_c(int *isub, int *ifix, int *lifix, double *fixv, int *lfixv,
                                    double *rhs, int *lrhs, int *is_rhs_complete, double *sol,
                                    int *lsol);

/* reference src/bddcml_interface_c.h:57, src/module_bddcml.f90:721-728.
 * After bddcml_change_subdomain_data_c: new eliminated Dirichlet values and
 * reduced right-hand side on the kept factorization
 * (src/module_levels.f90:852-903). */
void bddcml_setup_new_data_c();

/* reference src/bddcml_interface_c.h:59, src/module_bddcml.f90:731-758 */
void bddcml_dotprod_subdomain_c(int *isub, double *vec1, int *lvec1, double *vec2, int *lvec2,
                                double *dotprod);

/* reference src/bddcml_interface_c.h:61, src/module_bddcml.f90:761-770 */
void bddcml_finalize_c();

#ifdef __cplusplus
}
#endif
#endif
