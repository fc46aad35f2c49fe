/*
 * bddcml_amd.h -- extra entry points of the MI355X BDDC library.
 *
 * These have no counterpart in the reference interface: they expose the
 * device operators of the interface problem for parity tests, timings for
 * benchmarks and a few options.  Vectors over the interface are ordered by
 * increasing global dof index (see bddcml_amd_interface_dofs).  All
 * functions return 0 on success and a negative value on misuse.
 */
#ifndef BDDCML_AMD_H
#define BDDCML_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* out[0]=ncorner out[1]=nedge out[2]=nface out[3]=ndofc out[4]=interface dofs */
int bddcml_amd_counts(int *out);
/* Three or more levels (bddcml_init_c with nl >= 3, nsublev[1] >= 2): the
 * second level's out[0]=subdomains out[1]=dofs (the first level's coarse dofs)
 * out[2]=ncorner out[3]=nedge out[4]=nface out[5]=ndofc out[6]=interface
 * dofs; returns -1 without a second level.  Inspection only; the second
 * level follows src/module_levels.f90:1297-1530 (setup) and 3849-4249
 * (apply). */
int bddcml_amd_level2_counts(int *out);
/* the same for level ilevel >= 2 of a setup with more levels (-1 if that
 * level does not exist: fewer levels, or a level of one subdomain above it) */
int bddcml_amd_level_counts(int ilevel, int *out);
/* Adaptive constraints (bddcml_setup_preconditioner_c with
 * use_adaptive_constraints_int = 1): with ipair < 0 the number of face
 * pairs; otherwise for pair ipair out[0] = its face's coarse node (0-based),
 * out[1], out[2] = its subdomains (numbase), out[3] = eigenvalues sought,
 * out[4] = adaptive constraints made, out[5] = LOBPCG iterations, and
 * eigs (out[3] entries, largest first) its eigenvalues; the eigenproblems of
 * src/module_adaptivity.f90:296-3024 */
int bddcml_amd_adaptive_pairs(int ipair, int *out, double *eigs);
/* sorted global dof ids (0-based) of the interface, length out[4] above */
int bddcml_amd_interface_dofs(int *out);
/* y = S x, assembled Schur complement on the interface (device) */
int bddcml_amd_apply_schur(const double *x, double *y);
/* z = M r, BDDC preconditioner (device) */
int bddcml_amd_apply_pc(const double *r, double *z);
/* reduced right-hand side g = f_G - sum A21 A11^-1 f_I (device) */
int bddcml_amd_reduced_rhs(double *g);
/* interface solution of the last solve */
int bddcml_amd_interface_solution(double *u);
/* relative residual history of the last solve; returns its length */
int bddcml_amd_residual_history(double *out, int maxn);
/* up to 29 values: wall seconds of topology plan factor krylov postprocess
 * setup_total solve_total rhs; device event milliseconds of the interior
 * factorisation, the Schur-block GEMM launch, the interface phase, the
 * coarse phase; algorithmic flops of the Schur-block launch, the interior
 * factorisation and the interface phase; event milliseconds and algorithmic
 * bytes of one batched Schur GEMV; last iteration count; event milliseconds
 * and algorithmic bytes of one local preconditioner launch; seconds of the
 * next levels' setup; the GEMM launches of the interior factorization
 * (k_tile_gemm on FP64 tiles): summed event milliseconds, summed algorithmic
 * flops (2*64^3 per tile product, post multiplies included) and their number;
 * with refinement: corrections solved, final relative interface residual,
 * wall seconds of the refinement, event milliseconds and algorithmic bytes of
 * one FP64 residual launch */
int bddcml_amd_timings(double *out, int n);
/* options (set after bddcml_init_c, which resets them):
 *   "verbose" (0/1/2);
 *   "plan_only" (1: setup stops after the host plan, no device work; for
 *     inspecting topology without a GPU);
 *   "throw_on_error" (1: errors are recorded for bddcml_amd_last_error
 *     instead of aborting);
 *   "interior_ordering" (0: tile-aligned nested dissection by coordinate
 *     bisection, the default; 1: reverse Cuthill-McKee);
 *   "nd_leaf" (largest part, in dofs, that nested dissection leaves whole;
 *     default 512);
 *   "precision" (64: FP64 throughout, the default; 32: the augmented
 *     interface factor S', its triangular sweeps and Z = S'^{-1} C^T in FP32,
 *     while the Schur operator, the Krylov method and the coarse problem
 *     stay FP64);
 *   "refinement" (1: the interior factorization, the Schur complements and
 *     S' in FP32 on the first level -- FP32 tiles, FP32 products accumulated
 *     in FP32 -- and every solve refined in FP64: the residual f - A u of the
 *     subdomain problems from the uploaded triplets in FP64, condensed by the
 *     forward sweep to the interface residual of the exact Schur complements,
 *     on which the solve stops; each correction is solved on the FP32 Schur
 *     complements.  The iteration count returned sums every correction's.
 *     Krylov recycling is not used with it, and adaptive constraints are
 *     refused.  0: off, the default);
 *   "refinement_first_tol" (with refinement: relative tolerance of the
 *     first solve; 0, the default, takes the square root of the solve's
 *     tolerance);
 *   "refinement_max" (with refinement: most corrections after the first
 *     solve, 10 by default);
 *   "task_order" (1: each XCD runs one contiguous range of a tile-task
 *     launch, the default; 0: plain workgroup order);
 *   "sweep_waves" (waves per workgroup of the interior sweeps: 0, the
 *     default, 16 below 1,024 subdomains and 4 above; 4 or 16 forced);
 *   "pc_form" (local preconditioner launch: 0, the default, the deep form
 *     below 128 subdomains or when the largest subdomain's sweep is more than
 *     twice the mean, else the shallow one; 1 shallow; 2 deep);
 *   "sweep_in_global" (1: the interior sweeps, the local preconditioner and
 *     the coarse residual keep their per-subdomain working vector in global
 *     memory -- the forms used automatically for subdomains whose vector does
 *     not fit in LDS (about 20,000 dofs, or interfaces of about 19,000);
 *     for testing);
 *   "adaptive_threshold", "adaptive_tol", "adaptive_maxit",
 *     "adaptive_nvec" (adaptive constraints: eigenvalues at or above the
 *     threshold become constraints, 1.5 by default; LOBPCG stops when every
 *     residual norm is below tol times the pair's mean weight coefficient,
 *     1e-9 by default, or after maxit iterations, 15; at most nvec
 *     eigenvectors per face, 10; the defaults of
 *     src/module_adaptivity.f90:35-46);
 *   "pcg_graph" (1, the default: PCG iterations from the third on replay
 *     one captured graph of the iteration, with the iteration index read
 *     from device memory and r.r written to pinned host memory for the
 *     stopping test; 0: every iteration as separate launches; the two give
 *     the same arithmetic in the same order);
 *   "launch_forms" (0: the interior sweeps and the local preconditioner
 *     launch take their few-subdomain forms (16 waves per sweep workgroup,
 *     the deep preconditioner form) below 128 subdomains, the default; 1:
 *     always the many-subdomain forms; 2: always the few-subdomain forms;
 *     for testing);
 *   "explicit_inverse" (S'^{-1} formed, W = L^{-1} and W^T W on the tile
 *     GEMM, so that each preconditioner apply reads it once in a symmetric
 *     GEMV instead of sweeping the factor twice: 0 never; 1 at setup; 2, the
 *     default, at the start of the first solve after the preconditioner has
 *     been applied 40 times on the same setup, where the saved apply time
 *     has about paid for forming it; a solve never changes its
 *     preconditioner while it runs) */
int bddcml_amd_set_option(const char *key, double value);
/* per subdomain (isub in numbase): out[0]=ndofi out[1]=nconstr out[2]=ndofo */
int bddcml_amd_sub_info(int isub, int *out);
/* global dof of each local interface dof (reference interface order) and
 * global coarse dof of each constraint row */
int bddcml_amd_sub_maps(int isub, int *gdofs, int *indrowc);
/* dense Schur complement of one subdomain (which = 'S'), row major, in the
 * reference interface order; Phi (ndofi x nconstr) and Ac (nconstr x nconstr) */
int bddcml_amd_sub_matrix(int isub, int which, double *out);
int bddcml_amd_sub_coarse(int isub, double *phi, double *ac);
/* message of the last error raised while option "throw_on_error" is set
 * (then errors are reported instead of aborting); returns its length and
 * clears it */
int bddcml_amd_last_error(char *buf, int len);
/* runs one batched tile task on the device on 64x64 column-major tiles:
 * op 0 NT, 1 NN, 2 TN product sums (+10 post C X^T, +20 post X C,
 * +30 post X^T C), op + 100: the same in FP32,
 * op 3 Cholesky of C - sum A B^T with inverse; out = result tile, and for
 * op 3 the inverse in out[4096..8191] */
int bddcml_amd_selftest_tile(int op, int npairs, const double *A, const double *B, const double *C,
                             const double *X, double alpha, double beta, double *out);
/* host symmetric eigendecomposition used for the Rayleigh-Ritz steps of the
 * adaptive constraints' LOBPCG (Householder tridiagonalization + implicit
 * QR): a (n x n, row major) in, eigenvalues w (unordered) and eigenvectors
 * as the columns of v (row major) out; runs on the CPU only */
int bddcml_amd_selftest_symeig(int n, const double *a, double *w, double *v);
/* host generalized inverse used for the pair coarse problems of the
 * adaptive constraints: g (n x n, row major) with k g k = k for a symmetric
 * positive semidefinite k, by Cholesky with diagonal pivoting stopped at
 * pivots below rel_tol times the largest diagonal entry (rank returned);
 * g = k^-1 when k is definite.  Runs on the CPU only. */
int bddcml_amd_selftest_ginv(int n, const double *k, double rel_tol, double *g, int *rank);
/* Element coefficient gradient for the adjoint of A(k) u = f with
 * A = sum_e k_e A_e (no counterpart in the reference; for gradient-based
 * callers): grad[isegn(e)] = -lam_e^T A_e u_e for every element e of every
 * subdomain, where u is the solution, lam the adjoint solution (A lam =
 * dJ/du, homogeneous Dirichlet values; solved with
 * bddcml_change_subdomain_data_c + bddcml_setup_new_data_c + bddcml_solve_c
 * on the same factorization), both global vectors of ndof entries.  A_e is
 * the nke x nke element matrix per unit coefficient (row major, element dofs
 * in node order, nndf dofs per node), one block shared by all elements
 * (per_element = 0) or one per global element (per_element = 1).  grad has
 * nelem entries.  Needs isegn at upload. */
int bddcml_amd_element_gradient(const double *ae, int nke, int per_element, const double *u,
                                const double *lam, double *grad);
/* The same element gradient without device allocations.  No counterpart in
 * the reference: for callers that cannot allocate device memory (loops of
 * gradient evaluations under a no-hipMalloc rule).  Its device buffers (the
 * element dof table and ids, the element matrices, u, lam and the gradient)
 * live in a fixed array of 64 MiB per GPU, held until the process ends;
 * problems that need more (about 4 x ndof + (nke + 2) x nelem doubles, with
 * nelem x nke^2 more for per_element = 1) are rejected with an error.  Calls
 * that overlap on one GPU share the array and race. */
int bddcml_amd_element_gradient_noalloc(const double *ae, int nke, int per_element,
                                        const double *u, const double *lam, double *grad);
/* Krylov recycling (bddcml_solve_c with recycling_int = 1, PCG): number
 * of directions kept for later solves on the same factorization and
 * ||V^T W - I||_F of the stored basis (V A-orthonormal, W = S V), the
 * invariant of src/module_krylov.f90:346-367; the basis is dropped by a new
 * setup or bddcml_finalize_c */
int bddcml_amd_recycling_basis(int *nvec, double *orth);
/* plain device copy of nbytes (multiple of 16) through kernel
 * k_copy_calibration, timed with events: the reference point for HBM
 * counters and bandwidth; out[0] = milliseconds of the copy */
int bddcml_amd_selftest_copy(long long nbytes, double *out);

#ifdef __cplusplus
}
#endif
#endif
