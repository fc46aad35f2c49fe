// solver.hpp -- process-wide solver state behind the bddcml_*_c interface.
// The reference keeps the same kind of module state (levels(:), the coarse
// MUMPS instance, Krylov recycling basis; src/module_levels.f90:158-175,
// src/module_krylov.f90:34-43): one solver per process, not reentrant.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "plan.hpp"

// std::vector whose resize leaves new elements uninitialized (the large
// node tables are filled in parallel right after)
template <class T, class A = std::allocator<T>>
struct NoInitAlloc : A {
  using A::A;
  template <class U> struct rebind {
    using other = NoInitAlloc<U, typename std::allocator_traits<A>::template rebind_alloc<U>>;
  };
  template <class U> void construct(U *p) noexcept { ::new (static_cast<void *>(p)) U; }
  template <class U, class... Args> void construct(U *p, Args &&...args) {
    std::allocator_traits<A>::construct(static_cast<A &>(*this), p, std::forward<Args>(args)...);
  }
};
template <class T> using RawVec = std::vector<T, NoInitAlloc<T>>;
using RawInts = RawVec<int>;

namespace bddc {

struct Device;       // device-side plan and buffers (device.hip)
struct DeviceInput;  // assembled subdomain triplets in HBM (device.hip)

struct Timings {
  double topology = 0, plan = 0, upload = 0, factor = 0, schur_block = 0, interface = 0,
         coarse = 0, rhs = 0, krylov = 0, postprocess = 0, setup_total = 0, solve_total = 0;
  // device event times (ms) and algorithmic work of the measured launches
  double ev_interior_ms = 0, ev_schur_ms = 0, ev_interface_ms = 0, ev_coarse_ms = 0;
  double schur_flops = 0, interior_flops = 0, interface_flops = 0;
  double ev_gemv_ms = 0, gemv_bytes = 0;      // one batched S GEMV launch
  double ev_kapply_ms = 0, kapply_bytes = 0;  // one local preconditioner launch
  // GEMM launches of the interior factorization (k_tile_gemm<double>): their
  // summed event time, summed algorithmic flops and number
  double ev_gemm_ms = 0, gemm_flops = 0;
  int gemm_launches = 0;
  double level2 = 0;  // second level: mesh, host setup and device setup (s)
  int iterations = 0;
  // FP64 iterative refinement of the last solve: steps (corrections solved),
  // final relative interface residual, wall time of the refinement part (s)
  // and device time (ms) of one FP64 residual launch
  int refinements = 0;
  double refine_relres = 0, refine = 0, ev_resid_ms = 0, resid_bytes = 0;
};

struct Solver {
  // init
  int nl = 0, nsub = 0, numbase = 1, verbose = 0;
  int nsub2 = 0;  // subdomains of the next level (three or more levels)
  std::vector<int> nsublev;  // subdomains of every level, this one first
  bool initialised = false;
  // global sizes
  int nelem_g = 0, nnod_g = 0, ndof_g = 0, ndim = 3, meshdim = 3;
  int matrixtype = -1;
  std::vector<Sub> subs;
  // options
  int use_corner = 1, use_arith = 1, weights_type = 0;
  int use_user = 0;  // user constraints on edges (first level)
  int precision = 64;  // 64: FP64 everywhere; 32: FP32 preconditioner factors
  // FP32 interior factorization, Schur complements and S' factor on the
  // first level, and an outer FP64 iterative refinement of every solve
  // (residuals from the FP64 subdomain matrices): 0 off, 1 on.  The first
  // correction is solved to refine_first (0: the square root of the
  // tolerance), at most refine_max corrections follow.
  int refine = 0;
  double refine_first = 0.0;
  int refine_max = 10;
  int ordering = 0;    // interior ordering: 0 tile-aligned nested dissection, 1 RCM
  // nested dissection: largest part (dofs) left undivided (512: 128^3
  // Poisson with 16^3 subdomains 128 -> 120.5 ms per setup + solve against
  // 256; the autotuner tries 128, 256 and 512 on a caller's problem)
  int nd_leaf = 512;
  // local preconditioner through an explicit S'^{-1} (applied by a symmetric
  // tile GEMV) instead of two sweeps over the factor of S': 0 never, 1 formed
  // at setup, 2 formed once the preconditioner has been applied 40 times on
  // the same setup (the default: easy problems never pay for it)
  int explicit_inverse = 2;
  // levels below the first: interior solves of the level's apply through the
  // explicit inverse of the interior factor (tile GEMVs) instead of sweeps
  bool explicit_interior = false;
  // form of the local preconditioner launch: 0 chosen from the subdomain
  // count and sizes, 1 shallow (two tiles in flight), 2 deep (four)
  int pc_form = 0;
  // adaptive constraints: eigenvalue threshold, LOBPCG tolerance relative to
  // the pair's mean weight coefficient, iteration limit and block size
  // (src/module_adaptivity.f90:35-46)
  int use_adaptive = 0;
  double adapt_threshold = 1.5, adapt_tol = 1e-9;
  int adapt_maxit = 15, adapt_nvec = 10;
  // set while the first (arithmetic-constraint) setup of an adaptive run is
  // built: that setup forms the explicit S'^{-1}, which the preconditioner
  // of the face eigenproblems applies
  bool adapt_pass = false;
  // results per pair: face coarse node, subdomains, eigenvalues sought,
  // adaptive constraints, LOBPCG iterations; eigenvalues (largest first)
  std::vector<std::array<int, 6>> adapt_pairs;
  std::vector<std::vector<double>> adapt_eigs;
  double adapt_est = -1.0;
  // classes of subdomains with identical structure (ordering, symbolic)
  std::vector<int> ord_cls, ord_reps;
  // topology results
  int ncorner = 0, nedge = 0, nface = 0, nnodc = 0, ndofc = 0;
  std::vector<int> nndfc, kdofc;
  std::vector<double> xyzc;  // coarse node coordinates, nnodc x ndim column major
  bool bc_global = false;
  // node -> subdomains
  RawInts n2s_ptr, n2s_idx;
  // unique interface
  int nig = 0;
  std::vector<int> igdof;  // sorted global dofs
  std::vector<int> gpos;   // global dof -> unique position or -1
  std::vector<char> ig_fixed;
  // every copy of each unique interface dof: (subdomain, index in its
  // interface list), subdomains ascending (CSR over unique positions)
  std::vector<int> ig_ptr, ig_sub, ig_k;
  // state
  bool is_setup = false, solved = false, plan_only = false, reactions_ready = false;
  // the last solution (local vectors and interface) is in device memory
  // only; device_fetch_solution brings it to subs[*].sol and uig
  bool sol_on_device = false;
  // Krylov recycling: A-orthonormal directions kept across solves on the
  // same factorization (count in use, capacity fixed by the first request)
  int rec_n = 0, rec_max = 0;
  std::vector<double> uig;  // interface solution
  // global loading: the problem was uploaded as one mesh and divided here;
  // its IFIX is kept because new data must keep the fixed dofs
  bool global_loaded = false;
  std::vector<int> global_ifix;
  std::unique_ptr<Device> dev;
  std::unique_ptr<DeviceInput> din;
  // three or more levels: the next level, whose elements are this level's
  // subdomains and whose nodes are its coarse nodes; one application of its
  // BDDC replaces the coarse solve (src/module_levels.f90:3849-4249).  It
  // may have a next level of its own.
  std::unique_ptr<Solver> lev2;
  Timings t;
  int last_iters = 0, last_reason = 0;
  double last_cond = 0;
  std::vector<double> hist;

  Solver();
  ~Solver();
};

Solver &solver();
void reset_solver();

// host-side setup steps (plan.cpp)
void build_topology(Solver &S);  // node to subdomains, interior / interface split
void build_globs(Solver &S);     // components, neighbours, globs
void apply_boundary_conditions(Solver &S);
void refresh_boundary_conditions(Solver &S);  // new fixv, same fixed dofs
void build_reactions(Solver &S);              // from the solution of the last solve
void build_weights(Solver &S);
// weight coefficients rho of one subdomain at its interface dofs (reference
// interface order) for the weights type; false when its data is missing
bool weight_coefficients(const Solver &S, int is, std::vector<double> &rho);
void build_constraints(Solver &S);
void build_condensation_rhs(Solver &S, int isub_only = -1);
void build_unique_interface(Solver &S);
void build_orderings(Solver &S);  // interior/interface device orderings
// three levels (levels.cpp): coarse node coordinates, and the second level
// built from this one's coarse nodes and coarse matrices and set up
void build_coarse_nodes(Solver &S);
void setup_second_level(Solver &S);
// recursive coordinate bisection of n points (cent: n x nd, row major) into
// nparts parts: the axis of largest extent (the first on ties), points
// ordered by (coordinate, index), the first n * floor(p/2) / p to the lower
// half, parts numbered depth first; returns the part of every point
std::vector<int> bisect_points(const std::vector<double> &cent, int nd, int n, int nparts);
// global loading (global.cpp): the mesh divided into this level's
// subdomains, each uploaded as bddcml_upload_subdomain_data_c would; new
// global Dirichlet values / right-hand side / initial solution
void upload_global_data(int nelem, int nnod, int ndof, int ndim, int meshdim, const int *inet,
                        int linet, const int *nnet, int lnnet, const int *nndf, int lnndf,
                        const double *xyz, int lxyz1, int lxyz2, const int *ifix, int lifix,
                        const double *fixv, int lfixv, const double *rhs, int lrhs,
                        const double *sol, int lsol, int idelm, int neighbouring,
                        bool load_division);
void change_global_data(const int *ifix, int lifix, const double *fixv, int lfixv, const double *rhs,
                        int lrhs, const double *sol, int lsol);

// device steps (device.hip)
void device_upload_entries(Solver &S, int is);
void device_release_input(Solver &S);
// triplet patterns of subdomain i and cand[i] compared in HBM (eq[i] = 1:
// equal); false when not available (then eq is untouched)
bool device_same_patterns(Solver &S, const std::vector<int> &cand, std::vector<char> &eq);
// setup in three stages: begin builds the tile plan and launches the
// interior factorization; interface (after the constraints) plans and
// launches the interface phase and the coarse problem; finish (after the
// unique interface, Dirichlet data, weights and condensation rhs) builds the
// solve maps and lists, queues the reduced rhs and waits
// order of tile-task launches (kernels: 1 = one contiguous range of tasks
// per XCD, the default; 0 = blockIdx order)
void set_task_order(int chunked);
// interior sweeps always with the subdomain vector in global memory (testing;
// by default only subdomains whose vector does not fit in LDS use that form)
void set_trsv_global(bool on);
// forms of the sweeps and of the local preconditioner launch: 0 by the number
// of subdomains, 1 the many-subdomain forms, 2 the few-subdomain forms (testing)
void set_launch_forms(int forms);
// waves per workgroup of the interior sweeps (0 automatic, 4 or 16)
void set_sweep_waves(int waves);
// PCG iterations from the third on as replays of one captured graph (the
// default) or as separate launches
void set_pcg_graph(bool on);
// copies the solution of the last solve to the host (once; no-op otherwise)
void device_fetch_solution(Solver &S);
void device_setup_begin(Solver &S);
// after the constraints: the interface phase planned and launched
void device_setup_interface(Solver &S);
// after the weights and the condensation rhs: the rest, then the wait
void device_setup_finish(Solver &S);
// the subdomain pool holding the packed lower tiles of every S_i, with the
// first tile and the tile count of each (device.hip)
const double *device_schur_tiles(Solver &S, std::vector<int> &slow, std::vector<int> &nti);
// adaptive constraints (adapt.hip): face eigenproblems of every pair of
// subdomains sharing a face, solved on the device by batched LOBPCG on the
// set-up arithmetic-constraint preconditioner, then the eigenvectors above
// the threshold loaded as constraints of the two faces, each face block
// orthonormalized, the coarse dofs renumbered (src/module_adaptivity.f90:
// 296-3024, src/module_dd.f90:4386-4457, 4637-4802); the interface phase is
// then set up again (device_setup_interface, device_setup_finish)
void adaptive_constraints(Solver &S);
// constraints.cpp: orthonormal rows spanning the rows of M (r x n) by QR with
// column pivoting, keeping |R_kk| >= 0.01 |R_11| (src/module_dd.f90:
// 4637-4802); the rows of whole coarse nodes replaced in every subdomain
// (each node's rows over its sorted global dofs) with the coarse dofs
// renumbered; the user constraints on edges (src/module_dd.f90:4459-4565)
std::vector<std::vector<double>> orthonormal_rows(const std::vector<std::vector<double>> &M, int n);
void replace_coarse_node_rows(Solver &S, const std::map<int, std::vector<std::vector<double>>> &rows,
                              const std::map<int, std::vector<int>> &gdofs);
void add_user_constraints(Solver &S);
// symmetric eigendecomposition, host (adapt.hip): a (row major) overwritten,
// eigenvalues in w, eigenvectors in the columns of v (row major)
void sym_eig(int n, std::vector<double> &a, std::vector<double> &w, std::vector<double> &v);
// generalized inverse of a symmetric positive semidefinite matrix by
// truncated pivoted Cholesky, host (adapt.hip); returns the numerical rank
int psd_ginv(int n, const double *K, double rel_tol, double *G);
// what the preconditioner of the face eigenproblems reads from the device
// (device.hip): the explicit S'^{-1} tiles (factor pool, FP32 in the
// mixed-precision mode; packed lower by columns from each subdomain's
// pbase), Z and Phi (NI x nc column major at zoff) and host copies of the
// coarse matrices Ac_i = Phi_i^T S_i Phi_i (nconstr^2, row major)
struct PairPcSource {
  const void *inv = nullptr;
  bool inv_float = false;
  const double *Z = nullptr, *Phi = nullptr;
  std::vector<int> pbase;
  std::vector<int64_t> zoff;
  std::vector<std::vector<double>> Ac;
};
void device_pair_pc_source(Solver &S, PairPcSource &o);
void device_reduced_rhs(Solver &S);
// coarse matrices Ac_i of every subdomain (nconstr^2 each, row major) from HBM
void device_coarse_matrices(Solver &S, std::vector<std::vector<double>> &acs);
// maps between a level and its second level (which then shares its stream)
void device_link_levels(Solver &S);
void device_upload_rhs(Solver &S);  // re-upload the condensation rhs of all subdomains
void device_solve(Solver &S, int method, double tol, int maxit, int ndecrmax, bool recycling,
                  int maxvec, int *num_iter, int *reason, double *cond);
// stored recycling directions and ||V^T W - I||_F of the basis
void device_recycling_check(Solver &S, int *n, double *orth);
void device_apply_schur(Solver &S, const double *x, double *y);
void device_apply_pc(Solver &S, const double *r, double *z);
void device_reduced_rhs_download(Solver &S, double *g);
void device_sub_matrix(Solver &S, int is, int which, double *out);
void device_sub_coarse(Solver &S, int is, double *phi, double *ac);
double device_selftest_copy(long long nbytes);  // milliseconds
// grad[isegn(e)] = -lam_e^T A_e u_e over all elements (global u, lam)
void device_element_gradient(Solver &S, const double *ae, int nke, bool per_element,
                             const double *u, const double *lam, double *grad);
// the same without device allocations: its buffers live in a fixed device
// array of the library (noalloc.hip); fails on problems that exceed it
void device_element_gradient_noalloc(Solver &S, const double *ae, int nke, bool per_element,
                                     const double *u, const double *lam, double *grad);
void device_selftest_tile(int op, int npairs, const double *A, const double *B, const double *C,
                          const double *X, double alpha, double beta, double *out);
void device_release(Solver &S);

double wall_time();
// ask the runtime to spin (not sleep) in host waits on the device; called
// before the library's first device call
void device_prefer_spin_wait();

}  // namespace bddc
