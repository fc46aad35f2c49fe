// api.cpp -- the bddcml_*_c entry points (reference src/bddcml_interface_c.h,
// src/module_bddcml.f90, src/module_levels.f90:180-903) on top of the
// host plan and the device pipelines.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>

#include "../../include/bddcml_amd.h"
#include "../../include/bddcml_interface_c.h"
#include "solver.hpp"

namespace bddc {

static Solver *g_solver = nullptr;
Solver &solver() {
  if (!g_solver) g_solver = new Solver();
  return *g_solver;
}
void reset_solver() {
  if (g_solver) {
    device_release(*g_solver);
    delete g_solver;
  }
  g_solver = nullptr;
}

static std::string g_last_error;
static void record_error(const Error &e) { g_last_error = e.what(); }

static Sub &local_sub(Solver &S, int isub, const char *routine) {
  int is = isub - S.numbase;
  if (is < 0 || is >= S.nsub)
    fatal(routine, "Index of subdomain not found among local subdomains: " + std::to_string(isub));
  return S.subs[is];
}

}  // namespace bddc

using namespace bddc;

extern "C" {

void bddcml_init_c(int *nl, int *nsublev, int *lnsublev, int *nsub_loc_1, int *comm_init,
                   int *verbose_level, int *numbase, int *just_direct_solve_int) try {
  (void)comm_init;
  // the reference's global direct solve (src/module_levels.f90:469-478) is
  // not built: the problem is solved by BDDC, whose single-subdomain case
  // (nsublev(1) = 1, where the reference also falls back to the direct
  // solve) is the interior factorization and its sweeps, a direct solve
  if (*just_direct_solve_int != 0 && *verbose_level > 0)
    std::fprintf(stderr, "WARNING in BDDCML_INIT: the global direct solve is not built; "
                         "the problem is solved by BDDC\n");
  device_prefer_spin_wait();
  reset_solver();
  set_task_order(1);
  set_trsv_global(false);
  set_launch_forms(0);
  set_sweep_waves(0);
  set_pcg_graph(true);
  Solver &S = solver();
  if (*nl < 2) fatal("LEVELS_INIT", "Number of levels must be at least 2");
  if (*lnsublev < *nl) fatal("LEVELS_INIT", "lnsublev smaller than number of levels");
  if (nsublev[*nl - 1] != 1) fatal("LEVELS_INIT", "Number of subdomains at last level must be 1.");
  for (int l = 1; l < *nl - 1; l++)
    if (nsublev[l] < 1 || nsublev[l] > nsublev[l - 1])
      fatal("LEVELS_INIT", "Number of subdomains on level " + std::to_string(l + 1) +
                               " must be between 1 and that of level " + std::to_string(l));
  S.nl = *nl;
  S.nsub = nsublev[0];
  S.nsublev.assign(nsublev, nsublev + *nl);
  S.nsub2 = *nl >= 3 ? nsublev[1] : 0;
  S.numbase = *numbase;
  S.verbose = *verbose_level;
  S.subs.assign(S.nsub, Sub());
  *nsub_loc_1 = S.nsub;
  S.initialised = true;
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_upload_global_data_c(int *nelem, int *nnod, int *ndof, int *ndim, int *meshdim,
                                 int *inet, int *linet, int *nnet, int *lnnet, int *nndf,
                                 int *lnndf, double *xyz, int *lxyz1, int *lxyz2, int *ifix,
                                 int *lifix, double *fixv, int *lfixv, double *rhs, int *lrhs,
                                 double *sol, int *lsol, int *idelm, int *neighbouring,
                                 int *load_division_int) try {
  upload_global_data(*nelem, *nnod, *ndof, *ndim, *meshdim, inet, *linet, nnet, *lnnet, nndf, *lnndf,
                     xyz, *lxyz1, *lxyz2, ifix, *lifix, fixv, *lfixv, rhs, *lrhs, sol, *lsol, *idelm,
                     *neighbouring, *load_division_int != 0);
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_upload_subdomain_data_c(
    int *nelem, int *nnod, int *ndof, int *ndim, int *meshdim, int *isub, int *nelems,
    int *nnods, int *ndofs, int *inet, int *linet, int *nnet, int *lnnet, int *nndf, int *lnndf,
    int *isngn, int *lisngn, int *isvgvn, int *lisvgvn, int *isegn, int *lisegn, double *xyz,
    int *lxyz1, int *lxyz2, int *ifix, int *lifix, double *fixv, int *lfixv, double *rhs,
    int *lrhs, int *is_rhs_complete, double *sol, int *lsol, int *matrixtype, int *i_sparse,
    int *j_sparse, double *a_sparse, int *la, int *is_assembled_int, double *user_constraints,
    int *luser_constraints1, int *luser_constraints2, double *element_data,
    int *lelement_data1, int *lelement_data2, double *dof_data, int *ldof_data,
    int *find_components_int, int *use_dual_mesh_graph_int, int *neighbouring) try {
  (void)element_data;
  (void)lelement_data1;
  (void)lelement_data2;
  (void)is_assembled_int;
  const char *R = "LEVELS_UPLOAD_SUBDOMAIN_DATA";
  Solver &S = solver();
  if (!S.initialised) fatal(R, "bddcml_init was not called");
  S.nelem_g = *nelem;
  S.nnod_g = *nnod;
  S.ndof_g = *ndof;
  S.ndim = *ndim;
  S.meshdim = *meshdim;
  // matrixtype 1 (SPD) and 2 (symmetric general) in upper-triangle
  // storage; 0 (general) in full storage is accepted when the matrix is
  // symmetric and folded to its upper triangle.  The factorizations are
  // Cholesky: an indefinite matrix stops the setup with an error.
  if (*matrixtype < 0 || *matrixtype > 2) fatal(R, "unknown matrixtype " + std::to_string(*matrixtype));
  if (S.matrixtype >= 0 && S.matrixtype != *matrixtype) fatal(R, "matrixtype differs between subdomains");
  S.matrixtype = *matrixtype;
  Sub &s = local_sub(S, *isub, R);
  if (s.uploaded) fatal(R, "It appears that mesh was already loaded for subdomain " + std::to_string(*isub));
  const int b = S.numbase;
  s = Sub();
  s.isub = *isub - b;
  s.nelems = *nelems;
  s.nnods = *nnods;
  s.ndofs = *ndofs;
  if (*lnnet != *nelems || *lnndf != *nnods || *lisngn != *nnods || *lisvgvn != *ndofs)
    fatal(R, "array length mismatch");
  if (*lxyz1 != *nnods || *lxyz2 != *ndim) fatal(R, "coordinate array dimensions mismatch");
  if (*lifix != *ndofs || *lfixv != *ndofs || *lrhs != *ndofs || *lsol != *ndofs)
    fatal(R, "length of IFIX, FIXV, RHS or SOL differs from NDOFS");
  s.inet.resize(*linet);
  for (int k = 0; k < *linet; k++) s.inet[k] = inet[k] - b;
  s.nnet.assign(nnet, nnet + *nelems);
  if (*lisegn == *nelems) {
    s.isegn.resize(*nelems);
    for (int k = 0; k < *nelems; k++) s.isegn[k] = isegn[k] - b;
  }
  s.nndf.assign(nndf, nndf + *nnods);
  s.isngn.resize(*nnods);
  for (int k = 0; k < *nnods; k++) s.isngn[k] = isngn[k] - b;
  s.isvgvn.resize(*ndofs);
  for (int k = 0; k < *ndofs; k++) s.isvgvn[k] = isvgvn[k] - b;
  for (int g : s.isngn)
    if (g < 0 || g >= *nnod) fatal("BDDCML_UPLOAD_SUBDOMAIN_DATA", "global node index out of range");
  for (int g : s.isvgvn)
    if (g < 0 || g >= *ndof) fatal("BDDCML_UPLOAD_SUBDOMAIN_DATA", "global dof index out of range");
  s.xyz.assign(xyz, xyz + (size_t)*nnods * *ndim);
  s.ifix.assign(ifix, ifix + *ndofs);
  s.fixv.assign(fixv, fixv + *ndofs);
  s.rhs.assign(rhs, rhs + *ndofs);
  s.sol.assign(sol, sol + *ndofs);
  s.has_fix = std::any_of(s.ifix.begin(), s.ifix.end(), [](int v) { return v != 0; });
  s.rhs_complete = *is_rhs_complete != 0;
  s.dof_data.assign(dof_data, dof_data + *ldof_data);
  // user constraints: luser_constraints1 rows, one value per subdomain node
  // (src/module_dd.f90:3188-3232)
  if (*luser_constraints1 > 0 && *luser_constraints2 > 0) {
    if (*luser_constraints2 != *nnods) fatal(R, "Array USER_CONSTRAINTS dimension mismatch.");
    s.nuser_c = *luser_constraints1;
    s.user_c.assign(user_constraints, user_constraints + (size_t)*luser_constraints1 * *luser_constraints2);
  }
  s.find_components = *find_components_int;
  s.use_dual = *use_dual_mesh_graph_int;
  s.neighbouring = *neighbouring;
  s.kdof.assign(*nnods + 1, 0);
  for (int k = 0; k < *nnods; k++) s.kdof[k + 1] = s.kdof[k] + nndf[k];
  if (s.kdof[*nnods] != *ndofs) fatal(R, "sum of NNDF differs from NDOFS");
  // assemble triplets: sort by (row, col), sum duplicates (src/module_sm.f90:630-685).
  // A stable counting sort by row, then a stable insertion sort of each
  // row's short list by column: O(la) for element-assembled input, and
  // duplicates are summed in input order
  const int n = *la;
  const int nd = *ndofs;
  const bool full = *matrixtype == 0;
  for (int k = 0; k < n; k++) {
    if (!full && i_sparse[k] > j_sparse[k])
      fatal(R, "Matrix in wrong format - only upper triangle is expected for symmetric problems, matrixtype = " +
                   std::to_string(*matrixtype));
    const int i = i_sparse[k] - b, j = j_sparse[k] - b;
    if (i < 0 || j < 0 || i >= nd || j >= nd) fatal(R, "matrix index out of range");
  }
  std::vector<int> rowp(nd + 1, 0), ord(n);
  for (int k = 0; k < n; k++) rowp[i_sparse[k] - b + 1]++;
  for (int r = 0; r < nd; r++) rowp[r + 1] += rowp[r];
  {
    std::vector<int> fill(rowp.begin(), rowp.end() - 1);
    for (int k = 0; k < n; k++) ord[fill[i_sparse[k] - b]++] = k;
  }
  for (int r = 0; r < nd; r++) {
    if (rowp[r + 1] - rowp[r] > 64) {  // long rows: merge sort
      std::stable_sort(ord.begin() + rowp[r], ord.begin() + rowp[r + 1],
                       [&](int x, int y) { return j_sparse[x] < j_sparse[y]; });
      continue;
    }
    for (int q = rowp[r] + 1; q < rowp[r + 1]; q++) {
      const int k = ord[q], jk = j_sparse[k];
      int p = q;
      for (; p > rowp[r] && j_sparse[ord[p - 1]] > jk; p--) ord[p] = ord[p - 1];
      ord[p] = k;
    }
  }
  s.ai.reserve(n);
  s.aj.reserve(n);
  s.av.reserve(n);
  for (int q = 0; q < n; q++) {
    const int k = ord[q];
    const int i = i_sparse[k] - b, j = j_sparse[k] - b;
    if (!s.ai.empty() && s.ai.back() == i && s.aj.back() == j)
      s.av.back() += a_sparse[k];
    else
      s.ai.push_back(i), s.aj.push_back(j), s.av.push_back(a_sparse[k]);
  }
  if (full) {
    // both triangles assembled: every (i, j) needs its mirror (j, i) with the
    // same value (to 1e-12 of the largest entry: sums of the same element
    // contributions in another order round differently), then the upper
    // triangle is kept
    double amax = 0.0;
    for (double v : s.av) amax = std::max(amax, std::fabs(v));
    std::vector<int> rp(nd + 1, 0);
    for (int i : s.ai) rp[i + 1]++;
    for (int r = 0; r < nd; r++) rp[r + 1] += rp[r];
    for (size_t q = 0; q < s.ai.size(); q++) {
      const int i = s.ai[q], j = s.aj[q];
      if (i == j) continue;
      auto bgn = s.aj.begin() + rp[j], end = s.aj.begin() + rp[j + 1];
      auto it = std::lower_bound(bgn, end, i);
      const double v = s.av[q];
      const double m = (it == end || *it != i) ? 0.0 : s.av[it - s.aj.begin()];
      if (std::fabs(v - m) > 1e-12 * amax)
        fatal(R, "matrixtype 0: the matrix is not symmetric; the device path factors symmetric "
                 "positive definite matrices");
    }
    size_t o = 0;
    for (size_t q = 0; q < s.ai.size(); q++)
      if (s.ai[q] <= s.aj[q]) s.ai[o] = s.ai[q], s.aj[o] = s.aj[q], s.av[o] = s.av[q], o++;
    s.ai.resize(o), s.aj.resize(o), s.av.resize(o);
  }
  // a fixed dof without a diagonal entry gets an explicit zero one; the
  // Dirichlet elimination turns it into 1 (src/module_sm.f90:556-585)
  {
    std::vector<char> has_diag(s.ndofs, 0);
    for (size_t k = 0; k < s.ai.size(); k++)
      if (s.ai[k] == s.aj[k]) has_diag[s.ai[k]] = 1;
    bool added = false;
    for (int d = 0; d < s.ndofs; d++)
      if (s.ifix[d] && !has_diag[d]) {
        s.ai.push_back(d), s.aj.push_back(d), s.av.push_back(0.0);
        added = true;
      }
    if (added) {
      std::vector<size_t> o(s.ai.size());
      std::iota(o.begin(), o.end(), 0);
      std::sort(o.begin(), o.end(), [&](size_t x, size_t y) {
        return s.ai[x] != s.ai[y] ? s.ai[x] < s.ai[y] : s.aj[x] < s.aj[y];
      });
      std::vector<int> ai(o.size()), aj(o.size());
      std::vector<double> av(o.size());
      for (size_t q = 0; q < o.size(); q++) ai[q] = s.ai[o[q]], aj[q] = s.aj[o[q]], av[q] = s.av[o[q]];
      s.ai.swap(ai), s.aj.swap(aj), s.av.swap(av);
    }
  }
  // the assembled values go to HBM now; setup permutes and eliminates them there
  if (!S.plan_only) device_upload_entries(S, s.isub);
  s.uploaded = true;
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_setup_preconditioner_c(int *matrixtype, int *use_defaults_int,
                                   int *parallel_division_int, int *use_corner_constraints_int,
                                   int *use_arithmetic_constraints_int,
                                   int *use_adaptive_constraints_int,
                                   int *use_user_constraints_int, int *weights_type) try {
  (void)parallel_division_int;
  const char *R = "BDDCML_SETUP_PRECONDITIONER";
  Solver &S = solver();
  double t0 = wall_time();
  if (*matrixtype < 0 || *matrixtype > 2) fatal(R, "unknown matrixtype " + std::to_string(*matrixtype));
  for (auto &s : S.subs)
    if (!s.uploaded) fatal("LEVELS_PC_SETUP", "not all subdomains were uploaded");
  if (*use_defaults_int) {
    S.use_corner = 1;
    S.use_arith = 1;
    S.weights_type = 0;
    S.use_adaptive = 0;
    S.use_user = 0;
  } else {
    S.use_corner = *use_corner_constraints_int != 0;
    // adaptive constraints are an add-on to the arithmetic ones
    // (src/module_levels.f90:2699-2706)
    S.use_adaptive = *use_adaptive_constraints_int != 0;
    S.use_arith = *use_arithmetic_constraints_int != 0 || S.use_adaptive;
    S.weights_type = *weights_type;
    S.use_user = *use_user_constraints_int != 0;
    if (S.use_adaptive && S.nl >= 3 && S.nsub2 >= 2)
      fatal(R, "adaptive constraints are supported with two levels");
  }
  if (S.use_adaptive && S.refine)
    fatal(R, "adaptive constraints need the FP64 Schur complements (option refinement = 0)");
  double tp = wall_time();
  auto phase = [&](const char *name) {
    double now = wall_time();
    if (S.verbose >= 2) std::printf(" BDDCML_SETUP_PRECONDITIONER: %-22s %8.4f s\n", name, now - tp);
    tp = now;
  };
  // what the interior factorization needs: the interior / interface split
  // and the orderings; globs and constraints follow while the device factors
  build_topology(S);
  phase("topology");
  build_orderings(S);
  phase("orderings");
  S.t.topology += wall_time() - t0;
  if (!S.plan_only) {
    device_setup_begin(S);  // returns with the interior factorization running
    tp = wall_time();
  }
  // host work that overlaps the device factorization
  double th = wall_time();
  build_globs(S);
  phase("globs");
  build_constraints(S);
  if (S.use_user) add_user_constraints(S);
  phase("constraints");
  // the interface phase needs the constraints only: planned and launched
  // now, while the host builds the interface numbering and the weights
  S.adapt_pass = S.use_adaptive != 0;
  if (!S.plan_only) {
    const double tdi = wall_time();
    device_setup_interface(S);
    th += wall_time() - tdi;  // counted in the plan and factor timings
    phase("interface phase launched");
  }
  build_unique_interface(S);  // the exchange sums below run over it
  phase("unique interface");
  apply_boundary_conditions(S);
  phase("boundary conditions");
  build_weights(S);
  phase("weights");
  build_condensation_rhs(S);
  phase("condensation rhs");
  S.t.topology += wall_time() - th;
  if (S.plan_only) {
    if (S.nl >= 3 && S.nsub2 >= 2) setup_second_level(S);  // its host plan
    S.is_setup = true;
    return;
  }
  device_setup_finish(S);
  // adaptive constraints: face eigenproblems on the preconditioner just set
  // up, then the interface phase again with the new constraints
  // (src/module_levels.f90:2826-2945)
  if (S.use_adaptive) {
    adaptive_constraints(S);
    S.adapt_pass = false;
    phase("adaptive constraints");
    device_setup_interface(S);
    device_setup_finish(S);
    phase("interface phase again");
  }
  // three or more levels: the next level from the coarse nodes and matrices,
  // recursively (a next level of one subdomain: the coarse problem is solved
  // directly)
  if (S.nl >= 3 && S.nsub2 >= 2) {
    setup_second_level(S);
    phase("second level");
  }
  double t1 = wall_time();
  device_reduced_rhs(S);
  S.t.rhs += wall_time() - t1;
  S.is_setup = true;
  S.t.setup_total = wall_time() - t0;
  if (S.verbose > 0)
    std::printf(" BDDCML_SETUP_PRECONDITIONER: ncorner %d nedge %d nface %d ndofc %d interface %d\n",
                S.ncorner, S.nedge, S.nface, S.ndofc, S.nig);
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_solve_c(int *comm_all, int *method, double *tol, int *maxit, int *ndecrmax,
                    int *recycling_int, int *max_number_of_stored_vectors, int *num_iter,
                    int *converged_reason, double *condition_number) try {
  (void)comm_all;
  Solver &S = solver();
  if (!S.is_setup) fatal("BDDCML_SOLVE", "preconditioner not set up");
  // defaults of src/module_bddcml.f90:515-520
  int m = 1;
  double tl = 1e-6;
  int mi = 1000, nd = 30, maxvec = 100;
  bool rec = false;
  if (*method != -1) {
    m = *method;
    tl = *tol;
    mi = *maxit;
    nd = *ndecrmax;
    rec = *recycling_int != 0;
    maxvec = *max_number_of_stored_vectors;
  }
  if (m != 0 && m != 1) fatal("BDDCML_SOLVE", "unknown iterative method " + std::to_string(m));
  if (rec && maxvec < 0) fatal("BDDCML_SOLVE", "negative number of stored vectors");
  double t0 = wall_time();
  device_solve(S, m, tl, mi, nd, rec, maxvec, num_iter, converged_reason, condition_number);
  S.t.solve_total = wall_time() - t0;
  S.last_iters = *num_iter;
  S.last_reason = *converged_reason;
  S.last_cond = *condition_number;
  S.solved = true;
  S.reactions_ready = false;
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_download_local_solution_c(int *isub, double *sols, int *lsols) try {
  Solver &S = solver();
  Sub &s = local_sub(S, *isub, "LEVELS_DD_DOWNLOAD_SOLUTION");
  if (*lsols != s.ndofs) fatal("LEVELS_DD_DOWNLOAD_SOLUTION", "solution size mismatch");
  device_fetch_solution(S);
  std::copy(s.sol.begin(), s.sol.end(), sols);
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_download_global_solution_c(double *sol, int *lsol) try {
  Solver &S = solver();
  if (*lsol != S.ndof_g) fatal("LEVELS_GET_GLOBAL_SOLUTION", "solution size mismatch");
  device_fetch_solution(S);
  for (auto &s : S.subs)
    for (int d = 0; d < s.ndofs; d++) sol[s.isvgvn[d]] = s.sol[d];
}
catch (const Error &e) {
  record_error(e);
}

// reactions of the last solve, computed on the first request after it
// (the reference computes them in levels_postprocess_solution)
static void ensure_reactions(Solver &S, const char *R) {
  if (!S.solved) fatal(R, "Reactions are not computed: no solution yet");
  device_fetch_solution(S);
  if (!S.reactions_ready) {
    build_reactions(S);
    S.reactions_ready = true;
  }
}

void bddcml_download_local_reactions_c(int *isub, double *reas, int *lreas) try {
  const char *R = "LEVELS_DD_DOWNLOAD_REACTIONS";
  Solver &S = solver();
  Sub &s = local_sub(S, *isub, R);
  if (*lreas != s.ndofs) fatal("DD_DOWNLOAD_REACTIONS", "Array REA dimension mismatch.");
  ensure_reactions(S, R);
  std::copy(s.rea.begin(), s.rea.end(), reas);
}
catch (const Error &e) {
  record_error(e);
}
void bddcml_download_global_reactions_c(double *rea, int *lrea) try {
  const char *R = "LEVELS_GET_GLOBAL_REACTIONS";
  Solver &S = solver();
  if (*lrea != S.ndof_g) fatal(R, "Space for global reactions not properly allocated on root.");
  ensure_reactions(S, R);
  std::fill(rea, rea + S.ndof_g, 0.0);
  for (auto &s : S.subs)
    for (int d = 0; d < s.ndofs; d++) rea[s.isvgvn[d]] = s.rea[d];
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_change_global_data_c(int *ifix, int *lifix, double *fixv, int *lfixv, double *rhs,
                                 int *lrhs, double *sol, int *lsol) try {
  change_global_data(ifix, *lifix, fixv, *lfixv, rhs, *lrhs, sol, *lsol);
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_change_subdomain_data_c(int *isub, int *ifix, int *lifix, double *fixv, int *lfixv,
                                    double *rhs, int *lrhs, int *is_rhs_complete, double *sol,
                                    int *lsol) try {
  const char *R = "LEVELS_CHANGE_SUBDOMAIN_DATA";
  Solver &S = solver();
  if (!S.is_setup) fatal(R, "Cannot change data at this moment, level is not ready.");
  device_fetch_solution(S);
  Sub &s = local_sub(S, *isub, R);
  if (*lifix != s.ndofs || *lfixv != s.ndofs || *lrhs != s.ndofs || *lsol != s.ndofs)
    fatal(R, "array length mismatch");
  for (int d = 0; d < s.ndofs; d++)
    if ((ifix[d] != 0) != (s.ifix[d] != 0))
      fatal("DD_UPLOAD_BC", "Indices of fixed variables IFIX cannot be changed for subdomain " +
                                std::to_string(*isub));
  s.fixv.assign(fixv, fixv + s.ndofs);
  s.rhs.assign(rhs, rhs + s.ndofs);
  s.sol.assign(sol, sol + s.ndofs);
  s.rhs_complete = *is_rhs_complete != 0;
  s.uploaded = true;
  s.bc.clear();  // marks the subdomain for BC recomputation
}
catch (const Error &e) {
  record_error(e);
}

// reference src/module_levels.f90:852-903: new eliminated Dirichlet values
// and reduced right-hand side; the factorizations are kept
void bddcml_setup_new_data_c() try {
  const char *R = "LEVELS_SETUP_NEW_DATA";
  Solver &S = solver();
  if (!S.is_setup) fatal(R, "Cannot change data at this moment, level is not ready.");
  double t0 = wall_time();
  refresh_boundary_conditions(S);
  build_condensation_rhs(S);
  if (!S.plan_only) {
    device_upload_rhs(S);
    device_reduced_rhs(S);
  }
  S.t.rhs += wall_time() - t0;
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_dotprod_subdomain_c(int *isub, double *vec1, int *lvec1, double *vec2, int *lvec2,
                                double *dotprod) try {
  // reference src/module_dd.f90:12736-12788: interface dofs (weighted, in
  // interface order) first, then interior dofs
  const char *R = "DD_DOTPROD_SUBDOMAIN_LOCAL";
  Solver &S = solver();
  Sub &s = local_sub(S, *isub, "LEVELS_DD_DOTPROD_SUBDOMAIN_LOCAL");
  if (!S.is_setup || s.w.size() != s.ifc_dof.size()) fatal(R, "Weights not ready.");
  if (*lvec1 != *lvec2) fatal(R, "Dimensions mismatch.");
  if (*lvec1 != s.ndofs) fatal(R, "Dimensions mismatch with local subdomain size.");
  double sum = 0.0;
  for (size_t k = 0; k < s.ifc_dof.size(); k++) {
    const int d = s.ifc_dof[k];
    sum += vec1[d] * s.w[k] * vec2[d];
  }
  for (int d : s.int_dof) sum += vec1[d] * vec2[d];
  *dotprod = sum;
}
catch (const Error &e) {
  record_error(e);
}

void bddcml_finalize_c() try { reset_solver(); }
catch (const Error &e) {
  record_error(e);
}

// ------------------------------------------------------------- extra ABI
int bddcml_amd_level_counts(int ilevel, int *out) try {
  Solver &S = solver();
  if (!S.is_setup || ilevel < 2) return -1;
  const Solver *L = &S;
  for (int l = 1; l < ilevel; l++) {
    L = L->lev2.get();
    if (!L) return -1;
  }
  const Solver &T = *L;
  out[0] = T.nsub;
  out[1] = T.ndof_g;
  out[2] = T.ncorner;
  out[3] = T.nedge;
  out[4] = T.nface;
  out[5] = T.ndofc;
  out[6] = T.nig;
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -1;
}

int bddcml_amd_adaptive_pairs(int ipair, int *out, double *eigs) try {
  Solver &S = solver();
  if (!S.is_setup) return -1;
  if (ipair < 0) return (int)S.adapt_pairs.size();
  if (ipair >= (int)S.adapt_pairs.size()) return -2;
  const auto &q = S.adapt_pairs[ipair];
  for (int k = 0; k < 6; k++) out[k] = q[k];
  out[1] += S.numbase;
  out[2] += S.numbase;
  if (eigs)
    for (size_t k = 0; k < S.adapt_eigs[ipair].size(); k++) eigs[k] = S.adapt_eigs[ipair][k];
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -100;
}

int bddcml_amd_level2_counts(int *out) try {
  Solver &S = solver();
  if (!S.is_setup || !S.lev2) return -1;
  const Solver &T = *S.lev2;
  out[0] = T.nsub;
  out[1] = T.ndof_g;
  out[2] = T.ncorner;
  out[3] = T.nedge;
  out[4] = T.nface;
  out[5] = T.ndofc;
  out[6] = T.nig;
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -1;
}

int bddcml_amd_counts(int *out) try {
  Solver &S = solver();
  if (!S.is_setup) return -1;
  out[0] = S.ncorner;
  out[1] = S.nedge;
  out[2] = S.nface;
  out[3] = S.ndofc;
  out[4] = S.nig;
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_interface_dofs(int *out) try {
  Solver &S = solver();
  if (!S.is_setup) return -1;
  std::copy(S.igdof.begin(), S.igdof.end(), out);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_apply_schur(const double *x, double *y) try {
  Solver &S = solver();
  if (!S.is_setup || S.plan_only || !S.dev) return -1;
  device_apply_schur(S, x, y);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_apply_pc(const double *r, double *z) try {
  Solver &S = solver();
  if (!S.is_setup || S.plan_only || !S.dev) return -1;
  device_apply_pc(S, r, z);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_reduced_rhs(double *g) try {
  Solver &S = solver();
  if (!S.is_setup || S.plan_only || !S.dev) return -1;
  device_reduced_rhs_download(S, g);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_recycling_basis(int *nvec, double *orth) try {
  Solver &S = solver();
  if (!S.is_setup || S.plan_only || !S.dev) return -1;
  device_recycling_check(S, nvec, orth);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}

int bddcml_amd_interface_solution(double *u) try {
  Solver &S = solver();
  if (!S.solved) return -1;
  device_fetch_solution(S);
  std::copy(S.uig.begin(), S.uig.end(), u);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_residual_history(double *out, int maxn) try {
  Solver &S = solver();
  int n = (int)S.hist.size();
  if (out)
    for (int k = 0; k < std::min(n, maxn); k++) out[k] = S.hist[k];
  return n;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_timings(double *out, int n) try {
  Solver &S = solver();
  double v[29] = {S.t.topology,        S.t.plan,           S.t.factor,          S.t.krylov,
                  S.t.postprocess,     S.t.setup_total,    S.t.solve_total,     S.t.rhs,
                  S.t.ev_interior_ms,  S.t.ev_schur_ms,    S.t.ev_interface_ms, S.t.ev_coarse_ms,
                  S.t.schur_flops,     S.t.interior_flops, S.t.interface_flops, S.t.ev_gemv_ms,
                  S.t.gemv_bytes,      (double)S.last_iters, S.t.ev_kapply_ms, S.t.kapply_bytes,
                  S.t.level2,          S.t.ev_gemm_ms,     S.t.gemm_flops,     (double)S.t.gemm_launches,
                  (double)S.t.refinements, S.t.refine_relres, S.t.refine,     S.t.ev_resid_ms,
                  S.t.resid_bytes};
  for (int k = 0; k < std::min(n, 29); k++) out[k] = v[k];
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_set_option(const char *key, double value) try {
  Solver &S = solver();
  if (!std::strcmp(key, "throw_on_error")) {
    g_throw_on_error = value != 0.0;
    return 0;
  }
  if (!std::strcmp(key, "verbose")) {
    S.verbose = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "sweep_in_global")) {
    set_trsv_global(value != 0.0);
    return 0;
  }
  if (!std::strcmp(key, "adaptive_threshold")) {
    if (!(value > 0.0)) return -2;
    S.adapt_threshold = value;
    return 0;
  }
  if (!std::strcmp(key, "adaptive_tol")) {
    if (!(value > 0.0)) return -2;
    S.adapt_tol = value;
    return 0;
  }
  if (!std::strcmp(key, "adaptive_maxit")) {
    if (value < 1.0) return -2;
    S.adapt_maxit = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "adaptive_nvec")) {
    if (value < 1.0 || value > 16.0) return -2;
    S.adapt_nvec = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "pcg_graph")) {
    if (value != 0.0 && value != 1.0) return -2;
    set_pcg_graph(value != 0.0);
    return 0;
  }
  if (!std::strcmp(key, "launch_forms")) {
    if (value != 0.0 && value != 1.0 && value != 2.0) return -2;
    set_launch_forms((int)value);
    return 0;
  }
  if (!std::strcmp(key, "sweep_waves")) {
    if (value != 0.0 && value != 4.0 && value != 16.0) return -2;
    set_sweep_waves((int)value);
    return 0;
  }
  if (!std::strcmp(key, "pc_form")) {
    if (value != 0.0 && value != 1.0 && value != 2.0) return -2;
    S.pc_form = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "task_order")) {
    bddc::set_task_order(value != 0.0);
    return 0;
  }
  if (!std::strcmp(key, "plan_only")) {
    S.plan_only = value != 0.0;
    return 0;
  }
  if (!std::strcmp(key, "explicit_inverse")) {
    if (value != 0.0 && value != 1.0 && value != 2.0) return -1;
    S.explicit_inverse = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "precision")) {
    if (value != 32.0 && value != 64.0) return -2;
    S.precision = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "refinement")) {
    if (value != 0.0 && value != 1.0) return -2;
    S.refine = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "refinement_first_tol")) {
    if (!(value >= 0.0 && value < 1.0)) return -2;
    S.refine_first = value;
    return 0;
  }
  if (!std::strcmp(key, "refinement_max")) {
    if (value < 0.0) return -2;
    S.refine_max = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "interior_ordering")) {
    if (value != 0.0 && value != 1.0) return -2;
    S.ordering = (int)value;
    return 0;
  }
  if (!std::strcmp(key, "nd_leaf")) {
    if (value < 1.0) return -2;
    S.nd_leaf = (int)value;
    return 0;
  }
  return -1;
}
catch (const Error &e) {
  record_error(e); return -100;
}

static bool device_ready(Solver &S) { return S.is_setup && !S.plan_only && S.dev; }

int bddcml_amd_sub_info(int isub, int *out) try {
  Solver &S = solver();
  if (!S.is_setup) return -1;
  int is = isub - S.numbase;
  if (is < 0 || is >= S.nsub) return -2;
  const Sub &s = S.subs[is];
  out[0] = (int)s.ifc_dof.size();
  out[1] = s.nconstr;
  out[2] = (int)s.int_dof.size();
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_sub_maps(int isub, int *gdofs, int *indrowc) try {
  Solver &S = solver();
  if (!S.is_setup) return -1;
  int is = isub - S.numbase;
  if (is < 0 || is >= S.nsub) return -2;
  const Sub &s = S.subs[is];
  for (size_t k = 0; k < s.ifc_dof.size(); k++) gdofs[k] = s.isvgvn[s.ifc_dof[k]];
  for (int c = 0; c < s.nconstr; c++) indrowc[c] = s.indrowc[c];
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_sub_matrix(int isub, int which, double *out) try {
  Solver &S = solver();
  if (!device_ready(S)) return -1;
  int is = isub - S.numbase;
  if (is < 0 || is >= S.nsub || which != 'S') return -2;
  device_sub_matrix(S, is, which, out);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}
int bddcml_amd_sub_coarse(int isub, double *phi, double *ac) try {
  Solver &S = solver();
  if (!device_ready(S)) return -1;
  int is = isub - S.numbase;
  if (is < 0 || is >= S.nsub) return -2;
  device_sub_coarse(S, is, phi, ac);
  return 0;
}
catch (const Error &e) {
  record_error(e); return -100;
}

int bddcml_amd_last_error(char *buf, int len) {
  int n = (int)g_last_error.size();
  if (buf && len > 0) {
    int m = std::min(n, len - 1);
    std::memcpy(buf, g_last_error.data(), m);
    buf[m] = 0;
  }
  g_last_error.clear();
  return n;
}

int bddcml_amd_selftest_tile(int op, int npairs, const double *A, const double *B, const double *C,
                             const double *X, double alpha, double beta, double *out) try {
  device_prefer_spin_wait();
  device_selftest_tile(op, npairs, A, B, C, X, alpha, beta, out);
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -100;
}

int bddcml_amd_selftest_symeig(int n, const double *a, double *w, double *v) try {
  if (n < 0) fatal("BDDCML_AMD_SELFTEST_SYMEIG", "negative order");
  std::vector<double> A(a, a + (size_t)n * n), W, V;
  sym_eig(n, A, W, V);
  std::copy(W.begin(), W.end(), w);
  std::copy(V.begin(), V.end(), v);
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -100;
}

int bddcml_amd_selftest_ginv(int n, const double *k, double rel_tol, double *g, int *rank) try {
  if (n < 0) fatal("BDDCML_AMD_SELFTEST_GINV", "negative order");
  *rank = psd_ginv(n, k, rel_tol, g);
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -100;
}

int bddcml_amd_element_gradient(const double *ae, int nke, int per_element, const double *u,
                                const double *lam, double *grad) try {
  Solver &S = solver();
  if (!S.is_setup || S.plan_only) fatal("BDDCML_AMD_ELEMENT_GRADIENT", "solver not set up");
  device_element_gradient(S, ae, nke, per_element != 0, u, lam, grad);
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -100;
}

int bddcml_amd_element_gradient_noalloc(const double *ae, int nke, int per_element, const double *u,
                                        const double *lam, double *grad) try {
  Solver &S = solver();
  if (!S.is_setup || S.plan_only) fatal("BDDCML_AMD_ELEMENT_GRADIENT_NOALLOC", "solver not set up");
  device_element_gradient_noalloc(S, ae, nke, per_element != 0, u, lam, grad);
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -100;
}

int bddcml_amd_selftest_copy(long long nbytes, double *out) try {
  device_prefer_spin_wait();
  if (nbytes <= 0 || nbytes % 16 != 0) fatal("BDDCML_SELFTEST", "copy size must be a positive multiple of 16");
  out[0] = device_selftest_copy(nbytes);
  return 0;
}
catch (const Error &e) {
  record_error(e);
  return -100;
}

}  // extern "C"
