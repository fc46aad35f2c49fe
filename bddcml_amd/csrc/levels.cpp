// levels.cpp -- multilevel BDDC: the next level's mesh and its setup.
//
// The reference builds level l+1 from level l (src/module_levels.f90:
// 1484-1530): its elements are the subdomains of level l, its nodes the
// coarse nodes of level l (corners, edges, faces) with their coarse dofs and
// coordinates (src/module_dd.f90:4941-4992), its element matrices the
// subdomain coarse matrices Ac_i (dd_gather_matrix, src/module_dd.f90:
// 1382-1842).  Level l+1 is divided into subdomains (the reference calls
// METIS, pp_divide_mesh; this library bisects the element centroids
// recursively), each subdomain's nodes sorted by global id as
// pp_create_submesh does (src/module_pp.f90:1443-1503), and set up with
// the same pipeline as level 1: no Dirichlet data, right-hand sides
// complete, neighbours sharing any node (neighbouring 1).
//
// During the solve one application of the second level's BDDC replaces
// the coarse solve of level 1 (levels_corsub_standard_level,
// levels_add_standard_level, src/module_levels.f90:3849-4249): interior
// solves on the restricted coarse residual, the preconditioner on the
// reduced interface residual, interior recovery.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "solver.hpp"

namespace bddc {

// Coordinates of every coarse node: a corner's node, the mean of a glob's
// nodes summed in ascending global node order (the order fixes the rounding,
// and the bisection below and the corner search of the next level compare
// coordinates).
void build_coarse_nodes(Solver &S) {
  const int nc = S.nnodc, nd = S.ndim;
  S.xyzc.assign((size_t)nc * nd, 0.0);
  std::vector<char> done(nc, 0);
  for (const Sub &s : S.subs) {
    const size_t ncorner = s.corners.size();
    if (s.cnodes.size() != ncorner + s.globs.size())
      fatal("LEVELS_PREPARE_STANDARD_LEVEL", "coarse nodes of a subdomain missing");
    for (size_t c = 0; c < s.cnodes.size(); c++) {
      const int id = s.cnodes[c];
      if (done[id]) continue;
      done[id] = 1;
      if (c < ncorner) {
        auto it = std::lower_bound(s.g2l.begin(), s.g2l.end(), std::make_pair(s.corners[c], -1));
        if (it == s.g2l.end() || it->first != s.corners[c])
          fatal("LEVELS_PREPARE_STANDARD_LEVEL", "corner node not in its subdomain");
        for (int d = 0; d < nd; d++) S.xyzc[id + (size_t)nc * d] = s.xyz[it->second + (size_t)s.nnods * d];
      } else {
        const Sub::Glob &gl = s.globs[c - ncorner];
        std::vector<std::pair<int, int>> v;  // (global node, local node)
        for (int i : gl.nodes) v.push_back({s.isngn[s.iin[i]], s.iin[i]});
        std::sort(v.begin(), v.end());
        for (int d = 0; d < nd; d++) {
          double sum = 0.0;
          for (auto &p : v) sum += s.xyz[p.second + (size_t)s.nnods * d];
          S.xyzc[id + (size_t)nc * d] = sum / (double)v.size();
        }
      }
    }
  }
  for (int k = 0; k < nc; k++)
    if (!done[k]) fatal("LEVELS_PREPARE_STANDARD_LEVEL", "coarse node without a subdomain");
}

// Recursive coordinate bisection of n points into nparts parts: the axis of
// largest extent (the first on ties), points ordered by (coordinate, index),
// the first n * floor(p/2) / p of them to the lower half.
static void bisect(const std::vector<double> &cent, int nd, std::vector<int> ids, int p, int base,
                   std::vector<int> &part) {
  if (p == 1) {
    for (int i : ids) part[i] = base;
    return;
  }
  int ax = 0;
  double best = -1.0;
  for (int d = 0; d < nd; d++) {
    double lo = INFINITY, hi = -INFINITY;
    for (int i : ids) {
      lo = std::min(lo, cent[(size_t)i * nd + d]);
      hi = std::max(hi, cent[(size_t)i * nd + d]);
    }
    if (hi - lo > best) best = hi - lo, ax = d;
  }
  std::sort(ids.begin(), ids.end(), [&](int a, int b) {
    const double x = cent[(size_t)a * nd + ax], y = cent[(size_t)b * nd + ax];
    return x != y ? x < y : a < b;
  });
  const int pl = p / 2;
  const size_t nl = ids.size() * (size_t)pl / (size_t)p;
  bisect(cent, nd, std::vector<int>(ids.begin(), ids.begin() + nl), pl, base, part);
  bisect(cent, nd, std::vector<int>(ids.begin() + nl, ids.end()), p - pl, base + pl, part);
}

std::vector<int> bisect_points(const std::vector<double> &cent, int nd, int n, int nparts) {
  std::vector<int> part(n, -1), ids(n);
  std::iota(ids.begin(), ids.end(), 0);
  bisect(cent, nd, ids, nparts, 0, part);
  return part;
}

// Division of level 1's subdomains among nsub2 second-level subdomains by
// their centroids (mean of their coarse nodes, ascending coarse node order).
static std::vector<int> divide_subdomains(const Solver &S, int nsub2) {
  const int nd = S.ndim;
  std::vector<double> cent((size_t)S.nsub * nd, 0.0);
  for (int is = 0; is < S.nsub; is++) {
    std::vector<int> cn(S.subs[is].cnodes);
    if (cn.empty()) fatal("LEVELS_PREPARE_STANDARD_LEVEL", "a subdomain without coarse nodes");
    std::sort(cn.begin(), cn.end());
    for (int d = 0; d < nd; d++) {
      double sum = 0.0;
      for (int c : cn) sum += S.xyzc[c + (size_t)S.nnodc * d];
      cent[(size_t)is * nd + d] = sum / (double)cn.size();
    }
  }
  return bisect_points(cent, nd, S.nsub, nsub2);
}

// The second level's subdomain data (0-based, as uploaded), with the
// assembled upper triangle of its matrix from the coarse matrices; false if
// a coarse dof of an element is missing from the subdomain (inconsistent
// coarse numbering).
static bool level2_subdomain(const Solver &S, const std::vector<int> &elems,
                             const std::vector<std::vector<double>> &acs, int j, Sub &t) {
  t = Sub();
  t.isub = j;
  std::vector<int> nodes;
  for (int e : elems) nodes.insert(nodes.end(), S.subs[e].cnodes.begin(), S.subs[e].cnodes.end());
  std::sort(nodes.begin(), nodes.end());
  nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
  auto local = [&](int g) { return int(std::lower_bound(nodes.begin(), nodes.end(), g) - nodes.begin()); };
  t.nelems = (int)elems.size();
  t.nnods = (int)nodes.size();
  for (int e : elems) {
    for (int c : S.subs[e].cnodes) t.inet.push_back(local(c));
    t.nnet.push_back((int)S.subs[e].cnodes.size());
    t.isegn.push_back(e);
  }
  t.isngn = nodes;
  t.kdof.assign(t.nnods + 1, 0);
  for (int k = 0; k < t.nnods; k++) {
    t.nndf.push_back(S.nndfc[nodes[k]]);
    t.kdof[k + 1] = t.kdof[k] + S.nndfc[nodes[k]];
    for (int d = 0; d < S.nndfc[nodes[k]]; d++) t.isvgvn.push_back(S.kdofc[nodes[k]] + d);
  }
  t.ndofs = t.kdof[t.nnods];
  const int nd = S.ndim;
  t.xyz.resize((size_t)t.nnods * nd);
  for (int d = 0; d < nd; d++)
    for (int k = 0; k < t.nnods; k++) t.xyz[k + (size_t)t.nnods * d] = S.xyzc[nodes[k] + (size_t)S.nnodc * d];
  t.ifix.assign(t.ndofs, 0);
  t.fixv.assign(t.ndofs, 0.0);
  t.rhs.assign(t.ndofs, 0.0);
  t.sol.assign(t.ndofs, 0.0);
  t.rhs_complete = true;
  t.find_components = S.subs[0].find_components;
  t.use_dual = S.subs[0].use_dual;
  t.neighbouring = 1;
  // element matrices through the global coarse dofs of their rows (isvgvn
  // ascends with the sorted nodes); upper triangle, duplicates summed
  struct Tr {
    int i, j;
    double a;
  };
  std::vector<Tr> tr;
  for (int e : elems) {
    const Sub &s = S.subs[e];
    const int nc = s.nconstr;
    std::vector<int> loc(nc);
    for (int a = 0; a < nc; a++) {
      auto it = std::lower_bound(t.isvgvn.begin(), t.isvgvn.end(), s.indrowc[a]);
      if (it == t.isvgvn.end() || *it != s.indrowc[a]) return false;
      loc[a] = int(it - t.isvgvn.begin());
    }
    const std::vector<double> &A = acs[e];
    for (int a = 0; a < nc; a++)
      for (int b = 0; b < nc; b++)
        if (loc[a] <= loc[b]) tr.push_back({loc[a], loc[b], A[(size_t)a * nc + b]});
  }
  std::stable_sort(tr.begin(), tr.end(), [](const Tr &x, const Tr &y) { return x.i != y.i ? x.i < y.i : x.j < y.j; });
  for (const Tr &x : tr) {
    if (!t.ai.empty() && t.ai.back() == x.i && t.aj.back() == x.j)
      t.av.back() += x.a;
    else
      t.ai.push_back(x.i), t.aj.push_back(x.j), t.av.push_back(x.a);
  }
  t.uploaded = true;
  return true;
}

void setup_second_level(Solver &S) {
  const char *R = "LEVELS_PREPARE_STANDARD_LEVEL";
  double t0 = wall_time();
  if (S.nsub2 < 2) fatal(R, "the second level needs at least two subdomains");
  if (S.nsub2 > S.nsub) fatal(R, "more subdomains on the second level than on the first");
  build_coarse_nodes(S);
  std::vector<std::vector<double>> acs;
  if (S.plan_only) {
    // host plan only (no device): the second level's topology, division,
    // globs and corners do not depend on the values; identity blocks
    acs.resize(S.nsub);
    for (int is = 0; is < S.nsub; is++) {
      const int nc = S.subs[is].nconstr;
      acs[is].assign((size_t)nc * nc, 0.0);
      for (int a = 0; a < nc; a++) acs[is][(size_t)a * nc + a] = 1.0;
    }
  } else {
    device_coarse_matrices(S, acs);
  }
  const std::vector<int> part = divide_subdomains(S, S.nsub2);
  std::vector<std::vector<int>> elems(S.nsub2);
  for (int is = 0; is < S.nsub; is++) elems[part[is]].push_back(is);
  for (int j = 0; j < S.nsub2; j++)
    if (elems[j].empty()) fatal(R, "Partition does not contain subdomain " + std::to_string(j));
  S.lev2.reset(new Solver());
  Solver &T = *S.lev2;
  T.nl = S.nl - 1;
  T.nsub = S.nsub2;
  T.nsublev.assign(S.nsublev.begin() + 1, S.nsublev.end());
  T.nsub2 = T.nl >= 3 ? T.nsublev[1] : 0;
  T.numbase = 0;
  T.verbose = S.verbose >= 3 ? S.verbose : 0;
  T.initialised = true;
  T.nelem_g = S.nsub;
  T.nnod_g = S.nnodc;
  T.ndof_g = S.ndofc;
  T.ndim = S.ndim;
  T.meshdim = S.meshdim;
  T.matrixtype = 1;
  T.use_corner = S.use_corner;
  T.use_arith = S.use_arith;
  T.weights_type = S.weights_type;
  T.precision = 64;  // the coarse levels stay in FP64
  T.ordering = S.ordering;
  T.nd_leaf = S.nd_leaf;
  // levels below the first form S'^{-1} and the inverse of the interior
  // factor at setup unless the caller turned the explicit inverse off: their
  // subdomains are few and small, so the inverses cost little, and their
  // tile GEMVs spread each apply across the chip where the factor sweeps
  // would run one workgroup per subdomain
  T.explicit_inverse = S.explicit_inverse == 0 ? 0 : 1;
  T.explicit_interior = S.explicit_inverse != 0;
  T.plan_only = S.plan_only;
  T.subs.resize(T.nsub);
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int j = 0; j < T.nsub; j++)
    if (!level2_subdomain(S, elems[j], acs, j, T.subs[j])) bad |= 1;
  if (bad) fatal(R, "coarse dof outside the second-level subdomain");
  if (!T.plan_only)
    for (int j = 0; j < T.nsub; j++) device_upload_entries(T, j);
  // the pipeline of bddcml_setup_preconditioner on the second level
  build_topology(T);
  build_orderings(T);
  if (!T.plan_only) device_setup_begin(T);
  build_globs(T);
  build_constraints(T);
  if (!T.plan_only) device_setup_interface(T);
  build_unique_interface(T);
  apply_boundary_conditions(T);
  build_weights(T);
  build_condensation_rhs(T);
  if (!T.plan_only) device_setup_finish(T);
  // more levels: the next one is built from this one's coarse nodes (on its
  // stream, which the link below redirects to this level's)
  if (T.nl >= 3 && T.nsub2 >= 2) setup_second_level(T);
  if (!T.plan_only) device_link_levels(S);
  T.is_setup = true;
  S.t.level2 = wall_time() - t0;
  if (S.verbose > 0)
    std::printf(" LEVELS_PREPARE_STANDARD_LEVEL: level %d of %d: %d subdomains, %d dofs, ndofc %d, "
                "interface %d, %.4f s\n",
                (int)(S.nsublev.size() - T.nsublev.size()) + 1, (int)S.nsublev.size(), T.nsub, T.ndof_g,
                T.ndofc, T.nig, S.t.level2);
}

}  // namespace bddc
