// constraints.cpp -- constraint rows beyond corners and arithmetic
// averages: the orthonormalization of a coarse node's block of rows, the
// replacement of whole coarse nodes' rows in every subdomain with the coarse
// renumbering, and the user constraints on edges.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "solver.hpp"

namespace bddc {

// Householder QR with column pivoting of M^T (rows of M as columns), the
// columns kept while |R_kk| >= 0.01 |R_11|, and Q's first columns as
// orthonormal rows spanning those of M (LAPACK's DGEQP3 + DORGQR, sign
// conventions included; src/module_dd.f90:4637-4802).  M: r x n row major.
std::vector<std::vector<double>> orthonormal_rows(const std::vector<std::vector<double>> &M, int n) {
  const int r = (int)M.size();
  std::vector<std::vector<double>> a(r, std::vector<double>(n));  // columns of M^T
  for (int c = 0; c < r; c++) a[c] = M[c];
  std::vector<std::vector<double>> vs;  // Householder vectors
  std::vector<double> taus, rdiag;
  std::vector<double> norms(r);
  for (int c = 0; c < r; c++) {
    double s = 0.0;
    for (double x : a[c]) s += x * x;
    norms[c] = s;
  }
  const int kmax = std::min(r, n);
  for (int k = 0; k < kmax; k++) {
    // pivot: the remaining column of largest norm (first on ties)
    int piv = k;
    double best = -1.0;
    for (int c = k; c < r; c++) {
      double s = 0.0;
      for (int i = k; i < n; i++) s += a[c][i] * a[c][i];
      norms[c] = s;
      if (s > best) best = s, piv = c;
    }
    std::swap(a[k], a[piv]);
    // reflector for a[k][k..n)
    double alpha = a[k][k], xn = 0.0;
    for (int i = k + 1; i < n; i++) xn += a[k][i] * a[k][i];
    std::vector<double> v(n, 0.0);
    double tau = 0.0, beta = alpha;
    if (xn > 0.0) {
      beta = -std::copysign(std::sqrt(alpha * alpha + xn), alpha);
      tau = (beta - alpha) / beta;
      v[k] = 1.0;
      for (int i = k + 1; i < n; i++) v[i] = a[k][i] / (alpha - beta);
    }
    rdiag.push_back(beta);
    // apply H = I - tau v v^T to the remaining columns
    for (int c = k + 1; c < r; c++) {
      double s = 0.0;
      for (int i = k; i < n; i++) s += v[i] * a[c][i];
      for (int i = k; i < n; i++) a[c][i] -= tau * v[i] * s;
    }
    vs.push_back(v);
    taus.push_back(tau);
  }
  int nvalid = 0;
  if (!rdiag.empty() && std::fabs(rdiag[0]) > 1e-12) {
    for (size_t k = 0; k < rdiag.size(); k++) {
      if (std::fabs(rdiag[k]) < 0.01 * std::fabs(rdiag[0])) break;
      nvalid = (int)k + 1;
    }
  }
  // Q e_c = H_0 H_1 ... H_{nvalid-1} e_c
  std::vector<std::vector<double>> q(nvalid, std::vector<double>(n, 0.0));
  for (int c = 0; c < nvalid; c++) {
    std::vector<double> &x = q[c];
    x[c] = 1.0;
    for (int k = nvalid - 1; k >= 0; k--) {
      double s = 0.0;
      for (int i = k; i < n; i++) s += vs[k][i] * x[i];
      for (int i = k; i < n; i++) x[i] -= taus[k] * vs[k][i] * s;
    }
  }
  return q;
}

// The rows of the coarse nodes in `rows` (each row over that node's global
// dofs `gdofs`, sorted) replace the node's rows in every subdomain holding
// it; the coarse dofs are renumbered with the new counts per node (other
// nodes keep their rows and count).
void replace_coarse_node_rows(Solver &S, const std::map<int, std::vector<std::vector<double>>> &rows,
                              const std::map<int, std::vector<int>> &gdofs) {
  auto node_of = [&](int cdof) {
    return int(std::upper_bound(S.kdofc.begin(), S.kdofc.end(), cdof) - S.kdofc.begin()) - 1;
  };
  std::vector<int> nndfc(S.nndfc);
  for (auto &f : rows) nndfc[f.first] = (int)f.second.size();
  std::vector<int> kdofc(nndfc.size() + 1, 0);
  for (size_t k = 0; k < nndfc.size(); k++) kdofc[k + 1] = kdofc[k] + nndfc[k];
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    std::map<int, int> g2r;
    for (int r = 0; r < s.ni; r++) g2r[s.isvgvn[s.ifc_dof[r]]] = r;
    std::vector<int> cptr{0}, cidx, indrowc;
    std::vector<double> cval;
    for (int id : s.cnodes) {
      auto rq = rows.find(id);
      if (rq != rows.end()) {
        const auto &Q = rq->second;
        const auto &gd = gdofs.at(id);
        for (size_t r = 0; r < Q.size(); r++) {
          for (size_t k = 0; k < gd.size(); k++) {
            if (Q[r][k] == 0.0) continue;
            auto it = g2r.find(gd[k]);
            if (it == g2r.end()) {
              bad = 1;
              continue;
            }
            cidx.push_back(it->second), cval.push_back(Q[r][k]);
          }
          cptr.push_back((int)cidx.size());
          indrowc.push_back(kdofc[id] + (int)r);
        }
      } else {
        for (int c = 0; c < s.nconstr; c++) {
          if (node_of(s.indrowc[c]) != id) continue;
          for (int q = s.cptr[c]; q < s.cptr[c + 1]; q++) cidx.push_back(s.cidx[q]), cval.push_back(s.cval[q]);
          cptr.push_back((int)cidx.size());
          indrowc.push_back(kdofc[id] + (s.indrowc[c] - S.kdofc[id]));
        }
      }
    }
    s.cptr.swap(cptr);
    s.cidx.swap(cidx);
    s.cval.swap(cval);
    s.indrowc.swap(indrowc);
    s.nconstr = (int)s.indrowc.size();
  }
  if (bad) fatal("DD_LOAD_CONSTRAINTS", "a coarse node's dof is not on the interface of a subdomain holding it");
  S.nndfc = nndfc;
  S.kdofc = kdofc;
  S.ndofc = kdofc.back();
}

// User constraints on edges, first level (src/module_levels.f90:2709-2716):
// after its arithmetic rows, every edge gets one row per user constraint and
// nodal component, with the user's value of each edge node at that
// component's dof (src/module_dd.f90:4459-4565; a node's value in
// user_constraints(iconstr, node), the subdomain's local node numbering);
// the block is orthonormalized by QR with column pivoting keeping |R_kk| >=
// 0.01 |R_11| (src/module_dd.f90:4637-4802).  The block is built once per
// edge, over its dofs in global order, from the lowest-numbered subdomain
// holding it, and the same rows go to every subdomain of the edge (the
// subdomains' user values must agree on shared nodes, as in the reference);
// the coarse dofs are renumbered.
void add_user_constraints(Solver &S) {
  const char *R = "DD_LOAD_USER_CONSTRAINTS";
  std::map<int, std::pair<int, int>> owner;  // edge coarse node -> (subdomain, glob)
  for (int is = 0; is < S.nsub; is++) {
    const Sub &s = S.subs[is];
    const int nc = (int)s.corners.size();
    for (int gi = 0; gi < (int)s.globs.size(); gi++)
      if (s.globs[gi].type == 2) owner.emplace(s.cnodes[nc + gi], std::make_pair(is, gi));
  }
  std::map<int, std::vector<std::vector<double>>> rows;
  std::map<int, std::vector<int>> gdofs;
  for (auto &e : owner) {
    const int id = e.first;
    const Sub &s = S.subs[e.second.first];
    const Sub::Glob &gl = s.globs[e.second.second];
    // the edge's interface dofs: (global dof, interface index, node, component)
    struct Dof {
      int g, k, in, d;
    };
    std::vector<Dof> dofs;
    int ncdof = 0;
    for (int in : gl.nodes) {
      ncdof = std::max(ncdof, s.kdofi[in + 1] - s.kdofi[in]);
      for (int k = s.kdofi[in]; k < s.kdofi[in + 1]; k++)
        dofs.push_back({s.isvgvn[s.ifc_dof[k]], k, in, k - s.kdofi[in]});
    }
    std::sort(dofs.begin(), dofs.end(), [](const Dof &a, const Dof &b) { return a.g < b.g; });
    const int n = (int)dofs.size();
    std::map<int, int> col_of_k;
    for (int c = 0; c < n; c++) col_of_k[dofs[c].k] = c;
    std::vector<std::vector<double>> M;
    for (int c = 0; c < s.nconstr; c++) {
      if (s.indrowc[c] < S.kdofc[id] || s.indrowc[c] >= S.kdofc[id + 1]) continue;
      std::vector<double> v(n, 0.0);
      for (int q = s.cptr[c]; q < s.cptr[c + 1]; q++) {
        auto it = col_of_k.find(s.cidx[q]);
        if (it == col_of_k.end()) fatal(R, "arithmetic row outside its edge");
        v[it->second] = s.cval[q];
      }
      M.push_back(std::move(v));
    }
    std::vector<std::vector<double>> U;
    bool nonzero = false;
    for (int ic = 0; ic < s.nuser_c; ic++)
      for (int d = 0; d < ncdof; d++) {
        std::vector<double> v(n, 0.0);
        for (int c = 0; c < n; c++)
          if (dofs[c].d == d) {
            v[c] = s.user_c[(size_t)ic + (size_t)s.nuser_c * s.iin[dofs[c].in]];
            nonzero |= v[c] != 0.0;
          }
        U.push_back(std::move(v));
      }
    if (nonzero) M.insert(M.end(), U.begin(), U.end());
    else if (S.verbose > 0 && s.nuser_c > 0)
      std::printf(" %s: user constraints are plain zeros on coarse node %d\n", R, id);
    std::vector<int> gd(n);
    for (int c = 0; c < n; c++) gd[c] = dofs[c].g;
    rows[id] = orthonormal_rows(M, n);
    gdofs[id] = gd;
  }
  replace_coarse_node_rows(S, rows, gdofs);
}

}  // namespace bddc
