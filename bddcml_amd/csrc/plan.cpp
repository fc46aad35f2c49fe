// plan.cpp -- host prerequisites of the device path: neighbours, interface,
// globs and corners, Dirichlet elimination, weights, constraints, coarse
// numbering, condensation right-hand side and the device dof orderings.
//
// Semantics follow the reference (paths relative to its root):
//   interface, shared nodes ..... src/module_dd.f90:8071-8647
//   components .................. src/module_dd.f90:2662-2764
//   globs, corner selection ..... src/module_dd.f90:8925-10554
//   Dirichlet elimination ....... src/module_sm.f90:431-600, src/module_dd.f90:1081-1256
//   weights ..................... src/module_dd.f90:11554-11725
//   constraints ................. src/module_dd.f90:4298-4384, 4869-5145
//   condensation rhs ............ src/module_dd.f90:6995-7136
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <numeric>
#include <queue>
#include <unordered_map>

#include "solver.hpp"

namespace bddc {

bool g_throw_on_error = false;

[[noreturn]] void fatal(const char *routine, const std::string &msg) {
  std::fprintf(stderr, "ERROR in %s: %s\n", routine, msg.c_str());
  std::fflush(stderr);
  if (g_throw_on_error) throw Error(std::string("ERROR in ") + routine + ": " + msg);
  std::abort();
}

// ------------------------------------------------------------ components
static int components(int n, const std::vector<int> &xadj, const std::vector<int> &adj,
                      std::vector<int> &comp) {
  comp.assign(n, 0);
  int nc = 0;
  std::vector<int> stack;
  for (int v = 0; v < n; v++) {
    if (comp[v]) continue;
    comp[v] = ++nc;
    stack.push_back(v);
    while (!stack.empty()) {
      int u = stack.back();
      stack.pop_back();
      for (int k = xadj[u]; k < xadj[u + 1]; k++)
        if (!comp[adj[k]]) comp[adj[k]] = nc, stack.push_back(adj[k]);
    }
  }
  return nc;
}

static void nodal_components(Sub &s) {
  s.ncomp.assign(s.nnods, 1);
  if (!s.find_components) return;
  std::vector<int> kinet(s.nelems + 1, 0);
  for (int e = 0; e < s.nelems; e++) kinet[e + 1] = kinet[e] + s.nnet[e];
  // node -> element incidence (CSR, elements ascending)
  std::vector<int> nptr(s.nnods + 1, 0), nel(kinet[s.nelems]);
  for (int k = 0; k < kinet[s.nelems]; k++) nptr[s.inet[k] + 1]++;
  for (int v = 0; v < s.nnods; v++) nptr[v + 1] += nptr[v];
  {
    std::vector<int> fill(nptr.begin(), nptr.end() - 1);
    for (int e = 0; e < s.nelems; e++)
      for (int k = kinet[e]; k < kinet[e + 1]; k++) nel[fill[s.inet[k]]++] = e;
  }
  auto node_el = [&](int v) { return std::make_pair(nel.data() + nptr[v], nel.data() + nptr[v + 1]); };
  std::vector<int> xadj, adj;
  if (s.use_dual) {
    xadj.assign(s.nelems + 1, 0);
    std::vector<int> cnt(s.nelems, 0), touched;
    for (int e = 0; e < s.nelems; e++) {
      touched.clear();
      for (int k = kinet[e]; k < kinet[e + 1]; k++) {
        auto r = node_el(s.inet[k]);
        for (const int *f = r.first; f != r.second; f++)
          if (*f != e) {
            if (cnt[*f] == 0) touched.push_back(*f);
            cnt[*f]++;
          }
      }
      for (int f : touched) {
        if (cnt[f] >= s.neighbouring) adj.push_back(f);
        cnt[f] = 0;
      }
      xadj[e + 1] = (int)adj.size();
    }
    std::vector<int> ec;
    components(s.nelems, xadj, adj, ec);
    for (int e = 0; e < s.nelems; e++)
      for (int k = kinet[e]; k < kinet[e + 1]; k++) s.ncomp[s.inet[k]] = ec[e];
  } else {
    xadj.assign(s.nnods + 1, 0);
    std::vector<int> mark(s.nnods, -1);
    for (int v = 0; v < s.nnods; v++) {
      auto r = node_el(v);
      for (const int *e = r.first; e != r.second; e++)
        for (int k = kinet[*e]; k < kinet[*e + 1]; k++) {
          int w = s.inet[k];
          if (w != v && mark[w] != v) mark[w] = v, adj.push_back(w);
        }
      xadj[v + 1] = (int)adj.size();
    }
    components(s.nnods, xadj, adj, s.ncomp);
  }
}

// -------------------------------------------------------------- topology
// interface nodes (shared with another subdomain) and the interior /
// interface split of the local dofs
static void split_of(Solver &S, int is) {
  Sub &s = S.subs[is];
  s.iin.clear();
  s.ifc_dof.clear();
  s.int_dof.clear();
  s.kdofi.clear();
  for (int l = 0; l < s.nnods; l++) {
    const int g = s.isngn[l];
    if (S.n2s_ptr[g + 1] - S.n2s_ptr[g] >= 2) {
      s.iin.push_back(l);
      s.kdofi.push_back((int)s.ifc_dof.size());
      for (int d = s.kdof[l]; d < s.kdof[l + 1]; d++) s.ifc_dof.push_back(d);
    } else
      for (int d = s.kdof[l]; d < s.kdof[l + 1]; d++) s.int_dof.push_back(d);
  }
  s.kdofi.push_back((int)s.ifc_dof.size());
}

// neighbours of a subdomain and, per neighbour, the shared interface nodes
// (by global node) with their component inside the neighbour
static void neighbours_of(Solver &S, int is) {
  Sub &s = S.subs[is];
  struct Sh {
    int t, g, l;  // neighbour, global node, interface node
  };
  std::vector<Sh> sh;
  for (int i = 0; i < (int)s.iin.size(); i++) {
    const int g = s.isngn[s.iin[i]];
    for (int q = S.n2s_ptr[g]; q < S.n2s_ptr[g + 1]; q++) {
      int t = S.n2s_idx[q];
      if (t != is) sh.push_back({t, g, i});
    }
  }
  std::sort(sh.begin(), sh.end(), [](const Sh &a, const Sh &b) {
    return a.t != b.t ? a.t < b.t : a.g < b.g;
  });
  s.adj.clear();
  s.shptr.assign(1, 0);
  s.shnode.clear();
  s.shcomp.clear();
  for (size_t k = 0; k < sh.size(); k++) {
    if (s.adj.empty() || s.adj.back() != sh[k].t) {
      if (!s.adj.empty()) s.shptr.push_back((int)s.shnode.size());
      s.adj.push_back(sh[k].t);
    }
    s.shnode.push_back(sh[k].l);
    // component of the node inside the neighbour
    const Sub &u = S.subs[sh[k].t];
    int cu = 1;
    if (u.find_components) {
      auto it = std::lower_bound(u.g2l.begin(), u.g2l.end(), std::make_pair(sh[k].g, -1));
      cu = u.ncomp[it->second];
    }
    s.shcomp.push_back(cu);
  }
  if (!s.adj.empty()) s.shptr.push_back((int)s.shnode.size());
}

// glob classification and face-triangle corner selection
// (src/module_dd.f90:9130-9300 and 10017-10327)
// The nodes of `nodes` (ascending) grouped by equal glob keys: each group
// ascending; the groups in no particular order (their users sort or only
// classify them).
static std::vector<std::vector<int>> key_groups(const Sub &s, std::vector<int> nodes) {
  auto less = [&](int a, int b) {
    const int *pa = s.nkval.data() + s.nkptr[a], *pb = s.nkval.data() + s.nkptr[b];
    const int la = s.nkptr[a + 1] - s.nkptr[a], lb = s.nkptr[b + 1] - s.nkptr[b];
    if (la != lb) return la < lb;
    for (int k = 0; k < la; k++)
      if (pa[k] != pb[k]) return pa[k] < pb[k];
    return a < b;
  };
  auto same = [&](int a, int b) {
    const int la = s.nkptr[a + 1] - s.nkptr[a];
    return la == s.nkptr[b + 1] - s.nkptr[b] &&
           std::equal(s.nkval.begin() + s.nkptr[a], s.nkval.begin() + s.nkptr[a + 1],
                      s.nkval.begin() + s.nkptr[b]);
  };
  std::sort(nodes.begin(), nodes.end(), less);
  std::vector<std::vector<int>> out;
  for (size_t k = 0; k < nodes.size(); k++) {
    if (k == 0 || !same(nodes[k - 1], nodes[k])) out.emplace_back();
    out.back().push_back(nodes[k]);
  }
  return out;
}

static void globs_first_pass(Solver &S, int is) {
  Sub &s = S.subs[is];
  int nnodi = (int)s.iin.size();
  s.nsubnode.assign(nnodi, 0);
  for (size_t a = 0; a < s.adj.size(); a++)
    for (int q = s.shptr[a]; q < s.shptr[a + 1]; q++) s.nsubnode[s.shnode[q]]++;
  s.nkptr.assign(nnodi + 1, 0);
  for (int i = 0; i < nnodi; i++) s.nkptr[i + 1] = s.nkptr[i] + 3 * s.nsubnode[i];
  s.nkval.resize(s.nkptr[nnodi]);
  {
    std::vector<int> f(s.nkptr.begin(), s.nkptr.end() - 1);
    for (size_t a = 0; a < s.adj.size(); a++)
      for (int q = s.shptr[a]; q < s.shptr[a + 1]; q++) {
        const int in = s.shnode[q];
        int *k = s.nkval.data() + f[in];
        f[in] += 3;
        k[0] = s.adj[a];
        k[1] = s.shcomp[q];
        k[2] = s.ncomp[s.iin[in]];
      }
  }
  // group nodes with equal keys
  std::vector<int> all(nnodi);
  for (int i = 0; i < nnodi; i++) all[i] = i;
  std::vector<int> gt(nnodi, 0);
  for (const std::vector<int> &grp : key_groups(s, std::move(all))) {
    const int nsubn = (s.nkptr[grp[0] + 1] - s.nkptr[grp[0]]) / 3;
    const int t = nsubn == 1 ? 1 : (grp.size() > 1 ? 2 : 3);
    for (int i : grp) gt[i] = t;
  }
  for (int i = 0; i < nnodi; i++)
    if (s.nsubnode[i] == 1) gt[i] = 1;
  if (S.use_corner) {
    const int ndim = S.ndim, meshdim = S.meshdim;
    for (size_t a = 0; a < s.adj.size(); a++) {
      int q0 = s.shptr[a], n = s.shptr[a + 1] - q0;
      bool face = false;
      for (int q = 0; q < n && !face; q++) face = gt[s.shnode[q0 + q]] == 1;
      if (!face) continue;
      // combined interface components, renumbered densely in value order
      auto densify = [](std::vector<int> &v) {
        std::vector<int> u(v);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        for (int &x : v) x = int(std::lower_bound(u.begin(), u.end(), x) - u.begin()) + 1;
        return (int)u.size();
      };
      std::vector<int> mine(n), theirs(n), comb(n);
      for (int q = 0; q < n; q++) {
        mine[q] = s.ncomp[s.iin[s.shnode[q0 + q]]];
        theirs[q] = s.shcomp[q0 + q];
      }
      densify(mine);
      int nth = densify(theirs);
      for (int q = 0; q < n; q++) comb[q] = (mine[q] - 1) * nth + theirs[q];
      int ncomb = densify(comb);
      for (int c = 1; c <= ncomb; c++) {
        std::vector<int> idx;
        for (int q = 0; q < n; q++)
          if (comb[q] == c) idx.push_back(q);
        int m = (int)idx.size();
        std::vector<double> X(3 * m, 0.0);
        for (int k = 0; k < m; k++) {
          int ln = s.iin[s.shnode[q0 + idx[k]]];
          for (int d = 0; d < ndim; d++) X[3 * k + d] = s.xyz[(size_t)d * s.nnods + ln];
        }
        auto dist2 = [&](int i, int j) {
          double r = 0;
          for (int d = 0; d < ndim; d++) r += (X[3 * i + d] - X[3 * j + d]) * (X[3 * i + d] - X[3 * j + d]);
          return r;
        };
        int sel[4] = {-1, -1, -1, -1};
        // maxloc keeps the first maximum
        auto argmax_from = [&](int base) {
          int best = 0;
          double bv = dist2(0, base);
          for (int k = 1; k < m; k++) {
            double v = dist2(k, base);
            if (v > bv) bv = v, best = k;
          }
          return best;
        };
        if (m > 0) sel[0] = argmax_from(0);
        if (meshdim > 1 && m > 1) {
          sel[1] = argmax_from(sel[0]);
          if (sel[1] == sel[0]) sel[1] = sel[0] == 0 ? 1 : 0;
        }
        if (meshdim > 2 && m > 2) {
          const double *p1 = &X[3 * sel[0]], *p2 = &X[3 * sel[1]];
          std::vector<double> area(m);
          for (int k = 0; k < m; k++) {
            const double *p = &X[3 * k];
            double c1 = (p2[1] - p1[1]) * (p[2] - p1[2]) - (p[1] - p1[1]) * (p2[2] - p1[2]);
            double c2 = (p2[0] - p1[0]) * (p[2] - p1[2]) - (p[0] - p1[0]) * (p2[2] - p1[2]);
            double c3 = (p2[0] - p1[0]) * (p[1] - p1[1]) - (p[0] - p1[0]) * (p2[1] - p1[1]);
            area[k] = c1 * c1 + c2 * c2 + c3 * c3;
          }
          // strict maximum searched forwards and backwards, starting from 0
          double mv = 0.0;
          for (int k = 0; k < m; k++)
            if (area[k] > mv) mv = area[k], sel[2] = k;
          mv = 0.0;
          for (int k = m - 1; k >= 0; k--)
            if (area[k] > mv) mv = area[k], sel[3] = k;
          if ((sel[2] == sel[0] && sel[3] == sel[0]) || (sel[2] == sel[1] && sel[3] == sel[1])) {
            int k = 0;
            while (k < m && (k == sel[0] || k == sel[1])) k++;
            sel[2] = k < m ? k : -1;
            k = 0;
            while (k < m && (k == sel[0] || k == sel[1] || k == sel[2])) k++;
            sel[3] = k < m ? k : -1;
          }
        }
        int nb = meshdim <= 2 ? 2 : 4;
        for (int k = 0; k < nb; k++)
          if (sel[k] >= 0) gt[s.shnode[q0 + idx[sel[k]]]] = 3;
      }
    }
  }
  s.lcorner.clear();
  for (int i = 0; i < nnodi; i++)
    if (gt[i] == 3) s.lcorner.push_back(s.isngn[s.iin[i]]);
}

// corners merged with neighbours, edges, faces, Dirichlet removal
// (src/module_dd.f90:9322-9660)
static void globs_second_pass(Solver &S, int is) {
  Sub &s = S.subs[is];
  int nnodi = (int)s.iin.size();
  std::vector<int> merged(s.lcorner);
  for (int t : s.adj) merged.insert(merged.end(), S.subs[t].lcorner.begin(), S.subs[t].lcorner.end());
  std::sort(merged.begin(), merged.end());
  merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
  std::vector<int> kind(nnodi, 0);  // 3 corner, -2 removed
  s.corners.clear();
  for (int i = 0; i < nnodi; i++) {
    int g = s.isngn[s.iin[i]];
    if (std::binary_search(merged.begin(), merged.end(), g)) kind[i] = 3, s.corners.push_back(g);
  }
  std::sort(s.corners.begin(), s.corners.end());
  if (s.has_fix)
    for (int i = 0; i < nnodi; i++) {
      if (kind[i] == 3) continue;
      int ln = s.iin[i];
      for (int d = s.kdof[ln]; d < s.kdof[ln + 1]; d++)
        if (s.ifix[d]) kind[i] = -2;
    }
  std::vector<int> free_nodes;
  for (int i = 0; i < nnodi; i++)
    if (kind[i] == 0) free_nodes.push_back(i);
  s.globs.clear();
  for (std::vector<int> &grp : key_groups(s, std::move(free_nodes))) {
    Sub::Glob g;
    g.type = (s.nkptr[grp[0] + 1] - s.nkptr[grp[0]]) / 3 == 1 ? 1 : 2;
    g.nodes = std::move(grp);
    g.key = INT32_MAX;
    for (int i : g.nodes) g.key = std::min(g.key, s.isngn[s.iin[i]]);
    s.globs.push_back(std::move(g));
  }
  std::sort(s.globs.begin(), s.globs.end(), [](const Sub::Glob &a, const Sub::Glob &b) {
    return a.type != b.type ? a.type > b.type : a.key < b.key;
  });
}

// ------------------------------------------------------ structure classes
// Subdomains whose local structure is identical (translated copies on a
// regular decomposition) share per-subdomain analyses.  hash(i) digests the
// integer structure; same(r, i) decides.  Candidates are compared with the
// first subdomain of their hash bucket in parallel; the rare mismatches are
// resolved serially.
// pre(cand) runs once the candidates are known (before any same(cand[i], i)).
template <class H, class E, class P>
static void structure_classes(int n, H hash, E same, std::vector<int> &cls, std::vector<int> &reps,
                              P pre) {
  std::vector<uint64_t> h(n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) h[i] = hash(i);
  std::unordered_map<uint64_t, int> first;
  std::vector<int> cand(n);
  for (int i = 0; i < n; i++) cand[i] = first.emplace(h[i], i).first->second;
  pre(cand);
  std::vector<char> ok(n, 0);
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; i++) ok[i] = cand[i] == i || same(cand[i], i);
  cls.assign(n, -1);
  reps.clear();
  for (int i = 0; i < n; i++) {
    if (ok[i] && cls[cand[i]] >= 0) {
      cls[i] = cls[cand[i]];
      continue;
    }
    if (ok[i] && cand[i] == i) {
      cls[i] = (int)reps.size();
      reps.push_back(i);
      continue;
    }
    for (int r : reps)  // mismatch inside a hash bucket
      if (h[r] == h[i] && same(r, i)) {
        cls[i] = cls[r];
        break;
      }
    if (cls[i] < 0) {
      cls[i] = (int)reps.size();
      reps.push_back(i);
    }
  }
}
template <class H, class E>
static void structure_classes(int n, H hash, E same, std::vector<int> &cls, std::vector<int> &reps) {
  structure_classes(n, hash, same, cls, reps, [](const std::vector<int> &) {});
}

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(uint64_t v) { h = (h ^ v) * 1099511628211ull; }
  template <class T> void add(const std::vector<T> &v) {
    add((uint64_t)v.size());
    for (const T &x : v) add((uint64_t)x);
  }
  // size and every 61st element: enough to separate classes cheaply, since
  // candidates are confirmed by a full comparison
  template <class T> void sample(const std::vector<T> &v) {
    add((uint64_t)v.size());
    for (size_t k = 0; k < v.size(); k += 61) add((uint64_t)v[k]);
    if (!v.empty()) add((uint64_t)v.back());
  }
};

// relative node coordinates equal up to a tolerance of the extent
static bool same_shape(const Sub &a, const Sub &b, int ndim) {
  if (a.xyz.size() != b.xyz.size()) return false;
  const int n = a.nnods;
  for (int d = 0; d < ndim; d++) {
    const double *x = a.xyz.data() + (size_t)d * n, *y = b.xyz.data() + (size_t)d * n;
    double lo = x[0], hi = x[0];
    for (int i = 0; i < n; i++) lo = std::min(lo, x[i]), hi = std::max(hi, x[i]);
    const double tol = 1e-9 * std::max(hi - lo, 1e-300);
    for (int i = 0; i < n; i++)
      if (std::fabs((x[i] - x[0]) - (y[i] - y[0])) > tol) return false;
  }
  return true;
}

void build_topology(Solver &S) {
  const double t_n2s = wall_time();
  // node -> subdomains (CSR, subdomains ascending per node): counts and
  // slots by relaxed atomics, a blocked prefix sum, then a sort of each
  // node's short list
  const int ng = S.nnod_g;
  S.n2s_ptr.resize(ng + 1);
  int *cnt = S.n2s_ptr.data();
#pragma omp parallel for schedule(static)
  for (int g = 0; g <= ng; g++) cnt[g] = 0;
#pragma omp parallel for schedule(dynamic, 4)
  for (int is = 0; is < S.nsub; is++)
    for (int g : S.subs[is].isngn) __atomic_fetch_add(&cnt[g + 1], 1, __ATOMIC_RELAXED);
  {
    const int nb = 64, blk = (ng + nb) / nb;
    std::vector<int> part(nb + 1, 0);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; b++) {
      int acc = 0;
      for (int k = b * blk; k < std::min(ng + 1, (b + 1) * blk); k++) acc += cnt[k], cnt[k] = acc;
      part[b + 1] = acc;
    }
    for (int b = 0; b < nb; b++) part[b + 1] += part[b];
#pragma omp parallel for schedule(static)
    for (int b = 1; b < nb; b++)
      for (int k = b * blk; k < std::min(ng + 1, (b + 1) * blk); k++) cnt[k] += part[b];
  }
  double tp = wall_time();
  if (S.verbose >= 2) std::printf("   topology: %-20s %8.4f s\n", "node counts", tp - t_n2s);
  // the split needs only the counts; the lists themselves are filled by
  // build_node_subdomains, after the first device launch
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) split_of(S, is);
  S.n2s_idx.clear();
  if (S.verbose >= 2) std::printf("   topology: %-20s %8.4f s\n", "interface split", wall_time() - tp);
}

// the subdomains of every global node (CSR over n2s_ptr, ascending): for
// the neighbour lists, the globs and the interface copies
static void build_node_subdomains(Solver &S) {
  const double t0 = wall_time();
  const int ng = S.nnod_g;
  const int *cnt = S.n2s_ptr.data();
  S.n2s_idx.resize(S.n2s_ptr[ng]);  // every slot is written below
  RawInts fill(ng);
  int *f = fill.data(), *idx = S.n2s_idx.data();
#pragma omp parallel for schedule(static)
  for (int g = 0; g < ng; g++) f[g] = cnt[g];
  const double t1 = wall_time();
  // a node of one subdomain has one writer: a plain store, no atomic
#pragma omp parallel for schedule(dynamic, 4)
  for (int is = 0; is < S.nsub; is++)
    for (int g : S.subs[is].isngn) {
      if (cnt[g + 1] - cnt[g] == 1) idx[cnt[g]] = is;
      else idx[__atomic_fetch_add(&f[g], 1, __ATOMIC_RELAXED)] = is;
    }
  const double t2 = wall_time();
#pragma omp parallel for schedule(static, 4096)
  for (int g = 0; g < ng; g++)
    if (cnt[g + 1] - cnt[g] > 1) std::sort(idx + cnt[g], idx + cnt[g + 1]);
  if (S.verbose >= 3) std::printf("   topology: n2s alloc %.4f fill %.4f sort %.4f s\n", t1 - t0, t2 - t1, wall_time() - t2);
  if (S.verbose >= 2) std::printf("   topology: %-20s %8.4f s\n", "node to subdomains", wall_time() - t0);
}

// Components, neighbour lists and globs: needed by the constraints and the
// interface phase only, so they run while the device factors the interiors.
void build_globs(Solver &S) {
  build_node_subdomains(S);
  double tp = wall_time();
  auto phase = [&](const char *name) {
    double now = wall_time();
    if (S.verbose >= 2) std::printf("   globs: %-22s %8.4f s\n", name, now - tp);
    tp = now;
  };
  // nodal components depend only on the element connectivity: once per class
  std::vector<int> ccls, creps;
  structure_classes(
      S.nsub,
      [&](int i) {
        const Sub &s = S.subs[i];
        Fnv f;
        f.add((uint64_t)s.nnods), f.add((uint64_t)s.find_components), f.add((uint64_t)s.use_dual);
        f.add((uint64_t)s.neighbouring), f.sample(s.nnet), f.sample(s.inet);
        return f.h;
      },
      [&](int a, int b) {
        const Sub &x = S.subs[a], &y = S.subs[b];
        return x.nnods == y.nnods && x.find_components == y.find_components &&
               x.use_dual == y.use_dual && x.neighbouring == y.neighbouring && x.nnet == y.nnet &&
               x.inet == y.inet;
      },
      ccls, creps);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < (int)creps.size(); k++) nodal_components(S.subs[creps[k]]);
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    if (creps[ccls[is]] != is) s.ncomp = S.subs[creps[ccls[is]]].ncomp;
    s.g2l.resize(s.nnods);
    for (int l = 0; l < s.nnods; l++) s.g2l[l] = {s.isngn[l], l};
    std::sort(s.g2l.begin(), s.g2l.end());
  }
  phase("components");
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) neighbours_of(S, is);
  phase("neighbours");
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) globs_first_pass(S, is);
  phase("globs pass 1");
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) globs_second_pass(S, is);
  phase("globs pass 2");
}

// ------------------------------------------------------------ BC, weights
// Sum interface values over all subdomains sharing a dof.
// Sums over the copies of each interface dof in subdomain order, through
// the unique-interface CSR (build_unique_interface): parallel over dofs and
// identical to a serial pass over the subdomains.
static void exchange_sum(Solver &S, std::vector<std::vector<double>> &v) {
  if ((int)S.ig_ptr.size() != S.nig + 1) fatal("DD_EXCHANGE", "interface numbering missing");
#pragma omp parallel for schedule(static, 1024)
  for (int u = 0; u < S.nig; u++) {
    double acc = 0.0;
    for (int q = S.ig_ptr[u]; q < S.ig_ptr[u + 1]; q++) acc += v[S.ig_sub[q]][S.ig_k[q]];
    for (int q = S.ig_ptr[u]; q < S.ig_ptr[u + 1]; q++) v[S.ig_sub[q]][S.ig_k[q]] = acc;
  }
}

// bc = eliminated Dirichlet values from the stored couplings: -A(i, fixed) g
// at free dofs, a_ii g_i (diagonal kept, 1 if tiny) at fixed dofs
static void bc_from_stored(Sub &s) {
  s.bc.assign(s.ndofs, 0.0);
  for (size_t q = 0; q < s.bc_k.size(); q++) {
    const size_t k = s.bc_k[q];
    const int i = s.ai[k], j = s.aj[k];
    double a = s.bc_a[q];
    if (i != j) {
      if (s.ifix[i] && !s.ifix[j]) s.bc[j] -= a * s.fixv[i];
      if (s.ifix[j] && !s.ifix[i]) s.bc[i] -= a * s.fixv[j];
    } else {
      if (std::fabs(a) < 1e-12) a = 1.0;
      s.bc[i] += a * s.fixv[i];
    }
  }
}

// Dirichlet elimination keeping the diagonal (src/module_sm.f90:431-600):
// couplings to fixed dofs are zeroed after their values are stored
static void eliminate_bc(Sub &s) {
  s.bc.assign(s.ndofs, 0.0);
  s.bc_k.clear();
  s.bc_a.clear();
  if (!s.has_fix) return;
  // every fixed dof needs a diagonal entry
  std::vector<char> diag_seen(s.ndofs, 0);
  for (size_t k = 0; k < s.av.size(); k++)
    if (s.ai[k] == s.aj[k]) diag_seen[s.ai[k]] = 1;
  bool added = false;
  for (int i = 0; i < s.ndofs; i++)
    if (s.ifix[i] && !diag_seen[i]) {
      s.ai.push_back(i);
      s.aj.push_back(i);
      s.av.push_back(1.0);
      added = true;
    }
  if (added) {
    std::vector<size_t> o(s.av.size());
    std::iota(o.begin(), o.end(), 0);
    std::sort(o.begin(), o.end(), [&](size_t a, size_t b) {
      return s.ai[a] != s.ai[b] ? s.ai[a] < s.ai[b] : s.aj[a] < s.aj[b];
    });
    std::vector<int> ai(o.size()), aj(o.size());
    std::vector<double> av(o.size());
    for (size_t k = 0; k < o.size(); k++) ai[k] = s.ai[o[k]], aj[k] = s.aj[o[k]], av[k] = s.av[o[k]];
    s.ai.swap(ai);
    s.aj.swap(aj);
    s.av.swap(av);
  }
  for (size_t k = 0; k < s.av.size(); k++) {
    const int i = s.ai[k], j = s.aj[k];
    if (!s.ifix[i] && !s.ifix[j]) continue;
    s.bc_k.push_back(k);
    s.bc_a.push_back(s.av[k]);
    if (i != j) s.av[k] = 0.0;
    else if (std::fabs(s.av[k]) < 1e-12) s.av[k] = 1.0;
  }
  bc_from_stored(s);
}

// sum of the eliminated values over the subdomains sharing an interface dof
static void exchange_bc(Solver &S) {
  std::vector<std::vector<double>> b(S.nsub);
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    b[is].resize(s.ifc_dof.size());
    for (size_t k = 0; k < s.ifc_dof.size(); k++) b[is][k] = s.bc[s.ifc_dof[k]];
  }
  exchange_sum(S, b);
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    for (size_t k = 0; k < s.ifc_dof.size(); k++) s.bc[s.ifc_dof[k]] = b[is][k];
  }
}

// Reactions (src/module_dd.f90:1258-1380): rea = A_f u over the kept
// couplings to fixed dofs (upper triangle applied symmetrically), summed
// over the subdomains sharing an interface dof, minus the right-hand side at
// fixed dofs (before the sum when the rhs is not a complete restriction,
// after it when it is).
void build_reactions(Solver &S) {
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    s.rea.assign(s.ndofs, 0.0);
    if (!s.has_fix) continue;
    for (size_t q = 0; q < s.bc_k.size(); q++) {
      const size_t k = s.bc_k[q];
      const int i = s.ai[k], j = s.aj[k];
      const double a = s.bc_a[q];
      s.rea[i] += a * s.sol[j];
      if (i != j) s.rea[j] += a * s.sol[i];
    }
    if (!s.rhs_complete)
      for (int d = 0; d < s.ndofs; d++)
        if (s.ifix[d]) s.rea[d] -= s.rhs[d];
  }
  if (S.bc_global) {
    std::vector<std::vector<double>> b(S.nsub);
    for (int is = 0; is < S.nsub; is++) {
      const Sub &s = S.subs[is];
      b[is].resize(s.ifc_dof.size());
      for (size_t k = 0; k < s.ifc_dof.size(); k++) b[is][k] = s.has_fix ? s.rea[s.ifc_dof[k]] : 0.0;
    }
    exchange_sum(S, b);
    for (int is = 0; is < S.nsub; is++) {
      Sub &s = S.subs[is];
      for (size_t k = 0; k < s.ifc_dof.size(); k++) s.rea[s.ifc_dof[k]] = b[is][k];
      if (s.has_fix && s.rhs_complete)
        for (int d = 0; d < s.ndofs; d++)
          if (s.ifix[d]) s.rea[d] -= s.rhs[d];
    }
  }
}

// new Dirichlet values after bddcml_change_subdomain_data (the reference
// recomputes bc from its stored fixed-rows matrix, module_dd.f90:1081-1256)
void refresh_boundary_conditions(Solver &S) {
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) bc_from_stored(S.subs[is]);
  if (S.bc_global) exchange_bc(S);
}

void apply_boundary_conditions(Solver &S) {
  S.bc_global = false;
  for (auto &s : S.subs) S.bc_global = S.bc_global || s.has_fix;
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) eliminate_bc(S.subs[is]);
  if (S.bc_global) exchange_bc(S);
}

// rho of subdomain is at its interface dofs (reference interface order) for
// the solver's weights type (dd_get_my_coefficients_for_weights,
// src/module_dd.f90:11660-11725); false when type 3 lacks its dof data
bool weight_coefficients(const Solver &S, int is, std::vector<double> &rho) {
  const Sub &s = S.subs[is];
  const size_t ni = s.ifc_dof.size();
  rho.assign(ni, 1.0);
  if (S.weights_type == 1) {
    std::vector<double> dg(s.ndofs, 0.0);
    for (size_t k = 0; k < s.av.size(); k++)
      if (s.ai[k] == s.aj[k]) dg[s.ai[k]] += s.av[k];
    for (size_t k = 0; k < ni; k++) {
      rho[k] = dg[s.ifc_dof[k]];
      if (rho[k] == 0.0) rho[k] = 1.0;
    }
  } else if (S.weights_type == 3) {
    if ((int)s.dof_data.size() != s.ndofs) return false;
    for (size_t k = 0; k < ni; k++) rho[k] = s.dof_data[s.ifc_dof[k]];
  }
  return true;
}

void build_weights(Solver &S) {
  if (S.weights_type != 0 && S.weights_type != 1 && S.weights_type != 3)
    fatal("DD_GET_MY_COEFFICIENTS_FOR_WEIGHTS",
          "Type of weights not supported: " + std::to_string(S.weights_type));
  std::vector<std::vector<double>> rho(S.nsub), tot(S.nsub);
  int bad_data = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad_data)
  for (int is = 0; is < S.nsub; is++) {
    if (!weight_coefficients(S, is, rho[is])) {
      bad_data = 1;
      continue;
    }
    tot[is] = rho[is];
  }
  if (bad_data)
    fatal("DD_GET_MY_COEFFICIENTS_FOR_WEIGHTS", "dof_data must have ndofs entries for weights type 3");
  exchange_sum(S, tot);
  int zero_sub = S.nsub;
#pragma omp parallel for schedule(dynamic) reduction(min : zero_sub)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    s.w.resize(s.ifc_dof.size());
    for (size_t k = 0; k < s.w.size(); k++) {
      if (tot[is][k] == 0.0) zero_sub = std::min(zero_sub, is);
      s.w[k] = rho[is][k] / tot[is][k];
    }
  }
  if (zero_sub < S.nsub)
    fatal("DD_WEIGHTS_PREPARE", "Zero in denominator in computation of weights for subdomain: " +
                                    std::to_string(S.subs[zero_sub].isub + S.numbase));
}

// ------------------------------------------------------------ constraints
void build_constraints(Solver &S) {
  // global coarse nodes: corners by node id, then edges, faces by smallest
  // node id (a permutation of the reference numbering; the coarse space is
  // identical)
  struct CN {
    int order, key, nd;
  };
  std::vector<CN> all;
  for (auto &s : S.subs) {
    for (int g : s.corners) {
      int ln = -1;
      for (int l : s.iin)
        if (s.isngn[l] == g) {
          ln = l;
          break;
        }
      all.push_back({0, g, s.nndf[ln]});
    }
    for (auto &gl : s.globs) {
      int nd = 0;
      for (int i : gl.nodes) nd = std::max(nd, s.nndf[s.iin[i]]);
      all.push_back({gl.type == 2 ? 1 : 2, gl.key, S.use_arith ? nd : 0});
    }
  }
  std::sort(all.begin(), all.end(), [](const CN &a, const CN &b) {
    return a.order != b.order ? a.order < b.order : a.key < b.key;
  });
  std::vector<CN> u;
  for (auto &c : all) {
    if (!u.empty() && u.back().order == c.order && u.back().key == c.key) {
      if (u.back().nd != c.nd) fatal("DD_EMBED_CNODES", "different number of dof at a coarse node from different subdomains");
      continue;
    }
    u.push_back(c);
  }
  S.nnodc = (int)u.size();
  S.ncorner = S.nedge = S.nface = 0;
  S.nndfc.resize(u.size());
  S.kdofc.assign(u.size() + 1, 0);
  for (size_t k = 0; k < u.size(); k++) {
    (u[k].order == 0 ? S.ncorner : (u[k].order == 1 ? S.nedge : S.nface))++;
    S.nndfc[k] = u[k].nd;
    S.kdofc[k + 1] = S.kdofc[k] + u[k].nd;
  }
  S.ndofc = S.kdofc[u.size()];
  auto find = [&](int order, int key) {
    auto it = std::lower_bound(u.begin(), u.end(), CN{order, key, 0}, [](const CN &a, const CN &b) {
      return a.order != b.order ? a.order < b.order : a.key < b.key;
    });
    return int(it - u.begin());
  };
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    s.cptr.assign(1, 0);
    s.cidx.clear();
    s.cval.clear();
    s.indrowc.clear();
    s.cnodes.clear();
    std::vector<int> g2i;  // interface node of a global node (linear search on small sets)
    for (int g : s.corners) {
      int in = -1;
      for (size_t k = 0; k < s.iin.size(); k++)
        if (s.isngn[s.iin[k]] == g) {
          in = (int)k;
          break;
        }
      int id = find(0, g);
      s.cnodes.push_back(id);
      for (int d = 0; d < S.nndfc[id]; d++) {
        s.cidx.push_back(s.kdofi[in] + d);
        s.cval.push_back(1.0);
        s.cptr.push_back((int)s.cidx.size());
        s.indrowc.push_back(S.kdofc[id] + d);
      }
    }
    for (auto &gl : s.globs) {
      int id = find(gl.type == 2 ? 1 : 2, gl.key);
      s.cnodes.push_back(id);
      for (int d = 0; d < S.nndfc[id]; d++) {
        for (int in : gl.nodes)
          if (d < s.kdofi[in + 1] - s.kdofi[in]) {
            s.cidx.push_back(s.kdofi[in] + d);
            s.cval.push_back(1.0);
          }
        s.cptr.push_back((int)s.cidx.size());
        s.indrowc.push_back(S.kdofc[id] + d);
      }
    }
    s.nconstr = (int)s.indrowc.size();
  }
}

// rhs used for condensation: weighted at the interface when it is a
// restriction of a global rhs, Dirichlet values eliminated
// (src/module_dd.f90:7068-7087, src/module_sm.f90:602-628)
void build_condensation_rhs(Solver &S, int isub_only) {
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    if (isub_only >= 0 && is != isub_only) continue;
    Sub &s = S.subs[is];
    s.fmod = s.rhs;
    if (s.rhs_complete)
      for (size_t k = 0; k < s.ifc_dof.size(); k++) s.fmod[s.ifc_dof[k]] *= s.w[k];
    if (S.bc_global) {
      std::vector<double> bc(s.bc);
      for (size_t k = 0; k < s.ifc_dof.size(); k++) bc[s.ifc_dof[k]] *= s.w[k];
      if (s.has_fix)
        for (int d = 0; d < s.ndofs; d++)
          if (s.ifix[d]) s.fmod[d] = 0.0;
      for (int d = 0; d < s.ndofs; d++) s.fmod[d] += bc[d];
    }
  }
}

// Unique interface: sorted global dofs of every subdomain's interface, the
// position of each global dof in it, fixed flags, and the CSR of the copies
// of each unique dof (subdomain, interface index), subdomains ascending.
void build_unique_interface(Solver &S) {
  const int ng = S.ndof_g;
  std::vector<char> on(ng, 0), fixed(ng, 0);
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    const Sub &s = S.subs[is];
    for (int d : s.ifc_dof) {
      __atomic_store_n(&on[s.isvgvn[d]], (char)1, __ATOMIC_RELAXED);
      if (s.has_fix && s.ifix[d]) __atomic_store_n(&fixed[s.isvgvn[d]], (char)1, __ATOMIC_RELAXED);
    }
  }
  S.gpos.assign(ng, -1);
  {
    const int nb = 64, blk = (ng + nb - 1) / nb;
    std::vector<int> part(nb + 1, 0);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; b++) {
      int c = 0;
      for (int g = b * blk; g < std::min(ng, (b + 1) * blk); g++) c += on[g];
      part[b + 1] = c;
    }
    for (int b = 0; b < nb; b++) part[b + 1] += part[b];
    S.nig = part[nb];
    S.igdof.resize(S.nig);
    S.ig_fixed.resize(S.nig);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; b++) {
      int u = part[b];
      for (int g = b * blk; g < std::min(ng, (b + 1) * blk); g++)
        if (on[g]) S.gpos[g] = u, S.igdof[u] = g, S.ig_fixed[u] = fixed[g], u++;
    }
  }
  // copies: counts, then each subdomain fills its slots; a copy's rank among
  // the copies of its dof is the number of lower-numbered subdomains that
  // hold the dof
  std::vector<int> cnt(S.nig + 1, 0);
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++)
    for (int d : S.subs[is].ifc_dof) __atomic_fetch_add(&cnt[S.gpos[S.subs[is].isvgvn[d]] + 1], 1, __ATOMIC_RELAXED);
  for (int u = 0; u < S.nig; u++) cnt[u + 1] += cnt[u];
  S.ig_ptr = cnt;
  S.ig_sub.assign(cnt[S.nig], -1);
  S.ig_k.assign(cnt[S.nig], -1);
  // rank through the node-to-subdomain lists (every holder of a node holds
  // all of its dofs); checked below, with a serial fill as the fallback
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int is = 0; is < S.nsub; is++) {
    const Sub &s = S.subs[is];
    for (size_t i = 0; i < s.iin.size(); i++) {
      const int g = s.isngn[s.iin[i]];
      int rank = -1;
      for (int q = S.n2s_ptr[g]; q < S.n2s_ptr[g + 1]; q++)
        if (S.n2s_idx[q] == is) rank = q - S.n2s_ptr[g];
      for (int k = s.kdofi[i]; k < s.kdofi[i + 1]; k++) {
        const int u = S.gpos[s.isvgvn[s.ifc_dof[k]]];
        const int q = S.ig_ptr[u] + rank;
        if (rank < 0 || q >= S.ig_ptr[u + 1]) {
          bad = 1;
          continue;
        }
        S.ig_sub[q] = is;
        S.ig_k[q] = k;
      }
    }
  }
  for (size_t q = 0; q < S.ig_sub.size() && !bad; q++)
    if (S.ig_sub[q] < 0) bad = 1;
  if (bad) {
    std::vector<int> fill(S.ig_ptr.begin(), S.ig_ptr.end() - 1);
    for (int is = 0; is < S.nsub; is++) {
      const Sub &s = S.subs[is];
      for (size_t k = 0; k < s.ifc_dof.size(); k++) {
        const int q = fill[S.gpos[s.isvgvn[s.ifc_dof[k]]]]++;
        S.ig_sub[q] = is;
        S.ig_k[q] = (int)k;
      }
    }
  }
}

// ------------------------------------------------------------- orderings
// Interior dofs: reverse Cuthill-McKee on the interior graph from a
// pseudo-peripheral start (George & Liu 1979).  Interface dofs: sorted by
// the first interior position they couple to, so that the rows of the
// interface block of the factor start as late as possible.
// Reverse Cuthill-McKee order of the interior dofs (pseudo-peripheral start,
// neighbours by ascending degree).
static std::vector<int> order_rcm(const Sub &s, const std::vector<int> &loc, int no) {
  std::vector<std::vector<int>> g(no);
  for (size_t k = 0; k < s.av.size(); k++) {
    int i = loc[s.ai[k]], j = loc[s.aj[k]];
    if (i >= 0 && j >= 0 && i != j) g[i].push_back(j), g[j].push_back(i);
  }
  std::vector<int> order;
  order.reserve(no);
  std::vector<char> seen(no, 0);
  auto bfs_last = [&](int start, std::vector<int> &level) {
    level.assign(no, -1);
    std::queue<int> q;
    q.push(start);
    level[start] = 0;
    int last = start;
    while (!q.empty()) {
      int v = q.front();
      q.pop();
      last = v;
      for (int w : g[v])
        if (level[w] < 0) level[w] = level[v] + 1, q.push(w);
    }
    return last;
  };
  std::vector<int> level;
  for (int root = 0; root < no; root++) {
    if (seen[root]) continue;
    int start = root;
    for (int it = 0; it < 2; it++) {
      int far = bfs_last(start, level);
      // among the last level pick minimum degree
      int lv = level[far], best = far;
      for (int v = 0; v < no; v++)
        if (level[v] == lv && g[v].size() < g[best].size()) best = v;
      start = best;
    }
    std::queue<int> q;
    q.push(start);
    seen[start] = 1;
    while (!q.empty()) {
      int v = q.front();
      q.pop();
      order.push_back(v);
      std::vector<int> nb;
      for (int w : g[v])
        if (!seen[w]) seen[w] = 1, nb.push_back(w);
      std::sort(nb.begin(), nb.end(), [&](int a, int b) { return g[a].size() < g[b].size(); });
      for (int w : nb) q.push(w);
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Nested dissection of the interior nodes by recursive coordinate bisection.
// A part is cut at the median coordinate of its longest extent; the nodes of
// the upper half adjacent to the lower half form the separator.  Parts with
// at most leaf_dofs dofs stay whole and are swept along their longest axis.
// Blocks come out in elimination order: lower part, upper part, separator.
namespace {
struct Dissection {
  const std::vector<int> &ptr, &adj;  // interior node graph (CSR)
  const std::vector<double> &xyz;     // coordinates, column major nn x ndim
  const std::vector<int> &w;          // dofs per interior node
  int nn, ndim, leaf_dofs;
  std::vector<int> side;              // scratch: 1 = lower half
  std::vector<std::vector<int>> blocks;
  double coord(int v, int d) const { return xyz[(size_t)d * nn + v]; }
  int longest_axis(const std::vector<int> &part) const {
    int ax = 0;
    double best = -1.0;
    for (int d = 0; d < ndim; d++) {
      double lo = coord(part[0], d), hi = lo;
      for (int v : part) lo = std::min(lo, coord(v, d)), hi = std::max(hi, coord(v, d));
      if (hi - lo > best) best = hi - lo, ax = d;
    }
    return ax;
  }
  void leaf(std::vector<int> part) {
    const int ax = longest_axis(part);
    std::stable_sort(part.begin(), part.end(), [&](int a, int b) {
      for (int k = 0; k < ndim; k++) {
        int d = (ax + k) % ndim;
        if (coord(a, d) != coord(b, d)) return coord(a, d) < coord(b, d);
      }
      return a < b;
    });
    blocks.push_back(std::move(part));
  }
  void split(std::vector<int> part) {
    if (part.empty()) return;
    long dofs = 0;
    for (int v : part) dofs += w[v];
    if (dofs <= leaf_dofs || ndim == 0) return leaf(std::move(part));
    const int ax = longest_axis(part);
    std::vector<double> c(part.size());
    for (size_t k = 0; k < part.size(); k++) c[k] = coord(part[k], ax);
    std::vector<double> sc(c);
    std::nth_element(sc.begin(), sc.begin() + sc.size() / 2, sc.end());
    const double m = sc[sc.size() / 2];
    std::vector<int> lower, upper, sep;
    for (size_t k = 0; k < part.size(); k++) (c[k] < m ? lower : upper).push_back(part[k]);
    if (lower.empty()) return leaf(std::move(part));
    for (int v : lower) side[v] = 1;
    std::vector<int> rest;
    for (int v : upper) {
      bool touches = false;
      for (int q = ptr[v]; q < ptr[v + 1] && !touches; q++) touches = side[adj[q]] == 1;
      (touches ? sep : rest).push_back(v);
    }
    for (int v : lower) side[v] = 0;
    split(std::move(lower));
    split(std::move(rest));
    if (!sep.empty()) blocks.push_back(std::move(sep));
  }
};
}  // namespace

static void order_subdomain(Sub &s, int method, int leaf_dofs, int ndim) {
  const int n = s.ndofs;
  std::vector<int> loc(n, -1);  // interior index
  for (size_t k = 0; k < s.int_dof.size(); k++) loc[s.int_dof[k]] = (int)k;
  const int no = (int)s.int_dof.size(), ni = (int)s.ifc_dof.size();
  // device position of every interior dof; blocks start on tile boundaries
  std::vector<int> ipos(no, -1);
  int NO = 0;
  s.pad.clear();
  if (method == 1 || ndim <= 0 || s.xyz.size() < (size_t)s.nnods * ndim) {
    std::vector<int> order = order_rcm(s, loc, no);
    for (int k = 0; k < no; k++) ipos[order[k]] = k;
    NO = (no + TILE - 1) / TILE * TILE;
    for (int p = no; p < NO; p++) s.pad.push_back(p);
  } else {
    // interior nodes (all dofs of a node share its interior/interface role)
    std::vector<int> dnode(n);
    for (int l = 0; l < s.nnods; l++)
      for (int d = s.kdof[l]; d < s.kdof[l + 1]; d++) dnode[d] = l;
    std::vector<int> nloc(s.nnods, -1), nodes;
    for (int d : s.int_dof)
      if (nloc[dnode[d]] < 0) nloc[dnode[d]] = (int)nodes.size(), nodes.push_back(dnode[d]);
    const int nn = (int)nodes.size();
    // symmetric node adjacency by counting sort (repeated neighbours of
    // multi-dof nodes are harmless to the separator test)
    std::vector<int> ptr(nn + 1, 0);
    for (size_t k = 0; k < s.av.size(); k++) {
      int a = nloc[dnode[s.ai[k]]], b = nloc[dnode[s.aj[k]]];
      if (a >= 0 && b >= 0 && a != b) ptr[a + 1]++, ptr[b + 1]++;
    }
    for (int v = 0; v < nn; v++) ptr[v + 1] += ptr[v];
    std::vector<int> adj(ptr[nn]);
    {
      std::vector<int> fill(ptr.begin(), ptr.end() - 1);
      for (size_t k = 0; k < s.av.size(); k++) {
        int a = nloc[dnode[s.ai[k]]], b = nloc[dnode[s.aj[k]]];
        if (a >= 0 && b >= 0 && a != b) adj[fill[a]++] = b, adj[fill[b]++] = a;
      }
    }
    std::vector<double> xyz((size_t)nn * ndim);
    std::vector<int> w(nn);
    for (int v = 0; v < nn; v++) {
      for (int d = 0; d < ndim; d++) xyz[(size_t)d * nn + v] = s.xyz[(size_t)d * s.nnods + nodes[v]];
      w[v] = s.kdof[nodes[v] + 1] - s.kdof[nodes[v]];
    }
    Dissection nd{ptr, adj, xyz, w, nn, ndim, std::max(leaf_dofs, 1), std::vector<int>(nn, 0), {}};
    std::vector<int> all(nn);
    std::iota(all.begin(), all.end(), 0);
    nd.split(std::move(all));
    int p = 0;
    for (const auto &blk : nd.blocks) {
      for (int v : blk)
        for (int d = s.kdof[nodes[v]]; d < s.kdof[nodes[v] + 1]; d++) ipos[loc[d]] = p++;
      int q = (p + TILE - 1) / TILE * TILE;
      for (; p < q; p++) s.pad.push_back(p);
    }
    NO = p;
  }
  // interface: by first interior coupling position
  std::vector<int> iloc(n, -1);
  for (int k = 0; k < ni; k++) iloc[s.ifc_dof[k]] = k;
  std::vector<int> first(ni, NO);
  for (size_t k = 0; k < s.av.size(); k++) {
    int i = s.ai[k], j = s.aj[k];
    if (loc[i] >= 0 && iloc[j] >= 0) first[iloc[j]] = std::min(first[iloc[j]], ipos[loc[i]]);
    if (loc[j] >= 0 && iloc[i] >= 0) first[iloc[i]] = std::min(first[iloc[i]], ipos[loc[j]]);
  }
  std::vector<int> io(ni);
  std::iota(io.begin(), io.end(), 0);
  std::stable_sort(io.begin(), io.end(), [&](int a, int b) { return first[a] < first[b]; });
  s.no = no;
  s.ni = ni;
  s.NO = NO;
  s.NI = (ni + TILE - 1) / TILE * TILE;
  if (s.NI == 0) s.NI = TILE;
  s.nto = s.NO / TILE;
  s.nti = s.NI / TILE;
  s.perm.assign(n, -1);
  for (int k = 0; k < no; k++) s.perm[s.int_dof[k]] = ipos[k];
  s.dev_ifc.resize(ni);
  for (int k = 0; k < ni; k++) {
    s.dev_ifc[k] = io[k];
    s.perm[s.ifc_dof[io[k]]] = s.NO + k;
  }
}

// Orderings (and the symbolic factorization built on them) depend on the
// matrix pattern, the interior/interface split and the node coordinates up
// to a translation: computed once per class of identical subdomains.
void build_orderings(Solver &S) {
  double tp = wall_time();
  auto phase = [&](const char *name) {
    double now = wall_time();
    if (S.verbose >= 3) std::printf("   orderings: %-19s %8.4f s\n", name, now - tp);
    tp = now;
  };
  // the triplet patterns, the bulk of the comparison, are compared in HBM
  // where the uploads put them (each subdomain against its candidate)
  std::vector<int> cand_of;
  std::vector<char> pat_eq;
  structure_classes(
      S.nsub,
      [&](int i) {
        const Sub &s = S.subs[i];
        Fnv f;
        f.add((uint64_t)s.ndofs), f.add((uint64_t)s.nnods), f.sample(s.kdof);
        f.sample(s.ai), f.sample(s.aj), f.add(s.int_dof), f.add(s.ifc_dof);
        return f.h;
      },
      [&](int a, int b) {
        const Sub &x = S.subs[a], &y = S.subs[b];
        const bool known = !pat_eq.empty() && cand_of[b] == a;
        return x.ndofs == y.ndofs && x.nnods == y.nnods && x.kdof == y.kdof &&
               x.int_dof == y.int_dof && x.ifc_dof == y.ifc_dof &&
               (known ? pat_eq[b] != 0 : (x.ai == y.ai && x.aj == y.aj)) && same_shape(x, y, S.ndim);
      },
      S.ord_cls, S.ord_reps,
      [&](const std::vector<int> &cand) {
        cand_of = cand;
        if (!device_same_patterns(S, cand, pat_eq)) pat_eq.clear();
      });
  phase("classes");
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < (int)S.ord_reps.size(); k++)
    order_subdomain(S.subs[S.ord_reps[k]], S.ordering, S.nd_leaf, S.ndim);
  phase("representatives");
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < S.nsub; is++) {
    const Sub &r = S.subs[S.ord_reps[S.ord_cls[is]]];
    if (&r == &S.subs[is]) continue;
    Sub &s = S.subs[is];
    s.no = r.no, s.ni = r.ni, s.NO = r.NO, s.NI = r.NI, s.nto = r.nto, s.nti = r.nti;
    s.perm = r.perm;
    s.dev_ifc = r.dev_ifc;
    s.pad = r.pad;
  }
  phase("copies");
}

}  // namespace bddc
