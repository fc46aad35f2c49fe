// global.cpp -- global loading: one mesh for the whole problem, divided into
// subdomains here (reference bddcml_upload_global_data /
// bddcml_change_global_data, src/module_bddcml.f90:104-183, 644-679;
// src/module_levels.f90:422-612, 1560-1600, 1835-1870, 3576-3625).
//
// - Element matrices come from the reference's ELM file: an unformatted
//   sequential file of one record per element, in element order, each
//   holding lelm (a 4-byte integer) and lelm doubles: the element matrix by
//   columns, for symmetric storage the upper triangle only (column j, rows
//   1..j), rows and columns running over the element's nodes in INET order
//   with each node's dofs in turn (sm_pmd_load, sm_pmd_make_element_numbering,
//   src/module_sm.f90:170-214, 272-375; read in dd_load_matrix_from_file,
//   src/module_dd.f90:1030-1060).  The reference reads it from the Fortran
//   unit IDELM; this library, which has no Fortran runtime, reads it from
//   the POSIX file descriptor IDELM (a Fortran caller passes FNUM(unit)),
//   rewound first as the reference rewinds the unit.  Records carry the
//   4-byte length markers gfortran writes before and after each record.
// - The division into subdomains: with load_division the file
//   partition_l1.ES in the working directory (NELEM subdomain numbers,
//   1-based, free format), as the reference reads it; otherwise, where the
//   reference calls METIS, recursive coordinate bisection of the element
//   centroids (the mean of the element's node coordinates, summed in INET
//   order), the same rule as for the next levels.
// - Each subdomain as pp_create_submesh builds it (src/module_pp.f90:
//   1443-1503): its elements in ascending global order, its nodes the
//   sorted distinct global nodes of those elements; dofs numbered node by
//   node; IFIX, FIXV, SOL restricted and the right-hand side restricted as a
//   complete one (levels_localize_new_data); components detected on the
//   nodal mesh graph (find_components true, use_dual_mesh_graph false: the
//   reference's level defaults, src/module_levels.f90:142-143).
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>

#include "../../include/bddcml_amd.h"
#include "../../include/bddcml_interface_c.h"
#include "solver.hpp"

namespace bddc {

namespace {

const char *RG = "LEVELS_UPLOAD_GLOBAL_DATA";

// reads exactly n bytes (short reads retried)
bool read_full(int fd, void *p, size_t n) {
  char *c = (char *)p;
  while (n > 0) {
    ssize_t r = ::read(fd, c, n);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    c += r;
    n -= (size_t)r;
  }
  return true;
}

// The element matrices of the ELM file: values of element e at
// vals[off[e] .. off[e+1]).
void read_element_matrices(int fd, int nelem, std::vector<int64_t> &off, std::vector<double> &vals) {
  if (fd < 0) fatal(RG, "Cannot access file with element matrices, unit IDELM.");
  if (::lseek(fd, 0, SEEK_SET) != 0) fatal(RG, "Cannot rewind the file with element matrices (IDELM).");
  off.assign(nelem + 1, 0);
  vals.clear();
  for (int e = 0; e < nelem; e++) {
    int32_t head = 0, tail = 0, lelm = 0;
    if (!read_full(fd, &head, 4) || !read_full(fd, &lelm, 4))
      fatal(RG, "The file with element matrices ends before element " + std::to_string(e + 1));
    if (lelm < 0 || head != 4 + 8 * (int64_t)lelm)
      fatal(RG, "Record of element " + std::to_string(e + 1) +
                    " is not LELM followed by LELM doubles with 4-byte record markers");
    const size_t at = vals.size();
    vals.resize(at + (size_t)lelm);
    if (!read_full(fd, vals.data() + at, 8 * (size_t)lelm) || !read_full(fd, &tail, 4) || tail != head)
      fatal(RG, "Record of element " + std::to_string(e + 1) + " is truncated");
    off[e + 1] = (int64_t)vals.size();
  }
}

// subdomain (0-based) of every element
std::vector<int> divide_elements(const Solver &S, int nelem, const int *inet, const int *nnet,
                                 const std::vector<int64_t> &eptr, const double *xyz, int lxyz1,
                                 int ndim, bool load_division) {
  std::vector<int> iets(nelem, -1);
  if (load_division) {
    std::ifstream f("partition_l1.ES");
    if (!f) fatal(RG, "Cannot open the division file partition_l1.ES");
    for (int e = 0; e < nelem; e++) {
      long v = 0;
      if (!(f >> v)) fatal(RG, "partition_l1.ES holds fewer than NELEM subdomain numbers");
      if (v < 1 || v > S.nsub)
        fatal(RG, "partition_l1.ES: subdomain number out of range for element " + std::to_string(e + 1));
      iets[e] = (int)v - 1;
    }
    return iets;
  }
  std::vector<double> cent((size_t)nelem * ndim, 0.0);
  for (int e = 0; e < nelem; e++)
    for (int d = 0; d < ndim; d++) {
      double sum = 0.0;
      for (int64_t k = eptr[e]; k < eptr[e + 1]; k++) sum += xyz[inet[k] + (size_t)lxyz1 * d];
      cent[(size_t)e * ndim + d] = sum / (double)nnet[e];
    }
  return bisect_points(cent, ndim, nelem, S.nsub);
}

}  // namespace

void upload_global_data(int nelem, int nnod, int ndof, int ndim, int meshdim, const int *inet_in,
                        int linet, const int *nnet, int lnnet, const int *nndf, int lnndf,
                        const double *xyz, int lxyz1, int lxyz2, const int *ifix, int lifix,
                        const double *fixv, int lfixv, const double *rhs, int lrhs,
                        const double *sol, int lsol, int idelm, int neighbouring,
                        bool load_division) {
  Solver &S = solver();
  if (!S.initialised) fatal(RG, "bddcml_init was not called");
  if (S.nsub > nelem)
    fatal(RG, "Number of subdomains at first level must not be larger than number of elements.");
  if (lnnet != nelem || lnndf != nnod || lxyz1 != nnod || lxyz2 != ndim)
    fatal(RG, "array length mismatch");
  if (lifix != ndof || lfixv != ndof || lrhs != ndof || lsol != ndof)
    fatal(RG, "length of IFIX, FIXV, RHS or SOL differs from NDOF");
  const int b = S.numbase;
  std::vector<int64_t> eptr(nelem + 1, 0);
  for (int e = 0; e < nelem; e++) eptr[e + 1] = eptr[e] + nnet[e];
  if (eptr[nelem] != linet) fatal(RG, "sum of NNET differs from LINET");
  std::vector<int> inet(linet);
  for (int k = 0; k < linet; k++) {
    inet[k] = inet_in[k] - b;
    if (inet[k] < 0 || inet[k] >= nnod) fatal(RG, "node index in INET out of range");
  }
  std::vector<int> kdof(nnod + 1, 0);
  for (int n = 0; n < nnod; n++) kdof[n + 1] = kdof[n] + nndf[n];
  if (kdof[nnod] != ndof) fatal(RG, "sum of NNDF differs from NDOF");
  std::vector<int64_t> moff;
  std::vector<double> mval;
  read_element_matrices(idelm, nelem, moff, mval);
  const std::vector<int> iets =
      divide_elements(S, nelem, inet.data(), nnet, eptr, xyz, lxyz1, ndim, load_division);
  std::vector<std::vector<int>> elems(S.nsub);
  for (int e = 0; e < nelem; e++) elems[iets[e]].push_back(e);
  for (int is = 0; is < S.nsub; is++)
    if (elems[is].empty())
      fatal(RG, "Partition does not contain subdomain " + std::to_string(is + b));
  S.global_ifix.assign(ifix, ifix + ndof);
  for (int is = 0; is < S.nsub; is++) {
    const std::vector<int> &el = elems[is];
    std::vector<int> nodes;
    for (int e : el) nodes.insert(nodes.end(), inet.begin() + eptr[e], inet.begin() + eptr[e + 1]);
    std::sort(nodes.begin(), nodes.end());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
    auto local = [&](int g) { return int(std::lower_bound(nodes.begin(), nodes.end(), g) - nodes.begin()); };
    int nelems = (int)el.size(), nnods = (int)nodes.size();
    std::vector<int> linet_s, lnnet_s, lnndf_s(nnods), isngn(nnods), isegn(nelems), lkdof(nnods + 1, 0);
    for (int k = 0; k < nnods; k++) {
      lnndf_s[k] = nndf[nodes[k]];
      isngn[k] = nodes[k] + b;
      lkdof[k + 1] = lkdof[k] + lnndf_s[k];
    }
    int ndofs = lkdof[nnods];
    std::vector<int> isvgvn(ndofs), sifix(ndofs);
    std::vector<double> sxyz((size_t)nnods * ndim), sfixv(ndofs), srhs(ndofs), ssol(ndofs);
    for (int k = 0; k < nnods; k++) {
      for (int d = 0; d < lnndf_s[k]; d++) {
        const int g = kdof[nodes[k]] + d, l = lkdof[k] + d;
        isvgvn[l] = g + b;
        sifix[l] = ifix[g];
        sfixv[l] = fixv[g];
        srhs[l] = rhs[g];
        ssol[l] = sol[g];
      }
      for (int d = 0; d < ndim; d++) sxyz[k + (size_t)nnods * d] = xyz[nodes[k] + (size_t)lxyz1 * d];
    }
    // element matrices: column j of the upper triangle, rows 1..j; an entry
    // below the diagonal of the subdomain numbering is mirrored
    std::vector<int> ti, tj;
    std::vector<double> ta;
    std::vector<int> kdofe;
    for (int q = 0; q < nelems; q++) {
      const int e = el[q];
      isegn[q] = e + b;
      lnnet_s.push_back(nnet[e]);
      kdofe.clear();
      for (int64_t k = eptr[e]; k < eptr[e + 1]; k++) {
        const int ln = local(inet[k]);
        linet_s.push_back(ln + b);
        for (int d = 0; d < lnndf_s[ln]; d++) kdofe.push_back(lkdof[ln] + d);
      }
      const int64_t nve = (int64_t)kdofe.size();
      if (moff[e + 1] - moff[e] != nve * (nve + 1) / 2)
        fatal(RG, "Element " + std::to_string(e + 1) +
                      ": LELM differs from the upper triangle of its element matrix (symmetric storage)");
      int64_t p = moff[e];
      for (int j = 0; j < nve; j++)
        for (int i = 0; i <= j; i++, p++) {
          const int r = std::min(kdofe[i], kdofe[j]), c = std::max(kdofe[i], kdofe[j]);
          ti.push_back(r + b), tj.push_back(c + b), ta.push_back(mval[p]);
        }
    }
    int isub = is + b, ndim_ = ndim, meshdim_ = meshdim, nelem_ = nelem, nnod_ = nnod, ndof_ = ndof;
    int linet_ = (int)linet_s.size(), lnnet_ = nelems, lnndf_ = nnods, lisngn = nnods, lisvgvn = ndofs,
        lisegn = nelems, lx1 = nnods, lx2 = ndim, lfix = ndofs, lfv = ndofs, lrh = ndofs, lso = ndofs;
    int complete = 1, mtype = 1, la = (int)ta.size(), assembled = 0, luc1 = 0, luc2 = 0, led1 = 0,
        led2 = 0, ldd = 0, fc = 1, dual = 0, nb = neighbouring;
    if (ta.size() > (size_t)INT32_MAX) fatal(RG, "too many matrix entries in one subdomain");
    bddcml_upload_subdomain_data_c(&nelem_, &nnod_, &ndof_, &ndim_, &meshdim_, &isub, &nelems, &nnods,
                                   &ndofs, linet_s.data(), &linet_, lnnet_s.data(), &lnnet_,
                                   lnndf_s.data(), &lnndf_, isngn.data(), &lisngn, isvgvn.data(),
                                   &lisvgvn, isegn.data(), &lisegn, sxyz.data(), &lx1, &lx2,
                                   sifix.data(), &lfix, sfixv.data(), &lfv, srhs.data(), &lrh,
                                   &complete, ssol.data(), &lso, &mtype, ti.data(), tj.data(),
                                   ta.data(), &la, &assembled, nullptr, &luc1, &luc2, nullptr, &led1,
                                   &led2, nullptr, &ldd, &fc, &dual, &nb);
    if (!S.subs[is].uploaded) {  // reported instead of aborted (throw_on_error)
      char msg[512] = {0};
      bddcml_amd_last_error(msg, sizeof msg);
      fatal(RG, "subdomain " + std::to_string(isub) + " was not loaded: " + msg);
    }
  }
  S.global_loaded = true;
}

// new Dirichlet values, right-hand side and initial solution for a globally
// loaded problem; the fixed dofs must stay (src/module_levels.f90:555-612)
void change_global_data(const int *ifix, int lifix, const double *fixv, int lfixv, const double *rhs,
                        int lrhs, const double *sol, int lsol) {
  const char *R = "LEVELS_CHANGE_GLOBAL_DATA";
  Solver &S = solver();
  if (!S.global_loaded) fatal(R, "Global loading should be used.");
  if (!S.is_setup) fatal(R, "Cannot change data at this moment, level is not ready.");
  if ((size_t)lifix != S.global_ifix.size()) fatal(R, "Dimension mismatch in IFIX");
  if (lfixv != lifix || lrhs != lifix || lsol != lifix) fatal(R, "Dimension mismatch in FIXV, RHS or SOL");
  for (int g = 0; g < lifix; g++)
    if (S.global_ifix[g] != ifix[g]) fatal(R, "IFIX cannot be changed when loading new data");
  const int b = S.numbase;
  for (int is = 0; is < S.nsub; is++) {
    const Sub &s = S.subs[is];
    int n = s.ndofs, isub = is + b, complete = 1;
    std::vector<int> sifix(n);
    std::vector<double> sfixv(n), srhs(n), ssol(n);
    for (int d = 0; d < n; d++) {
      const int g = s.isvgvn[d];
      sifix[d] = ifix[g], sfixv[d] = fixv[g], srhs[d] = rhs[g], ssol[d] = sol[g];
    }
    // cannot fail: set up, lengths and fixed dofs checked above
    bddcml_change_subdomain_data_c(&isub, sifix.data(), &n, sfixv.data(), &n, srhs.data(), &n,
                                   &complete, ssol.data(), &n);
  }
}

}  // namespace bddc
