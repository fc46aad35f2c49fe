// adapt.hip -- adaptive BDDC constraints: the face eigenproblems of every pair
// of subdomains sharing a face, solved together by a batched LOBPCG on the
// device, and their eigenvectors loaded as constraints on the face.
//
// Reference: src/module_adaptivity.f90:296-3024 (pairs, operators, LOBPCG
// call and constraint creation), src/lobpcg_driver.c:169-304 (BLOPEX LOBPCG
// with an absolute tolerance), src/module_dd.f90:4386-4457 (loading on the
// face glob) and 4637-4802 (orthonormalizing each face block),
// src/module_dd.f90:10785-11009 (pairs = faces).
//
// For a pair (i, j) the vectors live on both interfaces, side i then side j,
// each in its device interface order and padded to its tile multiple (the
// padding stays zero).  With S = diag(S_i, S_j) (the packed lower tiles of
// the explicit Schur complements),
//   A = P (I - E)^T S (I - E) P,   B = P S P,
// P = I - sum_c q_c q_c^T over the shared constraint rows [C_i; -C_j]
// normalized (their supports are disjoint for corner and arithmetic rows,
// so this is the QR projection of the reference), (I - E) x = the weighted
// jump on the dofs common to i and j (w_j (x_a - x_b) on the i copy a,
// w_i (x_b - x_a) on the j copy b, 0 elsewhere), (I - E^T) its transpose.
// The reference runs LOBPCG for the smallest eigenvalues of (-A, B); here
// the largest of (A, B) are sought directly (the same eigenpairs).
//
// Each LOBPCG step runs batched over all pairs on the device: the block
// operator applies (projection, jump, the products with S for the A and B
// inputs together), the residuals and their norms, the Gram matrices of
// [X W P] with A and B, and the block updates.  The products with S_s (and
// with S'_s^{-1} in the preconditioner) are gathered per subdomain: one
// tile task list multiplies each subdomain's matrix by the columns of all
// pairs it belongs to on the f64 matrix cores, so each S_s is read once per
// 64 columns rather than once per pair.  The Rayleigh-Ritz
// step of each pair (a generalized eigenproblem of order at most 3k = 30)
// runs on the host between them, by two symmetric eigendecompositions: the
// directions where the B-Gram matrix is numerically singular are dropped, so
// converged columns and the floating null space of B need no special case.
// The residuals are preconditioned, as in the reference
// (lobpcg_preconditioner = 1, src/module_adaptivity.f90:2760-2905), by a
// BDDC of the pair itself: T = P (H + Phi K^+ Phi^T) P with, per side,
// H_s = S'_s^{-1} - Z_s G_s^{-1} Z_s^T (the solve with the subdomain's
// constraints held at zero: S'^{-1} applied from its explicit tiles, formed
// by the first setup) and the pair coarse problem K = sum_s R_s^T (Phi_s^T
// S_s Phi_s) R_s over the union of the two subdomains' coarse dofs,
// pseudo-inverted on the host (K is singular when the pair floats).  On the
// space where the shared constraints agree this is the inverse of S, so
// LOBPCG converges in a few iterations where a diagonal preconditioner
// stalls on coefficient jumps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "solver.hpp"

namespace bddc {

#define AD_CHECK(x)                                                                       \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      fatal("ADAPTIVITY_SOLVE_EIGENVECTORS", std::string(#x) + " failed: " +             \
                                                 hipGetErrorString(e_));                  \
  } while (0)

struct PairDesc {
  int NIi, n;        // padded interface size of i; n = NIi + NIj
  int k;             // columns of the block (eigenvectors sought)
  int active;        // still iterating
  double tol;        // residual norm below which a column has converged
  int64_t off, woff; // block offsets: n x kmax and n x 2 kmax per pair
  int q0, q1;        // projection columns (qptr range)
  int m0, m1;        // common dofs (range in the common-dof arrays)
  int nci, ncj, ncp; // constraints of i, j; coarse dofs of the pair
  int cm0;           // pair coarse dof of each constraint (cmap range, i then j)
  int64_t zoi, zoj;  // Z / Phi offsets of i, j (NI x nc column major)
  int64_t kc0;       // K^+ of the pair (ncp x ncp)
};

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ double *pcol(double *b, const PairDesc &p, bool wide, int c) {
  return b + (wide ? p.woff : p.off) + (int64_t)c * p.n;
}

// x <- x - sum_c q_c (q_c . x) on columns [c0, c0 + ncol) of every active
// pair; one thread per projection column (disjoint supports)
__global__ void k_ad_project(const PairDesc *__restrict__ pairs, const int *__restrict__ qptr,
                             const int *__restrict__ qpos, const double *__restrict__ qval,
                             double *buf, int wide, int c0) {
  const PairDesc p = pairs[blockIdx.x];
  const int col = blockIdx.y;
  if (!p.active || col >= p.k) return;
  double *x = pcol(buf, p, wide, c0 + col);
  for (int c = p.q0 + threadIdx.x; c < p.q1; c += blockDim.x) {
    double t = 0.0;
    for (int e = qptr[c]; e < qptr[c + 1]; e++) t += qval[e] * x[qpos[e]];
    for (int e = qptr[c]; e < qptr[c + 1]; e++) x[qpos[e]] -= t * qval[e];
  }
}

// columns [d0, d0 + k) of dst = 0, then on the common dofs the jump of
// columns [s0, s0 + k) of src: (I - E) (transpose 0) or (I - E^T) (1)
__global__ void k_ad_jump(const PairDesc *__restrict__ pairs, const int *__restrict__ ma,
                          const int *__restrict__ mb, const double *__restrict__ mwa,
                          const double *__restrict__ mwb, double *src, int src_wide, int s0,
                          double *dst, int dst_wide, int d0, int transpose) {
  const PairDesc p = pairs[blockIdx.x];
  const int col = blockIdx.y;
  if (!p.active || col >= p.k) return;
  const double *x = pcol(src, p, src_wide, s0 + col);
  double *y = pcol(dst, p, dst_wide, d0 + col);
  for (int i = threadIdx.x; i < p.n; i += blockDim.x) y[i] = 0.0;
  __syncthreads();
  for (int e = p.m0 + threadIdx.x; e < p.m1; e += blockDim.x) {
    const int a = ma[e], b = mb[e];
    const double wa = mwa[e], wb = mwb[e], xa = x[a], xb = x[b];
    if (transpose) {
      y[a] = wb * xa - wa * xb;
      y[b] = wa * xb - wb * xa;
    } else {
      y[a] = wb * (xa - xb);
      y[b] = wa * (xb - xa);
    }
  }
}

// columns [d0, d0 + k) of dst = columns [s0, s0 + k) of src
__global__ void k_ad_copy(const PairDesc *__restrict__ pairs, double *src, int src_wide, int s0,
                          double *dst, int dst_wide, int d0) {
  const PairDesc p = pairs[blockIdx.x];
  const int col = blockIdx.y;
  if (!p.active || col >= p.k) return;
  const double *x = pcol(src, p, src_wide, s0 + col);
  double *y = pcol(dst, p, dst_wide, d0 + col);
  for (int i = threadIdx.x; i < p.n; i += blockDim.x) y[i] = x[i];
}

// R = AX - BX diag(lam) into W and its squared column norms; a column
// whose residual is below the pair's tolerance contributes no direction
// (W = 0)
__global__ __launch_bounds__(256) void k_ad_resid(const PairDesc *__restrict__ pairs,
                                                  const double *__restrict__ lam, int kmax,
                                                  double *AX, double *BX, double *W,
                                                  double *__restrict__ nrm2) {
  __shared__ double red[256];
  const PairDesc p = pairs[blockIdx.x];
  const int col = blockIdx.y;
  if (!p.active || col >= p.k) return;
  const double l = lam[blockIdx.x * kmax + col];
  const double *ax = pcol(AX, p, false, col), *bx = pcol(BX, p, false, col);
  double *w = pcol(W, p, false, col);
  double s = 0.0;
  for (int i = threadIdx.x; i < p.n; i += 256) {
    const double v = ax[i] - l * bx[i];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double nr2 = red[0];
  const bool keep = nr2 > p.tol * p.tol;
  for (int i = threadIdx.x; i < p.n; i += 256) w[i] = keep ? ax[i] - l * bx[i] : 0.0;
  if (threadIdx.x == 0) nrm2[blockIdx.x * kmax + col] = nr2;
}

// Pair coarse correction of the preconditioner.  With r the first k wide
// columns of in and y = S'^{-1} r already in out, per side s = i, j:
//   y_s += Phi_s (R_s u - Z_s^T r_s),   u = K^+ sum_s R_s^T Phi_s^T r_s,
// since Z_s G_s^{-1} Z_s^T r_s = Phi_s Z_s^T r_s.  R_s maps the constraints
// of s to the pair's coarse dofs (cmap); K^+ is symmetric.  LDS: the coarse
// vectors g, u (ncp each) and Phi^T r, Z^T r (nci + ncj each).
__global__ __launch_bounds__(256) void k_ad_coarse_pc(const PairDesc *__restrict__ pairs,
                                                      const double *__restrict__ Z,
                                                      const double *__restrict__ Phi,
                                                      const int *__restrict__ cmap,
                                                      const double *__restrict__ Kp,
                                                      const double *__restrict__ in,
                                                      double *__restrict__ out) {
  extern __shared__ double sh[];
  const PairDesc p = pairs[blockIdx.x];
  const int col = blockIdx.y;
  if (!p.active || col >= p.k) return;
  const int nc2 = p.nci + p.ncj, NIj = p.n - p.NIi;
  double *g = sh, *u = g + p.ncp, *gp = u + p.ncp, *zr = gp + nc2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double *x = in + p.woff + (int64_t)col * p.n;
  // Phi^T r and Z^T r, one wave per constraint
  for (int e = wv; e < nc2; e += 4) {
    const int side = e >= p.nci, c = side ? e - p.nci : e;
    const int NI = side ? NIj : p.NIi;
    const int64_t zo = (side ? p.zoj : p.zoi) + (int64_t)c * NI;
    const double *xs = x + (side ? p.NIi : 0);
    double a = 0.0, b = 0.0;
    for (int r = lane; r < NI; r += 64) {
      a += Phi[zo + r] * xs[r];
      b += Z[zo + r] * xs[r];
    }
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_down(a, o);
      b += __shfl_down(b, o);
    }
    if (lane == 0) gp[e] = a, zr[e] = b;
  }
  for (int q = threadIdx.x; q < p.ncp; q += 256) g[q] = 0.0;
  __syncthreads();
  // assemble side by side: within a side the coarse dofs are distinct
  const int *cm = cmap + p.cm0;
  for (int c = threadIdx.x; c < p.nci; c += 256) g[cm[c]] += gp[c];
  __syncthreads();
  for (int c = threadIdx.x; c < p.ncj; c += 256) g[cm[p.nci + c]] += gp[p.nci + c];
  __syncthreads();
  const double *K = Kp + p.kc0;
  for (int q = threadIdx.x; q < p.ncp; q += 256) {
    double s = 0.0;
    for (int t = 0; t < p.ncp; t++) s += K[(int64_t)t * p.ncp + q] * g[t];
    u[q] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nc2; e += 256) gp[e] = u[cm[e]] - zr[e];
  __syncthreads();
  double *y = out + p.woff + (int64_t)col * p.n;
  for (int i = threadIdx.x; i < p.n; i += 256) {
    const int side = i >= p.NIi, r = side ? i - p.NIi : i;
    const int NI = side ? NIj : p.NIi, nc = side ? p.ncj : p.nci;
    const double *ph = Phi + (side ? p.zoj : p.zoi) + r;
    const double *w = gp + (side ? p.nci : 0);
    double s = 0.0;
    for (int c = 0; c < nc; c++) s += ph[(int64_t)c * NI] * w[c];
    y[i] += s;
  }
}

struct Blocks {
  double *v[3], *av[3], *bv[3];  // [X W P], A times them, B times them
};

// Gram matrices GA = V^T (A V), GB = V^T (B V) of the nb blocks of V, each
// of k columns (row major, order nb k, stride 3 kmax); one workgroup per
// pair, one wave per entry at a time
__global__ __launch_bounds__(256) void k_ad_gram(const PairDesc *__restrict__ pairs, Blocks b,
                                                 int nb, int kmax, double *__restrict__ GA,
                                                 double *__restrict__ GB) {
  const PairDesc p = pairs[blockIdx.x];
  if (!p.active) return;
  const int m = nb * p.k, ld = 3 * kmax;
  double *ga = GA + (int64_t)blockIdx.x * ld * ld, *gb = GB + (int64_t)blockIdx.x * ld * ld;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int e = wave; e < m * m; e += 4) {
    const int a = e / m, c = e % m;
    if (c < a) continue;
    const double *va = pcol(b.v[a / p.k], p, false, a % p.k);
    const double *ac = pcol(b.av[c / p.k], p, false, c % p.k);
    const double *bc = pcol(b.bv[c / p.k], p, false, c % p.k);
    double sa = 0.0, sb = 0.0;
    for (int i = lane; i < p.n; i += 64) {
      sa += va[i] * ac[i];
      sb += va[i] * bc[i];
    }
    for (int o = 32; o > 0; o >>= 1) {
      sa += __shfl_xor(sa, o);
      sb += __shfl_xor(sb, o);
    }
    if (lane == 0) {
      ga[a * ld + c] = ga[c * ld + a] = sa;
      gb[a * ld + c] = gb[c * ld + a] = sb;
    }
  }
}

// X' = V C, P' = V C without the X block (the W and P parts), and the same
// combinations of A V and B V; C is (nb k) x k per pair, row major with
// stride kmax; one thread per row
__global__ __launch_bounds__(256) void k_ad_combine(const PairDesc *__restrict__ pairs, Blocks b,
                                                    int nb, int kmax, const double *__restrict__ C,
                                                    Blocks o) {
  const PairDesc p = pairs[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n) return;
  if (!p.active) {  // converged: the blocks carry over unchanged
    for (int which = 0; which < 3; which++) {
      double *const *src = which == 0 ? b.v : (which == 1 ? b.av : b.bv);
      double *const *dst = which == 0 ? o.v : (which == 1 ? o.av : o.bv);
      for (int c = 0; c < p.k; c++) {
        pcol(dst[0], p, false, c)[i] = pcol(src[0], p, false, c)[i];
        pcol(dst[2], p, false, c)[i] = pcol(src[2], p, false, c)[i];
      }
    }
    return;
  }
  const double *cf = C + (int64_t)blockIdx.y * 3 * kmax * kmax;
  for (int which = 0; which < 3; which++) {
    double *const *src = which == 0 ? b.v : (which == 1 ? b.av : b.bv);
    double *const *dst = which == 0 ? o.v : (which == 1 ? o.av : o.bv);
    for (int c = 0; c < p.k; c++) {
      double sx = 0.0, sp = 0.0;
      for (int blk = 0; blk < nb; blk++)
        for (int j = 0; j < p.k; j++) {
          const double v = pcol(src[blk], p, false, j)[i] * cf[(blk * p.k + j) * kmax + c];
          sx += v;
          if (blk > 0) sp += v;
        }
      pcol(dst[0], p, false, c)[i] = sx;  // new X (or AX, BX)
      pcol(dst[2], p, false, c)[i] = sp;  // new P (or AP, BP)
    }
  }
}

// ------------------------------------------------------- host eigen-solves
// Symmetric eigendecomposition of an n x n matrix (row major, overwritten)
// (the Rayleigh-Ritz problems of LOBPCG): Householder reduction to
// tridiagonal form T = Q^T A Q, then implicit symmetric QR steps with
// Wilkinson shifts on T, the rotations accumulated into Q (Golub and Van
// Loan, Matrix Computations, algorithms 8.3.1 and 8.3.3).  Eigenvalues in
// w (unordered), eigenvectors in the columns of v.
void sym_eig(int n, std::vector<double> &a, std::vector<double> &w, std::vector<double> &v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
  w.assign(n, 0.0);
  if (n == 0) return;
  auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
  auto Q = [&](int i, int j) -> double & { return v[(size_t)i * n + j]; };
  std::vector<double> hv(n), pv(n), e(n, 0.0);
  // reduction: step k annihilates column k below the subdiagonal
  for (int k = 0; k + 2 < n; k++) {
    const int m = n - k - 1;  // rows k+1 .. n-1
    double nrm = 0.0;
    for (int i = 0; i < m; i++) nrm += A(k + 1 + i, k) * A(k + 1 + i, k);
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
    const double x0 = A(k + 1, k);
    const double alpha = x0 >= 0.0 ? -nrm : nrm;
    for (int i = 0; i < m; i++) hv[i] = A(k + 1 + i, k);
    hv[0] -= alpha;
    double hn = 0.0;
    for (int i = 0; i < m; i++) hn += hv[i] * hv[i];
    hn = std::sqrt(hn);
    if (hn == 0.0) continue;
    for (int i = 0; i < m; i++) hv[i] /= hn;
    // H = I - 2 h h^T on the trailing block: A22 -= 2 (h q^T + q h^T),
    // p = A22 h, q = p - (h^T p) h
    double K = 0.0;
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      for (int j = 0; j < m; j++) s += A(k + 1 + i, k + 1 + j) * hv[j];
      pv[i] = s;
      K += hv[i] * s;
    }
    for (int i = 0; i < m; i++) pv[i] -= K * hv[i];
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) A(k + 1 + i, k + 1 + j) -= 2.0 * (hv[i] * pv[j] + pv[i] * hv[j]);
    A(k + 1, k) = A(k, k + 1) = alpha;
    for (int i = 1; i < m; i++) A(k + 1 + i, k) = A(k, k + 1 + i) = 0.0;
    // Q <- Q H
    for (int r = 0; r < n; r++) {
      double s = 0.0;
      for (int j = 0; j < m; j++) s += Q(r, k + 1 + j) * hv[j];
      for (int j = 0; j < m; j++) Q(r, k + 1 + j) -= 2.0 * s * hv[j];
    }
  }
  for (int i = 0; i < n; i++) w[i] = A(i, i);
  for (int i = 0; i + 1 < n; i++) e[i] = A(i + 1, i);
  // implicit QR on the unreduced blocks, from the bottom
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, guard = 0;
  while (hi > 0 && guard < 60 * n) {
    if (std::fabs(e[hi - 1]) <= eps * (std::fabs(w[hi - 1]) + std::fabs(w[hi]))) {
      e[hi - 1] = 0.0;
      hi--;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(w[lo - 1]) + std::fabs(w[lo]))) lo--;
    guard++;
    const double dl = 0.5 * (w[hi - 1] - w[hi]), eh = e[hi - 1];
    const double mu = w[hi] - eh * eh / (dl + (dl >= 0.0 ? 1.0 : -1.0) * std::hypot(dl, eh));
    double x = w[lo] - mu, z = e[lo];
    for (int k = lo; k < hi; k++) {
      const double r = std::hypot(x, z);
      const double c = r > 0.0 ? x / r : 1.0, s = r > 0.0 ? z / r : 0.0;
      if (k > lo) e[k - 1] = r;
      const double ak = w[k], bk = e[k], dk = w[k + 1];
      w[k] = c * c * ak + 2.0 * c * s * bk + s * s * dk;
      w[k + 1] = s * s * ak - 2.0 * c * s * bk + c * c * dk;
      e[k] = c * s * (dk - ak) + (c * c - s * s) * bk;
      if (k + 1 < hi) {
        z = s * e[k + 1];  // the bulge at (k + 2, k)
        e[k + 1] *= c;
        x = e[k];
      }
      for (int r2 = 0; r2 < n; r2++) {
        const double qk = Q(r2, k), qk1 = Q(r2, k + 1);
        Q(r2, k) = c * qk + s * qk1;
        Q(r2, k + 1) = -s * qk + c * qk1;
      }
    }
  }
}

// Generalized inverse G (K G K = K, symmetric) of a symmetric positive
// semidefinite n x n matrix K (row major): Cholesky with diagonal pivoting
// stopped where the largest remaining pivot falls below rel_tol times the
// largest diagonal entry (numerical rank r), K = P^T [L11; L21][L11; L21]^T P,
// and G = P^T [L11^-T L11^-1, 0; 0, 0] P.  Then K G K = K, and G = K^-1
// when K is definite.  n^3/3 flops for the factor, r^3/3 each for the
// triangular inverse and the product.  Returns r.
int psd_ginv(int n, const double *K, double rel_tol, double *G) {
  // lower triangle, row major: row i holds A(i, 0..i) contiguously
  std::vector<double> a((size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int k = 0; k <= i; k++) a[(size_t)i * n + k] = K[(size_t)i * n + k];
  std::vector<int> piv(n);
  for (int i = 0; i < n; i++) piv[i] = i;
  auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
  double dmax = 0.0;
  for (int i = 0; i < n; i++) dmax = std::max(dmax, A(i, i));
  std::vector<double> colj(n);
  int r = 0;
  for (int j = 0; j < n; j++) {
    int q = j;
    for (int i = j + 1; i < n; i++)
      if (A(i, i) > A(q, q)) q = i;
    if (!(A(q, q) > rel_tol * dmax)) break;
    if (q != j) {  // symmetric swap of j < q in lower storage
      for (int k = 0; k < j; k++) std::swap(A(j, k), A(q, k));
      std::swap(A(j, j), A(q, q));
      for (int k = j + 1; k < q; k++) std::swap(A(k, j), A(q, k));
      for (int k = q + 1; k < n; k++) std::swap(A(k, j), A(k, q));
      std::swap(piv[j], piv[q]);
    }
    const double d = std::sqrt(A(j, j));
    A(j, j) = d;
    const double dinv = 1.0 / d;
    for (int i = j + 1; i < n; i++) colj[i] = A(i, j) *= dinv;
    // trailing update of the lower triangle, row by row
    for (int i = j + 1; i < n; i++) {
      const double lij = colj[i];
      double *ri = &A(i, 0);
      for (int k = j + 1; k <= i; k++) ri[k] -= lij * colj[k];
    }
    r = j + 1;
  }
  // X = L11^-1 (lower, row major): row i = -(1/L_ii) sum_{k<i} L_ik X(k, :)
  std::vector<double> X((size_t)r * r, 0.0);
  for (int i = 0; i < r; i++) {
    double *xi = &X[(size_t)i * r];
    for (int k = 0; k < i; k++) {
      const double lik = A(i, k);
      const double *xk = &X[(size_t)k * r];
      for (int c = 0; c <= k; c++) xi[c] += lik * xk[c];
    }
    const double inv = 1.0 / A(i, i);
    for (int c = 0; c < i; c++) xi[c] *= -inv;
    xi[i] = inv;
  }
  // H = X^T X (lower, row major) as rank-one updates by the rows of X
  std::vector<double> H((size_t)r * r, 0.0);
  for (int k = 0; k < r; k++) {
    const double *xk = &X[(size_t)k * r];
    for (int x = 0; x <= k; x++) {
      const double v = xk[x];
      double *hx = &H[(size_t)x * r];
      for (int y = 0; y <= x; y++) hx[y] += v * xk[y];
    }
  }
  std::fill(G, G + (size_t)n * n, 0.0);
  for (int x = 0; x < r; x++)
    for (int y = 0; y <= x; y++) {
      const double v = H[(size_t)x * r + y];
      G[(size_t)piv[x] * n + piv[y]] = v;
      G[(size_t)piv[y] * n + piv[x]] = v;
    }
  return r;
}

// Rayleigh-Ritz of one pair: the k largest eigenpairs of (GA, GB) (order m,
// leading dimension ld) on the range where GB is numerically positive;
// coefficients into C (m x k, stride kmax) and values into lam.  Returns the
// number of pairs found (fewer than k only if the block lost rank).
static int rayleigh_ritz(int m, int k, int ld, const double *ga, const double *gb, int kmax,
                         double *C, double *lam) {
  // the Gram matrices scaled to a unit B-diagonal first, so that the drop
  // threshold measures linear dependence, not column size
  std::vector<double> B((size_t)m * m), A((size_t)m * m), bw, bv, w, y, dsc(m);
  for (int i = 0; i < m; i++) {
    const double g = gb[i * ld + i];
    dsc[i] = g > 0.0 ? 1.0 / std::sqrt(g) : 0.0;
  }
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) {
      B[(size_t)i * m + j] = 0.5 * (gb[i * ld + j] + gb[j * ld + i]) * dsc[i] * dsc[j];
      A[(size_t)i * m + j] = 0.5 * (ga[i * ld + j] + ga[j * ld + i]) * dsc[i] * dsc[j];
    }
  sym_eig(m, B, bw, bv);
  double bmax = 0.0;
  for (double x : bw) bmax = std::max(bmax, x);
  std::vector<int> keep;
  for (int i = 0; i < m; i++)
    if (bw[i] > 1e-8 * bmax) keep.push_back(i);
  const int r = (int)keep.size();
  // T = D U_keep diag(bw)^{-1/2};  M = T^T A T (on the unscaled Gram matrices)
  std::vector<double> T((size_t)m * r), M((size_t)r * r, 0.0), AT((size_t)m * r, 0.0);
  for (int i = 0; i < m; i++)
    for (int c = 0; c < r; c++) T[(size_t)i * r + c] = bv[(size_t)i * m + keep[c]] / std::sqrt(bw[keep[c]]);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) {
      A[(size_t)i * m + j] = 0.5 * (ga[i * ld + j] + ga[j * ld + i]);
    }
  for (int i = 0; i < m; i++)
    for (int c = 0; c < r; c++) T[(size_t)i * r + c] *= dsc[i];
  for (int i = 0; i < m; i++)
    for (int c = 0; c < r; c++) {
      double s = 0.0;
      for (int j = 0; j < m; j++) s += A[(size_t)i * m + j] * T[(size_t)j * r + c];
      AT[(size_t)i * r + c] = s;
    }
  for (int a = 0; a < r; a++)
    for (int c = 0; c < r; c++) {
      double s = 0.0;
      for (int i = 0; i < m; i++) s += T[(size_t)i * r + a] * AT[(size_t)i * r + c];
      M[(size_t)a * r + c] = s;
    }
  sym_eig(r, M, w, y);
  std::vector<int> ord(r);
  for (int i = 0; i < r; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] > w[b]; });
  const int kk = std::min(k, r);
  for (int c = 0; c < k; c++) {
    lam[c] = c < kk ? w[ord[c]] : 0.0;
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      if (c < kk)
        for (int a = 0; a < r; a++) s += T[(size_t)i * r + a] * y[(size_t)a * r + ord[c]];
      C[(size_t)i * kmax + c] = s;
    }
  }
  return kk;
}

static double unit_random(uint64_t seed, uint64_t idx) {
  uint64_t z = (seed ^ idx) + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <class T> struct ABuf {
  T *p = nullptr;
  size_t n = 0;
  ABuf() = default;
  ABuf(const ABuf &) = delete;
  ABuf &operator=(const ABuf &) = delete;
  ~ABuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t cnt) {
    if (p) (void)hipFree(p);
    n = cnt;
    AD_CHECK(hipMalloc(&p, std::max<size_t>(cnt, 1) * sizeof(T)));
  }
  void swap(ABuf &o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
  void upload(const std::vector<T> &v, hipStream_t st) {
    alloc(v.size());
    if (!v.empty()) AD_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  }
};

// ------------------------------------------------------------------ driver
hipStream_t device_stream(Solver &S);  // compute stream (device.hip)

void adaptive_constraints(Solver &S) {
  const char *R = "ADAPTIVITY_SOLVE_EIGENVECTORS";
  const int nsub = S.nsub;
  double tph = wall_time();
  auto ph = [&](const char *name) {
    const double now = wall_time();
    if (S.verbose >= 2) std::printf("   adaptivity: %-28s %8.4f s\n", name, now - tph);
    tph = now;
  };
  hipStream_t st = device_stream(S);
  S.adapt_pairs.clear();
  S.adapt_eigs.clear();
  std::vector<int> slow, ntis;
  const double *pool = device_schur_tiles(S, slow, ntis);
  PairPcSource pcs;
  device_pair_pc_source(S, pcs);
  const int first_face = S.ncorner + S.nedge;
  // coarse node of every constraint row, and the subdomains of every face
  auto node_of = [&](int cdof) {
    return int(std::upper_bound(S.kdofc.begin(), S.kdofc.end(), cdof) - S.kdofc.begin()) - 1;
  };
  std::map<int, std::vector<int>> face_subs;
  for (int is = 0; is < nsub; is++)
    for (int id : S.subs[is].cnodes)
      if (id >= first_face) face_subs[id].push_back(is);
  // per subdomain: device position of each reference interface index, and
  // the weight coefficients
  std::vector<std::vector<int>> ref2dev(nsub);
  std::vector<std::vector<double>> rho(nsub);
  for (int is = 0; is < nsub; is++) {
    const Sub &s = S.subs[is];
    ref2dev[is].assign(s.ni, 0);
    for (int k = 0; k < s.ni; k++) ref2dev[is][s.dev_ifc[k]] = k;
    if (!weight_coefficients(S, is, rho[is])) fatal(R, "dof_data missing for weights type 3");
  }
  // ---- pairs: projection columns, common dofs, block widths
  struct HostPair {
    int face, i, j, k, neig;
    double tol;
  };
  std::vector<HostPair> hp;
  std::vector<PairDesc> pd;
  std::vector<int> qptr{0}, qpos, ma, mb, cmap;
  std::vector<double> qval, mwa, mwb;
  int64_t off = 0, woff = 0;
  int kmax = 0;
  for (auto &f : face_subs) {
    if (f.second.size() != 2) continue;
    const int i = f.second[0], j = f.second[1];
    const Sub &si = S.subs[i], &sj = S.subs[j];
    PairDesc d{};
    d.NIi = si.NI;
    d.n = si.NI + sj.NI;
    // common dofs by global dof
    std::map<int, int> gj;
    for (int r = 0; r < sj.ni; r++) gj[sj.isvgvn[sj.ifc_dof[r]]] = r;
    d.m0 = (int)ma.size();
    for (int r = 0; r < si.ni; r++) {
      auto it = gj.find(si.isvgvn[si.ifc_dof[r]]);
      if (it == gj.end()) continue;
      const double den = rho[i][r] + rho[j][it->second];
      ma.push_back(ref2dev[i][r]);
      mb.push_back(si.NI + ref2dev[j][it->second]);
      mwa.push_back(rho[i][r] / den);
      mwb.push_back(rho[j][it->second] / den);
    }
    d.m1 = (int)ma.size();
    const int ndofcomm = d.m1 - d.m0;
    // shared constraint rows [C_i; -C_j], normalized (disjoint supports)
    d.q0 = (int)qptr.size() - 1;
    int ncommon = 0;
    for (int ci = 0; ci < si.nconstr; ci++)
      for (int cj = 0; cj < sj.nconstr; cj++) {
        if (si.indrowc[ci] != sj.indrowc[cj]) continue;
        ncommon++;
        double nrm = 0.0;
        for (int q = si.cptr[ci]; q < si.cptr[ci + 1]; q++) nrm += si.cval[q] * si.cval[q];
        for (int q = sj.cptr[cj]; q < sj.cptr[cj + 1]; q++) nrm += sj.cval[q] * sj.cval[q];
        nrm = std::sqrt(nrm);
        for (int q = si.cptr[ci]; q < si.cptr[ci + 1]; q++)
          qpos.push_back(ref2dev[i][si.cidx[q]]), qval.push_back(si.cval[q] / nrm);
        for (int q = sj.cptr[cj]; q < sj.cptr[cj + 1]; q++)
          qpos.push_back(si.NI + ref2dev[j][sj.cidx[q]]), qval.push_back(-sj.cval[q] / nrm);
        qptr.push_back((int)qpos.size());
      }
    d.q1 = (int)qptr.size() - 1;
    const int neig = std::max(0, std::min(S.adapt_nvec, ndofcomm - ncommon));
    if (neig <= 0) {
      S.adapt_pairs.push_back({f.first, i, j, 0, 0, 0});
      S.adapt_eigs.push_back({});
      continue;
    }
    d.k = neig;
    d.active = 1;
    double trA = 0.0;
    for (double x : rho[i]) trA += x;
    for (double x : rho[j]) trA += x;
    d.tol = S.adapt_tol * trA / (si.ni + sj.ni);
    // preconditioner data: S'^{-1} tiles, Z / Phi, pair coarse dofs
    d.zoi = pcs.zoff[i], d.zoj = pcs.zoff[j];
    d.nci = si.nconstr, d.ncj = sj.nconstr;
    std::vector<int> cd(si.indrowc.begin(), si.indrowc.begin() + si.nconstr);
    cd.insert(cd.end(), sj.indrowc.begin(), sj.indrowc.begin() + sj.nconstr);
    std::sort(cd.begin(), cd.end());
    cd.erase(std::unique(cd.begin(), cd.end()), cd.end());
    d.ncp = (int)cd.size();
    d.cm0 = (int)cmap.size();
    for (const Sub *sb : {&si, &sj})
      for (int c = 0; c < sb->nconstr; c++)
        cmap.push_back(int(std::lower_bound(cd.begin(), cd.end(), sb->indrowc[c]) - cd.begin()));
    hp.push_back({f.first, i, j, neig, neig, d.tol});
    pd.push_back(d);
    kmax = std::max(kmax, neig);
  }
  const int np = (int)pd.size();
  int64_t kcsz = 0;
  size_t pc_lds = 0;
  for (int p = 0; p < np; p++) {
    pd[p].off = off, pd[p].woff = woff;
    off += (int64_t)pd[p].n * kmax;
    woff += (int64_t)pd[p].n * 2 * kmax;
    pd[p].kc0 = kcsz;
    kcsz += (int64_t)pd[p].ncp * pd[p].ncp;
    pc_lds = std::max(pc_lds, sizeof(double) * (2 * (size_t)pd[p].ncp + 2 * (size_t)(pd[p].nci + pd[p].ncj)));
  }
  if (pc_lds > 65536) fatal(R, "too many constraints on a pair for the pair coarse problem");
  // K^+ of every pair: a generalized inverse of the pair's coarse matrix
  // K = sum_s R_s^T Ac_s R_s (Ac_s = Phi_s^T S_s Phi_s; a floating pair has
  // rigid modes in its null space), by pivoted Cholesky truncated at pivots
  // below 1e-12 of the largest diagonal entry.  The residuals LOBPCG
  // preconditions have no rigid component, so any generalized inverse gives
  // the same correction up to rigid modes, which lie in the null space of B
  // and drop out of the Rayleigh-Ritz step.
  std::vector<double> kp((size_t)kcsz, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; p++) {
    const PairDesc &d = pd[p];
    const int m = d.ncp;
    std::vector<double> K((size_t)m * m, 0.0);
    const int sub2[2] = {hp[p].i, hp[p].j}, nc2[2] = {d.nci, d.ncj};
    const int *cm = cmap.data() + d.cm0;
    for (int sd = 0; sd < 2; sd++) {
      const std::vector<double> &Ac = pcs.Ac[sub2[sd]];
      const int nc = nc2[sd];
      for (int a = 0; a < nc; a++)
        for (int b = 0; b < nc; b++) K[(size_t)cm[a] * m + cm[b]] += Ac[(size_t)a * nc + b];
      cm += nc;
    }
    psd_ginv(m, K.data(), 1e-12, kp.data() + d.kc0);
  }
  if (S.verbose >= 2 && np > 0) {
    double mean = 0.0;
    int mx = 0;
    for (const PairDesc &d : pd) mean += d.ncp, mx = std::max(mx, d.ncp);
    std::printf("   adaptivity: %d pairs, pair coarse order mean %.1f max %d, block %d\n", np, mean / np,
                mx, kmax);
  }
  ph("pair coarse inverses");
  ph("pairs");
  // ---- the products with S_s (operator applies: the k A inputs and the k
  // B inputs of every pair side) and with S'_s^{-1} (preconditioner: k
  // columns) gathered per subdomain: each subdomain's matrix multiplies all
  // of its pairs' columns in one tile task list (subdomain, tile row, 64
  // columns), the columns addressed in place in the wide pair buffers
  std::vector<std::vector<int64_t>> capply(nsub), cpc(nsub);
  for (int p = 0; p < np; p++)
    for (int side = 0; side < 2; side++) {
      const PairDesc &d = pd[p];
      const int is = side ? hp[p].j : hp[p].i;
      const int64_t so = d.woff + (side ? d.NIi : 0);
      for (int c = 0; c < d.k; c++) {
        capply[is].push_back(so + (int64_t)c * d.n);
        cpc[is].push_back(so + (int64_t)c * d.n);
      }
      for (int c = 0; c < d.k; c++) capply[is].push_back(so + (int64_t)(kmax + c) * d.n);
    }
  std::vector<int64_t> colA, colP;
  std::vector<SymmSub> subA(nsub), subP(nsub);
  std::vector<SymmTask> taskA, taskP;
  for (int is = 0; is < nsub; is++) {
    for (int pass = 0; pass < 2; pass++) {
      std::vector<int64_t> &c = pass ? cpc[is] : capply[is];
      std::vector<int64_t> &all = pass ? colP : colA;
      std::vector<SymmTask> &tasks = pass ? taskP : taskA;
      SymmSub &sd = pass ? subP[is] : subA[is];
      c.resize((c.size() + 63) / 64 * 64, -1);
      sd.tile0 = pass ? pcs.pbase[is] : slow[is];
      sd.nt = ntis[is];
      sd.col0 = (int)all.size();
      sd.ncol = (int)c.size();
      all.insert(all.end(), c.begin(), c.end());
      for (int I = 0; I < sd.nt; I++)
        for (int c0 = 0; c0 < sd.ncol; c0 += 64) tasks.push_back({is, I, c0, 0});
    }
  }
  // ---- device buffers
  ABuf<PairDesc> dpairs;
  ABuf<int> dqptr, dqpos, dma, dmb, dcmap;
  ABuf<double> dqval, dmwa, dmwb, dkp;
  ABuf<int64_t> dcolA, dcolP;
  ABuf<SymmSub> dsubA, dsubP;
  ABuf<SymmTask> dtaskA, dtaskP;
  dpairs.upload(pd, st);
  dqptr.upload(qptr, st), dqpos.upload(qpos, st), dqval.upload(qval, st);
  dma.upload(ma, st), dmb.upload(mb, st), dmwa.upload(mwa, st), dmwb.upload(mwb, st);
  dcolA.upload(colA, st), dcolP.upload(colP, st);
  dsubA.upload(subA, st), dsubP.upload(subP, st);
  dtaskA.upload(taskA, st), dtaskP.upload(taskP, st);
  dcmap.upload(cmap, st), dkp.upload(kp, st);
  enum { X, AX, BX, W, AW, BW, P, AP, BP, XN, AXN, BXN, PN, APN, BPN, NBLK };
  std::vector<ABuf<double>> blk(NBLK);
  for (auto &b : blk) b.alloc(std::max<int64_t>(off, 1));
  ABuf<double> win, wout, dlam, dnrm, dGA, dGB, dC;
  win.alloc(std::max<int64_t>(woff, 1));
  wout.alloc(std::max<int64_t>(woff, 1));
  // the tile products read every pair's columns, also those of pairs that
  // have stopped iterating: start from zeros
  AD_CHECK(hipMemsetAsync(win.p, 0, std::max<int64_t>(woff, 1) * sizeof(double), st));
  dlam.alloc((size_t)np * kmax);
  dnrm.alloc((size_t)np * kmax);
  const int ld = 3 * kmax;
  dGA.alloc((size_t)np * ld * ld);
  dGB.alloc((size_t)np * ld * ld);
  dC.alloc((size_t)np * ld * kmax);
  const dim3 pk(std::max(np, 1), std::max(kmax, 1));
  // block operator: A x and B x for the k columns of block b
  auto apply = [&](int b, int ab, int bb) {
    if (np == 0) return;
    hipLaunchKernelGGL(k_ad_copy, pk, dim3(256), 0, st, dpairs.p, blk[b].p, 0, 0, win.p, 1, kmax);
    hipLaunchKernelGGL(k_ad_project, pk, dim3(64), 0, st, dpairs.p, dqptr.p, dqpos.p, dqval.p,
                       win.p, 1, kmax);
    hipLaunchKernelGGL(k_ad_jump, pk, dim3(256), 0, st, dpairs.p, dma.p, dmb.p, dmwa.p, dmwb.p,
                       win.p, 1, kmax, win.p, 1, 0, 0);
    launch_symm_cols<double>(st, pool, dsubA.p, dtaskA.p, (int)taskA.size(), dcolA.p, win.p, wout.p);
    hipLaunchKernelGGL(k_ad_jump, pk, dim3(256), 0, st, dpairs.p, dma.p, dmb.p, dmwa.p, dmwb.p,
                       wout.p, 1, 0, blk[ab].p, 0, 0, 1);
    hipLaunchKernelGGL(k_ad_project, pk, dim3(64), 0, st, dpairs.p, dqptr.p, dqpos.p, dqval.p,
                       blk[ab].p, 0, 0);
    hipLaunchKernelGGL(k_ad_copy, pk, dim3(256), 0, st, dpairs.p, wout.p, 1, kmax, blk[bb].p, 0, 0);
    hipLaunchKernelGGL(k_ad_project, pk, dim3(64), 0, st, dpairs.p, dqptr.p, dqpos.p, dqval.p,
                       blk[bb].p, 0, 0);
  };
  // preconditioner: W = P T W on the k columns of W (W already projected)
  double t_pc_coarse = 0.0;  // verbose 3: the pair coarse correction's share
  auto precondition = [&]() {
    if (np == 0) return;
    hipLaunchKernelGGL(k_ad_copy, pk, dim3(256), 0, st, dpairs.p, blk[W].p, 0, 0, win.p, 1, 0);
    if (pcs.inv_float)
      launch_symm_cols<float>(st, (const float *)pcs.inv, dsubP.p, dtaskP.p, (int)taskP.size(), dcolP.p,
                              win.p, wout.p);
    else
      launch_symm_cols<double>(st, (const double *)pcs.inv, dsubP.p, dtaskP.p, (int)taskP.size(),
                               dcolP.p, win.p, wout.p);
    double t0 = 0.0;
    if (S.verbose >= 3) {
      AD_CHECK(hipStreamSynchronize(st));
      t0 = wall_time();
    }
    hipLaunchKernelGGL(k_ad_coarse_pc, pk, dim3(256), pc_lds, st, dpairs.p, pcs.Z, pcs.Phi, dcmap.p,
                       dkp.p, win.p, wout.p);
    if (S.verbose >= 3) {
      AD_CHECK(hipStreamSynchronize(st));
      t_pc_coarse += wall_time() - t0;
    }
    hipLaunchKernelGGL(k_ad_copy, pk, dim3(256), 0, st, dpairs.p, wout.p, 1, 0, blk[W].p, 0, 0);
    hipLaunchKernelGGL(k_ad_project, pk, dim3(64), 0, st, dpairs.p, dqptr.p, dqpos.p, dqval.p,
                       blk[W].p, 0, 0);
  };
  std::vector<double> hGA((size_t)np * ld * ld), hGB((size_t)np * ld * ld), hC((size_t)np * ld * kmax, 0.0),
      hlam((size_t)np * kmax, 0.0), hnrm((size_t)np * kmax);
  std::vector<int> iters(np, 0);
  // Rayleigh-Ritz of every active pair on nb blocks, then the update
  double t_rr = 0.0;  // host part of the Rayleigh-Ritz steps
  auto rr_update = [&](int nb) {
    Blocks in{{blk[X].p, blk[W].p, blk[P].p}, {blk[AX].p, blk[AW].p, blk[AP].p}, {blk[BX].p, blk[BW].p, blk[BP].p}};
    hipLaunchKernelGGL(k_ad_gram, dim3(std::max(np, 1)), dim3(256), 0, st, dpairs.p, in, nb, kmax,
                       dGA.p, dGB.p);
    AD_CHECK(hipMemcpyAsync(hGA.data(), dGA.p, hGA.size() * 8, hipMemcpyDeviceToHost, st));
    AD_CHECK(hipMemcpyAsync(hGB.data(), dGB.p, hGB.size() * 8, hipMemcpyDeviceToHost, st));
    AD_CHECK(hipStreamSynchronize(st));
    const double t0 = wall_time();
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < np; p++) {
      if (!pd[p].active) continue;
      const int m = nb * pd[p].k;
      std::fill(hC.begin() + (size_t)p * ld * kmax, hC.begin() + (size_t)(p + 1) * ld * kmax, 0.0);
      rayleigh_ritz(m, pd[p].k, ld, hGA.data() + (size_t)p * ld * ld, hGB.data() + (size_t)p * ld * ld,
                    kmax, hC.data() + (size_t)p * ld * kmax, hlam.data() + (size_t)p * kmax);
    }
    t_rr += wall_time() - t0;
    AD_CHECK(hipMemcpyAsync(dC.p, hC.data(), hC.size() * 8, hipMemcpyHostToDevice, st));
    AD_CHECK(hipMemcpyAsync(dlam.p, hlam.data(), hlam.size() * 8, hipMemcpyHostToDevice, st));
    Blocks out{{blk[XN].p, nullptr, blk[PN].p}, {blk[AXN].p, nullptr, blk[APN].p}, {blk[BXN].p, nullptr, blk[BPN].p}};
    int maxn = 1;
    for (const PairDesc &d : pd) maxn = std::max(maxn, d.n);
    hipLaunchKernelGGL(k_ad_combine, dim3((maxn + 255) / 256, std::max(np, 1)), dim3(256), 0, st, dpairs.p,
                       in, nb, kmax, dC.p, out);
    blk[X].swap(blk[XN]), blk[AX].swap(blk[AXN]), blk[BX].swap(blk[BXN]);
    blk[P].swap(blk[PN]), blk[AP].swap(blk[APN]), blk[BP].swap(blk[BPN]);
  };
  if (np > 0) {
    // random start in range(P), padding zero
    std::vector<double> x0((size_t)off, 0.0);
    for (int p = 0; p < np; p++) {
      const PairDesc &d = pd[p];
      const Sub &si = S.subs[hp[p].i], &sj = S.subs[hp[p].j];
      for (int c = 0; c < d.k; c++)
        for (int r = 0; r < d.n; r++) {
          const bool real = r < d.NIi ? r < si.ni : r - d.NIi < sj.ni;
          if (real)
            x0[d.off + (int64_t)c * d.n + r] =
                2.0 * unit_random(0x5eedull + (uint64_t)hp[p].face, (uint64_t)c * 1000003ull + r) - 1.0;
        }
    }
    AD_CHECK(hipMemcpyAsync(blk[X].p, x0.data(), x0.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ad_project, pk, dim3(64), 0, st, dpairs.p, dqptr.p, dqpos.p, dqval.p,
                       blk[X].p, 0, 0);
    apply(X, AX, BX);
    rr_update(1);
    int nb = 2;
    // verbose 3: device time per section (each section synchronized)
    double sec[4] = {0, 0, 0, 0};  // refresh applies, preconditioner, apply W, Rayleigh-Ritz + update
    auto timed = [&](int k, auto &&f) {
      if (S.verbose < 3) return f();
      AD_CHECK(hipStreamSynchronize(st));
      const double t0 = wall_time();
      f();
      AD_CHECK(hipStreamSynchronize(st));
      sec[k] += wall_time() - t0;
    };
    for (int it = 1; it <= S.adapt_maxit; it++) {
      // the A and B images of X and P are carried by the block updates;
      // every ten iterations they are recomputed, so that rounding carried
      // through the recurrences does not leave the Gram matrices
      // inconsistent with the blocks
      if (it % 10 == 0) {
        timed(0, [&] {
          apply(X, AX, BX);
          apply(P, AP, BP);
        });
      }
      hipLaunchKernelGGL(k_ad_resid, pk, dim3(256), 0, st, dpairs.p, dlam.p, kmax, blk[AX].p,
                         blk[BX].p, blk[W].p, dnrm.p);
      AD_CHECK(hipMemcpyAsync(hnrm.data(), dnrm.p, hnrm.size() * 8, hipMemcpyDeviceToHost, st));
      AD_CHECK(hipStreamSynchronize(st));
      int nactive = 0;
      for (int p = 0; p < np; p++) {
        if (!pd[p].active) continue;
        bool conv = true;
        for (int c = 0; c < pd[p].k; c++)
          if (std::sqrt(hnrm[(size_t)p * kmax + c]) > hp[p].tol) conv = false;
        if (conv) pd[p].active = 0;
        else iters[p] = it, nactive++;
      }
      if (nactive == 0) break;
      AD_CHECK(hipMemcpyAsync(dpairs.p, pd.data(), pd.size() * sizeof(PairDesc), hipMemcpyHostToDevice, st));
      // W = P T P R (the residual kernel wrote R)
      hipLaunchKernelGGL(k_ad_project, pk, dim3(64), 0, st, dpairs.p, dqptr.p, dqpos.p, dqval.p,
                         blk[W].p, 0, 0);
      timed(1, [&] { precondition(); });
      timed(2, [&] { apply(W, AW, BW); });
      timed(3, [&] { rr_update(nb); });
      nb = 3;
    }
    if (S.verbose >= 3)
      std::printf("   adaptivity: LOBPCG sections: refresh %.4f s, preconditioner %.4f s, apply %.4f s, "
                  "Rayleigh-Ritz and update %.4f s (coarse correction %.4f s of the preconditioner)\n",
                  sec[0], sec[1], sec[2], sec[3], t_pc_coarse);
    // A X once more on the final X of every pair: the constraints
    for (PairDesc &d : pd) d.active = 1;
    AD_CHECK(hipMemcpyAsync(dpairs.p, pd.data(), pd.size() * sizeof(PairDesc), hipMemcpyHostToDevice, st));
    apply(X, AX, BX);
  }
  std::vector<double> hAX((size_t)off);
  if (np > 0) {
    AD_CHECK(hipMemcpyAsync(hAX.data(), blk[AX].p, hAX.size() * 8, hipMemcpyDeviceToHost, st));
    AD_CHECK(hipStreamSynchronize(st));
  }
  if (S.verbose >= 2) std::printf("   adaptivity: host Rayleigh-Ritz %.4f s of the LOBPCG\n", t_rr);
  ph("LOBPCG");
  // ---- every face: the support of its arithmetic rows (global dofs,
  // sorted) and those rows over it
  std::vector<int> faces;
  for (auto &f : face_subs) faces.push_back(f.first);
  const int nfaces = (int)faces.size();
  std::map<int, int> face_idx;
  for (int q = 0; q < nfaces; q++) face_idx[faces[q]] = q;
  std::vector<std::vector<int>> face_dofs(nfaces);
  std::vector<std::vector<std::vector<double>>> face_M(nfaces);
#pragma omp parallel for schedule(dynamic)
  for (int q = 0; q < nfaces; q++) {
    const int face = faces[q];
    const Sub &s = S.subs[face_subs.at(face)[0]];
    std::vector<int> &gd = face_dofs[q];
    for (int c = 0; c < s.nconstr; c++) {
      if (node_of(s.indrowc[c]) != face) continue;
      for (int e = s.cptr[c]; e < s.cptr[c + 1]; e++) gd.push_back(s.isvgvn[s.ifc_dof[s.cidx[e]]]);
    }
    std::sort(gd.begin(), gd.end());
    gd.erase(std::unique(gd.begin(), gd.end()), gd.end());
    for (int c = 0; c < s.nconstr; c++) {
      if (node_of(s.indrowc[c]) != face) continue;
      std::vector<double> v(gd.size(), 0.0);
      for (int e = s.cptr[c]; e < s.cptr[c + 1]; e++) {
        const int g = s.isvgvn[s.ifc_dof[s.cidx[e]]];
        v[std::lower_bound(gd.begin(), gd.end(), g) - gd.begin()] = s.cval[e];
      }
      face_M[q].push_back(std::move(v));
    }
  }
  // ---- the adaptive rows of every pair after its face's arithmetic rows,
  // largest eigenvalue first: c = -P (I - E^T) S (I - E) P v (the
  // reference's operator) on side i, restricted to the face dofs
  S.adapt_est = -1.0;
  std::vector<int> nadv(np, 0);
  for (int p = 0; p < np; p++) {
    const PairDesc &d = pd[p];
    const double *lam = hlam.data() + (size_t)p * kmax;
    int nad = 0;
    while (nad < d.k && lam[nad] >= S.adapt_threshold) nad++;
    const double est = nad < d.k ? lam[nad] : lam[nad - 1];
    S.adapt_est = std::max(S.adapt_est, est);
    S.adapt_pairs.push_back({hp[p].face, hp[p].i, hp[p].j, d.k, nad, iters[p]});
    S.adapt_eigs.push_back(std::vector<double>(lam, lam + d.k));
    nadv[p] = nad;
  }
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; p++) {
    const PairDesc &d = pd[p];
    const int nad = nadv[p];
    if (nad == 0) continue;
    const int i = hp[p].i, q = face_idx.at(hp[p].face);
    const Sub &si = S.subs[i];
    // interface position of every face dof on side i
    std::vector<std::pair<int, int>> g2r(si.ni);
    for (int r = 0; r < si.ni; r++) g2r[r] = {si.isvgvn[si.ifc_dof[r]], r};
    std::sort(g2r.begin(), g2r.end());
    const std::vector<int> &gd = face_dofs[q];
    std::vector<int> dev(gd.size());
    for (size_t k = 0; k < gd.size(); k++) {
      auto it = std::lower_bound(g2r.begin(), g2r.end(), std::make_pair(gd[k], -1));
      dev[k] = (it != g2r.end() && it->first == gd[k]) ? ref2dev[i][it->second] : -1;
    }
    for (int c = 0; c < nad; c++) {
      const double *a = hAX.data() + d.off + (int64_t)c * d.n;
      std::vector<double> v(gd.size(), 0.0);
      for (size_t k = 0; k < gd.size(); k++) v[k] = dev[k] >= 0 ? -a[dev[k]] : 0.0;
      face_M[q].push_back(std::move(v));
    }
  }
  // ---- every face block orthonormalized once (both subdomains get the
  // same rows); coarse dofs renumbered with the new counts
  std::vector<std::vector<std::vector<double>>> face_qv(nfaces);
#pragma omp parallel for schedule(dynamic)
  for (int q = 0; q < nfaces; q++) face_qv[q] = orthonormal_rows(face_M[q], (int)face_dofs[q].size());
  std::map<int, std::vector<std::vector<double>>> face_q;
  std::map<int, std::vector<int>> face_dof_map;
  for (int q = 0; q < nfaces; q++) {
    face_q[faces[q]] = std::move(face_qv[q]);
    face_dof_map[faces[q]] = face_dofs[q];
  }
  ph("face rows and orthonormalization");
  replace_coarse_node_rows(S, face_q, face_dof_map);
  ph("new constraint rows");
  if (S.verbose > 0) {
    int nad = 0;
    for (auto &q : S.adapt_pairs) nad += q[4];
    std::printf(" ADAPTIVITY_SOLVE_EIGENVECTORS: %d pairs, %d adaptive constraints, estimate %.4g, "
                "%d coarse dofs\n", (int)S.adapt_pairs.size(), nad, S.adapt_est, S.ndofc);
  }
}

}  // namespace bddc
