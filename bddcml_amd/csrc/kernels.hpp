// kernels.hpp -- host launchers of the HIP kernels (defined in *.hip).
// Tiles are 64x64 FP64, column major inside the tile, addressed by an int32
// tile id into one device pool: element (r,c) of tile t is
// pool[t*4096 + c*64 + r].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.hpp"

namespace bddc {

// ---- tile algebra -------------------------------------------------------
// op kinds of the batched GEMM-sum: C = alpha*C + beta * sum_p opA(A_p) opB(B_p)
enum GemmKind { GEMM_NT = 0, GEMM_NN = 1, GEMM_TN = 2 };
// post operations on E after the sum
enum GemmPost { POST_NONE = 0, POST_RIGHT_XT = 1, POST_LEFT_X = 2, POST_LEFT_XT = 3 };

// T = double or float (tile pool element type)
template <typename T>
void launch_tile_gemm(hipStream_t st, T *pool, const TileTask *tasks, const TilePair *pairs,
                      int ntasks, int kind, int post, double alpha, double beta);
// a symmetric matrix per subdomain (packed lower tiles from tile0, nt tile
// rows) times a set of columns scattered in memory (col0: first entry of the
// subdomain's column offsets)
struct SymmSub {
  int tile0, nt, col0, ncol;
};
struct SymmTask {
  int sub, I, c0, pad;
};
template <typename TP>
void launch_symm_cols(hipStream_t st, const TP *pool, const SymmSub *subs, const SymmTask *tasks,
                      int ntasks, const int64_t *cols, const double *in, double *out);
// C = C - sum A B^T as its own kernel symbol (Schur-complement block update)
// 1: tile-task launches hand each XCD one contiguous range of their task
// list (default); 0: plain blockIdx order
void set_task_order(int chunked);
// S = A22 - sum W^T W over the Schur task list (FP64 or FP32 tiles)
void launch_schur_gemm(hipStream_t st, double *pool, const TileTask *tasks, const TilePair *pairs,
                       int ntasks);
void launch_schur_gemm(hipStream_t st, float *pool, const TileTask *tasks, const TilePair *pairs,
                       int ntasks);
// C = C - sum A B^T ; C = chol(C) (lower, upper zeroed) ; X = inv(C)
template <typename T>
void launch_tile_potrf(hipStream_t st, T *pool, const TileTask *tasks, const TilePair *pairs,
                       int ntasks, int *err);
// tasks of mixed kinds in one launch: tile products with their own mode
// (task_mode) and tile factorizations (mode kTaskPotrf, C and inverse tile x)
void launch_tile_step(hipStream_t st, double *pool, const TileTask *tasks, const TilePair *pairs,
                      int ntasks, int *err);
// dst tiles (full storage) from lower tiles: for each job (src, dst, transpose)
struct MirrorJob {
  int src, dst, trans;
};
void launch_tile_mirror(hipStream_t st, double *pool, const MirrorJob *jobs, int njobs);
template <typename T> void launch_zero_tiles(hipStream_t st, T *pool, const int *ids, int n);
// clear ranges of consecutive tiles: ranges[k] = (first tile, count)
void launch_zero_ranges(hipStream_t st, double *pool, const int2 *ranges, int nranges);
// assembled triplets of each subdomain -> tile pool (see kernels_tile.hip)
struct ScatterSub {
  const int *ai, *aj;
  const double *av;
  int nnz;
  int foff;    // offset of the subdomain's Dirichlet flags (dofs)
  int poff;    // offset of its class's permutation (dofs)
  int toff;    // offset of its class's packed lower tile table (tile ids relative to base)
  int padoff;  // offset of its class's padding diagonals (element index relative to base)
  int npad;
  int base;    // first tile of the subdomain
};
template <typename T>
void launch_scatter_sub(hipStream_t st, T *pool, const ScatterSub *subs, int nsub, const int *perm,
                        const char *fix, const int *ttab, const int *pads, int *err);
// pool[idx[k]] = 1.0
void launch_scatter_ones(hipStream_t st, double *pool, const int64_t *idx, int64_t n);

// diff[i] = 1 where the triplet pattern (ai, aj: device pointers) of
// subdomain i differs from that of cand[i] (equal lengths cnt)
void launch_same_pattern(hipStream_t st, const uint64_t *pai, const uint64_t *paj, const int *cand,
                         const int64_t *cnt, int *diff, int nsub);

// ---- interface preconditioner pieces ------------------------------------
// Descriptor of one subdomain's dense interface matrices.
struct SubDesc {
  int ni, NI, nti, nc;    // interface dofs, padded, tile count, constraints
  int slow;               // first tile of S (lower tiles, packed by columns)
  int pbase;              // first tile of the packed lower factor of S' (factor pool)
  int plinv;              // first of the nti inverse diagonal factor tiles (factor pool)
  int ybase;              // first tile of Z = S'^{-1} C^T (nti x ncT, row-major; factor pool)
  int ncT;                // constraint tile columns, ceil(nc/64)
  int gtbase;             // G^{-1} as ncT x ncT tiles (factor pool)
  int phibase;            // Phi = Z G^{-1} as nti x ncT tiles (factor pool)
  int64_t loff;           // offset into local interface vectors
  int64_t zoff;           // offset into Z / Phi (NI x nc column major)
  int64_t goff;           // offset into G / Ginv (nc x nc)
  int coff;               // offset into constraint arrays (row pointers)
  double rho;             // augmented-Lagrangian scale
};
// S' = S + sum_c (rho/|D_c|) 1_c 1_c^T into the lower tiles of the factor
// pool ppool (element type T; FP32 in the mixed-precision mode) from S in
// pool (element type P: FP32 with the FP32 interior factorization)
template <typename P, typename T>
void launch_sprime(hipStream_t st, const P *pool, T *ppool, const SubDesc *subs,
                   const int *jobs_sub, const int *jobs_ij, const int *jobs_dst, int njobs,
                   const int *cid, const int *csize);
// right-hand sides C^T of the multi-RHS solve S' Z = C^T, written as the
// nti x ncT tile block at ybase: B(i, c) = 1 if interface dof i is in D_c
template <typename T>
void launch_ct_init(hipStream_t st, T *pool, const SubDesc *subs, int nsub, const int *cid,
                    int maxnti);
// Z (which 0) or Phi (which 1), column major NI x nc at zoff, from the
// tile block at ybase or phibase
template <typename T>
void launch_y_to_z(hipStream_t st, const T *pool, const SubDesc *subs, int nsub, double *Z,
                   int maxNI, int which);
// G^{-1} into the ncT x ncT tile block at gtbase (zero padded), the right
// operand of Phi = Z G^{-1} on the tile GEMM
template <typename T>
void launch_ginv_to_tiles(hipStream_t st, T *pool, const SubDesc *subs, int nsub, const double *G,
                          int maxncT);
// general constraint rows (several dense rows per glob, as adaptive
// constraints give): per interface dof (local flat index) the list
// dptr/dcon/dval of (global constraint index, value), constraints ascending;
// S' = S + sum_k omega_k c_k c_k^T and C^T tiles from those lists
template <typename P, typename T>
void launch_sprime_gen(hipStream_t st, const P *pool, T *ppool, const SubDesc *subs,
                       const int *jobs_sub, const int *jobs_ij, const int *jobs_dst, int njobs,
                       const int *dptr, const int *dcon, const double *dval, const double *omega);
template <typename T>
void launch_ct_init_gen(hipStream_t st, T *pool, const SubDesc *subs, int nsub, const int *dptr,
                        const int *dcon, const double *dval, int maxnti);
// G = C Z (row values cval, 1 when null), Ginv = G^{-1} (in place
// Gauss-Jordan, SPD), Acs = Ginv - Omega (Omega_c = rho/|D_c|, or omega[c])
void launch_g(hipStream_t st, const SubDesc *subs, int nsub, const int *cptr, const int *cdof,
              const double *cval, const double *Z, double *G);
void launch_ginv(hipStream_t st, const SubDesc *subs, int nsub, double *G, double *work, int *err);

void launch_acs(hipStream_t st, const SubDesc *subs, int nsub, const int *cptr,
                const double *omega, const double *G, double *Acs);
// coarse assembly: dst[k] = sum_{q in ptr[k]..ptr[k+1]} Acs[src[q]] into pool
void launch_coarse_assemble(hipStream_t st, double *pool, const int64_t *dst, const int *ptr,
                            const int64_t *src, int n, const double *Acs);

// ---- triangular sweeps of the partial factor ----------------------------
struct TrsvDesc {
  int nto, nti;          // interior / interface tile rows
  int rptr_off;          // offset into row lists
  int cptr_off;          // offset into column lists
  int linv;              // first Linv tile (nto consecutive)
  int64_t voff;          // offset into the device-ordered subdomain vector
};
// maxnt = largest nto+nti of the batch (sizes the vector held in LDS)
// the interior sweeps, the local preconditioner and the coarse residual keep
// their per-subdomain vector in LDS when it fits and in global memory
// otherwise; on = always in global memory (for testing)
void set_trsv_global(bool on);
// 0: launch forms chosen by the number of subdomains; 1: the many-subdomain
// forms; 2: the few-subdomain forms (testing)
void set_launch_forms(int forms);
// waves per workgroup of the interior sweeps: 0 chosen by the number of
// subdomains, 4 or 16 forced (a launch shape the autotuner tries)
void set_sweep_waves(int waves);
int launch_forms();
// P: element type of the interior factor tiles (double or float)
template <typename P>
void launch_trsv_fwd(hipStream_t st, const P *pool, const TrsvDesc *d, int nsub, const int *rptr,
                     const int *rk, const int *rt, double *v, int maxnt);
template <typename P>
void launch_trsv_bwd(hipStream_t st, const P *pool, const TrsvDesc *d, int nsub, const int *cptr,
                     const int *ck, const int *ct, double *v, int maxnt);
// largest dynamic LDS allocation a kernel may request on the current device
size_t max_dynamic_lds();

// ---- subdomain vectors between local dof order and device order ---------
struct OrderDesc {
  int64_t poff;  // offset of the subdomain's dofs in local-order arrays
  int64_t voff;  // offset of its device-ordered vector
  int ndofs;
  int permoff;   // offset of its permutation in perm
};
// vec[voff + perm[permoff + i]] = src[poff + i]
void launch_to_device_order(hipStream_t st, const OrderDesc *d, int nsub, const int *perm,
                            const double *src, double *vec, int maxndofs);
// dst[poff + i] = vec[voff + perm[permoff + i]]
void launch_from_device_order(hipStream_t st, const OrderDesc *d, int nsub, const int *perm,
                              const double *vec, double *dst, int maxndofs);
// r = f - A u per subdomain in local dof order, FP64, A from the uploaded
// triplets (values av) with the Dirichlet elimination of launch_scatter_sub;
// rowptr/ent (offsets rp_off/ent_off, shared by a class of identically
// structured subdomains): per local row the (triplet index, column) pairs
// of both triangles, columns ascending
struct ResidSub {
  const double *av;
  int64_t poff;     // offset of the subdomain's dofs in local-order arrays
  int64_t ent_off;  // offset of its class's entries
  int rp_off;       // offset of its class's row pointers
  int foff;         // offset of its Dirichlet flags
  int ndofs;
  int pad;
};
void launch_resid_rows(hipStream_t st, const ResidSub *subs, int nsub, const int *rowptr,
                       const int2 *ent, const char *fix, const double *f, const double *u,
                       double *r, int maxndofs);
// every local copy of unique interface dof g: vec[pos[q]] = u[g], q in ptr[g]..ptr[g+1]
// three levels: vec[p] = w[p] * rc[src[p]] (0 where src < 0), p < n; and
// uc[g] = vec[pos[g]], g < ndofc
void launch_level_fill(hipStream_t st, int64_t n, const int *src, const double *w, const double *rc,
                       double *vec);
void launch_level_gather(hipStream_t st, int ndofc, const int *pos, const double *vec, double *uc);
void launch_ifc_scatter(hipStream_t st, int n, const int *ptr, const int64_t *pos, const double *u,
                        double *vec);

// ---- iteration kernels ---------------------------------------------------
// Symmetric tile GEMV over the packed lower tiles of each subdomain's S:
// y_local[loff + I*64 + r] = sum_J S(I,J) x_J with x_J[r] = xin[map[loff + J*64 + r]]
// (0 where map < 0).  One workgroup per tile column (SymvCol) writes partial
// slots (I, J) of `part` (64 doubles each, packed like the tiles); the reduce
// sums each row's slots in a fixed order.
struct SymvCol {
  int tile0;  // tile (J, J); the column's tiles (I, J), I = J..nti-1, follow
  int nt;     // nti - J
  int mat;    // subdomain
  int J;
  int pslot;  // partial slot of (J, J)
};
struct SymvRow {
  int mat, I;
  int psub;   // partial slot of the subdomain's (0, 0)
  int nti;
};
// block-sparse tile GEMV: per task y[yoff..yoff+64) = (b ? b[boff..] : 0) +
// sign * sum over its pairs of op(tile) x[xoff..xoff+64), op the tile or
// (trans) its transpose; tiles column major in pool.  A task may write the
// segment its base reads (in place) but no segment another task reads.
struct GemvTask {
  int64_t yoff, boff;
  int p0, p1;
};
struct GemvPair {
  int tile, pad;
  int64_t xoff;
};
void launch_tile_gemv(hipStream_t st, const double *pool, const GemvTask *tasks, int ntasks,
                      const GemvPair *pairs, const double *x, const double *b, double *y, int trans,
                      double sign);
template <typename T>
void launch_symv(hipStream_t st, const T *pool, const SymvCol *cols, int ncols,
                 const SymvRow *rows, int nrows, const int64_t *loff, const int *map,
                 const double *wts, const double *xin, double *part, double *yout);
// y = A x for a dense nt x nt tile matrix in full storage (tile (I, J) at
// base + I*nt + J); part holds nt*nt*64 doubles (coarse inverse)
void launch_dense_gemv(hipStream_t st, const double *pool, int base, int nt, const double *x,
                       double *part, double *y);
// unique y[g] = sum_{q} (w? w[src]:1) * (ylocal[src] + (phiu? phiu[src]:0))
void launch_assemble(hipStream_t st, int n, const int *ptr, const int64_t *src, const double *w,
                     const double *ylocal, double *y);
// rc_s = Phi_s^T (w .* x_gathered)  -> coarse residual assembled by ptr/src
void launch_phit(hipStream_t st, const SubDesc *subs, int nsub, const double *Phi,
                 const int *map, const double *w, const double *x, double *rcs, int maxnc,
                 int maxNI);
void launch_coarse_gather(hipStream_t st, int n, const int *ptr, const int *src,
                          const double *rcs, double *rc);
// Local part of the BDDC preconditioner on every subdomain, one workgroup
// each: x = w .* r(map); z = S'^{-1} x through the packed factor of S'
// (forward and backward tile sweeps); zl = z + Phi (uc(indrowc) - Z^T x).
// deep: more loads in flight per workgroup at lower occupancy, for few or
// uneven subdomains (same arithmetic, same results).  ysym (may be null):
// S'^{-1} x already applied by an explicit inverse; the sweeps are skipped.
template <typename T>
void launch_kapply(hipStream_t st, const T *pool, const SubDesc *subs, int nsub, const int *map,
                   const double *w, const double *r, const double *Phi, const double *Z,
                   const int *cmap, const double *uc, const double *ysym, double *zl, int maxNI,
                   int maxnc, bool deep);

// deterministic reductions and vector updates on unique vectors
constexpr int DOT_BLOCKS = 240;
// out[d] = sum_i a_d[i] * b_d[i] for up to 4 pairs; partial has DOT_BLOCKS*4 doubles
struct DotArgs {
  const double *a[4];
  const double *b[4];
  int ndot;
};
void launch_dots(hipStream_t st, int n, DotArgs args, double *partial, double *out);
// y += a*x ; y2 += a2*x2 (y2 may be null)
void launch_axpy2(hipStream_t st, int n, double a, const double *x, double *y, double a2,
                  const double *x2, double *y2);
// y = x + b*y
void launch_xpay(hipStream_t st, int n, double b, const double *x, double *y);
// PCG steps with the scalars in device memory (iteration k >= 1):
// u += alpha p, r -= alpha ap with alpha = rz[k-1] / pap[k]
void launch_pcg_update(hipStream_t st, int n, const double *rz, const double *pap, int k,
                       const double *p, const double *ap, double *u, double *r);
// p = z + beta p with beta = rz[k] / rz[k-1]
void launch_pcg_dir(hipStream_t st, int n, const double *rz, int k, const double *z, double *p);
// forms with the iteration index k read from device memory (*kp), for a
// PCG iteration captured once as a graph: base[*kp] = a.b (and host[*kp],
// pinned host memory, when host is not null); the two updates at *kp; the
// counter set and stepped by one
void launch_dots_at(hipStream_t st, int n, const double *a, const double *b, double *partial,
                    double *base, const int *kp, double *host);
void launch_pcg_update_at(hipStream_t st, int n, const double *rz, const double *pap, const int *kp,
                          const double *p, const double *ap, double *u, double *r);
void launch_pcg_dir_at(hipStream_t st, int n, const double *rz, const int *kp, const double *z,
                       double *p);
void launch_count_set(hipStream_t st, int *kp, int v);
void launch_count_inc(hipStream_t st, int *kp);
// Krylov recycling basis (columns of length n, leading dimension ld):
// out[j] = V(:, j) . x for j < m (deterministic two-stage sums; partial
// holds DOT_BLOCKS * m doubles)
void launch_multidot(hipStream_t st, int n, const double *V, int64_t ld, int m, const double *x,
                     double *partial, double *out);
// y += sign * V(:, 0:m) c, c in device memory
void launch_basis_update(hipStream_t st, int n, const double *V, int64_t ld, int m,
                         const double *c, double sign, double *y);
// V(:, j) = p / sqrt(*pap), W(:, j) = ap / sqrt(*pap)
void launch_basis_store(hipStream_t st, int n, const double *pap, const double *p,
                        const double *ap, double *vcol, double *wcol);
// y = x + b*(y - w*v)   (BiCGSTAB direction update)
void launch_bicg_p(hipStream_t st, int n, double b, double w, const double *r, const double *v,
                   double *p);
// z = x - a*y
void launch_sub_scaled(hipStream_t st, int n, const double *x, double a, const double *y,
                       double *z);
void launch_mask(hipStream_t st, int n, const char *mask, double *v);
// grad[eid[e]] = -sum_ij lam[dof_ei] A_e(i,j) u[dof_ej], A_e = ae (+ e*nke^2
// when per_element); dofs of element e at edof[e*nke ...]
void launch_element_gradient(hipStream_t st, int nel, int nke, const int *edof, const int *eid,
                             const double *ae, bool per_element, const double *u,
                             const double *lam, double *grad);
// dst = src over n double2 elements (HBM copy reference)
void launch_copy_calibration(hipStream_t st, const double2 *src, double2 *dst, int64_t n);

}  // namespace bddc
