// kernels_misc.hip -- interface preconditioner assembly (C^T blocks, Z, G,
// Phi, coarse element matrices), triangular sweeps of the partial factor,
// and the per-iteration kernels of the Krylov methods (tile GEMV, the local
// preconditioner apply, weighted gather/scatter of the interface exchange,
// coarse gather, dot products).
#include "kernels.hpp"

namespace bddc {

// ------------------------------------------------- ordering class checks
// The uploaded triplet pattern (ai, aj) of subdomain i against that of its
// class candidate cand[i] (same length, checked by the caller): diff[i] = 1
// on any difference.  All writers of diff[i] write 1.
__global__ void k_same_pattern(const uint64_t *__restrict__ pai, const uint64_t *__restrict__ paj,
                               const int *__restrict__ cand, const int64_t *__restrict__ cnt,
                               int *__restrict__ diff) {
  const int i = blockIdx.y, c = cand[i];
  if (c == i) return;
  const int64_t n = cnt[i];
  const int *a0 = reinterpret_cast<const int *>(pai[i]), *a1 = reinterpret_cast<const int *>(pai[c]);
  const int *b0 = reinterpret_cast<const int *>(paj[i]), *b1 = reinterpret_cast<const int *>(paj[c]);
  int bad = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    bad |= (a0[k] != a1[k]) | (b0[k] != b1[k]);
  if (bad) diff[i] = 1;
}
void launch_same_pattern(hipStream_t st, const uint64_t *pai, const uint64_t *paj, const int *cand,
                         const int64_t *cnt, int *diff, int nsub) {
  if (nsub <= 0) return;
  hipLaunchKernelGGL(k_same_pattern, dim3(16, nsub), dim3(256), 0, st, pai, paj, cand, cnt, diff);
}

// --------------------------------------------------------- S' = S + rho C^T C
__device__ __forceinline__ void store2(float *p, double x, double y) {
  *reinterpret_cast<float2 *>(p) = make_float2((float)x, (float)y);
}
__device__ __forceinline__ void store2(double *p, double x, double y) {
  *reinterpret_cast<double2 *>(p) = make_double2(x, y);
}
__device__ __forceinline__ double2 load2(const double *p) {
  return *reinterpret_cast<const double2 *>(p);
}
__device__ __forceinline__ double2 load2(const float *p) {
  const float2 f = *reinterpret_cast<const float2 *>(p);
  return make_double2((double)f.x, (double)f.y);
}

// One job per lower tile (I, J): thread t handles rows 2(t%32), 2(t%32)+1
// of the columns t/32 + 8k (16-byte loads); the constraint ids of the rows
// are loop invariants, those of the columns are staged in LDS.  P: element
// type of S (FP64, or FP32 with the FP32 interior factorization), T: of S'.
template <typename P, typename T>
__global__ __launch_bounds__(256) void k_sprime(const P *__restrict__ pool,
                                                T *__restrict__ ppool,
                                                const SubDesc *__restrict__ subs,
                                                const int *__restrict__ jsub,
                                                const int *__restrict__ jij,
                                                const int *__restrict__ jdst,
                                                const int *__restrict__ cid,
                                                const int *__restrict__ csize) {
  __shared__ int ccol[64];
  const int job = blockIdx.x;
  const SubDesc d = subs[jsub[job]];
  const int I = jij[job] >> 16, J = jij[job] & 0xffff;
  // jobs cover I >= J: tile (I, J) of the packed lower S
  const P *src = pool + (size_t)(d.slow + J * d.nti - J * (J - 1) / 2 + (I - J)) * TILE_ELEMS;
  T *dst = ppool + (size_t)jdst[job] * TILE_ELEMS;
  const int *ci = cid + d.loff;
  const int tid = threadIdx.x;
  const int r0 = 2 * (tid & 31), c0 = tid >> 5;
  // the tile's loads go out before the constraint ids are staged
  double2 v[8];
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = load2(src + (c0 + 8 * k) * 64 + r0);
  const int a0 = ci[I * 64 + r0], a1 = ci[I * 64 + r0 + 1];
  if (tid < 64) ccol[tid] = ci[J * 64 + tid];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int c = c0 + 8 * k;
    double x = v[k].x, y = v[k].y;
    const int b = ccol[c];
    if (b >= 0 && (a0 == b || a1 == b)) {
      const double add = d.rho / (double)csize[b];
      if (a0 == b) x += add;
      if (a1 == b) y += add;
    }
    store2(dst + c * 64 + r0, x, y);
  }
}

template <typename P, typename T>
void launch_sprime(hipStream_t st, const P *pool, T *ppool, const SubDesc *subs,
                   const int *jobs_sub, const int *jobs_ij, const int *jobs_dst, int njobs,
                   const int *cid, const int *csize) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL((k_sprime<P, T>), dim3(njobs), dim3(256), 0, st, pool, ppool, subs, jobs_sub,
                     jobs_ij, jobs_dst, cid, csize);
}
template void launch_sprime<double, double>(hipStream_t, const double *, double *, const SubDesc *,
                                            const int *, const int *, const int *, int,
                                            const int *, const int *);
template void launch_sprime<double, float>(hipStream_t, const double *, float *, const SubDesc *,
                                           const int *, const int *, const int *, int, const int *,
                                           const int *);
template void launch_sprime<float, float>(hipStream_t, const float *, float *, const SubDesc *,
                                          const int *, const int *, const int *, int, const int *,
                                          const int *);
template void launch_sprime<float, double>(hipStream_t, const float *, double *, const SubDesc *,
                                           const int *, const int *, const int *, int, const int *,
                                           const int *);

// General constraint rows (adaptive constraints: several dense rows on one
// face): every interface dof carries the list of (constraint, value) of the
// rows it appears in, sorted by constraint; S'(a, b) = S(a, b) + sum over
// the constraints k of both lists of omega_k C_k(a) C_k(b).  Same jobs and
// thread layout as k_sprime.
__device__ __forceinline__ double shared_rows(const int *__restrict__ dptr,
                                              const int *__restrict__ dcon,
                                              const double *__restrict__ dval,
                                              const double *__restrict__ omega, int64_t a,
                                              int64_t b) {
  int p = dptr[a], pe = dptr[a + 1], q = dptr[b], qe = dptr[b + 1];
  double s = 0.0;
  while (p < pe && q < qe) {
    const int ka = dcon[p], kb = dcon[q];
    if (ka < kb) p++;
    else if (kb < ka) q++;
    else {
      s += omega[ka] * dval[p] * dval[q];
      p++, q++;
    }
  }
  return s;
}
template <typename P, typename T>
__global__ __launch_bounds__(256) void k_sprime_gen(const P *__restrict__ pool,
                                                    T *__restrict__ ppool,
                                                    const SubDesc *__restrict__ subs,
                                                    const int *__restrict__ jsub,
                                                    const int *__restrict__ jij,
                                                    const int *__restrict__ jdst,
                                                    const int *__restrict__ dptr,
                                                    const int *__restrict__ dcon,
                                                    const double *__restrict__ dval,
                                                    const double *__restrict__ omega) {
  const int job = blockIdx.x;
  const SubDesc d = subs[jsub[job]];
  const int I = jij[job] >> 16, J = jij[job] & 0xffff;
  const P *src = pool + (size_t)(d.slow + J * d.nti - J * (J - 1) / 2 + (I - J)) * TILE_ELEMS;
  T *dst = ppool + (size_t)jdst[job] * TILE_ELEMS;
  const int tid = threadIdx.x;
  const int r0 = 2 * (tid & 31), c0 = tid >> 5;
  const int64_t a0 = d.loff + I * 64 + r0;
#pragma unroll 2
  for (int k = 0; k < 8; k++) {
    const int c = c0 + 8 * k;
    const double2 v = load2(src + c * 64 + r0);
    const int64_t b = d.loff + J * 64 + c;
    const double x = v.x + shared_rows(dptr, dcon, dval, omega, a0, b);
    const double y = v.y + shared_rows(dptr, dcon, dval, omega, a0 + 1, b);
    store2(dst + c * 64 + r0, x, y);
  }
}

template <typename P, typename T>
void launch_sprime_gen(hipStream_t st, const P *pool, T *ppool, const SubDesc *subs,
                       const int *jobs_sub, const int *jobs_ij, const int *jobs_dst, int njobs,
                       const int *dptr, const int *dcon, const double *dval, const double *omega) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL((k_sprime_gen<P, T>), dim3(njobs), dim3(256), 0, st, pool, ppool, subs,
                     jobs_sub, jobs_ij, jobs_dst, dptr, dcon, dval, omega);
}
template void launch_sprime_gen<double, double>(hipStream_t, const double *, double *,
                                                const SubDesc *, const int *, const int *,
                                                const int *, int, const int *, const int *,
                                                const double *, const double *);
template void launch_sprime_gen<double, float>(hipStream_t, const double *, float *,
                                               const SubDesc *, const int *, const int *,
                                               const int *, int, const int *, const int *,
                                               const double *, const double *);
template void launch_sprime_gen<float, float>(hipStream_t, const float *, float *, const SubDesc *,
                                              const int *, const int *, const int *, int,
                                              const int *, const int *, const double *,
                                              const double *);
template void launch_sprime_gen<float, double>(hipStream_t, const float *, double *,
                                               const SubDesc *, const int *, const int *,
                                               const int *, int, const int *, const int *,
                                               const double *, const double *);

// C^T tiles for general rows: B(i, c) = value of dof i in constraint c
template <typename T>
__global__ void k_ct_init_gen(T *__restrict__ pool, const SubDesc *__restrict__ subs,
                              const int *__restrict__ dptr, const int *__restrict__ dcon,
                              const double *__restrict__ dval) {
  const SubDesc d = subs[blockIdx.y];
  const int I = blockIdx.x;
  if (I >= d.nti) return;
  for (int cT = 0; cT < d.ncT; cT++) {
    T *Y = pool + (size_t)(d.ybase + I * d.ncT + cT) * TILE_ELEMS;
    for (int e = threadIdx.x; e < TILE_ELEMS; e += blockDim.x) {
      const int r = e & 63, c = cT * 64 + (e >> 6);
      double v = 0.0;
      if (c < d.nc) {
        const int64_t a = d.loff + I * 64 + r;
        for (int q = dptr[a]; q < dptr[a + 1]; q++)
          if (dcon[q] == d.coff + c) v = dval[q];
      }
      Y[e] = (T)v;
    }
  }
}
template <typename T>
void launch_ct_init_gen(hipStream_t st, T *pool, const SubDesc *subs, int nsub, const int *dptr,
                        const int *dcon, const double *dval, int maxnti) {
  if (nsub <= 0 || maxnti <= 0) return;
  hipLaunchKernelGGL(k_ct_init_gen<T>, dim3(maxnti, nsub), dim3(256), 0, st, pool, subs, dptr, dcon,
                     dval);
}
template void launch_ct_init_gen<double>(hipStream_t, double *, const SubDesc *, int, const int *,
                                         const int *, const double *, int);
template void launch_ct_init_gen<float>(hipStream_t, float *, const SubDesc *, int, const int *,
                                        const int *, const double *, int);

// --------------------------------------------- C^T blocks and Z = S'^{-1} C^T
template <typename T>
__global__ void k_ct_init(T *__restrict__ pool, const SubDesc *__restrict__ subs,
                          const int *__restrict__ cid) {
  const SubDesc d = subs[blockIdx.y];
  const int I = blockIdx.x;
  if (I >= d.nti) return;
  const int *ci = cid + d.loff + I * 64;
  for (int cT = 0; cT < d.ncT; cT++) {
    T *Y = pool + (size_t)(d.ybase + I * d.ncT + cT) * TILE_ELEMS;
    for (int e = threadIdx.x; e < TILE_ELEMS; e += blockDim.x) {
      const int r = e & 63, c = cT * 64 + (e >> 6);
      Y[e] = (c < d.nc && ci[r] == d.coff + c) ? (T)1 : (T)0;
    }
  }
}

template <typename T>
__global__ void k_y_to_z(const T *__restrict__ pool, const SubDesc *__restrict__ subs,
                         double *__restrict__ Z, int which) {
  const SubDesc d = subs[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.NI) return;
  const int b0 = which == 0 ? d.ybase : d.phibase;
  const T *Y = pool + (size_t)(b0 + (i >> 6) * d.ncT) * TILE_ELEMS + (i & 63);
  for (int c = 0; c < d.nc; c++)
    Z[d.zoff + (size_t)c * d.NI + i] = (double)Y[(size_t)(c >> 6) * TILE_ELEMS + (c & 63) * 64];
}

// G = C Z over the constraint rows (values 1 when cval is null)
__global__ void k_g(const SubDesc *__restrict__ subs, const int *__restrict__ cptr,
                    const int *__restrict__ cdof, const double *__restrict__ cval,
                    const double *__restrict__ Z, double *__restrict__ G) {
  const SubDesc d = subs[blockIdx.x];
  const int nc = d.nc;
  for (int e = threadIdx.x; e < nc * nc; e += blockDim.x) {
    int c = e / nc, k = e % nc;
    double s = 0.0;
    for (int q = cptr[d.coff + c]; q < cptr[d.coff + c + 1]; q++)
      s += (cval ? cval[q] : 1.0) * Z[d.zoff + (size_t)k * d.NI + cdof[q]];
    G[d.goff + (size_t)k * nc + c] = s;
  }
}

// in-place Gauss-Jordan inverse of the SPD matrices G (no pivoting)
__global__ void k_ginv(const SubDesc *__restrict__ subs, double *__restrict__ G,
                       int *__restrict__ err) {
  __shared__ double F[1024];
  __shared__ double inv;
  const SubDesc d = subs[blockIdx.x];
  const int n = d.nc;
  double *A = G + d.goff;  // A[i*n + j]
  for (int k = 0; k < n; k++) {
    __syncthreads();
    if (threadIdx.x == 0) {
      double p = A[k * n + k];
      if (!(p > 0.0)) {
        atomicOr(err, 2);
        p = 1.0;
      }
      inv = 1.0 / p;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      double v = (j == k) ? 1.0 : A[k * n + j];
      A[k * n + j] = v * inv;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) F[i] = (i == k) ? 0.0 : A[i * n + k];
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      int i = e / n, j = e % n;
      if (i == k) continue;
      double v = (j == k) ? 0.0 : A[i * n + j];
      A[i * n + j] = v - F[i] * A[k * n + j];
    }
  }
}

// Ac_i = G^{-1} - Omega; Omega_c = rho / |D_c| for unit rows, or omega[c]
__global__ void k_acs(const SubDesc *__restrict__ subs, const int *__restrict__ cptr,
                      const double *__restrict__ omega, const double *__restrict__ G,
                      double *__restrict__ Acs) {
  const SubDesc d = subs[blockIdx.x];
  const int nc = d.nc;
  for (int e = threadIdx.x; e < nc * nc; e += blockDim.x) {
    int c = e / nc, k = e % nc;
    double v = G[d.goff + e];
    if (c == k)
      v -= omega ? omega[d.coff + c] : d.rho / (double)(cptr[d.coff + c + 1] - cptr[d.coff + c]);
    Acs[d.goff + e] = v;
  }
}

__global__ void k_coarse_assemble(double *__restrict__ pool, const int64_t *__restrict__ dst,
                                  const int *__restrict__ ptr, const int64_t *__restrict__ src,
                                  int n, const double *__restrict__ Acs) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s = 0.0;
  for (int q = ptr[k]; q < ptr[k + 1]; q++) s += Acs[src[q]];
  pool[dst[k]] = s;
}

template <typename T>
void launch_ct_init(hipStream_t st, T *pool, const SubDesc *subs, int nsub, const int *cid,
                    int maxnti) {
  if (nsub <= 0 || maxnti <= 0) return;
  hipLaunchKernelGGL(k_ct_init<T>, dim3(maxnti, nsub), dim3(256), 0, st, pool, subs, cid);
}
template <typename T>
void launch_y_to_z(hipStream_t st, const T *pool, const SubDesc *subs, int nsub, double *Z,
                   int maxNI, int which) {
  if (nsub <= 0 || maxNI <= 0) return;
  hipLaunchKernelGGL(k_y_to_z<T>, dim3((maxNI + 255) / 256, nsub), dim3(256), 0, st, pool, subs, Z,
                     which);
}

// G^{-1} (nc x nc) into the ncT x ncT tile block at gtbase, zero padded
template <typename T>
__global__ void k_ginv_to_tiles(T *__restrict__ pool, const SubDesc *__restrict__ subs,
                                const double *__restrict__ G) {
  const SubDesc d = subs[blockIdx.y];
  const int t = blockIdx.x;
  if (t >= d.ncT * d.ncT) return;
  const int KT = t / d.ncT, CT = t % d.ncT;
  T *dst = pool + (size_t)(d.gtbase + t) * TILE_ELEMS;
  for (int e = threadIdx.x; e < TILE_ELEMS; e += blockDim.x) {
    const int k = KT * 64 + (e & 63), c = CT * 64 + (e >> 6);
    dst[e] = (k < d.nc && c < d.nc) ? (T)G[d.goff + (size_t)c * d.nc + k] : (T)0;
  }
}
template <typename T>
void launch_ginv_to_tiles(hipStream_t st, T *pool, const SubDesc *subs, int nsub, const double *G,
                          int maxncT) {
  if (nsub <= 0 || maxncT <= 0) return;
  hipLaunchKernelGGL(k_ginv_to_tiles<T>, dim3(maxncT * maxncT, nsub), dim3(256), 0, st, pool, subs,
                     G);
}
template void launch_ginv_to_tiles<double>(hipStream_t, double *, const SubDesc *, int,
                                           const double *, int);
template void launch_ginv_to_tiles<float>(hipStream_t, float *, const SubDesc *, int, const double *,
                                          int);
template void launch_ct_init<double>(hipStream_t, double *, const SubDesc *, int, const int *, int);
template void launch_ct_init<float>(hipStream_t, float *, const SubDesc *, int, const int *, int);
template void launch_y_to_z<double>(hipStream_t, const double *, const SubDesc *, int, double *, int,
                                    int);
template void launch_y_to_z<float>(hipStream_t, const float *, const SubDesc *, int, double *, int,
                                   int);
void launch_g(hipStream_t st, const SubDesc *subs, int nsub, const int *cptr, const int *cdof,
              const double *cval, const double *Z, double *G) {
  if (nsub <= 0) return;
  hipLaunchKernelGGL(k_g, dim3(nsub), dim3(256), 0, st, subs, cptr, cdof, cval, Z, G);
}
void launch_ginv(hipStream_t st, const SubDesc *subs, int nsub, double *G, double *work, int *err) {
  (void)work;
  if (nsub <= 0) return;
  hipLaunchKernelGGL(k_ginv, dim3(nsub), dim3(256), 0, st, subs, G, err);
}
void launch_acs(hipStream_t st, const SubDesc *subs, int nsub, const int *cptr,
                const double *omega, const double *G, double *Acs) {
  if (nsub <= 0) return;
  hipLaunchKernelGGL(k_acs, dim3(nsub), dim3(256), 0, st, subs, cptr, omega, G, Acs);
}
void launch_coarse_assemble(hipStream_t st, double *pool, const int64_t *dst, const int *ptr,
                            const int64_t *src, int n, const double *Acs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_coarse_assemble, dim3((n + 255) / 256), dim3(256), 0, st, pool, dst, ptr,
                     src, n, Acs);
}

// ------------------------------------------------------- triangular sweeps
// One workgroup per subdomain keeps the whole device-ordered vector in LDS.
// A tile row accumulates all its tile products per thread and reduces once.
// Forward:  v_I = Linv_I (b_I - sum_k L_Ik v_k)      for the nto factored rows,
//           v_I = b_I - sum_k L_Ik v_k               for the nti trailing rows
// Backward: v_I = Linv_I^T (y_I - sum_k L_kI^T v_k)  for the nto factored rows
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// NW waves (64 lanes each) per workgroup: 4 when the subdomains fill the
// chip, 16 when there are few of them (a second level) so that one
// workgroup keeps more loads in flight.
template <int NW>
__device__ __forceinline__ double row_reduce(double s, double *red) {
  const int g = threadIdx.x >> 6, r = threadIdx.x & 63;
  red[g * 64 + r] = s;
  __syncthreads();
  double v = 0.0;
  if (threadIdx.x < 64) {
#pragma unroll
    for (int q = 0; q < NW; q += 4)
      v += (red[q * 64 + r] + red[(q + 1) * 64 + r]) + (red[(q + 2) * 64 + r] + red[(q + 3) * 64 + r]);
  }
  return v;
}

// GV: the subdomain vector stays in global memory (L2-cached) instead of
// LDS, for subdomains whose vector does not fit there; rows written by one
// wave are published to the others with a workgroup fence before the barrier.
// P: element type of the factor tiles (FP64, or FP32 with the FP32 interior
// factorization); the vector and the sums stay FP64.
template <typename P, int NW, bool GV>
__global__ __launch_bounds__(64 * NW) void k_trsv_fwd(const P *__restrict__ pool,
                                                  const TrsvDesc *__restrict__ ds,
                                                  const int *__restrict__ rptr,
                                                  const int *__restrict__ rk,
                                                  const int *__restrict__ rt,
                                                  double *vec) {
  extern __shared__ double smem[];
  const TrsvDesc d = ds[blockIdx.x];
  const int nt = d.nto + d.nti;
  double *gv = vec + d.voff;
  double *v = GV ? gv : smem;                    // nt*64
  double *red = GV ? smem : smem + nt * 64;      // NW*64
  double *t = red + NW * 64;                     // 64
  const int g = threadIdx.x >> 6, r = threadIdx.x & 63;
  if (!GV)
    for (int k = threadIdx.x; k < nt * 64; k += blockDim.x) v[k] = gv[k];
  __syncthreads();
  for (int I = 0; I < nt; I++) {
    double s = 0.0;
    for (int q = rptr[d.rptr_off + I]; q < rptr[d.rptr_off + I + 1]; q++) {
      const P *T = pool + (size_t)rt[q] * TILE_ELEMS + r;
      const double *x = v + rk[q] * 64;
#pragma unroll 4
      for (int c = g; c < 64; c += NW) s += T[c * 64] * x[c];
    }
    double tv = row_reduce<NW>(s, red);
    if (threadIdx.x < 64) t[r] = v[I * 64 + r] - tv;
    __syncthreads();
    if (I < d.nto) {
      const P *L = pool + (size_t)(d.linv + I) * TILE_ELEMS + r;
      double u = 0.0;
#pragma unroll 4
      for (int c = g; c < 64; c += NW) u += L[c * 64] * t[c];
      double w = row_reduce<NW>(u, red);
      if (threadIdx.x < 64) v[I * 64 + r] = w;
    } else if (threadIdx.x < 64) {
      v[I * 64 + r] = t[r];
    }
    if (GV) __threadfence_block();
    __syncthreads();
  }
  if (!GV)
    for (int k = threadIdx.x; k < nt * 64; k += blockDim.x) gv[k] = v[k];
}

template <typename P, int NW, bool GV>
__global__ __launch_bounds__(64 * NW) void k_trsv_bwd(const P *__restrict__ pool,
                                                  const TrsvDesc *__restrict__ ds,
                                                  const int *__restrict__ cptr,
                                                  const int *__restrict__ ck,
                                                  const int *__restrict__ ct,
                                                  double *vec) {
  constexpr int CW = 64 / NW;  // tile columns per wave
  extern __shared__ double smem[];
  const TrsvDesc d = ds[blockIdx.x];
  const int nt = d.nto + d.nti;
  double *gv = vec + d.voff;
  double *v = GV ? gv : smem;
  double *red = GV ? smem : smem + nt * 64;
  double *t = red + NW * 64;
  const int g = threadIdx.x >> 6, r = threadIdx.x & 63;
  if (!GV)
    for (int k = threadIdx.x; k < nt * 64; k += blockDim.x) v[k] = gv[k];
  __syncthreads();
  if constexpr (NW == 16) {
    // lane = row c of the contiguous column: wave g produces the outputs
    // g + 16 m (m < 4) by wave reductions, every load a 512-byte column
    const int l = r;
    for (int I = d.nto - 1; I >= 0; I--) {
      double a[4] = {0.0, 0.0, 0.0, 0.0};
      for (int q = cptr[d.cptr_off + I]; q < cptr[d.cptr_off + I + 1]; q++) {
        const P *T = pool + (size_t)ct[q] * TILE_ELEMS + l;
        const double xv = v[ck[q] * 64 + l];
#pragma unroll
        for (int m = 0; m < 4; m++) a[m] += T[(g + 16 * m) * 64] * xv;
      }
#pragma unroll
      for (int m = 0; m < 4; m++) a[m] = wave_sum(a[m]);
      if (l == 0) {
#pragma unroll
        for (int m = 0; m < 4; m++) t[g + 16 * m] = v[I * 64 + g + 16 * m] - a[m];
      }
      __syncthreads();
      const P *L = pool + (size_t)(d.linv + I) * TILE_ELEMS + l;
      const double tv = t[l];
#pragma unroll
      for (int m = 0; m < 4; m++) a[m] = wave_sum(L[(g + 16 * m) * 64] * tv);
      if (l == 0) {
#pragma unroll
        for (int m = 0; m < 4; m++) v[I * 64 + g + 16 * m] = a[m];
      }
      if (GV) __threadfence_block();
      __syncthreads();
    }
  } else {
  for (int I = d.nto - 1; I >= 0; I--) {
    double s = 0.0;
    for (int q = cptr[d.cptr_off + I]; q < cptr[d.cptr_off + I + 1]; q++) {
      // (L_kI^T v_k)[r] = sum_c L_kI(c, r) v_k[c], column r of the tile is contiguous
      const P *T = pool + (size_t)ct[q] * TILE_ELEMS + r * 64 + g * CW;
      const double *x = v + ck[q] * 64 + g * CW;
#pragma unroll
      for (int c = 0; c < CW; c++) s += T[c] * x[c];
    }
    double tv = row_reduce<NW>(s, red);
    if (threadIdx.x < 64) t[r] = v[I * 64 + r] - tv;
    __syncthreads();
    const P *L = pool + (size_t)(d.linv + I) * TILE_ELEMS + r * 64 + g * CW;
    double u = 0.0;
#pragma unroll
    for (int c = 0; c < CW; c++) u += L[c] * t[g * CW + c];
    double w = row_reduce<NW>(u, red);
    if (threadIdx.x < 64) v[I * 64 + r] = w;
    if (GV) __threadfence_block();
    __syncthreads();
  }
  }
  if (!GV)
    for (int k = threadIdx.x; k < d.nto * 64; k += blockDim.x) gv[k] = v[k];
}

size_t max_dynamic_lds() {
  static size_t cached = 0;
  if (!cached) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    cached = (size_t)v;
  }
  return cached;
}

// dynamic LDS request of a launch, checked against the device limit
static void lds_request(const void *kernel, size_t bytes, const char *what) {
  if (bytes > max_dynamic_lds()) fatal("BDDCML_DEVICE_PLAN", std::string(what) + ": subdomain too large for LDS");
  if (bytes > 65536) {
    if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      fatal("BDDCML_DEVICE_PLAN", std::string(what) + ": cannot raise the LDS limit");
  }
}

// 16 waves per workgroup below four subdomains per CU: a sweep is a chain
// of dependent tile rows, and more waves keep more of each row's loads in
// flight (512 subdomains of 16^3: forward sweep 2.1 -> 1.3 ms); the
// subdomain vector in LDS when it fits, else in global memory
static constexpr int kTrsvFewSubs = 1024;
// force the global-memory forms of the per-subdomain kernels (testing)
static bool g_trsv_global = false;
void set_trsv_global(bool on) { g_trsv_global = on; }
// forms chosen by the number of subdomains: 0 by that number, 1 always the
// many-subdomain forms, 2 always the few-subdomain forms (testing)
static int g_launch_forms = 0;
void set_launch_forms(int forms) { g_launch_forms = forms; }
int launch_forms() { return g_launch_forms; }
// waves per workgroup of the interior sweeps: 0 by the rule above, else 4 or 16
static int g_sweep_waves = 0;
void set_sweep_waves(int waves) { g_sweep_waves = waves; }
static bool few_subdomains(int nsub) {
  if (g_sweep_waves) return g_sweep_waves == 16;
  return g_launch_forms == 0 ? nsub < kTrsvFewSubs : g_launch_forms == 2;
}
template <int NW, typename P, class K1, class K2>
static void trsv_launch(hipStream_t st, int nsub, int maxnt, K1 lds_kernel, K2 gv_kernel,
                        const P *pool, const TrsvDesc *d, const int *p, const int *k,
                        const int *t, double *v) {
  const size_t sh = (size_t)(maxnt * 64 + NW * 64 + 64) * sizeof(double);
  if (sh <= max_dynamic_lds() && !g_trsv_global) {
    lds_request((const void *)lds_kernel, sh, "triangular sweep");
    hipLaunchKernelGGL(lds_kernel, dim3(nsub), dim3(64 * NW), sh, st, pool, d, p, k, t, v);
  } else {
    const size_t sg = (size_t)(NW * 64 + 64) * sizeof(double);
    hipLaunchKernelGGL(gv_kernel, dim3(nsub), dim3(64 * NW), sg, st, pool, d, p, k, t, v);
  }
}
template <typename P>
void launch_trsv_fwd(hipStream_t st, const P *pool, const TrsvDesc *d, int nsub, const int *rptr,
                     const int *rk, const int *rt, double *v, int maxnt) {
  if (nsub <= 0) return;
  if (few_subdomains(nsub))
    trsv_launch<16>(st, nsub, maxnt, k_trsv_fwd<P, 16, false>, k_trsv_fwd<P, 16, true>, pool, d,
                    rptr, rk, rt, v);
  else
    trsv_launch<4>(st, nsub, maxnt, k_trsv_fwd<P, 4, false>, k_trsv_fwd<P, 4, true>, pool, d, rptr,
                   rk, rt, v);
}
template <typename P>
void launch_trsv_bwd(hipStream_t st, const P *pool, const TrsvDesc *d, int nsub, const int *cptr,
                     const int *ck, const int *ct, double *v, int maxnt) {
  if (nsub <= 0) return;
  if (few_subdomains(nsub))
    trsv_launch<16>(st, nsub, maxnt, k_trsv_bwd<P, 16, false>, k_trsv_bwd<P, 16, true>, pool, d,
                    cptr, ck, ct, v);
  else
    trsv_launch<4>(st, nsub, maxnt, k_trsv_bwd<P, 4, false>, k_trsv_bwd<P, 4, true>, pool, d, cptr,
                   ck, ct, v);
}
template void launch_trsv_fwd<double>(hipStream_t, const double *, const TrsvDesc *, int,
                                      const int *, const int *, const int *, double *, int);
template void launch_trsv_fwd<float>(hipStream_t, const float *, const TrsvDesc *, int, const int *,
                                     const int *, const int *, double *, int);
template void launch_trsv_bwd<double>(hipStream_t, const double *, const TrsvDesc *, int,
                                      const int *, const int *, const int *, double *, int);
template void launch_trsv_bwd<float>(hipStream_t, const float *, const TrsvDesc *, int, const int *,
                                     const int *, const int *, double *, int);

// -------------------------------------------------------- iteration kernels
// y = S x for symmetric S held as its lower tiles.  Workgroup = tile column
// J of one subdomain; wave w reads columns 16w..16w+15 of each tile (lane =
// row, one coalesced 512-byte column per load), the next tile prefetched.
// Slot (I, J), I > J, receives S_IJ x_J; slot (J, J) receives S_JJ x_J plus
// sum_{I>J} S_IJ^T x_I (lane products summed over rows by butterflies).
// T: element type of the tiles (double: S; float or double: an explicit
// S'^{-1} in the factor space).  wts (may be null): per local entry weight
// applied to the gathered input.
template <typename T>
__global__ __launch_bounds__(256) void k_symv(const T *__restrict__ pool,
                                              const SymvCol *__restrict__ cols,
                                              const int64_t *__restrict__ loff,
                                              const int *__restrict__ map,
                                              const double *__restrict__ wts,
                                              const double *__restrict__ xin,
                                              double *__restrict__ part) {
  __shared__ double xj[64];
  __shared__ double red[2][4][64];
  __shared__ double dg[64];
  const SymvCol col = cols[blockIdx.x];
  const int64_t lo = loff[col.mat];
  const int tid = threadIdx.x, w = tid >> 6, r = tid & 63;
  const int *mp = map + lo;
  if (tid < 64) {
    const int g = mp[col.J * 64 + tid];
    xj[tid] = g >= 0 ? (wts ? wts[lo + col.J * 64 + tid] * xin[g] : xin[g]) : 0.0;
  }
  __syncthreads();
  double xc[16], acc[16], cur[16], nxt[16];
#pragma unroll
  for (int c = 0; c < 16; c++) xc[c] = xj[16 * w + c], acc[c] = 0.0;
  const T *T0 = pool + (size_t)col.tile0 * TILE_ELEMS + (16 * w) * 64 + r;
#pragma unroll
  for (int c = 0; c < 16; c++) cur[c] = (double)T0[c * 64];
  double xi = 0.0;  // x_I[r] of the current tile (I > J)
  for (int k = 0; k < col.nt; k++) {
    double xn = 0.0;
    if (k + 1 < col.nt) {
      const T *Tn = T0 + (size_t)(k + 1) * TILE_ELEMS;
#pragma unroll
      for (int c = 0; c < 16; c++) nxt[c] = (double)Tn[c * 64];
      const int g = mp[(col.J + k + 1) * 64 + r];
      xn = g >= 0 ? (wts ? wts[lo + (col.J + k + 1) * 64 + r] * xin[g] : xin[g]) : 0.0;
    }
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int c = 0; c < 16; c += 2) s0 += cur[c] * xc[c], s1 += cur[c + 1] * xc[c + 1];
    if (k > 0) {
#pragma unroll
      for (int c = 0; c < 16; c++) acc[c] += cur[c] * xi;
    }
    red[k & 1][w][r] = s0 + s1;
    __syncthreads();
    if (w == 0) {
      const double v = (red[k & 1][0][r] + red[k & 1][1][r]) + (red[k & 1][2][r] + red[k & 1][3][r]);
      if (k == 0)
        dg[r] = v;
      else
        part[(size_t)(col.pslot + k) * 64 + r] = v;
    }
#pragma unroll
    for (int c = 0; c < 16; c++) cur[c] = nxt[c];
    xi = xn;
  }
  double mine = 0.0;
#pragma unroll
  for (int c = 0; c < 16; c++) {
    double v = acc[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (r == c) mine = v;
  }
  __syncthreads();
  if (r < 16) part[(size_t)col.pslot * 64 + 16 * w + r] = dg[16 * w + r] + mine;
}

// y_local[loff + I*64 + r] = sum_{J <= I} slot(I, J)[r], J ascending; one
// wave per row
__global__ __launch_bounds__(256) void k_symv_reduce(const SymvRow *__restrict__ rows, int nrows,
                                                     const int64_t *__restrict__ loff,
                                                     const double *__restrict__ part,
                                                     double *__restrict__ yout) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), r = threadIdx.x & 63;
  if (row >= nrows) return;
  const SymvRow R = rows[row];
  double s = 0.0;
#pragma unroll 8
  for (int J = 0; J <= R.I; J++)
    s += part[(size_t)(R.psub + J * R.nti - J * (J - 1) / 2 + (R.I - J)) * 64 + r];
  yout[loff[R.mat] + R.I * 64 + r] = s;
}

template <typename T>
void launch_symv(hipStream_t st, const T *pool, const SymvCol *cols, int ncols,
                 const SymvRow *rows, int nrows, const int64_t *loff, const int *map,
                 const double *wts, const double *xin, double *part, double *yout) {
  if (ncols <= 0 || nrows <= 0) return;
  hipLaunchKernelGGL(k_symv<T>, dim3(ncols), dim3(256), 0, st, pool, cols, loff, map, wts, xin,
                     part);
  hipLaunchKernelGGL(k_symv_reduce, dim3((nrows + 3) / 4), dim3(256), 0, st, rows, nrows, loff,
                     part, yout);
}
template void launch_symv<double>(hipStream_t, const double *, const SymvCol *, int,
                                  const SymvRow *, int, const int64_t *, const int *,
                                  const double *, const double *, double *, double *);
template void launch_symv<float>(hipStream_t, const float *, const SymvCol *, int, const SymvRow *,
                                 int, const int64_t *, const int *, const double *, const double *,
                                 double *, double *);

// Block-sparse tile GEMV, one workgroup per output segment of 64:
// y[yoff..] = (b ? b[boff..] : 0) + sign * sum_p op(T_p) x[xoff_p..], op the
// tile (trans 0: lane = row, each wave 16 columns, 512-byte column loads) or
// its transpose (trans 1: wave w the output rows 16w..16w+15, lane = the
// summed index, 512-byte row loads); partial sums meet in LDS in a fixed
// order.  The explicit interior operators of the levels below the first.
__global__ __launch_bounds__(256) void k_tile_gemv(const double *__restrict__ pool,
                                                   const GemvTask *__restrict__ tasks,
                                                   const GemvPair *__restrict__ pairs,
                                                   const double *__restrict__ x,
                                                   const double *b, double *y, int trans,
                                                   double sign) {
  __shared__ double red[64][65];
  const GemvTask t = tasks[blockIdx.x];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (!trans) {
    double acc = 0.0;
    for (int q = t.p0; q < t.p1; q++) {
      const GemvPair p = pairs[q];
      const double *T = pool + (size_t)p.tile * TILE_ELEMS + (16 * w) * 64 + l;
      const double *xv = x + p.xoff + 16 * w;
#pragma unroll
      for (int c = 0; c < 16; c++) acc += T[c * 64] * xv[c];
    }
    red[w][l] = acc;
    __syncthreads();
    if (threadIdx.x < 64) {
      const double s = (red[0][l] + red[1][l]) + (red[2][l] + red[3][l]);
      const double base = b ? b[t.boff + l] : 0.0;
      y[t.yoff + l] = base + sign * s;
    }
  } else {
    double acc[16];
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0;
    for (int q = t.p0; q < t.p1; q++) {
      const GemvPair p = pairs[q];
      // (T^T x)[r] = sum_c T(c, r) x[c]; column r of the tile is contiguous
      const double *T = pool + (size_t)p.tile * TILE_ELEMS + (16 * w) * 64 + l;
      const double xl = x[p.xoff + l];
#pragma unroll
      for (int r = 0; r < 16; r++) acc[r] += T[r * 64] * xl;
    }
#pragma unroll
    for (int r = 0; r < 16; r++) red[16 * w + r][l] = acc[r];
    __syncthreads();
    if (threadIdx.x < 64) {
      double s = 0.0;
      for (int c = 0; c < 64; c++) s += red[l][c];
      const double base = b ? b[t.boff + l] : 0.0;
      y[t.yoff + l] = base + sign * s;
    }
  }
}
void launch_tile_gemv(hipStream_t st, const double *pool, const GemvTask *tasks, int ntasks,
                      const GemvPair *pairs, const double *x, const double *b, double *y, int trans,
                      double sign) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_tile_gemv, dim3(ntasks), dim3(256), 0, st, pool, tasks, pairs, x, b, y,
                     trans, sign);
}

// y = A x for a dense matrix of nt x nt tiles in full storage (tile (I, J)
// at base + I*nt + J), x read from global memory: one workgroup per tile
// writes the tile's product to partial slot (J, I); the reduce sums the
// slots of each row in a fixed order.  For the coarse inverse, whose tiles
// are few but large in total.
__global__ __launch_bounds__(256) void k_dense_gemv(const double *__restrict__ pool, int base,
                                                    int nt, const double *__restrict__ x,
                                                    double *__restrict__ part) {
  __shared__ double red[4][64];
  const int I = blockIdx.x / nt, J = blockIdx.x % nt;
  const int w = threadIdx.x >> 6, r = threadIdx.x & 63;
  const double *T = pool + (size_t)(base + I * nt + J) * TILE_ELEMS + r;
  const double *xj = x + J * 64;
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int c = 16 * w; c < 16 * w + 16; c += 2) {
    s0 += T[c * 64] * xj[c];
    s1 += T[(c + 1) * 64] * xj[c + 1];
  }
  red[w][r] = s0 + s1;
  __syncthreads();
  if (w == 0) part[((size_t)J * nt + I) * 64 + r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
}
__global__ void k_dense_gemv_reduce(int nt, const double *__restrict__ part, double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt * 64) return;
  // unrolled: the loads go out ahead of the (unchanged) chain of sums
  double s = 0.0;
#pragma unroll 8
  for (int J = 0; J < nt; J++) s += part[(size_t)J * nt * 64 + i];
  y[i] = s;
}
void launch_dense_gemv(hipStream_t st, const double *pool, int base, int nt, const double *x,
                       double *part, double *y) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_dense_gemv, dim3(nt * nt), dim3(256), 0, st, pool, base, nt, x, part);
  hipLaunchKernelGGL(k_dense_gemv_reduce, dim3((nt * 64 + 255) / 256), dim3(256), 0, st, nt, part, y);
}

__global__ void k_assemble(int n, const int *__restrict__ ptr, const int64_t *__restrict__ src,
                           const double *__restrict__ w, const double *__restrict__ yl,
                           double *__restrict__ y) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  double s = 0.0;
#pragma unroll 4
  for (int q = ptr[g]; q < ptr[g + 1]; q++) {
    int64_t k = src[q];
    s += w ? w[k] * yl[k] : yl[k];
  }
  y[g] = s;
}

__global__ void k_coarse_gather(int n, const int *__restrict__ ptr, const int *__restrict__ src,
                                const double *__restrict__ rcs, double *__restrict__ rc) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s = 0.0;
  for (int q = ptr[k]; q < ptr[k + 1]; q++) s += rcs[src[q]];
  rc[k] = s;
}

// ------------------------------------------------- local preconditioner apply
constexpr int KA_THR = 1024;  // 16 waves per subdomain

__device__ __forceinline__ int packed_lower(int n, int I, int J) {
  return J * n - J * (J - 1) / 2 + (I - J);
}


// rcs = Phi_i^T (w .* r(map)) per subdomain: the weighted interface vector is
// gathered into LDS once per block (STAGE; otherwise, for interfaces too
// large for LDS, per column from global memory), then each wave reduces
// PHIT_COLS / 4 of the block's PHIT_COLS columns of Phi_i (contiguous,
// lane-strided).
constexpr int PHIT_COLS = 16;
template <bool STAGE>
__global__ __launch_bounds__(256) void k_phit(const SubDesc *__restrict__ subs,
                                              const double *__restrict__ Phi,
                                              const int *__restrict__ map,
                                              const double *__restrict__ w,
                                              const double *__restrict__ x,
                                              double *__restrict__ rcs) {
  extern __shared__ double xw[];
  const SubDesc d = subs[blockIdx.y];
  const int c0 = blockIdx.x * PHIT_COLS;
  if (c0 >= d.nc) return;
  const int NI = d.NI;
  if (STAGE) {
    for (int i = threadIdx.x; i < NI; i += 256) {
      const int g = map[d.loff + i];
      xw[i] = g >= 0 ? w[d.loff + i] * x[g] : 0.0;
    }
    __syncthreads();
  }
  // the wave's (up to) four columns c0 + wave + 4k side by side, one
  // accumulator each (the order of each column's sum is the plain one):
  // four independent loads per step instead of one dependent chain
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int CPW = PHIT_COLS / 4;
  const int cend = min(c0 + PHIT_COLS, d.nc);
  const double *pc[CPW];
  double s[CPW];
#pragma unroll
  for (int k = 0; k < CPW; k++) {
    const int c = min(c0 + wave + 4 * k, cend - 1);  // a column past the end reads a valid one
    pc[k] = Phi + d.zoff + (size_t)c * NI;
    s[k] = 0.0;
  }
  for (int i = lane; i < NI; i += 64) {
    double xv;
    if (STAGE) {
      xv = xw[i];
    } else {
      const int g = map[d.loff + i];
      xv = g >= 0 ? w[d.loff + i] * x[g] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < CPW; k++) s[k] += pc[k][i] * xv;
  }
#pragma unroll
  for (int k = 0; k < CPW; k++) {
    const double t = wave_sum(s[k]);
    const int c = c0 + wave + 4 * k;
    if (lane == 0 && c < cend) rcs[d.coff + c] = t;
  }
}

// DEEP: four tiles of a row in flight and the inverse diagonal tile loaded
// ahead of the row sum (more loads per workgroup, 96 VGPRs: one workgroup
// per CU) -- for few or uneven subdomains, where the largest subdomain's
// sweep is the tail.  Otherwise two tiles in flight (56 VGPRs, two
// workgroups per CU), for many subdomains of similar size, where the
// launch is HBM-bound as a whole.
// barrier that also publishes global-memory writes to the other waves of
// the workgroup when the working vector lives there
template <bool GV>
__device__ __forceinline__ void publish_sync() {
  if (GV) __threadfence_block();
  __syncthreads();
}

template <typename T, bool DEEP, bool GV>
__global__ __launch_bounds__(KA_THR) void k_kapply(
    const T *__restrict__ pool, const SubDesc *__restrict__ subs, const int *__restrict__ map,
    const double *__restrict__ w, const double *__restrict__ rin, const double *__restrict__ Phi,
    const double *__restrict__ Z, const int *__restrict__ cmap, const double *__restrict__ uc,
    const double *__restrict__ ysym, double *zl) {
  extern __shared__ double smem[];
  const SubDesc d = subs[blockIdx.x];
  const int NI = d.NI, nc = d.nc, n = d.nti;
  // the subdomain interface vector: in LDS, or (GV, interfaces too large for
  // LDS) in the subdomain's own range of the output, updated in place
  double *xs = GV ? zl + d.loff : smem;
  double *red = GV ? smem : xs + NI;  // 16 x 64 partial sums
  double *t = red + 1024;   // 64
  double *cc = t + 64;      // nc: uc(indrowc) - Z^T x
  const int tid = threadIdx.x, g = tid >> 6, l = tid & 63;
  for (int k = tid; k < NI; k += KA_THR) {
    const int m = map[d.loff + k];
    xs[k] = m >= 0 ? w[d.loff + k] * rin[m] : 0.0;
  }
  publish_sync<GV>();
  for (int c = g; c < nc; c += 16) {
    const double *zc = Z + d.zoff + (size_t)c * NI;
    double s = 0.0;
    for (int i = l; i < NI; i += 64) s += zc[i] * xs[i];
    s = wave_sum(s);
    if (l == 0) cc[c] = uc[cmap[d.coff + c]] - s;
  }
  if (ysym) {
    // S'^{-1} x already applied (explicit inverse, symmetric GEMV): no sweeps
    publish_sync<GV>();
    for (int i = tid; i < NI; i += KA_THR) xs[i] = ysym[d.loff + i];
    publish_sync<GV>();
  } else {
  // forward sweep x_I <- Linv_I (x_I - sum_{k<I} L_Ik x_k); lane = tile row.
  // The Linv_I entries are loaded before the row sum (they do not depend on
  // it), and four tiles of the row are in flight at a time.
  for (int I = 0; I < n; I++) {
    const T *Li = pool + (size_t)(d.plinv + I) * TILE_ELEMS + l;
    double lv[4];
    if (DEEP) {
#pragma unroll
      for (int m = 0; m < 4; m++) lv[m] = (double)Li[(g + 16 * m) * 64];
    }
    double s = 0.0;
    if (DEEP) {
#pragma unroll 4
      for (int k = 0; k < I; k++) {
        const T *L = pool + (size_t)(d.pbase + packed_lower(n, I, k)) * TILE_ELEMS + l;
        const double *x = xs + k * 64;
#pragma unroll
        for (int m = 0; m < 4; m++) s += (double)L[(g + 16 * m) * 64] * x[g + 16 * m];
      }
    } else {
#pragma unroll 2
      for (int k = 0; k < I; k++) {
        const T *L = pool + (size_t)(d.pbase + packed_lower(n, I, k)) * TILE_ELEMS + l;
        const double *x = xs + k * 64;
#pragma unroll
        for (int m = 0; m < 4; m++) s += (double)L[(g + 16 * m) * 64] * x[g + 16 * m];
      }
    }
    red[g * 64 + l] = s;
    publish_sync<GV>();
    if (tid < 64) {
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 16; q++) v += red[q * 64 + l];
      t[l] = xs[I * 64 + l] - v;
    }
    publish_sync<GV>();
    {
      double u = 0.0;
      if (!DEEP) {
#pragma unroll
        for (int m = 0; m < 4; m++) lv[m] = (double)Li[(g + 16 * m) * 64];
      }
#pragma unroll
      for (int m = 0; m < 4; m++) u += lv[m] * t[g + 16 * m];
      red[g * 64 + l] = u;
    }
    publish_sync<GV>();
    if (tid < 64) {
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 16; q++) v += red[q * 64 + l];
      xs[I * 64 + l] = v;
    }
    publish_sync<GV>();
  }
  // backward sweep x_I <- Linv_I^T (x_I - sum_{k>I} L_kI^T x_k); lane = tile
  // row of L_kI (contiguous), wave g owns outputs g, g+16, g+32, g+48
  for (int I = n - 1; I >= 0; I--) {
    const T *Li = pool + (size_t)(d.plinv + I) * TILE_ELEMS + g * 64 + l;
    double lv[4];
    if (DEEP) {
#pragma unroll
      for (int m = 0; m < 4; m++) lv[m] = (double)Li[m * 16 * 64];
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (DEEP) {
#pragma unroll 4
      for (int k = I + 1; k < n; k++) {
        const T *L = pool + (size_t)(d.pbase + packed_lower(n, k, I)) * TILE_ELEMS + g * 64 + l;
        const double xv = xs[k * 64 + l];
#pragma unroll
        for (int m = 0; m < 4; m++) a[m] += (double)L[m * 16 * 64] * xv;
      }
    } else {
#pragma unroll 2
      for (int k = I + 1; k < n; k++) {
        const T *L = pool + (size_t)(d.pbase + packed_lower(n, k, I)) * TILE_ELEMS + g * 64 + l;
        const double xv = xs[k * 64 + l];
#pragma unroll
        for (int m = 0; m < 4; m++) a[m] += (double)L[m * 16 * 64] * xv;
      }
    }
#pragma unroll
    for (int m = 0; m < 4; m++) a[m] = wave_sum(a[m]);
    if (l == 0) {
#pragma unroll
      for (int m = 0; m < 4; m++) t[g + 16 * m] = xs[I * 64 + g + 16 * m] - a[m];
    }
    publish_sync<GV>();
    const double tv = t[l];
    if (!DEEP) {
#pragma unroll
      for (int m = 0; m < 4; m++) lv[m] = (double)Li[m * 16 * 64];
    }
#pragma unroll
    for (int m = 0; m < 4; m++) a[m] = wave_sum(lv[m] * tv);
    if (l == 0) {
#pragma unroll
      for (int m = 0; m < 4; m++) xs[I * 64 + g + 16 * m] = a[m];
    }
    publish_sync<GV>();
  }
  }
  publish_sync<GV>();
  for (int i = tid; i < NI; i += KA_THR) {
    double s = xs[i];
    for (int c = 0; c < nc; c++) s += Phi[d.zoff + (size_t)c * NI + i] * cc[c];
    zl[d.loff + i] = s;
  }
}

void launch_assemble(hipStream_t st, int n, const int *ptr, const int64_t *src, const double *w,
                     const double *ylocal, double *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_assemble, dim3((n + 255) / 256), dim3(256), 0, st, n, ptr, src, w, ylocal, y);
}
void launch_phit(hipStream_t st, const SubDesc *subs, int nsub, const double *Phi, const int *map,
                 const double *w, const double *x, double *rcs, int maxnc, int maxNI) {
  if (nsub <= 0 || maxnc <= 0) return;
  const size_t sh = (size_t)std::max(maxNI, 1) * sizeof(double);
  const dim3 grid((maxnc + PHIT_COLS - 1) / PHIT_COLS, nsub);
  if (sh <= max_dynamic_lds() && !g_trsv_global) {
    lds_request((const void *)k_phit<true>, sh, "coarse residual");
    hipLaunchKernelGGL(k_phit<true>, grid, dim3(256), sh, st, subs, Phi, map, w, x, rcs);
  } else {
    hipLaunchKernelGGL(k_phit<false>, grid, dim3(256), 0, st, subs, Phi, map, w, x, rcs);
  }
}
void launch_coarse_gather(hipStream_t st, int n, const int *ptr, const int *src,
                          const double *rcs, double *rc) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_coarse_gather, dim3((n + 255) / 256), dim3(256), 0, st, n, ptr, src, rcs, rc);
}
template <typename T>
void launch_kapply(hipStream_t st, const T *pool, const SubDesc *subs, int nsub, const int *map,
                   const double *w, const double *r, const double *Phi, const double *Z,
                   const int *cmap, const double *uc, const double *ysym, double *zl, int maxNI,
                   int maxnc, bool deep) {
  if (nsub <= 0) return;
  const size_t sh = (size_t)(maxNI + 1024 + 64 + maxnc) * sizeof(double);
  const size_t sg = (size_t)(1024 + 64 + maxnc) * sizeof(double);
  const bool gv = sh > max_dynamic_lds() || g_trsv_global;
  auto go = [&](auto kernel, size_t bytes) {
    lds_request((const void *)kernel, bytes, "preconditioner apply");
    hipLaunchKernelGGL(kernel, dim3(nsub), dim3(KA_THR), bytes, st, pool, subs, map, w, r, Phi, Z,
                       cmap, uc, ysym, zl);
  };
  if (gv) {
    if (deep && !ysym) go(k_kapply<T, true, true>, sg);
    else go(k_kapply<T, false, true>, sg);
  } else {
    if (deep && !ysym) go(k_kapply<T, true, false>, sh);
    else go(k_kapply<T, false, false>, sh);
  }
}
template void launch_kapply<double>(hipStream_t, const double *, const SubDesc *, int, const int *,
                                    const double *, const double *, const double *,
                                    const double *, const int *, const double *, const double *,
                                    double *, int, int, bool);
template void launch_kapply<float>(hipStream_t, const float *, const SubDesc *, int, const int *,
                                   const double *, const double *, const double *, const double *,
                                   const int *, const double *, const double *, double *, int, int,
                                   bool);

// ------------------------------------------------------------ vector ops
__global__ __launch_bounds__(256) void k_dots(int n, DotArgs a, double *__restrict__ partial) {
  __shared__ double red[4][256];
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    for (int d = 0; d < a.ndot; d++) s[d] += a.a[d][i] * a.b[d][i];
  for (int d = 0; d < 4; d++) red[d][threadIdx.x] = s[d];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h)
      for (int d = 0; d < 4; d++) red[d][threadIdx.x] += red[d][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

__global__ void k_dots_final(int nb, int ndot, const double *__restrict__ partial,
                             double *__restrict__ out) {
  __shared__ double red[256];
  for (int d = 0; d < ndot; d++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += partial[d * nb + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[d] = red[0];
    __syncthreads();
  }
}

void launch_dots(hipStream_t st, int n, DotArgs args, double *partial, double *out) {
  hipLaunchKernelGGL(k_dots, dim3(DOT_BLOCKS), dim3(256), 0, st, n, args, partial);
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(256), 0, st, DOT_BLOCKS, args.ndot, partial, out);
}

// The same single dot product with its result stored at base[*kp] (and,
// when host is not null, at host[*kp], pinned host memory the host polls):
// the slot index is read on the device, so that one captured graph of the
// PCG iteration serves every iteration.
__global__ void k_dots_final_at(int nb, const double *__restrict__ partial,
                                double *__restrict__ base, const int *__restrict__ kp,
                                double *host) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int k = *kp;
    base[k] = red[0];
    if (host) {
      host[k] = red[0];
      __threadfence_system();
    }
  }
}
void launch_dots_at(hipStream_t st, int n, const double *a, const double *b, double *partial,
                    double *base, const int *kp, double *host) {
  DotArgs args{};
  args.ndot = 1;
  args.a[0] = a;
  args.b[0] = b;
  hipLaunchKernelGGL(k_dots, dim3(DOT_BLOCKS), dim3(256), 0, st, n, args, partial);
  hipLaunchKernelGGL(k_dots_final_at, dim3(1), dim3(256), 0, st, DOT_BLOCKS, partial, base, kp,
                     host);
}

__global__ void k_axpy2(int n, double a, const double *__restrict__ x, double *__restrict__ y,
                        double a2, const double *__restrict__ x2, double *__restrict__ y2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] += a * x[i];
  if (y2) y2[i] += a2 * x2[i];
}
void launch_axpy2(hipStream_t st, int n, double a, const double *x, double *y, double a2,
                  const double *x2, double *y2) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy2, dim3((n + 255) / 256), dim3(256), 0, st, n, a, x, y, a2, x2, y2);
}
__global__ void k_xpay(int n, double b, const double *__restrict__ x, double *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + b * y[i];
}
void launch_xpay(hipStream_t st, int n, double b, const double *x, double *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_xpay, dim3((n + 255) / 256), dim3(256), 0, st, n, b, x, y);
}
__global__ void k_pcg_update(int n, const double *__restrict__ rz, const double *__restrict__ pap,
                             int k, const double *__restrict__ p, const double *__restrict__ ap,
                             double *__restrict__ u, double *__restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = rz[k - 1] / pap[k];
  u[i] += alpha * p[i];
  r[i] -= alpha * ap[i];
}
void launch_pcg_update(hipStream_t st, int n, const double *rz, const double *pap, int k,
                       const double *p, const double *ap, double *u, double *r) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pcg_update, dim3((n + 255) / 256), dim3(256), 0, st, n, rz, pap, k, p, ap,
                     u, r);
}
__global__ void k_pcg_dir(int n, const double *__restrict__ rz, int k,
                          const double *__restrict__ z, double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double beta = rz[k] / rz[k - 1];
  p[i] = z[i] + beta * p[i];
}
void launch_pcg_dir(hipStream_t st, int n, const double *rz, int k, const double *z, double *p) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pcg_dir, dim3((n + 255) / 256), dim3(256), 0, st, n, rz, k, z, p);
}

// iteration-index forms of the two PCG updates (index read from *kp) and
// the counter step that ends a graph replay
__global__ void k_pcg_update_at(int n, const double *__restrict__ rz, const double *__restrict__ pap,
                                const int *__restrict__ kp, const double *__restrict__ p,
                                const double *__restrict__ ap, double *__restrict__ u,
                                double *__restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = *kp;
  const double alpha = rz[k - 1] / pap[k];
  u[i] += alpha * p[i];
  r[i] -= alpha * ap[i];
}
void launch_pcg_update_at(hipStream_t st, int n, const double *rz, const double *pap, const int *kp,
                          const double *p, const double *ap, double *u, double *r) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pcg_update_at, dim3((n + 255) / 256), dim3(256), 0, st, n, rz, pap, kp, p,
                     ap, u, r);
}
__global__ void k_pcg_dir_at(int n, const double *__restrict__ rz, const int *__restrict__ kp,
                             const double *__restrict__ z, double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = *kp;
  const double beta = rz[k] / rz[k - 1];
  p[i] = z[i] + beta * p[i];
}
void launch_pcg_dir_at(hipStream_t st, int n, const double *rz, const int *kp, const double *z,
                       double *p) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pcg_dir_at, dim3((n + 255) / 256), dim3(256), 0, st, n, rz, kp, z, p);
}
__global__ void k_count_set(int *kp, int v, int add) {
  if (threadIdx.x == 0) *kp = add ? *kp + v : v;
}
void launch_count_set(hipStream_t st, int *kp, int v) {
  hipLaunchKernelGGL(k_count_set, dim3(1), dim3(64), 0, st, kp, v, 0);
}
void launch_count_inc(hipStream_t st, int *kp) {
  hipLaunchKernelGGL(k_count_set, dim3(1), dim3(64), 0, st, kp, 1, 1);
}

// ------------------------------------------------ Krylov recycling basis
__global__ __launch_bounds__(256) void k_multidot(int n, const double *__restrict__ V, int64_t ld,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ partial) {
  __shared__ double red[256];
  const double *v = V + (size_t)blockIdx.y * ld;
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += v[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_multidot_final(int nb, const double *__restrict__ partial,
                                                        double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[(size_t)blockIdx.x * nb + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
void launch_multidot(hipStream_t st, int n, const double *V, int64_t ld, int m, const double *x,
                     double *partial, double *out) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_multidot, dim3(DOT_BLOCKS, m), dim3(256), 0, st, n, V, ld, x, partial);
  hipLaunchKernelGGL(k_multidot_final, dim3(m), dim3(256), 0, st, DOT_BLOCKS, partial, out);
}
__global__ void k_basis_update(int n, const double *__restrict__ V, int64_t ld, int m,
                               const double *__restrict__ c, double sign, double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < m; j++) s += V[(size_t)j * ld + i] * c[j];
  y[i] += sign * s;
}
void launch_basis_update(hipStream_t st, int n, const double *V, int64_t ld, int m,
                         const double *c, double sign, double *y) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_basis_update, dim3((n + 255) / 256), dim3(256), 0, st, n, V, ld, m, c, sign,
                     y);
}
__global__ void k_basis_store(int n, const double *__restrict__ pap, const double *__restrict__ p,
                              const double *__restrict__ ap, double *__restrict__ vcol,
                              double *__restrict__ wcol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double sc = 1.0 / sqrt(*pap);
  vcol[i] = sc * p[i];
  wcol[i] = sc * ap[i];
}
void launch_basis_store(hipStream_t st, int n, const double *pap, const double *p,
                        const double *ap, double *vcol, double *wcol) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_basis_store, dim3((n + 255) / 256), dim3(256), 0, st, n, pap, p, ap, vcol,
                     wcol);
}
__global__ void k_bicg_p(int n, double b, double w, const double *__restrict__ r,
                         const double *__restrict__ v, double *__restrict__ p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = r[i] + b * (p[i] - w * v[i]);
}
void launch_bicg_p(hipStream_t st, int n, double b, double w, const double *r, const double *v,
                   double *p) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_bicg_p, dim3((n + 255) / 256), dim3(256), 0, st, n, b, w, r, v, p);
}
__global__ void k_sub_scaled(int n, const double *__restrict__ x, double a,
                             const double *__restrict__ y, double *__restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x[i] - a * y[i];
}
void launch_sub_scaled(hipStream_t st, int n, const double *x, double a, const double *y,
                       double *z) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sub_scaled, dim3((n + 255) / 256), dim3(256), 0, st, n, x, a, y, z);
}
__global__ void k_mask(int n, const char *__restrict__ m, double *__restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && m[i]) v[i] = 0.0;
}
void launch_mask(hipStream_t st, int n, const char *mask, double *v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mask, dim3((n + 255) / 256), dim3(256), 0, st, n, mask, v);
}

// ------------------------------------------------ coefficient gradient
// one thread per element: g_e = -lam_e^T A_e u_e
__global__ void k_element_gradient(int nel, int nke, const int *__restrict__ edof,
                                   const int *__restrict__ eid, const double *__restrict__ ae,
                                   int per_element, const double *__restrict__ u,
                                   const double *__restrict__ lam, double *__restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  const int *dof = edof + (size_t)e * nke;
  const double *A = ae + (per_element ? (size_t)eid[e] * nke * nke : 0);
  double g = 0.0;
  for (int i = 0; i < nke; i++) {
    const double li = lam[dof[i]];
    if (li == 0.0) continue;
    double s = 0.0;
    for (int j = 0; j < nke; j++) s += A[i * nke + j] * u[dof[j]];
    g += li * s;
  }
  grad[eid[e]] = -g;
}
void launch_element_gradient(hipStream_t st, int nel, int nke, const int *edof, const int *eid,
                             const double *ae, bool per_element, const double *u,
                             const double *lam, double *grad) {
  if (nel <= 0) return;
  hipLaunchKernelGGL(k_element_gradient, dim3((nel + 255) / 256), dim3(256), 0, st, nel, nke, edof,
                     eid, ae, per_element ? 1 : 0, u, lam, grad);
}

// ------------------------------------------- local order <-> device order
__global__ void k_to_device_order(const OrderDesc *__restrict__ ds, const int *__restrict__ perm,
                                  const double *__restrict__ src, double *__restrict__ vec) {
  const OrderDesc d = ds[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.ndofs) vec[d.voff + perm[d.permoff + i]] = src[d.poff + i];
}
__global__ void k_from_device_order(const OrderDesc *__restrict__ ds, const int *__restrict__ perm,
                                    const double *__restrict__ vec, double *__restrict__ dst) {
  const OrderDesc d = ds[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.ndofs) dst[d.poff + i] = vec[d.voff + perm[d.permoff + i]];
}
__global__ void k_ifc_scatter(int n, const int *__restrict__ ptr, const int64_t *__restrict__ pos,
                              const double *__restrict__ u, double *__restrict__ vec) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double v = u[g];
  for (int q = ptr[g]; q < ptr[g + 1]; q++) vec[pos[q]] = v;
}
void launch_to_device_order(hipStream_t st, const OrderDesc *d, int nsub, const int *perm,
                            const double *src, double *vec, int maxndofs) {
  if (nsub <= 0 || maxndofs <= 0) return;
  hipLaunchKernelGGL(k_to_device_order, dim3((maxndofs + 255) / 256, nsub), dim3(256), 0, st, d,
                     perm, src, vec);
}
void launch_from_device_order(hipStream_t st, const OrderDesc *d, int nsub, const int *perm,
                              const double *vec, double *dst, int maxndofs) {
  if (nsub <= 0 || maxndofs <= 0) return;
  hipLaunchKernelGGL(k_from_device_order, dim3((maxndofs + 255) / 256, nsub), dim3(256), 0, st, d,
                     perm, vec, dst);
}
// r = f - A u of every subdomain in local dof order (FP64), A the subdomain
// matrix as the tile scatter builds it (couplings of fixed dofs dropped, a
// fixed diagonal below 1e-12 read as 1), from its assembled upper-triangle
// triplets through a row-segmented index shared by the subdomain's class:
// row i lists (triplet, column) of both triangles, columns ascending.
// RESID_LANES lanes per row take every RESID_LANES-th entry and their sums
// meet in a fixed butterfly, so the result does not depend on scheduling.
constexpr int RESID_LANES = 8;
__global__ __launch_bounds__(256) void k_resid_rows(const ResidSub *__restrict__ subs,
                                                    const int *__restrict__ rowptr,
                                                    const int2 *__restrict__ ent,
                                                    const char *__restrict__ fix,
                                                    const double *__restrict__ f,
                                                    const double *__restrict__ u,
                                                    double *__restrict__ r) {
  const ResidSub d = subs[blockIdx.y];
  const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) / RESID_LANES);
  const int lane = threadIdx.x % RESID_LANES;
  double s = 0.0;
  if (row < d.ndofs) {
    const int *rp = rowptr + d.rp_off;
    const int2 *E = ent + d.ent_off;
    const char *fx = fix + d.foff;
    const double *uu = u + d.poff;
    const bool fr = fx[row] != 0;
    for (int e = rp[row] + lane; e < rp[row + 1]; e += RESID_LANES) {
      const int2 q = E[e];
      double a = d.av[q.x];
      if (q.y != row) {
        if (fr || fx[q.y]) continue;
      } else if (fr && fabs(a) < 1e-12) {
        a = 1.0;
      }
      s += a * uu[q.y];
    }
  }
#pragma unroll
  for (int m = 1; m < RESID_LANES; m <<= 1) s += __shfl_xor(s, m, RESID_LANES);
  if (row < d.ndofs && lane == 0) r[d.poff + row] = f[d.poff + row] - s;
}
void launch_resid_rows(hipStream_t st, const ResidSub *subs, int nsub, const int *rowptr,
                       const int2 *ent, const char *fix, const double *f, const double *u,
                       double *r, int maxndofs) {
  if (nsub <= 0 || maxndofs <= 0) return;
  const int64_t nthr = (int64_t)maxndofs * RESID_LANES;
  hipLaunchKernelGGL(k_resid_rows, dim3((unsigned)((nthr + 255) / 256), nsub), dim3(256), 0, st,
                     subs, rowptr, ent, fix, f, u, r);
}

__global__ void k_level_fill(int64_t n, const int *__restrict__ src, const double *__restrict__ w,
                             const double *__restrict__ rc, double *__restrict__ vec) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int g = src[p];
  vec[p] = g >= 0 ? w[p] * rc[g] : 0.0;
}
__global__ void k_level_gather(int n, const int *__restrict__ pos, const double *__restrict__ vec,
                               double *__restrict__ uc) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) uc[g] = vec[pos[g]];
}
void launch_level_fill(hipStream_t st, int64_t n, const int *src, const double *w, const double *rc,
                       double *vec) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_level_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, src, w, rc,
                     vec);
}
void launch_level_gather(hipStream_t st, int ndofc, const int *pos, const double *vec, double *uc) {
  if (ndofc <= 0) return;
  hipLaunchKernelGGL(k_level_gather, dim3((ndofc + 255) / 256), dim3(256), 0, st, ndofc, pos, vec, uc);
}
void launch_ifc_scatter(hipStream_t st, int n, const int *ptr, const int64_t *pos, const double *u,
                        double *vec) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_ifc_scatter, dim3((n + 255) / 256), dim3(256), 0, st, n, ptr, pos, u, vec);
}

// four double2 per thread, all loads issued before the stores
__global__ __launch_bounds__(256) void k_copy_calibration(const double2 *__restrict__ src,
                                                          double2 *__restrict__ dst, int64_t n) {
  const int64_t i0 = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  double2 v[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (i0 + 256 * k < n) v[k] = src[i0 + 256 * k];
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (i0 + 256 * k < n) dst[i0 + 256 * k] = v[k];
}
void launch_copy_calibration(hipStream_t st, const double2 *src, double2 *dst, int64_t n) {
  if (n <= 0) return;
  const int64_t nb = (n + 1023) / 1024;
  hipLaunchKernelGGL(k_copy_calibration, dim3((unsigned)nb), dim3(256), 0, st, src, dst, n);
}

}  // namespace bddc
