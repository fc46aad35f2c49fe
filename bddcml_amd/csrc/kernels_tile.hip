// kernels_tile.hip -- batched 64x64 tile algebra on CDNA4 matrix cores, FP64
// (and FP32 for the preconditioner factors of the mixed-precision mode).
//
// One workgroup (4 waves) owns one output tile.  Each wave computes a 32x32
// quadrant as 2x2 blocks of 16x16x4 MFMA accumulators; operand tiles
// are staged through LDS with a padded leading dimension.  These kernels
// carry every dense factorisation of the path: the partial Cholesky of the
// subdomain matrices (Schur complements), the Cholesky / inverse of the
// augmented interface matrices and of the coarse matrix.
#include "kernels.hpp"

namespace bddc {

constexpr int LDL = 65;  // LDS leading dimension (elements)
constexpr int NTHR = 256;

// Element-type traits of the 16x16x4 matrix-core instruction.  Lane l
// supplies A(l%16, l/16) and B(l/16, l%16); register v of lane l holds
// D(row(l, v), l%16), with row = l/16 + 4v for f64 and 4(l/16) + v for f32
// (both measured on gfx950).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef double2 vec;  // 16-byte global access
  static constexpr int VE = 2;
  __device__ static __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef float4 vec;
  static constexpr int VE = 4;
  __device__ static __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <typename T> __device__ __forceinline__ T vget(const typename Mfma<T>::vec &v, int i);
template <> __device__ __forceinline__ double vget<double>(const double2 &v, int i) {
  return i == 0 ? v.x : v.y;
}
template <> __device__ __forceinline__ float vget<float>(const float4 &v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
template <typename T> __device__ __forceinline__ void vset(typename Mfma<T>::vec &v, int i, T x);
template <> __device__ __forceinline__ void vset<double>(double2 &v, int i, double x) {
  if (i == 0) v.x = x;
  else v.y = x;
}
template <> __device__ __forceinline__ void vset<float>(float4 &v, int i, float x) {
  if (i == 0) v.x = x;
  else if (i == 1) v.y = x;
  else if (i == 2) v.z = x;
  else v.w = x;
}

// a tile held in registers: thread tid owns vector elements tid + NTHR*i
template <typename T> struct TileRegs {
  static constexpr int N = TILE_ELEMS / Mfma<T>::VE / NTHR;
  typename Mfma<T>::vec v[N];
};
template <typename T>
__device__ __forceinline__ void tile_load(const T *__restrict__ t, TileRegs<T> &r) {
  const typename Mfma<T>::vec *V = reinterpret_cast<const typename Mfma<T>::vec *>(t);
#pragma unroll
  for (int i = 0; i < TileRegs<T>::N; i++) r.v[i] = V[threadIdx.x + NTHR * i];
}
// lds[q*LDL + r] = T(r,q) (direct) or T(q,r) (transposed)
template <typename T>
__device__ __forceinline__ void tile_store_lds(const TileRegs<T> &r, T *lds, bool trans) {
  constexpr int VE = Mfma<T>::VE;
#pragma unroll
  for (int i = 0; i < TileRegs<T>::N; i++) {
    const int e = VE * (threadIdx.x + NTHR * i);
    const int row = e & 63, col = e >> 6;
#pragma unroll
    for (int j = 0; j < VE; j++) {
      const T x = vget<T>(r.v[i], j);
      if (!trans) lds[col * LDL + row + j] = x;
      else lds[(row + j) * LDL + col] = x;
    }
  }
}
template <typename T>
__device__ __forceinline__ void tile_to_lds(const T *__restrict__ t, T *lds, bool trans) {
  TileRegs<T> r;
  tile_load(t, r);
  tile_store_lds(r, lds, trans);
}
// global tile (column major) from lds holding element (r,c) at c*LDL + r
template <typename T> __device__ __forceinline__ void lds_to_tile(const T *lds, T *__restrict__ t) {
  constexpr int VE = Mfma<T>::VE;
  typename Mfma<T>::vec *V = reinterpret_cast<typename Mfma<T>::vec *>(t);
#pragma unroll
  for (int i = 0; i < TileRegs<T>::N; i++) {
    const int e2 = threadIdx.x + NTHR * i, e = VE * e2;
    const int row = e & 63, col = e >> 6;
    typename Mfma<T>::vec v;
#pragma unroll
    for (int j = 0; j < VE; j++) vset<T>(v, j, lds[col * LDL + row + j]);
    V[e2] = v;
  }
}

// acc(r,c) += sum_q As[q*LDL + r] * Bs[q*LDL + c] over this wave's 32x32
// quadrant, as 2x2 blocks of 16x16
template <typename T>
__device__ __forceinline__ void mfma_accumulate(const T *As, const T *Bs,
                                                typename Mfma<T>::v4 acc[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int R0 = (wave & 1) * 32, C0 = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int q = 0; q < 64; q += 4) {
    const T *ar = As + (q + lk) * LDL + R0 + li;
    const T *br = Bs + (q + lk) * LDL + C0 + li;
    const T a0 = ar[0], a1 = ar[16];
    const T b0 = br[0], b1 = br[16];
    acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
  }
}

// tile coordinates of accumulator (bi, bj, v) of this lane
template <typename T> __device__ __forceinline__ void acc_coord(int bi, int bj, int v, int &r, int &c) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  r = (wave & 1) * 32 + bi * 16 + Mfma<T>::row(lane, v);
  c = (wave >> 1) * 32 + bj * 16 + (lane & 15);
}

template <typename T> __device__ __forceinline__ void zero_acc(typename Mfma<T>::v4 acc[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (typename Mfma<T>::v4){0, 0, 0, 0};
}

// acc = sum over pairs of opA(A) opB(B) in the "r,q x c,q" form.  Software
// pipelined two pairs deep: the tiles of pairs p+1 and p+2 are in flight in
// registers while the matrix cores work on pair p from LDS.
template <typename T>
__device__ __forceinline__ void gemm_sum(const T *__restrict__ pool, const TilePair *pairs, int p0,
                                         int p1, int base, int kind, T *As, T *Bs,
                                         typename Mfma<T>::v4 acc[2][2]) {
  const bool ta = (kind == GEMM_TN);
  const bool tb = (kind == GEMM_NN || kind == GEMM_TN);
  const T *pb = pool + (size_t)base * TILE_ELEMS;
  if (p0 >= p1) return;
  TileRegs<T> ra0, rb0, ra1, rb1;
  auto load = [&](int p, TileRegs<T> &ra, TileRegs<T> &rb) {
    const TilePair pr = pairs[p];
    tile_load(pb + (size_t)pr.a * TILE_ELEMS, ra);
    tile_load(pb + (size_t)pr.b * TILE_ELEMS, rb);
  };
  // one pair from registers: to LDS, refill the registers with pair p + 2,
  // multiply
  auto step = [&](int p, TileRegs<T> &ra, TileRegs<T> &rb) {
    __syncthreads();
    tile_store_lds(ra, As, ta);
    tile_store_lds(rb, Bs, tb);
    __syncthreads();
    if (p + 2 < p1) load(p + 2, ra, rb);
    mfma_accumulate<T>(As, Bs, acc);
  };
  load(p0, ra0, rb0);
  if (p0 + 1 < p1) load(p0 + 1, ra1, rb1);
  int p = p0;
  for (; p + 1 < p1; p += 2) {
    step(p, ra0, rb0);
    step(p + 1, ra1, rb1);
  }
  if (p < p1) step(p, ra0, rb0);
}

// E = alpha*C + beta*acc, left in As as element (r,c) at c*LDL + r
template <typename T>
__device__ __forceinline__ void form_e(const T *__restrict__ C, double alpha, double beta,
                                       typename Mfma<T>::v4 acc[2][2], T *As) {
  __syncthreads();
  if (alpha != 0.0) tile_to_lds(C, As, false);
  __syncthreads();
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int r, c;
        acc_coord<T>(bi, bj, v, r, c);
        T e = (T)beta * acc[bi][bj][v];
        if (alpha != 0.0) e += (T)alpha * As[c * LDL + r];
        As[c * LDL + r] = e;
      }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void acc_to_lds(typename Mfma<T>::v4 acc[2][2], T *As) {
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int r, c;
        acc_coord<T>(bi, bj, v, r, c);
        As[c * LDL + r] = acc[bi][bj][v];
      }
}

template <typename T>
__device__ __forceinline__ void tile_gemm_task_lds(T *__restrict__ pool, const TileTask t,
                                                   const TilePair *__restrict__ pairs, int kind,
                                                   int post, double alpha, double beta, T *As,
                                                   T *Bs) {
  typename Mfma<T>::v4 acc[2][2];
  zero_acc<T>(acc);
  gemm_sum<T>(pool, pairs, t.p0, t.p1, t.base, kind, As, Bs, acc);
  T *C = pool + (size_t)t.c * TILE_ELEMS;
  form_e<T>(C, t.fresh ? 0.0 : alpha, beta, acc, As);  // E in As
  if (post == POST_NONE) {
    lds_to_tile(As, C);
    return;
  }
  const T *X = pool + (size_t)t.x * TILE_ELEMS;
  if (post == POST_RIGHT_XT) {
    // C_new(r,c) = sum_q E(r,q) X(c,q): A' = E (already As), B'(c,q) = X(c,q)
    tile_to_lds(X, Bs, false);
  } else {
    // C_new(r,c) = sum_q X(r,q) E(q,c) (LEFT_X) or X(q,r) E(q,c) (LEFT_XT):
    // B'(c,q) = E(q,c) = As[c*LDL + q]
    for (int e = threadIdx.x; e < 4096; e += NTHR) {
      int q = e & 63, c = e >> 6;
      Bs[q * LDL + c] = As[c * LDL + q];
    }
    __syncthreads();
    tile_to_lds(X, As, post == POST_LEFT_XT);
  }
  __syncthreads();
  zero_acc<T>(acc);
  mfma_accumulate<T>(As, Bs, acc);
  __syncthreads();
  acc_to_lds<T>(acc, As);
  __syncthreads();
  lds_to_tile(As, C);
}
template <typename T>
__device__ __forceinline__ void tile_gemm_task(T *__restrict__ pool, const TileTask t,
                                               const TilePair *__restrict__ pairs, int kind,
                                               int post, double alpha, double beta) {
  __shared__ T As[64 * LDL];
  __shared__ T Bs[64 * LDL];
  tile_gemm_task_lds<T>(pool, t, pairs, kind, post, alpha, beta, As, Bs);
}

// Logical task of this workgroup.  Workgroups reach the 8 XCDs round-robin
// (blockIdx % 8); with `chunked` set each XCD receives one contiguous range
// of the task list instead, so that neighbouring tasks, which share a row of
// operand tiles, run under the same L2.  A bijection of [0, gridDim.x).
__device__ __forceinline__ int task_index(int chunked) {
  const int b = blockIdx.x;
  if (!chunked) return b;
  const int n = gridDim.x, q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

static int g_task_order = 1;
void set_task_order(int chunked) { g_task_order = chunked != 0; }

template <typename T>
__global__ __launch_bounds__(NTHR) void k_tile_gemm(T *__restrict__ pool,
                                                    const TileTask *__restrict__ tasks,
                                                    const TilePair *__restrict__ pairs, int kind,
                                                    int post, double alpha, double beta,
                                                    int chunked) {
  const TileTask t = tasks[task_index(chunked)];
  if (t.mode >= 0)
    tile_gemm_task<T>(pool, t, pairs, t.mode & 3, (t.mode >> 2) & 3, 1.0, (t.mode & 16) ? -1.0 : 0.0);
  else
    tile_gemm_task<T>(pool, t, pairs, kind, post, alpha, beta);
}

// the Schur-complement block update S = A22 - sum W^T W in a launch of its
// own symbol, so that profiles separate it from the factorization steps
__global__ __launch_bounds__(NTHR) void k_schur_gemm(double *__restrict__ pool,
                                                     const TileTask *__restrict__ tasks,
                                                     const TilePair *__restrict__ pairs,
                                                     int chunked) {
  tile_gemm_task<double>(pool, tasks[task_index(chunked)], pairs, GEMM_NT, POST_NONE, 1.0, -1.0);
}
// the same update on FP32 tiles (FP32 interior factorization)
__global__ __launch_bounds__(NTHR) void k_schur_gemm_f32(float *__restrict__ pool,
                                                         const TileTask *__restrict__ tasks,
                                                         const TilePair *__restrict__ pairs,
                                                         int chunked) {
  tile_gemm_task<float>(pool, tasks[task_index(chunked)], pairs, GEMM_NT, POST_NONE, 1.0, -1.0);
}

// acc += sum_{k < nk} A(r, k) B'(c, k) for one 16x16 block in the lane layout
// of the 16x16x4 MFMA: lane l supplies A(l%16, q + l/16) and B'(l%16, q + l/16);
// A(r, k) = pa[k*ska + r*sra], B'(c, k) = pb[k*skb + c*srb]
template <typename T>
__device__ __forceinline__ typename Mfma<T>::v4 mfma16(typename Mfma<T>::v4 acc, const T *pa,
                                                       int sra, int ska, const T *pb, int srb,
                                                       int skb, int nk) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  for (int q = 0; q < nk; q += 4) {
    const T a = pa[(q + lk) * ska + li * sra];
    const T b = pb[(q + lk) * skb + li * srb];
    acc = Mfma<T>::op(a, b, acc);
  }
  return acc;
}

// value of x in lane l (uniform l) for every lane of the wave
__device__ __forceinline__ double lane_bcast(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_bcast(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// 1/sqrt(d) for d > 0: the hardware estimate and Newton steps
// y <- y + y (1/2 - d y^2 / 2) (two for FP64, one for FP32)
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
  y = fma(y, fma(-h * y, y, 0.5), y);
  return fma(y, fma(-h * y, y, 0.5), y);
}
__device__ __forceinline__ float rsqrt_nr(float d) {
  const float y = __builtin_amdgcn_rsqf(d);
  return fmaf(y, fmaf(-0.5f * d * y, y, 0.5f), y);
}

// Cholesky factor of block column J (16 columns from c0, rows c0..63) by
// one wave in registers: lane l holds row c0 + l, right-looking over the
// block's columns, so the rows below the diagonal block come out as L(R, J)
// with no separate panel solve.  Every cross-lane value is a broadcast from
// one lane (lane reads into scalar registers).  One reciprocal square root
// per pivot gives both L(j,j) = d / sqrt(d) and 1/L(j,j) (kept in dv); the
// updates run on every lane unmasked (entries above the diagonal are never
// read as L and are zeroed).
template <typename T> __device__ __forceinline__ void chol_panel16(T *As, int c0, T *dv, int &bad) {
  const int l = threadIdx.x & 63;
  const int r = c0 + l;
  const bool live = r < 64;
  const int rr = live ? r : 63;  // lanes past the last row compute a copy
  T a[16];
#pragma unroll
  for (int c = 0; c < 16; c++) a[c] = As[(c0 + c) * LDL + rr];
  T dinv = (T)0;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    T d = lane_bcast(a[j], j);
    if (!(d > (T)0)) {
      bad = 1;
      d = (T)1;
    }
    const T q = rsqrt_nr(d);
    a[j] = (l == j) ? d * q : a[j] * q;
    if (l == j) dinv = q;
#pragma unroll
    for (int k = j + 1; k < 16; k++) a[k] -= a[j] * lane_bcast(a[j], k);
  }
  if (l < 16) {
#pragma unroll
    for (int c = 0; c < 16; c++)
      if (c > l) a[c] = (T)0;
    dv[c0 + l] = dinv;
  }
  if (live) {
#pragma unroll
    for (int c = 0; c < 16; c++) As[(c0 + c) * LDL + r] = a[c];
  }
}

// inverse of the 16x16 lower diagonal block at (c0, c0) (L in As, 1/L(j,j)
// in dv) into the same block of Bs, by one wave: lane i (mod 16) computes
// column i by forward substitution; L(r, k) is a[k] of lane r
template <typename T>
__device__ __forceinline__ void inv16(const T *As, T *Bs, const T *dv, int c0) {
  const int lane = threadIdx.x & 63, i = lane & 15;
  T a[16];
#pragma unroll
  for (int c = 0; c < 16; c++) a[c] = As[(c0 + c) * LDL + c0 + i];
  T x[16];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    T sacc = (r == i) ? (T)1 : (T)0;
#pragma unroll
    for (int k = 0; k < r; k++) sacc -= lane_bcast(a[k], r) * x[k];
    x[r] = sacc * dv[c0 + r];
  }
  if (lane < 16) {
#pragma unroll
    for (int r = 0; r < 16; r++) Bs[(c0 + i) * LDL + c0 + r] = x[r];
  }
}

// C = C - sum of the task's pairs A B^T, then its Cholesky factor L into C
// and inv(L) into tile t.x; a non-positive pivot sets *err.  Blocked by 16:
// per block column the update by the columns before it on the matrix
// cores, then its panel Cholesky by one wave; then the four diagonal-block
// inverses, one per wave, and the off-diagonal blocks of the inverse
// (X_IJ = -inv(L_II) sum_{J<=K<I} L_IK X_KJ) on the matrix cores.
template <typename T>
__device__ __forceinline__ void tile_potrf_task(T *__restrict__ pool, const TileTask t,
                                                const TilePair *__restrict__ pairs,
                                                int *__restrict__ err, T *As, T *Bs,
                                                T (*Ts)[16 * 17], T *dv, int &bad) {
  typedef typename Mfma<T>::v4 v4;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  v4 acc[2][2];
  zero_acc<T>(acc);
  gemm_sum<T>(pool, pairs, t.p0, t.p1, t.base, GEMM_NT, As, Bs, acc);
  T *C = pool + (size_t)t.c * TILE_ELEMS;
  form_e<T>(C, t.fresh ? 0.0 : 1.0, -1.0, acc, As);  // ends with a barrier
  for (int e = tid; e < 64 * LDL; e += NTHR) Bs[e] = (T)0;
  if (tid == 0) bad = 0;
  __syncthreads();
  const int ocol = lane & 15;  // accumulator element (Mfma<T>::row(lane, v), ocol)
  for (int J = 0; J < 4; J++) {
    const int c0 = 16 * J;
    // block column J -= L(:, 0:c0) L(J rows, 0:c0)^T, one row block per wave
    if (J > 0) {
      if (wave < 4 - J) {
        const int R = J + wave;
        v4 u = (v4){0, 0, 0, 0};
        u = mfma16<T>(u, As + 16 * R, 1, LDL, As + c0, 1, LDL, c0);
#pragma unroll
        for (int v = 0; v < 4; v++) As[(c0 + ocol) * LDL + 16 * R + Mfma<T>::row(lane, v)] -= u[v];
      }
      __syncthreads();
    }
    if (wave == 0) {
      int b = 0;
      chol_panel16<T>(As, c0, dv, b);
      if (b && lane == 0) bad = 1;
    }
    __syncthreads();
  }
  // zero the strict upper triangle of L
  for (int e = tid; e < 4096; e += NTHR) {
    int r = e & 63, c = e >> 6;
    if (r < c) As[c * LDL + r] = (T)0;
  }
  inv16<T>(As, Bs, dv, 16 * wave);
  __syncthreads();
  // off-diagonal blocks of inv(L), one block row at a time
  for (int I = 1; I < 4; I++) {
    const int Jb = wave;
    v4 u = (v4){0, 0, 0, 0};
    if (Jb < I) {
      for (int K = Jb; K < I; K++)
        u = mfma16<T>(u, As + (16 * K) * LDL + 16 * I, 1, LDL, Bs + (16 * Jb) * LDL + 16 * K, LDL, 1, 16);
#pragma unroll
      for (int v = 0; v < 4; v++) Ts[wave][ocol * 17 + Mfma<T>::row(lane, v)] = u[v];
    }
    __syncthreads();
    if (Jb < I) {
      u = (v4){0, 0, 0, 0};
      u = mfma16<T>(u, Bs + (16 * I) * LDL + 16 * I, 1, LDL, Ts[wave], 17, 1, 16);
#pragma unroll
      for (int v = 0; v < 4; v++) Bs[(16 * Jb + ocol) * LDL + 16 * I + Mfma<T>::row(lane, v)] = -u[v];
    }
    __syncthreads();
  }
  lds_to_tile(As, C);
  lds_to_tile(Bs, pool + (size_t)t.x * TILE_ELEMS);
  if (tid == 0 && bad) atomicOr(err, 1);
}

template <typename T>
__global__ __launch_bounds__(NTHR) void k_tile_potrf(T *__restrict__ pool,
                                                     const TileTask *__restrict__ tasks,
                                                     const TilePair *__restrict__ pairs,
                                                     int *__restrict__ err, int chunked) {
  __shared__ T As[64 * LDL];
  __shared__ T Bs[64 * LDL];
  __shared__ T Ts[4][16 * 17];
  __shared__ T dv[64];
  __shared__ int bad;
  tile_potrf_task<T>(pool, tasks[task_index(chunked)], pairs, err, As, Bs, Ts, dv, bad);
}

// A launch whose tasks are tile products with their own mode (task_mode) or
// tile factorizations (mode kTaskPotrf): the trailing update of one column
// of a right-looking Cholesky together with the factorization of the next
// diagonal tile, whose update is among the tasks
__global__ __launch_bounds__(NTHR) void k_tile_step(double *__restrict__ pool,
                                                    const TileTask *__restrict__ tasks,
                                                    const TilePair *__restrict__ pairs,
                                                    int *__restrict__ err, int chunked) {
  __shared__ double As[64 * LDL];
  __shared__ double Bs[64 * LDL];
  __shared__ double Ts[4][16 * 17];
  __shared__ double dv[64];
  __shared__ int bad;
  const TileTask t = tasks[task_index(chunked)];
  if (t.mode == kTaskPotrf)
    tile_potrf_task<double>(pool, t, pairs, err, As, Bs, Ts, dv, bad);
  else
    tile_gemm_task_lds<double>(pool, t, pairs, t.mode & 3, (t.mode >> 2) & 3, 1.0,
                               (t.mode & 16) ? -1.0 : 0.0, As, Bs);
}

__global__ __launch_bounds__(NTHR) void k_tile_mirror(double *__restrict__ pool,
                                                      const MirrorJob *__restrict__ jobs) {
  __shared__ double As[64 * LDL];
  const MirrorJob j = jobs[blockIdx.x];
  tile_to_lds(pool + (size_t)j.src * TILE_ELEMS, As, j.trans != 0);
  __syncthreads();
  lds_to_tile(As, pool + (size_t)j.dst * TILE_ELEMS);
}

template <typename T>
__global__ void k_zero_tiles(T *__restrict__ pool, const int *__restrict__ ids) {
  uint4 *Q = reinterpret_cast<uint4 *>(pool + (size_t)ids[blockIdx.x] * TILE_ELEMS);
  constexpr int n16 = TILE_ELEMS * (int)sizeof(T) / 16;
  for (int e = threadIdx.x; e < n16; e += blockDim.x) Q[e] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_scatter_ones(double *__restrict__ pool, const int64_t *__restrict__ idx,
                               int64_t n) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) pool[idx[k]] = 1.0;
}

// Assembled subdomain triplets (upper triangle, local dof indices) into the
// tile pool: Dirichlet elimination keeping the diagonal
// (src/module_sm.f90:431-600), device permutation, both triangles inside
// diagonal tiles; then 1.0 on the diagonal of every padding slot (slots no
// triplet reaches).
template <typename T>
__global__ void k_scatter_sub(T *__restrict__ pool, const ScatterSub *__restrict__ subs,
                              const int *__restrict__ perm, const char *__restrict__ fix,
                              const int *__restrict__ ttab, const int *__restrict__ pads,
                              int *__restrict__ err) {
  const ScatterSub d = subs[blockIdx.y];
  T *P = pool + (size_t)d.base * TILE_ELEMS;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d.nnz + d.npad; e += gridDim.x * blockDim.x) {
    if (e >= d.nnz) {
      P[pads[d.padoff + e - d.nnz]] = 1.0;
      continue;
    }
    int i = d.ai[e], j = d.aj[e];
    double v = d.av[e];
    const bool fi = fix[d.foff + i] != 0, fj = fix[d.foff + j] != 0;
    if (i != j) {
      if (fi || fj) v = 0.0;
    } else if (fi && fabs(v) < 1e-12) {
      v = 1.0;
    }
    int a = perm[d.poff + i], b = perm[d.poff + j];
    if (a < b) {
      int t = a;
      a = b;
      b = t;
    }
    const int I = a >> 6, J = b >> 6;
    const int t = ttab[d.toff + I * (I + 1) / 2 + J];
    if (t < 0) {
      atomicOr(err, 4);
      continue;
    }
    P[(size_t)t * TILE_ELEMS + (b & 63) * 64 + (a & 63)] = (T)v;
    if (I == J && a != b) P[(size_t)t * TILE_ELEMS + (a & 63) * 64 + (b & 63)] = (T)v;
  }
}

__global__ void k_zero_ranges(double *__restrict__ pool, const int2 *__restrict__ ranges) {
  const int2 rg = ranges[blockIdx.y];  // first tile, count
  const size_t n2 = (size_t)rg.y * (TILE_ELEMS / 2);
  double2 *T = reinterpret_cast<double2 *>(pool + (size_t)rg.x * TILE_ELEMS);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x)
    T[e] = make_double2(0.0, 0.0);
}

template <typename T>
void launch_scatter_sub(hipStream_t st, T *pool, const ScatterSub *subs, int nsub, const int *perm,
                        const char *fix, const int *ttab, const int *pads, int *err) {
  if (nsub <= 0) return;
  hipLaunchKernelGGL(k_scatter_sub<T>, dim3(64, nsub), dim3(256), 0, st, pool, subs, perm, fix, ttab,
                     pads, err);
}
template void launch_scatter_sub<double>(hipStream_t, double *, const ScatterSub *, int, const int *,
                                         const char *, const int *, const int *, int *);
template void launch_scatter_sub<float>(hipStream_t, float *, const ScatterSub *, int, const int *,
                                        const char *, const int *, const int *, int *);
void launch_zero_ranges(hipStream_t st, double *pool, const int2 *ranges, int nranges) {
  if (nranges <= 0) return;
  hipLaunchKernelGGL(k_zero_ranges, dim3(64, nranges), dim3(256), 0, st, pool, ranges);
}

template <typename T>
void launch_tile_gemm(hipStream_t st, T *pool, const TileTask *tasks, const TilePair *pairs,
                      int ntasks, int kind, int post, double alpha, double beta) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_tile_gemm<T>, dim3(ntasks), dim3(NTHR), 0, st, pool, tasks, pairs, kind,
                     post, alpha, beta, g_task_order);
}
template void launch_tile_gemm<double>(hipStream_t, double *, const TileTask *, const TilePair *,
                                       int, int, int, double, double);
template void launch_tile_gemm<float>(hipStream_t, float *, const TileTask *, const TilePair *, int,
                                      int, int, double, double);
void launch_schur_gemm(hipStream_t st, double *pool, const TileTask *tasks, const TilePair *pairs,
                       int ntasks) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_schur_gemm, dim3(ntasks), dim3(NTHR), 0, st, pool, tasks, pairs,
                     g_task_order);
}
void launch_schur_gemm(hipStream_t st, float *pool, const TileTask *tasks, const TilePair *pairs,
                       int ntasks) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_schur_gemm_f32, dim3(ntasks), dim3(NTHR), 0, st, pool, tasks, pairs,
                     g_task_order);
}
template <typename T>
void launch_tile_potrf(hipStream_t st, T *pool, const TileTask *tasks, const TilePair *pairs,
                       int ntasks, int *err) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_tile_potrf<T>, dim3(ntasks), dim3(NTHR), 0, st, pool, tasks, pairs, err,
                     g_task_order);
}
template void launch_tile_potrf<double>(hipStream_t, double *, const TileTask *, const TilePair *,
                                        int, int *);
template void launch_tile_potrf<float>(hipStream_t, float *, const TileTask *, const TilePair *, int,
                                       int *);
void launch_tile_step(hipStream_t st, double *pool, const TileTask *tasks, const TilePair *pairs,
                      int ntasks, int *err) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_tile_step, dim3(ntasks), dim3(NTHR), 0, st, pool, tasks, pairs, err,
                     g_task_order);
}
void launch_tile_mirror(hipStream_t st, double *pool, const MirrorJob *jobs, int njobs) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(k_tile_mirror, dim3(njobs), dim3(NTHR), 0, st, pool, jobs);
}
template <typename T> void launch_zero_tiles(hipStream_t st, T *pool, const int *ids, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_tiles<T>, dim3(n), dim3(256), 0, st, pool, ids);
}
template void launch_zero_tiles<double>(hipStream_t, double *, const int *, int);
template void launch_zero_tiles<float>(hipStream_t, float *, const int *, int);
void launch_scatter_ones(hipStream_t st, double *pool, const int64_t *idx, int64_t n) {
  if (n <= 0) return;
  int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_scatter_ones, dim3((unsigned)nb), dim3(256), 0, st, pool, idx, n);
}

// ------------------------------------------- symmetric tiles x column sets
// out(:, cols) = M_s in(:, cols) for every subdomain s of a batch, M_s a
// symmetric matrix in packed lower tiles (tile (I, J), I >= J, at tile0 +
// J*nt - J(J-1)/2 + I - J; element type TP, FP32 staged as FP64) and the
// columns scattered in the in / out buffers: column c of subdomain s starts
// at cols[col0 + c] (-1: a padding column, read as zero and not written).
// One workgroup per (subdomain, tile row I, 64 columns): out_I = sum_J M_IJ
// in_J on the f64 matrix cores, the M tile and the 64 x 64 input block
// staged in LDS (the upper tiles as transposes of the stored lower ones).
template <typename TP>
__device__ __forceinline__ void tile_to_lds_f64(const TP *__restrict__ t, double *lds, bool trans) {
  for (int e = threadIdx.x; e < TILE_ELEMS; e += NTHR) {
    const int row = e & 63, col = e >> 6;
    const double x = (double)t[e];
    if (!trans) lds[col * LDL + row] = x;
    else lds[row * LDL + col] = x;
  }
}
template <typename TP>
__global__ __launch_bounds__(NTHR) void k_symm_cols(const TP *__restrict__ pool,
                                                    const SymmSub *__restrict__ subs,
                                                    const SymmTask *__restrict__ tasks,
                                                    const int64_t *__restrict__ cols,
                                                    const double *__restrict__ in,
                                                    double *__restrict__ out) {
  __shared__ double As[64 * LDL];
  __shared__ double Bs[64 * LDL];
  __shared__ int64_t cl[64];
  const SymmTask t = tasks[blockIdx.x];
  const SymmSub d = subs[t.sub];
  if (threadIdx.x < 64) cl[threadIdx.x] = cols[d.col0 + t.c0 + threadIdx.x];
  typename Mfma<double>::v4 acc[2][2];
  zero_acc<double>(acc);
  for (int J = 0; J < d.nt; J++) {
    const bool low = J <= t.I;
    const int I0 = low ? t.I : J, J0 = low ? J : t.I;
    const TP *T = pool + (size_t)(d.tile0 + J0 * d.nt - J0 * (J0 - 1) / 2 + (I0 - J0)) * TILE_ELEMS;
    __syncthreads();
    // As[q][r] = M(I, J)(r, q): the stored tile itself (I >= J) or the
    // transpose of tile (J, I)
    tile_to_lds_f64(T, As, !low);
    // Bs[q][c] = in(J*64 + q, column c)
    for (int e = threadIdx.x; e < TILE_ELEMS; e += NTHR) {
      const int q = e & 63, c = e >> 6;
      const int64_t o = cl[c];
      Bs[q * LDL + c] = o >= 0 ? in[o + J * 64 + q] : 0.0;
    }
    __syncthreads();
    mfma_accumulate<double>(As, Bs, acc);
  }
  __syncthreads();
  acc_to_lds<double>(acc, As);  // As[c*LDL + r] = out(I*64 + r, column c)
  __syncthreads();
  for (int e = threadIdx.x; e < TILE_ELEMS; e += NTHR) {
    const int r = e & 63, c = e >> 6;
    const int64_t o = cl[c];
    if (o >= 0) out[o + t.I * 64 + r] = As[c * LDL + r];
  }
}
template <typename TP>
void launch_symm_cols(hipStream_t st, const TP *pool, const SymmSub *subs, const SymmTask *tasks,
                      int ntasks, const int64_t *cols, const double *in, double *out) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_symm_cols<TP>, dim3(ntasks), dim3(NTHR), 0, st, pool, subs, tasks, cols, in,
                     out);
}
template void launch_symm_cols<double>(hipStream_t, const double *, const SymmSub *, const SymmTask *,
                                       int, const int64_t *, const double *, double *);
template void launch_symm_cols<float>(hipStream_t, const float *, const SymmSub *, const SymmTask *,
                                      int, const int64_t *, const double *, double *);

}  // namespace bddc
