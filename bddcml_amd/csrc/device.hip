// device.hip -- device plan and pipelines of the MI355X BDDC path.
//
// All subdomains of the level live on one GPU as batched slots.  Setup
// (the reference's bddcml_setup_preconditioner numerics,
// src/module_levels.f90:2500-3169):
//   1. scatter the eliminated subdomain matrices into a block-sparse tile
//      pool (interior dofs first, interface last),
//   2. batched left-looking tile Cholesky of the interior block; the trailing
//      interface block becomes the Schur complement S_i
//      (dd_prepare_schur, src/module_dd.f90:4175-4296),
//   3. augmented interface matrix S'_i = S_i + C_i^T Omega C_i, its Cholesky
//      factor and inverse give K_i (interface block of [S C^T; C 0]^-1),
//      the coarse basis Phi_i and the coarse element matrix Ac_i
//      (dd_prepare_aug / dd_prepare_coarse, src/module_dd.f90:5211-5800),
//   4. coarse matrix assembly, Cholesky and inverse
//      (levels_prepare_last_level, src/module_levels.f90:3001-3169).
// Solve: PCG / BiCGSTAB on the unique interface vector with the explicit
// S_i and K_i applied as batched tile GEMVs, the interface exchange as a
// deterministic gather/scatter (src/module_levels.f90:3627-4507,
// src/module_krylov.f90:48-1505).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <utility>
#include <map>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "kernels.hpp"
#include "solver.hpp"

namespace bddc {

#define HIPCHECK(x)                                                                       \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      fatal("BDDCML_DEVICE", std::string(#x) + " failed: " + hipGetErrorString(e_));      \
  } while (0)

// Host waits on the device spin instead of sleeping: a sleeping wait woke
// milliseconds after short launches finished (about 6 ms per wait inside
// setups and solves of tens of ms).  Effective only before the process's
// device context exists; a later call is refused by the runtime and ignored.
void device_prefer_spin_wait() {
  static bool done = false;
  if (done) return;
  done = true;
  (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
  (void)hipGetLastError();
}

double wall_time() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// host seconds spent in device allocations and uploads (setup timers)
static double g_alloc_s = 0.0, g_upload_s = 0.0;

// A device array: its own allocation, or (carve) a range of another
// DBuf's allocation, which that one frees
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  bool owned = true;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), owned(o.owned) { o.p = nullptr, o.n = 0, o.owned = true; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p, n = o.n, owned = o.owned;
      o.p = nullptr, o.n = 0, o.owned = true;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = true;
  }
  // cnt elements from *off (in elements, kept 256-byte aligned) of block
  void carve(DBuf<T> &block, size_t &off, size_t cnt) {
    release();
    const size_t align = 256 / sizeof(T);
    off = (off + align - 1) / align * align;
    if (off + std::max<size_t>(cnt, 1) > block.n) {
      std::fprintf(stderr, "ERROR in DEVICE_ALLOC: carved range outside its block\n");
      std::abort();
    }
    p = block.p + off;
    n = cnt;
    owned = false;
    off += std::max<size_t>(cnt, 1);
  }
  void alloc(size_t cnt) {
    const double t = wall_time();
    release();
    n = cnt;
    HIPCHECK(hipMalloc(&p, std::max<size_t>(cnt, 1) * sizeof(T)));
    g_alloc_s += wall_time() - t;
  }
  void upload(const std::vector<T> &v, hipStream_t st) {
    alloc(v.size());
    const double t = wall_time();
    if (!v.empty()) HIPCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    g_upload_s += wall_time() - t;
  }
  void zero(hipStream_t st) {
    if (n) HIPCHECK(hipMemsetAsync(p, 0, n * sizeof(T), st));
  }
};

// A device array inside an Arena.
template <class T> struct View {
  T *p = nullptr;
  size_t n = 0;
};

// Many small plan arrays gathered into one host buffer and moved to the
// device with one allocation and one copy.  Views get their device
// addresses at commit; they must not move between add and commit.
// Pinned host staging for the plan uploads, kept for the life of the
// process and grown to the largest plan seen (tens of MB), so that uploads
// run at PCIe speed without pinning pages on every setup.  Used by one
// arena at a time; every commit waits for its copy before returning.
struct Staging {
  char *p = nullptr;
  size_t cap = 0, used = 0;
  void grow(size_t n) {
    if (n <= cap) return;
    const size_t c = std::max(n, cap + cap / 2 + (4u << 20));
    char *q = nullptr;
    HIPCHECK(hipHostMalloc((void **)&q, c, hipHostMallocDefault));
    if (used) std::memcpy(q, p, used);
    if (p) (void)hipHostFree(p);
    p = q;
    cap = c;
  }
};
static Staging &staging() {
  static Staging *s = new Staging();  // never freed: outlives the HIP runtime's teardown order
  return *s;
}

// copy of a large plan array into the staging area by all host threads
static void copy_par(char *dst, const char *src, size_t n) {
  constexpr size_t chunk = (size_t)1 << 18;
  if (n < 4 * chunk) {
    std::memcpy(dst, src, n);
    return;
  }
  const int64_t nc = (int64_t)((n + chunk - 1) / chunk);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; c++) {
    const size_t o = (size_t)c * chunk;
    std::memcpy(dst + o, src + o, std::min(chunk, n - o));
  }
}

// Plan arrays of one setup half, packed into the pinned staging area and
// sent to one device allocation with one copy; each View is fixed up to
// its place in that allocation at commit.
struct Arena {
  DBuf<char> dev;
  std::vector<std::pair<size_t, void **>> fix;
  int verbose = 0;
  bool open = false;
  size_t bytes() const { return open ? staging().used : 0; }
  template <class T, class A> void add(View<T> &v, const std::vector<T, A> &src) {
    if (verbose >= 4) std::printf("   device plan: arena %10zu bytes\n", src.size() * sizeof(T));
    Staging &h = staging();
    if (!open) h.used = 0, open = true;
    const size_t off = (h.used + 255) / 256 * 256;
    h.grow(off + src.size() * sizeof(T));
    if (!src.empty()) copy_par(h.p + off, reinterpret_cast<const char *>(src.data()), src.size() * sizeof(T));
    h.used = off + src.size() * sizeof(T);
    v.n = src.size();
    fix.push_back({off, reinterpret_cast<void **>(&v.p)});
  }
  void commit(hipStream_t st) {
    Staging &h = staging();
    const size_t n = open ? h.used : 0;
    dev.alloc(std::max<size_t>(n, 1));
    if (n) HIPCHECK(hipMemcpyAsync(dev.p, h.p, n, hipMemcpyHostToDevice, st));
    for (auto &f : fix) *f.second = dev.p + f.first;
    fix.clear();
    HIPCHECK(hipStreamSynchronize(st));  // the staging area is reused by the next arena
    h.used = 0;
    open = false;
  }
};

struct Phase {
  View<TileTask> tasks;
  View<TilePair> pairs;
  int ntask = 0;
  void upload(const TaskList &l, Arena &a) {
    ntask = (int)l.tasks.size();
    a.add(tasks, l.tasks);
    a.add(pairs, l.pairs);
  }
};

// one launch of a fixed launch sequence (tile Cholesky or GEMM-sum)
struct SeqLaunch {
  TaskList l;
  int kind = GEMM_NT, post = POST_NONE;
  double alpha = 1.0, beta = -1.0;
  bool potrf = false;
  bool step = false;  // mixed products and factorizations (k_tile_step)
};
struct SeqPhase {
  Phase ph;
  int kind = GEMM_NT, post = POST_NONE;
  double alpha = 1.0, beta = -1.0;
  bool potrf = false;
  bool step = false;
};

static inline int lowidx(int n, int I, int J) { return J * n - J * (J - 1) / 2 + (I - J); }

struct SubTiles {
  int nto = 0, nti = 0;
  TilePattern pat;   // A pattern (lower), tile ids absolute
  std::vector<std::vector<uint64_t>> col;  // column bitsets
  int base = 0, span = 0;  // first tile id and tile count of the subdomain
  // factor space (S' factor, its inverse diagonal tiles, Z, G^{-1}, Phi):
  // a pool of its own (FP64, or FP32 in the mixed-precision mode), laid out
  // per subdomain by FactorLayout from fbase
  int fbase = 0, fspan = 0;
  // slow: tile (nto, nto), the first of S's lower tiles; the interface tile
  // columns are dense, so tile (nto + I, nto + J) is slow + lowidx(nti, I, J)
  int linvA = 0, slow = 0, pbase = 0, linvP = 0, ybase = 0, ncT = 0, gtbase = 0, phibase = 0;
  // explicit interior inverse M = L11^{-1} (levels below the first): its
  // strictly lower tiles are nonzero where row I is an ancestor of column J
  // in the tile elimination tree; by interior column J, rows mrow[mptr[J] ..
  // mptr[J+1]) ascending, tile mbase + q.  The diagonal tiles are Linv_A.
  int mbase = 0;
  std::vector<int> mptr, mrow;
  int M(int I, int J) const {  // tile of M(I, J), I > J, or -1
    auto b = mrow.begin() + mptr[J], e = mrow.begin() + mptr[J + 1];
    auto it = std::lower_bound(b, e, I);
    return (it == e || *it != I) ? -1 : mbase + (int)(it - mrow.begin());
  }
  int P(int I, int J) const { return pbase + lowidx(nti, I, J); }
  int Y(int I, int cT) const { return ybase + I * ncT + cT; }
};

int TilePattern::find(int I, int J) const {
  auto b = rowidx.begin() + colptr[J], e = rowidx.begin() + colptr[J + 1];
  auto it = std::lower_bound(b, e, I);
  if (it == e || *it != I) return -1;
  return tile[it - rowidx.begin()];
}

struct Device {
  hipStream_t st = nullptr;  // compute stream
  bool borrowed_st = false;  // st belongs to the level above (second level)
  hipStream_t up = nullptr;  // uploads of the second setup half
  hipEvent_t ev[5] = {};     // setup phase boundaries (timing)
  // start and end of every GEMM launch of the interior factorization (the
  // launches that take most of the setup's device time)
  std::vector<hipEvent_t> ev_gemm;
  // around one S GEMV, one local preconditioner launch and one FP64
  // residual launch (refinement) of a solve
  hipEvent_t ev_apply[6] = {};
  bool gemv_timed = false, kapply_timed = false;
  std::vector<double> gemm_flops;  // algorithmic flops of each of them
  hipEvent_t ev_up = nullptr;
  hipEvent_t ev_rhs = nullptr;  // reduced right-hand side done (upload stream)
  bool rhs_pending = false;     // it was launched there by the setup
  DBuf<double> pool;      // subdomain tiles: pattern of A (factor, S), inverse diagonal tiles
  // the same tiles in FP32 (interior factor and S in FP32, refined solves)
  DBuf<float> ipool32;
  bool f32i = false;
  int64_t ntiles = 0;
  bool mixed = false;     // S' factor, its inverse diagonal tiles and Z in FP32
  DBuf<double> fpool;     // factor space, FP64 mode
  DBuf<float> pool32;     // factor space, mixed-precision mode
  int64_t ntilesf = 0;
  DBuf<double> cpool;     // coarse matrix: factor, inverse diagonal, W, full inverse
  int64_t ntilesc = 0;
  Arena plan2;            // task lists, job lists and constraint maps of the interface stage
  Arena plan2b;           // interface maps, assembly and sweep lists, vectors
  int nco = 0;            // constraints over all subdomains
  Arena plan3;            // coarse factorization lists (built during the interface phase)
  Arena plan4;            // residual index of the refined form (built during it too)
  DBuf<int> err;
  std::vector<SubTiles> sub;
  // setup phases
  std::vector<Phase> a_diag, a_off, p_diag, p_off, z_bwd, w_rows;
  // explicit interior inverse (levels below the first): its formation steps,
  // the four GEMV lists of the interior solves, their work vectors
  bool xint = false;
  std::vector<Phase> m_rows;
  View<GemvTask> gx_t[4];
  View<GemvPair> gx_p[4];
  int64_t gx_vt = 0;
  DBuf<double> ybuf, zbuf;
  Phase a_schur, cinv, phi, kinv;
  // explicit inverse of S': planned (W tiles in the factor space, task
  // lists, GEMV columns), formed (W = L^{-1}, W^T W: the apply is a GEMV),
  // automatic (formed once the preconditioner has been applied
  // kInverseAfterApplies times on this setup)
  bool inv_planned = false, explicit_inv = false, inv_auto = false;
  int pc_count = 0;
  std::vector<SeqPhase> cseq;  // coarse factorization and inverse
  Arena plan0;  // scatter tables and tiles to clear (uploaded first)
  Arena plan;  // task lists and job lists of the first half
  View<int> sc_perm, sc_ttab, sc_pads;  // shared per class of identically ordered subdomains
  std::vector<int> perm_off;            // offset of each subdomain's permutation in sc_perm
  View<char> sc_fix;
  View<ScatterSub> sc_subs;
  // FP64 residual of the refinement: row-segmented triplet index per class
  View<int> rs_ptr;
  View<int2> rs_ent;
  View<ResidSub> rs_subs;
  View<int> clear_ids;  // subdomain tiles cleared before the scatter
  View<int2> c_zero;                    // coarse tiles cleared before assembly
  View<int64_t> c_pads;                 // padding diagonals of the coarse matrix
  View<MirrorJob> mir_c;
  View<int> sp_sub, sp_ij, sp_dst;
  int n_sp = 0;
  // interface data
  std::vector<SubDesc> hdesc;
  // uploaded with the interface stage's arena (plan2; the interface dofs
  // and weights with plan2b)
  View<SubDesc> desc;
  View<int64_t> loff;
  View<int> NIs, map, cid, csize, cptr, cdof, cmap;
  // general constraint rows (overlapping supports or values other than 1,
  // as adaptive constraints give): row values, omega of every constraint and
  // per interface dof the (constraint, value) lists
  bool gen_constraints = false;
  View<double> cvals, comega, dval;
  View<int> dptr, dcon;
  View<double> w;
  DBuf<double> Z, Phi, G, Acs;
  int maxNI = 0, maxnc = 0, maxnti = 0, maxnt = 1, maxncT = 0;
  bool kapply_deep = false;  // form of the local preconditioner launch
  int64_t LV = 0;
  View<SymvCol> scols;
  View<SymvCol> pcols;  // the same columns over the explicit S'^{-1} tiles
  View<SymvRow> srows;
  DBuf<double> spart;  // partial slots of the symmetric S GEMV
  DBuf<double> cpart;  // partial slots of the coarse GEMV
  int nscols = 0, nsrows = 0;
  View<int> aptr;
  View<int64_t> asrc;
  View<int> cgptr, cgsrc;
  // coarse
  int ntc = 0, NC = 0, cbase = 0, clinv = 0, cwbase = 0, cfull = 0;
  View<int64_t> ca_dst, ca_src;
  View<int> ca_ptr;
  int nca = 0;
  // triangular sweeps
  std::vector<TrsvDesc> htrsv;
  View<TrsvDesc> trsv;
  View<int> rptr, rk, rt, tcptr, tck, tct;
  DBuf<double> vec;
  int64_t VT = 0;
  // vec holds the forward sweep of the current right-hand side (the
  // backward sweep of the solution overwrites it)
  bool vec_y = false;
  // refinement: local-order residual / correction, reduced right-hand side
  // of the residual, interface solution before the correction
  DBuf<double> rloc, gk, uacc;
  View<int64_t> rhs_src;
  // local dof order <-> device order of every subdomain
  std::vector<OrderDesc> hodesc;
  View<OrderDesc> odesc;
  int maxndofs = 0;
  int64_t ndofs_total = 0;
  View<double> fmod;  // right-hand side after Dirichlet elimination (rewritten in place)
  DBuf<double> sol;
  // vectors
  DBuf<double> vblock;  // the allocation the solve vectors are carved from
  DBuf<double> yl, zl, rcs, rc, uc, partial, dots;
  DBuf<double> pcg;           // PCG scalars per iteration: r.z, p.Ap, r.r
  DBuf<double> recV, recW;    // recycling basis V and W = S V (columns of length nig)
  DBuf<double> recc, recpart; // its projection coefficients and partial sums
  double *hrr = nullptr;      // pinned host copies of r.r per iteration
  double *hrr_dev = nullptr;  // the same memory as the device addresses it
  size_t hrr_n = 0;
  // PCG iteration captured as a graph (iterations from the third on):
  // replayed with the iteration index in kcount; rebuilt when the buffers
  // it was captured with or any level's preconditioner form change
  hipGraph_t pcg_graph = nullptr;
  hipGraphExec_t pcg_exec = nullptr;
  std::vector<const void *> pcg_key;
  DBuf<int> kcount;
  DBuf<double> u, r, p, z, ap, rs, v, y, s, t, g, xin, xout;
  View<char> fixmask;
  int nig = 0;
  // three levels: no dense coarse matrix here; maps to the second level's
  // device-ordered vectors (source coarse dof and weight of each entry, and
  // one entry of every coarse dof)
  bool dense_coarse = true;
  DBuf<int> l2src, l2pos;
  DBuf<double> l2w;
  int64_t l2n = 0;
};

// a second level replaces the dense coarse solve (with one second-level
// subdomain the coarse problem stays dense: the same coarse solve)
static bool has_level2(const Solver &S) { return S.nl >= 3 && S.nsub2 >= 2; }

// Assembled subdomain triplets kept in HBM from upload time on.
struct DeviceInput {
  std::vector<DBuf<int>> ai, aj;
  std::vector<DBuf<double>> av;
  DBuf<char> cmp;  // pointer tables and flags of the pattern comparison
};

// Ordering classes: the triplet pattern of each subdomain i compared with
// that of its candidate cand[i] where the uploaded copies are (HBM), instead
// of reading both patterns on the host.  eq[i] = 1 when equal.  False (eq
// untouched) when the host arrays are not the uploaded ones (a Dirichlet
// elimination added diagonal entries) or in the plan-only mode.
bool device_same_patterns(Solver &S, const std::vector<int> &cand, std::vector<char> &eq) {
  if (S.plan_only || !S.din) return false;
  DeviceInput &I = *S.din;
  const int n = S.nsub;
  if ((int)I.ai.size() != n || (int)I.aj.size() != n || (int)cand.size() != n) return false;
  std::vector<uint64_t> pa(n), pb(n);
  std::vector<int64_t> cnt(n);
  std::vector<int> c2(cand);
  for (int i = 0; i < n; i++) {
    const Sub &s = S.subs[i];
    if (I.ai[i].n != s.ai.size() || I.aj[i].n != s.aj.size() || s.ai.size() != s.aj.size()) return false;
    pa[i] = (uint64_t)(uintptr_t)I.ai[i].p;
    pb[i] = (uint64_t)(uintptr_t)I.aj[i].p;
    cnt[i] = (int64_t)s.ai.size();
  }
  eq.assign(n, 1);
  for (int i = 0; i < n; i++)
    if (c2[i] < 0 || c2[i] >= n) return false;
    else if (c2[i] != i && cnt[c2[i]] != cnt[i]) eq[i] = 0, c2[i] = i;  // lengths differ
  // one buffer: pointers, lengths, candidates, flags
  const size_t o1 = 8 * (size_t)n, o2 = 16 * (size_t)n, o3 = 24 * (size_t)n, o4 = 28 * (size_t)n,
               bytes = 32 * (size_t)n;
  if (I.cmp.n < bytes) I.cmp.alloc(bytes);
  std::vector<char> h(bytes, 0);
  std::memcpy(h.data(), pa.data(), 8 * (size_t)n);
  std::memcpy(h.data() + o1, pb.data(), 8 * (size_t)n);
  std::memcpy(h.data() + o2, cnt.data(), 8 * (size_t)n);
  std::memcpy(h.data() + o3, c2.data(), 4 * (size_t)n);
  char *d = I.cmp.p;
  HIPCHECK(hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice));
  launch_same_pattern(nullptr, reinterpret_cast<const uint64_t *>(d),
                      reinterpret_cast<const uint64_t *>(d + o1), reinterpret_cast<const int *>(d + o3),
                      reinterpret_cast<const int64_t *>(d + o2), reinterpret_cast<int *>(d + o4), n);
  HIPCHECK(hipGetLastError());
  std::vector<int> diff(n);
  HIPCHECK(hipMemcpy(diff.data(), d + o4, 4 * (size_t)n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++)
    if (diff[i]) eq[i] = 0;
  return true;
}

Solver::Solver() = default;
Solver::~Solver() = default;

void device_upload_entries(Solver &S, int is) {
  if (!S.din) {
    S.din.reset(new DeviceInput());
    S.din->ai.resize(S.nsub);
    S.din->aj.resize(S.nsub);
    S.din->av.resize(S.nsub);
  }
  const Sub &s = S.subs[is];
  DeviceInput &I = *S.din;
  I.ai[is].alloc(s.ai.size());
  I.aj[is].alloc(s.aj.size());
  I.av[is].alloc(s.av.size());
  if (!s.ai.empty()) {
    HIPCHECK(hipMemcpy(I.ai[is].p, s.ai.data(), s.ai.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(I.aj[is].p, s.aj.data(), s.aj.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(I.av[is].p, s.av.data(), s.av.size() * sizeof(double), hipMemcpyHostToDevice));
  }
}

void device_release_input(Solver &S) { S.din.reset(); }

void device_release(Solver &S) {
  if (S.lev2) {  // borrows this level's stream: released first
    device_release(*S.lev2);
    S.lev2.reset();
  }
  S.sol_on_device = false;
  S.rec_n = S.rec_max = 0;  // stored directions belong to the released operator
  if (S.dev) {
    Device &D = *S.dev;
    if (D.st) (void)hipStreamSynchronize(D.st);
    if (D.up) (void)hipStreamSynchronize(D.up);
    if (D.hrr) (void)hipHostFree(D.hrr), D.hrr = nullptr;
    if (D.pcg_exec) (void)hipGraphExecDestroy(D.pcg_exec), D.pcg_exec = nullptr;
    if (D.pcg_graph) (void)hipGraphDestroy(D.pcg_graph), D.pcg_graph = nullptr;
    for (hipEvent_t &e : D.ev)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    for (hipEvent_t &e : D.ev_gemm)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    for (hipEvent_t &e : D.ev_apply)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    if (D.ev_up) (void)hipEventDestroy(D.ev_up), D.ev_up = nullptr;
    if (D.ev_rhs) (void)hipEventDestroy(D.ev_rhs), D.ev_rhs = nullptr;
    if (D.st && !D.borrowed_st) (void)hipStreamDestroy(D.st);
    if (D.up) (void)hipStreamDestroy(D.up);
    D.st = D.up = nullptr;
    S.dev.reset();
  }
}

// ------------------------------------------------------- symbolic planning
static void bs_set(std::vector<uint64_t> &b, int i) { b[i >> 6] |= (1ull << (i & 63)); }
static bool bs_get(const std::vector<uint64_t> &b, int i) { return (b[i >> 6] >> (i & 63)) & 1ull; }

// Block-sparse lower pattern of the device-ordered subdomain matrix and its
// fill under elimination of the interior tile columns.
static void symbolic(const Sub &s, SubTiles &T) {
  T.nto = s.nto;
  T.nti = s.nti;
  const int nt = s.nto + s.nti, nw = (nt + 63) / 64;
  T.col.assign(nt, std::vector<uint64_t>(nw, 0ull));
  for (int J = 0; J < nt; J++) bs_set(T.col[J], J);
  for (size_t k = 0; k < s.av.size(); k++) {
    int a = s.perm[s.ai[k]], b = s.perm[s.aj[k]];
    int I = std::max(a, b) / TILE, J = std::min(a, b) / TILE;
    bs_set(T.col[J], I);
  }
  for (int J = s.nto; J < nt; J++)
    for (int I = J; I < nt; I++) bs_set(T.col[J], I);
  for (int J = 0; J < s.nto; J++) {
    std::vector<int> R;
    for (int I = J + 1; I < nt; I++)
      if (bs_get(T.col[J], I)) R.push_back(I);
    for (size_t a = 0; a < R.size(); a++)
      for (size_t b = a; b < R.size(); b++) bs_set(T.col[R[a]], R[b]);
  }
}

static void assign_tiles(SubTiles &T, int &next, bool explicit_interior) {
  const int nt = T.nto + T.nti;
  T.base = next;
  T.pat.nt = nt;
  T.pat.ntf = T.nto;
  T.pat.colptr.assign(nt + 1, 0);
  T.pat.rowidx.clear();
  T.pat.tile.clear();
  for (int J = 0; J < nt; J++) {
    for (int I = J; I < nt; I++)
      if (bs_get(T.col[J], I)) {
        T.pat.rowidx.push_back(I);
        T.pat.tile.push_back(next++);
      }
    T.pat.colptr[J + 1] = (int)T.pat.rowidx.size();
  }
  T.linvA = next;
  next += T.nto;
  T.mptr.assign(T.nto + 1, 0);
  T.mrow.clear();
  if (explicit_interior) {
    // ancestors of each interior tile column in the elimination tree (the
    // parent: the first interior row below the diagonal of the filled column)
    std::vector<int> parent(T.nto, -1);
    for (int J = 0; J < T.nto; J++)
      for (int I = J + 1; I < T.nto; I++)
        if (bs_get(T.col[J], I)) {
          parent[J] = I;
          break;
        }
    for (int J = 0; J < T.nto; J++) {
      for (int I = parent[J]; I >= 0; I = parent[I]) T.mrow.push_back(I);
      T.mptr[J + 1] = (int)T.mrow.size();
    }
    T.mbase = next;
    next += (int)T.mrow.size();
  }
  T.slow = T.pat.find(T.nto, T.nto);
  T.span = next - T.base;
}

// Factor space of one subdomain relative to its first tile: P (S' factor,
// packed lower), LinvP (inverse diagonal tiles), Y (Z = S'^{-1} C^T, nti x
// ncT), Gt (G^{-1}, ncT x ncT), Phi (nti x ncT); with the explicit inverse,
// W (the strictly lower tiles of L^{-1}, packed by columns), after which P
// holds S'^{-1} = W^T W (lower tiles)
struct FactorLayout {
  int p, linv, y, gt, phi, w, span;
  FactorLayout(int nti, int ncT, bool explicit_inv) {
    p = 0;
    linv = nti * (nti + 1) / 2;
    y = linv + nti;
    gt = y + nti * ncT;
    phi = gt + ncT * ncT;
    w = phi + nti * ncT;
    span = w + (explicit_inv ? nti * (nti - 1) / 2 : 0);
  }
  // strictly lower tile (I, J), I > J, of W
  int W(int n, int I, int J) const { return w + J * (n - 1) - J * (J - 1) / 2 + (I - J - 1); }
};

static void assign_factor(SubTiles &T, int nc, int64_t &nf, bool explicit_inv) {
  T.ncT = (nc + TILE - 1) / TILE;
  const FactorLayout L(T.nti, T.ncT, explicit_inv);
  T.fbase = (int)nf;
  T.fspan = L.span;
  T.pbase = T.fbase + L.p;
  T.linvP = T.fbase + L.linv;
  T.ybase = T.fbase + L.y;
  T.gtbase = T.fbase + L.gt;
  T.phibase = T.fbase + L.phi;
  nf += L.span;
}

// tasks of a dense lower Cholesky on tiles P(I,J) with inverse tiles Linv(J).
// The diagonal update of column J by the columns k < J - 1 runs as a task of
// the product launch of column J - 1 (first in its list; those tiles are
// final before it), so that the factorization launch of column J, one
// workgroup per diagonal tile, adds only the pair of column J - 1.
template <class FP, class FL>
static void dense_chol_tasks(int n, FP Ptile, FL Linv, std::vector<TaskList> &diag,
                             std::vector<TaskList> &off) {
  if ((int)diag.size() < n) diag.resize(n), off.resize(n);
  for (int J = 0; J < n; J++) {
    TileTask t{Ptile(J, J), Linv(J), (int)diag[J].pairs.size(), 0};
    if (J >= 2) {
      TaskList &P = off[J - 1];
      TileTask e{Ptile(J, J), -1, (int)P.pairs.size(), 0};
      for (int k = 0; k < J - 1; k++) P.pairs.push_back({Ptile(J, k), Ptile(J, k)});
      e.p1 = (int)P.pairs.size();
      e.mode = task_mode(GEMM_NT, POST_NONE, true);
      P.tasks.insert(P.tasks.begin(), e);
    }
    for (int k = std::max(0, J - 1); k < J; k++) diag[J].pairs.push_back({Ptile(J, k), Ptile(J, k)});
    t.p1 = (int)diag[J].pairs.size();
    diag[J].tasks.push_back(t);
    for (int I = J + 1; I < n; I++) {
      TileTask u{Ptile(I, J), Linv(J), (int)off[J].pairs.size(), 0};
      for (int k = 0; k < J; k++) off[J].pairs.push_back({Ptile(I, k), Ptile(J, k)});
      u.p1 = (int)off[J].pairs.size();
      off[J].tasks.push_back(u);
    }
  }
}

// Host-side bounds check of every tile id a task list hands to the kernels.
static void check_tasks(const TaskList &l, int64_t ntiles, bool uses_x, const char *what) {
  auto bad = [&](int id) { return id < 0 || id >= ntiles; };
  for (const TileTask &t : l.tasks) {
    const bool need_x =
        uses_x || t.mode == kTaskPotrf || (t.mode >= 0 && ((t.mode >> 2) & 3) != 0);
    if (bad(t.c) || (need_x && bad(t.x)) || t.p0 < 0 || t.p1 < t.p0 || t.p1 > (int)l.pairs.size())
      fatal("BDDCML_DEVICE_PLAN", std::string("invalid task in ") + what);
  }
  for (const TilePair &p : l.pairs)
    if (bad(p.a) || bad(p.b)) fatal("BDDCML_DEVICE_PLAN", std::string("invalid tile pair in ") + what);
}

static void upload_phases(std::vector<Phase> &dst, const std::vector<TaskList> &src, Arena &a) {
  dst.resize(src.size());
  for (size_t k = 0; k < src.size(); k++) dst[k].upload(src[k], a);
}

// Makes every tile id of a class list relative to the subdomain's first
// tile b0; false if an id leaves [b0, b0 + nb) or a pair range is invalid.
static bool relative_ids(TaskList &L, int b0, int nb) {
  auto inside = [&](int id) { return id >= b0 && id < b0 + nb; };
  bool ok = true;
  for (TileTask &t : L.tasks) {
    if (!inside(t.c) || (t.x != -1 && !inside(t.x)) || t.p0 < 0 || t.p1 < t.p0 ||
        t.p1 > (int)L.pairs.size())
      ok = false;
    t.c -= b0;
    if (t.x >= 0) t.x -= b0;
  }
  for (TilePair &p : L.pairs) {
    if (!inside(p.a) || !inside(p.b)) ok = false;
    p.a -= b0;
    p.b -= b0;
  }
  return ok;
}

// One launch over every subdomain from the class lists (tile ids relative
// to the subdomain's base): the pairs once per class, the tasks per
// subdomain with its base added, subdomain-major.
template <class Q, class Get, class Base>
static TaskList materialize(const std::vector<Q> &sts, Get get, const std::vector<int> &cls,
                            Base base_of) {
  const int ncls = (int)sts.size(), nsub = (int)cls.size();
  TaskList out;
  std::vector<size_t> poff(ncls + 1, 0), toff(nsub + 1, 0);
  for (int ic = 0; ic < ncls; ic++) {
    const TaskList *L = get(sts[ic]);
    poff[ic + 1] = poff[ic] + (L ? L->pairs.size() : 0);
  }
  for (int is = 0; is < nsub; is++) {
    const TaskList *L = get(sts[cls[is]]);
    toff[is + 1] = toff[is] + (L ? L->tasks.size() : 0);
  }
  out.pairs.resize(poff[ncls]);
  out.tasks.resize(toff[nsub]);
  for (int ic = 0; ic < ncls; ic++)
    if (const TaskList *L = get(sts[ic]))
      std::copy(L->pairs.begin(), L->pairs.end(), out.pairs.begin() + poff[ic]);
#pragma omp parallel for schedule(static)
  for (int is = 0; is < nsub; is++) {
    const TaskList *L = get(sts[cls[is]]);
    if (!L) continue;
    const int b = base_of(is), po = (int)poff[cls[is]];
    TileTask *o = out.tasks.data() + toff[is];
    for (const TileTask &t : L->tasks)
      *o++ = TileTask{t.c + b, t.x >= 0 ? t.x + b : -1, t.p0 + po, t.p1 + po, b, t.fresh, t.mode};
  }
  return out;
}

// the same for every step of a multi-step member (steps of different
// classes aligned by index)
template <class Q, class Base>
static std::vector<TaskList> merge_steps(const std::vector<Q> &sts, std::vector<TaskList> Q::*member,
                                         const std::vector<int> &cls, Base base_of) {
  size_t nsteps = 0;
  for (auto &q : sts) nsteps = std::max(nsteps, (q.*member).size());
  std::vector<TaskList> out(nsteps);
  for (size_t k = 0; k < nsteps; k++)
    out[k] = materialize(
        sts,
        [&](const Q &q) -> const TaskList * { return k < (q.*member).size() ? &(q.*member)[k] : nullptr; },
        cls, base_of);
  return out;
}

// algorithmic work: 2*64^3 flops per tile product, 64^3/3 per tile Cholesky
// plus 64^3/3 for its inverse, 2*64^3 per post multiply; pair lists are
// shared between subdomains, so the pairs of each task are counted
static const double kTileGemm = 2.0 * 64 * 64 * 64, kTileChol = 64.0 * 64 * 64 / 3.0;
static double tile_products(const TaskList &l) {
  double n = 0;
  for (const TileTask &t : l.tasks) n += t.p1 - t.p0;
  return n;
}
// tasks that end with a post multiply or a factorization (a task of its own
// mode may have neither)
static double posted_tasks(const TaskList &l) {
  double n = 0;
  for (const TileTask &t : l.tasks) n += t.mode < 0 || t.mode == kTaskPotrf || ((t.mode >> 2) & 3) != 0;
  return n;
}
static double tile_work(const std::vector<TaskList> &v, double per_task) {
  double f = 0;
  for (auto &l : v) f += kTileGemm * tile_products(l) + per_task * posted_tasks(l);
  return f;
}

// ------------------------------------------------------------------ setup
// Setup runs in three stages.  device_setup_begin needs only the interface
// split and the orderings: it builds the subdomain tile layout and the task
// lists of the interior factorization and the Schur complements, and
// launches them without waiting.  The host then classifies globs and builds
// the constraints while the device factors; device_setup_interface lays out
// the factor space and the coarse tiles (pools of their own, sized by the
// constraints), builds their task lists and the constraint maps (uploaded
// on a second stream) and launches the interface phase and the coarse
// problem; the host builds the interface numbering, Dirichlet data, weights
// and the condensation right-hand side meanwhile, and device_setup_finish
// adds the solve maps and lists, queues the reduced right-hand side and
// waits.
void device_setup_begin(Solver &S) {
  device_fetch_solution(S);  // a new setup keeps the last solution as the initial guess
  device_release(S);
  S.dev.reset(new Device());
  Device &D = *S.dev;
  HIPCHECK(hipStreamCreateWithFlags(&D.st, hipStreamNonBlocking));
  HIPCHECK(hipStreamCreateWithFlags(&D.up, hipStreamNonBlocking));
  hipStream_t st = D.st;
  const int nsub = S.nsub;
  double t0 = wall_time();
  double tph = t0;
  auto ph = [&](const char *name) {
    double now = wall_time();
    if (S.verbose >= 2) std::printf("   device plan: %-24s %8.4f s\n", name, now - tph);
    tph = now;
  };

  // ---- symbolic patterns and tile ids
  D.plan.verbose = S.verbose;
  D.sub.resize(nsub);
  // symbolic factorization once per class of identical subdomains
  if ((int)S.ord_cls.size() != nsub) fatal("BDDCML_DEVICE_PLAN", "orderings missing");
  D.f32i = S.refine != 0;
  D.mixed = S.precision == 32 || D.f32i;
  D.xint = S.explicit_interior;
  // one tile layout per class (ids relative to the subdomain's first tile),
  // then every subdomain takes its class's layout at its own base
  const int nrep0 = (int)S.ord_reps.size();
  std::vector<SubTiles> layout(nrep0);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nrep0; k++) {
    symbolic(S.subs[S.ord_reps[k]], layout[k]);
    int rel = 0;
    assign_tiles(layout[k], rel, D.xint);
    layout[k].col.clear();
    layout[k].col.shrink_to_fit();
  }
  std::vector<int64_t> sbase(nsub + 1, 0);
  for (int is = 0; is < nsub; is++) sbase[is + 1] = sbase[is] + layout[S.ord_cls[is]].span;
  if (sbase[nsub] >= INT32_MAX) fatal("BDDCML_DEVICE_PLAN", "too many subdomain tiles");
#pragma omp parallel for schedule(static)
  for (int is = 0; is < nsub; is++) {
    SubTiles &T = D.sub[is];
    T = layout[S.ord_cls[is]];
    const int b = (int)sbase[is];
    T.base = b;
    for (int &t : T.pat.tile) t += b;
    T.linvA += b;
    T.mbase += b;
    if (T.slow >= 0) T.slow += b;
  }
  D.ntiles = sbase[nsub];

  ph("symbolic");
  // ---- the tile pool is allocated on a helper thread while the host builds
  // the task lists and tables below (a large hipMalloc takes milliseconds);
  // joined before the first launch, also on error paths.
  struct PoolAlloc {
    std::thread th;
    std::exception_ptr err;
    ~PoolAlloc() {
      if (th.joinable()) th.join();
    }
    void wait() {
      if (th.joinable()) th.join();
      if (err) std::rethrow_exception(std::exchange(err, nullptr));
    }
  } pool_alloc;
  {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    pool_alloc.th = std::thread([&D, &pool_alloc, dev] {
      try {
        HIPCHECK(hipSetDevice(dev));
        if (D.f32i) D.ipool32.alloc((size_t)D.ntiles * TILE_ELEMS);
        else D.pool.alloc((size_t)D.ntiles * TILE_ELEMS);
      } catch (...) {
        pool_alloc.err = std::current_exception();
      }
    });
  }
  // ---- task lists of the interior factorization and the Schur complements.
  // Identically ordered subdomains have identical tile patterns: each class
  // shares one list of tile pairs, kept relative to the subdomain's first
  // tile; the tasks (output tile, inverse tile, pair range, tile base) are
  // materialised per subdomain, subdomain-major within each step.
  const std::vector<int> &cls = S.ord_cls, &reps = S.ord_reps;
  const int ncls = (int)reps.size();
  struct SubTasks {
    std::vector<TaskList> a_diag, a_off;
    std::vector<TaskList> m_rows;  // explicit interior inverse, one step per tile row
    TaskList a_schur;
    std::vector<int> clear;  // pattern tiles (relative) that start as zeros
  };
  std::vector<SubTasks> sts(ncls);
  int bad_ids = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad_ids)
  for (int ic = 0; ic < ncls; ic++) {
    const SubTiles &T = D.sub[reps[ic]];
    SubTasks &Q = sts[ic];
    const int nto = T.nto, nt = T.nto + T.nti;
    // tile id of (I, J), I >= J, or -1: a dense table (the lookups below run
    // once per tile pair)
    std::vector<int> tid((size_t)nt * nt, -1);
    for (int J = 0; J < nt; J++)
      for (int q = T.pat.colptr[J]; q < T.pat.colptr[J + 1]; q++)
        tid[(size_t)T.pat.rowidx[q] * nt + J] = T.pat.tile[q];
    auto find = [&](int I, int J) { return tid[(size_t)I * nt + J]; };
    // row sets restricted to factored columns
    std::vector<std::vector<int>> rowk(nt);
    for (int J = 0; J < nto; J++)
      for (int q = T.pat.colptr[J]; q < T.pat.colptr[J + 1]; q++) rowk[T.pat.rowidx[q]].push_back(J);
    // level of each factored tile column in the tile elimination tree:
    // columns of one level depend only on lower levels and share a step
    std::vector<int> lvl(nto, 0);
    int nlev = 0;
    for (int J = 0; J < nto; J++) {
      for (int k : rowk[J])
        if (k < J) lvl[J] = std::max(lvl[J], lvl[k] + 1);
      nlev = std::max(nlev, lvl[J] + 1);
    }
    Q.a_diag.resize(nlev);
    Q.a_off.resize(nlev);
    // pairs (L(I,k), L(J,k)) over the columns k < kmax both rows reach
    // (and that `keep` accepts)
    auto common = [&](int I, int J, int kmax, TaskList &L, auto keep) {
      const auto &a = rowk[I], &b = rowk[J];
      size_t x = 0, y = 0;
      while (x < a.size() && y < b.size()) {
        if (a[x] >= kmax || b[y] >= kmax) break;
        if (a[x] < b[y]) x++;
        else if (b[y] < a[x]) y++;
        else {
          if (keep(a[x])) L.pairs.push_back({find(I, a[x]), find(J, b[y])});
          x++, y++;
        }
      }
    };
    auto all = [](int) { return true; };
    // The diagonal update of a column at level l by columns of levels below
    // l - 1 runs in the product launch of level l - 1 (those tiles are final
    // before it; the tasks go first in that list); the factorization launch
    // adds only the columns of level l - 1.  `split` records the
    // factorization tasks that then continue a partial sum.
    std::vector<std::vector<TileTask>> early(nlev);
    std::vector<std::pair<int, int>> split;
    for (int J = 0; J < nto; J++) {
      const int lv = lvl[J];
      TaskList &Dg = Q.a_diag[lv], &Of = Q.a_off[lv];
      TileTask t{find(J, J), T.linvA + J, (int)Dg.pairs.size(), 0};
      if (lv >= 1) {
        TaskList &P = Q.a_off[lv - 1];
        TileTask e{find(J, J), -1, (int)P.pairs.size(), 0};
        common(J, J, J, P, [&](int k) { return lvl[k] < lv - 1; });
        e.p1 = (int)P.pairs.size();
        e.mode = task_mode(GEMM_NT, POST_NONE, true);
        if (e.p1 > e.p0) {
          early[lv - 1].push_back(e);
          split.push_back({lv, (int)Dg.tasks.size()});
        }
        common(J, J, J, Dg, [&](int k) { return lvl[k] == lv - 1; });
      } else {
        common(J, J, J, Dg, all);
      }
      t.p1 = (int)Dg.pairs.size();
      Dg.tasks.push_back(t);
      for (int q = T.pat.colptr[J]; q < T.pat.colptr[J + 1]; q++) {
        int I = T.pat.rowidx[q];
        if (I == J) continue;
        TileTask u{T.pat.tile[q], T.linvA + J, (int)Of.pairs.size(), 0};
        common(I, J, J, Of, all);
        u.p1 = (int)Of.pairs.size();
        Of.tasks.push_back(u);
      }
    }
    for (int l = 0; l < nlev; l++)
      Q.a_off[l].tasks.insert(Q.a_off[l].tasks.begin(), early[l].begin(), early[l].end());
    for (int J = nto; J < nt; J++)
      for (int I = J; I < nt; I++) {
        TileTask t{find(I, J), -1, (int)Q.a_schur.pairs.size(), 0};
        common(I, J, nto, Q.a_schur, all);
        t.p1 = (int)Q.a_schur.pairs.size();
        if (t.p1 > t.p0) Q.a_schur.tasks.push_back(t);
      }
    if (D.xint) {
      // M = L11^{-1} row by row: M(I, J) = -Linv_I sum_{J <= K < I} L(I, K) M(K, J),
      // M(J, J) = Linv_J; every task starts from zero
      Q.m_rows.resize(nto);
      for (int I = 0; I < nto; I++)
        for (int J = 0; J < I; J++) {
          const int c = T.M(I, J);
          if (c < 0) continue;
          TaskList &L = Q.m_rows[I];
          TileTask t{c, T.linvA + I, (int)L.pairs.size(), 0};
          t.fresh = 1;
          for (int K = J; K < I; K++) {
            const int lik = find(I, K);
            const int mkj = K == J ? T.linvA + J : T.M(K, J);
            if (lik >= 0 && mkj >= 0) L.pairs.push_back({lik, mkj});
          }
          t.p1 = (int)L.pairs.size();
          L.tasks.push_back(t);
        }
      for (TaskList &L : Q.m_rows)
        if (!relative_ids(L, T.base, T.span)) bad_ids |= 1;
    }
    // Tiles without matrix entries whose task writes them whole are
    // "fresh": the task starts from zero and the tile is never cleared.
    // Entry tiles (and the padding diagonals) are cleared and scattered;
    // tiles no task writes are cleared.
    {
      const Sub &s = S.subs[reps[ic]];
      const int npat = (int)T.pat.tile.size();
      std::vector<char> entry(npat, 0), written(npat, 0);
      auto mark = [&](int a, int b) {
        const int I = std::max(a, b) / TILE, J = std::min(a, b) / TILE;
        const int q = find(I, J) - T.base;
        if (q < 0 || q >= npat) bad_ids |= 1;
        else entry[q] = 1;
      };
      for (size_t k = 0; k < s.ai.size(); k++) mark(s.perm[s.ai[k]], s.perm[s.aj[k]]);
      for (int p : s.pad) mark(p, p);
      for (int p = s.NO + s.ni; p < s.NO + s.NI; p++) mark(p, p);
      auto visit = [&](TaskList &L) {
        for (TileTask &t : L.tasks) {
          const int q = t.c - T.base;
          if (q < 0 || q >= npat) {
            bad_ids |= 1;
            continue;
          }
          written[q] = 1;
          t.fresh = entry[q] ? 0 : 1;
        }
      };
      for (auto *v : {&Q.a_diag, &Q.a_off})
        for (TaskList &L : *v) visit(L);
      visit(Q.a_schur);
      // a factorization after an early diagonal update reads that sum
      for (auto &sp : split) Q.a_diag[sp.first].tasks[sp.second].fresh = 0;
      for (int q = 0; q < npat; q++)
        if (entry[q] || !written[q]) Q.clear.push_back(q);
    }
    for (auto *v : {&Q.a_diag, &Q.a_off})
      for (TaskList &L : *v)
        if (!relative_ids(L, T.base, T.span)) bad_ids |= 1;
    if (!relative_ids(Q.a_schur, T.base, T.span)) bad_ids |= 1;
  }
  if (bad_ids) fatal("BDDCML_DEVICE_PLAN", "tile task outside its subdomain's tiles");
  if (S.verbose >= 3) ph("  class task lists");
  for (int is = 0; is < nsub; is++) {
    const SubTiles &T = D.sub[is], &R = D.sub[reps[cls[is]]];
    if ((int64_t)T.base + T.span > D.ntiles || T.span != R.span)
      fatal("BDDCML_DEVICE_PLAN", "tile range");
  }
  {
    // tiles to clear before the scatter, per subdomain
    std::vector<size_t> off(nsub + 1, 0);
    for (int is = 0; is < nsub; is++) off[is + 1] = off[is] + sts[cls[is]].clear.size();
    std::vector<int> ids(off[nsub]);
#pragma omp parallel for schedule(static)
    for (int is = 0; is < nsub; is++) {
      const std::vector<int> &c = sts[cls[is]].clear;
      for (size_t k = 0; k < c.size(); k++) ids[off[is] + k] = D.sub[is].base + c[k];
    }
    for (int id : ids)
      if (id < 0 || id >= D.ntiles) fatal("BDDCML_DEVICE_PLAN", "invalid tile to clear");
    D.plan0.add(D.clear_ids, ids);
  }
  // ---- matrix entries: the assembled triplets are already on the device
  // (uploaded with each subdomain); here only the device permutation, the
  // Dirichlet flags and the tile tables go up.  The permutation, the packed
  // lower tile table (tile ids relative to the subdomain's first tile) and
  // the padding diagonals (1.0) are shared by each class of identically
  // ordered subdomains; the Dirichlet flags are per subdomain.
  if (!S.din)
    for (int is = 0; is < nsub; is++) device_upload_entries(S, is);
  {
    const std::vector<int> &ocls = S.ord_cls, &oreps = S.ord_reps;
    const int nrep = (int)oreps.size();
    std::vector<size_t> foff(nsub + 1, 0), poff(nrep + 1, 0), toff(nrep + 1, 0), doff(nrep + 1, 0);
    for (int is = 0; is < nsub; is++) foff[is + 1] = foff[is] + S.subs[is].ndofs;
    for (int k = 0; k < nrep; k++) {
      const Sub &s = S.subs[oreps[k]];
      const SubTiles &T = D.sub[oreps[k]];
      const size_t nt = T.nto + T.nti;
      poff[k + 1] = poff[k] + s.ndofs;
      toff[k + 1] = toff[k] + nt * (nt + 1) / 2;
      doff[k + 1] = doff[k] + s.pad.size() + (s.NI - s.ni);
    }
    std::vector<int> perm(poff[nrep]), ttab(toff[nrep], -1), pads(doff[nrep]);
    std::vector<char> fix(foff[nsub]);
    std::vector<ScatterSub> ss(nsub);
    if (S.verbose >= 3) ph("  scatter allocation");
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int k = 0; k < nrep; k++) {
      const Sub &s = S.subs[oreps[k]];
      const SubTiles &T = D.sub[oreps[k]];
      const int nt = T.nto + T.nti;
      if ((int64_t)T.pat.tile.size() * TILE_ELEMS > INT32_MAX) bad |= 4;
      for (int d = 0; d < s.ndofs; d++) {
        perm[poff[k] + d] = s.perm[d];
        if (s.perm[d] < 0 || s.perm[d] >= s.NO + s.NI) bad |= 2;
      }
      int *tt = ttab.data() + toff[k];
      for (int J = 0; J < nt; J++)
        for (int q = T.pat.colptr[J]; q < T.pat.colptr[J + 1]; q++) {
          int I = T.pat.rowidx[q];
          tt[(size_t)I * (I + 1) / 2 + J] = T.pat.tile[q] - T.base;
        }
      int *pd = pads.data() + doff[k];
      auto put_diag = [&](int p) {
        int tid = T.pat.find(p / TILE, p / TILE) - T.base;
        *pd++ = tid * TILE_ELEMS + (p % TILE) * TILE + (p % TILE);
      };
      for (int p : s.pad) put_diag(p);
      for (int p = s.NO + s.ni; p < s.NO + s.NI; p++) put_diag(p);
    }
    D.perm_off.assign(nsub, 0);
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int is = 0; is < nsub; is++) {
      const Sub &s = S.subs[is];
      const SubTiles &T = D.sub[is];
      const int c = ocls[is];
      ss[is] = ScatterSub{S.din->ai[is].p, S.din->aj[is].p, S.din->av[is].p,
                          (int)S.din->ai[is].n, (int)foff[is], (int)poff[c], (int)toff[c],
                          (int)doff[c], (int)(doff[c + 1] - doff[c]), T.base};
      D.perm_off[is] = (int)poff[c];
      if ((int64_t)S.din->ai[is].n != (int64_t)s.ai.size()) bad |= 1;
      if (T.base != T.pat.tile.front() || s.ndofs != S.subs[oreps[c]].ndofs) bad |= 4;
      if ((int64_t)T.base + (int64_t)T.pat.tile.size() > D.ntiles) bad |= 4;
      for (int d = 0; d < s.ndofs; d++) fix[foff[is] + d] = s.ifix[d] != 0;
    }
    if (S.verbose >= 3) ph("  scatter tables");
    if (bad & 1) fatal("BDDCML_DEVICE_PLAN", "device triplets out of date");
    if (bad & 2) fatal("BDDCML_DEVICE_PLAN", "invalid permutation");
    if (bad & 4) fatal("BDDCML_DEVICE_PLAN", "invalid subdomain tile range");
    for (int k = 0; k < nrep; k++) {
      const int span = (int)D.sub[oreps[k]].pat.tile.size();
      for (size_t q = toff[k]; q < toff[k + 1]; q++)
        if (ttab[q] >= span) fatal("BDDCML_DEVICE_PLAN", "invalid tile table entry");
      for (size_t q = doff[k]; q < doff[k + 1]; q++)
        if (pads[q] < 0 || pads[q] >= span * TILE_ELEMS) fatal("BDDCML_DEVICE_PLAN", "invalid padding entry");
    }
    if (S.verbose >= 3) ph("  scatter checks");
    D.plan0.add(D.sc_perm, perm);
    D.plan0.add(D.sc_fix, fix);
    D.plan0.add(D.sc_ttab, ttab);
    D.plan0.add(D.sc_pads, pads);
    D.plan0.add(D.sc_subs, ss);
  }

  ph("scatter entries");
  // the scatter's tables go up first, in an arena of their own, and the
  // tiles are cleared and scattered while the host materializes the task
  // lists of the factorization
  D.plan0.commit(st);
  pool_alloc.wait();
  D.err.alloc(1);
  D.err.zero(st);
  hipEvent_t *ev = D.ev;
  for (int k = 0; k < 5; k++) HIPCHECK(hipEventCreate(&ev[k]));
  HIPCHECK(hipEventRecord(ev[0], st));
  if (D.f32i) {
    launch_zero_tiles(st, D.ipool32.p, D.clear_ids.p, (int)D.clear_ids.n);
    launch_scatter_sub(st, D.ipool32.p, D.sc_subs.p, nsub, D.sc_perm.p, D.sc_fix.p, D.sc_ttab.p,
                       D.sc_pads.p, D.err.p);
  } else {
    launch_zero_tiles(st, D.pool.p, D.clear_ids.p, (int)D.clear_ids.n);
    launch_scatter_sub(st, D.pool.p, D.sc_subs.p, nsub, D.sc_perm.p, D.sc_fix.p, D.sc_ttab.p,
                       D.sc_pads.p, D.err.p);
  }
  HIPCHECK(hipGetLastError());
  ph("scatter launched");
  auto base_of = [&](int is) { return D.sub[is].base; };
  std::vector<TaskList> a_diag = merge_steps(sts, &SubTasks::a_diag, cls, base_of),
                        a_off = merge_steps(sts, &SubTasks::a_off, cls, base_of);
  TaskList a_schur = materialize(
      sts, [](const SubTasks &Q) -> const TaskList * { return &Q.a_schur; }, cls, base_of);
  std::vector<TaskList> m_rows;
  if (D.xint) m_rows = merge_steps(sts, &SubTasks::m_rows, cls, base_of);
  sts.clear();
  if (S.verbose >= 3) ph("  materialized");
  if (S.verbose >= 2)
    std::printf("   device plan: %d structure classes over %d subdomains, %zu interior steps\n", ncls,
                nsub, a_diag.size());
  upload_phases(D.a_diag, a_diag, D.plan);
  upload_phases(D.a_off, a_off, D.plan);
  D.a_schur.upload(a_schur, D.plan);
  if (D.xint) {
    upload_phases(D.m_rows, m_rows, D.plan);
    // the four GEMVs of the interior solves, per subdomain (absolute tile
    // ids, offsets into the device-ordered subdomain vectors):
    // 0: y1 = M f1; 1: f2 -= L21 y1; 2: z1 = y1 - L21^T u2; 3: u1 = M^T z1
    std::vector<GemvTask> gt[4];
    std::vector<GemvPair> gp[4];
    int64_t vo = 0;
    for (int is = 0; is < nsub; is++) {
      const SubTiles &T = D.sub[is];
      const int nto = T.nto, nt = T.nto + T.nti;
      auto task = [&](int k, int row) {
        gt[k].push_back(GemvTask{vo + 64 * row, vo + 64 * row, (int)gp[k].size(), 0});
      };
      auto pair = [&](int k, int tile, int col) { gp[k].push_back(GemvPair{tile, 0, vo + 64 * col}); };
      auto close = [&](int k) { gt[k].back().p1 = (int)gp[k].size(); };
      for (int I = 0; I < nto; I++) {
        task(0, I);
        for (int J = 0; J < I; J++)
          if (T.M(I, J) >= 0) pair(0, T.M(I, J), J);
        pair(0, T.linvA + I, I);
        close(0);
      }
      for (int I = nto; I < nt; I++) {
        task(1, I);
        for (int K = 0; K < nto; K++)
          if (T.pat.find(I, K) >= 0) pair(1, T.pat.find(I, K), K);
        close(1);
      }
      for (int J = 0; J < nto; J++) {
        task(2, J);
        for (int I = nto; I < nt; I++)
          if (T.pat.find(I, J) >= 0) pair(2, T.pat.find(I, J), I);
        close(2);
        task(3, J);
        pair(3, T.linvA + J, J);
        for (int q = T.mptr[J]; q < T.mptr[J + 1]; q++) pair(3, T.mbase + q, T.mrow[q]);
        close(3);
      }
      vo += (int64_t)S.subs[is].NO + S.subs[is].NI;
    }
    for (int k = 0; k < 4; k++) {
      for (const GemvPair &p : gp[k])
        if (p.tile < 0 || p.tile >= D.ntiles || p.xoff < 0 || p.xoff + 64 > vo)
          fatal("BDDCML_DEVICE_PLAN", "interior GEMV pair out of range");
      for (const GemvTask &t : gt[k])
        if (t.yoff < 0 || t.yoff + 64 > vo) fatal("BDDCML_DEVICE_PLAN", "interior GEMV task out of range");
      D.plan.add(D.gx_t[k], gt[k]);
      D.plan.add(D.gx_p[k], gp[k]);
    }
    D.gx_vt = vo;
  }

  ph("task lists");
  // S stays in its lower tiles, read by the symmetric GEMV and by S'
  int bad_s = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad_s)
  for (int k = 0; k < ncls; k++) {
    const SubTiles &T = D.sub[reps[k]];
    for (int J = 0; J < T.nti; J++)
      for (int I = J; I < T.nti; I++)
        if (T.pat.find(T.nto + I, T.nto + J) != T.slow + lowidx(T.nti, I, J)) bad_s = 1;
  }
  for (int is = 0; is < nsub; is++) {
    const SubTiles &T = D.sub[is];
    const SubTiles &R = D.sub[reps[cls[is]]];
    if (T.slow - T.base != R.slow - R.base || T.slow < 0 ||
        (int64_t)T.slow + (int64_t)T.nti * (T.nti + 1) / 2 > D.ntiles)
      bad_s = 1;
  }
  if (bad_s) fatal("BDDCML_DEVICE_PLAN", "interface tiles of S are not packed by columns");
  ph("packed S check");
  const size_t plan_bytes = D.plan.bytes();
  // on the upload stream: the compute stream is busy with the scatter
  D.plan.commit(D.up);
  if (S.verbose >= 3) {
    std::printf("   device plan: %.1f MB of plan arrays\n", plan_bytes / 1e6);
    ph("  plan upload");
  }
  if (S.verbose >= 2)
    std::printf("   device plan: %lld subdomain tiles (%.1f GB), %zu padding slots\n",
                (long long)D.ntiles, D.ntiles * (D.f32i ? 16384.0 : 32768.0) / 1e9,
                (size_t)D.sc_pads.n);

  S.t.schur_flops = kTileGemm * tile_products(a_schur);
  S.t.interior_flops = tile_work(a_diag, 2 * kTileChol) + tile_work(a_off, kTileGemm) +
                       tile_work(m_rows, kTileGemm) + S.t.schur_flops;
  // ---- numerics, first half: Dirichlet elimination into the tiles,
  // interior factorization, Schur complements
  D.ev_gemm.assign(2 * D.a_off.size(), nullptr);
  for (hipEvent_t &e : D.ev_gemm) HIPCHECK(hipEventCreate(&e));
  D.gemm_flops.clear();
  for (const TaskList &l : a_off) D.gemm_flops.push_back(kTileGemm * (tile_products(l) + posted_tasks(l)));
  HIPCHECK(hipEventCreateWithFlags(&D.ev_up, hipEventDisableTiming));
  HIPCHECK(hipEventCreateWithFlags(&D.ev_rhs, hipEventDisableTiming));
  // element type of the subdomain tiles: FP64, or FP32 with refinement
  auto interior = [&](auto *pool) {
    for (size_t J = 0; J < D.a_diag.size(); J++) {
      launch_tile_potrf(st, pool, D.a_diag[J].tasks.p, D.a_diag[J].pairs.p, D.a_diag[J].ntask,
                        D.err.p);
      HIPCHECK(hipEventRecord(D.ev_gemm[2 * J], st));
      launch_tile_gemm(st, pool, D.a_off[J].tasks.p, D.a_off[J].pairs.p, D.a_off[J].ntask, GEMM_NT,
                       POST_RIGHT_XT, 1.0, -1.0);
      HIPCHECK(hipEventRecord(D.ev_gemm[2 * J + 1], st));
    }
    HIPCHECK(hipEventRecord(ev[1], st));
  };
  if (D.f32i) {
    interior(D.ipool32.p);
    launch_schur_gemm(st, D.ipool32.p, D.a_schur.tasks.p, D.a_schur.pairs.p, D.a_schur.ntask);
  } else {
    interior(D.pool.p);
    launch_schur_gemm(st, D.pool.p, D.a_schur.tasks.p, D.a_schur.pairs.p, D.a_schur.ntask);
  }
  for (const Phase &m : D.m_rows)
    launch_tile_gemm(st, D.pool.p, m.tasks.p, m.pairs.p, m.ntask, GEMM_NN, POST_LEFT_X, 1.0, -1.0);
  HIPCHECK(hipEventRecord(ev[2], st));
  HIPCHECK(hipGetLastError());
  S.t.plan += wall_time() - t0;
}

// Explicit inverse of S' on the compute stream: W = L^{-1} row by row, then
// S'^{-1} = W^T W over the factor tiles (the factor is not needed after
// that: later applies use the GEMV).  Launches queued behind work already on
// the stream, which still sees the factor.
static void form_inverse(Solver &S) {
  Device &D = *S.dev;
  hipStream_t st = D.st;
  auto run = [&](auto *fp) {
    for (size_t i = 0; i < D.w_rows.size(); i++)
      launch_tile_gemm(st, fp, D.w_rows[i].tasks.p, D.w_rows[i].pairs.p, D.w_rows[i].ntask, GEMM_NN,
                       POST_LEFT_X, 1.0, -1.0);
    launch_tile_gemm(st, fp, D.kinv.tasks.p, D.kinv.pairs.p, D.kinv.ntask, GEMM_TN, POST_NONE, 1.0,
                     1.0);
  };
  if (D.mixed) run(D.pool32.p);
  else run(D.fpool.p);
  D.explicit_inv = true;
}

// condensation rhs (src: local order, the Dirichlet-eliminated rhs when
// null) into the device-ordered subdomain vectors (padding 0), forward sweep
// over the interior factor, sum of the interface parts into g (D.g when null)
static void reduced_rhs_launches(Solver &S, hipStream_t st, const double *src = nullptr,
                                 double *g = nullptr) {
  Device &D = *S.dev;
  HIPCHECK(hipMemsetAsync(D.vec.p, 0, D.VT * sizeof(double), st));
  launch_to_device_order(st, D.odesc.p, S.nsub, D.sc_perm.p, src ? src : D.fmod.p, D.vec.p,
                         D.maxndofs);
  if (D.f32i)
    launch_trsv_fwd(st, D.ipool32.p, D.trsv.p, S.nsub, D.rptr.p, D.rk.p, D.rt.p, D.vec.p, D.maxnt);
  else
    launch_trsv_fwd(st, D.pool.p, D.trsv.p, S.nsub, D.rptr.p, D.rk.p, D.rt.p, D.vec.p, D.maxnt);
  launch_assemble(st, S.nig, D.aptr.p, D.rhs_src.p, nullptr, D.vec.p, g ? g : D.g.p);
  D.vec_y = true;
}

// Row-segmented index of the triplets for the FP64 residual of the refined
// form, built once per class (identical triplet structure): row r lists the
// entries (k, i) with aj[k] = r, i < r, then (k, j) with ai[k] = r.  Built
// and uploaded (own arena, upload stream) while the device runs the
// interface phase; the solve waits for the upload.
static void build_resid_index(Solver &S) {
  Device &D = *S.dev;
  const int nsub = S.nsub;
  const std::vector<int> &ocls = S.ord_cls, &oreps = S.ord_reps;
  const int nrep = (int)oreps.size();
  std::vector<size_t> rpo(nrep + 1, 0), eno(nrep + 1, 0);
  for (int k = 0; k < nrep; k++) {
    const Sub &s = S.subs[oreps[k]];
    rpo[k + 1] = rpo[k] + s.ndofs + 1;
    size_t ne = 0;
    for (size_t q = 0; q < s.ai.size(); q++) ne += s.ai[q] == s.aj[q] ? 1 : 2;
    eno[k + 1] = eno[k] + ne;
  }
  if (rpo[nrep] > (size_t)INT32_MAX) fatal("BDDCML_DEVICE_PLAN", "residual index too large");
  RawVec<int> rptr(rpo[nrep]);  // every entry is written below
  RawVec<int2> ent(eno[nrep]);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nrep; k++) {
    const Sub &s = S.subs[oreps[k]];
    int *rp = rptr.data() + rpo[k];
    int2 *E = ent.data() + eno[k];
    std::vector<int> nt(s.ndofs, 0), nr(s.ndofs, 0);
    for (size_t q = 0; q < s.ai.size(); q++) {
      nr[s.ai[q]]++;
      if (s.ai[q] != s.aj[q]) nt[s.aj[q]]++;
    }
    rp[0] = 0;
    for (int r = 0; r < s.ndofs; r++) rp[r + 1] = rp[r] + nt[r] + nr[r];
    std::vector<int> ct(s.ndofs), cr(s.ndofs);
    for (int r = 0; r < s.ndofs; r++) ct[r] = rp[r], cr[r] = rp[r] + nt[r];
    // triplets sorted by (row, column): both passes visit columns ascending
    for (size_t q = 0; q < s.ai.size(); q++) {
      const int i = s.ai[q], j = s.aj[q];
      E[cr[i]++] = make_int2((int)q, j);
      if (i != j) E[ct[j]++] = make_int2((int)q, i);
    }
  }
  std::vector<ResidSub> rs(nsub);
  int64_t foff = 0;
  for (int is = 0; is < nsub; is++) {
    const int c = ocls[is];
    if (S.subs[is].ai.size() != S.subs[oreps[c]].ai.size() ||
        (int64_t)S.din->av[is].n != (int64_t)S.subs[is].ai.size())
      fatal("BDDCML_DEVICE_PLAN", "triplet structure differs within a class");
    if (foff != D.hodesc[is].poff) fatal("BDDCML_DEVICE_PLAN", "local-order offsets");
    rs[is] = ResidSub{S.din->av[is].p, foff, (int64_t)eno[c], (int)rpo[c], (int)foff,
                      S.subs[is].ndofs, 0};
    foff += S.subs[is].ndofs;
  }
  D.plan4.add(D.rs_ptr, rptr);
  D.plan4.add(D.rs_ent, ent);
  D.plan4.add(D.rs_subs, rs);
  D.plan4.commit(D.up);
}

// rho of the S' augmentation: the largest interface diagonal entry of the
// subdomain matrix as the Dirichlet elimination leaves it (a fixed dof's
// diagonal below 1e-12, or missing, becomes 1.0), read from the uploaded
// triplets before that elimination runs on the host
static double interface_rho(const Sub &s) {
  double rho = 0.0;
  std::vector<char> seen(s.has_fix ? s.ndofs : 0, 0);
  for (size_t k = 0; k < s.av.size(); k++) {
    const int i = s.ai[k];
    if (i != s.aj[k] || s.perm[i] < s.NO) continue;
    double v = std::fabs(s.av[k]);
    if (s.has_fix && s.ifix[i]) {
      seen[i] = 1;
      if (v < 1e-12) v = 1.0;
    }
    rho = std::max(rho, v);
  }
  if (s.has_fix)
    for (int i = 0; i < s.ndofs; i++)
      if (s.ifix[i] && !seen[i] && s.perm[i] >= s.NO) rho = std::max(rho, 1.0);
  return rho > 0.0 ? rho : 1.0;
}

// Second half of the setup, first part: everything that needs the
// constraints but not the interface numbering, the weights or the
// right-hand side -- the factor space, the S' / Z / Phi lists, the
// constraint maps, the coarse lists -- uploaded and launched (S' assembly,
// its factorization, Z, G, Phi, the coarse matrices and the coarse
// factorization) while the host goes on with the unique interface, the
// boundary conditions and the weights.
void device_setup_interface(Solver &S) {
  Device &D = *S.dev;
  hipStream_t st = D.up;  // uploads overlap the factorization running on D.st
  const int nsub = S.nsub;
  double t0 = wall_time();
  double tph = t0;
  g_alloc_s = g_upload_s = 0.0;
  auto ph = [&](const char *name) {
    double now = wall_time();
    if (S.verbose >= 2) std::printf("   device plan: %-24s %8.4f s\n", name, now - tph);
    if (S.verbose >= 3)
      std::printf("   device plan:   (allocations %.4f s, uploads %.4f s so far)\n", g_alloc_s, g_upload_s);
    tph = now;
  };
  // ---- interface stage of the plan, now that the constraints are known: the
  // factor space (its own pool, FP64 or FP32 in the mixed-precision mode),
  // the coarse tiles (FP64 pool), the task lists of the S' factorization, Z
  // and Phi, the coarse factorization, the S' assembly jobs and the coarse
  // mirror
  {
    int64_t nf = 0;
    // the first setup of an adaptive run forms S'^{-1} at once: the face
    // eigenproblems' preconditioner applies it
    D.inv_planned = S.explicit_inverse != 0 || S.adapt_pass;
    D.inv_auto = S.explicit_inverse == 2 && !S.adapt_pass;
    D.explicit_inv = false;
    for (int is = 0; is < nsub; is++) assign_factor(D.sub[is], S.subs[is].nconstr, nf, D.inv_planned);
    if (nf >= INT32_MAX) fatal("BDDCML_DEVICE_PLAN", "too many factor tiles");
    D.ntilesf = nf;
    D.dense_coarse = !has_level2(S);
    D.NC = std::max(TILE, (S.ndofc + TILE - 1) / TILE * TILE);
    D.ntc = D.dense_coarse ? D.NC / TILE : 0;
    const int n = D.ntc;
    int next = 0;
    D.cbase = next;
    next += n * (n + 1) / 2;
    D.clinv = next;
    next += n;
    D.cwbase = next;
    next += n * (n - 1) / 2;
    D.cfull = next;
    next += n * n;
    D.ntilesc = next;
    if (D.mixed)
      D.pool32.alloc((size_t)std::max<int64_t>(nf, 1) * TILE_ELEMS);
    else
      D.fpool.alloc((size_t)std::max<int64_t>(nf, 1) * TILE_ELEMS);
    if (D.dense_coarse) D.cpool.alloc((size_t)D.ntilesc * TILE_ELEMS);
  }
  {
    // the factor-space lists depend only on (nti, ncT): one class per pair
    std::map<std::pair<int, int>, int> key;
    std::vector<int> fcls(nsub);
    std::vector<std::pair<int, int>> fkeys;
    for (int is = 0; is < nsub; is++) {
      auto k = std::make_pair(D.sub[is].nti, D.sub[is].ncT);
      auto it = key.emplace(k, (int)fkeys.size()).first;
      if (it->second == (int)fkeys.size()) fkeys.push_back(k);
      fcls[is] = it->second;
    }
    struct FactorTasks {
      std::vector<TaskList> p_diag, p_off, z_bwd, w_rows;
      TaskList phi, kinv;
    };
    std::vector<FactorTasks> fts(fkeys.size());
    int bad_ids = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad_ids)
    for (int k = 0; k < (int)fkeys.size(); k++) {
      const int n = fkeys[k].first, ncT = fkeys[k].second;
      const FactorLayout L(n, ncT, D.inv_planned);
      FactorTasks &Q = fts[k];
      auto Pt = [&](int I, int J) { return L.p + lowidx(n, I, J); };
      auto Yt = [&](int I, int cT) { return L.y + I * ncT + cT; };
      dense_chol_tasks(n, Pt, [&](int J) { return L.linv + J; }, Q.p_diag, Q.p_off);
      // Z = S'^{-1} C^T by tile sweeps over the factor.  Forward steps
      // Y_I = Linv_I (Y_I - sum_{k<I} L_Ik Y_k) need the factor's columns
      // k <= I only: they run in the product launch of column I of the S'
      // factorization (tasks of their own mode, Y initialized to C^T
      // before it).  Backward steps Y_I = Linv_I^T (Y_I - sum_{k>I} L_kI^T Y_k)
      // follow, one launch per tile column.
      Q.z_bwd.resize(ncT ? n : 0);
      for (int cT = 0; cT < ncT; cT++)
        for (int I = 0; I < n; I++) {
          TaskList &F = Q.p_off[I];
          TileTask t{Yt(I, cT), L.linv + I, (int)F.pairs.size(), 0};
          for (int q = 0; q < I; q++) F.pairs.push_back({Pt(I, q), Yt(q, cT)});
          t.p1 = (int)F.pairs.size();
          t.mode = task_mode(GEMM_NN, POST_LEFT_X, true);
          F.tasks.push_back(t);
          TaskList &B = Q.z_bwd[n - 1 - I];
          TileTask u{Yt(I, cT), L.linv + I, (int)B.pairs.size(), 0};
          for (int q = I + 1; q < n; q++) B.pairs.push_back({Pt(q, I), Yt(q, cT)});
          u.p1 = (int)B.pairs.size();
          B.tasks.push_back(u);
        }
      // Phi = Z G^{-1} on tiles: task (I, CT) sums Z(I, KT) Ginv(KT, CT)
      for (int I = 0; I < n; I++)
        for (int CT = 0; CT < ncT; CT++) {
          TileTask t{L.phi + I * ncT + CT, -1, (int)Q.phi.pairs.size(), 0};
          for (int KT = 0; KT < ncT; KT++) Q.phi.pairs.push_back({Yt(I, KT), L.gt + KT * ncT + CT});
          t.p1 = (int)Q.phi.pairs.size();
          Q.phi.tasks.push_back(t);
        }
      if (D.inv_planned && n > 0) {
        // W = L^{-1} by rows: W_IJ = -Linv_I sum_{J <= K < I} L_IK W_KJ with
        // W_JJ = Linv_J (one launch per row I, tasks start from zero);
        // then S'^{-1} = W^T W into the P tiles: (I, J) = sum_{K >= I} W_KI^T W_KJ
        auto Wf = [&](int K, int J) { return K == J ? L.linv + J : L.W(n, K, J); };
        Q.w_rows.resize(n);
        for (int I = 1; I < n; I++) {
          TaskList &R = Q.w_rows[I];
          for (int J = 0; J < I; J++) {
            TileTask t{L.W(n, I, J), L.linv + I, (int)R.pairs.size(), 0};
            for (int K = J; K < I; K++) R.pairs.push_back({Pt(I, K), Wf(K, J)});
            t.p1 = (int)R.pairs.size();
            t.fresh = 1;
            R.tasks.push_back(t);
          }
        }
        for (int J = 0; J < n; J++)
          for (int I = J; I < n; I++) {
            TileTask t{Pt(I, J), -1, (int)Q.kinv.pairs.size(), 0};
            for (int K = I; K < n; K++) Q.kinv.pairs.push_back({Wf(K, I), Wf(K, J)});
            t.p1 = (int)Q.kinv.pairs.size();
            t.fresh = 1;
            Q.kinv.tasks.push_back(t);
          }
      }
      for (auto *v : {&Q.p_diag, &Q.p_off, &Q.z_bwd, &Q.w_rows})
        for (TaskList &T : *v)
          if (!relative_ids(T, 0, L.span)) bad_ids |= 1;
      if (!relative_ids(Q.phi, 0, L.span)) bad_ids |= 1;
      if (!relative_ids(Q.kinv, 0, L.span)) bad_ids |= 1;
    }
    if (bad_ids) fatal("BDDCML_DEVICE_PLAN", "tile task outside its subdomain's factor tiles");
    for (int is = 0; is < nsub; is++)
      if (D.sub[is].fbase < 0 || (int64_t)D.sub[is].fbase + D.sub[is].fspan > D.ntilesf)
        fatal("BDDCML_DEVICE_PLAN", "factor tile range");
    auto fbase_of = [&](int is) { return D.sub[is].fbase; };
    std::vector<TaskList> p_diag = merge_steps(fts, &FactorTasks::p_diag, fcls, fbase_of),
                          p_off = merge_steps(fts, &FactorTasks::p_off, fcls, fbase_of),
                          z_bwd = merge_steps(fts, &FactorTasks::z_bwd, fcls, fbase_of);
    TaskList phi = materialize(
        fts, [](const FactorTasks &Q) -> const TaskList * { return &Q.phi; }, fcls, fbase_of);
    std::vector<TaskList> w_rows = merge_steps(fts, &FactorTasks::w_rows, fcls, fbase_of);
    TaskList kinv = materialize(
        fts, [](const FactorTasks &Q) -> const TaskList * { return &Q.kinv; }, fcls, fbase_of);
    // the explicit inverse's rows and product count only where the setup
    // forms it (option 1); the automatic mode forms it during a solve
    S.t.interface_flops = tile_work(p_diag, 2 * kTileChol) + tile_work(p_off, kTileGemm) +
                          tile_work(z_bwd, kTileGemm) +
                          kTileGemm * tile_products(phi);
    if (D.inv_planned && !D.inv_auto)
      S.t.interface_flops += tile_work(w_rows, kTileGemm) + kTileGemm * tile_products(kinv);
    D.plan2.verbose = S.verbose;
    upload_phases(D.p_diag, p_diag, D.plan2);
    upload_phases(D.p_off, p_off, D.plan2);
    upload_phases(D.z_bwd, z_bwd, D.plan2);
    D.phi.upload(phi, D.plan2);
    upload_phases(D.w_rows, w_rows, D.plan2);
    D.kinv.upload(kinv, D.plan2);
  }
  {
    // S' = S + C^T W C: one job per lower tile, into the factor space
    std::vector<int> spsub, spij, spdst;
    for (int is = 0; is < nsub; is++) {
      const SubTiles &T = D.sub[is];
      const int n = T.nti;
      for (int J = 0; J < n; J++)
        for (int I = J; I < n; I++) {
          spsub.push_back(is);
          spij.push_back((I << 16) | J);
          spdst.push_back(T.P(I, J));
        }
    }
    for (int d : spdst)
      if (d < 0 || d >= D.ntilesf) fatal("BDDCML_DEVICE_PLAN", "invalid augmented tile");
    D.plan2.add(D.sp_sub, spsub);
    D.plan2.add(D.sp_ij, spij);
    D.plan2.add(D.sp_dst, spdst);
    D.n_sp = (int)spsub.size();
  }
  ph("factor and coarse lists");
  if (S.verbose >= 2)
    std::printf("   device plan: %lld factor tiles (%.1f GB), %lld coarse tiles\n",
                (long long)D.ntilesf, D.ntilesf * (D.mixed ? 16384.0 : 32768.0) / 1e9,
                (long long)D.ntilesc);
  // ---- interface descriptors and maps: offsets first, then every
  // subdomain fills its own ranges in parallel
  D.hdesc.resize(nsub);
  std::vector<int64_t> loff(nsub), cdoff(nsub + 1, 0);
  std::vector<int> NIs(nsub);
  int64_t lv = 0, zo = 0, go = 0;
  int co = 0;
  for (int is = 0; is < nsub; is++) {
    const Sub &s = S.subs[is];
    const SubTiles &T = D.sub[is];
    SubDesc d;
    d.ni = s.ni;
    d.NI = s.NI;
    d.nti = s.nti;
    d.nc = s.nconstr;
    d.slow = T.slow;
    d.pbase = T.pbase;
    d.plinv = T.linvP;
    d.ybase = T.ybase;
    d.gtbase = T.gtbase;
    d.phibase = T.phibase;
    d.ncT = T.ncT;
    d.loff = lv;
    d.zoff = zo;
    d.goff = go;
    d.coff = co;
    d.rho = 1.0;
    D.hdesc[is] = d;
    loff[is] = lv;
    NIs[is] = s.NI;
    cdoff[is + 1] = cdoff[is] + (s.nconstr ? s.cptr[s.nconstr] - s.cptr[0] : 0);
    lv += s.NI;
    zo += (int64_t)s.NI * s.nconstr;
    go += (int64_t)s.nconstr * s.nconstr;
    co += s.nconstr;
    D.maxNI = std::max(D.maxNI, s.NI);
    D.maxnc = std::max(D.maxnc, s.nconstr);
    D.maxnti = std::max(D.maxnti, s.nti);
    D.maxncT = std::max(D.maxncT, T.ncT);
  }
  std::vector<int> cidv(lv), csz(co), cptrv(co + 1, 0), cdofv(cdoff[nsub]), cmapv(co);
  std::vector<double> cvalv(cdoff[nsub]), omegav(co);
  D.nco = co;
  int two_constraints = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : two_constraints)
  for (int is = 0; is < nsub; is++) {
    const Sub &s = S.subs[is];
    SubDesc &d = D.hdesc[is];
    d.rho = interface_rho(s);
    std::vector<int> ref2dev(s.ni);
    for (int k = 0; k < s.ni; k++) ref2dev[s.dev_ifc[k]] = k;
    int *cid_loc = cidv.data() + d.loff;
    std::fill(cid_loc, cid_loc + s.NI, -1);
    int64_t q2 = cdoff[is];
    for (int c = 0; c < s.nconstr; c++) {
      double nrm2 = 0.0;
      for (int q = s.cptr[c]; q < s.cptr[c + 1]; q++) {
        int k = ref2dev[s.cidx[q]];
        if (cid_loc[k] >= 0 || s.cval[q] != 1.0) two_constraints |= 1;
        cid_loc[k] = d.coff + c;
        cvalv[q2] = s.cval[q];
        cdofv[q2++] = k;
        nrm2 += s.cval[q] * s.cval[q];
      }
      cptrv[d.coff + c + 1] = (int)q2;
      csz[d.coff + c] = s.cptr[c + 1] - s.cptr[c];
      cmapv[d.coff + c] = s.indrowc[c];
      // omega_c c c^T: rho / |D_c| for a unit row, as the indicator form
      omegav[d.coff + c] = nrm2 > 0.0 ? d.rho / nrm2 : d.rho;
    }
  }
  // rows with overlapping supports or values other than 1 (adaptive
  // constraints) take the general forms of the S' and C^T kernels
  D.gen_constraints = two_constraints != 0;
  if (D.gen_constraints) {
    std::vector<int> cnt(lv + 1, 0);
    for (int is = 0; is < nsub; is++) {
      const SubDesc &d = D.hdesc[is];
      for (int c = 0; c < d.nc; c++)
        for (int q = cptrv[d.coff + c]; q < cptrv[d.coff + c + 1]; q++) cnt[d.loff + cdofv[q] + 1]++;
    }
    for (int64_t k = 0; k < lv; k++) cnt[k + 1] += cnt[k];
    std::vector<int> dconv(cnt[lv]), fillp(cnt.begin(), cnt.end() - 1);
    std::vector<double> dvalv(cnt[lv]);
    // constraints visited in ascending global index: every list sorted
    for (int is = 0; is < nsub; is++) {
      const SubDesc &d = D.hdesc[is];
      for (int c = 0; c < d.nc; c++)
        for (int q = cptrv[d.coff + c]; q < cptrv[d.coff + c + 1]; q++) {
          const int64_t a = d.loff + cdofv[q];
          dconv[fillp[a]] = d.coff + c;
          dvalv[fillp[a]++] = cvalv[q];
        }
    }
    std::fill(cidv.begin(), cidv.end(), -1);
    D.plan2.add(D.dptr, cnt);
    D.plan2.add(D.dcon, dconv);
    D.plan2.add(D.dval, dvalv);
    D.plan2.add(D.cvals, cvalv);
    D.plan2.add(D.comega, omegav);
  }
  // limits of the kernels: G^{-1} in place with one LDS row buffer (k_ginv),
  // the whole interface vector of a subdomain in LDS (k_kapply), the coarse
  // matrix as full tiles
  if (D.maxnc > 1024) fatal("BDDCML_SETUP_PRECONDITIONER", "more than 1024 constraints per subdomain");
  // (interfaces too large for LDS keep their working vector in global memory)
  if ((size_t)(1024 + 64 + D.maxnc) * sizeof(double) > max_dynamic_lds())
    fatal("BDDCML_SETUP_PRECONDITIONER", "too many constraints per subdomain for the preconditioner apply");
  if ((int64_t)D.ntc * D.ntc > ((int64_t)1 << 24))
    fatal("BDDCML_SETUP_PRECONDITIONER", "coarse problem too large for its dense factorization; "
                                         "use three levels");
  D.LV = lv;
  {
    // the deep form when subdomains are few or the largest sweep is well
    // above the average one (its workgroup is then the tail of the launch)
    double mx = 0.0, sum = 0.0;
    for (const SubDesc &d : D.hdesc) {
      const double wk = (double)d.nti * d.nti;
      mx = std::max(mx, wk);
      sum += wk;
    }
    D.kapply_deep = S.pc_form ? S.pc_form == 2
                    : launch_forms() == 0 ? nsub < 128 || (sum > 0.0 && mx > 2.0 * sum / nsub)
                                          : launch_forms() == 2;
    if (S.verbose >= 2)
      std::printf("   device plan: local preconditioner %s form (largest sweep %.2fx the mean)\n",
                  D.kapply_deep ? "deep" : "shallow", sum > 0.0 ? mx * nsub / sum : 0.0);
  }
  D.plan2.add(D.desc, D.hdesc);
  D.plan2.add(D.loff, loff);
  D.plan2.add(D.NIs, NIs);
  D.plan2.add(D.cid, cidv);
  D.plan2.add(D.csize, csz);
  D.plan2.add(D.cptr, cptrv);
  D.plan2.add(D.cdof, cdofv);
  D.plan2.add(D.cmap, cmapv);
  D.Z.alloc(std::max<int64_t>(zo, 1));
  D.Phi.alloc(std::max<int64_t>(zo, 1));
  D.G.alloc(std::max<int64_t>(go, 1));
  D.Acs.alloc(std::max<int64_t>(go, 1));
  // symmetric S GEMV: tile columns, longest first, and reduce rows; coarse
  // GEMV rows
  std::vector<SymvCol> sc;
  std::vector<SymvRow> sr;
  {
    int64_t ps = 0;
    for (int is = 0; is < nsub; is++) {
      const SubTiles &T = D.sub[is];
      for (int I = 0; I < T.nti; I++) sr.push_back({is, I, (int)ps, T.nti});
      ps += (int64_t)T.nti * (T.nti + 1) / 2;
    }
    if (ps >= INT32_MAX) fatal("BDDCML_DEVICE_PLAN", "too many interface tiles");
    for (int J = 0; J < D.maxnti; J++)
      for (int is = 0; is < nsub; is++) {
        const SubTiles &T = D.sub[is];
        if (J < T.nti)
          sc.push_back({T.slow + lowidx(T.nti, J, J), T.nti - J, is, J,
                        0});  // pslot set below
      }
    std::vector<int> psub(nsub);
    for (const SymvRow &r : sr) psub[r.mat] = r.psub;
    for (SymvCol &c : sc) c.pslot = psub[c.mat] + lowidx(D.sub[c.mat].nti, c.J, c.J);
    std::stable_sort(sc.begin(), sc.end(), [](const SymvCol &a, const SymvCol &b) { return a.nt > b.nt; });
    for (const SymvCol &c : sc)
      if (c.tile0 < 0 || (int64_t)c.tile0 + c.nt > D.ntiles) fatal("BDDCML_DEVICE_PLAN", "invalid GEMV column");
    if (D.inv_planned) {
      std::vector<SymvCol> pc(sc);
      for (SymvCol &c : pc) {
        c.tile0 = D.sub[c.mat].pbase + lowidx(D.sub[c.mat].nti, c.J, c.J);
        if (c.tile0 < 0 || (int64_t)c.tile0 + c.nt > D.ntilesf) fatal("BDDCML_DEVICE_PLAN", "invalid GEMV column");
      }
      D.plan2.add(D.pcols, pc);
    }
    D.spart.alloc((size_t)std::max<int64_t>(ps, 1) * 64);
  }
  // coarse inverse GEMV over the full-storage tiles
  if (D.dense_coarse) {
    if (D.cfull < 0 || (int64_t)D.cfull + (int64_t)D.ntc * D.ntc > D.ntilesc)
      fatal("BDDCML_DEVICE_PLAN", "invalid coarse GEMV tiles");
    D.cpart.alloc((size_t)D.ntc * D.ntc * 64);
  }
  D.plan2.add(D.scols, sc);
  D.plan2.add(D.srows, sr);
  D.nsrows = (int)sr.size();
  D.nscols = (int)sc.size();
  ph("interface maps");
  // coarse gather CSR: coarse dof -> constraint rows (subdomain order)
  {
    std::vector<int> cnt(S.ndofc + 1, 0);
    for (int v : cmapv) cnt[v + 1]++;
    for (int k = 0; k < S.ndofc; k++) cnt[k + 1] += cnt[k];
    std::vector<int> src(cnt[S.ndofc]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < cmapv.size(); k++) src[fill[cmapv[k]]++] = (int)k;
    D.plan2.add(D.cgptr, cnt);
    D.plan2.add(D.cgsrc, src);
  }
  D.plan2.commit(st);  // every array of this part in one copy
  ph("coarse gather, upload");
  HIPCHECK(hipEventRecord(D.ev_up, D.up));
  S.t.plan += wall_time() - t0;
  // ---- numerics of the interface stage: augmented interface matrices, Z, Phi, coarse
  t0 = wall_time();
  st = D.st;
  HIPCHECK(hipStreamWaitEvent(st, D.ev_up, 0));
  hipEvent_t *ev = D.ev;
  // S' assembly from S (FP64 tiles, or FP32 with refinement)
  auto sprime = [&](const auto *pool, auto *fp) {
    if (D.gen_constraints)
      launch_sprime_gen(st, pool, fp, D.desc.p, D.sp_sub.p, D.sp_ij.p, D.sp_dst.p, D.n_sp, D.dptr.p,
                        D.dcon.p, D.dval.p, D.comega.p);
    else
      launch_sprime(st, pool, fp, D.desc.p, D.sp_sub.p, D.sp_ij.p, D.sp_dst.p, D.n_sp, D.cid.p,
                    D.csize.p);
  };
  // S' factor and Z = S'^{-1} C^T in the factor space (FP32 when mixed);
  // then G = C Z, Phi = Z G^{-1} in FP64
  auto factor_phase = [&](auto *fp) {
    if (D.f32i) sprime((const float *)D.ipool32.p, fp);
    else sprime((const double *)D.pool.p, fp);
    // Y = C^T, then the factorization with the forward steps of Z in its
    // product launches
    if (D.gen_constraints)
      launch_ct_init_gen(st, fp, D.desc.p, nsub, D.dptr.p, D.dcon.p, D.dval.p, D.maxnti);
    else
      launch_ct_init(st, fp, D.desc.p, nsub, D.cid.p, D.maxnti);
    for (size_t J = 0; J < D.p_diag.size(); J++) {
      launch_tile_potrf(st, fp, D.p_diag[J].tasks.p, D.p_diag[J].pairs.p, D.p_diag[J].ntask, D.err.p);
      launch_tile_gemm(st, fp, D.p_off[J].tasks.p, D.p_off[J].pairs.p, D.p_off[J].ntask, GEMM_NT,
                       POST_RIGHT_XT, 1.0, -1.0);
    }
    for (size_t i = 0; i < D.z_bwd.size(); i++)
      launch_tile_gemm(st, fp, D.z_bwd[i].tasks.p, D.z_bwd[i].pairs.p, D.z_bwd[i].ntask, GEMM_TN,
                       POST_LEFT_XT, 1.0, -1.0);
    launch_y_to_z(st, fp, D.desc.p, nsub, D.Z.p, D.maxNI, 0);
  };
  // Phi = Z G^{-1} on the tile GEMM, in the factor space
  auto phi_phase = [&](auto *fp) {
    launch_ginv_to_tiles(st, fp, D.desc.p, nsub, D.G.p, D.maxncT);
    launch_tile_gemm(st, fp, D.phi.tasks.p, D.phi.pairs.p, D.phi.ntask, GEMM_NN, POST_NONE, 0.0,
                     1.0);
    launch_y_to_z(st, fp, D.desc.p, nsub, D.Phi.p, D.maxNI, 1);
  };
  if (D.mixed) factor_phase(D.pool32.p);
  else factor_phase(D.fpool.p);
  launch_g(st, D.desc.p, nsub, D.cptr.p, D.cdof.p, D.gen_constraints ? D.cvals.p : nullptr, D.Z.p,
           D.G.p);
  launch_ginv(st, D.desc.p, nsub, D.G.p, nullptr, D.err.p);
  if (D.mixed) phi_phase(D.pool32.p);
  else phi_phase(D.fpool.p);
  if (D.inv_planned && !D.inv_auto) form_inverse(S);
  launch_acs(st, D.desc.p, nsub, D.cptr.p, D.gen_constraints ? D.comega.p : nullptr, D.G.p,
             D.Acs.p);
  HIPCHECK(hipEventRecord(ev[3], st));
  ph("interface launches");
  // coarse problem (with three levels the second level is set up instead):
  // its task lists and assembly map are built on the host while the device
  // runs the interface phase launched above, and uploaded on the upload
  // stream (synchronized there, so the compute stream is not waited on)
  if (D.dense_coarse) {
    const double tc = wall_time();
    {
      // coarse problem (one dense matrix, tiles in the coarse pool):
      // right-looking tile Cholesky and the inverse factor W = L^{-1} by
      // right-looking forward substitution, so that every task is a single tile
      // product and each step is a few short launches; then Ac^{-1} = W^T W
      std::vector<SeqLaunch> cseq;
      TaskList cinv;
      const int n = D.ntc, cb = D.cbase, cl = D.clinv, cw = D.cwbase;
      auto Pc = [=](int I, int J) { return cb + lowidx(n, I, J); };
      auto Wc = [=](int I, int J) { return I == J ? cl + I : cw + lowidx(n - 1, I - 1, J); };
      // per column k two launches: the panel L_ik = A_ik Linv_k^T together
      // with the W row W_kj = Linv_k B_kj; the trailing updates
      // A_IK -= L_Ik L_Kk^T together with B_ij -= L_ik W_kj, where the task
      // of the next diagonal tile also factors it (L_k+1, Linv_k+1); only
      // the first diagonal tile has a launch of its own (tasks of one
      // launch carry their own kind)
      const int panel = task_mode(GEMM_NT, POST_RIGHT_XT, false),
                wrow = task_mode(GEMM_NN, POST_LEFT_X, false),
                trail = task_mode(GEMM_NT, POST_NONE, true), bupd = task_mode(GEMM_NN, POST_NONE, true);
      auto one = [](TaskList &l, int c, int a, int b, int mode) {
        const int p0 = (int)l.pairs.size();
        l.pairs.push_back({a, b});
        TileTask t{c, -1, p0, p0 + 1};
        t.mode = mode;
        l.tasks.push_back(t);
      };
      auto post = [](TaskList &l, int c, int x, int mode) {
        TileTask t{c, x, 0, 0};
        t.mode = mode;
        l.tasks.push_back(t);
      };
      for (int k = 0; k < n; k++) {
        if (k == 0) {
          SeqLaunch f;  // L_00 = chol(A_00), Linv_0
          f.potrf = true;
          f.l.tasks.push_back({Pc(0, 0), cl, 0, 0});
          cseq.push_back(std::move(f));
        }
        SeqLaunch pw;
        for (int i = k + 1; i < n; i++) post(pw.l, Pc(i, k), cl + k, panel);
        for (int j = 0; j < k; j++) post(pw.l, Wc(k, j), cl + k, wrow);
        if (!pw.l.tasks.empty()) cseq.push_back(std::move(pw));
        SeqLaunch up;
        up.step = true;
        for (int K = k + 1; K < n; K++)
          for (int I = K; I < n; I++) {
            one(up.l, Pc(I, K), Pc(I, k), Pc(K, k), trail);
            if (I == k + 1) {  // A_k+1,k+1 -= L L^T, then L_k+1 and Linv_k+1
              up.l.tasks.back().x = cl + k + 1;
              up.l.tasks.back().mode = kTaskPotrf;
            }
          }
        for (int i = k + 1; i < n; i++)
          for (int j = 0; j <= k; j++) one(up.l, Wc(i, j), Pc(i, k), Wc(k, j), bupd);
        if (!up.l.tasks.empty()) cseq.push_back(std::move(up));
      }
      for (int j = 0; j < n; j++)
        for (int i = j; i < n; i++) {
          TileTask t{Pc(i, j), -1, (int)cinv.pairs.size(), 0};
          for (int k = i; k < n; k++) cinv.pairs.push_back({Wc(k, i), Wc(k, j)});
          t.p1 = (int)cinv.pairs.size();
          cinv.tasks.push_back(t);
        }
      for (auto &q : cseq) check_tasks(q.l, D.ntilesc, q.potrf || q.post != POST_NONE, "coarse factorization");
      check_tasks(cinv, D.ntilesc, false, "coarse W^T W");
      D.cseq.resize(cseq.size());
      for (size_t k = 0; k < cseq.size(); k++) {
        SeqPhase &q = D.cseq[k];
        q.ph.upload(cseq[k].l, D.plan3);
        q.kind = cseq[k].kind, q.post = cseq[k].post, q.alpha = cseq[k].alpha, q.beta = cseq[k].beta;
        q.potrf = cseq[k].potrf;
        q.step = cseq[k].step;
      }
      D.cinv.upload(cinv, D.plan3);
      // full-storage mirror of Ac^{-1} for its GEMV; the factor, inverse
      // diagonal and W tiles start from zero; padding diagonals 1.0
      std::vector<MirrorJob> mc;
      for (int J = 0; J < n; J++)
        for (int I = J; I < n; I++) {
          int src = D.cbase + lowidx(n, I, J);
          mc.push_back({src, D.cfull + I * n + J, 0});
          if (I > J) mc.push_back({src, D.cfull + J * n + I, 1});
        }
      for (const MirrorJob &j : mc)
        if (j.src < 0 || j.src >= D.ntilesc || j.dst < 0 || j.dst >= D.ntilesc)
          fatal("BDDCML_DEVICE_PLAN", "invalid mirror job");
      std::vector<int2> cz{make_int2(D.cbase, n * (n + 1) / 2 + n + n * (n - 1) / 2)};
      if ((int64_t)cz[0].x + cz[0].y > D.ntilesc) fatal("BDDCML_DEVICE_PLAN", "invalid tile range to clear");
      std::vector<int64_t> cpads;
      for (int p = S.ndofc; p < D.NC; p++)
        cpads.push_back((int64_t)(D.cbase + lowidx(n, p / TILE, p / TILE)) * TILE_ELEMS +
                        (p % TILE) * TILE + (p % TILE));
      for (int64_t k : cpads)
        if (k < 0 || k >= D.ntilesc * (int64_t)TILE_ELEMS) fatal("BDDCML_DEVICE_PLAN", "scatter index out of range");
      D.plan3.add(D.mir_c, mc);
      D.plan3.add(D.c_zero, cz);
      D.plan3.add(D.c_pads, cpads);
    }
    ph("coarse factorization lists");
    // coarse assembly CSR over the lower tiles of the coarse matrix
    {
      struct E {
        int64_t dst, src;
      };
      const int n = D.ntc;
      // entries in subdomain order: counts, then each subdomain fills its range
      std::vector<size_t> eo(nsub + 1, 0);
      int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
      for (int is = 0; is < nsub; is++) {
        const Sub &s = S.subs[is];
        size_t c = 0;
        for (int a = 0; a < s.nconstr; a++)
          for (int b = 0; b < s.nconstr; b++) {
            const int ga = s.indrowc[a], gb = s.indrowc[b];
            if (ga < 0 || ga >= S.ndofc || gb < 0 || gb >= S.ndofc) bad = 1;
            else c += ga / TILE >= gb / TILE;
          }
        eo[is + 1] = c;
      }
      if (bad) fatal("BDDCML_DEVICE_PLAN", "coarse dof out of range");
      for (int is = 0; is < nsub; is++) eo[is + 1] += eo[is];
      std::vector<E> ent(eo[nsub]);
#pragma omp parallel for schedule(dynamic)
      for (int is = 0; is < nsub; is++) {
        const Sub &s = S.subs[is];
        const SubDesc &d = D.hdesc[is];
        size_t k = eo[is];
        for (int a = 0; a < s.nconstr; a++)
          for (int b = 0; b < s.nconstr; b++) {
            const int ga = s.indrowc[a], gb = s.indrowc[b];
            const int I = ga / TILE, J = gb / TILE;
            if (I < J) continue;
            const int64_t dst = (int64_t)(D.cbase + lowidx(n, I, J)) * TILE_ELEMS + (gb % TILE) * TILE + ga % TILE;
            ent[k++] = E{dst, d.goff + (int64_t)a * s.nconstr + b};
          }
      }
      // stable sort by destination: least-significant-digit radix passes
      {
        const int64_t maxd = (int64_t)D.ntilesc * TILE_ELEMS;
        int bits = 1;
        while (bits < 62 && ((int64_t)1 << bits) < maxd) bits++;
        constexpr int RB = 12, NB = 1 << RB;
        std::vector<E> tmp(ent.size());
        std::vector<size_t> c(NB + 1);
        for (int sh = 0; sh < bits; sh += RB) {
          std::fill(c.begin(), c.end(), 0);
          for (const E &e : ent) c[((e.dst >> sh) & (NB - 1)) + 1]++;
          for (int b = 0; b < NB; b++) c[b + 1] += c[b];
          for (const E &e : ent) tmp[c[(e.dst >> sh) & (NB - 1)]++] = e;
          ent.swap(tmp);
        }
      }
      std::vector<int64_t> dst, src;
      std::vector<int> ptr(1, 0);
      for (size_t k = 0; k < ent.size(); k++) {
        if (dst.empty() || dst.back() != ent[k].dst) {
          if (!dst.empty()) ptr.push_back((int)k);
          dst.push_back(ent[k].dst);
        }
        src.push_back(ent[k].src);
      }
      ptr.push_back((int)ent.size());
      D.plan3.add(D.ca_dst, dst);
      D.plan3.add(D.ca_src, src);
      D.plan3.add(D.ca_ptr, ptr);
      D.nca = (int)dst.size();
    }
    ph("coarse assembly lists");
    D.plan3.commit(D.up);
    S.t.plan += wall_time() - tc;
    // coarse tiles cleared (they accumulate), padding diagonals 1
    launch_zero_ranges(st, D.cpool.p, D.c_zero.p, (int)D.c_zero.n);
    launch_scatter_ones(st, D.cpool.p, D.c_pads.p, (int64_t)D.c_pads.n);
    double *cpool = D.cpool.p;
    launch_coarse_assemble(st, cpool, D.ca_dst.p, D.ca_ptr.p, D.ca_src.p, D.nca, D.Acs.p);
    for (const SeqPhase &q : D.cseq) {
      if (q.potrf)
        launch_tile_potrf(st, cpool, q.ph.tasks.p, q.ph.pairs.p, q.ph.ntask, D.err.p);
      else if (q.step)
        launch_tile_step(st, cpool, q.ph.tasks.p, q.ph.pairs.p, q.ph.ntask, D.err.p);
      else
        launch_tile_gemm(st, cpool, q.ph.tasks.p, q.ph.pairs.p, q.ph.ntask, q.kind, q.post,
                         q.alpha, q.beta);
    }
    launch_tile_gemm(st, cpool, D.cinv.tasks.p, D.cinv.pairs.p, D.cinv.ntask, GEMM_TN, POST_NONE,
                     0.0, 1.0);
    launch_tile_mirror(st, cpool, D.mir_c.p, (int)D.mir_c.n);
  }
  HIPCHECK(hipEventRecord(ev[4], st));
  HIPCHECK(hipGetLastError());
  S.t.factor += wall_time() - t0;
}

// Second half of the setup, second part (after the unique interface, the
// boundary conditions, the weights and the condensation right-hand side):
// the interface maps and weights, the assembly and sweep lists, the solve
// vectors; the reduced right-hand side on the upload stream; then the wait
// for the whole setup.
void device_setup_finish(Solver &S) {
  Device &D = *S.dev;
  hipStream_t st = D.up;
  const int nsub = S.nsub;
  double t0 = wall_time();
  double tph = t0;
  auto ph = [&](const char *name) {
    double now = wall_time();
    if (S.verbose >= 2) std::printf("   device plan: %-24s %8.4f s\n", name, now - tph);
    tph = now;
  };
  const int64_t lv = D.LV;
  const int co = D.nco;
  std::vector<int64_t> loff(nsub);
  for (int is = 0; is < nsub; is++) loff[is] = D.hdesc[is].loff;
  // interface dof of each local interface entry (-1: padding) and its weight
  // devpos: device position of each reference interface index
  // (every entry read below is written first)
  RawVec<int> map(lv), devpos(lv);
  RawVec<double> wv(lv);
#pragma omp parallel for schedule(dynamic)
  for (int is = 0; is < nsub; is++) {
    const Sub &s = S.subs[is];
    const int64_t o = loff[is];
    for (int k = 0; k < s.NI; k++) {
      if (k < s.ni) {
        int kr = s.dev_ifc[k];
        map[o + k] = S.gpos[s.isvgvn[s.ifc_dof[kr]]];
        wv[o + k] = s.w[kr];
        devpos[o + kr] = k;
      } else {
        map[o + k] = -1;
        wv[o + k] = 0.0;
      }
    }
  }
  D.plan2b.verbose = S.verbose;
  D.plan2b.add(D.map, map);
  D.plan2b.add(D.w, wv);
  ph("interface dofs, weights");
  // assembly CSR: unique interface dof -> local flat indices (subdomain
  // order), from the copy lists of the unique interface (ig_ptr, ig_sub,
  // ig_k: the copies of each dof in subdomain order, by reference
  // interface index) through each subdomain's device positions
  {
    const std::vector<int> &cnt = S.ig_ptr;
    if ((int)cnt.size() != S.nig + 1) fatal("BDDCML_DEVICE_PLAN", "interface numbering missing");
    const int64_t ncopy = cnt[S.nig];
    RawVec<int64_t> src(ncopy);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t q = 0; q < ncopy; q++) {
      const int is = S.ig_sub[q], kr = S.ig_k[q];
      if (is < 0 || is >= nsub || kr < 0 || kr >= S.subs[is].ni) {
        bad = 1;
        continue;
      }
      src[q] = loff[is] + devpos[loff[is] + kr];
    }
    if (bad) fatal("BDDCML_DEVICE_PLAN", "interface copy lists");
    D.plan2b.add(D.aptr, cnt);
    D.plan2b.add(D.asrc, src);
    // the same contributors in the device-ordered subdomain vectors
    std::vector<int64_t> voff(nsub);
    int64_t vt = 0;
    for (int is = 0; is < nsub; is++) voff[is] = vt, vt += S.subs[is].NO + S.subs[is].NI;
    D.VT = vt;
    if (D.xint) {
      if (D.gx_vt != vt) fatal("BDDCML_DEVICE_PLAN", "interior GEMV offsets");
      D.ybuf.alloc(std::max<int64_t>(vt, 1));
      D.zbuf.alloc(std::max<int64_t>(vt, 1));
      D.ybuf.zero(D.up);
      D.zbuf.zero(D.up);
    }
    D.hodesc.resize(nsub);
    {
      int64_t po = 0;
      for (int is = 0; is < nsub; is++) {
        D.hodesc[is] = OrderDesc{po, voff[is], S.subs[is].ndofs, D.perm_off[is]};
        po += S.subs[is].ndofs;
        D.maxndofs = std::max(D.maxndofs, S.subs[is].ndofs);
      }
      D.ndofs_total = po;
    }
    D.plan2b.add(D.odesc, D.hodesc);
    RawVec<int64_t> vpos(lv);
#pragma omp parallel for schedule(static)
    for (int is = 0; is < nsub; is++)
      for (int k = 0; k < S.subs[is].NI; k++) vpos[loff[is] + k] = voff[is] + S.subs[is].NO + k;
    RawVec<int64_t> rsrc(src.size());
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < src.size(); q++) rsrc[q] = vpos[src[q]];
    D.plan2b.add(D.rhs_src, rsrc);
    ph("assembly lists");
    // triangular sweep lists of the interior factor: row lists (forward) and
    // column lists (backward) of the strictly lower factored tiles
    D.htrsv.resize(nsub);
    std::vector<size_t> roff(nsub + 1, 0), eoff(nsub + 1, 0), coff_(nsub + 1, 0);
    int maxnt = 1;
    for (int is = 0; is < nsub; is++) {
      const SubTiles &T = D.sub[is];
      const int nt = T.nto + T.nti;
      maxnt = std::max(maxnt, nt);
      size_t ne = 0;
      for (int J = 0; J < T.nto; J++) ne += T.pat.colptr[J + 1] - T.pat.colptr[J] - 1;
      roff[is + 1] = roff[is] + nt + 1;
      coff_[is + 1] = coff_[is] + T.nto + 1;
      eoff[is + 1] = eoff[is] + ne;
      D.htrsv[is] = TrsvDesc{T.nto, T.nti, (int)roff[is], (int)coff_[is], T.linvA, voff[is]};
    }
    RawVec<int> rp(roff[nsub]), rkv(eoff[nsub]), rtv(eoff[nsub]), cp(coff_[nsub]),
        ckv(eoff[nsub]), ctv(eoff[nsub]);  // every entry is written below
#pragma omp parallel for schedule(dynamic)
    for (int is = 0; is < nsub; is++) {
      const SubTiles &T = D.sub[is];
      const int nt = T.nto + T.nti;
      const int e0 = (int)eoff[is];
      // column lists straight from the CSC pattern (diagonal skipped)
      int e = e0;
      for (int J = 0; J < T.nto; J++) {
        cp[coff_[is] + J] = e;
        for (int q = T.pat.colptr[J] + 1; q < T.pat.colptr[J + 1]; q++)
          ckv[e] = T.pat.rowidx[q], ctv[e] = T.pat.tile[q], e++;
      }
      cp[coff_[is] + T.nto] = e;
      // row lists by counting sort over rows, columns ascending
      std::vector<int> cnt(nt + 1, 0);
      for (int J = 0; J < T.nto; J++)
        for (int q = T.pat.colptr[J] + 1; q < T.pat.colptr[J + 1]; q++) cnt[T.pat.rowidx[q] + 1]++;
      for (int I = 0; I < nt; I++) cnt[I + 1] += cnt[I];
      for (int I = 0; I <= nt; I++) rp[roff[is] + I] = e0 + cnt[I];
      for (int J = 0; J < T.nto; J++)
        for (int q = T.pat.colptr[J] + 1; q < T.pat.colptr[J + 1]; q++) {
          int k = e0 + cnt[T.pat.rowidx[q]]++;
          rkv[k] = J, rtv[k] = T.pat.tile[q];
        }
    }
    // (subdomains whose vector does not fit in LDS are swept with it in
    // global memory, see launch_trsv_fwd)
    for (int id : rtv)
      if (id < 0 || id >= D.ntiles) fatal("BDDCML_DEVICE_PLAN", "invalid sweep tile");
    for (int id : ctv)
      if (id < 0 || id >= D.ntiles) fatal("BDDCML_DEVICE_PLAN", "invalid sweep tile");
    D.maxnt = maxnt;
    D.plan2b.add(D.trsv, D.htrsv);
    D.plan2b.add(D.rptr, rp);
    D.plan2b.add(D.rk, rkv);
    D.plan2b.add(D.rt, rtv);
    D.plan2b.add(D.tcptr, cp);
    D.plan2b.add(D.tck, ckv);
    D.plan2b.add(D.tct, ctv);
  }
  ph("sweep lists");
  // vectors, carved from one allocation
  D.nig = S.nig;
  {
    DBuf<double> *krylov[] = {&D.u, &D.r, &D.p, &D.z, &D.ap, &D.rs, &D.v, &D.y, &D.s, &D.t, &D.g,
                              &D.xin, &D.xout};
    const size_t nk = sizeof(krylov) / sizeof(krylov[0]);
    const size_t sizes[] = {(size_t)lv, (size_t)lv, (size_t)std::max(co, 1), (size_t)D.NC,
                            (size_t)D.NC, (size_t)DOT_BLOCKS * 4, 4, (size_t)D.VT};
    DBuf<double> *fixed[] = {&D.yl, &D.zl, &D.rcs, &D.rc, &D.uc, &D.partial, &D.dots, &D.vec};
    size_t total = 0;
    for (size_t k = 0; k < 8; k++) total += sizes[k] + 32;
    total += nk * ((size_t)std::max(S.nig, 1) + 32);
    for (DBuf<double> *b : fixed) b->release();
    for (DBuf<double> *b : krylov) b->release();
    D.vblock.alloc(total);
    size_t off = 0;
    for (size_t k = 0; k < 8; k++) fixed[k]->carve(D.vblock, off, sizes[k]);
    for (DBuf<double> *b : krylov) b->carve(D.vblock, off, std::max(S.nig, 1));
  }
  D.plan2b.add(D.fixmask, std::vector<char>(S.ig_fixed.begin(), S.ig_fixed.end()));
  D.rc.zero(st);
  D.uc.zero(st);
  {
    RawVec<double> fm(D.ndofs_total);
#pragma omp parallel for schedule(static)
    for (int is = 0; is < nsub; is++)
      std::copy(S.subs[is].fmod.begin(), S.subs[is].fmod.end(), fm.begin() + D.hodesc[is].poff);
    D.plan2b.add(D.fmod, fm);
    D.sol.alloc(D.ndofs_total);
  }
  ph("vectors");
  D.plan2b.commit(st);
  ph("upload");
  // the residual index of the refined form first: its upload synchronizes
  // the upload stream, which then takes the reduced right-hand side
  if (D.f32i) {
    build_resid_index(S);
    ph("residual index");
  }
  // the reduced right-hand side needs only the interior factor (event 2):
  // on the upload stream, concurrent with the interface phase
  HIPCHECK(hipStreamWaitEvent(D.up, D.ev[2], 0));
  reduced_rhs_launches(S, D.up);
  HIPCHECK(hipEventRecord(D.ev_rhs, D.up));
  D.rhs_pending = true;
  S.t.plan += wall_time() - t0;
  t0 = wall_time();
  st = D.st;
  hipEvent_t *ev = D.ev;
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
  {
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[2]));
    S.t.ev_interior_ms = ms;
    HIPCHECK(hipEventElapsedTime(&ms, ev[1], ev[2]));
    S.t.ev_schur_ms = ms;
    HIPCHECK(hipEventElapsedTime(&ms, ev[2], ev[3]));
    S.t.ev_interface_ms = ms;
    HIPCHECK(hipEventElapsedTime(&ms, ev[3], ev[4]));
    S.t.ev_coarse_ms = ms;
    // the interior GEMM launches, skipping empty ones
    S.t.ev_gemm_ms = S.t.gemm_flops = 0.0;
    S.t.gemm_launches = 0;
    for (size_t J = 0; J < D.gemm_flops.size(); J++) {
      if (D.gemm_flops[J] <= 0.0) continue;
      HIPCHECK(hipEventElapsedTime(&ms, D.ev_gemm[2 * J], D.ev_gemm[2 * J + 1]));
      S.t.ev_gemm_ms += ms;
      S.t.gemm_flops += D.gemm_flops[J];
      S.t.gemm_launches++;
      if (S.verbose >= 2)
        std::printf("   interior GEMM launch %zu: %.4e flops, %.3f ms, %.1f TFLOP/s\n", J,
                    D.gemm_flops[J], ms, D.gemm_flops[J] / (ms * 1e9));
    }
  }
  int herr = 0;
  HIPCHECK(hipMemcpy(&herr, D.err.p, sizeof(int), hipMemcpyDeviceToHost));
  if (herr & 1) fatal("BDDCML_SETUP_PRECONDITIONER", "matrix not positive definite in a tile Cholesky factorization");
  if (herr & 2) fatal("BDDCML_SETUP_PRECONDITIONER", "singular constraint matrix C S'^{-1} C^T");
  if (herr & 4) fatal("BDDCML_SETUP_PRECONDITIONER", "matrix entry outside the symbolic tile pattern");
  S.t.factor += wall_time() - t0;
}

// ------------------------------------------------------------- right-hand side
void device_reduced_rhs(Solver &S) {
  Device &D = *S.dev;
  if (D.rhs_pending) {  // launched by the setup on the upload stream
    HIPCHECK(hipStreamWaitEvent(D.st, D.ev_rhs, 0));
    D.rhs_pending = false;
  } else {
    reduced_rhs_launches(S, D.st);
  }
  HIPCHECK(hipStreamSynchronize(D.st));
}

// new condensation rhs of every subdomain (after new right-hand sides or
// Dirichlet values) uploaded for the next device_reduced_rhs
void device_upload_rhs(Solver &S) {
  Device &D = *S.dev;
  std::vector<double> fm(D.ndofs_total);
#pragma omp parallel for schedule(static)
  for (int is = 0; is < S.nsub; is++)
    std::copy(S.subs[is].fmod.begin(), S.subs[is].fmod.end(), fm.begin() + D.hodesc[is].poff);
  HIPCHECK(hipMemcpyAsync(D.fmod.p, fm.data(), fm.size() * sizeof(double), hipMemcpyHostToDevice, D.st));
  HIPCHECK(hipStreamSynchronize(D.st));
}

void device_reduced_rhs_download(Solver &S, double *g) {
  Device &D = *S.dev;
  HIPCHECK(hipMemcpy(g, D.g.p, S.nig * sizeof(double), hipMemcpyDeviceToHost));
}

// ------------------------------------------------------------- debug probes
void device_selftest_tile(int op, int npairs, const double *A, const double *B, const double *C,
                          const double *X, double alpha, double beta, double *out) {
  if (npairs < 0 || npairs > 16) fatal("BDDCML_SELFTEST", "npairs out of range");
  // tiles: 0 = C, 1 = X, 2.. = A_p, B_p
  const int nt = 2 + 2 * npairs;
  DBuf<double> pool;
  pool.alloc((size_t)nt * TILE_ELEMS);
  std::vector<double> h((size_t)nt * TILE_ELEMS, 0.0);
  std::copy(C, C + TILE_ELEMS, h.begin());
  if (X) std::copy(X, X + TILE_ELEMS, h.begin() + TILE_ELEMS);
  for (int p = 0; p < npairs; p++) {
    std::copy(A + (size_t)p * TILE_ELEMS, A + (size_t)(p + 1) * TILE_ELEMS,
              h.begin() + (size_t)(2 + 2 * p) * TILE_ELEMS);
    std::copy(B + (size_t)p * TILE_ELEMS, B + (size_t)(p + 1) * TILE_ELEMS,
              h.begin() + (size_t)(3 + 2 * p) * TILE_ELEMS);
  }
  const bool f32 = op >= 100;
  if (f32) op -= 100;
  DBuf<float> pool32;
  if (f32) {
    std::vector<float> hf(h.begin(), h.end());
    pool32.alloc(hf.size());
    HIPCHECK(hipMemcpy(pool32.p, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  HIPCHECK(hipMemcpy(pool.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  TaskList l;
  l.tasks.push_back({0, 1, 0, npairs});
  for (int p = 0; p < npairs; p++) l.pairs.push_back({2 + 2 * p, 3 + 2 * p});
  DBuf<TileTask> t;
  DBuf<TilePair> pr;
  DBuf<int> err;
  t.upload(l.tasks, nullptr);
  pr.upload(l.pairs, nullptr);
  err.alloc(1);
  err.zero(nullptr);
  if (op == 3) {
    if (f32) launch_tile_potrf(nullptr, pool32.p, t.p, pr.p, 1, err.p);
    else launch_tile_potrf(nullptr, pool.p, t.p, pr.p, 1, err.p);
  } else {
    int kind = op % 10, post = op / 10;
    if (kind < 0 || kind > 2 || post < 0 || post > 3) fatal("BDDCML_SELFTEST", "bad op");
    if (f32) launch_tile_gemm(nullptr, pool32.p, t.p, pr.p, 1, kind, post, alpha, beta);
    else launch_tile_gemm(nullptr, pool.p, t.p, pr.p, 1, kind, post, alpha, beta);
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipDeviceSynchronize());
  if (f32) {
    std::vector<float> hf(2 * TILE_ELEMS);
    HIPCHECK(hipMemcpy(hf.data(), pool32.p, hf.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::copy(hf.begin(), hf.begin() + (op == 3 ? 2 : 1) * TILE_ELEMS, out);
    return;
  }
  HIPCHECK(hipMemcpy(out, pool.p, TILE_ELEMS * sizeof(double), hipMemcpyDeviceToHost));
  if (op == 3)
    HIPCHECK(hipMemcpy(out + TILE_ELEMS, pool.p + TILE_ELEMS, TILE_ELEMS * sizeof(double),
                       hipMemcpyDeviceToHost));
}

double device_selftest_copy(long long nbytes) {
  DBuf<double2> a, b;
  const int64_t n = nbytes / 16;
  a.alloc(n);
  b.alloc(n);
  HIPCHECK(hipMemset(a.p, 0, n * sizeof(double2)));
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  launch_copy_calibration(nullptr, a.p, b.p, n);  // warm
  HIPCHECK(hipEventRecord(e0, nullptr));
  launch_copy_calibration(nullptr, a.p, b.p, n);
  HIPCHECK(hipEventRecord(e1, nullptr));
  HIPCHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

// the packed lower tiles of every S_i: the subdomain pool, and per
// subdomain the first tile of S and its tile count
const double *device_schur_tiles(Solver &S, std::vector<int> &slow, std::vector<int> &nti) {
  if (!S.dev || !S.dev->pool.p) fatal("BDDCML_DEVICE", "solver not set up on the device");
  slow.resize(S.nsub);
  nti.resize(S.nsub);
  for (int is = 0; is < S.nsub; is++) slow[is] = S.dev->sub[is].slow, nti[is] = S.dev->sub[is].nti;
  return S.dev->pool.p;
}

hipStream_t device_stream(Solver &S) {
  if (!S.dev || !S.dev->st) fatal("BDDCML_DEVICE", "solver not set up on the device");
  return S.dev->st;
}

void device_element_gradient(Solver &S, const double *ae, int nke, bool per_element,
                             const double *u, const double *lam, double *grad) {
  const char *R = "BDDCML_AMD_ELEMENT_GRADIENT";
  Device &D = *S.dev;
  // element dof lists in global numbering and global element ids
  std::vector<int> edof, eid;
  for (const Sub &s : S.subs) {
    if ((int)s.isegn.size() != s.nelems) fatal(R, "global element numbers (isegn) were not uploaded");
    int k = 0;
    for (int e = 0; e < s.nelems; e++) {
      int nd = 0;
      for (int q = 0; q < s.nnet[e]; q++) {
        const int l = s.inet[k + q];
        for (int d = s.kdof[l]; d < s.kdof[l + 1]; d++, nd++) edof.push_back(s.isvgvn[d]);
      }
      k += s.nnet[e];
      if (nd != nke) fatal(R, "element dof count differs from nke");
      if (s.isegn[e] < 0 || s.isegn[e] >= S.nelem_g) fatal(R, "element number out of range");
      eid.push_back(s.isegn[e]);
    }
  }
  for (int d : edof)
    if (d < 0 || d >= S.ndof_g) fatal(R, "dof number out of range");
  const int nel = (int)eid.size();
  const size_t nae = per_element ? (size_t)S.nelem_g * nke * nke : (size_t)nke * nke;
  DBuf<int> dedof, deid;
  DBuf<double> dae, du, dl, dg;
  dedof.upload(edof, D.st);
  deid.upload(eid, D.st);
  dae.upload(std::vector<double>(ae, ae + nae), D.st);
  du.upload(std::vector<double>(u, u + S.ndof_g), D.st);
  dl.upload(std::vector<double>(lam, lam + S.ndof_g), D.st);
  dg.alloc(S.nelem_g);
  HIPCHECK(hipMemsetAsync(dg.p, 0, S.nelem_g * sizeof(double), D.st));
  launch_element_gradient(D.st, nel, nke, dedof.p, deid.p, dae.p, per_element, du.p, dl.p, dg.p);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(grad, dg.p, S.nelem_g * sizeof(double), hipMemcpyDeviceToHost, D.st));
  HIPCHECK(hipStreamSynchronize(D.st));
}

void device_sub_matrix(Solver &S, int is, int which, double *out) {
  Device &D = *S.dev;
  const Sub &s = S.subs[is];
  const SubTiles &T = D.sub[is];
  if (which != 'S') fatal("BDDCML_AMD_SUB_MATRIX", "only the Schur complement is stored explicitly");
  const int n = T.nti;
  std::vector<double> tiles((size_t)n * (n + 1) / 2 * TILE_ELEMS);
  if (D.f32i) {
    std::vector<float> t32(tiles.size());
    HIPCHECK(hipMemcpy(t32.data(), D.ipool32.p + (size_t)T.slow * TILE_ELEMS,
                       t32.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::copy(t32.begin(), t32.end(), tiles.begin());
  } else {
    HIPCHECK(hipMemcpy(tiles.data(), D.pool.p + (size_t)T.slow * TILE_ELEMS,
                       tiles.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  for (int a = 0; a < s.ni; a++)
    for (int b = 0; b < s.ni; b++) {
      const int p = std::max(a, b), q = std::min(a, b);  // lower triangle entry (p, q)
      const size_t t = lowidx(n, p / TILE, q / TILE);
      out[(size_t)s.dev_ifc[a] * s.ni + s.dev_ifc[b]] =
          tiles[t * TILE_ELEMS + (q % TILE) * TILE + p % TILE];
    }
}

void device_sub_coarse(Solver &S, int is, double *phi, double *ac) {
  Device &D = *S.dev;
  const Sub &s = S.subs[is];
  const SubDesc &d = D.hdesc[is];
  const int nc = s.nconstr;
  std::vector<double> P((size_t)d.NI * nc), A((size_t)nc * nc);
  if (nc) {
    HIPCHECK(hipMemcpy(P.data(), D.Phi.p + d.zoff, P.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(A.data(), D.Acs.p + d.goff, A.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  for (int a = 0; a < s.ni; a++)
    for (int c = 0; c < nc; c++) phi[(size_t)s.dev_ifc[a] * nc + c] = P[(size_t)c * d.NI + a];
  for (size_t k = 0; k < A.size(); k++) ac[k] = A[k];
}

void device_coarse_matrices(Solver &S, std::vector<std::vector<double>> &acs) {
  Device &D = *S.dev;
  std::vector<double> all(D.Acs.n);
  HIPCHECK(hipMemcpy(all.data(), D.Acs.p, all.size() * sizeof(double), hipMemcpyDeviceToHost));
  acs.resize(S.nsub);
  for (int is = 0; is < S.nsub; is++) {
    const int64_t o = D.hdesc[is].goff, m = (int64_t)S.subs[is].nconstr * S.subs[is].nconstr;
    if (o < 0 || o + m > (int64_t)all.size()) fatal("LEVELS_PREPARE_STANDARD_LEVEL", "coarse matrix range");
    acs[is].assign(all.begin() + o, all.begin() + o + m);
  }
}

void device_pair_pc_source(Solver &S, PairPcSource &o) {
  Device &D = *S.dev;
  if (!D.explicit_inv) fatal("ADAPTIVITY_SOLVE_EIGENVECTORS", "explicit S'^-1 not formed by the setup");
  HIPCHECK(hipStreamSynchronize(D.st));
  o.inv = D.mixed ? (const void *)D.pool32.p : (const void *)D.fpool.p;
  o.inv_float = D.mixed;
  o.Z = D.Z.p;
  o.Phi = D.Phi.p;
  o.pbase.resize(S.nsub);
  o.zoff.resize(S.nsub);
  for (int is = 0; is < S.nsub; is++) o.pbase[is] = D.hdesc[is].pbase, o.zoff[is] = D.hdesc[is].zoff;
  device_coarse_matrices(S, o.Ac);
}

// Entry p of the second level's device-ordered subdomain vectors takes the
// level-1 coarse residual at its coarse dof, weighted at the interface (the
// residual is a complete right-hand side: src/module_levels.f90:3914-3926,
// src/module_dd.f90:7068-7087); pads take 0.  Each coarse dof reads its
// solution back from one of its copies (all copies agree after the
// interface assembly).  The second level then runs on this level's stream.
void device_link_levels(Solver &S) {
  const char *R = "LEVELS_PREPARE_STANDARD_LEVEL";
  Solver &T = *S.lev2;
  Device &D = *S.dev, &E = *T.dev;
  const int64_t n = E.VT;
  if (n >= INT32_MAX) fatal(R, "second level too large");
  std::vector<int> src(n, -1), pos(std::max(S.ndofc, 1), -1);
  std::vector<double> w(n, 0.0);
  for (int j = 0; j < T.nsub; j++) {
    const Sub &t = T.subs[j];
    std::vector<double> wd(t.ndofs, 1.0);
    for (size_t k = 0; k < t.ifc_dof.size(); k++) wd[t.ifc_dof[k]] = t.w[k];
    for (int d = 0; d < t.ndofs; d++) {
      const int64_t p = E.hodesc[j].voff + t.perm[d];
      const int g = t.isvgvn[d];
      if (p < 0 || p >= n || src[p] != -1 || g < 0 || g >= S.ndofc) fatal(R, "invalid second-level map");
      src[p] = g;
      w[p] = wd[d];
      if (pos[g] < 0) pos[g] = (int)p;
    }
  }
  for (int g = 0; g < S.ndofc; g++)
    if (pos[g] < 0) fatal(R, "coarse dof missing on the second level");
  D.l2n = n;
  D.l2src.upload(src, D.st);
  D.l2w.upload(w, D.st);
  D.l2pos.upload(pos, D.st);
  HIPCHECK(hipStreamSynchronize(D.st));
  HIPCHECK(hipStreamSynchronize(E.st));
  HIPCHECK(hipStreamSynchronize(E.up));  // its own reduced rhs (unused) is done
  E.rhs_pending = false;
  if (!E.borrowed_st) HIPCHECK(hipStreamDestroy(E.st));
  E.st = D.st;
  E.borrowed_st = true;
  // the levels below it borrowed its stream: they run on this one too
  for (Solver *c = T.lev2.get(); c && c->dev; c = c->lev2.get()) {
    c->dev->st = D.st;
    c->dev->borrowed_st = true;
  }
}

// --------------------------------------------------------------- operators
// events around one S GEMV launch and one local preconditioner launch of
// a solve, read after its final synchronize (recording them does not stop
// the host)
static void timing_events(Device &D) {
  for (hipEvent_t &e : D.ev_apply)
    if (!e) HIPCHECK(hipEventCreate(&e));
}

static void op_schur(Solver &S, const double *x, double *y, bool timed = false) {
  Device &D = *S.dev;
  if (timed) {
    timing_events(D);
    HIPCHECK(hipEventRecord(D.ev_apply[0], D.st));
  }
  if (D.f32i)
    launch_symv(D.st, (const float *)D.ipool32.p, D.scols.p, D.nscols, D.srows.p, D.nsrows,
                D.loff.p, D.map.p, nullptr, x, D.spart.p, D.yl.p);
  else
    launch_symv(D.st, (const double *)D.pool.p, D.scols.p, D.nscols, D.srows.p, D.nsrows,
                D.loff.p, D.map.p, nullptr, x, D.spart.p, D.yl.p);
  if (timed) {
    HIPCHECK(hipEventRecord(D.ev_apply[1], D.st));
    double tiles = 0;
    for (const SubTiles &T : D.sub) tiles += (double)T.nti * (T.nti + 1) / 2;
    // lower matrix tiles (8 or 4 bytes an entry) + gathered input + written
    // output (8 bytes each)
    S.t.gemv_bytes = tiles * TILE_ELEMS * (D.f32i ? 4.0 : 8.0) + 16.0 * (double)D.LV;
    D.gemv_timed = true;
  }
  launch_assemble(D.st, S.nig, D.aptr.p, D.asrc.p, nullptr, D.yl.p, y);
}

// event times of the timed applies of the last solve (after its synchronize)
static void read_apply_timings(Solver &S) {
  Device &D = *S.dev;
  float ms = 0;
  if (D.gemv_timed) {
    HIPCHECK(hipEventElapsedTime(&ms, D.ev_apply[0], D.ev_apply[1]));
    S.t.ev_gemv_ms = ms;
  }
  if (D.kapply_timed) {
    HIPCHECK(hipEventElapsedTime(&ms, D.ev_apply[2], D.ev_apply[3]));
    S.t.ev_kapply_ms = ms;
  }
  D.gemv_timed = D.kapply_timed = false;
}

static void op_pc(Solver &S, const double *r, double *z, bool timed = false);

// uc = one application of the second level's BDDC to the coarse residual rc
// (src/module_levels.f90:3849-4026, 4064-4249): its subdomain right-hand
// sides from rc, interior forward sweep and reduced interface residual,
// preconditioner, interface values back, interior backward sweep; then the
// coarse dofs read back.  All on this level's stream.
static void level2_apply(Solver &S, const double *rc, double *uc) {
  Solver &T = *S.lev2;
  Device &D = *S.dev, &E = *T.dev;
  hipStream_t st = D.st;
  launch_level_fill(st, D.l2n, D.l2src.p, D.l2w.p, rc, E.vec.p);
  const double *ep = E.pool.p;
  auto gemv = [&](int k, const double *x, const double *b, double *y, int trans, double sign) {
    launch_tile_gemv(st, ep, E.gx_t[k].p, (int)E.gx_t[k].n, E.gx_p[k].p, x, b, y, trans, sign);
  };
  // interior solves: by the explicit inverse (four GEMVs across the chip),
  // or by the forward and backward sweeps
  if (E.xint) {
    gemv(0, E.vec.p, nullptr, E.ybuf.p, 0, 1.0);
    gemv(1, E.ybuf.p, E.vec.p, E.vec.p, 0, -1.0);
  } else {
    launch_trsv_fwd(st, E.pool.p, E.trsv.p, T.nsub, E.rptr.p, E.rk.p, E.rt.p, E.vec.p, E.maxnt);
  }
  launch_assemble(st, T.nig, E.aptr.p, E.rhs_src.p, nullptr, E.vec.p, E.g.p);
  op_pc(T, E.g.p, E.u.p);
  launch_ifc_scatter(st, T.nig, E.aptr.p, E.rhs_src.p, E.u.p, E.vec.p);
  if (E.xint) {
    gemv(2, E.vec.p, E.ybuf.p, E.zbuf.p, 1, -1.0);
    gemv(3, E.zbuf.p, nullptr, E.vec.p, 1, 1.0);
  } else {
    launch_trsv_bwd(st, E.pool.p, E.trsv.p, T.nsub, E.tcptr.p, E.tck.p, E.tct.p, E.vec.p,
                    E.maxnt);
  }
  launch_level_gather(st, S.ndofc, D.l2pos.p, E.vec.p, uc);
}

// applies after which the automatic mode forms the explicit inverse: about
// the number whose saving (roughly half an apply each) pays for forming it
constexpr int kInverseAfterApplies = 40;

// automatic explicit inverse, at a solve boundary only (so that one Krylov
// run sees one preconditioner): formed on every level whose preconditioner
// has been applied kInverseAfterApplies times on this setup
static void inverse_at_solve_start(Solver &S) {
  for (Solver *L = &S; L && L->dev; L = L->lev2.get()) {
    Device &D = *L->dev;
    if (D.inv_auto && !D.explicit_inv && D.pc_count >= kInverseAfterApplies) form_inverse(*L);
  }
}

static void op_pc(Solver &S, const double *r, double *z, bool timed) {
  Device &D = *S.dev;
  hipStream_t st = D.st;
  if (D.inv_auto && !D.explicit_inv) ++D.pc_count;
  // coarse correction uc = Ac^{-1} sum_i R_ci^T Phi_i^T W_i r_i (three
  // levels: the second level's BDDC in place of Ac^{-1})
  launch_phit(st, D.desc.p, S.nsub, D.Phi.p, D.map.p, D.w.p, r, D.rcs.p, D.maxnc, D.maxNI);
  launch_coarse_gather(st, S.ndofc, D.cgptr.p, D.cgsrc.p, D.rcs.p, D.rc.p);
  if (S.lev2)
    level2_apply(S, D.rc.p, D.uc.p);
  else
    launch_dense_gemv(st, D.cpool.p, D.cfull, D.ntc, D.rc.p, D.cpart.p, D.uc.p);
  // zl_i = K_i W_i r_i + Phi_i uc(indrowc), then z = sum_i R_i^T W_i zl_i
  if (timed) {
    timing_events(D);
    HIPCHECK(hipEventRecord(D.ev_apply[2], st));
  }
  // explicit inverse: yl = S'^{-1} (w .* r) by the symmetric tile GEMV
  const double *ysym = nullptr;
  if (D.explicit_inv) {
    if (D.mixed)
      launch_symv(st, (const float *)D.pool32.p, D.pcols.p, D.nscols, D.srows.p, D.nsrows,
                  D.loff.p, D.map.p, D.w.p, r, D.spart.p, D.yl.p);
    else
      launch_symv(st, (const double *)D.fpool.p, D.pcols.p, D.nscols, D.srows.p, D.nsrows,
                  D.loff.p, D.map.p, D.w.p, r, D.spart.p, D.yl.p);
    ysym = D.yl.p;
  }
  if (D.mixed)
    launch_kapply(st, (const float *)D.pool32.p, D.desc.p, S.nsub, D.map.p, D.w.p, r, D.Phi.p,
                  D.Z.p, D.cmap.p, D.uc.p, ysym, D.zl.p, D.maxNI, D.maxnc, D.kapply_deep);
  else
    launch_kapply(st, (const double *)D.fpool.p, D.desc.p, S.nsub, D.map.p, D.w.p, r, D.Phi.p,
                  D.Z.p, D.cmap.p, D.uc.p, ysym, D.zl.p, D.maxNI, D.maxnc, D.kapply_deep);
  if (timed) {
    HIPCHECK(hipEventRecord(D.ev_apply[3], st));
    D.kapply_timed = true;
    // packed S' factor read by both sweeps, Z and Phi columns, map (4 B),
    // weight, gathered input and output (8 B each) per local interface entry
    double tiles = 0, zb = 0;
    for (const SubDesc &d : D.hdesc) {
      tiles += 0.5 * d.nti * (d.nti + 1.0);
      zb += 2.0 * d.NI * d.nc;
    }
    const double fe = D.mixed ? 4.0 : 8.0;  // factor element size
    // the explicit inverse is read once (symmetric GEMV), the factor twice
    const double passes = D.explicit_inv ? 1.0 : 2.0;
    S.t.kapply_bytes = passes * tiles * TILE_ELEMS * fe + 8.0 * zb + 28.0 * (double)D.LV;
  }
  launch_assemble(st, S.nig, D.aptr.p, D.asrc.p, D.w.p, D.zl.p, z);
}

// dot products into device memory (no host synchronization)
static void dots_dev(Solver &S, const double *a, const double *b, double *dst) {
  DotArgs args{};
  args.ndot = 1;
  args.a[0] = a;
  args.b[0] = b;
  launch_dots(S.dev->st, S.nig, args, S.dev->partial.p, dst);
}

static void dots(Solver &S, std::initializer_list<std::pair<const double *, const double *>> prs,
                 double *out) {
  Device &D = *S.dev;
  DotArgs a{};
  a.ndot = 0;
  for (auto &p : prs) {
    a.a[a.ndot] = p.first;
    a.b[a.ndot] = p.second;
    a.ndot++;
  }
  launch_dots(D.st, S.nig, a, D.partial.p, D.dots.p);
  HIPCHECK(hipMemcpyAsync(out, D.dots.p, a.ndot * sizeof(double), hipMemcpyDeviceToHost, D.st));
  HIPCHECK(hipStreamSynchronize(D.st));
}

void device_recycling_check(Solver &S, int *nvec, double *orth) {
  *nvec = S.rec_n;
  *orth = 0.0;
  if (!S.dev || S.rec_n == 0) return;
  Device &D = *S.dev;
  const int n = S.nig, m = S.rec_n;
  std::vector<double> c(m);
  double f = 0.0;
  for (int j = 0; j < m; j++) {
    launch_multidot(D.st, n, D.recV.p, n, m, D.recW.p + (size_t)j * n, D.recpart.p, D.recc.p);
    HIPCHECK(hipMemcpyAsync(c.data(), D.recc.p, m * sizeof(double), hipMemcpyDeviceToHost, D.st));
    HIPCHECK(hipStreamSynchronize(D.st));
    for (int i = 0; i < m; i++) {
      const double e = c[i] - (i == j ? 1.0 : 0.0);
      f += e * e;
    }
  }
  *orth = std::sqrt(f);
}

void device_apply_schur(Solver &S, const double *x, double *y) {
  Device &D = *S.dev;
  HIPCHECK(hipMemcpy(D.xin.p, x, S.nig * sizeof(double), hipMemcpyHostToDevice));
  op_schur(S, D.xin.p, D.xout.p);
  HIPCHECK(hipStreamSynchronize(D.st));
  HIPCHECK(hipMemcpy(y, D.xout.p, S.nig * sizeof(double), hipMemcpyDeviceToHost));
}

void device_apply_pc(Solver &S, const double *r, double *z) {
  Device &D = *S.dev;
  HIPCHECK(hipMemcpy(D.xin.p, r, S.nig * sizeof(double), hipMemcpyHostToDevice));
  op_pc(S, D.xin.p, D.xout.p);
  HIPCHECK(hipStreamSynchronize(D.st));
  HIPCHECK(hipMemcpy(z, D.xout.p, S.nig * sizeof(double), hipMemcpyDeviceToHost));
}

// eigenvalues of the Lanczos tridiagonal: largest one by Sturm bisection
static double tridiag_max_eig(const std::vector<double> &d, const std::vector<double> &e, int n) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; i++) {
    double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto count_below = [&](double x) {  // number of eigenvalues < x
    int c = 0;
    double q = 1.0;
    for (int i = 0; i < n; i++) {
      double ee = i > 0 ? e[i - 1] * e[i - 1] : 0.0;
      q = d[i] - x - (i > 0 ? ee / q : 0.0);
      if (q == 0.0) q = -1e-300;
      if (q < 0) c++;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-14 * std::max(1.0, std::fabs(hi)); it++) {
    double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= n) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// ------------------------------------------------------------------- Krylov
// The solve ends with the local solutions in device memory (the timed
// result); the host copies are made on first use.
// interface values of every subdomain copy <- u, backward sweep for the
// interior (u_o = A11^{-1}(f_o - A12 u_G)) over the forward sweep left in
// vec, back to local dof order into dst
static void back_substitute(Solver &S, double *dst) {
  Device &D = *S.dev;
  hipStream_t st = D.st;
  if (!D.vec_y) fatal("BDDCML_SOLVE", "forward sweep of the right-hand side missing");
  launch_ifc_scatter(st, S.nig, D.aptr.p, D.rhs_src.p, D.u.p, D.vec.p);
  if (D.f32i)
    launch_trsv_bwd(st, D.ipool32.p, D.trsv.p, S.nsub, D.tcptr.p, D.tck.p, D.tct.p, D.vec.p,
                    D.maxnt);
  else
    launch_trsv_bwd(st, D.pool.p, D.trsv.p, S.nsub, D.tcptr.p, D.tck.p, D.tct.p, D.vec.p, D.maxnt);
  D.vec_y = false;
  launch_from_device_order(st, D.odesc.p, S.nsub, D.sc_perm.p, D.vec.p, dst, D.maxndofs);
  HIPCHECK(hipGetLastError());
}
static void postprocess(Solver &S) {
  Device &D = *S.dev;
  back_substitute(S, D.sol.p);
  HIPCHECK(hipStreamSynchronize(D.st));
  S.sol_on_device = true;
}

void device_fetch_solution(Solver &S) {
  if (!S.sol_on_device || !S.dev) return;
  Device &D = *S.dev;
  hipStream_t st = D.st;
  Staging &h = staging();
  const size_t nb = (size_t)D.ndofs_total * sizeof(double);
  h.used = 0;
  h.grow(nb);
  S.uig.resize(S.nig);
  HIPCHECK(hipMemcpyAsync(S.uig.data(), D.u.p, S.nig * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(h.p, D.sol.p, nb, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  const double *hs = reinterpret_cast<const double *>(h.p);
#pragma omp parallel for schedule(static)
  for (int is = 0; is < S.nsub; is++) {
    Sub &s = S.subs[is];
    s.sol.assign(hs + D.hodesc[is].poff, hs + D.hodesc[is].poff + s.ndofs);
  }
  S.sol_on_device = false;
}

// u0 on the unique interface: each dof's value from its copy in the
// lowest-numbered subdomain holding it (the fixed value at Dirichlet dofs),
// gathered in parallel into the pinned staging area and sent in one copy;
// returns whether u0 is zero
static bool initial_guess(Solver &S) {
  Device &D = *S.dev;
  device_fetch_solution(S);
  if ((int)S.ig_ptr.size() != S.nig + 1) fatal("BDDCML_SOLVE", "interface numbering missing");
  Staging &h = staging();
  h.used = 0;
  h.grow((size_t)S.nig * sizeof(double));
  double *u0 = reinterpret_cast<double *>(h.p);
  int nonzero = 0;
#pragma omp parallel for schedule(static, 4096) reduction(| : nonzero)
  for (int g = 0; g < S.nig; g++) {
    double v = 0.0;
    if (S.ig_ptr[g] < S.ig_ptr[g + 1]) {
      const int q = S.ig_ptr[g];
      const Sub &s = S.subs[S.ig_sub[q]];
      const int d = s.ifc_dof[S.ig_k[q]];
      v = (s.has_fix && s.ifix[d]) ? s.fixv[d] : s.sol[d];
    }
    u0[g] = v;
    nonzero |= v != 0.0;
  }
  HIPCHECK(hipMemcpyAsync(D.u.p, u0, S.nig * sizeof(double), hipMemcpyHostToDevice, D.st));
  // the staging area is reused by the next host transfer
  HIPCHECK(hipStreamSynchronize(D.st));
  return nonzero == 0;
}

// ---------------------------------------------------- PCG iteration graph
static bool g_pcg_graph = true;
void set_pcg_graph(bool on) { g_pcg_graph = on; }

// preconditioner applies counted per level (the automatic explicit inverse
// looks at them); a graph replay runs one apply on every level
static void count_pc_applies(Solver &S, int add) {
  for (Solver *L = &S; L && L->dev; L = L->lev2.get())
    if (L->dev->inv_auto && !L->dev->explicit_inv) L->dev->pc_count += add;
}

// what a captured iteration depends on beyond the solver's fixed buffers
static std::vector<const void *> pcg_graph_key(Solver &S, const double *rz) {
  std::vector<const void *> key{rz, S.dev->hrr_dev};
  for (Solver *L = &S; L && L->dev; L = L->lev2.get())
    key.push_back(L->dev->explicit_inv ? (const void *)L->dev->fpool.p : nullptr);
  return key;
}

// body of one PCG iteration k = *kcount, as the launches of the eager loop:
// ap = S p, p.Ap, u and r updated, r.r (also to the pinned host slot), the
// next direction z = M r, r.z, p = z + beta p, k stepped by one
static void pcg_graph_capture(Solver &S, double *rz, double *pap, double *rr) {
  Device &D = *S.dev;
  hipStream_t st = D.st;
  const int n = S.nig;
  if (D.pcg_exec) {
    HIPCHECK(hipGraphExecDestroy(D.pcg_exec));
    D.pcg_exec = nullptr;
  }
  if (D.pcg_graph) {
    HIPCHECK(hipGraphDestroy(D.pcg_graph));
    D.pcg_graph = nullptr;
  }
  if (!D.kcount.p) D.kcount.alloc(1);
  // op_pc counts its applies on the host when it is called; the capture is
  // not an apply
  std::vector<int> counts;
  for (Solver *L = &S; L && L->dev; L = L->lev2.get()) counts.push_back(L->dev->pc_count);
  HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  op_schur(S, D.p.p, D.ap.p);
  launch_dots_at(st, n, D.p.p, D.ap.p, D.partial.p, pap, D.kcount.p, nullptr);
  launch_pcg_update_at(st, n, rz, pap, D.kcount.p, D.p.p, D.ap.p, D.u.p, D.r.p);
  launch_dots_at(st, n, D.r.p, D.r.p, D.partial.p, rr, D.kcount.p, D.hrr_dev);
  op_pc(S, D.r.p, D.z.p);
  launch_dots_at(st, n, D.r.p, D.z.p, D.partial.p, rz, D.kcount.p, nullptr);
  launch_pcg_dir_at(st, n, rz, D.kcount.p, D.z.p, D.p.p);
  launch_count_inc(st, D.kcount.p);
  HIPCHECK(hipStreamEndCapture(st, &D.pcg_graph));
  HIPCHECK(hipGraphInstantiate(&D.pcg_exec, D.pcg_graph, nullptr, nullptr, 0));
  int q = 0;
  for (Solver *L = &S; L && L->dev; L = L->lev2.get()) L->dev->pc_count = counts[q++];
  D.pcg_key = pcg_graph_key(S, rz);
}

// iteration k by the graph; returns once r.r of iteration k is in hrr[k]
// (the rest of the iteration, the next direction, may still run)
static void pcg_graph_replay(Solver &S, int k, double *rz, double *pap, double *rr) {
  Device &D = *S.dev;
  hipStream_t st = D.st;
  if (!D.pcg_exec || D.pcg_key != pcg_graph_key(S, rz)) pcg_graph_capture(S, rz, pap, rr);
  if (k == 3) launch_count_set(st, D.kcount.p, k);  // later replays step it themselves
  volatile double *h = D.hrr + k;
  *h = -1.0;  // r.r >= 0 replaces it
  HIPCHECK(hipGraphLaunch(D.pcg_exec, st));
  count_pc_applies(S, 1);
  // wait for the value; if the stream has drained without writing it,
  // something failed: stop instead of waiting forever
  for (long polls = 1; *h < 0.0; polls++) {
    if ((polls & 255) == 0) {
      const hipError_t q = hipStreamQuery(st);
      if (q == hipSuccess && *h < 0.0)
        fatal("KRYLOV_BDDCPCG", "PCG iteration graph finished without its residual");
      if (q != hipSuccess && q != hipErrorNotReady) HIPCHECK(q);
    }
  }
}

// Krylov solve of S u = rhs on the interface from the u in D.u (fixed dofs
// keep their values; u_zero: it is zero, so the initial residual is the
// right-hand side without an S apply); appends the relative residuals to
// S.hist and returns the norm of the initial residual in normres0_out (0
// when it vanishes)
static void krylov(Solver &S, const double *rhs, int method, double tol, int maxit, int ndecrmax,
                   bool recycling, int maxvec, bool u_zero, int *num_iter, int *reason,
                   double *cond, double *normres0_out) {
  Device &D = *S.dev;
  hipStream_t st = D.st;
  const int n = S.nig;
  *num_iter = 0;
  *reason = 0;
  *cond = (method == 0) ? 1.0 : -1.0;
  *normres0_out = 0.0;
  HIPCHECK(hipMemcpyAsync(D.r.p, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  launch_mask(st, n, D.fixmask.p, D.r.p);
  double sc[4];
  if (method == 0) {
    dots(S, {{D.r.p, D.r.p}}, sc);
    if (sc[0] == 0.0) goto finish;
    if (!u_zero) {
      op_schur(S, D.u.p, D.ap.p);
      launch_mask(st, n, D.fixmask.p, D.ap.p);
      launch_axpy2(st, n, -1.0, D.ap.p, D.r.p, 0.0, nullptr, nullptr);
      dots(S, {{D.r.p, D.r.p}}, sc);
    }
    double normres0 = std::sqrt(sc[0]);
    if (normres0 == 0.0) goto finish;
    *normres0_out = normres0;
    // Krylov recycling (src/module_krylov.f90:186-216, 308-455, 547-613,
    // 663-726, 1859-2036): the kept directions V (A-orthonormal, W = S V)
    // correct the initial guess by V V^T r0, every new direction is
    // A-orthogonalized against them by two classical Gram-Schmidt passes,
    // and new directions are stored while there is room
    const bool rec = recycling && maxvec > 0 && method == 0;
    if (rec && S.rec_max == 0) {
      S.rec_max = maxvec;
      S.rec_n = 0;
      D.recV.alloc((size_t)n * maxvec);
      D.recW.alloc((size_t)n * maxvec);
      D.recc.alloc(maxvec);
      D.recpart.alloc((size_t)DOT_BLOCKS * maxvec);
    }
    if (rec && S.rec_n > 0) {
      launch_multidot(st, n, D.recV.p, n, S.rec_n, D.r.p, D.recpart.p, D.recc.p);
      HIPCHECK(hipMemsetAsync(D.z.p, 0, n * sizeof(double), st));
      launch_basis_update(st, n, D.recV.p, n, S.rec_n, D.recc.p, 1.0, D.z.p);
      launch_axpy2(st, n, 1.0, D.z.p, D.u.p, 0.0, nullptr, nullptr);
      op_schur(S, D.z.p, D.ap.p);
      launch_axpy2(st, n, -1.0, D.ap.p, D.r.p, 0.0, nullptr, nullptr);
      launch_mask(st, n, D.fixmask.p, D.r.p);
      dots(S, {{D.r.p, D.r.p}}, sc);
      if (std::sqrt(sc[0]) / normres0 < tol) {
        *num_iter = 0;
        *reason = 0;
        *cond = 1.0;
        goto finish;
      }
    }
    op_pc(S, D.r.p, D.p.p);
    // scalars stay on the device: rz[k] = r.z, pap[k] = p.Ap, rr[k] = r.r
    // of iteration k; the host reads only rr[k], and while it waits the
    // device already applies the preconditioner for the next direction
    // (work that is simply unused when the residual test stops)
    // (kept across solves while large enough, so that a captured iteration
    // graph stays valid)
    const size_t m = (size_t)maxit + 2;
    if (D.pcg.n < 3 * m) D.pcg.alloc(3 * m);
    const size_t mc = D.pcg.n / 3;
    double *rz = D.pcg.p, *pap = D.pcg.p + mc, *rr = D.pcg.p + 2 * mc;
    if (D.hrr_n < m) {
      if (D.hrr) (void)hipHostFree(D.hrr), D.hrr = nullptr;
      // mapped and coherent: the graph's dot kernel writes r.r here and the
      // host polls it
      HIPCHECK(hipHostMalloc((void **)&D.hrr, m * sizeof(double),
                             hipHostMallocMapped | hipHostMallocCoherent));
      HIPCHECK(hipHostGetDevicePointer((void **)&D.hrr_dev, D.hrr, 0));
      D.hrr_n = m;
    }
    dots_dev(S, D.r.p, D.p.p, rz);
    hipEvent_t evr = nullptr;
    HIPCHECK(hipEventCreateWithFlags(&evr, hipEventDisableTiming));
    int ndecr = 0, nw = 0, last = 0;
    double lastres = 1.0;
    // iterations from the third on replay one captured graph of the
    // iteration (recycling keeps the launches: its basis grows each step)
    const bool graph = g_pcg_graph && !rec;
    for (int iter = 1; iter <= maxit; iter++) {
      if (graph && iter >= 3) {
        pcg_graph_replay(S, iter, rz, pap, rr);
        goto residual_test;
      }
      if (rec && S.rec_n > 0)
        for (int pass = 0; pass < 2; pass++) {
          launch_multidot(st, n, D.recW.p, n, S.rec_n, D.p.p, D.recpart.p, D.recc.p);
          launch_basis_update(st, n, D.recV.p, n, S.rec_n, D.recc.p, -1.0, D.p.p);
        }
      op_schur(S, D.p.p, D.ap.p, iter == 2);
      dots_dev(S, D.p.p, D.ap.p, pap + iter);
      if (rec && S.rec_n < S.rec_max) {
        launch_basis_store(st, n, pap + iter, D.p.p, D.ap.p, D.recV.p + (size_t)S.rec_n * n,
                           D.recW.p + (size_t)S.rec_n * n);
        S.rec_n++;
      }
      launch_pcg_update(st, n, rz, pap, iter, D.p.p, D.ap.p, D.u.p, D.r.p);
      dots_dev(S, D.r.p, D.r.p, rr + iter);
      HIPCHECK(hipMemcpyAsync(D.hrr + iter, rr + iter, sizeof(double), hipMemcpyDeviceToHost, st));
      HIPCHECK(hipEventRecord(evr, st));
      // next direction, computed ahead of the residual test
      op_pc(S, D.r.p, D.z.p, iter == 2);
      dots_dev(S, D.r.p, D.z.p, rz + iter);
      launch_pcg_dir(st, n, rz, iter, D.z.p, D.p.p);
      HIPCHECK(hipEventSynchronize(evr));
    residual_test:
      last = iter;
      double relres = std::sqrt(D.hrr[iter]) / normres0;
      S.hist.push_back(relres);
      if (S.verbose > 0) std::printf(" KRYLOV_BDDCPCG: iteration: %d relative residual: %.6e\n", iter, relres);
      if (relres < tol) {
        nw = iter - 1;
        *num_iter = iter;
        *reason = 0;
        break;
      }
      if (iter == maxit) {
        nw = iter - 1;
        *num_iter = iter;
        *reason = -1;
        break;
      }
      if (relres < lastres) ndecr = 0;
      else if (++ndecr >= ndecrmax) {
        nw = iter - 1;
        *num_iter = iter;
        *reason = -2;
        break;
      }
      lastres = relres;
    }
    (void)hipEventDestroy(evr);
    // Lanczos tridiagonal of the finished iterations for the condition
    // number estimate (src/module_krylov.f90)
    std::vector<double> hrz(m, 0.0), hpap(m, 0.0);
    HIPCHECK(hipMemcpyAsync(hrz.data(), rz, m * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(hpap.data(), pap, m * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<double> diag(maxit + 2, 0.0), sub(maxit + 2, 0.0);
    bool pap_neg = false, rz_neg = false;
    for (int k = 1; k <= last; k++) {
      pap_neg |= hpap[k] <= 0.0;
      if (k <= nw) rz_neg |= hrz[k] <= 0.0;
    }
    if (S.verbose > 0 && pap_neg)
      std::fprintf(stderr, "WARNING in KRYLOV_BDDCPCG: System matrix not positive definite!\n");
    if (S.verbose > 0 && rz_neg)
      std::fprintf(stderr, "WARNING in KRYLOV_BDDCPCG: Preconditioner not positive definite!\n");
    for (int k = 1; k <= nw; k++) {
      const double alpha = hrz[k - 1] / hpap[k], beta = hrz[k] / hrz[k - 1];
      diag[k - 1] += 1.0 / alpha;
      diag[k] = beta / alpha;
      sub[k - 1] = beta >= 0.0 ? -std::sqrt(beta) / alpha : 0.0;
    }
    *cond = nw > 1 ? tridiag_max_eig(diag, sub, nw) : 1.0;
  } else {
    // BiCGSTAB (src/module_krylov.f90:948-1505)
    if (!u_zero) {
      op_schur(S, D.u.p, D.v.p);
      launch_axpy2(st, n, -1.0, D.v.p, D.r.p, 0.0, nullptr, nullptr);
      launch_mask(st, n, D.fixmask.p, D.r.p);
    }
    dots(S, {{D.r.p, D.r.p}}, sc);
    double normrhs = std::sqrt(sc[0]);
    if (normrhs == 0.0) goto finish;
    *normres0_out = normrhs;
    double rhoold = 1.0, alpha = 1.0, omega = 1.0;
    HIPCHECK(hipMemsetAsync(D.v.p, 0, n * sizeof(double), st));
    HIPCHECK(hipMemsetAsync(D.p.p, 0, n * sizeof(double), st));
    HIPCHECK(hipMemcpyAsync(D.rs.p, D.r.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    int ndecr = 0;
    double lastres = 1.0;
    for (int iter = 1; iter <= maxit; iter++) {
      dots(S, {{D.r.p, D.rs.p}}, sc);
      double rho = sc[0];
      double beta = rho * alpha / (rhoold * omega);
      launch_bicg_p(st, n, beta, omega, D.r.p, D.v.p, D.p.p);
      op_pc(S, D.p.p, D.y.p);
      op_schur(S, D.y.p, D.v.p);
      dots(S, {{D.v.p, D.rs.p}}, sc);
      alpha = rho / sc[0];
      launch_axpy2(st, n, alpha, D.y.p, D.u.p, 0.0, nullptr, nullptr);
      launch_sub_scaled(st, n, D.r.p, alpha, D.v.p, D.s.p);
      dots(S, {{D.s.p, D.s.p}}, sc);
      double relres = std::sqrt(sc[0]) / normrhs;
      S.hist.push_back(relres);
      if (relres < tol) {
        *num_iter = iter;
        *reason = 0;
        break;
      }
      if (relres < lastres) ndecr = 0;
      else if (++ndecr >= ndecrmax) {
        *num_iter = iter;
        *reason = -2;
        break;
      }
      lastres = relres;
      op_pc(S, D.s.p, D.z.p);
      op_schur(S, D.z.p, D.t.p);
      dots(S, {{D.s.p, D.t.p}, {D.t.p, D.t.p}}, sc);
      omega = sc[0] / sc[1];
      launch_axpy2(st, n, omega, D.z.p, D.u.p, 0.0, nullptr, nullptr);
      launch_sub_scaled(st, n, D.s.p, omega, D.t.p, D.r.p);
      dots(S, {{D.r.p, D.r.p}}, sc);
      relres = std::sqrt(sc[0]) / normrhs;
      S.hist.push_back(relres);
      if (relres < tol) {
        *num_iter = iter;
        *reason = 0;
        break;
      }
      if (iter == maxit) {
        *num_iter = iter;
        *reason = -1;
        break;
      }
      if (relres < lastres) ndecr = 0;
      else if (++ndecr >= ndecrmax) {
        *num_iter = iter;
        *reason = -2;
        break;
      }
      lastres = relres;
      rhoold = rho;
    }
  }
finish:
  HIPCHECK(hipStreamSynchronize(st));
}

// FP64 iterative refinement over the FP32 factorizations (the interior
// factors, the Schur complements S~ and the S' factors).  The FP64 residual
// of the subdomain problems r = f - A u (uploaded triplets, local order) is
// condensed by the forward sweep; its interface part g_k = g - S u_G is the
// residual of the interface problem with the exact Schur complements (the
// interior solves of the sweep cancel out of it), so ||g_k|| / ||r_0|| is
// the relative residual the FP64 method stops on.  Each step solves
// S~ d = g_k from d = 0 to the tolerance the step needs, back-substitutes
// the correction in the interior and adds it to the solution.
static void refine_solution(Solver &S, int method, double tol, int maxit, int ndecrmax,
                            double normres0, int *num_iter, int *reason) {
  Device &D = *S.dev;
  hipStream_t st = D.st;
  const int n = S.nig;
  if (D.rloc.n < (size_t)D.ndofs_total) D.rloc.alloc(D.ndofs_total);
  if (D.gk.n < (size_t)n) {
    D.gk.alloc(n);
    D.uacc.alloc(n);
  }
  if (D.ndofs_total > INT32_MAX) fatal("BDDCML_SOLVE", "too many subdomain dofs for the refinement");
  int total = *num_iter;
  double sc[4];
  double tph = wall_time();
  auto ph = [&](const char *name, int k) {
    if (S.verbose < 2) return;
    HIPCHECK(hipStreamSynchronize(st));
    const double now = wall_time();
    std::printf("   refinement step %d: %-28s %8.4f s\n", k, name, now - tph);
    tph = now;
  };
  for (int k = 0;; k++) {
    if (k == 0) {
      timing_events(D);
      HIPCHECK(hipEventRecord(D.ev_apply[4], st));
    }
    launch_resid_rows(st, D.rs_subs.p, S.nsub, D.rs_ptr.p, D.rs_ent.p, D.sc_fix.p, D.fmod.p,
                      D.sol.p, D.rloc.p, D.maxndofs);
    if (k == 0) HIPCHECK(hipEventRecord(D.ev_apply[5], st));
    reduced_rhs_launches(S, st, D.rloc.p, D.gk.p);
    HIPCHECK(hipMemcpyAsync(D.r.p, D.gk.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    launch_mask(st, n, D.fixmask.p, D.r.p);
    dots(S, {{D.r.p, D.r.p}}, sc);
    if (k == 0) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, D.ev_apply[4], D.ev_apply[5]));
      S.t.ev_resid_ms = ms;
      // triplet values, flags, f, u and r of every dof (the class indices
      // are shared by the subdomains of a class)
      double nnz = 0;
      for (int is = 0; is < S.nsub; is++) nnz += (double)S.subs[is].ai.size();
      S.t.resid_bytes = 8.0 * nnz + 25.0 * (double)D.ndofs_total;
    }
    const double rel = std::sqrt(sc[0]) / normres0;
    S.t.refine_relres = rel;
    ph("residual and condensation", k);
    if (S.verbose > 0)
      std::printf(" BDDCML_SOLVE: refinement step %d: relative residual %.6e\n", k, rel);
    if (rel < tol) {
      *reason = 0;
      break;
    }
    if (k >= S.refine_max || total >= maxit) {
      *reason = -1;
      break;
    }
    HIPCHECK(hipMemcpyAsync(D.uacc.p, D.u.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemsetAsync(D.u.p, 0, n * sizeof(double), st));
    const double inner = std::min(0.5, 0.5 * tol / rel);
    const size_t h0 = S.hist.size();
    int it = 0, rs = 0;
    double c = 0.0, nr = 0.0;
    krylov(S, D.gk.p, method, inner, maxit - total, ndecrmax, false, 0, true, &it, &rs, &c, &nr);
    for (size_t q = h0; q < S.hist.size(); q++) S.hist[q] *= nr / normres0;
    total += it;
    S.t.refinements++;
    if (S.verbose >= 2) std::printf("   refinement step %d: correction of %d iterations\n", k, it);
    ph("correction solve", k);
    back_substitute(S, D.rloc.p);
    launch_axpy2(st, (int)D.ndofs_total, 1.0, D.rloc.p, D.sol.p, 0.0, nullptr, nullptr);
    launch_axpy2(st, n, 1.0, D.uacc.p, D.u.p, 0.0, nullptr, nullptr);
    ph("back substitution and update", k);
    if (rs == -2) {
      *reason = -2;
      break;
    }
  }
  *num_iter = total;
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
  S.sol_on_device = true;
}

void device_solve(Solver &S, int method, double tol, int maxit, int ndecrmax, bool recycling,
                  int maxvec, int *num_iter, int *reason, double *cond) {
  Device &D = *S.dev;
  double t0 = wall_time();
  S.hist.clear();
  S.t.refinements = 0;
  S.t.refine_relres = 0.0;
  inverse_at_solve_start(S);
  // a repeated solve of the same right-hand side redoes its forward sweep
  // (the back substitution of the last one overwrote it)
  if (!D.vec_y) reduced_rhs_launches(S, D.st);
  const bool u_zero = initial_guess(S);
  // with refinement the first solve stops early (at the square root of the
  // tolerance unless set): the FP32 operator's own error caps what more
  // iterations on it would give
  const bool refine = D.f32i;
  const double tol1 =
      refine ? std::max(tol, S.refine_first > 0.0 ? S.refine_first : std::sqrt(tol)) : tol;
  double normres0 = 0.0;
  krylov(S, D.g.p, method, tol1, maxit, ndecrmax, recycling && !refine, maxvec, u_zero, num_iter,
         reason, cond, &normres0);
  read_apply_timings(S);
  S.t.krylov += wall_time() - t0;
  t0 = wall_time();
  postprocess(S);
  S.t.postprocess += wall_time() - t0;
  // refined also after a stagnating first solve (a first solve from a
  // converged initial guess stagnates at rounding level): the corrections
  // bring the interior back-substitution to FP64 accuracy
  if (refine && *reason != -1 && normres0 > 0.0) {
    t0 = wall_time();
    refine_solution(S, method, tol, maxit, ndecrmax, normres0, num_iter, reason);
    S.t.refine += wall_time() - t0;
  }
}

}  // namespace bddc
