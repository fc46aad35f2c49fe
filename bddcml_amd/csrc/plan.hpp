// plan.hpp -- host-side data of the MI355X BDDC library.
//
// The host owns what the reference keeps in `subdomain_type`
// (src/module_dd.f90:65-372): mesh maps, interface, globs, constraints,
// weights.  It turns them into a "plan": flat index arrays that drive the
// device kernels (tile patterns of every subdomain matrix, batched task
// lists of the tile factorisations, gather/scatter maps of the interface).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <utility>
#include <string>
#include <vector>

namespace bddc {

constexpr int TILE = 64;                   // tile edge (doubles)
constexpr int TILE_ELEMS = TILE * TILE;    // 4096

// Errors print "ERROR in <ROUTINE>: <message>" and abort, as the reference
// does (src/module_utils.f90:265-299); with g_throw_on_error they throw
// Error instead and the C entry points record the message.
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
extern bool g_throw_on_error;
[[noreturn]] void fatal(const char *routine, const std::string &msg);

struct Sub {
  // ---- uploaded (0-based indices) -------------------------------------
  int isub = -1;
  int nelems = 0, nnods = 0, ndofs = 0;
  std::vector<int> inet, nnet, nndf, isngn, isvgvn;
  std::vector<int> isegn;   // global element ids (when uploaded)
  std::vector<double> xyz;  // nnods x ndim, column major
  std::vector<int> ifix;
  std::vector<double> fixv, rhs, sol;
  bool rhs_complete = false, has_fix = false, uploaded = false;
  std::vector<double> dof_data;
  // user constraints: nuser_c rows of one value per local node (column
  // major, nuser_c x nnods)
  std::vector<double> user_c;
  int nuser_c = 0;
  int find_components = 0, use_dual = 0, neighbouring = 1;
  // assembled upper triangle (sorted by row, col; duplicates summed)
  std::vector<int> ai, aj;
  std::vector<double> av;

  // ---- topology --------------------------------------------------------
  std::vector<int> kdof;    // first dof of each node (nnods+1)
  std::vector<int> ncomp;   // nodal components
  std::vector<std::pair<int, int>> g2l;  // (global node, local node), sorted
  std::vector<int> adj;     // neighbouring subdomains, ascending
  std::vector<int> iin;     // interface nodes (local node index, ascending)
  std::vector<int> ifc_dof; // interface dofs (local dof), reference order
  std::vector<int> int_dof; // interior dofs (local dof), reference order
  std::vector<int> kdofi;   // first interface dof of each interface node
  // shared interface nodes per neighbour, sorted by global node id
  std::vector<int> shptr, shnode, shcomp;
  std::vector<int> nsubnode;               // number of neighbours sharing a node
  // glob key of each interface node: its (neighbour, its comp, my comp)
  // triples in neighbour order, node i's at nkval[nkptr[i] .. nkptr[i+1])
  std::vector<int> nkptr, nkval;
  std::vector<int> lcorner;                // locally selected corners (global nodes)
  std::vector<int> corners;                // final corners (sorted global nodes)
  struct Glob {
    int type = 0;            // 2 edge, 1 face
    int key = 0;             // smallest global node id
    std::vector<int> nodes;  // interface node indices
  };
  std::vector<Glob> globs;

  // ---- constraints (sparse rows over local interface dofs) -------------
  int nconstr = 0;
  std::vector<int> cptr, cidx;  // CSR, cidx = interface dof index (reference order)
  std::vector<double> cval;
  std::vector<int> indrowc;     // global coarse dof of each row
  std::vector<int> cnodes;      // global coarse node of each local one (corners, then globs)

  std::vector<double> bc;   // eliminated Dirichlet values, ndofs
  // triplet entries touching fixed dofs with their values before the
  // elimination, kept to eliminate new Dirichlet values later
  std::vector<size_t> bc_k;
  std::vector<double> bc_a;
  std::vector<double> rea;  // reactions of the last solve (ndofs)
  std::vector<double> w;    // interface weights (reference order)
  std::vector<double> fmod; // right-hand side used for condensation, ndofs

  // ---- device ordering --------------------------------------------------
  int no = 0, ni = 0;       // interior / interface dof counts
  int NO = 0, NI = 0;       // padded to TILE multiples
  int nto = 0, nti = 0;     // tile counts (interior, interface)
  std::vector<int> perm;    // device position of each local dof
  std::vector<int> dev_ifc; // interface device index k (0..ni-1) -> reference interface index
  std::vector<int> pad;     // interior device positions holding no dof (unit diagonal)
};

// One batched tile task: C = alpha*C + beta*sum(op(A_p) op(B_p)), optional
// post multiply by tile X, optional Cholesky.  Pair tile ids are relative to
// base, so subdomains of identical structure share one pair list.
struct TileTask {
  int c;          // output tile id
  int x;          // post-multiply / inverse tile id (-1 none)
  int p0, p1;     // pair range
  int base = 0;   // added to the tile ids of the pairs
  int fresh = 0;  // 1: the output tile holds no matrix entries; the task
                  // starts from zero and never reads it (it is not cleared)
  int mode = -1;  // >= 0: this task's product kind, post step and beta
                  // instead of the launch's (task_mode), so that tasks of
                  // different kinds share a launch
};
// per-task mode: kind | post << 2 | (beta == -1) << 4; alpha is 1
constexpr int task_mode(int kind, int post, bool minus) { return kind | post << 2 | (minus ? 16 : 0); }
// task mode of a tile factorization in a mixed launch: C -= sum of its pairs
// A B^T, then C = chol(C) and tile x = inv(chol(C))
constexpr int kTaskPotrf = 64;
struct TilePair {
  int a, b;
};

struct TaskList {
  std::vector<TileTask> tasks;
  std::vector<TilePair> pairs;
  void clear() {
    tasks.clear();
    pairs.clear();
  }
};

// Symbolic tile pattern of one lower-triangular block-sparse matrix.
struct TilePattern {
  int nt = 0;                  // tile rows / cols
  int ntf = 0;                 // columns that get factored
  std::vector<int> colptr;     // CSC over lower tiles (row >= col)
  std::vector<int> rowidx;
  std::vector<int> tile;       // tile id of each entry
  int find(int I, int J) const;  // tile id or -1
};

}  // namespace bddc
