/*
 * bddc_oracle.c -- CPU FP64 restatement of the reference's two-level BDDC
 * hot path.  TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links it.
 *
 * What each block follows (reference paths relative to the repository root):
 *   upload / assembly ........ src/module_levels.f90:615-782, src/module_sm.f90:630-685
 *   Dirichlet elimination .... src/module_sm.f90:431-600, src/module_dd.f90:1081-1256
 *   interface / neighbours ... src/module_dd.f90:8071-8647
 *   globs and corners ........ src/module_dd.f90:8925-10554
 *   weights .................. src/module_dd.f90:11554-11725
 *   constraints .............. src/module_dd.f90:4298-4384, 4869-5145, 11011-11234
 *   local factorisations ..... src/module_dd.f90:4175-4296 (explicit Schur),
 *                              5211-5487 (aug, dense LU path 5330), 5489-5800 (coarse basis)
 *   coarse problem ........... src/module_levels.f90:3001-3169
 *   reduced rhs .............. src/module_dd.f90:6995-7258
 *   Schur / BDDC apply ....... src/module_levels.f90:3627-4346, 4443-4507
 *   PCG / BiCGSTAB / cond .... src/module_krylov.f90:48-945, 948-1505, 2061-2118
 *   interior recovery ........ src/module_dd.f90:6811-6932
 *
 * Pinning (what checks this file): tests/test_oracle_dense.py compares the
 * Schur complements, coarse basis, coarse matrices, preconditioner and the
 * converged solution against dense numpy algebra on the assembled global
 * matrix of small cubes; tests/test_oracle_globs.py checks the glob/corner
 * classification against an independent Python restatement and against
 * hand-counted structured-cube values; invariants (partition of unity of
 * weights, S*1 = 0 on floating subdomains, lambda_min(MS) >= 1) are tested too.
 *
 *   coarse nodes / level mesh . src/module_dd.f90:4869-5016 (cnode coordinates),
 *                              src/module_levels.f90:1484-1530 (next level's mesh)
 *   inexact coarse solve ..... src/module_levels.f90:3849-4249 (orc_apply_level)
 *
 * Restrictions: matrixtype 1 (SPD) only; weights types 0, 1 and 3;
 * corner and arithmetic (edge/face average) constraints.  Three levels are
 * composed by the caller (oracle/multilevel.py): a second instance holds the
 * next level's mesh and replaces the coarse solve through
 * orc_set_coarse_solver.
 */
#include "bddc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ONE_MAX(a, b) ((a) > (b) ? (a) : (b))
#define ONE_MIN(a, b) ((a) < (b) ? (a) : (b))

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) {
    fprintf(stderr, "ERROR in ORACLE: out of memory\n");
    abort();
  }
  return p;
}
static int *imalloc(size_t n) { return (int *)xcalloc(n, sizeof(int)); }
static double *dmalloc(size_t n) { return (double *)xcalloc(n, sizeof(double)); }

typedef struct {
  int uploaded;
  int nelems, nnods, ndofs;
  int *inet, linet, *nnet, *nndf, *isngn, *isvgvn;
  double *xyz; /* nnods x ndim, column major like xyz(lxyz1,lxyz2) */
  int *ifix;
  double *fixv, *rhs, *sol0;
  int is_rhs_complete, has_fix;
  double *dof_data;
  int ldof_data;
  /* assembled upper triangle (matrixtype 1) */
  int nnz, *mi, *mj;
  double *ma;
  int find_components, use_dual, neighbouring;
  int *ncomp; /* nodal components */
  int *kdof;  /* first dof of node */
  /* neighbours / interface */
  int nadj, *iadj, *kshn, *ishn, *ishncadj; /* CSR over neighbours */
  int nnodi, ndofi, ndofo, *iin, *iivsvn, *iovsvn;
  int *kdofi; /* first interface dof of interface node */
  int *nsubnode, *gsubs, nsubnx;
  int *globtypes, *kglobs;
  int nlcorner, *lcorner; /* locally selected corners (global node ids) */
  int ncorner, *icgn;     /* corners after merge (sorted global node ids) */
  int nedges, nfaces;
  /* globs: ng = nedges + nfaces, CSR of interface node indices */
  int ng, *gtype, *gptr, *gnod, *gkey;
  /* coarse nodes and constraints */
  int nconstr;
  int ncn, *cnid; /* global coarse node of each local one (corners, then globs) */
  double *C; /* nconstr x ndofi row major */
  int *indrowc;
  double *wi;
  double *bc; /* eliminated BC vector, ndofs */
  /* numerics */
  int *sky_first, *sky_off;
  double *sky; /* skyline Cholesky factor of A11 */
  double *S;   /* ndofi^2 */
  double *aug; /* LU of [S C^T; C 0] */
  int *augpiv;
  double *phi;  /* ndofi x nconstr */
  double *acs;  /* nconstr^2 */
  double *g;    /* reduced rhs (assembled) */
  double *sols; /* final local solution */
  int *gdofi;   /* global dof of each local interface dof */
} osub;

typedef struct {
  int nsub, nnod_g, ndof_g, ndim, meshdim, numbase;
  osub *s;
  int matrixtype;
  int *n2s_ptr, *n2s_idx; /* global node -> subdomains */
  int ncorner, nedge, nface, nnodc, ndofc;
  int *nndfc, *kdofc;
  double *Ac; /* dense coarse matrix, Cholesky factor (lower) */
  int is_setup, bc_global;
  int nig, *igdof, *gpos; /* unique interface dofs, gdof -> position */
  double *gbuf;
  double *uig;                /* last interface solution (unique) */
  double *hist;
  int nhist;
  /* coarse solve replaced by the next level's inexact one (three levels) */
  orc_coarse_fn cfn;
  void *cctx;
} oprob;

/* ------------------------------------------------------------------ utils */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}
typedef struct {
  int i, j;
  double a;
} trip;
static int cmp_trip(const void *a, const void *b) {
  const trip *x = (const trip *)a, *y = (const trip *)b;
  if (x->i != y->i) return (x->i > y->i) - (x->i < y->i);
  return (x->j > y->j) - (x->j < y->j);
}
static int lower_bound(const int *a, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int m = (lo + hi) / 2;
    if (a[m] < v)
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}

void *orc_create(int nsub, int nnod_global, int ndof_global, int ndim,
                 int meshdim, int numbase) {
  oprob *P = (oprob *)xcalloc(1, sizeof(oprob));
  P->nsub = nsub;
  P->nnod_g = nnod_global;
  P->ndof_g = ndof_global;
  P->ndim = ndim;
  P->meshdim = meshdim;
  P->numbase = numbase;
  P->s = (osub *)xcalloc(nsub, sizeof(osub));
  P->matrixtype = -1;
  return P;
}

static void free_sub(osub *s) {
  void *ptrs[] = {s->inet, s->nnet, s->nndf, s->isngn, s->isvgvn, s->xyz,
                  s->ifix, s->fixv, s->rhs, s->sol0, s->dof_data, s->mi,
                  s->mj, s->ma, s->ncomp, s->kdof, s->iadj, s->kshn, s->ishn,
                  s->ishncadj, s->iin, s->iivsvn, s->iovsvn, s->kdofi,
                  s->nsubnode, s->gsubs, s->globtypes, s->kglobs, s->lcorner,
                  s->icgn, s->gtype, s->gptr, s->gnod, s->gkey, s->C,
                  s->indrowc, s->wi, s->bc, s->sky_first, s->sky_off, s->sky,
                  s->S, s->aug, s->augpiv, s->phi, s->acs, s->g, s->sols,
                  s->gdofi, s->cnid};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
}

void orc_destroy(void *h) {
  oprob *P = (oprob *)h;
  if (!P) return;
  for (int i = 0; i < P->nsub; i++) free_sub(&P->s[i]);
  free(P->s);
  free(P->n2s_ptr);
  free(P->n2s_idx);
  free(P->nndfc);
  free(P->kdofc);
  free(P->Ac);
  free(P->igdof);
  free(P->gpos);
  free(P->gbuf);
  free(P->uig);
  free(P->hist);
  free(P);
}

/* --------------------------------------------------------------- upload */
/* assembly of triplets: sort by (i,j), sum duplicates (module_sm.f90:630-685) */
static int assemble_triplets(trip *t, int n) {
  if (n <= 0) return 0;
  qsort(t, n, sizeof(trip), cmp_trip);
  int w = 0;
  for (int k = 1; k < n; k++) {
    if (t[k].i == t[w].i && t[k].j == t[w].j)
      t[w].a += t[k].a;
    else
      t[++w] = t[k];
  }
  return w + 1;
}

int orc_upload_subdomain(void *h, int isub, int nelems, int nnods, int ndofs,
                         const int *inet, int linet, const int *nnet,
                         const int *nndf, const int *isngn, const int *isvgvn,
                         const double *xyz, int lxyz1, int lxyz2,
                         const int *ifix, const double *fixv,
                         const double *rhs, int is_rhs_complete,
                         const double *sol, int matrixtype,
                         const int *i_sparse, const int *j_sparse,
                         const double *a_sparse, int la, int is_assembled,
                         const double *dof_data, int ldof_data,
                         int find_components, int use_dual_mesh_graph,
                         int neighbouring) {
  oprob *P = (oprob *)h;
  int b = P->numbase;
  int is = isub - b;
  (void)is_assembled;
  if (is < 0 || is >= P->nsub) return -1;
  if (matrixtype != 1) {
    fprintf(stderr, "ERROR in ORACLE: only matrixtype 1 is supported\n");
    return -2;
  }
  if (lxyz1 != nnods || lxyz2 != P->ndim) return -3;
  P->matrixtype = matrixtype;
  osub *s = &P->s[is];
  if (s->uploaded) free_sub(s), memset(s, 0, sizeof(osub));
  s->uploaded = 1;
  s->nelems = nelems;
  s->nnods = nnods;
  s->ndofs = ndofs;
  s->linet = linet;
  s->inet = imalloc(linet);
  for (int k = 0; k < linet; k++) s->inet[k] = inet[k] - b;
  s->nnet = imalloc(nelems);
  memcpy(s->nnet, nnet, sizeof(int) * nelems);
  s->nndf = imalloc(nnods);
  memcpy(s->nndf, nndf, sizeof(int) * nnods);
  s->isngn = imalloc(nnods);
  for (int k = 0; k < nnods; k++) s->isngn[k] = isngn[k] - b;
  s->isvgvn = imalloc(ndofs);
  for (int k = 0; k < ndofs; k++) s->isvgvn[k] = isvgvn[k] - b;
  s->xyz = dmalloc((size_t)nnods * P->ndim);
  memcpy(s->xyz, xyz, sizeof(double) * nnods * P->ndim);
  s->ifix = imalloc(ndofs);
  s->fixv = dmalloc(ndofs);
  s->rhs = dmalloc(ndofs);
  s->sol0 = dmalloc(ndofs);
  s->has_fix = 0;
  for (int k = 0; k < ndofs; k++) {
    s->ifix[k] = ifix[k];
    s->fixv[k] = fixv[k];
    s->rhs[k] = rhs[k];
    s->sol0[k] = sol[k];
    if (ifix[k] != 0) s->has_fix = 1;
  }
  s->is_rhs_complete = is_rhs_complete;
  s->ldof_data = ldof_data;
  s->dof_data = dmalloc(ldof_data);
  if (ldof_data) memcpy(s->dof_data, dof_data, sizeof(double) * ldof_data);
  /* upper triangle check (module_levels.f90:748-752) */
  trip *t = (trip *)xcalloc(la + ndofs, sizeof(trip));
  for (int k = 0; k < la; k++) {
    t[k].i = i_sparse[k] - b;
    t[k].j = j_sparse[k] - b;
    t[k].a = a_sparse[k];
    if (t[k].i > t[k].j) {
      fprintf(stderr, "ERROR in ORACLE: only upper triangle is expected\n");
      free(t);
      return -4;
    }
  }
  int nnz = assemble_triplets(t, la);
  s->nnz = nnz;
  s->mi = imalloc(nnz + ndofs);
  s->mj = imalloc(nnz + ndofs);
  s->ma = dmalloc(nnz + ndofs);
  for (int k = 0; k < nnz; k++) {
    s->mi[k] = t[k].i;
    s->mj[k] = t[k].j;
    s->ma[k] = t[k].a;
  }
  free(t);
  s->find_components = find_components;
  s->use_dual = use_dual_mesh_graph;
  s->neighbouring = neighbouring;
  s->kdof = imalloc(nnods + 1);
  for (int k = 0; k < nnods; k++) s->kdof[k + 1] = s->kdof[k] + nndf[k];
  return 0;
}

/* ----------------------------------------------------------- components */
/* connected components by DFS over a CSR graph, numbered in order of the
 * first vertex (module_graph.f90:1052-1125) */
static int graph_components(int n, const int *xadj, const int *adj, int *comp) {
  int nc = 0;
  int *stack = imalloc(n);
  for (int v = 0; v < n; v++) comp[v] = 0;
  for (int v = 0; v < n; v++) {
    if (comp[v]) continue;
    nc++;
    int top = 0;
    stack[top++] = v;
    comp[v] = nc;
    while (top) {
      int u = stack[--top];
      for (int k = xadj[u]; k < xadj[u + 1]; k++) {
        int w = adj[k];
        if (!comp[w]) {
          comp[w] = nc;
          stack[top++] = w;
        }
      }
    }
  }
  free(stack);
  return nc;
}

/* nodal components (module_dd.f90:2662-2764) */
static void sub_components(osub *s) {
  s->ncomp = imalloc(s->nnods);
  if (!s->find_components) {
    for (int k = 0; k < s->nnods; k++) s->ncomp[k] = 1;
    return;
  }
  int ne = s->nelems, nn = s->nnods;
  int *kinet = imalloc(ne + 1);
  for (int e = 0; e < ne; e++) kinet[e + 1] = kinet[e] + s->nnet[e];
  /* node -> elements */
  int *nptr = imalloc(nn + 1), *nel = imalloc(s->linet);
  for (int k = 0; k < s->linet; k++) nptr[s->inet[k] + 1]++;
  for (int k = 0; k < nn; k++) nptr[k + 1] += nptr[k];
  int *fill = imalloc(nn);
  for (int e = 0; e < ne; e++)
    for (int k = kinet[e]; k < kinet[e + 1]; k++) {
      int v = s->inet[k];
      nel[nptr[v] + fill[v]++] = e;
    }
  if (s->use_dual) {
    /* dual graph: elements adjacent if they share >= neighbouring nodes */
    int cap = 16, cnt = 0;
    int *xadj = imalloc(ne + 1), *adj = imalloc(cap);
    int *shared = imalloc(ne), *touched = imalloc(ne);
    for (int e = 0; e < ne; e++) {
      int nt = 0;
      for (int k = kinet[e]; k < kinet[e + 1]; k++) {
        int v = s->inet[k];
        for (int q = nptr[v]; q < nptr[v + 1]; q++) {
          int f = nel[q];
          if (f == e) continue;
          if (shared[f] == 0) touched[nt++] = f;
          shared[f]++;
        }
      }
      for (int q = 0; q < nt; q++) {
        int f = touched[q];
        if (shared[f] >= s->neighbouring) {
          if (cnt == cap) {
            cap *= 2;
            adj = (int *)realloc(adj, sizeof(int) * cap);
          }
          adj[cnt++] = f;
        }
        shared[f] = 0;
      }
      xadj[e + 1] = cnt;
    }
    int *ec = imalloc(ne);
    graph_components(ne, xadj, adj, ec);
    for (int k = 0; k < nn; k++) s->ncomp[k] = 1;
    /* copy element components into nodes: last element wins */
    for (int e = 0; e < ne; e++)
      for (int k = kinet[e]; k < kinet[e + 1]; k++) s->ncomp[s->inet[k]] = ec[e];
    free(ec);
    free(xadj);
    free(adj);
    free(shared);
    free(touched);
  } else {
    /* primal graph: nodes adjacent if they share an element */
    int cap = 16, cnt = 0;
    int *xadj = imalloc(nn + 1), *adj = imalloc(cap);
    int *mark = imalloc(nn);
    for (int v = 0; v < nn; v++) mark[v] = -1;
    for (int v = 0; v < nn; v++) {
      for (int q = nptr[v]; q < nptr[v + 1]; q++) {
        int e = nel[q];
        for (int k = kinet[e]; k < kinet[e + 1]; k++) {
          int w = s->inet[k];
          if (w == v || mark[w] == v) continue;
          mark[w] = v;
          if (cnt == cap) {
            cap *= 2;
            adj = (int *)realloc(adj, sizeof(int) * cap);
          }
          adj[cnt++] = w;
        }
      }
      xadj[v + 1] = cnt;
    }
    graph_components(nn, xadj, adj, s->ncomp);
    free(xadj);
    free(adj);
    free(mark);
  }
  free(kinet);
  free(nptr);
  free(nel);
  free(fill);
}

/* ------------------------------------------------- neighbours, interface */
static void build_node_map(oprob *P) {
  int ng = P->nnod_g;
  P->n2s_ptr = imalloc(ng + 1);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->nnods; k++) P->n2s_ptr[s->isngn[k] + 1]++;
  }
  for (int k = 0; k < ng; k++) P->n2s_ptr[k + 1] += P->n2s_ptr[k];
  P->n2s_idx = imalloc(P->n2s_ptr[ng]);
  int *fill = imalloc(ng);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->nnods; k++) {
      int gn = s->isngn[k];
      P->n2s_idx[P->n2s_ptr[gn] + fill[gn]++] = is; /* ascending is */
    }
  }
  free(fill);
}

/* local node index of a global node in subdomain s (isngn may be unsorted) */
typedef struct {
  int g, l;
} gl;
static int cmp_gl(const void *a, const void *b) {
  return cmp_int(&((const gl *)a)->g, &((const gl *)b)->g);
}

/* interface and shared-node lists (module_dd.f90:8071-8647) */
static void sub_interface(oprob *P, int is) {
  osub *s = &P->s[is];
  int nn = s->nnods;
  /* neighbours: subdomains sharing at least one node, ascending */
  int *mark = imalloc(P->nsub);
  int nadj = 0;
  for (int k = 0; k < nn; k++) {
    int gn = s->isngn[k];
    for (int q = P->n2s_ptr[gn]; q < P->n2s_ptr[gn + 1]; q++) {
      int t = P->n2s_idx[q];
      if (t != is && !mark[t]) mark[t] = 1, nadj++;
    }
  }
  s->nadj = nadj;
  s->iadj = imalloc(nadj);
  nadj = 0;
  for (int t = 0; t < P->nsub; t++)
    if (mark[t]) s->iadj[nadj++] = t;
  free(mark);
  /* interface nodes: shared with some neighbour, in local node order */
  int *isint = imalloc(nn);
  for (int k = 0; k < nn; k++) {
    int gn = s->isngn[k];
    if (P->n2s_ptr[gn + 1] - P->n2s_ptr[gn] > 1) isint[k] = 1;
  }
  s->nnodi = 0;
  s->ndofi = 0;
  s->ndofo = 0;
  for (int k = 0; k < nn; k++) {
    if (isint[k]) {
      s->nnodi++;
      s->ndofi += s->nndf[k];
    } else
      s->ndofo += s->nndf[k];
  }
  s->iin = imalloc(s->nnodi);
  s->iivsvn = imalloc(s->ndofi);
  s->iovsvn = imalloc(s->ndofo);
  s->kdofi = imalloc(s->nnodi + 1);
  int ni = 0, di = 0, dof = 0;
  for (int k = 0; k < nn; k++) {
    if (isint[k]) {
      s->iin[ni] = k;
      s->kdofi[ni] = di;
      ni++;
      for (int d = 0; d < s->nndf[k]; d++) s->iivsvn[di++] = s->kdof[k] + d;
    } else
      for (int d = 0; d < s->nndf[k]; d++) s->iovsvn[dof++] = s->kdof[k] + d;
  }
  s->kdofi[ni] = di;
  free(isint);
  s->gdofi = imalloc(s->ndofi);
  for (int k = 0; k < s->ndofi; k++) s->gdofi[k] = s->isvgvn[s->iivsvn[k]];
  /* shared nodes with each neighbour, sorted by global node id */
  gl *mine = (gl *)xcalloc(nn, sizeof(gl));
  for (int k = 0; k < nn; k++) mine[k].g = s->isngn[k], mine[k].l = k;
  qsort(mine, nn, sizeof(gl), cmp_gl);
  /* local node -> interface index */
  int *l2i = imalloc(nn);
  for (int k = 0; k < nn; k++) l2i[k] = -1;
  for (int k = 0; k < s->nnodi; k++) l2i[s->iin[k]] = k;
  s->kshn = imalloc(s->nadj + 1);
  int cap = 0;
  for (int a = 0; a < s->nadj; a++) cap += s->nnodi;
  s->ishn = imalloc(cap);
  s->ishncadj = imalloc(cap);
  int cnt = 0;
  for (int a = 0; a < s->nadj; a++) {
    int t = s->iadj[a];
    osub *u = &P->s[t];
    for (int k = 0; k < nn; k++) {
      int gn = mine[k].g;
      int shared = 0;
      for (int q = P->n2s_ptr[gn]; q < P->n2s_ptr[gn + 1]; q++)
        if (P->n2s_idx[q] == t) shared = 1;
      if (!shared) continue;
      s->ishn[cnt] = l2i[mine[k].l];
      /* component of the node in the neighbour */
      int lu = -1;
      for (int q = 0; q < u->nnods; q++)
        if (u->isngn[q] == gn) {
          lu = q;
          break;
        }
      s->ishncadj[cnt] = u->ncomp[lu];
      cnt++;
    }
    s->kshn[a + 1] = cnt;
  }
  free(mine);
  free(l2i);
}

/* -------------------------------------------------------- globs, corners */
static int same_glob(const osub *s, int i, int j) {
  if (s->nsubnode[i] != s->nsubnode[j]) return 0;
  int w = s->nsubnx * 3;
  for (int m = 0; m < s->nsubnode[i] * 3; m++)
    if (s->gsubs[i * w + m] != s->gsubs[j * w + m]) return 0;
  return 1;
}

/* first pass of glob identification and local corner selection
 * (module_dd.f90:9130-9300, selection 10017-10327) */
static void sub_globs_pass1(oprob *P, int is, int use_corner) {
  osub *s = &P->s[is];
  int nnodi = s->nnodi;
  s->nsubnode = imalloc(nnodi);
  for (int a = 0; a < s->nadj; a++)
    for (int q = s->kshn[a]; q < s->kshn[a + 1]; q++) s->nsubnode[s->ishn[q]]++;
  int nsubnx = 0;
  for (int k = 0; k < nnodi; k++) nsubnx = ONE_MAX(nsubnx, s->nsubnode[k]);
  s->nsubnx = nsubnx;
  int w = nsubnx * 3;
  s->gsubs = imalloc((size_t)nnodi * (w ? w : 1));
  int *cnt = imalloc(nnodi);
  for (int a = 0; a < s->nadj; a++)
    for (int q = s->kshn[a]; q < s->kshn[a + 1]; q++) {
      int in = s->ishn[q];
      int m = cnt[in]++;
      s->gsubs[in * w + 3 * m + 0] = s->iadj[a];
      s->gsubs[in * w + 3 * m + 1] = s->ishncadj[q];
      s->gsubs[in * w + 3 * m + 2] = s->ncomp[s->iin[in]];
    }
  free(cnt);
  s->globtypes = imalloc(nnodi);
  s->kglobs = imalloc(nnodi);
  int *gt = s->globtypes, *kg = s->kglobs;
  for (int k = 0; k < nnodi; k++) kg[k] = -1, gt[k] = 0;
  int iglobs = 0;
  for (int i = 0; i < nnodi; i++) {
    if (kg[i] != -1) continue;
    kg[i] = ++iglobs;
    int nsubn = s->nsubnode[i], nglobn = 1;
    for (int j = i + 1; j < nnodi; j++) {
      if (kg[j] == -1 && gt[j] != 3 && same_glob(s, i, j)) {
        kg[j] = iglobs;
        nglobn++;
        gt[j] = (nsubn == 1) ? 1 : 2;
      }
    }
    if (nsubn == 1)
      gt[i] = 1;
    else if (nglobn > 1)
      gt[i] = 2;
    else
      gt[i] = 3;
  }
  for (int k = 0; k < nnodi; k++)
    if (s->nsubnode[k] == 1) gt[k] = 1;

  if (use_corner) {
    int ndim = P->ndim, meshdim = P->meshdim;
    for (int a = 0; a < s->nadj; a++) {
      int q0 = s->kshn[a], nshn = s->kshn[a + 1] - q0;
      int face = 0;
      for (int q = 0; q < nshn; q++)
        if (gt[s->ishn[q0 + q]] == 1) {
          face = 1;
          break;
        }
      if (!face) continue;
      /* interface components: combination of my and adjacent components */
      int *myc = imalloc(nshn), *adc = imalloc(nshn), *ic = imalloc(nshn);
      for (int q = 0; q < nshn; q++) {
        myc[q] = s->ncomp[s->iin[s->ishn[q0 + q]]];
        adc[q] = s->ishncadj[q0 + q];
      }
      /* renumber both from 1 densely in order of old value */
      int arrs = 2;
      int *arr[2] = {myc, adc};
      int nnew[2] = {0, 0};
      for (int r = 0; r < arrs; r++) {
        int mx = 0;
        for (int q = 0; q < nshn; q++) mx = ONE_MAX(mx, arr[r][q]);
        int *newid = imalloc(mx + 1);
        for (int q = 0; q < nshn; q++) newid[arr[r][q]] = 1;
        int c = 0;
        for (int v = 1; v <= mx; v++)
          if (newid[v]) newid[v] = ++c;
        for (int q = 0; q < nshn; q++) arr[r][q] = newid[arr[r][q]];
        nnew[r] = c;
        free(newid);
      }
      int mx = 0;
      for (int q = 0; q < nshn; q++) {
        ic[q] = (myc[q] - 1) * nnew[1] + adc[q];
        mx = ONE_MAX(mx, ic[q]);
      }
      int *newid = imalloc(mx + 1);
      for (int q = 0; q < nshn; q++) newid[ic[q]] = 1;
      int nic = 0;
      for (int v = 1; v <= mx; v++)
        if (newid[v]) newid[v] = ++nic;
      for (int q = 0; q < nshn; q++) ic[q] = newid[ic[q]];
      free(newid);
      for (int c = 1; c <= nic; c++) {
        int nc = 0;
        for (int q = 0; q < nshn; q++)
          if (ic[q] == c) nc++;
        int *compind = imalloc(nc);
        nc = 0;
        for (int q = 0; q < nshn; q++)
          if (ic[q] == c) compind[nc++] = q;
        double *xs = dmalloc((size_t)nc * 3);
        for (int m = 0; m < nc; m++) {
          int ln = s->iin[s->ishn[q0 + compind[m]]];
          for (int d = 0; d < ndim; d++) xs[m * 3 + d] = s->xyz[(size_t)d * s->nnods + ln];
        }
        int f[4] = {0, 0, 0, 0}; /* 1-based like the reference, 0 = unset */
        if (nc > 0) {
          double best = -1.0;
          for (int m = 0; m < nc; m++) {
            double dd = 0;
            for (int d = 0; d < ndim; d++) dd += (xs[m * 3 + d] - xs[d]) * (xs[m * 3 + d] - xs[d]);
            if (dd > best) best = dd, f[0] = m + 1; /* maxloc: first max */
          }
        }
        if (meshdim > 1 && nc > 1) {
          double best = -1.0;
          int b0 = f[0] - 1;
          for (int m = 0; m < nc; m++) {
            double dd = 0;
            for (int d = 0; d < ndim; d++)
              dd += (xs[m * 3 + d] - xs[b0 * 3 + d]) * (xs[m * 3 + d] - xs[b0 * 3 + d]);
            if (dd > best) best = dd, f[1] = m + 1;
          }
          if (f[1] == f[0]) {
            for (int m = 1; m <= nc; m++)
              if (m != f[0]) {
                f[1] = m;
                break;
              }
          }
        }
        if (meshdim > 2 && nc > 2) {
          double x1 = xs[(f[0] - 1) * 3], y1 = xs[(f[0] - 1) * 3 + 1], z1 = xs[(f[0] - 1) * 3 + 2];
          double x2 = xs[(f[1] - 1) * 3], y2 = xs[(f[1] - 1) * 3 + 1], z2 = xs[(f[1] - 1) * 3 + 2];
          double *area = dmalloc(nc);
          for (int m = 0; m < nc; m++) {
            double xi = xs[m * 3], yi = xs[m * 3 + 1], zi = xs[m * 3 + 2];
            double c1 = (y2 - y1) * (zi - z1) - (yi - y1) * (z2 - z1);
            double c2 = (x2 - x1) * (zi - z1) - (xi - x1) * (z2 - z1);
            double c3 = (x2 - x1) * (yi - y1) - (xi - x1) * (y2 - y1);
            area[m] = c1 * c1 + c2 * c2 + c3 * c3;
          }
          for (int round = 0; round < 2; round++) {
            double maxv = 0.0;
            int maxl = 0;
            if (round == 0) {
              for (int m = 0; m < nc; m++)
                if (area[m] > maxv) maxv = area[m], maxl = m + 1;
            } else {
              for (int m = nc - 1; m >= 0; m--)
                if (area[m] > maxv) maxv = area[m], maxl = m + 1;
            }
            f[2 + round] = maxl;
          }
          if ((f[2] == f[0] && f[3] == f[0]) || (f[2] == f[1] && f[3] == f[1])) {
            for (int m = 1; m <= nc; m++)
              if (m != f[0] && m != f[1]) {
                f[2] = m;
                break;
              }
            for (int m = 1; m <= nc; m++)
              if (m != f[0] && m != f[1] && m != f[2]) {
                f[3] = m;
                break;
              }
          }
          free(area);
        }
        int end_bound = (meshdim <= 2) ? 2 : 4;
        for (int k = 0; k < end_bound; k++)
          if (f[k] != 0) gt[s->ishn[q0 + compind[f[k] - 1]]] = 3;
        free(xs);
        free(compind);
      }
      free(myc);
      free(adc);
      free(ic);
    }
  }
  s->nlcorner = 0;
  for (int k = 0; k < nnodi; k++)
    if (gt[k] == 3) s->nlcorner++;
  s->lcorner = imalloc(s->nlcorner);
  int c = 0;
  for (int k = 0; k < nnodi; k++)
    if (gt[k] == 3) s->lcorner[c++] = s->isngn[s->iin[k]];
}

/* merge corners with neighbours, edges, faces, BC removal
 * (module_dd.f90:9322-9660) */
static void sub_globs_pass2(oprob *P, int is, int remove_bc) {
  osub *s = &P->s[is];
  int nnodi = s->nnodi;
  /* union of own and neighbours' local corners */
  int tot = s->nlcorner;
  for (int a = 0; a < s->nadj; a++) tot += P->s[s->iadj[a]].nlcorner;
  int *merge = imalloc(tot);
  int m = 0;
  for (int k = 0; k < s->nlcorner; k++) merge[m++] = s->lcorner[k];
  for (int a = 0; a < s->nadj; a++) {
    osub *u = &P->s[s->iadj[a]];
    for (int k = 0; k < u->nlcorner; k++) merge[m++] = u->lcorner[k];
  }
  qsort(merge, tot, sizeof(int), cmp_int);
  int *ign = imalloc(nnodi), *ord = imalloc(nnodi);
  for (int k = 0; k < nnodi; k++) ign[k] = s->isngn[s->iin[k]];
  /* intersection (sorted) */
  int *sorted_ign = imalloc(nnodi);
  memcpy(sorted_ign, ign, sizeof(int) * nnodi);
  qsort(sorted_ign, nnodi, sizeof(int), cmp_int);
  s->icgn = imalloc(nnodi);
  s->ncorner = 0;
  for (int k = 0; k < nnodi; k++) {
    int g = sorted_ign[k];
    int p = lower_bound(merge, tot, g);
    if (p < tot && merge[p] == g) s->icgn[s->ncorner++] = g;
  }
  free(merge);
  free(sorted_ign);
  free(ord);
  int *gt = s->globtypes, *kg = s->kglobs;
  for (int k = 0; k < nnodi; k++) kg[k] = -1, gt[k] = 0;
  for (int c = 0; c < s->ncorner; c++) {
    int g = s->icgn[c];
    for (int k = 0; k < nnodi; k++)
      if (ign[k] == g) {
        kg[k] = c + 1;
        gt[k] = 3;
        break;
      }
  }
  int nc = s->ncorner;
  int iedges = 0;
  for (int i = 0; i < nnodi; i++) {
    if (kg[i] != -1 || s->nsubnode[i] <= 1) continue;
    iedges++;
    kg[i] = nc + iedges;
    gt[i] = 2;
    for (int j = i + 1; j < nnodi; j++)
      if (kg[j] == -1 && same_glob(s, i, j)) kg[j] = nc + iedges, gt[j] = 2;
  }
  int nedges = iedges, ifaces = 0;
  for (int i = 0; i < nnodi; i++) {
    if (kg[i] != -1) continue;
    if (s->nsubnode[i] != 1) {
      fprintf(stderr, "ERROR in ORACLE: unassigned non-face node\n");
      abort();
    }
    ifaces++;
    kg[i] = nc + nedges + ifaces;
    gt[i] = 1;
    for (int j = i + 1; j < nnodi; j++)
      if (kg[j] == -1 && same_glob(s, i, j)) kg[j] = nc + nedges + ifaces, gt[j] = 1;
  }
  int nfaces = ifaces;
  if (remove_bc && s->has_fix) {
    for (int i = 0; i < nnodi; i++) {
      if (gt[i] == 3) continue;
      int ln = s->iin[i], fixed = 0;
      for (int d = s->kdof[ln]; d < s->kdof[ln + 1]; d++)
        if (s->ifix[d]) fixed = 1;
      if (fixed) kg[i] = -2, gt[i] = -2;
    }
    int ne2 = 0;
    for (int e = 1; e <= nedges; e++) {
      int old = nc + e, any = 0;
      for (int k = 0; k < nnodi; k++)
        if (kg[k] == old) any = 1;
      if (any) {
        ne2++;
        for (int k = 0; k < nnodi; k++)
          if (kg[k] == old) kg[k] = nc + ne2;
      }
    }
    int nf2 = 0;
    for (int f = 1; f <= nfaces; f++) {
      int old = nc + nedges + f, any = 0;
      for (int k = 0; k < nnodi; k++)
        if (kg[k] == old) any = 1;
      if (any) {
        nf2++;
        for (int k = 0; k < nnodi; k++)
          if (kg[k] == old) kg[k] = nc + ne2 + nf2;
      }
    }
    nedges = ne2;
    nfaces = nf2;
  }
  s->nedges = nedges;
  s->nfaces = nfaces;
  /* glob lists: edges then faces, nodes in interface order */
  s->ng = nedges + nfaces;
  s->gtype = imalloc(s->ng);
  s->gptr = imalloc(s->ng + 1);
  s->gkey = imalloc(s->ng);
  for (int k = 0; k < nnodi; k++)
    if (kg[k] > nc) s->gptr[kg[k] - nc]++;
  for (int q = 0; q < s->ng; q++) s->gptr[q + 1] += s->gptr[q];
  s->gnod = imalloc(s->gptr[s->ng]);
  int *fill = imalloc(s->ng);
  for (int q = 0; q < s->ng; q++) s->gkey[q] = -1;
  for (int k = 0; k < nnodi; k++)
    if (kg[k] > nc) {
      int q = kg[k] - nc - 1;
      s->gnod[s->gptr[q] + fill[q]++] = k;
      s->gtype[q] = gt[k];
      if (s->gkey[q] < 0 || ign[k] < s->gkey[q]) s->gkey[q] = ign[k];
    }
  free(fill);
  free(ign);
}

/* ------------------------------------------------------------- BC, weights */
/* sm_apply_bc (module_sm.f90:431-600) on the assembled upper triangle */
static void sub_apply_bc(osub *s) {
  s->bc = dmalloc(s->ndofs);
  if (!s->has_fix) return;
  int *fixed_diag = imalloc(s->ndofs);
  for (int k = 0; k < s->nnz; k++) {
    int i = s->mi[k], j = s->mj[k];
    double a = s->ma[k];
    if (i != j) {
      if (s->ifix[i] && s->fixv[i] != 0.0 && a != 0.0 && !s->ifix[j]) s->bc[j] -= a * s->fixv[i];
      if (s->ifix[j] && s->fixv[j] != 0.0 && a != 0.0 && !s->ifix[i]) s->bc[i] -= a * s->fixv[j];
      if (s->ifix[i] || s->ifix[j]) s->ma[k] = 0.0;
    } else if (s->ifix[i]) {
      if (fabs(a) < 1e-12) s->ma[k] = a = 1.0;
      if (s->fixv[i] != 0.0 && a != 0.0) s->bc[i] += a * s->fixv[i];
      fixed_diag[i] = 1;
    }
  }
  for (int i = 0; i < s->ndofs; i++)
    if (s->ifix[i] && !fixed_diag[i]) {
      s->mi[s->nnz] = i;
      s->mj[s->nnz] = i;
      s->ma[s->nnz] = 1.0;
      s->bc[i] += s->fixv[i];
      s->nnz++;
    }
  free(fixed_diag);
}

/* sum the values of interface copies over all subdomains sharing each dof
 * (dd_comm_upload/swapdata/download, module_dd.f90:11236-11552) */
static void exchange_sum(oprob *P, double **v) {
  memset(P->gbuf, 0, sizeof(double) * P->ndof_g);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->ndofi; k++) P->gbuf[s->gdofi[k]] += v[is][k];
  }
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->ndofi; k++) v[is][k] = P->gbuf[s->gdofi[k]];
  }
}

static double **alloc_rep(oprob *P) {
  double **v = (double **)xcalloc(P->nsub, sizeof(double *));
  for (int is = 0; is < P->nsub; is++) v[is] = dmalloc(P->s[is].ndofi);
  return v;
}
static void free_rep(oprob *P, double **v) {
  for (int is = 0; is < P->nsub; is++) free(v[is]);
  free(v);
}

/* weights (module_dd.f90:11554-11725) */
static int setup_weights(oprob *P, int weights_type) {
  double **rho = alloc_rep(P), **tot = alloc_rep(P);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    if (weights_type == 0) {
      for (int k = 0; k < s->ndofi; k++) rho[is][k] = 1.0;
    } else if (weights_type == 1) {
      double *dg = dmalloc(s->ndofs);
      for (int k = 0; k < s->nnz; k++)
        if (s->mi[k] == s->mj[k]) dg[s->mi[k]] += s->ma[k];
      for (int k = 0; k < s->ndofi; k++) {
        rho[is][k] = dg[s->iivsvn[k]];
        if (rho[is][k] == 0.0) rho[is][k] = 1.0;
      }
      free(dg);
    } else if (weights_type == 3) {
      if (s->ldof_data != s->ndofs) return -1;
      for (int k = 0; k < s->ndofi; k++) rho[is][k] = s->dof_data[s->iivsvn[k]];
    } else {
      fprintf(stderr, "ERROR in ORACLE: weights type %d not supported\n", weights_type);
      return -1;
    }
    memcpy(tot[is], rho[is], sizeof(double) * s->ndofi);
  }
  exchange_sum(P, tot);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    s->wi = dmalloc(s->ndofi);
    for (int k = 0; k < s->ndofi; k++) {
      if (tot[is][k] == 0.0) {
        fprintf(stderr, "ERROR in ORACLE: zero denominator in weights\n");
        return -1;
      }
      s->wi[k] = rho[is][k] / tot[is][k];
    }
  }
  free_rep(P, rho);
  free_rep(P, tot);
  return 0;
}

/* ---------------------------------------------------- coarse numbering, C */
typedef struct {
  int type, key, nd;
} cn;
static int cmp_cn(const void *a, const void *b) {
  const cn *x = (const cn *)a, *y = (const cn *)b;
  int ox = (x->type == 3) ? 0 : (x->type == 2 ? 1 : 2);
  int oy = (y->type == 3) ? 0 : (y->type == 2 ? 1 : 2);
  if (ox != oy) return ox - oy;
  return (x->key > y->key) - (x->key < y->key);
}

/* coarse nodes: corners (sorted by node id), edges, faces (sorted by their
 * smallest node id); constraints: unit rows at corners, unnormalised sums per
 * component on globs (module_dd.f90:4298-4384, 4869-5016, 5018-5145) */
static int setup_constraints(oprob *P, int use_arith) {
  int cap = 0;
  for (int is = 0; is < P->nsub; is++) cap += P->s[is].ncorner + P->s[is].ng;
  cn *all = (cn *)xcalloc(cap, sizeof(cn));
  int n = 0;
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int c = 0; c < s->ncorner; c++) {
      int g = s->icgn[c];
      /* number of dofs at the corner node */
      int ln = -1;
      for (int k = 0; k < s->nnodi; k++)
        if (s->isngn[s->iin[k]] == g) ln = s->iin[k];
      all[n].type = 3, all[n].key = g, all[n].nd = s->nndf[ln], n++;
    }
    for (int q = 0; q < s->ng; q++) {
      int nd = 0;
      for (int m = s->gptr[q]; m < s->gptr[q + 1]; m++)
        nd = ONE_MAX(nd, s->nndf[s->iin[s->gnod[m]]]);
      all[n].type = s->gtype[q], all[n].key = s->gkey[q], all[n].nd = use_arith ? nd : 0, n++;
    }
  }
  qsort(all, n, sizeof(cn), cmp_cn);
  int u = 0;
  for (int k = 0; k < n; k++) {
    if (u > 0 && all[u - 1].type == all[k].type && all[u - 1].key == all[k].key) {
      if (all[u - 1].nd != all[k].nd) {
        fprintf(stderr, "ERROR in ORACLE: inconsistent coarse node dofs\n");
        return -1;
      }
      continue;
    }
    all[u++] = all[k];
  }
  P->nnodc = u;
  P->ncorner = P->nedge = P->nface = 0;
  for (int k = 0; k < u; k++) {
    if (all[k].type == 3) P->ncorner++;
    else if (all[k].type == 2) P->nedge++;
    else P->nface++;
  }
  P->nndfc = imalloc(u);
  P->kdofc = imalloc(u + 1);
  for (int k = 0; k < u; k++) {
    P->nndfc[k] = all[k].nd;
    P->kdofc[k + 1] = P->kdofc[k] + all[k].nd;
  }
  P->ndofc = P->kdofc[u];
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int nconstr = 0;
    /* cnodes in reference order: corners (icnsin order = sorted ids), globs */
    int ncn = s->ncorner + s->ng;
    int *cid = imalloc(ncn);
    for (int c = 0; c < ncn; c++) {
      cn key;
      if (c < s->ncorner)
        key.type = 3, key.key = s->icgn[c];
      else
        key.type = s->gtype[c - s->ncorner], key.key = s->gkey[c - s->ncorner];
      int lo = 0, hi = u;
      while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (cmp_cn(&all[mid], &key) < 0) lo = mid + 1;
        else hi = mid;
      }
      cid[c] = lo;
      nconstr += P->nndfc[lo];
    }
    s->nconstr = nconstr;
    s->C = dmalloc((size_t)nconstr * s->ndofi);
    s->indrowc = imalloc(nconstr);
    int r = 0;
    for (int c = 0; c < ncn; c++) {
      int id = cid[c], nd = P->nndfc[id];
      if (c < s->ncorner) {
        int g = s->icgn[c], in = -1;
        for (int k = 0; k < s->nnodi; k++)
          if (s->isngn[s->iin[k]] == g) in = k;
        for (int d = 0; d < nd; d++) {
          s->C[(size_t)(r + d) * s->ndofi + s->kdofi[in] + d] = 1.0;
          s->indrowc[r + d] = P->kdofc[id] + d;
        }
      } else {
        int q = c - s->ncorner;
        for (int d = 0; d < nd; d++) {
          for (int mm = s->gptr[q]; mm < s->gptr[q + 1]; mm++) {
            int in = s->gnod[mm];
            if (d < s->kdofi[in + 1] - s->kdofi[in])
              s->C[(size_t)(r + d) * s->ndofi + s->kdofi[in] + d] = 1.0;
          }
          s->indrowc[r + d] = P->kdofc[id] + d;
        }
      }
      r += nd;
    }
    free(s->cnid);
    s->ncn = ncn;
    s->cnid = cid;
  }
  free(all);
  return 0;
}

/* --------------------------------------------------------- local algebra */
/* skyline Cholesky of A11 in the reference's interior order (iovsvn) */
static int sub_factor_a11(osub *s, int *o2l) {
  int no = s->ndofo;
  s->sky_first = imalloc(no + 1);
  s->sky_off = imalloc(no + 1);
  for (int i = 0; i < no; i++) s->sky_first[i] = i;
  for (int k = 0; k < s->nnz; k++) {
    int p = o2l[s->mi[k]], q = o2l[s->mj[k]];
    if (p < 0 || q < 0) continue;
    int r = ONE_MAX(p, q), c = ONE_MIN(p, q);
    if (c < s->sky_first[r]) s->sky_first[r] = c;
  }
  for (int i = 0; i < no; i++) s->sky_off[i + 1] = s->sky_off[i] + (i - s->sky_first[i] + 1);
  s->sky = dmalloc(s->sky_off[no]);
  for (int k = 0; k < s->nnz; k++) {
    int p = o2l[s->mi[k]], q = o2l[s->mj[k]];
    if (p < 0 || q < 0) continue;
    int r = ONE_MAX(p, q), c = ONE_MIN(p, q);
    s->sky[s->sky_off[r] + c - s->sky_first[r]] += s->ma[k];
  }
  for (int i = 0; i < no; i++) {
    double *Li = s->sky + s->sky_off[i] - s->sky_first[i];
    int fi = s->sky_first[i];
    for (int j = fi; j < i; j++) {
      double *Lj = s->sky + s->sky_off[j] - s->sky_first[j];
      int k0 = ONE_MAX(fi, s->sky_first[j]);
      double sum = Li[j];
      for (int k = k0; k < j; k++) sum -= Li[k] * Lj[k];
      Li[j] = sum / Lj[j];
    }
    double d = Li[i];
    for (int k = fi; k < i; k++) d -= Li[k] * Li[k];
    if (!(d > 0.0)) {
      fprintf(stderr, "ERROR in ORACLE: A11 not positive definite\n");
      return -1;
    }
    Li[i] = sqrt(d);
  }
  return 0;
}

static void sky_solve(const osub *s, double *x) {
  int no = s->ndofo;
  for (int i = 0; i < no; i++) {
    const double *Li = s->sky + s->sky_off[i] - s->sky_first[i];
    double sum = x[i];
    for (int k = s->sky_first[i]; k < i; k++) sum -= Li[k] * x[k];
    x[i] = sum / Li[i];
  }
  for (int i = no - 1; i >= 0; i--) {
    const double *Li = s->sky + s->sky_off[i] - s->sky_first[i];
    x[i] /= Li[i];
    double xi = x[i];
    for (int k = s->sky_first[i]; k < i; k++) x[k] -= Li[k] * xi;
  }
}

/* dense LU with partial pivoting, row major n x n */
static int lu_factor(double *A, int *piv, int n) {
  for (int k = 0; k < n; k++) {
    int p = k;
    double mx = fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(A[(size_t)i * n + k]) > mx) mx = fabs(A[(size_t)i * n + k]), p = i;
    piv[k] = p;
    if (mx == 0.0) return -1;
    if (p != k)
      for (int j = 0; j < n; j++) {
        double t = A[(size_t)k * n + j];
        A[(size_t)k * n + j] = A[(size_t)p * n + j];
        A[(size_t)p * n + j] = t;
      }
    double inv = 1.0 / A[(size_t)k * n + k];
    for (int i = k + 1; i < n; i++) {
      double l = A[(size_t)i * n + k] * inv;
      A[(size_t)i * n + k] = l;
      if (l != 0.0) {
        double *ri = A + (size_t)i * n, *rk = A + (size_t)k * n;
        for (int j = k + 1; j < n; j++) ri[j] -= l * rk[j];
      }
    }
  }
  return 0;
}
static void lu_solve(const double *A, const int *piv, int n, double *b) {
  for (int k = 0; k < n; k++) {
    int p = piv[k];
    if (p != k) {
      double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
  }
  for (int i = 0; i < n; i++) {
    double sum = b[i];
    const double *ri = A + (size_t)i * n;
    for (int j = 0; j < i; j++) sum -= ri[j] * b[j];
    b[i] = sum;
  }
  for (int i = n - 1; i >= 0; i--) {
    double sum = b[i];
    const double *ri = A + (size_t)i * n;
    for (int j = i + 1; j < n; j++) sum -= ri[j] * b[j];
    b[i] = sum / ri[i];
  }
}

/* y = A12 x  (x interface, y interior) and transpose, via the upper triplets */
static void sub_a12(const osub *s, const int *o2l, const int *i2l, const double *x,
                    double *y) {
  memset(y, 0, sizeof(double) * s->ndofo);
  for (int k = 0; k < s->nnz; k++) {
    int i = s->mi[k], j = s->mj[k];
    if (o2l[i] >= 0 && i2l[j] >= 0) y[o2l[i]] += s->ma[k] * x[i2l[j]];
    else if (i2l[i] >= 0 && o2l[j] >= 0) y[o2l[j]] += s->ma[k] * x[i2l[i]];
  }
}
static void sub_a21(const osub *s, const int *o2l, const int *i2l, const double *x,
                    double *y) {
  memset(y, 0, sizeof(double) * s->ndofi);
  for (int k = 0; k < s->nnz; k++) {
    int i = s->mi[k], j = s->mj[k];
    if (o2l[i] >= 0 && i2l[j] >= 0) y[i2l[j]] += s->ma[k] * x[o2l[i]];
    else if (i2l[i] >= 0 && o2l[j] >= 0) y[i2l[i]] += s->ma[k] * x[o2l[j]];
  }
}

static void sub_maps(const osub *s, int *o2l, int *i2l) {
  for (int k = 0; k < s->ndofs; k++) o2l[k] = i2l[k] = -1;
  for (int k = 0; k < s->ndofo; k++) o2l[s->iovsvn[k]] = k;
  for (int k = 0; k < s->ndofi; k++) i2l[s->iivsvn[k]] = k;
}

/* explicit Schur complement, aug LU, coarse basis (module_dd.f90:4175-4296,
 * 5211-5487, 5489-5800) */
static int sub_coarse_basis(osub *s);
static int sub_numerics(osub *s) {
  int *o2l = imalloc(s->ndofs), *i2l = imalloc(s->ndofs);
  sub_maps(s, o2l, i2l);
  int ni = s->ndofi, no = s->ndofo;
  if (no > 0 && sub_factor_a11(s, o2l)) return -1;
  s->S = dmalloc((size_t)ni * ni);
  /* A22 */
  for (int k = 0; k < s->nnz; k++) {
    int p = i2l[s->mi[k]], q = i2l[s->mj[k]];
    if (p < 0 || q < 0) continue;
    s->S[(size_t)p * ni + q] += s->ma[k];
    if (p != q) s->S[(size_t)q * ni + p] += s->ma[k];
  }
  if (no > 0) {
    double *e = dmalloc(ni), *x = dmalloc(no), *y = dmalloc(ni);
    for (int j = 0; j < ni; j++) {
      e[j] = 1.0;
      sub_a12(s, o2l, i2l, e, x);
      e[j] = 0.0;
      sky_solve(s, x);
      sub_a21(s, o2l, i2l, x, y);
      for (int i = 0; i < ni; i++) s->S[(size_t)i * ni + j] -= y[i];
    }
    free(e);
    free(x);
    free(y);
  }
  free(o2l);
  free(i2l);
  return sub_coarse_basis(s);
}

/* Phi_i and Ac_i from [S C^T; C 0] (dd_prepare_aug, dd_prepare_coarse:
 * module_dd.f90:5211-5800) for the current constraint rows C */
static int sub_coarse_basis(osub *s) {
  int ni = s->ndofi;
  free(s->aug);
  free(s->augpiv);
  free(s->phi);
  free(s->acs);
  int nc = s->nconstr, na = ni + nc;
  s->aug = dmalloc((size_t)na * na);
  s->augpiv = imalloc(na);
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < ni; j++) s->aug[(size_t)i * na + j] = s->S[(size_t)i * ni + j];
  for (int r = 0; r < nc; r++)
    for (int j = 0; j < ni; j++) {
      double c = s->C[(size_t)r * ni + j];
      s->aug[(size_t)(ni + r) * na + j] = c;
      s->aug[(size_t)j * na + ni + r] = c;
    }
  if (lu_factor(s->aug, s->augpiv, na)) {
    fprintf(stderr, "ERROR in ORACLE: singular augmented matrix\n");
    return -1;
  }
  s->phi = dmalloc((size_t)ni * nc);
  s->acs = dmalloc((size_t)nc * nc);
  double *b = dmalloc(na);
  for (int j = 0; j < nc; j++) {
    memset(b, 0, sizeof(double) * na);
    b[ni + j] = 1.0;
    lu_solve(s->aug, s->augpiv, na, b);
    for (int i = 0; i < ni; i++) s->phi[(size_t)i * nc + j] = b[i];
    for (int i = 0; i < nc; i++) s->acs[(size_t)i * nc + j] = -b[ni + i];
  }
  free(b);
  return 0;
}

/* dense Cholesky (lower, row major) */
static int chol_factor(double *A, int n) {
  for (int j = 0; j < n; j++) {
    double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; k++) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(d > 0.0)) return -1;
    d = sqrt(d);
    A[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; i++) {
      double sum = A[(size_t)i * n + j];
      const double *ri = A + (size_t)i * n, *rj = A + (size_t)j * n;
      for (int k = 0; k < j; k++) sum -= ri[k] * rj[k];
      A[(size_t)i * n + j] = sum / d;
    }
  }
  return 0;
}
static void chol_solve(const double *L, int n, double *b) {
  for (int i = 0; i < n; i++) {
    double sum = b[i];
    for (int k = 0; k < i; k++) sum -= L[(size_t)i * n + k] * b[k];
    b[i] = sum / L[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double sum = b[i];
    for (int k = i + 1; k < n; k++) sum -= L[(size_t)k * n + i] * b[k];
    b[i] = sum / L[(size_t)i * n + i];
  }
}

/* reduced right-hand side g = f2 - A21 A11^-1 f1 (module_dd.f90:6995-7258) */
static void sub_rhs_mod(const oprob *P, const osub *s, int weighted, double *f) {
  memcpy(f, s->rhs, sizeof(double) * s->ndofs);
  if (weighted && s->is_rhs_complete)
    for (int k = 0; k < s->ndofi; k++) f[s->iivsvn[k]] *= s->wi[k];
  if (P->bc_global) {
    double *bc = dmalloc(s->ndofs);
    memcpy(bc, s->bc, sizeof(double) * s->ndofs);
    if (weighted)
      for (int k = 0; k < s->ndofi; k++) bc[s->iivsvn[k]] *= s->wi[k];
    if (s->has_fix)
      for (int k = 0; k < s->ndofs; k++)
        if (s->ifix[k]) f[k] = 0.0;
    for (int k = 0; k < s->ndofs; k++) f[k] += bc[k];
    free(bc);
  }
}

static void setup_reduced_rhs(oprob *P) {
  double **g = alloc_rep(P);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int *o2l = imalloc(s->ndofs), *i2l = imalloc(s->ndofs);
    sub_maps(s, o2l, i2l);
    double *f = dmalloc(s->ndofs), *fo = dmalloc(s->ndofo + 1), *t = dmalloc(s->ndofi);
    sub_rhs_mod(P, s, 1, f);
    for (int k = 0; k < s->ndofo; k++) fo[k] = f[s->iovsvn[k]];
    if (s->ndofo > 0) {
      sky_solve(s, fo);
      sub_a21(s, o2l, i2l, fo, t);
    }
    for (int k = 0; k < s->ndofi; k++) g[is][k] = f[s->iivsvn[k]] - t[k];
    free(f);
    free(fo);
    free(t);
    free(o2l);
    free(i2l);
  }
  exchange_sum(P, g);
  for (int is = 0; is < P->nsub; is++) {
    free(P->s[is].g);
    P->s[is].g = g[is];
  }
  free(g);
}

/* coarse problem (module_levels.f90:3001-3169, dd_get_my_coarsem 6384-6464):
 * the Ac_i summed through indrowc and factored */
static int coarse_problem(oprob *P) {
  int nc = P->ndofc;
  free(P->Ac);
  P->Ac = dmalloc((size_t)nc * nc);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int i = 0; i < s->nconstr; i++)
      for (int j = 0; j < s->nconstr; j++)
        P->Ac[(size_t)s->indrowc[i] * nc + s->indrowc[j]] += s->acs[(size_t)i * s->nconstr + j];
  }
  if (nc > 0 && chol_factor(P->Ac, nc)) {
    fprintf(stderr, "ERROR in ORACLE: coarse matrix not positive definite\n");
    return -1;
  }
  return 0;
}

int orc_setup(void *h, int use_corner, int use_arith, int weights_type) {
  oprob *P = (oprob *)h;
  for (int is = 0; is < P->nsub; is++)
    if (!P->s[is].uploaded) {
      fprintf(stderr, "ERROR in ORACLE: subdomain %d not uploaded\n", is);
      return -1;
    }
  P->gbuf = dmalloc(P->ndof_g);
  build_node_map(P);
  for (int is = 0; is < P->nsub; is++) sub_components(&P->s[is]);
  for (int is = 0; is < P->nsub; is++) sub_interface(P, is);
  for (int is = 0; is < P->nsub; is++) sub_globs_pass1(P, is, use_corner);
  for (int is = 0; is < P->nsub; is++) sub_globs_pass2(P, is, 1);
  /* Dirichlet elimination (module_dd.f90:1081-1256) */
  P->bc_global = 0;
  for (int is = 0; is < P->nsub; is++)
    if (P->s[is].has_fix) P->bc_global = 1;
  for (int is = 0; is < P->nsub; is++) sub_apply_bc(&P->s[is]);
  if (P->bc_global) {
    double **b = alloc_rep(P), **own = alloc_rep(P);
    for (int is = 0; is < P->nsub; is++) {
      osub *s = &P->s[is];
      for (int k = 0; k < s->ndofi; k++) b[is][k] = own[is][k] = s->bc[s->iivsvn[k]];
    }
    exchange_sum(P, b);
    for (int is = 0; is < P->nsub; is++) {
      osub *s = &P->s[is];
      for (int k = 0; k < s->ndofi; k++) s->bc[s->iivsvn[k]] = b[is][k];
    }
    free_rep(P, b);
    free_rep(P, own);
  }
  /* re-assemble (entries added for missing fixed diagonals) */
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    trip *t = (trip *)xcalloc(s->nnz, sizeof(trip));
    for (int k = 0; k < s->nnz; k++) t[k].i = s->mi[k], t[k].j = s->mj[k], t[k].a = s->ma[k];
    s->nnz = assemble_triplets(t, s->nnz);
    for (int k = 0; k < s->nnz; k++) s->mi[k] = t[k].i, s->mj[k] = t[k].j, s->ma[k] = t[k].a;
    free(t);
  }
  if (setup_weights(P, weights_type)) return -1;
  if (setup_constraints(P, use_arith)) return -1;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
  for (int is = 0; is < P->nsub; is++) err |= sub_numerics(&P->s[is]) ? 1 : 0;
  if (err) return -1;
  if (coarse_problem(P)) return -1;
  setup_reduced_rhs(P);
  /* unique interface numbering for probes */
  int *isig = imalloc(P->ndof_g);
  for (int is = 0; is < P->nsub; is++)
    for (int k = 0; k < P->s[is].ndofi; k++) isig[P->s[is].gdofi[k]] = 1;
  P->nig = 0;
  for (int k = 0; k < P->ndof_g; k++) P->nig += isig[k];
  P->igdof = imalloc(P->nig);
  P->gpos = imalloc(P->ndof_g);
  int c = 0;
  for (int k = 0; k < P->ndof_g; k++) {
    P->gpos[k] = -1;
    if (isig[k]) P->gpos[k] = c, P->igdof[c++] = k;
  }
  free(isig);
  P->is_setup = 1;
  return 0;
}

/* -------------------------------------------------------------- operators */
/* S apply on replicated copies + exchange (module_levels.f90:4443-4507) */
static void op_schur(oprob *P, double **x, double **y) {
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int n = s->ndofi;
    for (int i = 0; i < n; i++) {
      double sum = 0;
      const double *r = s->S + (size_t)i * n;
      for (int j = 0; j < n; j++) sum += r[j] * x[is][j];
      y[is][i] = sum;
    }
  }
  exchange_sum(P, y);
}

/* BDDC apply (module_levels.f90:3627-3846, 4029-4061, 4252-4346) */
static void op_pc(oprob *P, double **r, double **z) {
  int nc = P->ndofc;
  double *rc = dmalloc(nc + 1);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int ni = s->ndofi, m = s->nconstr, na = ni + m;
    double *aux = dmalloc(na);
    for (int k = 0; k < ni; k++) aux[k] = s->wi[k] * r[is][k];
    for (int j = 0; j < m; j++) {
      double sum = 0;
      for (int k = 0; k < ni; k++) sum += s->phi[(size_t)k * m + j] * aux[k];
      rc[s->indrowc[j]] += sum;
    }
    lu_solve(s->aug, s->augpiv, na, aux);
    memcpy(z[is], aux, sizeof(double) * ni);
    free(aux);
  }
  if (nc > 0 && P->cfn) {
    double *uc = dmalloc(nc);
    P->cfn(P->cctx, rc, uc);
    memcpy(rc, uc, sizeof(double) * nc);
    free(uc);
  } else if (nc > 0)
    chol_solve(P->Ac, nc, rc);
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int ni = s->ndofi, m = s->nconstr;
    for (int k = 0; k < ni; k++) {
      double sum = 0;
      for (int j = 0; j < m; j++) sum += s->phi[(size_t)k * m + j] * rc[s->indrowc[j]];
      z[is][k] = s->wi[k] * (z[is][k] + sum);
    }
  }
  free(rc);
  exchange_sum(P, z);
}

static double wdot(oprob *P, double **a, double **b) {
  double tot = 0;
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    double sub = 0;
    for (int k = 0; k < s->ndofi; k++) sub += s->wi[k] * a[is][k] * b[is][k];
    tot += sub;
  }
  return tot;
}
static void fix_bc_dual(oprob *P, double **v) {
  if (!P->bc_global) return;
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    if (!s->has_fix) continue;
    for (int k = 0; k < s->ndofi; k++)
      if (s->ifix[s->iivsvn[k]]) v[is][k] = 0.0;
  }
}

/* eigenvalues of a symmetric tridiagonal matrix by the implicit QL method
 * (Bowdler, Martin, Reinsch, Wilkinson 1968), d overwritten, e destroyed */
static void tridiag_ql(int n, double *d, double *e) {
  if (n > 0) e[n - 1] = 0.0;
  for (int l = 0; l < n; l++) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; m++) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 1e-15 * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) break;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; i--) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = hypot(f, g));
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

static void record_hist(oprob *P, double v) {
  P->hist = (double *)realloc(P->hist, sizeof(double) * (P->nhist + 1));
  P->hist[P->nhist++] = v;
}

/* interior recovery (module_dd.f90:6811-6932) */
static void postprocess(oprob *P, double **u) {
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int *o2l = imalloc(s->ndofs), *i2l = imalloc(s->ndofs);
    sub_maps(s, o2l, i2l);
    free(s->sols);
    s->sols = dmalloc(s->ndofs);
    if (s->ndofo > 0) {
      double *f = dmalloc(s->ndofs), *x = dmalloc(s->ndofo);
      sub_rhs_mod(P, s, 0, f);
      sub_a12(s, o2l, i2l, u[is], x);
      for (int k = 0; k < s->ndofo; k++) x[k] = f[s->iovsvn[k]] - x[k];
      sky_solve(s, x);
      for (int k = 0; k < s->ndofo; k++) s->sols[s->iovsvn[k]] = x[k];
      free(f);
      free(x);
    }
    for (int k = 0; k < s->ndofi; k++) s->sols[s->iivsvn[k]] = u[is][k];
    free(o2l);
    free(i2l);
  }
  free(P->uig);
  P->uig = dmalloc(P->nig);
  for (int is = P->nsub - 1; is >= 0; is--) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->ndofi; k++) P->uig[P->gpos[s->gdofi[k]]] = u[is][k];
  }
}

static void initial_data(oprob *P, double **u, double **r) {
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->ndofi; k++) {
      int d = s->iivsvn[k];
      u[is][k] = s->ifix[d] ? s->fixv[d] : s->sol0[d];
      r[is][k] = s->g[k];
    }
  }
}

static int pcg(oprob *P, double tol, int maxit, int ndecrmax, int *num_iter, int *reason,
               double *cond) {
  double **u = alloc_rep(P), **r = alloc_rep(P), **p = alloc_rep(P), **ap = alloc_rep(P),
         **z = alloc_rep(P);
  double *diag = dmalloc(maxit + 2), *sub = dmalloc(maxit + 2);
  int ret = 0, nw = 0;
  *num_iter = 0;
  *reason = 0;
  *cond = 1.0;
  initial_data(P, u, r);
  fix_bc_dual(P, r);
  double normrhs = sqrt(wdot(P, r, r));
  if (normrhs == 0.0) goto done;
  op_schur(P, u, ap);
  fix_bc_dual(P, ap);
  for (int is = 0; is < P->nsub; is++)
    for (int k = 0; k < P->s[is].ndofi; k++) r[is][k] -= ap[is][k];
  double normres0 = sqrt(wdot(P, r, r));
  if (normres0 == 0.0) goto done;
  op_pc(P, r, p);
  double rmp = wdot(P, r, p);
  int ndecr = 0;
  double lastres = 1.0;
  for (int iter = 1; iter <= maxit; iter++) {
    op_schur(P, p, ap);
    double pap = wdot(P, p, ap);
    if (pap <= 0) fprintf(stderr, "WARNING in ORACLE: System matrix not positive definite!\n");
    double alpha = rmp / pap;
    for (int is = 0; is < P->nsub; is++)
      for (int k = 0; k < P->s[is].ndofi; k++) {
        u[is][k] += alpha * p[is][k];
        r[is][k] -= alpha * ap[is][k];
      }
    double relres = sqrt(wdot(P, r, r)) / normres0;
    record_hist(P, relres);
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
    if (relres < lastres)
      ndecr = 0;
    else if (++ndecr >= ndecrmax) {
      nw = iter - 1;
      *num_iter = iter;
      *reason = -2;
      break;
    }
    lastres = relres;
    op_pc(P, r, z);
    double rmpold = rmp;
    rmp = wdot(P, r, z);
    if (rmp <= 0) fprintf(stderr, "WARNING in ORACLE: Preconditioner not positive definite!\n");
    double beta = rmp / rmpold;
    for (int is = 0; is < P->nsub; is++)
      for (int k = 0; k < P->s[is].ndofi; k++) p[is][k] = z[is][k] + beta * p[is][k];
    diag[iter - 1] += 1.0 / alpha;
    diag[iter] = beta / alpha;
    sub[iter - 1] = beta >= 0 ? -sqrt(beta) / alpha : 0.0;
  }
  if (nw > 1) {
    tridiag_ql(nw, diag, sub);
    double mx = diag[0];
    for (int k = 1; k < nw; k++) mx = ONE_MAX(mx, diag[k]);
    *cond = mx;
  } else
    *cond = 1.0;
done:
  postprocess(P, u);
  free_rep(P, u);
  free_rep(P, r);
  free_rep(P, p);
  free_rep(P, ap);
  free_rep(P, z);
  free(diag);
  free(sub);
  return ret;
}

static int bicgstab(oprob *P, double tol, int maxit, int ndecrmax, int *num_iter, int *reason) {
  double **u = alloc_rep(P), **r = alloc_rep(P), **rs = alloc_rep(P), **v = alloc_rep(P),
         **p = alloc_rep(P), **y = alloc_rep(P), **z = alloc_rep(P), **s_ = alloc_rep(P),
         **t = alloc_rep(P);
  *num_iter = 0;
  *reason = 0;
  initial_data(P, u, r);
  fix_bc_dual(P, r);
  op_schur(P, u, v);
  for (int is = 0; is < P->nsub; is++)
    for (int k = 0; k < P->s[is].ndofi; k++) r[is][k] -= v[is][k];
  fix_bc_dual(P, r);
  double normrhs = sqrt(wdot(P, r, r));
  if (normrhs == 0.0) goto done;
  double rhoold = 1.0, alpha = 1.0, omega = 1.0;
  for (int is = 0; is < P->nsub; is++) {
    int n = P->s[is].ndofi;
    memset(v[is], 0, sizeof(double) * n);
    memset(p[is], 0, sizeof(double) * n);
    memcpy(rs[is], r[is], sizeof(double) * n);
  }
  int ndecr = 0;
  double lastres = 1.0;
  for (int iter = 1; iter <= maxit; iter++) {
    double rho = wdot(P, r, rs);
    double beta = rho * alpha / (rhoold * omega);
    for (int is = 0; is < P->nsub; is++)
      for (int k = 0; k < P->s[is].ndofi; k++)
        p[is][k] = r[is][k] + beta * (p[is][k] - omega * v[is][k]);
    op_pc(P, p, y);
    op_schur(P, y, v);
    double vrs = wdot(P, v, rs);
    alpha = rho / vrs;
    for (int is = 0; is < P->nsub; is++)
      for (int k = 0; k < P->s[is].ndofi; k++) {
        u[is][k] += alpha * y[is][k];
        s_[is][k] = r[is][k] - alpha * v[is][k];
      }
    double relres = sqrt(wdot(P, s_, s_)) / normrhs;
    record_hist(P, relres);
    if (relres < tol) {
      *num_iter = iter;
      *reason = 0;
      break;
    }
    if (relres < lastres)
      ndecr = 0;
    else if (++ndecr >= ndecrmax) {
      *num_iter = iter;
      *reason = -2;
      break;
    }
    lastres = relres;
    op_pc(P, s_, z);
    op_schur(P, z, t);
    double ts = wdot(P, s_, t), tt = wdot(P, t, t);
    omega = ts / tt;
    for (int is = 0; is < P->nsub; is++)
      for (int k = 0; k < P->s[is].ndofi; k++) {
        u[is][k] += omega * z[is][k];
        r[is][k] = s_[is][k] - omega * t[is][k];
      }
    relres = sqrt(wdot(P, r, r)) / normrhs;
    record_hist(P, relres);
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
    if (relres < lastres)
      ndecr = 0;
    else if (++ndecr >= ndecrmax) {
      *num_iter = iter;
      *reason = -2;
      break;
    }
    lastres = relres;
    rhoold = rho;
  }
done:
  postprocess(P, u);
  free_rep(P, u);
  free_rep(P, r);
  free_rep(P, rs);
  free_rep(P, v);
  free_rep(P, p);
  free_rep(P, y);
  free_rep(P, z);
  free_rep(P, s_);
  free_rep(P, t);
  return 0;
}

int orc_solve(void *h, int method, double tol, int maxit, int ndecrmax, int *num_iter,
              int *converged_reason, double *condition_number) {
  oprob *P = (oprob *)h;
  if (!P->is_setup) return -1;
  P->nhist = 0;
  if (method == -1) method = 1, tol = 1e-6, maxit = 1000, ndecrmax = 30;
  if (method == 0)
    return pcg(P, tol, maxit, ndecrmax, num_iter, converged_reason, condition_number);
  if (method == 1) {
    *condition_number = -1.0;
    return bicgstab(P, tol, maxit, ndecrmax, num_iter, converged_reason);
  }
  fprintf(stderr, "ERROR in ORACLE: unknown iterative method %d\n", method);
  return -1;
}

int orc_download_local_solution(void *h, int isub, double *sols, int lsols) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  if (lsols != s->ndofs || !s->sols) return -1;
  memcpy(sols, s->sols, sizeof(double) * lsols);
  return 0;
}

/* vec1' * W * vec2 over a whole subdomain vector (module_dd.f90:12736-12788) */
int orc_dotprod_subdomain(void *h, int isub, const double *v1, const double *v2, int len,
                          double *out) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  if (len != s->ndofs) return -1;
  /* dd_dotprod_subdomain_local (module_dd.f90:12736-12788): weighted
   * interface dofs in interface order, then the interior dofs */
  double sum = 0;
  for (int k = 0; k < s->ndofi; k++) sum += v1[s->iivsvn[k]] * s->wi[k] * v2[s->iivsvn[k]];
  for (int k = 0; k < s->ndofo; k++) sum += v1[s->iovsvn[k]] * v2[s->iovsvn[k]];
  *out = sum;
  return 0;
}

/* ----------------------------------------------------------------- probes */
void orc_global_counts(void *h, int *out) {
  oprob *P = (oprob *)h;
  out[0] = P->ncorner;
  out[1] = P->nedge;
  out[2] = P->nface;
  out[3] = P->ndofc;
  out[4] = P->nig;
}
void orc_interface_dofs(void *h, int *out) {
  oprob *P = (oprob *)h;
  memcpy(out, P->igdof, sizeof(int) * P->nig);
}
void orc_corner_nodes(void *h, int *out) {
  oprob *P = (oprob *)h;
  int *mark = imalloc(P->nnod_g);
  for (int is = 0; is < P->nsub; is++)
    for (int c = 0; c < P->s[is].ncorner; c++) mark[P->s[is].icgn[c]] = 1;
  int n = 0;
  for (int k = 0; k < P->nnod_g; k++)
    if (mark[k]) out[n++] = k;
  free(mark);
}
void orc_sub_sizes(void *h, int isub, int *out) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  out[0] = s->ndofs;
  out[1] = s->ndofi;
  out[2] = s->ndofo;
  out[3] = s->nconstr;
  out[4] = s->nnodi;
  out[5] = s->ncorner;
  out[6] = s->nedges;
  out[7] = s->nfaces;
}
void orc_sub_interface_gdofs(void *h, int isub, int *out) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  memcpy(out, s->gdofi, sizeof(int) * s->ndofi);
}
void orc_sub_weights(void *h, int isub, double *out) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  memcpy(out, s->wi, sizeof(double) * s->ndofi);
}
void orc_sub_schur(void *h, int isub, double *out) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  memcpy(out, s->S, sizeof(double) * s->ndofi * s->ndofi);
}
void orc_sub_constraints(void *h, int isub, double *C, int *indrowc) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  memcpy(C, s->C, sizeof(double) * s->nconstr * s->ndofi);
  memcpy(indrowc, s->indrowc, sizeof(int) * s->nconstr);
}
void orc_sub_phi(void *h, int isub, double *out) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  memcpy(out, s->phi, sizeof(double) * s->ndofi * s->nconstr);
}
void orc_sub_coarse(void *h, int isub, double *out) {
  oprob *P = (oprob *)h;
  osub *s = &P->s[isub - P->numbase];
  memcpy(out, s->acs, sizeof(double) * s->nconstr * s->nconstr);
}
static void unique_to_rep(oprob *P, const double *x, double **v) {
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->ndofi; k++) v[is][k] = x[P->gpos[s->gdofi[k]]];
  }
}
static void rep_to_unique(oprob *P, double **v, double *x) {
  for (int is = P->nsub - 1; is >= 0; is--) {
    osub *s = &P->s[is];
    for (int k = 0; k < s->ndofi; k++) x[P->gpos[s->gdofi[k]]] = v[is][k];
  }
}
void orc_apply_schur(void *h, const double *x, double *y) {
  oprob *P = (oprob *)h;
  double **a = alloc_rep(P), **b = alloc_rep(P);
  unique_to_rep(P, x, a);
  op_schur(P, a, b);
  rep_to_unique(P, b, y);
  free_rep(P, a);
  free_rep(P, b);
}
void orc_apply_pc(void *h, const double *r, double *z) {
  oprob *P = (oprob *)h;
  double **a = alloc_rep(P), **b = alloc_rep(P);
  unique_to_rep(P, r, a);
  op_pc(P, a, b);
  rep_to_unique(P, b, z);
  free_rep(P, a);
  free_rep(P, b);
}
void orc_reduced_rhs(void *h, double *g) {
  oprob *P = (oprob *)h;
  double **a = alloc_rep(P);
  for (int is = 0; is < P->nsub; is++) memcpy(a[is], P->s[is].g, sizeof(double) * P->s[is].ndofi);
  rep_to_unique(P, a, g);
  free_rep(P, a);
}
void orc_interface_solution(void *h, double *u) {
  oprob *P = (oprob *)h;
  if (P->uig) memcpy(u, P->uig, sizeof(double) * P->nig);
}
int orc_residual_history(void *h, double *out, int maxn) {
  oprob *P = (oprob *)h;
  int n = ONE_MIN(maxn, P->nhist);
  if (out) memcpy(out, P->hist, sizeof(double) * n);
  return P->nhist;
}

/* ---------------------------------------------------------- three levels */
typedef struct {
  int g, l;
} gnl;
static int cmp_gnl(const void *a, const void *b) {
  const gnl *x = (const gnl *)a, *y = (const gnl *)b;
  return (x->g > y->g) - (x->g < y->g);
}

void orc_coarse_nodes(void *h, int *nndfc, double *xyzc) {
  oprob *P = (oprob *)h;
  const int nc = P->nnodc, nd = P->ndim;
  char *done = (char *)xcalloc(nc > 0 ? nc : 1, 1);
  for (int k = 0; k < nc; k++) nndfc[k] = P->nndfc[k];
  for (int is = 0; is < P->nsub; is++) {
    const osub *s = &P->s[is];
    for (int c = 0; c < s->ncn; c++) {
      int id = s->cnid[c];
      if (done[id]) continue;
      done[id] = 1;
      if (c < s->ncorner) {
        int ln = -1;
        for (int k = 0; k < s->nnodi; k++)
          if (s->isngn[s->iin[k]] == s->icgn[c]) ln = s->iin[k];
        for (int d = 0; d < nd; d++) xyzc[id + (size_t)nc * d] = s->xyz[ln + (size_t)s->nnods * d];
      } else {
        int q = c - s->ncorner, m = s->gptr[q + 1] - s->gptr[q];
        gnl *v = (gnl *)xcalloc(m, sizeof(gnl));
        for (int k = 0; k < m; k++) {
          v[k].l = s->iin[s->gnod[s->gptr[q] + k]];
          v[k].g = s->isngn[v[k].l];
        }
        qsort(v, m, sizeof(gnl), cmp_gnl);
        for (int d = 0; d < nd; d++) {
          double sum = 0.0;
          for (int k = 0; k < m; k++) sum += s->xyz[v[k].l + (size_t)s->nnods * d];
          xyzc[id + (size_t)nc * d] = sum / m;
        }
        free(v);
      }
    }
  }
  free(done);
}

int orc_sub_cnodes(void *h, int isub, int *out) {
  oprob *P = (oprob *)h;
  const osub *s = &P->s[isub - P->numbase];
  if (out) memcpy(out, s->cnid, sizeof(int) * s->ncn);
  return s->ncn;
}

void orc_set_coarse_solver(void *h, orc_coarse_fn fn, void *ctx) {
  oprob *P = (oprob *)h;
  P->cfn = fn;
  P->cctx = ctx;
}

int orc_apply_level(void *h, const double *f, double *x) {
  oprob *P = (oprob *)h;
  if (P->bc_global) return -1;
  double **g = alloc_rep(P), **z = alloc_rep(P);
  double **fs = (double **)xcalloc(P->nsub, sizeof(double *));
  /* right-hand side of each subdomain, weighted at the interface (it is
   * complete), interior solve and reduced interface part */
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int *o2l = imalloc(s->ndofs), *i2l = imalloc(s->ndofs);
    sub_maps(s, o2l, i2l);
    fs[is] = dmalloc(s->ndofs);
    for (int k = 0; k < s->ndofs; k++) fs[is][k] = f[s->isvgvn[k]];
    double *fo = dmalloc(s->ndofo + 1), *t = dmalloc(s->ndofi + 1);
    for (int k = 0; k < s->ndofo; k++) fo[k] = fs[is][s->iovsvn[k]];
    if (s->ndofo > 0) {
      sky_solve(s, fo);
      sub_a21(s, o2l, i2l, fo, t);
    }
    for (int k = 0; k < s->ndofi; k++) g[is][k] = s->wi[k] * fs[is][s->iivsvn[k]] - t[k];
    free(fo);
    free(t);
    free(o2l);
    free(i2l);
  }
  exchange_sum(P, g);
  op_pc(P, g, z);
  /* interior recovery from the assembled interface values */
  for (int is = 0; is < P->nsub; is++) {
    osub *s = &P->s[is];
    int *o2l = imalloc(s->ndofs), *i2l = imalloc(s->ndofs);
    sub_maps(s, o2l, i2l);
    if (s->ndofo > 0) {
      double *xo = dmalloc(s->ndofo);
      sub_a12(s, o2l, i2l, z[is], xo);
      for (int k = 0; k < s->ndofo; k++) xo[k] = fs[is][s->iovsvn[k]] - xo[k];
      sky_solve(s, xo);
      for (int k = 0; k < s->ndofo; k++) x[s->isvgvn[s->iovsvn[k]]] = xo[k];
      free(xo);
    }
    for (int k = 0; k < s->ndofi; k++) x[s->isvgvn[s->iivsvn[k]]] = z[is][k];
    free(fs[is]);
    free(o2l);
    free(i2l);
  }
  free(fs);
  free_rep(P, g);
  free_rep(P, z);
  return 0;
}

/* ------------------------------------------------- general constraint rows */
/* Replace the constraint rows of subdomain isub by nc dense rows over its
 * interface dofs (C: nc x ndofi row major, the oracle's interface order) with
 * global coarse dofs indrowc (0-based), the new coarse problem having ndofc
 * dofs.  Used for adaptive constraints (dd_load_adaptive_constraints and
 * dd_orthogonalize_constraints, module_dd.f90:4386-4457, 4637-4802); the
 * coarse basis follows with orc_refresh_coarse. */
int orc_set_constraints(void *h, int isub, int nc, const double *C, const int *indrowc,
                        int ndofc) {
  oprob *P = (oprob *)h;
  int is = isub - P->numbase;
  if (is < 0 || is >= P->nsub || !P->is_setup) return -1;
  osub *s = &P->s[is];
  for (int r = 0; r < nc; r++)
    if (indrowc[r] < 0 || indrowc[r] >= ndofc) return -2;
  free(s->C);
  free(s->indrowc);
  s->nconstr = nc;
  s->C = dmalloc((size_t)nc * s->ndofi);
  memcpy(s->C, C, sizeof(double) * nc * s->ndofi);
  s->indrowc = imalloc(nc > 0 ? nc : 1);
  memcpy(s->indrowc, indrowc, sizeof(int) * nc);
  P->ndofc = ndofc;
  return 0;
}

/* Phi_i, Ac_i of every subdomain and the coarse factor for the current
 * constraint rows (dd_prepare_c, dd_prepare_aug, dd_prepare_coarse after the
 * adaptive constraints are loaded: module_levels.f90:2938-2945) */
int orc_refresh_coarse(void *h) {
  oprob *P = (oprob *)h;
  if (!P->is_setup) return -1;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
  for (int is = 0; is < P->nsub; is++) err |= sub_coarse_basis(&P->s[is]) ? 1 : 0;
  if (err) return -2;
  return coarse_problem(P) ? -3 : 0;
}
