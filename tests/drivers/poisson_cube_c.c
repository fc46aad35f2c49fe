/* C caller of the drop-in interface (include/bddcml_interface_c.h):
 * -laplace(u) = 1 on the unit cube, Q1 hexahedra, zero Dirichlet values,
 * split into ns^3 cubic subdomains of nes^3 elements, each built here and
 * uploaded as the reference's example harness does
 * (examples/poisson_on_cube.f90:404-519), with C numbering (numbase 0).
 *
 *   poisson_cube_c NES NS MODE OUT
 *
 * MODE 0: host plan only (no GPU: option plan_only), MODE 1: PCG solve to
 * 1e-10.  OUT receives the counts (corners, edges, faces, coarse dofs,
 * interface dofs), then iterations, converged reason, condition number and
 * the global solution (MODE 1). */
#include <stdio.h>
#include <stdlib.h>

#include "bddcml_amd.h"
#include "bddcml_interface_c.h"

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s NES NS MODE OUT\n", argv[0]);
    return 1;
  }
  int nes = atoi(argv[1]), ns = atoi(argv[2]), mode = atoi(argv[3]);
  int nel = nes * ns, n = nel + 1, nnod = n * n * n, ndof = nnod, nelem = nel * nel * nel;
  int nsub = ns * ns * ns;
  double h = 1.0 / nel, kref[8][8];
  static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++) {
      int q = (corner[a][0] != corner[b][0]) + (corner[a][1] != corner[b][1]) +
              (corner[a][2] != corner[b][2]);
      kref[a][b] = q == 0 ? h / 3.0 : (q == 1 ? 0.0 : -h / 12.0);
    }
  int nl = 2, nsublev[2] = {nsub, 1}, lnsublev = 2, nsub_loc = -1, comm = 0, verbose = 0,
      numbase = 0, jds = 0;
  bddcml_init_c(&nl, nsublev, &lnsublev, &nsub_loc, &comm, &verbose, &numbase, &jds);
  if (mode == 0 && bddcml_amd_set_option("plan_only", 1.0) != 0) return 2;
  int m = nes + 1, nnods = m * m * m, nelems = nes * nes * nes, ndofs = nnods, la = nelems * 36;
  int *inet = malloc(sizeof(int) * 8 * nelems), *nnet = malloc(sizeof(int) * nelems),
      *nndf = malloc(sizeof(int) * nnods), *isngn = malloc(sizeof(int) * nnods),
      *isegn = malloc(sizeof(int) * nelems), *ifix = malloc(sizeof(int) * ndofs),
      *ii = malloc(sizeof(int) * la), *jj = malloc(sizeof(int) * la);
  double *xyz = malloc(sizeof(double) * 3 * nnods), *fixv = calloc(ndofs, sizeof(double)),
         *rhs = malloc(sizeof(double) * ndofs), *sol = calloc(ndofs, sizeof(double)),
         *aa = malloc(sizeof(double) * la), none = 0.0;
  for (int isub = 0; isub < nsub; isub++) {
    int x0 = (isub % ns) * nes, y0 = (isub / ns % ns) * nes, z0 = isub / (ns * ns) * nes;
    for (int k = 0; k <= nes; k++)
      for (int j = 0; j <= nes; j++)
        for (int i = 0; i <= nes; i++) {
          int q = i + j * m + k * m * m, ig = x0 + i, jg = y0 + j, kg = z0 + k;
          isngn[q] = ig + jg * n + kg * n * n;
          xyz[q] = ig * h, xyz[nnods + q] = jg * h, xyz[2 * nnods + q] = kg * h;
          int onb = ig == 0 || ig == nel || jg == 0 || jg == nel || kg == 0 || kg == nel;
          ifix[q] = onb;
          rhs[q] = onb ? 0.0 : h * h * h;
          nndf[q] = 1;
        }
    int q = 0;
    for (int ek = 0; ek < nes; ek++)
      for (int ej = 0; ej < nes; ej++)
        for (int ei = 0; ei < nes; ei++) {
          int e = ei + ej * nes + ek * nes * nes, n1 = ei + ej * m + ek * m * m;
          int nod[8] = {n1, n1 + 1, n1 + 1 + m, n1 + m, n1 + m * m, n1 + 1 + m * m,
                        n1 + 1 + m + m * m, n1 + m + m * m};
          for (int a = 0; a < 8; a++) inet[8 * e + a] = nod[a];
          nnet[e] = 8;
          isegn[e] = (x0 + ei) + (y0 + ej) * nel + (z0 + ek) * nel * nel;
          for (int a = 0; a < 8; a++)
            for (int b = a; b < 8; b++, q++) {
              ii[q] = nod[a] < nod[b] ? nod[a] : nod[b];
              jj[q] = nod[a] < nod[b] ? nod[b] : nod[a];
              aa[q] = kref[a][b];
            }
        }
    int three = 3, isub_c = isub, linet = 8 * nelems, one = 1, zero = 0, mt = 1, nb = 3;
    bddcml_upload_subdomain_data_c(&nelem, &nnod, &ndof, &three, &three, &isub_c, &nelems, &nnods,
                                   &ndofs, inet, &linet, nnet, &nelems, nndf, &nnods, isngn,
                                   &nnods, isngn, &nnods, isegn, &nelems, xyz, &nnods, &three,
                                   ifix, &ndofs, fixv, &ndofs, rhs, &ndofs, &one, sol, &ndofs, &mt,
                                   ii, jj, aa, &la, &zero, &none, &zero, &zero, &none, &zero,
                                   &zero, &none, &zero, &one, &one, &nb);
  }
  int mt = 1, use_defaults = 0, pdiv = 1, corners = 1, arith = 1, adapt = 0, user = 0, wt = 0;
  bddcml_setup_preconditioner_c(&mt, &use_defaults, &pdiv, &corners, &arith, &adapt, &user, &wt);
  int counts[5];
  if (bddcml_amd_counts(counts) != 0) return 3;
  FILE *f = fopen(argv[4], "w");
  if (!f) return 4;
  fprintf(f, "%d %d %d %d %d\n", counts[0], counts[1], counts[2], counts[3], counts[4]);
  if (mode == 1) {
    int method = 0, maxit = 500, ndecr = 50, rec = 0, maxvec = 0, it = 0, reason = 0;
    double tol = 1e-10, cond = 0.0;
    bddcml_solve_c(&comm, &method, &tol, &maxit, &ndecr, &rec, &maxvec, &it, &reason, &cond);
    double *g = malloc(sizeof(double) * ndof);
    bddcml_download_global_solution_c(g, &ndof);
    fprintf(f, "%d %d %.17e\n", it, reason, cond);
    for (int d = 0; d < ndof; d++) fprintf(f, "%.17e\n", g[d]);
    free(g);
  }
  fclose(f);
  bddcml_finalize_c();
  free(inet), free(nnet), free(nndf), free(isngn), free(isegn), free(ifix), free(ii), free(jj);
  free(xyz), free(fixv), free(rhs), free(sol), free(aa);
  return 0;
}
