! module_bddcml -- Fortran interface of the MI355X BDDC library.
!
! Same module name, procedure names and dummy arguments as the reference's
! Fortran interface (src/module_bddcml.f90:36-770), so a Fortran caller of
! the reference rebuilds against this module (module files are compiler
! specific) and keeps its calls.  Every procedure hands its arguments by
! reference to the library's C entry point of the same name with suffix _c
! (include/bddcml_interface_c.h).  The reference converts its *_int flags to
! LOGICAL inside (integer2logical: nonzero is true); the C entry points take
! the integers themselves and apply the same rule, so they pass through.
! MPI communicators are accepted and ignored: one process drives one GPU
! holding every subdomain.
!
! Two extra procedures reach the library's own options and counts
! (include/bddcml_amd.h): bddcml_amd_set_option and bddcml_amd_counts.
module module_bddcml
   use, intrinsic :: iso_c_binding, only : c_int, c_double, c_char, c_null_char
   use, intrinsic :: iso_fortran_env, only : real64
   implicit none
   private

   integer, parameter :: kr = real64

   public :: bddcml_init, bddcml_upload_global_data, bddcml_upload_subdomain_data
   public :: bddcml_setup_preconditioner, bddcml_solve
   public :: bddcml_download_local_solution, bddcml_download_global_solution
   public :: bddcml_download_local_reactions, bddcml_download_global_reactions
   public :: bddcml_change_global_data, bddcml_change_subdomain_data
   public :: bddcml_setup_new_data, bddcml_dotprod_subdomain, bddcml_finalize
   public :: bddcml_amd_set_option, bddcml_amd_counts

   interface
      subroutine c_init(nl, nsublev, lnsublev, nsub_loc_1, comm_init, verbose_level, numbase, &
                        just_direct_solve_int) bind(c, name="bddcml_init_c")
         import :: c_int
         integer(c_int) :: nl, lnsublev, nsub_loc_1, comm_init, verbose_level, numbase
         integer(c_int) :: nsublev(*), just_direct_solve_int
      end subroutine

      subroutine c_upload_global_data(nelem, nnod, ndof, ndim, meshdim, inet, linet, nnet, lnnet, &
                                      nndf, lnndf, xyz, lxyz1, lxyz2, ifix, lifix, fixv, lfixv, &
                                      rhs, lrhs, sol, lsol, idelm, neighbouring, load_division_int) &
                                      bind(c, name="bddcml_upload_global_data_c")
         import :: c_int, c_double
         integer(c_int) :: nelem, nnod, ndof, ndim, meshdim, linet, lnnet, lnndf, lxyz1, lxyz2
         integer(c_int) :: lifix, lfixv, lrhs, lsol, idelm, neighbouring, load_division_int
         integer(c_int) :: inet(*), nnet(*), nndf(*), ifix(*)
         real(c_double) :: xyz(*), fixv(*), rhs(*), sol(*)
      end subroutine

      subroutine c_upload_subdomain_data(nelem, nnod, ndof, ndim, meshdim, isub, nelems, nnods, &
                                         ndofs, inet, linet, nnet, lnnet, nndf, lnndf, isngn, &
                                         lisngn, isvgvn, lisvgvn, isegn, lisegn, xyz, lxyz1, &
                                         lxyz2, ifix, lifix, fixv, lfixv, rhs, lrhs, &
                                         is_rhs_complete_int, sol, lsol, matrixtype, i_sparse, &
                                         j_sparse, a_sparse, la, is_assembled_int, &
                                         user_constraints, luser_constraints1, &
                                         luser_constraints2, element_data, lelement_data1, &
                                         lelement_data2, dof_data, ldof_data, &
                                         find_components_int, use_dual_mesh_graph_int, &
                                         neighbouring) &
                                         bind(c, name="bddcml_upload_subdomain_data_c")
         import :: c_int, c_double
         integer(c_int) :: nelem, nnod, ndof, ndim, meshdim, isub, nelems, nnods, ndofs
         integer(c_int) :: linet, lnnet, lnndf, lisngn, lisvgvn, lisegn, lxyz1, lxyz2
         integer(c_int) :: lifix, lfixv, lrhs, is_rhs_complete_int, lsol, matrixtype, la
         integer(c_int) :: is_assembled_int, luser_constraints1, luser_constraints2
         integer(c_int) :: lelement_data1, lelement_data2, ldof_data
         integer(c_int) :: find_components_int, use_dual_mesh_graph_int, neighbouring
         integer(c_int) :: inet(*), nnet(*), nndf(*), isngn(*), isvgvn(*), isegn(*), ifix(*)
         integer(c_int) :: i_sparse(*), j_sparse(*)
         real(c_double) :: xyz(*), fixv(*), rhs(*), sol(*), a_sparse(*)
         real(c_double) :: user_constraints(*), element_data(*), dof_data(*)
      end subroutine

      subroutine c_setup_preconditioner(matrixtype, use_defaults_int, parallel_division_int, &
                                        use_corner_constraints_int, &
                                        use_arithmetic_constraints_int, &
                                        use_adaptive_constraints_int, use_user_constraints_int, &
                                        weights_type) bind(c, name="bddcml_setup_preconditioner_c")
         import :: c_int
         integer(c_int) :: matrixtype, use_defaults_int, parallel_division_int
         integer(c_int) :: use_corner_constraints_int, use_arithmetic_constraints_int
         integer(c_int) :: use_adaptive_constraints_int, use_user_constraints_int, weights_type
      end subroutine

      subroutine c_solve(comm_all, method, tol, maxit, ndecrmax, recycling_int, &
                         max_number_of_stored_vectors, num_iter, converged_reason, &
                         condition_number) bind(c, name="bddcml_solve_c")
         import :: c_int, c_double
         integer(c_int) :: comm_all, method, maxit, ndecrmax, recycling_int
         integer(c_int) :: max_number_of_stored_vectors, num_iter, converged_reason
         real(c_double) :: tol, condition_number
      end subroutine

      subroutine c_download_local_solution(isub, sols, lsols) &
                                           bind(c, name="bddcml_download_local_solution_c")
         import :: c_int, c_double
         integer(c_int) :: isub, lsols
         real(c_double) :: sols(*)
      end subroutine

      subroutine c_download_global_solution(sol, lsol) &
                                            bind(c, name="bddcml_download_global_solution_c")
         import :: c_int, c_double
         integer(c_int) :: lsol
         real(c_double) :: sol(*)
      end subroutine

      subroutine c_download_local_reactions(isub, reas, lreas) &
                                            bind(c, name="bddcml_download_local_reactions_c")
         import :: c_int, c_double
         integer(c_int) :: isub, lreas
         real(c_double) :: reas(*)
      end subroutine

      subroutine c_download_global_reactions(rea, lrea) &
                                             bind(c, name="bddcml_download_global_reactions_c")
         import :: c_int, c_double
         integer(c_int) :: lrea
         real(c_double) :: rea(*)
      end subroutine

      subroutine c_change_global_data(ifix, lifix, fixv, lfixv, rhs, lrhs, sol, lsol) &
                                      bind(c, name="bddcml_change_global_data_c")
         import :: c_int, c_double
         integer(c_int) :: lifix, lfixv, lrhs, lsol
         integer(c_int) :: ifix(*)
         real(c_double) :: fixv(*), rhs(*), sol(*)
      end subroutine

      subroutine c_change_subdomain_data(isub, ifix, lifix, fixv, lfixv, rhs, lrhs, &
                                         is_rhs_complete_int, sol, lsol) &
                                         bind(c, name="bddcml_change_subdomain_data_c")
         import :: c_int, c_double
         integer(c_int) :: isub, lifix, lfixv, lrhs, is_rhs_complete_int, lsol
         integer(c_int) :: ifix(*)
         real(c_double) :: fixv(*), rhs(*), sol(*)
      end subroutine

      subroutine c_setup_new_data() bind(c, name="bddcml_setup_new_data_c")
      end subroutine

      subroutine c_dotprod_subdomain(isub, vec1, lvec1, vec2, lvec2, dotprod) &
                                     bind(c, name="bddcml_dotprod_subdomain_c")
         import :: c_int, c_double
         integer(c_int) :: isub, lvec1, lvec2
         real(c_double) :: vec1(*), vec2(*), dotprod
      end subroutine

      subroutine c_finalize() bind(c, name="bddcml_finalize_c")
      end subroutine

      function c_amd_set_option(key, value) result(rc) bind(c, name="bddcml_amd_set_option")
         import :: c_int, c_double, c_char
         character(kind=c_char) :: key(*)
         real(c_double), value :: value
         integer(c_int) :: rc
      end function

      function c_amd_counts(out) result(rc) bind(c, name="bddcml_amd_counts")
         import :: c_int
         integer(c_int) :: out(*)
         integer(c_int) :: rc
      end function
   end interface

contains

   subroutine bddcml_init(nl, nsublev, lnsublev, nsub_loc_1, comm_init, verbose_level, numbase, &
                          just_direct_solve_int)
      integer, intent(in) :: nl
      integer, intent(in) :: lnsublev
      integer, intent(in) :: nsublev(lnsublev)
      integer, intent(inout) :: nsub_loc_1
      integer, intent(in) :: comm_init
      integer, intent(in) :: verbose_level
      integer, intent(in) :: numbase
      integer, intent(in) :: just_direct_solve_int
      integer(c_int) :: a_nl, a_ns(lnsublev), a_lns, a_loc, a_comm, a_verb, a_base, a_jds
      a_nl = nl
      a_ns = nsublev
      a_lns = lnsublev
      a_loc = nsub_loc_1
      a_comm = comm_init
      a_verb = verbose_level
      a_base = numbase
      a_jds = just_direct_solve_int
      call c_init(a_nl, a_ns, a_lns, a_loc, a_comm, a_verb, a_base, a_jds)
      nsub_loc_1 = a_loc
   end subroutine

   subroutine bddcml_upload_global_data(nelem, nnod, ndof, ndim, meshdim, &
                                        inet, linet, nnet, lnnet, nndf, lnndf, xyz, lxyz1, lxyz2, &
                                        ifix, lifix, fixv, lfixv, rhs, lrhs, sol, lsol, idelm, &
                                        neighbouring, load_division_int)
      integer, intent(in) :: nelem, nnod, ndof, ndim, meshdim
      integer, intent(in) :: linet
      integer, intent(in) :: inet(linet)
      integer, intent(in) :: lnnet
      integer, intent(in) :: nnet(lnnet)
      integer, intent(in) :: lnndf
      integer, intent(in) :: nndf(lnndf)
      integer, intent(in) :: lxyz1, lxyz2
      real(kr), intent(in) :: xyz(lxyz1, lxyz2)
      integer, intent(in) :: lifix
      integer, intent(in) :: ifix(lifix)
      integer, intent(in) :: lfixv
      real(kr), intent(in) :: fixv(lfixv)
      integer, intent(in) :: lrhs
      real(kr), intent(in) :: rhs(lrhs)
      integer, intent(in) :: lsol
      real(kr), intent(in) :: sol(lsol)
      integer, intent(in) :: idelm
      integer, intent(in) :: neighbouring
      integer, intent(in) :: load_division_int
      call c_upload_global_data(ci(nelem), ci(nnod), ci(ndof), ci(ndim), ci(meshdim), &
                                ia(inet), ci(linet), ia(nnet), ci(lnnet), ia(nndf), ci(lnndf), &
                                da(reshape(xyz, [lxyz1 * lxyz2])), ci(lxyz1), ci(lxyz2), &
                                ia(ifix), ci(lifix), da(fixv), ci(lfixv), da(rhs), ci(lrhs), &
                                da(sol), ci(lsol), ci(idelm), ci(neighbouring), ci(load_division_int))
   end subroutine

   subroutine bddcml_upload_subdomain_data(nelem, nnod, ndof, ndim, meshdim, &
                                           isub, nelems, nnods, ndofs, &
                                           inet, linet, nnet, lnnet, nndf, lnndf, &
                                           isngn, lisngn, isvgvn, lisvgvn, isegn, lisegn, &
                                           xyz, lxyz1, lxyz2, &
                                           ifix, lifix, fixv, lfixv, &
                                           rhs, lrhs, is_rhs_complete_int, &
                                           sol, lsol, &
                                           matrixtype, i_sparse, j_sparse, a_sparse, la, &
                                           is_assembled_int, &
                                           user_constraints, luser_constraints1, &
                                           luser_constraints2, &
                                           element_data, lelement_data1, lelement_data2, &
                                           dof_data, ldof_data, &
                                           find_components_int, use_dual_mesh_graph_int, &
                                           neighbouring)
      integer, intent(in) :: nelem, nnod, ndof, ndim, meshdim
      integer, intent(in) :: isub, nelems, nnods, ndofs
      integer, intent(in) :: linet
      integer, intent(in) :: inet(linet)
      integer, intent(in) :: lnnet
      integer, intent(in) :: nnet(lnnet)
      integer, intent(in) :: lnndf
      integer, intent(in) :: nndf(lnndf)
      integer, intent(in) :: lisngn
      integer, intent(in) :: isngn(lisngn)
      integer, intent(in) :: lisvgvn
      integer, intent(in) :: isvgvn(lisvgvn)
      integer, intent(in) :: lisegn
      integer, intent(in) :: isegn(lisegn)
      integer, intent(in) :: lxyz1, lxyz2
      real(kr), intent(in) :: xyz(lxyz1, lxyz2)
      integer, intent(in) :: lifix
      integer, intent(in) :: ifix(lifix)
      integer, intent(in) :: lfixv
      real(kr), intent(in) :: fixv(lfixv)
      integer, intent(in) :: lrhs
      real(kr), intent(in) :: rhs(lrhs)
      integer, intent(in) :: is_rhs_complete_int
      integer, intent(in) :: lsol
      real(kr), intent(in) :: sol(lsol)
      integer, intent(in) :: matrixtype
      integer, intent(in) :: la
      integer, intent(in) :: i_sparse(la)
      integer, intent(in) :: j_sparse(la)
      real(kr), intent(in) :: a_sparse(la)
      integer, intent(in) :: is_assembled_int
      integer, intent(in) :: luser_constraints1, luser_constraints2
      real(kr), intent(in) :: user_constraints(luser_constraints1 * luser_constraints2)
      integer, intent(in) :: lelement_data1, lelement_data2
      real(kr), intent(in) :: element_data(lelement_data1 * lelement_data2)
      integer, intent(in) :: ldof_data
      real(kr), intent(in) :: dof_data(ldof_data)
      integer, intent(in) :: find_components_int
      integer, intent(in) :: use_dual_mesh_graph_int
      integer, intent(in) :: neighbouring
      real(c_double) :: none(1)
      none = 0.0_c_double
      ! empty optional arrays still need a valid address on the C side
      call c_upload_subdomain_data(ci(nelem), ci(nnod), ci(ndof), ci(ndim), ci(meshdim), &
                                   ci(isub), ci(nelems), ci(nnods), ci(ndofs), &
                                   ia(inet), ci(linet), ia(nnet), ci(lnnet), ia(nndf), ci(lnndf), &
                                   ia(isngn), ci(lisngn), ia(isvgvn), ci(lisvgvn), &
                                   ia(isegn), ci(lisegn), &
                                   da(reshape(xyz, [lxyz1 * lxyz2])), ci(lxyz1), ci(lxyz2), &
                                   ia(ifix), ci(lifix), da(fixv), ci(lfixv), da(rhs), ci(lrhs), &
                                   ci(is_rhs_complete_int), da(sol), ci(lsol), ci(matrixtype), &
                                   ia(i_sparse), ia(j_sparse), da(a_sparse), ci(la), &
                                   ci(is_assembled_int), &
                                   nonempty(user_constraints, none), ci(luser_constraints1), &
                                   ci(luser_constraints2), &
                                   nonempty(element_data, none), ci(lelement_data1), &
                                   ci(lelement_data2), &
                                   nonempty(dof_data, none), ci(ldof_data), &
                                   ci(find_components_int), ci(use_dual_mesh_graph_int), &
                                   ci(neighbouring))
   end subroutine

   subroutine bddcml_setup_preconditioner(matrixtype, use_defaults_int, parallel_division_int, &
                                          use_corner_constraints_int, &
                                          use_arithmetic_constraints_int, &
                                          use_adaptive_constraints_int, &
                                          use_user_constraints_int, weights_type)
      integer, intent(in) :: matrixtype
      integer, intent(in) :: use_defaults_int
      integer, intent(in) :: parallel_division_int
      integer, intent(in) :: use_corner_constraints_int
      integer, intent(in) :: use_arithmetic_constraints_int
      integer, intent(in) :: use_adaptive_constraints_int
      integer, intent(in) :: use_user_constraints_int
      integer, intent(in) :: weights_type
      call c_setup_preconditioner(ci(matrixtype), ci(use_defaults_int), ci(parallel_division_int), &
                                  ci(use_corner_constraints_int), &
                                  ci(use_arithmetic_constraints_int), &
                                  ci(use_adaptive_constraints_int), ci(use_user_constraints_int), &
                                  ci(weights_type))
   end subroutine

   subroutine bddcml_solve(comm_all, method, tol, maxit, ndecrmax, recycling_int, &
                           max_number_of_stored_vectors, num_iter, converged_reason, &
                           condition_number)
      integer, intent(in) :: comm_all
      integer, intent(in) :: method
      real(kr), intent(in) :: tol
      integer, intent(in) :: maxit
      integer, intent(in) :: ndecrmax
      integer, intent(in) :: recycling_int
      integer, intent(in) :: max_number_of_stored_vectors
      integer, intent(out) :: num_iter
      integer, intent(out) :: converged_reason
      real(kr), intent(out) :: condition_number
      integer(c_int) :: it, reason
      real(c_double) :: cond, t
      t = tol
      it = 0
      reason = 0
      cond = 0.0_c_double
      call c_solve(ci(comm_all), ci(method), t, ci(maxit), ci(ndecrmax), ci(recycling_int), &
                   ci(max_number_of_stored_vectors), it, reason, cond)
      num_iter = it
      converged_reason = reason
      condition_number = cond
   end subroutine

   subroutine bddcml_download_local_solution(isub, sols, lsols)
      integer, intent(in) :: isub
      integer, intent(in) :: lsols
      real(kr), intent(out) :: sols(lsols)
      call c_download_local_solution(ci(isub), sols, ci(lsols))
   end subroutine

   subroutine bddcml_download_global_solution(sol, lsol)
      integer, intent(in) :: lsol
      real(kr), intent(out) :: sol(lsol)
      call c_download_global_solution(sol, ci(lsol))
   end subroutine

   subroutine bddcml_download_local_reactions(isub, reas, lreas)
      integer, intent(in) :: isub
      integer, intent(in) :: lreas
      real(kr), intent(out) :: reas(lreas)
      call c_download_local_reactions(ci(isub), reas, ci(lreas))
   end subroutine

   subroutine bddcml_download_global_reactions(rea, lrea)
      integer, intent(in) :: lrea
      real(kr), intent(out) :: rea(lrea)
      call c_download_global_reactions(rea, ci(lrea))
   end subroutine

   subroutine bddcml_change_global_data(ifix, lifix, fixv, lfixv, rhs, lrhs, sol, lsol)
      integer, intent(in) :: lifix
      integer, intent(in) :: ifix(lifix)
      integer, intent(in) :: lfixv
      real(kr), intent(in) :: fixv(lfixv)
      integer, intent(in) :: lrhs
      real(kr), intent(in) :: rhs(lrhs)
      integer, intent(in) :: lsol
      real(kr), intent(in) :: sol(lsol)
      call c_change_global_data(ia(ifix), ci(lifix), da(fixv), ci(lfixv), da(rhs), ci(lrhs), &
                                da(sol), ci(lsol))
   end subroutine

   subroutine bddcml_change_subdomain_data(isub, ifix, lifix, fixv, lfixv, rhs, lrhs, &
                                           is_rhs_complete_int, sol, lsol)
      integer, intent(in) :: isub
      integer, intent(in) :: lifix
      integer, intent(in) :: ifix(lifix)
      integer, intent(in) :: lfixv
      real(kr), intent(in) :: fixv(lfixv)
      integer, intent(in) :: lrhs
      real(kr), intent(in) :: rhs(lrhs)
      integer, intent(in) :: lsol
      real(kr), intent(in) :: sol(lsol)
      integer, intent(in) :: is_rhs_complete_int
      call c_change_subdomain_data(ci(isub), ia(ifix), ci(lifix), da(fixv), ci(lfixv), da(rhs), &
                                   ci(lrhs), ci(is_rhs_complete_int), da(sol), ci(lsol))
   end subroutine

   subroutine bddcml_setup_new_data()
      call c_setup_new_data()
   end subroutine

   subroutine bddcml_dotprod_subdomain(isub, vec1, lvec1, vec2, lvec2, dotprod)
      integer, intent(in) :: isub
      integer, intent(in) :: lvec1
      real(kr), intent(in) :: vec1(lvec1)
      integer, intent(in) :: lvec2
      real(kr), intent(in) :: vec2(lvec2)
      real(kr), intent(out) :: dotprod
      real(c_double) :: d
      d = 0.0_c_double
      call c_dotprod_subdomain(ci(isub), da(vec1), ci(lvec1), da(vec2), ci(lvec2), d)
      dotprod = d
   end subroutine

   subroutine bddcml_finalize()
      call c_finalize()
   end subroutine

   ! library options (include/bddcml_amd.h); rc = 0 when accepted
   subroutine bddcml_amd_set_option(key, value, rc)
      character(*), intent(in) :: key
      real(kr), intent(in) :: value
      integer, intent(out) :: rc
      rc = c_amd_set_option(trim(key) // c_null_char, real(value, c_double))
   end subroutine

   ! counts(1:5) = corners, edges, faces, coarse dofs, interface dofs of the
   ! set-up problem; rc = 0 when the solver is set up
   subroutine bddcml_amd_counts(counts, rc)
      integer, intent(out) :: counts(5)
      integer, intent(out) :: rc
      integer(c_int) :: out(5)
      out = 0
      rc = c_amd_counts(out)
      counts = out
   end subroutine

   ! ---------------------------------------------------------------- helpers
   ! scalars and arrays as the C kinds (copies in temporaries when the
   ! default kinds differ; intent(in) data only)
   function ci(x) result(y)
      integer, intent(in) :: x
      integer(c_int) :: y
      y = int(x, c_int)
   end function

   function ia(x) result(y)
      integer, intent(in) :: x(:)
      integer(c_int) :: y(size(x))
      y = int(x, c_int)
   end function

   function da(x) result(y)
      real(kr), intent(in) :: x(:)
      real(c_double) :: y(size(x))
      y = real(x, c_double)
   end function

   function nonempty(x, fallback) result(y)
      real(kr), intent(in) :: x(:)
      real(c_double), intent(in) :: fallback(:)
      real(c_double), allocatable :: y(:)
      if (size(x) > 0) then
         y = real(x, c_double)
      else
         y = fallback
      end if
   end function

end module module_bddcml
