! Fortran caller of the drop-in interface (module_bddcml): -laplace(u) = 1 on
! the unit cube, Q1 hexahedra, zero Dirichlet values, split into ns^3 cubic
! subdomains of nes^3 elements, each built here and uploaded the way the
! reference's example harness does (examples/poisson_on_cube.f90:404-519:
! numbering x fastest, right-hand side h^3 per free node, complete).
!
!   poisson_cube_f NES NS MODE OUT
!
! MODE 0: host plan only (no GPU: option plan_only), MODE 1: PCG solve to
! 1e-10.  OUT receives the counts (corners, edges, faces, coarse dofs,
! interface dofs), then iterations, converged reason, condition number and
! the global solution (MODE 1).
program poisson_cube_f
   use module_bddcml
   implicit none
   integer, parameter :: kr = kind(1.0d0)
   integer :: nes, ns, mode, nel, n, nnod, ndof, nelem, nsub
   integer :: isub, sx, sy, sz, x0, y0, z0, m, nnods, nelems, ndofs, i, j, k, e, a, b, q
   integer :: ig, jg, kg, ei, ej, ek, n1, la, rc, nsub_loc, it, reason, counts(5)
   integer, allocatable :: inet(:), nnet(:), nndf(:), isngn(:), isegn(:), ifix(:)
   integer, allocatable :: i_sparse(:), j_sparse(:)
   real(kr), allocatable :: xyz(:, :), fixv(:), rhs(:), sol(:), a_sparse(:), gsol(:)
   real(kr) :: h, kref(8, 8), cond, none(0)
   integer :: corner(3, 8), nod(8), nsublev(2)
   character(len=256) :: arg, out

   call get_command_argument(1, arg); read (arg, *) nes
   call get_command_argument(2, arg); read (arg, *) ns
   call get_command_argument(3, arg); read (arg, *) mode
   call get_command_argument(4, out)
   nel = nes * ns
   n = nel + 1
   nnod = n**3
   ndof = nnod
   nelem = nel**3
   nsub = ns**3
   h = 1.0_kr / nel
   ! Q1 Laplacian on a cube of edge h: h/3 on the diagonal, 0 between nodes
   ! on one element edge, -h/12 across a face or body diagonal
   corner = reshape([0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 0, 0, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1], [3, 8])
   do a = 1, 8
      do b = 1, 8
         q = count(corner(:, a) /= corner(:, b))
         if (q == 0) then
            kref(a, b) = h / 3.0_kr
         else if (q == 1) then
            kref(a, b) = 0.0_kr
         else
            kref(a, b) = -h / 12.0_kr
         end if
      end do
   end do

   nsublev = [nsub, 1]
   nsub_loc = -1
   call bddcml_init(2, nsublev, 2, nsub_loc, 0, 0, 1, 0)
   if (mode == 0) then
      call bddcml_amd_set_option("plan_only", 1.0_kr, rc)
      if (rc /= 0) stop 2
   end if
   m = nes + 1
   nnods = m**3
   nelems = nes**3
   ndofs = nnods
   la = nelems * 36
   allocate (inet(8 * nelems), nnet(nelems), nndf(nnods), isngn(nnods), isegn(nelems))
   allocate (ifix(ndofs), xyz(nnods, 3), fixv(ndofs), rhs(ndofs), sol(ndofs))
   allocate (i_sparse(la), j_sparse(la), a_sparse(la))
   do isub = 1, nsub
      sx = mod(isub - 1, ns)
      sy = mod((isub - 1) / ns, ns)
      sz = (isub - 1) / (ns * ns)
      x0 = sx * nes
      y0 = sy * nes
      z0 = sz * nes
      do k = 0, nes
         do j = 0, nes
            do i = 0, nes
               q = 1 + i + j * m + k * m * m
               ig = x0 + i
               jg = y0 + j
               kg = z0 + k
               isngn(q) = 1 + ig + jg * n + kg * n * n
               xyz(q, 1) = ig * h
               xyz(q, 2) = jg * h
               xyz(q, 3) = kg * h
               if (ig == 0 .or. ig == nel .or. jg == 0 .or. jg == nel .or. kg == 0 .or. kg == nel) then
                  ifix(q) = 1
                  rhs(q) = 0.0_kr
               else
                  ifix(q) = 0
                  rhs(q) = h**3
               end if
            end do
         end do
      end do
      nndf = 1
      nnet = 8
      fixv = 0.0_kr
      sol = 0.0_kr
      q = 0
      do ek = 0, nes - 1
         do ej = 0, nes - 1
            do ei = 0, nes - 1
               e = 1 + ei + ej * nes + ek * nes * nes
               n1 = ei + ej * m + ek * m * m
               nod = 1 + [n1, n1 + 1, n1 + 1 + m, n1 + m, n1 + m * m, n1 + 1 + m * m, &
                          n1 + 1 + m + m * m, n1 + m + m * m]
               inet(8 * (e - 1) + 1:8 * e) = nod
               isegn(e) = 1 + (x0 + ei) + (y0 + ej) * nel + (z0 + ek) * nel * nel
               do a = 1, 8
                  do b = a, 8
                     q = q + 1
                     i_sparse(q) = min(nod(a), nod(b))
                     j_sparse(q) = max(nod(a), nod(b))
                     a_sparse(q) = kref(a, b)
                  end do
               end do
            end do
         end do
      end do
      call bddcml_upload_subdomain_data(nelem, nnod, ndof, 3, 3, isub, nelems, nnods, ndofs, &
                                        inet, size(inet), nnet, size(nnet), nndf, size(nndf), &
                                        isngn, size(isngn), isngn, size(isngn), isegn, size(isegn), &
                                        xyz, nnods, 3, ifix, ndofs, fixv, ndofs, rhs, ndofs, 1, &
                                        sol, ndofs, 1, i_sparse, j_sparse, a_sparse, la, 0, &
                                        none, 0, 0, none, 0, 0, none, 0, 1, 1, 3)
   end do
   call bddcml_setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
   call bddcml_amd_counts(counts, rc)
   if (rc /= 0) stop 3
   it = 0
   reason = 0
   cond = 0.0_kr
   open (unit=11, file=trim(out), status="replace", action="write")
   write (11, '(5i10)') counts
   if (mode == 1) then
      call bddcml_solve(0, 0, 1.0e-10_kr, 500, 50, 0, 0, it, reason, cond)
      allocate (gsol(ndof))
      call bddcml_download_global_solution(gsol, ndof)
      write (11, '(2i10, es26.17)') it, reason, cond
      do q = 1, ndof
         write (11, '(es26.17)') gsol(q)
      end do
   end if
   close (11)
   call bddcml_finalize()
end program poisson_cube_f
