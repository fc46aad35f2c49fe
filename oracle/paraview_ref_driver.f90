! Test infrastructure only: writes one subdomain VTU file and the umbrella
! PVD file with the reference's own ParaView module (src/module_paraview.f90,
! compiled from the reference sources into oracle/_ref by the Makefile), in
! the sequence examples/poisson_on_cube.f90:925-965 exports a subdomain's
! solution.  Input: a list-directed file holding the prefix, then isub nsub
! nelem nnod meshdim ndim linet, then inet, nnet, xyz (column major) and the
! solution.  The library's VTU writer is compared with its output.
program paraview_ref_driver
   use module_paraview
   implicit none
   integer, parameter :: kr = kind(1.d0)
   character(len=256) :: infile, prefix
   integer :: isub, nsub, nelem, nnod, meshdim, ndim, linet, idin, idvtu
   integer, allocatable :: inet(:), nnet(:), subdomain(:)
   real(kr), allocatable :: xyz(:,:), sol(:)

   call get_command_argument(1, infile)
   open(newunit=idin, file=trim(infile), status='old', form='formatted')
   read(idin,*) prefix
   read(idin,*) isub, nsub, nelem, nnod, meshdim, ndim, linet
   allocate(inet(linet), nnet(nelem), xyz(nnod, ndim), sol(nnod), subdomain(nelem))
   read(idin,*) inet
   read(idin,*) nnet
   read(idin,*) xyz
   read(idin,*) sol
   close(idin)

   call paraview_open_subdomain_file(trim(prefix), isub, idvtu)
   call paraview_write_mesh(idvtu, nelem, nnod, meshdim, inet, linet, nnet, nelem, xyz, nnod, ndim)
   call paraview_open_celldata(idvtu)
   subdomain = isub
   call paraview_write_dataarray(idvtu, 1, 'subdomain', subdomain, nelem)
   call paraview_close_celldata(idvtu)
   call paraview_open_pointdata(idvtu)
   call paraview_write_dataarray(idvtu, 1, 'Solution', sol, nnod)
   call paraview_close_pointdata(idvtu)
   call paraview_finalize_file(idvtu)
   call paraview_close_subdomain_file(idvtu)
   call paraview_export_pvd_file(trim(prefix), nsub)
end program paraview_ref_driver
