! Reads the problem files NAME.PAR, NAME.ML, NAME.GMIS, NAME.FVS and
! NAME.RHS with the READ statements of the reference's file-driven driver
! (list-directed, one statement per array; the RHS unformatted), then writes
! what it read to OUT.PAR ... OUT.RHS with list-directed WRITE statements and
! an unformatted record, as a Fortran preprocessor would produce them.
! Usage: problem_files_f NAME OUT
program problem_files_f
   implicit none
   integer, parameter :: kr = kind(1.d0)
   character(len=256) :: name, out
   integer :: ndim, nsub, nelem, ndof, nnod, nnodc, linet, maxit, ndecrmax, meshdim, nlevels
   real(kr) :: tol
   integer :: id
   integer, allocatable :: inet(:), nnet(:), nndf(:), ifix(:), nsublev(:)
   real(kr), allocatable :: xyz(:,:), fixv(:), rhs(:)

   call get_command_argument(1, name)
   call get_command_argument(2, out)

   open(newunit=id, file=trim(name)//'.PAR', status='old', form='formatted')
   read(id,*) ndim, nsub, nelem, ndof, nnod, nnodc, linet, tol, maxit, ndecrmax
   read(id,*,err=18,end=18) meshdim
   goto 19
18 meshdim = ndim
19 close(id)

   open(newunit=id, file=trim(name)//'.ML', status='old', form='formatted')
   read(id,*) nlevels
   allocate(nsublev(nlevels))
   read(id,*) nsublev
   close(id)

   allocate(inet(linet), nnet(nelem), nndf(nnod), xyz(nnod, ndim), ifix(ndof), fixv(ndof), rhs(ndof))
   open(newunit=id, file=trim(name)//'.GMIS', status='old', form='formatted')
   read(id,*) inet
   read(id,*) nnet
   read(id,*) nndf
   read(id,*) xyz
   close(id)

   open(newunit=id, file=trim(name)//'.FVS', status='old', form='formatted')
   read(id,*) ifix
   read(id,*) fixv
   close(id)

   open(newunit=id, file=trim(name)//'.RHS', status='old', form='unformatted')
   read(id) rhs
   close(id)

   open(newunit=id, file=trim(out)//'.PAR', status='replace', form='formatted')
   write(id,*) ndim, nsub, nelem, ndof, nnod, nnodc, linet, tol, maxit, ndecrmax
   write(id,*) meshdim
   close(id)
   open(newunit=id, file=trim(out)//'.ML', status='replace', form='formatted')
   write(id,*) nlevels
   write(id,*) nsublev
   close(id)
   open(newunit=id, file=trim(out)//'.GMIS', status='replace', form='formatted')
   write(id,*) inet
   write(id,*) nnet
   write(id,*) nndf
   write(id,*) xyz
   close(id)
   open(newunit=id, file=trim(out)//'.FVS', status='replace', form='formatted')
   write(id,*) ifix
   write(id,*) fixv
   close(id)
   open(newunit=id, file=trim(out)//'.RHS', status='replace', form='unformatted')
   write(id) rhs
   close(id)
end program problem_files_f
