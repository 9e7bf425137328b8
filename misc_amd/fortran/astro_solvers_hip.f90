! astro_solvers_hip.f90 -- drop-in replacement of the reference's solver module
! for a Fortran host.  It is a module named AStrO_solvers exporting
! getSparseLDLFact, solveSparseLDLFact and gMRes with the argument lists of
! Working/AStrO_solvers.f90:10-15, 83-90 and 128-134, forwarding each call to
! the C ABI of misc_amd/libastro_hip.so (include/astro_abi.h) through
! ISO_C_BINDING.  A host that compiles this file in place of
! Working/AStrO_solvers.f90 and links libastro_hip.so runs its factor, the
! preconditioner sweeps and GMRES on the GPU; nothing else changes.
!
! Differences the caller can see: a nonzero status from the library stops the
! run with its message (the reference has no status); gMRes ends on the
! stable rule |P^-1 r| <= 1e-12 |P^-1 r0| instead of the absolute 1e-12 of
! Working/AStrO_solvers.f90:184 (see the library documentation).
module AStrO_solvers
    use, intrinsic :: iso_c_binding
    use AStrO_constantVals
    implicit none

    interface
        integer(c_int) function astro_getSparseLDLFact(ALTri, ASize, ARange, n) bind(C, name='astro_getSparseLDLFact')
            import :: c_int, c_int64_t, c_double
            real(c_double), intent(inout) :: ALTri(*)
            integer(c_int64_t), value :: ASize
            integer(c_int), intent(in) :: ARange(*)
            integer(c_int), value :: n
        end function
        integer(c_int) function astro_solveSparseLDLFact(soln, rhs, Lmat, LSize, LRange, Ldim, zVec) &
                bind(C, name='astro_solveSparseLDLFact')
            import :: c_int, c_int64_t, c_double
            real(c_double), intent(out) :: soln(*)
            real(c_double), intent(in) :: rhs(*), Lmat(*)
            integer(c_int64_t), value :: LSize
            integer(c_int), intent(in) :: LRange(*)
            integer(c_int), value :: Ldim
            real(c_double), intent(out) :: zVec(*)
        end function
        integer(c_int) function astro_gMRes(xVec, Amat, Acols, ARange, ASize, ADim, Cmat, Ccols, CRange, CSize, &
                CDim, Pmat, PRange, PSize, bVec, numVecs, maxIt, rtol, atol) bind(C, name='astro_gMRes')
            import :: c_int, c_int64_t, c_double
            real(c_double), intent(inout) :: xVec(*)
            real(c_double), intent(in) :: Amat(*), Cmat(*), Pmat(*), bVec(*)
            integer(c_int), intent(in) :: Acols(*), ARange(*), Ccols(*), CRange(*), PRange(*)
            integer(c_int64_t), value :: ASize, CSize, PSize
            integer(c_int), value :: ADim, CDim, numVecs, maxIt
            real(c_double), value :: rtol, atol
        end function
        type(c_ptr) function astro_last_error() bind(C, name='astro_last_error')
            import :: c_ptr
        end function
    end interface

contains

    subroutine astro_stop_on(status, what)
        integer(c_int), intent(in) :: status
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: i
        if (status .eq. 0) return
        call c_f_pointer(astro_last_error(), msg, [512])
        write(*,'(A,A,I0,A)', advance='no') what, ': status ', status, ': '
        do i = 1, 512
            if (msg(i) .eq. c_null_char) exit
            write(*,'(A)', advance='no') msg(i)
        enddo
        write(*,*)
        stop 1
    end subroutine astro_stop_on

    ! Working/AStrO_solvers.f90:10-15
    subroutine getSparseLDLFact(ALTri, ASize, ARange, n)
        integer, intent(in) :: ASize, n
        integer, intent(in) :: ARange(0:n)
        real*8, intent(inout) :: ALTri(ASize)
        call astro_stop_on(astro_getSparseLDLFact(ALTri, int(ASize, c_int64_t), ARange, n), 'getSparseLDLFact')
    end subroutine getSparseLDLFact

    ! Working/AStrO_solvers.f90:83-90
    subroutine solveSparseLDLFact(soln, rhs, Lmat, LSize, LRange, Ldim, zVec)
        integer, intent(in) :: LSize, Ldim
        integer, intent(in) :: LRange(0:Ldim)
        real*8, intent(in) :: rhs(Ldim), Lmat(LSize)
        real*8, intent(out) :: soln(Ldim)
        real*8, intent(out) :: zVec(Ldim)
        call astro_stop_on(astro_solveSparseLDLFact(soln, rhs, Lmat, int(LSize, c_int64_t), LRange, Ldim, zVec), &
                           'solveSparseLDLFact')
    end subroutine solveSparseLDLFact

    ! Working/AStrO_solvers.f90:128-134
    subroutine gMRes(xVec,Amat,Acols,ARange,ASize,ADim,Cmat,Ccols,CRange,CSize,CDim,Pmat,PRange,PSize,bVec,numVecs,maxIt)
        integer, intent(in) :: ASize, ADim, CSize, CDim, PSize, numVecs, maxIt
        integer, intent(in) :: ACols(ASize), ARange(0:ADim), Ccols(CSize), CRange(0:CDim), PRange(0:ADim)
        real*8, intent(in) :: Amat(ASize), Cmat(CSize), Pmat(PSize), bVec(ADim)
        real*8, intent(out) :: xVec(ADim)
        call astro_stop_on(astro_gMRes(xVec, Amat, Acols, ARange, int(ASize, c_int64_t), ADim, Cmat, Ccols, CRange, &
                int(CSize, c_int64_t), CDim, Pmat, PRange, int(PSize, c_int64_t), bVec, numVecs, maxIt, &
                1.0d-12, 0.0d0), 'gMRes')
    end subroutine gMRes

end module AStrO_solvers
