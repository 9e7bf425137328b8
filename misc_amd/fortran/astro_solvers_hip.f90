! astro_solvers_hip.f90 -- drop-in replacement of the reference's solver module
! for a Fortran host.  It is a module named AStrO_solvers exporting
! getSparseLDLFact, solveSparseLDLFact and gMRes with the argument lists of
! Working/AStrO_solvers.f90:10-15, 83-90 and 128-134, forwarding each call to
! the C ABI of misc_amd/libastro_hip.so (include/astro_abi.h) through
! ISO_C_BINDING.  A host that compiles this file in place of
! Working/AStrO_solvers.f90 and links libastro_hip.so runs its factor, the
! preconditioner sweeps and GMRES on the GPU; nothing else changes.
!
! Stopping rule of gMRes: by default the reference's own, |P^-1 r| <= 1e-12
! absolute (the single-precision literal of Working/AStrO_solvers.f90:184)
! within maxIt/numVecs + 1 restart cycles.  Setting the environment variable
! ASTRO_GMRES_RTOL (e.g. 1e-12) selects the relative rule |P^-1 r| <= rtol
! |P^-1 r0| instead.  Either way a cycle ends as soon as the rotated residual
! estimate meets the target (the literal form keeps iterating a converged
! cycle and returns rounding noise, Working/AStrO_solvers.f90:187-245).
!
! Differences the caller can see: a device or argument error stops the run
! with its message (the reference has no status); running out of restart
! cycles prints a warning and returns the last iterate, as the reference's
! gMRes returns without a word.
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
        if (status .eq. 4) then
            write(*,'(A,A)', advance='no') what, ': warning: '
        else
            write(*,'(A,A,I0,A)', advance='no') what, ': status ', status, ': '
        endif
        do i = 1, 512
            if (msg(i) .eq. c_null_char) exit
            write(*,'(A)', advance='no') msg(i)
        enddo
        write(*,*)
        ! ASTRO_ERR_NO_CONVERGE: the iterate is returned, as the reference does
        if (status .eq. 4) return
        stop 1
    end subroutine astro_stop_on

    ! Relative tolerance from ASTRO_GMRES_RTOL, or 0 (the reference's rule).
    real(c_double) function gmres_rtol()
        character(len=64) :: v
        integer :: ln, st
        gmres_rtol = 0.0d0
        call get_environment_variable('ASTRO_GMRES_RTOL', v, ln, st)
        if (st .eq. 0 .and. ln .gt. 0) read(v, *, iostat=st) gmres_rtol
        if (st .ne. 0) gmres_rtol = 0.0d0
    end function gmres_rtol

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
        real(c_double) :: rtol, atol
        rtol = gmres_rtol()
        atol = 0.0d0
        if (rtol .le. 0.0d0) atol = dble(1e-12)
        call astro_stop_on(astro_gMRes(xVec, Amat, Acols, ARange, int(ASize, c_int64_t), ADim, Cmat, Ccols, CRange, &
                int(CSize, c_int64_t), CDim, Pmat, PRange, int(PSize, c_int64_t), bVec, numVecs, maxIt, &
                rtol, atol), 'gMRes')
    end subroutine gMRes

end module AStrO_solvers
