! ref_scalars.f90 -- test-only driver that runs the reference's scalar
! overloads r_sqrt and r_greater (Working/AStrO_r_overloadFunctions.f90:6-43)
! on inputs read from a binary file and writes their outputs, so that the
! tests can pin the IEEE sqrt and comparison that replace them.
!
! usage: astro_ref_scalars <in.f64> <out-prefix>
!   in.f64: n, then n sqrt arguments, then m, then m (a, b) pairs (all real*8,
!   counts stored as real*8); writes <out-prefix>sqrt.f64 (n) and
!   <out-prefix>greater.i32 (m).
program astro_ref_scalars
    use AStrO_constantVals
    use AStrO_r_overloadFunctions
    implicit none
    character(len=256) :: inFile, prefix
    real*8, allocatable :: x(:), s(:), ab(:,:)
    integer, allocatable :: g(:)
    real*8 :: cnt
    integer :: n, m, i, u
    call get_command_argument(1, inFile)
    call get_command_argument(2, prefix)
    open(newunit=u, file=trim(inFile), access='stream', form='unformatted', action='read')
    read(u) cnt
    n = int(cnt)
    allocate(x(n), s(n))
    read(u) x
    read(u) cnt
    m = int(cnt)
    allocate(ab(2, m), g(m))
    read(u) ab
    close(u)
    do i = 1, n
        call r_sqrt(s(i), x(i))
    enddo
    do i = 1, m
        call r_greater(g(i), ab(1, i), ab(2, i))
    enddo
    open(newunit=u, file=trim(prefix)//'sqrt.f64', access='stream', form='unformatted', status='replace')
    write(u) s
    close(u)
    open(newunit=u, file=trim(prefix)//'greater.i32', access='stream', form='unformatted', status='replace')
    write(u) g
    close(u)
end program astro_ref_scalars
