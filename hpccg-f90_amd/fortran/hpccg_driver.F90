! Stand-alone driver for the GPU build: `hpccg nx ny nz`.  Same flow and stdout lines as the
! reference's HPCCGmain (MPI/src/main.F90:33-135): build the 27-point problem, print the
! banner, run cg with max_iter = 150 and tolerance = 0.  The reference's own main.F90 also
! compiles unchanged against these modules (link-substitution test).
program hpccg_driver
    use constants
    use SparseMatrix
    use numerics
    use timer
    implicit none
    character(len=32) :: arg
    type(SparseMatrixT) :: A
    real(kind=dp), allocatable :: x(:), b(:), xexact(:)
    integer :: dims(3), i, niters
    real(kind=dp) :: normr

    if (command_argument_count() /= 3) then
        print *, 'USAGE: ./HPCCG <nx> <ny> <nz>'
        stop
    end if
    do i = 1, 3
        call get_command_argument(i, arg)
        read (arg, *) dims(i)
    end do
    call timer_init()
    call generate_matrix(dims(1), dims(2), dims(3), A, x, b, xexact)
    print *, "HPCCG Benchmark"
    print *, "Using double precision"
    print *, "Total size (N x N):", A%total_nrow
    print *, "Total nonzeroes:", A%total_nnz
    print *, "MPI disabled."
    print *, "OpenMP disabled."
    niters = 0
    normr = 0.0
    call cg(A, b, x, 150, 0.0_dp, niters, normr)
end program hpccg_driver
