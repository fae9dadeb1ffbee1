! Calls every drop-in procedure of the Fortran modules (SparseMatrix, vector_operations,
! numerics) on the 6x5x4 grid with the seeded inputs of tests/golden/matrix_and_kernels.json
! (x = U(-1,1) seed 101, y = seed 102, alpha = 0.7, beta = -1.3) and prints every result as
! 16-digit hex so that the Python test can compare bit for bit.
program dropin_selftest
    use iso_c_binding
    use constants
    use SparseMatrix
    use vector_operations
    use numerics
    implicit none
    interface
        integer(c_int) function hpccg_xstar_uniform(n, seed, out) bind(C, name='hpccg_xstar_uniform')
            import :: c_int, c_int64_t, c_double
            integer(c_int64_t), value :: n, seed
            real(c_double) :: out(*)
        end function
    end interface
    type(SparseMatrixT) :: A
    real(kind=dp), allocatable :: x0(:), b(:), xexact(:), x(:), y(:), w(:)
    real(kind=dp) :: d1, d2, normr, res
    integer :: n, niters, st

    call generate_matrix(6, 5, 4, A, x0, b, xexact)
    n = A%local_nrow
    allocate(x(n), y(n), w(n))
    st = hpccg_xstar_uniform(int(n, c_int64_t), 101_c_int64_t, x)
    st = hpccg_xstar_uniform(int(n, c_int64_t), 102_c_int64_t, y)
    call spmv(A, x, w)
    call emit('spmv', w, n)
    call waxpby(n, 0.7_dp, x, -1.3_dp, y, w)
    call emit('waxpby', w, n)
    w = y
    call wwpby(n, -1.3_dp, x, w)
    call emit('wwpby', w, n)
    w = y
    call wxbw(n, x, -1.3_dp, w)
    call emit('wxbw', w, n)
    call vcopy(n, x, w)
    call emit('vcopy', w, n)
    d1 = 0.0_dp
    d2 = 0.0_dp
    call ddot(n, x, y, d1)
    call ddot1(n, x, d2)
    write (*, '(a,1x,z16.16,1x,z16.16)') 'DOTS', d1, d2
    ! cg on b = A*1 from x0 = 0, then the reference's own error metric against x* = 1
    x0 = 0.0_dp
    call cg(A, b, x0, 40, 0.0_dp, niters, normr)
    call compute_residual(n, x0, xexact, res)
    write (*, '(a,1x,i0,1x,z16.16,1x,z16.16)') 'CG', niters, normr, res

contains
    subroutine emit(name, v, len)
        character(len=*), intent(in) :: name
        integer, intent(in) :: len
        real(kind=dp), intent(in) :: v(len)
        integer :: i
        write (*, '(a)', advance='no') 'VEC '//name
        do i = 1, len
            write (*, '(1x,z16.16)', advance='no') v(i)
        end do
        write (*, *)
    end subroutine emit
end program dropin_selftest
