! vector_operations module: the reference's vector kernels (MPI/src/vector_operations.F90:40-169)
! with unchanged argument lists, executed on the GPU by libhpccg_hip.  The elementwise
! kernels round exactly as the reference's loops; ddot/ddot1 follow HPCCG_DOT_MODE.
module vector_operations
    use constants
    use iso_c_binding
    use hpccg_capi
    implicit none
contains
    subroutine ddot(n, x, y, r)
        integer, intent(in) :: n
        real(kind=dp), intent(in) :: x(:)
        real(kind=dp), intent(in) :: y(:)
        real(kind=dp), intent(inout) :: r
        call hpccg_check(c_hpccg_ddot(hpccg_context(), int(n, c_int32_t), x, y, r), 'ddot')
    end subroutine ddot

    subroutine ddot1(n, x, r)
        integer, intent(in) :: n
        real(kind=dp), intent(in) :: x(:)
        real(kind=dp), intent(inout) :: r
        call hpccg_check(c_hpccg_ddot1(hpccg_context(), int(n, c_int32_t), x, r), 'ddot1')
    end subroutine ddot1

    subroutine waxpby(n, alpha, x, beta, y, w)
        integer, intent(in) :: n
        real(kind=dp), intent(in) :: alpha
        real(kind=dp), allocatable, intent(in) :: x(:)
        real(kind=dp), intent(in) :: beta
        real(kind=dp), allocatable, intent(in) :: y(:)
        real(kind=dp), allocatable, intent(inout) :: w(:)
        call hpccg_check(c_hpccg_waxpby(hpccg_context(), int(n, c_int32_t), alpha, x, beta, y, w), 'waxpby')
    end subroutine waxpby

    subroutine wwpby(n, beta, y, w)
        integer, intent(in) :: n
        real(kind=dp), allocatable, intent(in) :: y(:)
        real(kind=dp), intent(in) :: beta
        real(kind=dp), allocatable, intent(inout) :: w(:)
        call hpccg_check(c_hpccg_wwpby(hpccg_context(), int(n, c_int32_t), beta, y, w), 'wwpby')
    end subroutine wwpby

    subroutine wxbw(n, x, beta, w)
        integer, intent(in) :: n
        real(kind=dp), allocatable, intent(in) :: x(:)
        real(kind=dp), intent(in) :: beta
        real(kind=dp), allocatable, intent(inout) :: w(:)
        call hpccg_check(c_hpccg_wxbw(hpccg_context(), int(n, c_int32_t), x, beta, w), 'wxbw')
    end subroutine wxbw

    subroutine vcopy(n, x, y)
        integer, intent(in) :: n
        real(kind=dp), allocatable, intent(in) :: x(:)
        real(kind=dp), allocatable, intent(inout) :: y(:)
        call hpccg_check(c_hpccg_vcopy(hpccg_context(), int(n, c_int32_t), x, y), 'vcopy')
    end subroutine vcopy
end module vector_operations
